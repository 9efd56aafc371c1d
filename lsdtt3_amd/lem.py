"""LandscapeModel on the GPU, mirroring the reference's Python
``lsdtt3.LandscapeModel`` (python_bindings/bindings/landscape_model.cpp:
348-361, 430-546) for the D8 fluvial + ``roering_implicit`` hillslope
configuration.  Each ``step()`` is one ``run_lem_step``
(src/cpu/lsdtt_lem_cpu.cpp:285-732) executed by liblsdtt3_gpu.so on an FP64
elevation state that stays resident on the device between steps.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._lem_sig import LemOptions, LemStats


def default_options() -> LemOptions:
    o = LemOptions()
    N.lib().lsdtt_gpu_lem_options_init(C.byref(o))
    return o


def build_boundary_flats(mask, boundary_conditions="oooo"):
    """lsdtt::build_boundary_flats (src/cpu/lsdtt_lem_cpu.cpp:735-808)."""
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    nr, nc = m.shape
    out = np.empty(nr * nc, np.int64)
    k = C.c_int64()
    N.check(N.lib().lsdtt_gpu_build_boundary_flats(nr, nc, N.ptr(m), boundary_conditions.encode()[:4],
                                                   N.ptr(out), C.byref(k)))
    return out[:k.value].copy()


class LandscapeModel:
    """GPU landscape evolution model.

    Parameter names follow the reference binding; ``hillslope_model`` may be
    "roering_implicit" (Newton; linear when nonlinear_hillslopes is False) or
    "none".  Extra keyword options map onto the Newton/Krylov settings of
    ``lsdtt_gpu_lem_options`` (newton_rel_tol, newton_abs_tol,
    krylov_rel_tol, ...).
    """

    def __init__(self, dem, cell_size, timestep_years=100.0, erodibility=1.0e-3, m_exponent=0.45,
                 n_exponent=1.0, uplift_rate=1.0e-4, hillslope_diffusivity=1.0e-3,
                 hillslope_model="roering_implicit", nonlinear_hillslopes=False,
                 critical_hillslope_slope=0.8, incision_threshold_specific_area=0.0,
                 boundary_conditions="oooo", fluvial_method="d8", mask=None, fixed_pixels=None,
                 baselevel_fall_rate=0.0, erodibility_field=None, **solver):
        z = np.ascontiguousarray(dem, dtype=np.float64)
        if z.ndim != 2:
            raise ValueError("dem must be 2-D")
        if hillslope_model not in ("roering_implicit", "none"):
            raise ValueError(f"hillslope_model {hillslope_model!r} is not built on the GPU path")
        if fluvial_method not in ("d8", "none"):
            raise ValueError(f"fluvial_method {fluvial_method!r} is not built on the GPU path")
        self.n_rows, self.n_cols = z.shape
        self.cell_size = float(cell_size)
        self.mask = (np.ones(z.shape, np.uint8) if mask is None
                     else np.ascontiguousarray(mask, dtype=np.uint8))
        o = default_options()
        o.timestep = float(timestep_years)
        o.erodibility = float(erodibility)
        o.m_exponent = float(m_exponent)
        o.n_exponent = float(n_exponent)
        o.uplift_rate = float(uplift_rate)
        o.hillslope_diffusivity = float(hillslope_diffusivity)
        o.enable_hillslope_transport = int(hillslope_model == "roering_implicit")
        o.enable_nonlinear_hillslopes = int(bool(nonlinear_hillslopes))
        o.critical_hillslope_slope = float(critical_hillslope_slope)
        o.incision_threshold_area = float(incision_threshold_specific_area)
        o.boundary_conditions = boundary_conditions.encode()[:4]
        o.enable_d8_fluvial_erosion = int(fluvial_method == "d8")
        o.baselevel_fall_rate = float(baselevel_fall_rate)
        for k, v in solver.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown option {k!r}")
            setattr(o, k, v)
        self.options = o
        L = N.lib()
        h = L.lsdtt_gpu_lem_create(self.n_rows, self.n_cols, self.cell_size, N.ptr(self.mask), C.byref(o))
        if not h:
            raise N.GpuError(N.ERR_ARG, (L.lsdtt_gpu_last_error() or b"").decode())
        self._h = h
        N.check(L.lsdtt_gpu_lem_set_elevation(h, N.ptr(z)))
        if erodibility_field is not None:
            self.set_erodibility_field(erodibility_field)
        self.fixed_pixels = (build_boundary_flats(self.mask, boundary_conditions) if fixed_pixels is None
                             else np.ascontiguousarray(fixed_pixels, np.int64))
        self.iteration = 0
        self.time_years = 0.0
        self.last_stats = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().lsdtt_gpu_lem_destroy(h)
            self._h = None

    @property
    def elevation(self):
        z = np.empty((self.n_rows, self.n_cols), np.float64)
        N.check(N.lib().lsdtt_gpu_lem_get_elevation(self._h, N.ptr(z)))
        return z

    @elevation.setter
    def elevation(self, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        N.check(N.lib().lsdtt_gpu_lem_set_elevation(self._h, N.ptr(z)))

    def elevation_device_ptr(self) -> int:
        """device address of the resident FP64 elevation"""
        return int(N.lib().lsdtt_gpu_lem_elevation_device(self._h))

    def set_channel_mask(self, channel):
        """explicit_channel_mask: cells != 0 are fixed channel pixels for the
        hillslope step (None clears it)."""
        if channel is None:
            N.check(N.lib().lsdtt_gpu_lem_set_channel_mask(self._h, None))
            return
        ch = np.ascontiguousarray(channel, dtype=np.uint8)
        if ch.shape != (self.n_rows, self.n_cols):
            raise ValueError("channel mask must have the DEM's shape")
        N.check(N.lib().lsdtt_gpu_lem_set_channel_mask(self._h, N.ptr(ch)))

    def set_erodibility_field(self, K):
        Kd = None if K is None else np.ascontiguousarray(K, dtype=np.float64).ravel()
        N.check(N.lib().lsdtt_gpu_lem_set_erodibility_field(self._h, N.ptr(Kd)))

    def step(self) -> dict:
        st = LemStats()
        fp = self.fixed_pixels
        N.check(N.lib().lsdtt_gpu_lem_step(self._h, N.ptr(fp) if fp.size else None, fp.size, C.byref(st)))
        self.iteration += 1
        self.time_years += self.options.timestep
        self.last_stats = st.as_dict()
        return self.last_stats

    def run(self, n_steps: int) -> dict:
        for _ in range(int(n_steps)):
            self.step()
        return self.last_stats

    # ---- per-kernel event timing (roofline measurement) -------------------
    def profile(self, enable: bool = True) -> None:
        """start (clearing totals) or stop HIP-event timing of the kernels"""
        N.check(N.lib().lsdtt_gpu_lem_profile(self._h, int(enable)))

    def profile_totals(self) -> dict:
        """{kernel group: (ms, launches, algorithmic bytes)}"""
        L = N.lib()
        out = {}
        for i in range(L.lsdtt_gpu_lem_profile_count(self._h)):
            name = C.create_string_buffer(64)
            ms, nl, b = C.c_double(), C.c_int64(), C.c_double()
            N.check(L.lsdtt_gpu_lem_profile_get(self._h, i, name, 64, C.byref(ms), C.byref(nl), C.byref(b)))
            out[name.value.decode()] = (ms.value, nl.value, b.value)
        return out
