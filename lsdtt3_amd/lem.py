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
from ._lem_sig import ConvergenceOptions, LemOptions, LemStats, TimestepDiagnosis, TimestepOptions


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


def make_options(timestep_years=100.0, erodibility=1.0e-3, m_exponent=0.45, n_exponent=1.0, uplift_rate=1.0e-4,
                 hillslope_diffusivity=1.0e-3, hillslope_model="roering_implicit", nonlinear_hillslopes=False,
                 critical_hillslope_slope=0.8, incision_threshold_specific_area=0.0, boundary_conditions="oooo",
                 fluvial_method="d8", baselevel_fall_rate=0.0, **solver) -> LemOptions:
    """lsdtt_gpu_lem_options from the reference binding's parameter names;
    extra keywords set fields of the struct by name."""
    if hillslope_model not in ("roering_implicit", "none"):
        raise ValueError(f"hillslope_model {hillslope_model!r} is not built on the GPU path")
    if fluvial_method not in ("d8", "dinf", "none"):
        raise ValueError(f"fluvial_method {fluvial_method!r} is not built on the GPU path")
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
    o.enable_dinf_fluvial_erosion = int(fluvial_method == "dinf")
    o.baselevel_fall_rate = float(baselevel_fall_rate)
    for k, v in solver.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k!r}")
        setattr(o, k, v)
    return o


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
        o = make_options(timestep_years, erodibility, m_exponent, n_exponent, uplift_rate, hillslope_diffusivity,
                         hillslope_model, nonlinear_hillslopes, critical_hillslope_slope,
                         incision_threshold_specific_area, boundary_conditions, fluvial_method, baselevel_fall_rate,
                         **solver)
        self.n_rows, self.n_cols = z.shape
        self.cell_size = float(cell_size)
        self.mask = (np.ones(z.shape, np.uint8) if mask is None
                     else np.ascontiguousarray(mask, dtype=np.uint8))
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

    def run(self, n_steps: int, stats_csv=None) -> dict:
        """n_steps steps; with ``stats_csv`` (a path) the per-step statistics
        table of lsdtt-lem (``formats.ElevationStatsCSV``) is written there."""
        table = None
        if stats_csv is not None:
            from .formats import ElevationStatsCSV
            table = ElevationStatsCSV.for_model(self)
        for _ in range(int(n_steps)):
            st = self.step()
            if table is not None:
                table.add(self, st)
        if table is not None:
            table.write(stats_csv)
        return self.last_stats

    @property
    def channel_mask(self):
        """the hillslope's Dirichlet channel set of the last step (uint8):
        the explicit mask, the D-inf area threshold's cells, or none"""
        out = np.empty((self.n_rows, self.n_cols), np.uint8)
        N.check(N.lib().lsdtt_gpu_lem_get_channel_mask(self._h, N.ptr(out)))
        return out

    @property
    def erosion_rate(self):
        """lsdtt_lem_step_stats::erosion_rate of the last step (float32,
        m/yr, uplift backed out; needs compute_erosion_rate_field=1)"""
        out = np.empty((self.n_rows, self.n_cols), np.float32)
        N.check(N.lib().lsdtt_gpu_lem_get_erosion_rate(self._h, N.ptr(out)))
        return out

    # ---- timestep control (src/cpu/lsdtt_lem_timestep_cpu.hpp) --------------
    def median_interior_gradient(self) -> float:
        """median_interior_gradient of the current surface (device)"""
        g = C.c_double(0.0)
        N.check(N.lib().lsdtt_gpu_lem_median_interior_gradient(self._h, C.byref(g)))
        return g.value

    def diagnose_timestep(self, uplift_rate=None, timestep_years=None, target_ratio=0.7) -> dict:
        """diagnose_lem_timestep for this model's (or the given) U and dt"""
        d = TimestepDiagnosis()
        U = self.options.uplift_rate if uplift_rate is None else uplift_rate
        dt = self.options.timestep if timestep_years is None else timestep_years
        N.check(N.lib().lsdtt_gpu_lem_diagnose_timestep(self._h, float(U), float(dt), float(target_ratio),
                                                        C.byref(d)))
        out = {k: getattr(d, k) for k, _ in TimestepDiagnosis._fields_}
        out["too_large"] = bool(out["too_large"])
        out["warning"] = d.warning.decode()
        return out

    @property
    def timestep(self) -> float:
        return self.options.timestep

    @timestep.setter
    def timestep(self, dt: float) -> None:
        N.check(N.lib().lsdtt_gpu_lem_set_timestep(self._h, float(dt)))
        self.options.timestep = float(dt)

    # ---- tile shape of the routing sweeps (speed only) ----------------------
    @property
    def tile_height(self) -> int:
        """rows of the 32-column tiles of the count and stream-power sweeps"""
        return int(N.lib().lsdtt_gpu_lem_tile_height(self._h))

    @tile_height.setter
    def tile_height(self, th: int) -> None:
        N.check(N.lib().lsdtt_gpu_lem_set_tile_height(self._h, int(th)))

    def autotune_tiles(self, steps: int = 2):
        """times `steps` steps with each tile height (8, 16, 32) from the
        current surface (restored after each trial) and keeps the fastest;
        returns (height, {height: ms per step})"""
        th = C.c_int32(0)
        ms = (C.c_double * 3)()
        fp = self.fixed_pixels
        N.check(N.lib().lsdtt_gpu_lem_autotune_tiles(self._h, N.ptr(fp) if fp.size else None, fp.size,
                                                     int(steps), C.byref(th), ms))
        return th.value, {h: ms[i] for i, h in enumerate((8, 16, 32))}

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


class TimestepController:
    """lsdtt_lem_timestep_controller on a LandscapeModel's resident surface:
    ``update(model, max_local_dz)`` returns the next timestep."""

    def __init__(self, dt_initial, **options):
        o = TimestepOptions()
        N.lib().lsdtt_gpu_timestep_options_init(C.byref(o))
        for k, v in options.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown option {k!r}")
            setattr(o, k, v)
        self._h = N.lib().lsdtt_gpu_timestep_controller_create(float(dt_initial), C.byref(o))
        if not self._h:
            raise N.GpuError(N.ERR_ARG, (N.lib().lsdtt_gpu_last_error() or b"").decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().lsdtt_gpu_timestep_controller_destroy(h)
            self._h = None

    def update(self, model, max_local_dz):
        dt = C.c_double(0.0)
        N.check(N.lib().lsdtt_gpu_timestep_controller_update(self._h, model._h, float(max_local_dz), C.byref(dt)))
        return dt.value

    def state(self) -> dict:
        v = [C.c_double() for _ in range(3)] + [C.c_int64() for _ in range(3)]
        N.check(N.lib().lsdtt_gpu_timestep_controller_get(self._h, *[C.byref(x) for x in v]))
        names = ("dt", "last_ratio", "reference_slope", "n_shrinks", "n_grows", "n_changes")
        return {k: x.value for k, x in zip(names, v)}


class ConvergenceMonitor:
    """lsdtt_lem_convergence_monitor: ``update(max_dz, dt_used)`` returns
    "running", "converged" or "stalled"."""

    STATES = ("running", "converged", "stalled")

    def __init__(self, dt_initial, **options):
        o = ConvergenceOptions()
        N.lib().lsdtt_gpu_convergence_options_init(C.byref(o))
        for k, v in options.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown option {k!r}")
            setattr(o, k, v)
        self._h = N.lib().lsdtt_gpu_convergence_monitor_create(float(dt_initial), C.byref(o))
        if not self._h:
            raise N.GpuError(N.ERR_ARG, (N.lib().lsdtt_gpu_last_error() or b"").decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().lsdtt_gpu_convergence_monitor_destroy(h)
            self._h = None

    def update(self, max_dz, dt_used):
        st = C.c_int32(0)
        N.check(N.lib().lsdtt_gpu_convergence_monitor_update(self._h, float(max_dz), float(dt_used), C.byref(st)))
        return self.STATES[st.value]

    def state(self) -> dict:
        st, osc, ns = C.c_int32(), C.c_int32(), C.c_int64()
        ld, dec = C.c_double(), C.c_double()
        msg = C.create_string_buffer(512)
        N.check(N.lib().lsdtt_gpu_convergence_monitor_get(self._h, C.byref(st), C.byref(osc), C.byref(ns),
                                                          C.byref(ld), C.byref(dec), msg, 512))
        return dict(state=self.STATES[st.value], oscillating=bool(osc.value), n_steps=ns.value, last_dz=ld.value,
                    window_decay=dec.value, message=msg.value.decode())


class Ensemble:
    """Independent LEM instances (uneven grids, per-instance parameters)
    stepped together on the GPU (lsdtt_gpu_ensemble_*): every kernel of a
    step runs once for all instances and the step's launches replay from a
    captured hipGraph.  ``models`` is a list of dicts with ``dem``,
    ``cell_size``, optional ``mask`` and the LandscapeModel keyword
    options.  An instance's results do not depend on the other instances."""

    def __init__(self, models, workers=0):
        L = N.lib()
        self.shapes = []
        nr, nc, dx, masks = [], [], [], []
        opts = (LemOptions * len(models))()
        dems = []
        for i, m in enumerate(models):
            m = dict(m)
            z = np.ascontiguousarray(m.pop("dem"), np.float64)
            d = float(m.pop("cell_size"))
            mk = m.pop("mask", None)
            mk = np.ones(z.shape, np.uint8) if mk is None else np.ascontiguousarray(mk, np.uint8)
            opts[i] = make_options(**m)
            nr.append(z.shape[0])
            nc.append(z.shape[1])
            dx.append(d)
            masks.append(mk)
            dems.append(z)
            self.shapes.append(z.shape)
        self.n = len(models)
        mptr = (C.c_void_p * self.n)(*[mk.ctypes.data for mk in masks])
        h = L.lsdtt_gpu_ensemble_create(self.n, (C.c_int32 * self.n)(*nr), (C.c_int32 * self.n)(*nc),
                                        (C.c_double * self.n)(*dx), mptr, opts, int(workers))
        if not h:
            raise N.GpuError(N.ERR_ARG, (L.lsdtt_gpu_last_error() or b"").decode())
        self._h = h
        for i, z in enumerate(dems):
            N.check(L.lsdtt_gpu_ensemble_set_elevation(h, i, N.ptr(z)))
        self.n_valid = [int(mk.sum()) for mk in masks]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().lsdtt_gpu_ensemble_destroy(h)
            self._h = None

    def run(self, n_steps):
        """(list of last-step stats dicts, wall seconds)"""
        st = (LemStats * self.n)()
        sec = C.c_double(0.0)
        N.check(N.lib().lsdtt_gpu_ensemble_run(self._h, int(n_steps), st, C.byref(sec)))
        return [s.as_dict() for s in st], sec.value

    def set_graph(self, enable: bool) -> None:
        """graph replay of the step's launch pattern (default on)"""
        N.check(N.lib().lsdtt_gpu_ensemble_set_graph(self._h, int(bool(enable))))

    def set_warm_start(self, mode: int) -> None:
        """warm start of each instance's first Newton solve: 0 off, 3 the
        instance's previous step's hillslope change (WARM_PREVIOUS)"""
        N.check(N.lib().lsdtt_gpu_ensemble_set_warm_start(self._h, int(mode)))

    def pattern(self) -> dict:
        """launch-pattern counters: graph replays and captures, rounds and
        passes launched outside the pattern, restarted steps, current caps"""
        v = [C.c_int64() for _ in range(5)]
        caps = (C.c_int32 * 3)()
        N.check(N.lib().lsdtt_gpu_ensemble_pattern(self._h, *[C.byref(x) for x in v], caps))
        return dict(graph_launches=v[0].value, graph_captures=v[1].value, extra_rounds=v[2].value,
                    extra_passes=v[3].value, restarts=v[4].value, count_passes=caps[0], stream_power_passes=caps[1],
                    rounds=caps[2])

    @property
    def tile_height(self) -> int:
        return int(N.lib().lsdtt_gpu_ensemble_tile_height(self._h))

    @tile_height.setter
    def tile_height(self, th: int) -> None:
        N.check(N.lib().lsdtt_gpu_ensemble_set_tile_height(self._h, int(th)))

    def autotune_tiles(self, steps=2):
        """times `steps` steps of the ensemble per tile height (surfaces
        restored) and keeps the fastest; returns (height, {height: ms per step})"""
        th = C.c_int32(0)
        ms = (C.c_double * 3)()
        N.check(N.lib().lsdtt_gpu_ensemble_autotune_tiles(self._h, int(steps), C.byref(th), ms))
        return th.value, {h: ms[i] for i, h in enumerate((8, 16, 32))}

    def elevation(self, i):
        z = np.empty(self.shapes[i], np.float64)
        N.check(N.lib().lsdtt_gpu_ensemble_get_elevation(self._h, int(i), N.ptr(z)))
        return z

    def set_elevation(self, i, z):
        z = np.ascontiguousarray(z, np.float64)
        if z.shape != tuple(self.shapes[i]):
            raise ValueError(f"instance {i} is {self.shapes[i]}, got {z.shape}")
        N.check(N.lib().lsdtt_gpu_ensemble_set_elevation(self._h, int(i), N.ptr(z)))

    def erosion_rate(self, i):
        out = np.empty(self.shapes[i], np.float32)
        N.check(N.lib().lsdtt_gpu_ensemble_get_erosion_rate(self._h, int(i), N.ptr(out)))
        return out
