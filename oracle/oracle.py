"""ctypes front end of the CPU oracle (oracle/liblsdtt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the lsdtt3_amd package.  Each wrapper
names the C function it calls; the C sources cite the reference lines they
restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblsdtt_oracle.so")
_lib = None

i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_d8_receivers.argtypes = [C.c_int, C.c_int, C.c_double, f64p, u8p,
                                       C.c_char_p, C.c_int, C.c_int, C.c_void_p,
                                       i64p, u8p, C.POINTER(C.c_int64)]
        L.orc_d8_receivers.restype = C.c_int
        L.orc_donors_outlets.argtypes = [C.c_int, C.c_int, u8p, i64p, u8p, i64p]
        L.orc_donors_outlets.restype = C.c_int64
        L.orc_topo_kahn.argtypes = [C.c_int, C.c_int, u8p, i64p, u8p, i64p]
        L.orc_topo_kahn.restype = C.c_int64
        L.orc_s_vector.argtypes = [C.c_int, C.c_int, u8p, C.c_char_p, i64p, i64p,
                                   C.c_int64, i64p, i64p, u64p]
        L.orc_s_vector.restype = C.c_int64
        L.orc_accumulate.argtypes = [C.c_int64, u8p, i64p, i64p, C.c_int64,
                                     C.c_void_p, C.c_double, f64p]
        L.orc_accumulate.restype = None
        L.orc_upstream_counts.argtypes = [C.c_int64, u8p, i64p, i64p, C.c_int64, i64p]
        L.orc_upstream_counts.restype = None
        L.orc_fill_priority_flood.argtypes = [C.c_int, C.c_int, f32p, u8p,
                                              C.c_double, C.c_char_p, f32p]
        L.orc_fill_priority_flood.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        L.orc_max_threads.argtypes = []
        L.orc_max_threads.restype = C.c_int
        L.orc_fnv1a.argtypes = [C.c_uint64, C.c_void_p, C.c_int64]
        L.orc_fnv1a.restype = C.c_uint64
        L.orc_stream_power_d8.argtypes = [C.c_int64, u8p, i64p, u8p, i64p, C.c_int64, f64p,
                                          C.c_void_p, C.POINTER(SpParams), f64p, f64p]
        L.orc_stream_power_d8.restype = C.c_int64
        L.orc_hillslope_newton.argtypes = [C.c_int, C.c_int, C.c_double, u8p, C.c_char_p, C.c_void_p,
                                           C.POINTER(HsParams), C.c_double, f64p, f64p,
                                           C.POINTER(HsResult)]
        L.orc_hillslope_newton.restype = C.c_int
        L.orc_hillslope_transient.argtypes = [C.c_int, C.c_int, C.c_double, u8p, C.c_char_p,
                                              C.c_void_p, C.POINTER(HsParams), C.c_double,
                                              C.c_double, f64p, C.POINTER(HsResult)]
        L.orc_hillslope_transient.restype = C.c_int
        L.orc_boundary_flats.argtypes = [C.c_int, C.c_int, u8p, C.c_char_p, i64p]
        L.orc_boundary_flats.restype = C.c_int64
        L.orc_lem_step.argtypes = [C.c_int, C.c_int, C.c_double, u8p, C.POINTER(LemParams), i64p,
                                   C.c_int64, C.c_void_p, f64p, C.POINTER(LemStats)]
        L.orc_lem_step.restype = C.c_int
        L.orc_lem_step_ex.argtypes = [C.c_int, C.c_int, C.c_double, u8p, C.POINTER(LemParams), i64p,
                                      C.c_int64, C.c_void_p, f64p, C.POINTER(LemStats), C.c_void_p]
        L.orc_lem_step_ex.restype = C.c_int
        L.orc_picard_transient.argtypes = [C.c_int, C.c_int, C.c_double, u8p, C.c_char_p, C.c_void_p,
                                           C.POINTER(PicardParams), C.c_double, f64p, C.POINTER(PicardResult)]
        L.orc_picard_transient.restype = C.c_int
        L.orc_smooth_area.argtypes = [C.c_int, C.c_int, f64p, u8p, C.c_int, C.c_char_p, f64p]
        L.orc_smooth_area.restype = C.c_int
        L.orc_chi_raster.argtypes = [C.c_int, C.c_int, C.c_double, u8p, i64p, u8p, i64p, i64p, u64p,
                                     C.c_double, C.c_double, C.c_float, f32p, u8p]
        L.orc_chi_raster.restype = C.c_int
        L.orc_chi_threshold.argtypes = [C.c_int, C.c_int, C.c_double, u8p, i64p, u8p, i64p, i64p, u64p,
                                        C.c_double, C.c_double, C.c_int64, i64p, f32p, i64p, i64p,
                                        C.POINTER(C.c_int64)]
        L.orc_chi_threshold.restype = C.c_int64
        L.orc_dinf_graph.argtypes = [C.c_int, C.c_int, C.c_double, f32p, u8p, C.c_float, C.c_char_p, C.c_int,
                                     C.c_int, i64p, f32p, u8p, f32p, f32p]
        L.orc_dinf_graph.restype = C.c_int
        L.orc_dinf_topo.argtypes = [C.c_int64, u8p, i64p, f32p, u8p, i64p]
        L.orc_dinf_topo.restype = C.c_int64
        L.orc_dinf_accumulate.argtypes = [C.c_int64, i64p, f32p, u8p, i64p, C.c_int64, C.c_double, f64p]
        L.orc_dinf_accumulate.restype = None
        L.orc_channel_threshold_quantity.argtypes = [C.c_double, C.c_double, C.c_int]
        L.orc_channel_threshold_quantity.restype = C.c_double
        L.orc_dinf_channel_mask.argtypes = [C.c_int, C.c_int, C.c_double, f32p, u8p, C.c_char_p, C.c_double,
                                            C.c_int, C.c_int, u8p, C.c_void_p]
        L.orc_dinf_channel_mask.restype = C.c_int
        L.orc_stream_power_dinf.argtypes = [C.c_int64, u8p, i64p, f32p, u8p, i64p, C.c_int64, f64p, C.c_void_p,
                                            C.POINTER(SpParams), f64p, f64p]
        L.orc_stream_power_dinf.restype = C.c_int64
        L.orc_steady_state_d8.argtypes = [C.c_int, C.c_int, C.c_double, u8p, i64p, u8p, i64p, C.c_int64, f64p,
                                          C.c_void_p, C.POINTER(SsOptions), C.c_int, C.c_float, f32p,
                                          C.POINTER(SsResult)]
        L.orc_steady_state_d8.restype = C.c_int
        L.orc_steady_state_dinf.argtypes = [C.c_int, C.c_int, C.c_double, u8p, i64p, f32p, u8p, i64p, C.c_int64,
                                            f64p, C.c_void_p, C.POINTER(SsOptions), C.c_int, C.c_float, f32p,
                                            C.POINTER(SsResult)]
        L.orc_steady_state_dinf.restype = C.c_int
        L.orc_median_interior_gradient.argtypes = [C.c_int, C.c_int, C.c_double, f64p, u8p]
        L.orc_median_interior_gradient.restype = C.c_double
        L.orc_diagnose_timestep.argtypes = [C.c_int, C.c_int, C.c_double, f64p, u8p, C.c_double, C.c_double,
                                            C.c_double, C.POINTER(TsDiagnosis)]
        L.orc_diagnose_timestep.restype = None
        L.orc_ts_controller_init.argtypes = [C.POINTER(TsController), C.c_double, C.POINTER(TsOptions)]
        L.orc_ts_controller_init.restype = None
        L.orc_ts_controller_update.argtypes = [C.POINTER(TsController), C.c_double, C.c_double, C.c_double]
        L.orc_ts_controller_update.restype = C.c_double
        L.orc_ts_monitor_init.argtypes = [C.POINTER(TsMonitor), C.c_double, C.POINTER(TsMonitorOptions)]
        L.orc_ts_monitor_init.restype = None
        L.orc_ts_monitor_free.argtypes = [C.POINTER(TsMonitor)]
        L.orc_ts_monitor_free.restype = None
        L.orc_ts_monitor_update.argtypes = [C.POINTER(TsMonitor), C.c_double, C.c_double]
        L.orc_ts_monitor_update.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's parallel loops; returns the setting"""
    lib().orc_set_threads(int(n))
    return lib().orc_max_threads()


def fnv1a(h: int, data: bytes) -> int:
    """continue a 64-bit FNV-1a hash over `data`"""
    return lib().orc_fnv1a(h, data, len(data))


class OracleError(RuntimeError):
    pass


def d8_receivers(z, mask, dx, bc="oooo", interior_nodata_is_outlet=True,
                 allow_pits_as_outlets=False, forced=None):
    """orc_d8_receivers; z is float32 or float64, widened exactly to double."""
    nr, nc = z.shape
    zd = np.ascontiguousarray(z, dtype=np.float64)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    rcv = np.empty(nr * nc, np.int64)
    d = np.empty(nr * nc, np.uint8)
    fail = C.c_int64(-1)
    fplane = None
    fptr = None
    if forced is not None and len(forced):
        fplane = np.zeros(nr * nc, np.uint8)
        fplane[np.asarray(forced, np.int64)] = 1
        fptr = fplane.ctypes.data_as(C.c_void_p)
    st = lib().orc_d8_receivers(nr, nc, float(dx), zd.ravel(), m.ravel(),
                                bc.encode(), int(interior_nodata_is_outlet),
                                int(allow_pits_as_outlets), fptr, rcv, d,
                                C.byref(fail))
    if st != 0:
        f = fail.value
        raise OracleError(f"could not assign a receiver at row {f // nc}, col {f % nc}")
    return rcv, d


def flow_build(z, mask, dx, bc="oooo", interior_nodata_is_outlet=True,
               allow_pits_as_outlets=False, forced=None, topo=True, svector=True):
    """Full D8 flow object: dict of receiver, dir, donors, outlets, topo,
    s_vector, s_start, s_size (absent parts are None)."""
    nr, nc = z.shape
    n = nr * nc
    m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    rcv, d = d8_receivers(z, mask, dx, bc, interior_nodata_is_outlet,
                          allow_pits_as_outlets, forced)
    donors = np.empty(n, np.uint8)
    outlets = np.empty(max(n, 1), np.int64)
    no = lib().orc_donors_outlets(nr, nc, m, rcv, donors, outlets)
    if no < 0:
        raise OracleError("donor count failure")
    out = dict(receiver=rcv, dir=d, donors=donors, outlets=outlets[:no].copy(),
               topo=None, s_vector=None, s_start=None, s_size=None)
    if topo or svector:
        t = np.empty(max(n, 1), np.int64)
        nt = lib().orc_topo_kahn(nr, nc, m, rcv, donors, t)
        if nt != int(m.sum()):
            raise OracleError("routing graph is cyclic or incomplete")
        out["topo"] = t[:nt].copy()
    if svector:
        sv = np.empty(max(n, 1), np.int64)
        ss = np.empty(n, np.int64)
        sz = np.empty(n, np.uint64)
        ns = lib().orc_s_vector(nr, nc, m, bc.encode(), rcv, out["outlets"],
                                len(out["outlets"]), sv, ss, sz)
        if ns < 0:
            raise OracleError("s-vector failure")
        out["s_vector"], out["s_start"], out["s_size"] = sv[:ns].copy(), ss, sz
    return out


def chi_raster(flow, mask, dx, m_over_n=0.45, A0=1.0, nodata=-9999.0):
    """orc_chi_raster on a flow_build() dict: (float32 chi, uint8 chi_mask)."""
    nr, nc = mask.shape
    m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    chi = np.empty(nr * nc, np.float32)
    cm = np.empty(nr * nc, np.uint8)
    st = lib().orc_chi_raster(nr, nc, float(dx), m, flow["receiver"], flow["dir"], flow["s_vector"],
                              flow["s_start"], flow["s_size"], float(m_over_n), float(A0), float(nodata),
                              chi, cm)
    if st != 0:
        raise OracleError("chi extraction failed")
    return chi.reshape(nr, nc), cm.reshape(nr, nc)


def chi_threshold(flow, mask, dx, min_pixels, m_over_n=0.45, A0=1.0):
    """orc_chi_threshold: dict of flats, chi, receivers, basin_start."""
    nr, nc = mask.shape
    m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    nv = max(int(m.sum()), 1)
    flats = np.empty(nv, np.int64)
    chi = np.empty(nv, np.float32)
    rec = np.empty(nv, np.int64)
    bs = np.empty(nv, np.int64)
    nb = C.c_int64(0)
    k = lib().orc_chi_threshold(nr, nc, float(dx), m, flow["receiver"], flow["dir"], flow["s_vector"],
                                flow["s_start"], flow["s_size"], float(m_over_n), float(A0),
                                int(min_pixels), flats, chi, rec, bs, C.byref(nb))
    if k < 0:
        raise OracleError("thresholded chi extraction failed")
    return dict(flats=flats[:k].copy(), chi=chi[:k].copy(), receivers=rec[:k].copy(),
                basin_start=bs[:nb.value].copy())


def accumulate(mask, receiver, topo, dx, weights=None):
    m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    n = m.size
    acc = np.empty(n, np.float64)
    wptr = None
    if weights is not None:
        w = np.ascontiguousarray(weights, dtype=np.float64)
        wptr = w.ctypes.data_as(C.c_void_p)
    lib().orc_accumulate(n, m, receiver, topo, len(topo), wptr, float(dx), acc)
    return acc


def upstream_counts(mask, receiver, topo):
    m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    cnt = np.empty(m.size, np.int64)
    lib().orc_upstream_counts(m.size, m, receiver, topo, len(topo), cnt)
    return cnt


def fill(z, mask, nodata=-9999.0, bc="oooo"):
    """orc_fill_priority_flood: epsilon-free spill-level fill (float32)."""
    nr, nc = z.shape
    zf = np.ascontiguousarray(z, dtype=np.float32)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty_like(zf)
    st = lib().orc_fill_priority_flood(nr, nc, zf.ravel(), m.ravel(),
                                       float(nodata), bc.encode(), out.ravel())
    if st != 0:
        raise OracleError(f"fill failed (status {st}, bc {bc!r})")
    return out



# ---------------------------------------------------------------------------
# The reference's own fill, built from its sources by `make -C oracle ref`
# (oracle/ref_fill.cpp).  Absent unless the reference tree was present at build.
# ---------------------------------------------------------------------------
REF_FILL_PATH = os.path.join(_HERE, "_ref", "libref_fill.so")
_ref_lib = None


def ref_fill_available() -> bool:
    return os.path.exists(REF_FILL_PATH)


def ref_fill(z, mask, dx=1.0, bc="oooo", min_slope=0.0, n_threads=1, nodata=None, mode=0):
    """lsdtt::fill_pits(raster, min_slope, bc, n_threads) (mode 0, the call of
    src/cpu/lsdtt_lem_cpu.cpp:826) or fill_pits_priority_flood (mode 1) of the
    reference itself.  Cells equal to `nodata` join the invalid mask."""
    global _ref_lib
    if _ref_lib is None:
        L = C.CDLL(REF_FILL_PATH)
        L.ref_fill.argtypes = [C.c_int, C.c_int, C.c_double, f32p, u8p, C.c_float,
                               C.c_char_p, C.c_int, C.c_int, f32p]
        L.ref_fill.restype = C.c_int
        L.ref_fill_error.restype = C.c_char_p
        _ref_lib = L
    nr, nc = z.shape
    zf = np.ascontiguousarray(z, dtype=np.float32)
    m = np.ascontiguousarray(mask, dtype=np.uint8).copy()
    if nodata is not None:
        m[zf == np.float32(nodata)] = 0
    out = np.empty_like(zf)
    if _ref_lib.ref_fill(nr, nc, float(dx), zf.ravel(), m.ravel(), float(min_slope),
                         bc.encode(), int(n_threads), int(mode), out.ravel()) != 0:
        raise OracleError("reference fill: " + _ref_lib.ref_fill_error().decode())
    return out

# ---------------------------------------------------------------------------
# LEM pieces (oracle/orc_lem.c)
# ---------------------------------------------------------------------------
class SpParams(C.Structure):
    _fields_ = [("K", C.c_double), ("m", C.c_double), ("n", C.c_double), ("dt", C.c_double),
                ("dx", C.c_double), ("threshold", C.c_double), ("divide_area_by_dx", C.c_int),
                ("slope_clamp", C.c_double), ("width_factor", C.c_double)]


class HsParams(C.Structure):
    _fields_ = [("D", C.c_double), ("U", C.c_double), ("Sc", C.c_double), ("nonlinear", C.c_int),
                ("slope_ratio_cap", C.c_double), ("abs_tol", C.c_double), ("rel_tol", C.c_double),
                ("max_iterations", C.c_int), ("max_line_search", C.c_int)]


class HsResult(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("line_search_steps", C.c_int),
                ("substeps", C.c_int), ("rejected", C.c_int), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("tolerance", C.c_double),
                ("final_max_dz", C.c_double), ("max_slope_ratio", C.c_double)]


class PicardParams(C.Structure):
    _fields_ = [("D", C.c_double), ("U", C.c_double), ("Sc", C.c_double), ("nonlinear", C.c_int),
                ("picard_tol", C.c_double), ("picard_max_iter", C.c_int), ("relaxation_omega", C.c_double),
                ("slope_ratio_cap", C.c_double), ("dt_init", C.c_double), ("dt_min", C.c_double),
                ("dt_max", C.c_double), ("grow_factor", C.c_double), ("shrink_factor", C.c_double),
                ("dz_target", C.c_double)]


class PicardResult(C.Structure):
    _fields_ = [("substeps", C.c_int), ("rejected", C.c_int), ("picard_iters", C.c_int),
                ("limiter_fired", C.c_int), ("final_dt", C.c_double), ("max_slope_ratio", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def picard_params(D, Sc=0.8, nonlinear=False, U=0.0, picard_tol=1e-4, picard_max_iter=50, omega=1.0,
                  r_max=0.99, dt_init=0.0, dt_min=0.0, dt_max=0.0, grow=1.5, shrink=0.5, dz_target=0.0):
    return PicardParams(D, U, Sc, int(nonlinear), picard_tol, picard_max_iter, omega, r_max, dt_init, dt_min,
                        dt_max, grow, shrink, dz_target)


def picard_transient(z, mask, dx, bc, params, dt_total, channel=None):
    """orc_picard_transient on a copy of z: (z_new, result dict)"""
    zz = np.ascontiguousarray(z, np.float64).copy()
    nr, nc = zz.shape
    ch = None if channel is None else np.ascontiguousarray(channel, np.uint8).ravel()
    res = PicardResult()
    lib().orc_picard_transient(nr, nc, float(dx), np.ascontiguousarray(mask, np.uint8).ravel(), bc.encode(),
                               None if ch is None else ch.ctypes.data_as(C.c_void_p), C.byref(params),
                               float(dt_total), zz.ravel(), C.byref(res))
    return zz, res.as_dict()


class SsOptions(C.Structure):
    _fields_ = [("U", C.c_float), ("K", C.c_float), ("m", C.c_float), ("n", C.c_float), ("relief_mode", C.c_int),
                ("target_relief", C.c_float), ("baselevel_elevation", C.c_float), ("divide_area_by_dx", C.c_int),
                ("channel_threshold_area", C.c_float)]


class SsResult(C.Structure):
    _fields_ = [("K_used", C.c_float), ("max_unit_height", C.c_float), ("computed_relief", C.c_float),
                ("erodibility_field_scale", C.c_float), ("max_unit_height_at_channel_head", C.c_float),
                ("n_channel_pixels", C.c_int64), ("chi_c_reported", C.c_int)]


class TsDiagnosis(C.Structure):
    _fields_ = [("reference_slope", C.c_double), ("cell_relief_m", C.c_double), ("uplift_per_step_m", C.c_double),
                ("ratio", C.c_double), ("suggested_dt_yr", C.c_double), ("too_large", C.c_int)]


class TsOptions(C.Structure):
    _fields_ = [("target_ratio", C.c_double), ("dt_min_yr", C.c_double), ("dt_max_yr", C.c_double),
                ("grow_factor", C.c_double), ("shrink_factor", C.c_double), ("slope_refresh_interval", C.c_int),
                ("ladder_ratio", C.c_double), ("ratio_smoothing", C.c_double)]


def ts_options(**kw):
    """lsdtt_lem_timestep_controller_options defaults
    (src/cpu/lsdtt_lem_timestep_cpu.hpp:116-160) with overrides"""
    o = TsOptions(0.7, 0.0, 0.0, 1.15, 0.5, 25, 2.0, 0.15)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class TsController(C.Structure):
    _fields_ = [("o", TsOptions), ("dt", C.c_double), ("dt_raw", C.c_double), ("dt_initial", C.c_double),
                ("dt_min", C.c_double), ("dt_max", C.c_double), ("slope", C.c_double), ("last_ratio", C.c_double),
                ("ratio_ema", C.c_double), ("n_steps", C.c_long), ("n_shrinks", C.c_long), ("n_grows", C.c_long),
                ("n_changes", C.c_long), ("rung", C.c_int), ("lowest_rung", C.c_int)]


class TsMonitorOptions(C.Structure):
    _fields_ = [("dz_tolerance_m", C.c_double), ("dz_rate_tolerance_m_per_yr", C.c_double),
                ("patience_steps", C.c_int), ("min_dt_fraction", C.c_double), ("stall_window", C.c_int),
                ("stall_decay_threshold", C.c_double), ("oscillation_alternation_fraction", C.c_double)]


def ts_monitor_options(**kw):
    """lsdtt_lem_convergence_options defaults (.hpp:236-292) with overrides"""
    o = TsMonitorOptions(1e-4, 0.0, 5, 0.25, 1000, 0.99, 0.70)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class TsMonitor(C.Structure):
    _fields_ = [("o", TsMonitorOptions), ("dt_initial", C.c_double), ("last_dz", C.c_double), ("decay", C.c_double),
                ("n_steps", C.c_long), ("n_below", C.c_int), ("oscillating", C.c_int), ("state", C.c_int),
                ("window", C.c_int), ("n_hist", C.c_int), ("hist", C.c_void_p)]


class LemParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("K", C.c_double), ("m", C.c_double), ("n", C.c_double),
                ("threshold", C.c_double), ("slope_clamp", C.c_double), ("U", C.c_double),
                ("baselevel_fall", C.c_double), ("divide_area_by_dx", C.c_int),
                ("enable_fluvial", C.c_int), ("enable_hillslope", C.c_int),
                ("enable_uplift", C.c_int), ("bc", C.c_char * 4), ("hs", HsParams),
                ("area_kernel", C.c_int), ("width_factor", C.c_double), ("channel", C.c_void_p),
                ("force_picard", C.c_int), ("picard", PicardParams), ("regime", C.c_int),
                ("fluvial_dinf", C.c_int)]


class LemStats(C.Structure):
    _fields_ = [("delta_max_elevation", C.c_double), ("delta_mean_elevation", C.c_double),
                ("max_local_elevation_change", C.c_double),
                ("p999_local_elevation_change", C.c_double), ("mean_erosion_rate", C.c_double),
                ("max_erosion_rate", C.c_double), ("n_valid", C.c_int64),
                ("cells_eroded", C.c_int64), ("newton_iterations", C.c_int),
                ("newton_substeps", C.c_int), ("hillslope_residual", C.c_double),
                ("hillslope_initial_residual", C.c_double), ("hillslope_method", C.c_int),
                ("picard_iterations", C.c_int), ("picard_substeps", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def hs_params(D, Sc=0.8, nonlinear=False, U=0.0, rel_tol=1e-8, abs_tol=1e-12, cap=0.9999,
              max_iterations=50, max_line_search=30):
    return HsParams(D, U, Sc, int(nonlinear), cap, abs_tol, rel_tol, max_iterations, max_line_search)


def stream_power(z_old, z, mask, receiver, dirs, topo, area, K, m, n, dt, dx, threshold=0.0,
                 divide_area_by_dx=True, slope_clamp=1e-5, K_field=None):
    """orc_stream_power_d8 on a copy of z; returns (z_new, cells eroded)"""
    zz = np.ascontiguousarray(z, dtype=np.float64).ravel().copy()
    zo = np.ascontiguousarray(z_old, dtype=np.float64).ravel()
    m8 = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    p = SpParams(K, m, n, dt, dx, threshold, int(divide_area_by_dx), slope_clamp, 1.0)
    kf = None
    kptr = None
    if K_field is not None:
        kf = np.ascontiguousarray(K_field, dtype=np.float64).ravel()
        kptr = kf.ctypes.data_as(C.c_void_p)
    e = lib().orc_stream_power_d8(zz.size, m8, receiver, dirs, topo, len(topo),
                                  np.ascontiguousarray(area, np.float64).ravel(), kptr, C.byref(p), zo, zz)
    return zz.reshape(np.shape(z)), e


def stream_power_dinf(z_old, z, mask, graph, area, K, m, n, dt, dx, threshold=0.0, divide_area_by_dx=True,
                      slope_clamp=1e-5, K_field=None, width_factor=1.0):
    """orc_stream_power_dinf on a copy of z (graph from dinf_graph); returns
    (z_new, cells whose solve ran)"""
    zz = np.ascontiguousarray(z, dtype=np.float64).ravel().copy()
    zo = np.ascontiguousarray(z_old, dtype=np.float64).ravel()
    m8 = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    _, topo = dinf_accumulate(graph, m8, dx)
    p = SpParams(K, m, n, dt, dx, threshold, int(divide_area_by_dx), slope_clamp, float(width_factor))
    kf = None if K_field is None else np.ascontiguousarray(K_field, dtype=np.float64).ravel()
    e = lib().orc_stream_power_dinf(zz.size, m8, graph["rec"].ravel(), graph["w"].ravel(), graph["nrec"], topo,
                                    len(topo), np.ascontiguousarray(area, np.float64).ravel(),
                                    None if kf is None else kf.ctypes.data_as(C.c_void_p), C.byref(p), zo, zz)
    return zz.reshape(np.shape(z)), e


def hillslope_newton(z, zn, mask, dx, bc, params, inv_dt):
    """one backward-Euler Newton solve; returns (z_new, HsResult)"""
    zz = np.ascontiguousarray(z, dtype=np.float64).copy()
    nr, nc = zz.shape
    r = HsResult()
    lib().orc_hillslope_newton(nr, nc, float(dx), np.ascontiguousarray(mask, np.uint8).ravel(),
                               bc.encode(), None, C.byref(params), float(inv_dt),
                               np.ascontiguousarray(zn, np.float64).ravel(), zz.ravel(), C.byref(r))
    return zz, r


def hillslope_transient(z, mask, dx, bc, params, dt_total, dt_init=0.0):
    zz = np.ascontiguousarray(z, dtype=np.float64).copy()
    nr, nc = zz.shape
    r = HsResult()
    lib().orc_hillslope_transient(nr, nc, float(dx), np.ascontiguousarray(mask, np.uint8).ravel(),
                                  bc.encode(), None, C.byref(params), float(dt_total), float(dt_init),
                                  zz.ravel(), C.byref(r))
    return zz, r


def boundary_flats(mask, bc):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    nr, nc = m.shape
    out = np.empty(nr * nc, np.int64)
    k = lib().orc_boundary_flats(nr, nc, m.ravel(), bc.encode(), out)
    return out[:k].copy()


def lem_params(dt, K, m, n, U, D, Sc=0.8, nonlinear=False, bc="oooo", threshold=0.0,
               slope_clamp=1e-5, baselevel_fall=0.0, fluvial=True, hillslope=True, uplift=True,
               rel_tol=1e-8, abs_tol=1e-12, area_kernel=1, width_factor=1.0, channel=None,
               divide_area_by_dx=True, force_picard=False, picard=None, max_iterations=50, regime=0,
               fluvial_dinf=False):
    """channel: optional uint8 Dirichlet channel mask for the hillslope (kept
    alive on the returned structure).  regime: 0 automatic (channels as
    boundaries when a threshold or a mask is given), 1 coupled, 2 channels
    as boundaries; without a mask the channels are the cells whose D-inf
    drainage area passes the threshold."""
    ch = None if channel is None else np.ascontiguousarray(channel, np.uint8).ravel()
    p = LemParams(dt, K, m, n, threshold, slope_clamp, U, baselevel_fall, int(divide_area_by_dx), int(fluvial),
                  int(hillslope), int(uplift), bc.encode(),
                  hs_params(D, Sc, nonlinear, 0.0, rel_tol, abs_tol), int(area_kernel), float(width_factor),
                  None if ch is None else ch.ctypes.data_as(C.c_void_p), int(force_picard),
                  picard if picard is not None else picard_params(D, Sc, nonlinear))
    p.hs.max_iterations = int(max_iterations)
    p.regime = int(regime)
    p.fluvial_dinf = int(fluvial_dinf)
    p._channel_ref = ch
    return p


def smooth_area(area, mask, kernel_size, bc="oooo"):
    """orc_smooth_area: moving-average drainage area (float64)."""
    a = np.ascontiguousarray(area, np.float64)
    nr, nc = a.shape
    out = np.empty_like(a)
    st = lib().orc_smooth_area(nr, nc, a.ravel(), np.ascontiguousarray(mask, np.uint8).ravel(), int(kernel_size),
                               bc.encode(), out.ravel())
    if st != 0:
        raise OracleError("drainage area moving-average kernel size must be a positive odd number")
    return out


def lem_step(z, mask, dx, params, boundary=None, K_field=None, erosion_rate=None):
    """orc_lem_step on a copy of z; returns (z_new, stats dict, status).
    erosion_rate: optional float32 array of z's shape that receives the
    per-pixel erosion rate field."""
    zz = np.ascontiguousarray(z, dtype=np.float64).copy()
    nr, nc = zz.shape
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    bf = boundary_flats(m, params.bc.decode()) if boundary is None else np.ascontiguousarray(boundary, np.int64)
    st = LemStats()
    kptr = None
    kf = None
    if K_field is not None:
        kf = np.ascontiguousarray(K_field, np.float64).ravel()
        kptr = kf.ctypes.data_as(C.c_void_p)
    er = None
    if erosion_rate is not None:
        if erosion_rate.dtype != np.float32 or erosion_rate.shape != zz.shape or not erosion_rate.flags.c_contiguous:
            raise ValueError("erosion_rate must be a C-contiguous float32 array of the DEM's shape")
        er = erosion_rate.ctypes.data_as(C.c_void_p)
    rc = lib().orc_lem_step_ex(nr, nc, float(dx), m.ravel(), C.byref(params), bf, len(bf), kptr,
                               zz.ravel(), C.byref(st), er)
    return zz, st.as_dict(), rc


def dinf_graph(z, mask, dx, bc="oooo", nodata=-9999.0, interior_nodata_is_outlet=True, libm_atan2=False):
    """orc_dinf_graph on a float32 surface: dict of rec (n, 2) int64 (-1 =
    none), w (n, 2) float32, nrec, slope, bearing (float32 rasters)."""
    zf = np.ascontiguousarray(z, np.float32)
    nr, nc = zf.shape
    n = nr * nc
    rec = np.empty((n, 2), np.int64)
    w = np.empty((n, 2), np.float32)
    nrec = np.empty(n, np.uint8)
    slope = np.empty((nr, nc), np.float32)
    bearing = np.empty((nr, nc), np.float32)
    st = lib().orc_dinf_graph(nr, nc, float(dx), zf.ravel(), np.ascontiguousarray(mask, np.uint8).ravel(),
                              float(nodata), bc.lower().encode(), int(interior_nodata_is_outlet), int(libm_atan2),
                              rec.ravel(), w.ravel(), nrec, slope.ravel(), bearing.ravel())
    if st != 0:
        raise OracleError(f"dinf graph failed ({st})")
    return dict(rec=rec, w=w, nrec=nrec, slope=slope, bearing=bearing)


def dinf_accumulate(g, valid, dx):
    """Kahn order then accumulate (dx^2 per valid cell); (area, topo)"""
    v = np.ascontiguousarray(valid, np.uint8).ravel()
    n = v.size
    topo = np.empty(n, np.int64)
    nt = lib().orc_dinf_topo(n, v, g["rec"].ravel(), g["w"].ravel(), g["nrec"], topo)
    if nt < 0:
        raise OracleError("bad D-inf graph")
    acc = np.empty(n, np.float64)
    lib().orc_dinf_accumulate(n, g["rec"].ravel(), g["w"].ravel(), g["nrec"], topo, nt, float(dx), acc)
    return acc, topo[:nt]


def channel_threshold_quantity(area, dx, divide_area_by_dx=True):
    return lib().orc_channel_threshold_quantity(float(area), float(dx), int(divide_area_by_dx))


def dinf_channel_mask(z, mask, dx, threshold, bc="oooo", divide_area_by_dx=True, libm_atan2=False):
    """orc_dinf_channel_mask: (is_channel uint8 raster, area float64 raster)"""
    zf = np.ascontiguousarray(z, np.float32)
    nr, nc = zf.shape
    out = np.empty((nr, nc), np.uint8)
    area = np.empty((nr, nc), np.float64)
    st = lib().orc_dinf_channel_mask(nr, nc, float(dx), zf.ravel(), np.ascontiguousarray(mask, np.uint8).ravel(),
                                     bc.lower().encode(), float(threshold), int(divide_area_by_dx), int(libm_atan2),
                                     out.ravel(), area.ctypes.data_as(C.c_void_p))
    if st != 0:
        raise OracleError(f"dinf channel mask failed ({st})")
    return out, area


def median_interior_gradient(z, mask, dx):
    zz = np.ascontiguousarray(z, np.float64)
    nr, nc = zz.shape
    return lib().orc_median_interior_gradient(nr, nc, float(dx), zz.ravel(),
                                              np.ascontiguousarray(mask, np.uint8).ravel())


def diagnose_timestep(z, mask, dx, U, dt, target=0.7):
    zz = np.ascontiguousarray(z, np.float64)
    nr, nc = zz.shape
    d = TsDiagnosis()
    lib().orc_diagnose_timestep(nr, nc, float(dx), zz.ravel(), np.ascontiguousarray(mask, np.uint8).ravel(),
                                float(U), float(dt), float(target), C.byref(d))
    return {k: getattr(d, k) for k, _ in TsDiagnosis._fields_}


class TimestepController:
    """the oracle controller; update(max_dz, z, mask, dx) -> next dt"""

    def __init__(self, dt_initial, **opts):
        self.c = TsController()
        self.o = ts_options(**opts)
        lib().orc_ts_controller_init(C.byref(self.c), float(dt_initial), C.byref(self.o))

    def update(self, max_dz, z, mask, dx):
        slope = median_interior_gradient(z, mask, dx)
        return lib().orc_ts_controller_update(C.byref(self.c), float(max_dz), slope, float(dx))


class ConvergenceMonitor:
    """the oracle monitor; update(max_dz, dt_used) -> 0 running, 1 converged, 2 stalled"""

    def __init__(self, dt_initial, **opts):
        self.m = TsMonitor()
        self.o = ts_monitor_options(**opts)
        lib().orc_ts_monitor_init(C.byref(self.m), float(dt_initial), C.byref(self.o))

    def __del__(self):
        if getattr(self, "m", None) is not None and self.m.hist:
            lib().orc_ts_monitor_free(C.byref(self.m))

    def update(self, max_dz, dt_used):
        return lib().orc_ts_monitor_update(C.byref(self.m), float(max_dz), float(dt_used))


def steady_state_d8(z, mask, dx, bc="oooo", U=1e-4, K=1e-4, m=0.45, n=1.0, relief_mode=False,
                    target_relief=1000.0, baselevel=0.0, divide_area_by_dx=True, channel_threshold_area=0.0,
                    K_field=None, libm_pow=False, nodata=-9999.0, interior_nodata_is_outlet=True,
                    allow_pits_as_outlets=False):
    """compute_d8_fluvial_steady_state on the oracle's flow object of z
    (float32 surface): (z_out float32 raster, result dict)"""
    zf = np.ascontiguousarray(z, np.float32)
    nr, nc = zf.shape
    mk = np.ascontiguousarray(mask, np.uint8)
    fl = flow_build(zf.astype(np.float64), mk, dx, bc, interior_nodata_is_outlet, allow_pits_as_outlets,
                    svector=False)
    area = accumulate(mk, fl["receiver"], fl["topo"], dx)
    o = SsOptions(U, K, m, n, int(relief_mode), target_relief, baselevel, int(divide_area_by_dx),
                  channel_threshold_area)
    r = SsResult()
    out = np.empty((nr, nc), np.float32)
    kf = None if K_field is None else np.ascontiguousarray(K_field, np.float32).ravel()
    st = lib().orc_steady_state_d8(nr, nc, float(dx), mk.ravel(), fl["receiver"], fl["dir"], fl["topo"],
                                   len(fl["topo"]), np.ascontiguousarray(area, np.float64).ravel(),
                                   None if kf is None else kf.ctypes.data_as(C.c_void_p), C.byref(o),
                                   int(libm_pow), float(nodata), out.ravel(), C.byref(r))
    if st != 0:
        raise OracleError("steady state: bad erodibility field")
    res = {k: getattr(r, k) for k, _ in SsResult._fields_}
    res["area"] = np.asarray(area).reshape(nr, nc)
    res["flow"] = fl
    return out, res


def steady_state_dinf(z, mask, dx, bc="oooo", U=1e-4, K=1e-4, m=0.45, n=1.0, relief_mode=False,
                      target_relief=1000.0, baselevel=0.0, divide_area_by_dx=True, channel_threshold_area=0.0,
                      K_field=None, libm_pow=False, libm_atan2=False, nodata=-9999.0, interior_nodata_is_outlet=True,
                      area=None):
    """compute_dinf_fluvial_steady_state on the oracle's D-inf graph of z
    (float32): (z_out float32 raster, result dict with graph and area).
    area: accumulated areas to use in place of the oracle's own (the chi
    sweep is then checked alone)"""
    zf = np.ascontiguousarray(z, np.float32)
    nr, nc = zf.shape
    mk = np.ascontiguousarray(mask, np.uint8)
    g = dinf_graph(zf, mk, dx, bc=bc, nodata=nodata, interior_nodata_is_outlet=interior_nodata_is_outlet,
                   libm_atan2=libm_atan2)
    valid = ((mk != 0) & np.isfinite(zf) & (zf != np.float32(nodata))).astype(np.uint8)
    own, topo = dinf_accumulate(g, valid, dx)
    area = own if area is None else np.ascontiguousarray(area, np.float64).ravel()
    o = SsOptions(U, K, m, n, int(relief_mode), target_relief, baselevel, int(divide_area_by_dx),
                  channel_threshold_area)
    r = SsResult()
    out = np.empty((nr, nc), np.float32)
    kf = None if K_field is None else np.ascontiguousarray(K_field, np.float32).ravel()
    st = lib().orc_steady_state_dinf(nr, nc, float(dx), valid.ravel(), g["rec"].ravel(), g["w"].ravel(), g["nrec"],
                                     np.ascontiguousarray(topo, np.int64), len(topo), area,
                                     None if kf is None else kf.ctypes.data_as(C.c_void_p), C.byref(o),
                                     int(libm_pow), float(nodata), out.ravel(), C.byref(r))
    if st != 0:
        raise OracleError("steady state: bad erodibility field")
    res = {k: getattr(r, k) for k, _ in SsResult._fields_}
    res["area"] = area.reshape(nr, nc)
    res["graph"] = g
    res["valid"] = valid
    return out, res
