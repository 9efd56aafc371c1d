"""FlowInfo and fill_pits on the GPU, mirroring the reference's Python
``lsdtt3.FlowInfo`` (python_bindings/bindings/flow_info.cpp:315) and the C++
``lsdtt_flowinfo`` query surface (src/core/lsdtt_flow_base.hpp:126-491).

Arrays are numpy; every computation runs in liblsdtt3_gpu.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class FlowInfo:
    """D8 flow object built on the GPU.

    Parameters mirror ``lsdtt_flowinfo_options`` (src/core/lsdtt_flow_base.hpp:
    79-111): boundary_conditions (N,E,S,W of o/n/p/b), interior_nodata_is_outlet,
    allow_pits_as_outlets, build_topological_order, build_s_vector and
    forced_outlet_flats.  The DEM may be float32 (the reference raster type) or
    float64 (the LEM state); slopes are taken in double either way.
    """

    def __init__(self, dem, mask, cell_size, boundary_conditions="oooo",
                 interior_nodata_is_outlet=True, allow_pits_as_outlets=False,
                 build_topological_order=True, build_s_vector=True,
                 forced_outlet_flats=None):
        dem = np.asarray(dem)
        if dem.ndim != 2:
            raise ValueError("dem must be 2-D")
        self.n_rows, self.n_cols = dem.shape
        self.n = self.n_rows * self.n_cols
        self.cell_size = float(cell_size)
        L = N.lib()
        o = N.FlowOptions()
        L.lsdtt_gpu_flow_options_init(C.byref(o), self.n_rows, self.n_cols, self.cell_size)
        o.boundary_conditions = boundary_conditions.encode()[:4]
        o.interior_nodata_is_outlet = int(interior_nodata_is_outlet)
        o.allow_pits_as_outlets = int(allow_pits_as_outlets)
        o.build_topological_order = int(build_topological_order)
        o.build_s_vector = int(build_s_vector)
        self._has_topo = bool(build_topological_order)
        self._has_sv = bool(build_s_vector)
        h = L.lsdtt_gpu_flow_create(C.byref(o))
        if not h:
            raise N.GpuError(N.ERR_ARG, (L.lsdtt_gpu_last_error() or b"").decode())
        self._h = h
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        forced = None if forced_outlet_flats is None else np.ascontiguousarray(forced_outlet_flats, np.int64)
        nf = 0 if forced is None else forced.size
        if dem.dtype == np.float64:
            z = np.ascontiguousarray(dem)
            N.check(L.lsdtt_gpu_flow_build_f64(h, N.ptr(z), N.ptr(m), N.ptr(forced), nf))
        else:
            z = np.ascontiguousarray(dem, dtype=np.float32)
            N.check(L.lsdtt_gpu_flow_build_f32(h, N.ptr(z), N.ptr(m), N.ptr(forced), nf))
        nv, no, md = C.c_int64(), C.c_int64(), C.c_int64()
        N.check(L.lsdtt_gpu_flow_sizes(h, C.byref(nv), C.byref(no), C.byref(md)))
        self.n_valid, self.n_outlets, self.max_depth = nv.value, no.value, md.value
        self._cache = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().lsdtt_gpu_flow_destroy(h)
            self._h = None

    def _get(self):
        if "receiver" not in self._cache:
            n = self.n
            rcv = np.empty(n, np.int64)
            d = np.empty(n, np.uint8)
            dn = np.empty(n, np.uint8)
            outl = np.empty(max(self.n_outlets, 1), np.int64)
            topo = np.empty(max(self.n_valid, 1), np.int64) if self._has_topo else None
            sv = np.empty(max(self.n_valid, 1), np.int64) if self._has_sv else None
            ss = np.empty(n, np.int64) if self._has_sv else None
            sz = np.empty(n, np.uint64) if self._has_sv else None
            N.check(N.lib().lsdtt_gpu_flow_get(self._h, N.ptr(rcv), N.ptr(d), N.ptr(dn), N.ptr(outl),
                                               N.ptr(topo), N.ptr(sv), N.ptr(ss), N.ptr(sz)))
            self._cache.update(receiver=rcv, dir=d, donors=dn, outlets=outl[:self.n_outlets],
                               topo=None if topo is None else topo[:self.n_valid],
                               s_vector=None if sv is None else sv[:self.n_valid],
                               s_start=ss, s_size=sz)
        return self._cache

    # ---- the lsdtt_flow_base array queries ---------------------------------
    @property
    def receivers(self):
        return self._get()["receiver"]

    @property
    def receiver_directions(self):
        return self._get()["dir"]

    @property
    def donor_counts(self):
        return self._get()["donors"]

    def outlets(self):
        return self._get()["outlets"]

    def topological_order(self):
        if not self._has_topo:
            raise RuntimeError("topological_order not built - set build_topological_order=True")
        return self._get()["topo"]

    @property
    def s_vector(self):
        if not self._has_sv:
            raise RuntimeError("s_vector not built - set build_s_vector=True")
        return self._get()["s_vector"]

    @property
    def s_vector_start(self):
        return self._get()["s_start"]

    @property
    def s_vector_size(self):
        return self._get()["s_size"]

    def upslope_nodes(self, flat):
        c = self._get()
        st, sz = int(c["s_start"][flat]), int(c["s_size"][flat])
        return c["s_vector"][st:st + sz]

    def get_drainage_area_raster(self, no_data_value=-9999.0):
        """(area float32[n_rows, n_cols], mask uint8) with area =
        float(n_contributing * cell_size^2)."""
        a = np.empty(self.n, np.float32)
        m = np.empty(self.n, np.uint8)
        N.check(N.lib().lsdtt_gpu_flow_area(self._h, float(no_data_value), N.ptr(a), N.ptr(m), None))
        return a.reshape(self.n_rows, self.n_cols), m.reshape(self.n_rows, self.n_cols)

    def n_contributing_pixels(self):
        c = np.empty(self.n, np.int64)
        N.check(N.lib().lsdtt_gpu_flow_area(self._h, 0.0, None, None, N.ptr(c)))
        return c

    def accumulate(self, cell_weights=None):
        """double accumulation in topological order (cell_size^2 by default)."""
        out = np.empty(self.n, np.float64)
        w = None if cell_weights is None else np.ascontiguousarray(cell_weights, np.float64).ravel()
        if w is not None and w.size != self.n:
            raise ValueError("cell_weights size must equal n_rows * n_cols")
        N.check(N.lib().lsdtt_gpu_flow_accumulate(self._h, N.ptr(w), N.ptr(out)))
        return out


    def fluvial_steady_state(self, U=1e-4, K=1e-4, m=0.45, n=1.0, relief_mode=False, target_relief=1000.0,
                             baselevel_elevation=0.0, divide_area_by_dx=True, channel_threshold_area=0.0,
                             erodibility_field=None, no_data_value=-9999.0):
        """compute_d8_fluvial_steady_state(flow, opts)
        (src/cpu/lsdtt_steady_state_cpu.cpp:173-225): (float32 elevation
        raster, result dict with K_used, max_unit_height, computed_relief,
        erodibility_field_scale, chi_c fields)"""
        o = _steady_options(U, K, m, n, relief_mode, target_relief, baselevel_elevation, divide_area_by_dx,
                            channel_threshold_area)
        kf = _erodibility(erodibility_field, self.n)
        out = np.empty((self.n_rows, self.n_cols), np.float32)
        r = N.SteadyResult()
        N.check(N.lib().lsdtt_gpu_flow_steady_state(self._h, C.byref(o), N.ptr(kf), float(no_data_value),
                                                    N.ptr(out), C.byref(r)))
        return out, _steady_result(r)

    def get_chi_raster(self, m_over_n=0.45, A_0=1.0, no_data_value=-9999.0):
        """lsdtt::get_chi_raster(flow, {m_over_n, A_0})
        (src/cpu/lsdtt_chi_extraction_cpu.cpp:739-765): (chi float32
        [n_rows, n_cols], mask uint8), every outlet at chi = 0."""
        chi = np.empty(self.n, np.float32)
        m = np.empty(self.n, np.uint8)
        N.check(N.lib().lsdtt_gpu_flow_chi(self._h, float(m_over_n), float(A_0), float(no_data_value),
                                           N.ptr(chi), N.ptr(m)))
        return chi.reshape(self.n_rows, self.n_cols), m.reshape(self.n_rows, self.n_cols)

    def get_chi_from_all_outlets_threshold(self, minimum_contributing_pixels, m_over_n=0.45, A_0=1.0):
        """get_all_basin_outlet_flats(flow, min) followed by
        get_chi_from_outlets_zero_base_threshold (lsdtt_chi_extraction_cpu.cpp:
        368-497): list of basins, each a dict of outlet_flat, flats, chi and
        receiver_flats (the outlet's receiver is -2)."""
        npx, nb = C.c_int64(0), C.c_int64(0)
        N.check(N.lib().lsdtt_gpu_flow_chi_threshold(self._h, float(m_over_n), float(A_0),
                                                     int(minimum_contributing_pixels), C.byref(npx), C.byref(nb),
                                                     None, None, None, None))
        k, b = npx.value, nb.value
        flats = np.empty(max(k, 1), np.int64)
        chi = np.empty(max(k, 1), np.float32)
        rec = np.empty(max(k, 1), np.int64)
        bs = np.empty(max(b, 1), np.int64)
        N.check(N.lib().lsdtt_gpu_flow_chi_threshold(self._h, float(m_over_n), float(A_0),
                                                     int(minimum_contributing_pixels), C.byref(npx), C.byref(nb),
                                                     N.ptr(flats), N.ptr(chi), N.ptr(rec), N.ptr(bs)))
        bounds = list(bs[:b]) + [k]
        return [dict(outlet_flat=int(flats[bounds[i]]), flats=flats[bounds[i]:bounds[i + 1]].copy(),
                     chi=chi[bounds[i]:bounds[i + 1]].copy(), receiver_flats=rec[bounds[i]:bounds[i + 1]].copy())
                for i in range(b)]


def _steady_options(U, K, m, n, relief_mode, target_relief, baselevel_elevation, divide_area_by_dx,
                    channel_threshold_area):
    o = N.SteadyOptions()
    N.lib().lsdtt_gpu_steady_state_options_init(C.byref(o))
    o.U, o.K, o.m, o.n = float(U), float(K), float(m), float(n)
    o.relief_mode = int(relief_mode)
    o.target_relief = float(target_relief)
    o.baselevel_elevation = float(baselevel_elevation)
    o.divide_area_by_dx = int(divide_area_by_dx)
    o.channel_threshold_area = float(channel_threshold_area)
    return o


def _erodibility(field, n):
    if field is None:
        return None
    kf = np.ascontiguousarray(field, np.float32)
    if kf.size != n:
        raise ValueError("erodibility_field size must equal n_rows * n_cols")
    return kf


def _steady_result(r):
    res = {k: getattr(r, k) for k, _ in N.SteadyResult._fields_}
    res["chi_c_reported"] = bool(res["chi_c_reported"])
    return res


class FlowGraph:
    """D-infinity flow graph on the GPU, mirroring the reference's Python
    ``FlowGraph(dem, method="dinf")`` (python_bindings/bindings/flow_graph.cpp:
    90-101) over ``lsdtt_flowgraph_cpu`` (src/cpu/lsdtt_flowgraph_cpu.cpp).

    ``receivers`` (n, 2) int64 flat indices (-1 = none) and ``weights``
    (n, 2) float32, in the reference's CSR slot order; ``slope`` and
    ``bearing`` rasters; ``accumulate()`` the drainage area (cell_size^2 per
    valid cell, float64).  Only method "dinf" is built here.
    """

    def __init__(self, dem, mask, cell_size, method="dinf", boundary_conditions="oooo",
                 interior_nodata_is_outlet=True, no_data_value=-9999.0):
        if method not in ("dinf", "d_infinity"):
            raise ValueError("only the D-infinity method is built on the GPU")
        z = np.ascontiguousarray(dem, dtype=np.float32)
        if z.ndim != 2:
            raise ValueError("dem must be 2-D")
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.n_rows, self.n_cols = z.shape
        n = z.size
        self.cell_size = float(cell_size)
        self.receivers = np.empty((n, 2), np.int64)
        self.weights = np.empty((n, 2), np.float32)
        self.slope = np.empty(z.shape, np.float32)
        self.bearing = np.empty(z.shape, np.float32)
        self._area = np.empty(z.shape, np.float64)
        self._z, self._m = z, m
        self._bc = boundary_conditions.encode()[:4]
        self._ino = int(interior_nodata_is_outlet)
        self.no_data_value = float(no_data_value)
        N.check(N.lib().lsdtt_gpu_dinf_graph(N.ptr(z), N.ptr(m), self.n_rows, self.n_cols, self.cell_size,
                                             self._bc, self._ino, self.no_data_value, N.ptr(self.receivers),
                                             N.ptr(self.weights), N.ptr(self.slope), N.ptr(self.bearing),
                                             N.ptr(self._area)))

    def accumulate(self):
        """drainage area (m^2) in float64, 0 off the valid cells"""
        return self._area.copy()

    def fluvial_steady_state(self, U=1e-4, K=1e-4, m=0.45, n=1.0, relief_mode=False, target_relief=1000.0,
                             baselevel_elevation=0.0, divide_area_by_dx=True, channel_threshold_area=0.0,
                             erodibility_field=None):
        """compute_dinf_fluvial_steady_state(flow, opts)
        (src/cpu/lsdtt_steady_state_cpu.cpp:227-436) on this graph's
        surface: (float32 elevation raster, result dict as
        FlowInfo.fluvial_steady_state's; z is no_data_value off the valid
        cells)"""
        o = _steady_options(U, K, m, n, relief_mode, target_relief, baselevel_elevation, divide_area_by_dx,
                            channel_threshold_area)
        kf = _erodibility(erodibility_field, self.n_rows * self.n_cols)
        out = np.empty((self.n_rows, self.n_cols), np.float32)
        area = np.empty((self.n_rows, self.n_cols), np.float64)
        r = N.SteadyResult()
        N.check(N.lib().lsdtt_gpu_dinf_steady_state(N.ptr(self._z), N.ptr(self._m), self.n_rows, self.n_cols,
                                                    self.cell_size, self._bc, self._ino, self.no_data_value,
                                                    C.byref(o), N.ptr(kf), N.ptr(out), N.ptr(area), C.byref(r)))
        res = _steady_result(r)
        res["area"] = area
        return out, res

    def receivers_of(self, row, col):
        f = int(row) * self.n_cols + int(col)
        k = int((self.receivers[f] >= 0).sum())
        rec = self.receivers[f, :k]
        return dict(pixels=np.stack([rec // self.n_cols, rec % self.n_cols], axis=1),
                    weights=self.weights[f, :k].copy())


def fill_pits(dem, mask, no_data_value=-9999.0, boundary_conditions="oooo"):
    """Epsilon-free fill (lsdtt::fill_pits with min_slope <= 0), float32."""
    z = np.ascontiguousarray(dem, dtype=np.float32)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty_like(z)
    nr, nc = z.shape
    N.check(N.lib().lsdtt_gpu_fill(N.ptr(z), N.ptr(m), nr, nc, float(no_data_value),
                                   boundary_conditions.encode()[:4], N.ptr(out)))
    return out


def gpu_available() -> bool:
    return bool(N.lib().lsdtt_gpu_available())
