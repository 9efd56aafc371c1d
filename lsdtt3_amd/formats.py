"""File formats on either side of the LEM path.

* Rasters: ``read_raster`` / ``write_raster`` over the C ABI's GeoTIFF and
  ESRI ASCII reader and writer (``lsdtt_gpu_raster_*``), with the reference
  loader's semantics (src/core/lsdtt_data_loader.cpp:700-765, 1735-1808):
  float32 values, mask 0 where a value equals the no-data value or is not
  finite, north-up square pixels, no-data -9999 when the file names none.
* The per-step elevation statistics CSV that ``lsdtt-lem`` writes
  (``<fname>_model_elevation_stats.csv``, cli/lsdtt-lem.cpp:3505-3531,
  3688-3716, 3966-3977): the same header, the same number formatting
  (``std::to_string`` for the step values, seven significant figures for the
  process parameters) and the channelization index CI = K l^(m+n) /
  (D^n U^(1-n)) (src/cpu/lsdtt_valley_spacing_cpu.cpp:31-43) over the length
  between the fixed edges (cli/lsdtt-lem.cpp:3533-3549).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _native as N


class RasterMeta(C.Structure):
    """lsdtt_gpu_raster_meta (include/lsdtt3_gpu.h)."""
    _fields_ = [("n_rows", C.c_int32), ("n_cols", C.c_int32), ("x_min", C.c_double), ("y_min", C.c_double),
                ("cell_size", C.c_double), ("cell_size_y", C.c_double), ("no_data_value", C.c_double),
                ("epsg", C.c_int32), ("geographic", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _lib():
    L = N.lib()
    if not getattr(L, "_raster_sigs", False):
        N._sig(L, "lsdtt_gpu_raster_info", C.c_int, [C.c_char_p, C.POINTER(RasterMeta)])
        N._sig(L, "lsdtt_gpu_raster_read", C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.POINTER(RasterMeta)])
        N._sig(L, "lsdtt_gpu_raster_write", C.c_int,
               [C.c_char_p, C.c_void_p, C.c_void_p, C.POINTER(RasterMeta), C.c_int32])
        L._raster_sigs = True
    return L


def raster_info(path) -> dict:
    """The raster's header (shape, georeference, no-data, EPSG code)."""
    m = RasterMeta()
    N.check(_lib().lsdtt_gpu_raster_info(str(path).encode(), C.byref(m)))
    return m.as_dict()


def read_raster(path):
    """(z float32[n_rows, n_cols], mask uint8, meta dict), north row first."""
    L = _lib()
    m = RasterMeta()
    N.check(L.lsdtt_gpu_raster_info(str(path).encode(), C.byref(m)))
    z = np.empty((m.n_rows, m.n_cols), np.float32)
    mask = np.empty((m.n_rows, m.n_cols), np.uint8)
    N.check(L.lsdtt_gpu_raster_read(str(path).encode(), N.ptr(z), N.ptr(mask), C.byref(m)))
    return z, mask, m.as_dict()


def write_raster(path, z, meta, mask=None, compress=True):
    """Write a GeoTIFF (.tif/.tiff; DEFLATE with the floating-point
    predictor when ``compress``) or an ESRI ASCII grid (.asc).  ``meta``
    needs cell_size and may give x_min, y_min, cell_size_y, no_data_value,
    epsg and geographic; cells with mask 0 are written as the no-data value."""
    z = np.ascontiguousarray(z, dtype=np.float32)
    if z.ndim != 2:
        raise ValueError("z must be 2-D")
    m = RasterMeta()
    m.n_rows, m.n_cols = z.shape
    m.x_min = float(meta.get("x_min", 0.0))
    m.y_min = float(meta.get("y_min", 0.0))
    m.cell_size = float(meta["cell_size"])
    m.cell_size_y = float(meta.get("cell_size_y", 0.0))
    m.no_data_value = float(meta.get("no_data_value", -9999.0))
    m.epsg = int(meta.get("epsg", 0))
    m.geographic = int(meta.get("geographic", 0))
    mk = None
    if mask is not None:
        mk = np.ascontiguousarray(mask, dtype=np.uint8)
        if mk.shape != z.shape:
            raise ValueError("mask must have z's shape")
    N.check(_lib().lsdtt_gpu_raster_write(str(path).encode(), N.ptr(z), N.ptr(mk) if mk is not None else None,
                                          C.byref(m), int(bool(compress))))


def channelization_index(K, D, U, length_m, m, n) -> float:
    """CI = K l^(m+n) / (D^n U^(1-n)); 0 when K, D or l is not positive
    (src/cpu/lsdtt_valley_spacing_cpu.cpp:31-43)."""
    if not (K > 0.0) or not (D > 0.0) or not (length_m > 0.0):
        return 0.0
    num = K * math.pow(length_m, m + n)
    den = math.pow(D, n) * (1.0 if abs(n - 1.0) < 1e-12 else math.pow(U, 1.0 - n))
    return num / den if den > 0.0 else 0.0


def channelization_length(n_rows, n_cols, cell_size, boundary_conditions) -> float:
    """The extent between the fixed ('o'/'b') edges: N-S outlets give the row
    extent, E-W outlets the column extent, otherwise the longer side
    (cli/lsdtt-lem.cpp:3533-3549)."""
    bc = boundary_conditions

    def fixed(i):
        return len(bc) > i and bc[i] in "ob"

    ns = float(n_rows) * cell_size
    ew = float(n_cols) * cell_size
    if fixed(0) and fixed(2) and not (fixed(1) and fixed(3)):
        return ns
    if fixed(1) and fixed(3) and not (fixed(0) and fixed(2)):
        return ew
    return max(ns, ew)


def _to_string(v) -> str:
    """std::to_string of a float or double: printf "%f"."""
    return "%f" % float(v)


def _sci(v) -> str:
    """The CLI's sci(): seven significant figures in the default float
    format (std::setprecision(7))."""
    s = "%.7g" % float(v)
    return s


class ElevationStatsCSV:
    """The per-step statistics table of ``lsdtt-lem``.

    ``add(model, stats)`` after each step appends a row from the step's
    statistics and the model's options; the elapsed time accumulates in
    float32 like the CLI's ``float total_time`` (cli/lsdtt-lem.cpp:3567,
    3642), and every float option is formatted from its float32 value."""

    HEADER = ["iteration", "time_yr", "delta_max_z", "delta_mean_z", "max_local_delta_z", "mean_erosion_rate",
              "max_erosion_rate", "uplift_rate", "baselevel_fall_rate", "erodibility", "hillslope_diffusivity",
              "channelization_index"]

    def __init__(self, n_rows, n_cols, cell_size, boundary_conditions):
        self.length_m = channelization_length(n_rows, n_cols, cell_size, boundary_conditions)
        self.rows = []
        self.iteration = 0
        self.total_time = np.float32(0.0)

    @classmethod
    def for_model(cls, model):
        bc = model.options.boundary_conditions
        bc = bc.decode() if isinstance(bc, bytes) else str(bc)
        return cls(model.n_rows, model.n_cols, model.cell_size, bc)

    def add_row(self, stats, dt, uplift_rate, baselevel_fall_rate, erodibility, hillslope_diffusivity, m, n):
        f32 = np.float32
        self.total_time = f32(self.total_time + f32(dt))
        self.iteration += 1
        K, D, U = float(f32(erodibility)), float(f32(hillslope_diffusivity)), float(f32(uplift_rate))
        ci = channelization_index(K, D, U, self.length_m, float(f32(m)), float(f32(n)))
        self.rows.append([str(self.iteration), _to_string(self.total_time),
                          _to_string(f32(stats["delta_max_elevation"])),
                          _to_string(f32(stats["delta_mean_elevation"])),
                          _to_string(f32(stats["max_local_elevation_change"])),
                          _to_string(f32(stats["mean_erosion_rate"])),
                          _to_string(f32(stats["max_erosion_rate"])),
                          _to_string(f32(uplift_rate)), _to_string(f32(baselevel_fall_rate)),
                          _sci(K), _sci(D), _sci(ci)])

    def add(self, model, stats):
        o = model.options
        self.add_row(stats, o.timestep, o.uplift_rate, o.baselevel_fall_rate, o.erodibility,
                     o.hillslope_diffusivity, o.m_exponent, o.n_exponent)

    def text(self) -> str:
        def field(f):
            if any(ch in f for ch in ',"\n'):
                return '"' + f.replace('"', '""') + '"'
            return f
        lines = [",".join(field(f) for f in self.HEADER)]
        lines += [",".join(field(f) for f in r) for r in self.rows]
        return "\n".join(lines) + "\n"

    def write(self, path):
        with open(path, "w") as fh:
            fh.write(self.text())
