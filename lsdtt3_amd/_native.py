"""ctypes binding of liblsdtt3_gpu.so (declared in include/lsdtt3_gpu.h).

The library is built in-tree by ``python -m lsdtt3_amd.build`` (or
``__graft_entry__.build()``).  Loading fails loudly: there is no CPU fallback
inside this package -- a caller that wants one uses its own CPU path when a
call raises, as the reference's GPU dispatch rule prescribes
(docs/developer_notes/gpu_compute.md:51-62).
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liblsdtt3_gpu.so")

OK = 0
ERR_ARG = 1
ERR_RUNTIME = 2
ERR_UNRESOLVED = 3
ERR_NO_DEVICE = 4
ERR_NOT_CONVERGED = 5

_lib = None


class GpuError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class FlowOptions(C.Structure):
    _fields_ = [("n_rows", C.c_int32), ("n_cols", C.c_int32), ("cell_size", C.c_double),
                ("boundary_conditions", C.c_char * 4),
                ("interior_nodata_is_outlet", C.c_int32), ("allow_pits_as_outlets", C.c_int32),
                ("build_topological_order", C.c_int32), ("build_s_vector", C.c_int32)]


def _sig(L, name, res, args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = args


class SteadyOptions(C.Structure):
    _fields_ = [("U", C.c_double), ("K", C.c_double), ("m", C.c_double), ("n", C.c_double),
                ("relief_mode", C.c_int32), ("target_relief", C.c_double), ("baselevel_elevation", C.c_double),
                ("divide_area_by_dx", C.c_int32), ("channel_threshold_area", C.c_double)]


class SteadyResult(C.Structure):
    _fields_ = [("K_used", C.c_double), ("max_unit_height", C.c_double), ("computed_relief", C.c_double),
                ("erodibility_field_scale", C.c_double), ("max_unit_height_at_channel_head", C.c_double),
                ("n_channel_pixels", C.c_int64), ("chi_c_reported", C.c_int32)]


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m lsdtt3_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    _sig(L, "lsdtt_gpu_available", C.c_int, [])
    _sig(L, "lsdtt_gpu_last_error", C.c_char_p, [])
    _sig(L, "lsdtt_gpu_device_alloc", vp, [i64])
    _sig(L, "lsdtt_gpu_device_free", None, [vp])
    _sig(L, "lsdtt_gpu_memcpy", C.c_int, [vp, vp, i64, i32])
    _sig(L, "lsdtt_gpu_synchronize", C.c_int, [])
    _sig(L, "lsdtt_gpu_flow_options_init", None, [C.POINTER(FlowOptions), i32, i32, dbl])
    _sig(L, "lsdtt_gpu_flow_create", vp, [C.POINTER(FlowOptions)])
    _sig(L, "lsdtt_gpu_flow_destroy", None, [vp])
    _sig(L, "lsdtt_gpu_flow_build_f32", C.c_int, [vp, vp, vp, vp, i64])
    _sig(L, "lsdtt_gpu_flow_build_f64", C.c_int, [vp, vp, vp, vp, i64])
    _sig(L, "lsdtt_gpu_flow_build_device", C.c_int, [vp, vp, i32, vp, vp])
    _sig(L, "lsdtt_gpu_flow_sizes", C.c_int, [vp, vp, vp, vp])
    _sig(L, "lsdtt_gpu_flow_get", C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp])
    _sig(L, "lsdtt_gpu_flow_area", C.c_int, [vp, dbl, vp, vp, vp])
    _sig(L, "lsdtt_gpu_flow_accumulate", C.c_int, [vp, vp, vp])
    _sig(L, "lsdtt_gpu_steady_state_options_init", None, [C.POINTER(SteadyOptions)])
    _sig(L, "lsdtt_gpu_flow_steady_state", C.c_int, [vp, C.POINTER(SteadyOptions), vp, C.c_double, vp,
                                                     C.POINTER(SteadyResult)])
    _sig(L, "lsdtt_gpu_flow_chi", C.c_int, [vp, dbl, dbl, dbl, vp, vp])
    _sig(L, "lsdtt_gpu_flow_chi_device", C.c_int, [vp, dbl, dbl, dbl, vp, vp])
    _sig(L, "lsdtt_gpu_flow_chi_threshold", C.c_int, [vp, dbl, dbl, i64, vp, vp, vp, vp, vp, vp])
    _sig(L, "lsdtt_gpu_fill", C.c_int, [vp, vp, i32, i32, dbl, C.c_char_p, vp])
    _sig(L, "lsdtt_gpu_fill_device", C.c_int, [vp, vp, i32, i32, dbl, C.c_char_p, vp, vp])
    _sig(L, "lsdtt_gpu_fill_device_noalloc", C.c_int, [vp, vp, i32, i32, dbl, C.c_char_p, vp, vp])
    _sig(L, "lsdtt_gpu_fill_noalloc_max_cells", i64, [])
    _extend(L)
    _lib = L
    return L


def _extend(L):
    """Signatures of the LEM entry points (kept separate for readability)."""
    from . import _lem_sig
    _lem_sig.declare(L, _sig)


def check(status: int) -> None:
    if status != OK:
        msg = lib().lsdtt_gpu_last_error() or b""
        raise GpuError(status, msg.decode(errors="replace"))


def ptr(a) -> C.c_void_p:
    """Address of a numpy array (None -> NULL)."""
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


class DeviceArray:
    """Device memory of the library (for the *_device entry points): a
    numpy-shaped buffer with explicit uploads and downloads."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = lib().lsdtt_gpu_device_alloc(max(1, self.nbytes))
        if not p:
            raise GpuError(ERR_RUNTIME, (lib().lsdtt_gpu_last_error() or b"").decode())
        self.ptr = C.c_void_p(p)

    @classmethod
    def from_host(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(lib().lsdtt_gpu_memcpy(d.ptr, C.c_void_p(a.ctypes.data), d.nbytes, 1))
        return d

    def to_host(self):
        import numpy as np
        out = np.empty(self.shape, self.dtype)
        check(lib().lsdtt_gpu_memcpy(C.c_void_p(out.ctypes.data), self.ptr, self.nbytes, 2))
        return out

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p is not None and _lib is not None:
            _lib.lsdtt_gpu_device_free(p)
            self.ptr = None


def synchronize() -> None:
    check(lib().lsdtt_gpu_synchronize())
