"""lsdtt3_amd -- MI355X-native landscape-evolution timestep for lsdtt3.

Hand-written HIP kernels for gfx950 behind a C ABI (include/lsdtt3_gpu.h,
lib/liblsdtt3_gpu.so); this package is a thin Python front end over that ABI
that mirrors the reference's Python names (FlowInfo, LandscapeModel).
"""
from .flow import FlowGraph, FlowInfo, fill_pits, gpu_available  # noqa: F401
from .lem import (ConvergenceMonitor, Ensemble, LandscapeModel, TimestepController,  # noqa: F401
                  build_boundary_flats, make_options)

__all__ = ["FlowGraph", "FlowInfo", "fill_pits", "gpu_available", "LandscapeModel", "build_boundary_flats"]
