"""MI355X-native NLOS time-of-flight backprojection (drop-in for mjgalindo/backprojector's
GPU path).  Compute lives in libbpx.so (hand-written gfx950 HIP kernels behind the C ABI of
include/bpx.h); this package is its Python front.
"""
from . import nlosbpy
from ._native import BpxError, InvalidArgument, device_count
from .core import (Batch, Net, Plan, as_pairs, backprojection, octree_backprojection, phasor_pulse,
                   render_transient, row_windows, window_backprojection, PAIR_DTYPE)

__all__ = ["nlosbpy", "Batch", "Net", "Plan", "backprojection", "window_backprojection", "octree_backprojection",
           "phasor_pulse", "render_transient", "row_windows", "as_pairs", "PAIR_DTYPE", "BpxError",
           "InvalidArgument", "device_count"]
