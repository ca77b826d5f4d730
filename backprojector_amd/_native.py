"""ctypes binding of libbpx.so (the C ABI in include/bpx.h).

The library is loaded from this package directory (built in-tree by
backprojector_amd.build).  There is no fallback: if the library is missing or no gfx950
device is visible, calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbpx.so")

PAIR_DTYPE = np.dtype(
    [("camera_point", np.float32, (3,)), ("laser_point", np.float32, (3,)),
     ("camera_wall", np.float32), ("laser_wall", np.float32)]
)
assert PAIR_DTYPE.itemsize == 32

STATUS = {0: "BPX_OK", 1: "BPX_ERR_INVALID_ARGUMENT", 2: "BPX_ERR_OUT_OF_MEMORY",
          3: "BPX_ERR_DEVICE", 4: "BPX_ERR_UNSUPPORTED", 5: "BPX_ERR_NO_DEVICE"}

STORAGE = {"f32": 0, "fp32": 0, "float32": 0, "f16": 1, "fp16": 1, "float16": 1,
           "bf16": 2, "bfloat16": 2}


class BpxError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class InvalidArgument(BpxError, ValueError):
    pass


class Options(ctypes.Structure):
    _fields_ = [("storage", ctypes.c_uint32), ("kahan_block", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("variant", ctypes.c_int32)]


class Timing(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("per_exec_ms", ctypes.c_double),
                ("table_kernel_ms", ctypes.c_double), ("gather_kernel_ms", ctypes.c_double),
                ("iterations", ctypes.c_int32)]


class Pair(ctypes.Structure):
    _fields_ = [("camera_point", ctypes.c_float * 3), ("laser_point", ctypes.c_float * 3),
                ("camera_wall", ctypes.c_float), ("laser_wall", ctypes.c_float)]


class NlosPrepared(ctypes.Structure):
    _fields_ = [("chunk", ctypes.POINTER(ctypes.c_float)), ("rows", ctypes.c_uint64),
                ("Tc", ctypes.c_uint32), ("pairs", ctypes.POINTER(Pair)),
                ("num_pairs", ctypes.c_uint32), ("zero", ctypes.c_float * 3),
                ("inc", ctypes.c_float * 9), ("chunked_t0", ctypes.c_float),
                ("min_index", ctypes.c_float), ("max_index", ctypes.c_float)]


class OctreeOptions(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_uint32), ("threshold", ctypes.c_float),
                ("proposal_threshold", ctypes.c_float), ("flags", ctypes.c_uint32),
                ("level_proposal_threshold", ctypes.c_float * 8), ("keep_bound_ratio", ctypes.c_float),
                ("level_fallback_threshold", ctypes.c_float * 8), ("coarse_pair_stride", ctypes.c_uint32)]


def make_octree_options(levels=4, threshold=0.25, proposal_threshold=0.5, bounds_only=False,
                        keep_bound_ratio=0.0, fallback_threshold=0.0, coarse_pair_stride=1):
    """proposal_threshold / fallback_threshold: one probability, or one per coarse level."""
    o = OctreeOptions(int(levels), float(threshold), 0.5, OCTREE_BOUNDS_ONLY if bounds_only else 0)
    o.keep_bound_ratio = float(keep_bound_ratio)
    o.coarse_pair_stride = int(coarse_pair_stride)
    fb = [float(fallback_threshold)] * int(levels) if np.ndim(fallback_threshold) == 0 else \
        [float(x) for x in fallback_threshold]
    if len(fb) != int(levels):
        raise ValueError(f"expected {levels} level fallback thresholds, got {len(fb)}")
    for i, x in enumerate(fb):
        o.level_fallback_threshold[i] = x
    if np.ndim(proposal_threshold) == 0:
        o.proposal_threshold = float(proposal_threshold)
    else:
        per = [float(x) for x in proposal_threshold]
        if len(per) != int(levels):
            raise ValueError(f"expected {levels} level proposal thresholds, got {len(per)}")
        for i, x in enumerate(per):
            o.level_proposal_threshold[i] = x
    return o


class OctreeStats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_uint32), ("active", ctypes.c_uint32 * 9),
                ("running_max", ctypes.c_double * 8), ("bound_max", ctypes.c_double * 8),
                ("bound_norm", ctypes.c_double * 8), ("threshold", ctypes.c_double * 8),
                ("fine_updates", ctypes.c_uint64), ("bound_updates", ctypes.c_uint64),
                ("fallback_voxels", ctypes.c_uint32), ("coarse_pairs", ctypes.c_uint32)]

    def to_dict(self):
        L = self.levels
        return dict(levels=L, active=list(self.active[:L + 1]),
                    running_max=list(self.running_max[:L]), bound_max=list(self.bound_max[:L]),
                    bound_norm=list(self.bound_norm[:L]), threshold=list(self.threshold[:L]),
                    fine_updates=int(self.fine_updates), bound_updates=int(self.bound_updates),
                    fallback_voxels=int(self.fallback_voxels), coarse_pairs=int(self.coarse_pairs))


class CaptureDesc(ctypes.Structure):
    _fields_ = [("transient_chunk", ctypes.c_void_p), ("transient_size", ctypes.c_uint64),
                ("T", ctypes.c_uint32), ("scanned_pairs", ctypes.c_void_p),
                ("num_pairs", ctypes.c_uint32), ("voxels_per_side", ctypes.c_uint32 * 3),
                ("volume_zero_pos", ctypes.c_float * 3), ("voxel_inc", ctypes.c_float * 9),
                ("t0", ctypes.c_float), ("deltaT", ctypes.c_float)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("lane_voxels", ctypes.c_uint32), ("groups", ctypes.c_uint32),
                ("order", ctypes.c_uint32), ("lds_captures", ctypes.c_uint32), ("lds_depth", ctypes.c_uint32),
                ("launches", ctypes.c_uint32)]

    def to_dict(self):
        return dict(lane_voxels=int(self.lane_voxels), groups=int(self.groups),
                    order=int(self.order), lds_captures=int(self.lds_captures), lds_depth=int(self.lds_depth),
                    launches=int(self.launches))


class BatchTiming(ctypes.Structure):
    _fields_ = [("per_exec_ms", ctypes.c_double), ("updates_per_s", ctypes.c_double),
                ("updates", ctypes.c_uint64), ("iterations", ctypes.c_int32)]


FLAG_REFERENCE_PAIR_TRUNCATION = 0x1
FLAG_CENTRED_WALL_BASIS = 0x2
FLAG_COMPLEX_SAMPLES = 0x4
FLAG_KEEP_F32_ROWS = 0x8
OCTREE_BOUNDS_ONLY = 0x1

_lib = None

# Every symbol include/bpx.h declares (checked by the CPU test suite).
EXPORTED = [
    "bpx_last_error", "bpx_version", "bpx_abi_version", "bpx_device_count", "bpx_row_windows",
    "bpx_backprojection", "bpx_complex_backprojection", "bpx_backprojection_ex", "bpx_backprojection_noalloc",
    "bpx_window_backprojection", "bpx_phasor_pulse", "bpx_nlos_backproject", "bpx_nlos_phasor", "bpx_nlos_phasor_multi",
    "bpx_nlos_phasor_output",
    "bpx_plan_create", "bpx_nlos_plan_create", "bpx_plan_execute", "bpx_plan_synchronize", "bpx_plan_read_volume",
    "bpx_plan_info", "bpx_plan_kernel_name", "bpx_plan_row_storage", "bpx_plan_read_thresholds", "bpx_plan_benchmark", "bpx_plan_tune", "bpx_plan_destroy", "bpx_render_transient",
    "bpx_copy_benchmark", "bpx_octree_backprojection", "bpx_octree_plan_create",
    "bpx_octree_plan_stats", "bpx_nlos_prepare", "bpx_nlos_prepared_free",
    "bpx_net_param_count", "bpx_net_create", "bpx_net_destroy", "bpx_net_get_params",
    "bpx_net_set_params", "bpx_net_forward", "bpx_net_gradients", "bpx_net_train_octree",
    "bpx_octree_plan_attach_net", "bpx_octree_plan_set_options",
    "bpx_batch_create", "bpx_batch_size", "bpx_batch_tune", "bpx_batch_set_config",
    "bpx_batch_get_config", "bpx_batch_execute", "bpx_batch_synchronize", "bpx_batch_read_volume",
    "bpx_batch_latencies", "bpx_batch_benchmark", "bpx_batch_destroy",
]


def lib():
    """Loads libbpx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run backprojector_amd.build.build() first")
    L = ctypes.CDLL(LIB_PATH)
    c_p = ctypes.c_void_p
    u32, u64, i32, f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float
    L.bpx_last_error.restype = ctypes.c_char_p
    L.bpx_version.restype = ctypes.c_char_p
    L.bpx_abi_version.restype = ctypes.c_int
    L.bpx_device_count.restype = ctypes.c_int
    for name in ("bpx_backprojection", "bpx_complex_backprojection"):
        getattr(L, name).argtypes = [c_p, u64, u32, c_p, u32, c_p, c_p, c_p, c_p, f32, f32]
    L.bpx_backprojection_ex.argtypes = [c_p, u64, u32, c_p, u32, c_p, c_p, c_p, c_p, f32, f32,
                                        ctypes.c_int, c_p]
    L.bpx_backprojection_noalloc.argtypes = [c_p, u64, u32, c_p, u32, c_p, c_p, c_p, c_p, f32, f32, c_p]
    L.bpx_window_backprojection.argtypes = [c_p, u64, u32, c_p, u32, c_p, c_p, c_p, c_p, f32,
                                            f32, ctypes.c_int]
    L.bpx_phasor_pulse.argtypes = [c_p, u64, u32, f32, f32, u32, f32, c_p]
    L.bpx_nlos_backproject.argtypes = [c_p, c_p, ctypes.c_int, c_p, c_p, c_p, c_p, c_p, c_p, f32,
                                       f32, ctypes.c_int, c_p, c_p, c_p, c_p, ctypes.c_int,
                                       ctypes.c_int, c_p]
    L.bpx_nlos_phasor.argtypes = [c_p, c_p, ctypes.c_int, c_p, c_p, c_p, c_p, c_p, c_p, f32, f32,
                                  ctypes.c_int, c_p, c_p, c_p, f32, ctypes.c_int, ctypes.c_int,
                                  c_p]
    L.bpx_nlos_phasor_multi.argtypes = [c_p, c_p, ctypes.c_int, c_p, c_p, c_p, c_p, c_p, c_p, f32, f32,
                                        ctypes.c_int, c_p, c_p, c_p, c_p, u32, ctypes.c_int, ctypes.c_int,
                                        c_p]
    L.bpx_nlos_phasor_output.argtypes = [c_p, c_p, ctypes.c_int, c_p, c_p, c_p, c_p, c_p, c_p, f32, f32,
                                         ctypes.c_int, c_p, c_p, c_p, c_p, u32, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, c_p]
    L.bpx_plan_create.argtypes = [c_p, u64, u32, c_p, u32, c_p, c_p, c_p, f32, f32, ctypes.c_int,
                                  c_p, ctypes.POINTER(c_p)]
    L.bpx_nlos_plan_create.argtypes = [c_p, c_p, ctypes.c_int, c_p, c_p, c_p, c_p, c_p, c_p, f32,
                                       f32, ctypes.c_int, c_p, c_p, c_p, c_p, ctypes.c_int, c_p,
                                       ctypes.POINTER(c_p)]
    L.bpx_plan_execute.argtypes = [c_p, c_p]
    L.bpx_plan_synchronize.argtypes = [c_p]
    L.bpx_plan_read_volume.argtypes = [c_p, c_p]
    L.bpx_plan_read_thresholds.argtypes = [c_p, c_p, u64]
    L.bpx_plan_tune.argtypes = [c_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.bpx_row_windows.argtypes = [c_p, ctypes.c_uint32, c_p, c_p, c_p, f32, f32, ctypes.c_uint32, c_p]
    L.bpx_plan_kernel_name.argtypes = [c_p, ctypes.c_char_p, ctypes.c_uint64]
    L.bpx_plan_row_storage.argtypes = [c_p, ctypes.POINTER(ctypes.c_uint32)]
    L.bpx_plan_info.argtypes = [c_p, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(f32)]
    L.bpx_plan_benchmark.argtypes = [c_p, i32, i32, ctypes.POINTER(Timing)]
    L.bpx_plan_destroy.argtypes = [c_p]
    L.bpx_plan_destroy.restype = None
    L.bpx_render_transient.argtypes = [c_p, u32, u32, f32, f32, c_p, u64, f32, ctypes.c_int,
                                       u64, c_p]
    L.bpx_copy_benchmark.argtypes = [u64, i32, ctypes.POINTER(ctypes.c_double)]
    L.bpx_octree_backprojection.argtypes = [c_p, u64, u32, c_p, u32, c_p, c_p, c_p, c_p, f32,
                                            f32, c_p, c_p]
    L.bpx_octree_plan_create.argtypes = [c_p, u64, u32, c_p, u32, c_p, c_p, c_p, f32, f32, c_p,
                                         ctypes.POINTER(c_p)]
    L.bpx_octree_plan_stats.argtypes = [c_p, c_p]
    L.bpx_octree_plan_set_options.argtypes = [c_p, c_p]
    L.bpx_nlos_prepare.argtypes = [c_p, c_p, ctypes.c_int, c_p, c_p, c_p, c_p, c_p, c_p, f32, f32,
                                   ctypes.c_int, c_p, c_p, c_p, c_p, ctypes.c_int,
                                   ctypes.POINTER(NlosPrepared)]
    L.bpx_nlos_prepared_free.argtypes = [ctypes.POINTER(NlosPrepared)]
    L.bpx_nlos_prepared_free.restype = None
    L.bpx_net_param_count.restype = u32
    L.bpx_net_param_count.argtypes = []
    L.bpx_net_create.argtypes = [u64, ctypes.POINTER(c_p)]
    L.bpx_net_destroy.argtypes = [c_p]
    L.bpx_net_destroy.restype = None
    L.bpx_net_get_params.argtypes = [c_p, c_p]
    L.bpx_net_set_params.argtypes = [c_p, c_p]
    L.bpx_net_forward.argtypes = [c_p, c_p, c_p, c_p]
    L.bpx_net_gradients.argtypes = [c_p, c_p, c_p, c_p, c_p, f32,
                                    ctypes.POINTER(ctypes.c_double), c_p]
    L.bpx_net_train_octree.argtypes = [c_p, c_p, c_p, f32, f32, f32,
                                       ctypes.POINTER(ctypes.c_double)]
    L.bpx_octree_plan_attach_net.argtypes = [c_p, c_p]
    L.bpx_batch_create.argtypes = [ctypes.POINTER(CaptureDesc), u32, c_p, ctypes.POINTER(c_p)]
    L.bpx_batch_size.argtypes = [c_p]
    L.bpx_batch_size.restype = u32
    L.bpx_batch_tune.argtypes = [c_p, i32, ctypes.POINTER(BatchConfig)]
    L.bpx_batch_set_config.argtypes = [c_p, ctypes.POINTER(BatchConfig)]
    L.bpx_batch_get_config.argtypes = [c_p, ctypes.POINTER(BatchConfig)]
    L.bpx_batch_execute.argtypes = [c_p, c_p]
    L.bpx_batch_synchronize.argtypes = [c_p]
    L.bpx_batch_read_volume.argtypes = [c_p, u32, c_p]
    L.bpx_batch_latencies.argtypes = [c_p, c_p]
    L.bpx_batch_benchmark.argtypes = [c_p, i32, i32, ctypes.POINTER(BatchTiming)]
    L.bpx_batch_destroy.argtypes = [c_p]
    L.bpx_batch_destroy.restype = None
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        msg = lib().bpx_last_error().decode(errors="replace")
        if status == 1:
            raise InvalidArgument(status, msg)
        raise BpxError(status, msg)


def ptr(a: np.ndarray):
    """Pointer argument for a numpy array.  data_as() keeps a reference to the array inside the
    returned ctypes object, so temporaries such as ptr(_f32(x)) stay alive for the whole call
    (the library runs without the GIL, so a freed temporary could be reused meanwhile)."""
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def device_count() -> int:
    return int(lib().bpx_device_count())


def make_options(storage="f32", kahan_block=0, flags=0, variant=-1) -> Options:
    if isinstance(storage, str):
        storage = STORAGE[storage.lower()]
    return Options(int(storage), int(kahan_block), int(flags), int(variant))
