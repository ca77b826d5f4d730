"""Python front of the C ABI: numpy in, numpy out.  All compute runs in libbpx.so on the GPU.

Function names follow the reference's C++ interface (backproject_cuda.hpp:19-49,
backproject.hpp:236-247, 430-479); each docstring names the reference it replaces.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from ._native import PAIR_DTYPE, check, lib, ptr


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _vec3(a, name="vector"):
    """A float32 3-vector for the C ABI, which reads exactly 3 floats; any other size raises."""
    v = _f32(a).reshape(-1)
    if v.size != 3:
        raise ValueError(f"{name} must have 3 elements, got {v.size}")
    return v


def _mat9(a, name="voxel_inc"):
    """The 3x3 row-major voxel increment matrix as 9 float32 values."""
    v = _f32(a).reshape(-1)
    if v.size != 9:
        raise ValueError(f"{name} must have 9 elements, got {v.size}")
    return v


def as_pairs(pairs) -> np.ndarray:
    """Coerces to the CameraLaserPair record layout (backproject_cuda.hpp:9-13)."""
    pr = np.ascontiguousarray(pairs)
    if pr.dtype != PAIR_DTYPE:
        if pr.dtype.names and set(pr.dtype.names) >= set(PAIR_DTYPE.names):
            out = np.zeros(len(pr), PAIR_DTYPE)
            for n in PAIR_DTYPE.names:
                out[n] = pr[n]
            return out
        raise TypeError("pairs must be a structured array with the CameraLaserPair fields")
    return pr


def _samples(chunk, complex_samples):
    if complex_samples:
        c = np.ascontiguousarray(chunk, dtype=np.complex64)
        return c, c.size
    c = _f32(chunk)
    return c, c.size


def backprojection(transient_chunk, pairs, voxels_per_side, volume_zero_pos, voxel_inc, t0,
                   deltaT, storage="f32", kahan_block=0, reference_pair_truncation=False,
                   variant=-1):
    """call_cuda_backprojection / call_cuda_complex_backprojection
    (backproject_cuda.cu:370-500, 671-801): single-tap GPU-path volume [N0][N1][N2].
    Complex input selects the complex path (complex64 volume)."""
    cplx = np.iscomplexobj(transient_chunk)
    H, n = _samples(transient_chunk, cplx)
    pr = as_pairs(pairs)
    P = len(pr)
    T = (n // P) if P else (H.shape[-1] if H.ndim else 0)
    N = _u32(voxels_per_side).reshape(3)
    out = np.zeros(tuple(int(x) for x in N), np.complex64 if cplx else np.float32)
    flags = nat.FLAG_REFERENCE_PAIR_TRUNCATION if reference_pair_truncation else 0
    opt = nat.make_options(storage, kahan_block, flags, variant)
    check(lib().bpx_backprojection_ex(ptr(H), n, T, ptr(pr), P, ptr(out), ptr(N),
                                      ptr(_vec3(volume_zero_pos, "volume_zero_pos")), ptr(_mat9(voxel_inc)),
                                      float(t0), float(deltaT), 1 if cplx else 0,
                                      ctypes.byref(opt)))
    return out


def window_backprojection(transient_chunk, pairs, voxels_per_side, volume_position, volume_size,
                          t0, deltaT):
    """classic_backprojection (backproject.hpp:236-247): the CPU path's semantics (grid
    spacing size/N, voxel-diagonal window, double distances), computed on the GPU."""
    cplx = np.iscomplexobj(transient_chunk)
    H, n = _samples(transient_chunk, cplx)
    pr = as_pairs(pairs)
    P = len(pr)
    T = (n // P) if P else 0
    N = _u32(voxels_per_side).reshape(3)
    out = np.zeros(tuple(int(x) for x in N), np.complex64 if cplx else np.float32)
    check(lib().bpx_window_backprojection(ptr(H), n, T, ptr(pr), P, ptr(out), ptr(N),
                                          ptr(_vec3(volume_position, "volume_position")), ptr(_vec3(volume_size, "volume_size")),
                                          float(t0), float(deltaT), 1 if cplx else 0))
    return out


def phasor_pulse(transient_data, wavelength, deltaT, times=4, sigma=0.0):
    """phasor_pulse (backproject.hpp:430-479): complex FIR along the last axis on the GPU.
    sigma > 0 selects the explicit-sigma extension."""
    x = _f32(transient_data)
    T = x.shape[-1]
    rows = x.size // T if T else 0
    out = np.zeros(x.shape, np.complex64)
    check(lib().bpx_phasor_pulse(ptr(x), rows, T, float(wavelength), float(deltaT), int(times),
                                 float(sigma), ptr(out)))
    return out


def row_windows(pairs, voxels_per_side, volume_zero_pos, voxel_inc, t0, deltaT, Tc):
    """[P][2] uint32 element windows [lo, hi) of the fused rows that a confocal plan of this
    geometry stages at each execution (no reference counterpart; host only, no device call)."""
    pr = as_pairs(pairs)
    out = np.zeros((len(pr), 2), np.uint32)
    check(lib().bpx_row_windows(ptr(pr), len(pr), ptr(_u32(voxels_per_side).reshape(3)),
                                ptr(_vec3(volume_zero_pos, "volume_zero_pos")), ptr(_mat9(voxel_inc)),
                                float(t0), float(deltaT), int(Tc), ptr(out)))
    return out


def render_transient(pairs, T, t0, deltaT, samples, peak_count=1000.0, poisson=True, seed=0):
    """Synthetic capture [P][T] of hidden samples [n][7] = (x, y, z, nx, ny, nz, area)."""
    pr = as_pairs(pairs)
    s = _f32(samples).reshape(-1, 7)
    out = np.zeros((len(pr), int(T)), np.float32)
    check(lib().bpx_render_transient(ptr(pr), len(pr), int(T), float(t0), float(deltaT), ptr(s),
                                     len(s), float(peak_count), 1 if poisson else 0,
                                     int(seed) & ((1 << 64) - 1), ptr(out)))
    return out


def octree_backprojection(transient_chunk, pairs, voxels_per_side, volume_zero_pos, voxel_inc,
                          t0, deltaT, levels=4, threshold=0.25, proposal_threshold=0.5):
    """Coarse-to-fine single-tap reconstruction (replaces the reference's non-functional
    call_cuda_octree_backprojection, backproject_cuda.cu:503-669).  Returns (volume, stats)."""
    H, n = _samples(transient_chunk, False)
    pr = as_pairs(pairs)
    P = len(pr)
    T = (n // P) if P else 0
    N = _u32(voxels_per_side).reshape(3)
    out = np.zeros(tuple(int(x) for x in N), np.float32)
    o = nat.make_octree_options(levels, threshold, proposal_threshold)
    st = nat.OctreeStats()
    check(lib().bpx_octree_backprojection(ptr(H), n, T, ptr(pr), P, ptr(out), ptr(N),
                                          ptr(_vec3(volume_zero_pos, "volume_zero_pos")),
                                          ptr(_mat9(voxel_inc)), float(t0),
                                          float(deltaT), ctypes.byref(o), ctypes.byref(st)))
    return out, st.to_dict()


class Plan:
    """Device-resident capture + volume; execute() runs the whole reconstruction on the GPU."""

    def __init__(self, transient_chunk, pairs, voxels_per_side, volume_zero_pos, voxel_inc, t0,
                 deltaT, storage="f32", kahan_block=0, reference_pair_truncation=False,
                 variant=-1, keep_f32_rows=False):
        cplx = np.iscomplexobj(transient_chunk)
        H, n = _samples(transient_chunk, cplx)
        pr = as_pairs(pairs)
        P = len(pr)
        T = (n // P) if P else 0
        self.N = tuple(int(x) for x in _u32(voxels_per_side).reshape(3))
        self.complex = cplx
        flags = nat.FLAG_REFERENCE_PAIR_TRUNCATION if reference_pair_truncation else 0
        if keep_f32_rows:   # no exact fp16 staging of integer samples (see bpx.h)
            flags |= nat.FLAG_KEEP_F32_ROWS
        opt = nat.make_options(storage, kahan_block, flags, variant)
        h = ctypes.c_void_p()
        check(lib().bpx_plan_create(ptr(H), n, T, ptr(pr), P, ptr(_u32(self.N)),
                                    ptr(_vec3(volume_zero_pos, "volume_zero_pos")), ptr(_mat9(voxel_inc)),
                                    float(t0), float(deltaT), 1 if cplx else 0, ctypes.byref(opt),
                                    ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_nlos(cls, transient_data, camera_grid_positions, laser_grid_positions,
                  camera_position, laser_position, t0, deltaT, is_confocal, volume_position,
                  volume_size, voxels_per_side, wall_normal=(0, 0, 0), assume_row_major=True,
                  storage="f32", kahan_block=0, variant=-1, centred_wall_basis=False):
        """Plan of nlosbpy.backproject's GPU path (bp::backproject, backproject.hpp:762-891).
        centred_wall_basis=True builds a non-zero wall normal's volume basis right-handed and
        centred on volume_position (the reference's basis, backproject.hpp:825-839, points z
        away from the volume for the normal (0, -1, 0) and is NaN for normals parallel to z)."""
        H = _f32(transient_data)
        cg, lg = _f32(camera_grid_positions), _f32(laser_grid_positions)
        N = _u32(voxels_per_side).reshape(3)
        sh = lambda a: np.ascontiguousarray(np.asarray(a.shape, np.int64))
        ds, cs, ls = sh(H), sh(cg), sh(lg)
        opt = nat.make_options(storage, kahan_block, nat.FLAG_CENTRED_WALL_BASIS if centred_wall_basis else 0,
                               variant)
        self = cls.__new__(cls)
        self.N = tuple(int(x) for x in N)
        self.complex = False
        h = ctypes.c_void_p()
        check(lib().bpx_nlos_plan_create(ptr(H), ptr(ds), H.ndim, ptr(cg), ptr(cs), ptr(lg), ptr(ls),
                                         ptr(_vec3(camera_position, "camera_position")), ptr(_vec3(laser_position, "laser_position")),
                                         float(t0), float(deltaT), 1 if is_confocal else 0,
                                         ptr(_vec3(volume_position, "volume_position")), ptr(_vec3(volume_size, "volume_size")),
                                         ptr(N), ptr(_vec3(wall_normal, "wall_normal")),
                                         1 if assume_row_major else 0, ctypes.byref(opt),
                                         ctypes.byref(h)))
        self._h = h
        return self

    @classmethod
    def octree(cls, transient_chunk, pairs, voxels_per_side, volume_zero_pos, voxel_inc, t0,
               deltaT, levels=4, threshold=0.25, proposal_threshold=0.5, bounds_only=False):
        """Device-resident octree plan (see octree_backprojection).  bounds_only runs the coarse
        levels only and skips the exact finest level (training data for Net.train_octree)."""
        H, n = _samples(transient_chunk, False)
        pr = as_pairs(pairs)
        P = len(pr)
        T = (n // P) if P else 0
        self = cls.__new__(cls)
        self.N = tuple(int(x) for x in _u32(voxels_per_side).reshape(3))
        self.complex = False
        o = nat.make_octree_options(levels, threshold, proposal_threshold, bounds_only)
        self.levels = int(levels)
        h = ctypes.c_void_p()
        check(lib().bpx_octree_plan_create(ptr(H), n, T, ptr(pr), P, ptr(_u32(self.N)),
                                           ptr(_vec3(volume_zero_pos, "volume_zero_pos")),
                                           ptr(_mat9(voxel_inc)), float(t0),
                                           float(deltaT), ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        return self

    def set_octree_options(self, threshold=0.25, proposal_threshold=0.5, bounds_only=False,
                           keep_bound_ratio=0.0, fallback_threshold=0.0, coarse_pair_stride=1):
        """Replaces an octree plan's thresholds (proposal_threshold: one value or one per level;
        keep_bound_ratio > 0: children of blocks whose bound reaches that fraction of the running
        max are refined whatever an attached net says; fallback_threshold (one value or one per
        level): children the net does not refine but rates above this are computed at full
        resolution; coarse_pair_stride s: with a net attached, the coarse levels use every s-th
        wall point along each grid axis)."""
        o = nat.make_octree_options(self.levels, threshold, proposal_threshold, bounds_only, keep_bound_ratio,
                                    fallback_threshold, coarse_pair_stride)
        check(lib().bpx_octree_plan_set_options(self._h, ctypes.byref(o)))

    def attach_net(self, net):
        """Attaches a learned proposal Net to an octree plan (None detaches)."""
        check(lib().bpx_octree_plan_attach_net(self._h, net._h if net is not None else None))
        self._net = net

    def octree_stats(self):
        st = nat.OctreeStats()
        check(lib().bpx_octree_plan_stats(self._h, ctypes.byref(st)))
        return st.to_dict()

    def execute(self, stream=None):
        check(lib().bpx_plan_execute(self._h, stream))

    def synchronize(self):
        check(lib().bpx_plan_synchronize(self._h))

    def volume(self):
        out = np.zeros(self.N, np.complex64 if self.complex else np.float32)
        check(lib().bpx_plan_read_volume(self._h, ptr(out)))
        return out

    def thresholds(self):
        """The confocal threshold table of the last execution, [pairs_used][Tc + 1] float32."""
        inf = self.info()
        out = np.zeros((inf["pairs_used"], inf["Tc"] + 1), np.float32)
        check(lib().bpx_plan_read_thresholds(self._h, ptr(out), out.size))
        return out

    def info(self):
        path, k = ctypes.c_int32(), ctypes.c_int32()
        used, tc, sc = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_float()
        check(lib().bpx_plan_info(self._h, ctypes.byref(path), ctypes.byref(k), ctypes.byref(used),
                                  ctypes.byref(tc), ctypes.byref(sc)))
        names = {0: "confocal", 1: "grid", 2: "pairs", 3: "window"}
        kname = ctypes.create_string_buffer(160)
        check(lib().bpx_plan_kernel_name(self._h, kname, len(kname)))
        rows = ctypes.c_uint32()
        check(lib().bpx_plan_row_storage(self._h, ctypes.byref(rows)))
        return dict(path=names.get(path.value, path.value), lane_voxels=k.value,
                    pairs_used=used.value, Tc=tc.value, scale=sc.value, kernel=kname.value.decode(),
                    row_storage={0: "f32", 1: "f16", 2: "bf16"}.get(rows.value, rows.value))

    def benchmark(self, warmup=2, iterations=10):
        t = nat.Timing()
        check(lib().bpx_plan_benchmark(self._h, int(warmup), int(iterations), ctypes.byref(t)))
        return dict(total_ms=t.total_ms, per_exec_ms=t.per_exec_ms,
                    table_kernel_ms=t.table_kernel_ms, gather_kernel_ms=t.gather_kernel_ms,
                    iterations=t.iterations)

    def tune(self, reps=3):
        """bpx_plan_tune: a confocal plan with fused rows times every gather variant that applies
        to it on its own data and keeps the fastest (the same volume bit for bit); returns the
        chosen variant and its kernel.  Other plans keep their variant."""
        v = ctypes.c_int32()
        check(lib().bpx_plan_tune(self._h, int(reps), ctypes.byref(v)))
        return dict(variant=int(v.value), kernel=self.info()["kernel"])

    def close(self):
        if getattr(self, "_h", None):
            lib().bpx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Net:
    """Learned octree proposal net (4 x 3x3x3 conv, 16 channels, BF16 MFMA, FP32 accumulation).
    No reference counterpart: it narrows the refinement of octree plans it is attached to."""

    def __init__(self, seed=0, params=None):
        h = ctypes.c_void_p()
        check(lib().bpx_net_create(int(seed), ctypes.byref(h)))
        self._h = h
        if params is not None:
            self.set_params(params)

    @staticmethod
    def param_count():
        return int(lib().bpx_net_param_count())

    def params(self):
        out = np.zeros(self.param_count(), np.float32)
        check(lib().bpx_net_get_params(self._h, ptr(out)))
        return out

    def set_params(self, params):
        p = _f32(params).reshape(-1)
        if p.size != self.param_count():
            raise ValueError(f"expected {self.param_count()} parameters, got {p.size}")
        check(lib().bpx_net_set_params(self._h, ptr(p)))

    def forward(self, x):
        """x: [n0][n1][n2][16] float32 (channels >= 5 zero); returns logits [n0][n1][n2][16]."""
        x = _f32(x)
        dims = _u32(x.shape[:3])
        out = np.zeros(x.shape[:3] + (16,), np.float32)
        check(lib().bpx_net_forward(self._h, ptr(x), ptr(dims), ptr(out)))
        return out

    def gradients(self, x, labels, mask=None, pos_weight=1.0):
        """Weighted BCE loss over [..][8] labels and its gradient w.r.t. every parameter."""
        x = _f32(x)
        dims = _u32(x.shape[:3])
        lab = np.ascontiguousarray(labels, np.uint8)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        loss = ctypes.c_double()
        g = np.zeros(self.param_count(), np.float32)
        check(lib().bpx_net_gradients(self._h, ptr(x), ptr(dims), ptr(lab), ptr(m),
                                      float(pos_weight), ctypes.byref(loss), ptr(g)))
        return loss.value, g

    def train_octree(self, octree_plan, dense_plan, occupied_fraction=0.25, pos_weight=4.0,
                     lr=1e-3):
        """One Adam step on one scene: octree_plan (typically bounds_only=True) and dense_plan
        have both been executed on the same grid."""
        loss = ctypes.c_double()
        check(lib().bpx_net_train_octree(self._h, octree_plan._h, dense_plan._h,
                                         float(occupied_fraction), float(pos_weight), float(lr),
                                         ctypes.byref(loss)))
        return loss.value

    def close(self):
        if getattr(self, "_h", None):
            lib().bpx_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many confocal captures reconstructed by one persistent work-queue launch replayed from a
    hipGraph (no reference counterpart; each volume equals backprojection() of that capture
    alone with the same storage).  captures: iterable of (transient_chunk, pairs,
    voxels_per_side, volume_zero_pos, voxel_inc, t0, deltaT); storage "f32", "f16" or "bf16"
    for all of them.  Complex chunks (all or none) give complex volumes."""

    def __init__(self, captures, kahan_block=0, reference_pair_truncation=False, storage="f32"):
        self._keep = []
        descs = []
        self.shapes = []
        captures = list(captures)
        kinds = {bool(np.iscomplexobj(c[0])) for c in captures}
        if len(kinds) > 1:
            raise ValueError("a batch takes real or complex captures, not both")
        self.complex = kinds == {True}
        for (chunk, pairs, N, zero, inc, t0, dT) in captures:
            H, n = _samples(chunk, self.complex)
            pr = as_pairs(pairs)
            P = len(pr)
            Nv = _u32(N).reshape(3)
            d = nat.CaptureDesc()
            d.transient_chunk = H.ctypes.data
            d.transient_size = n
            d.T = (n // P) if P else 1     # an empty capture still needs T > 0
            d.scanned_pairs = pr.ctypes.data
            d.num_pairs = P
            d.voxels_per_side[:] = [int(x) for x in Nv]
            d.volume_zero_pos[:] = [float(x) for x in _vec3(zero, "volume_zero_pos")]
            d.voxel_inc[:] = [float(x) for x in _mat9(inc)]
            d.t0 = float(t0)
            d.deltaT = float(dT)
            self._keep += [H, pr]
            descs.append(d)
            self.shapes.append(tuple(int(x) for x in Nv))
        arr = (nat.CaptureDesc * max(len(descs), 1))(*descs)
        flags = nat.FLAG_REFERENCE_PAIR_TRUNCATION if reference_pair_truncation else 0
        if self.complex:
            flags |= nat.FLAG_COMPLEX_SAMPLES
        opt = nat.make_options(storage, kahan_block, flags, -1)
        h = ctypes.c_void_p()
        check(lib().bpx_batch_create(arr, len(descs), ctypes.byref(opt), ctypes.byref(h)))
        self._h = h
        self._keep = None   # the library staged every capture on the device

    def __len__(self):
        return int(lib().bpx_batch_size(self._h))

    def tune(self, reps=2):
        c = nat.BatchConfig()
        check(lib().bpx_batch_tune(self._h, int(reps), ctypes.byref(c)))
        return c.to_dict()

    def config(self):
        c = nat.BatchConfig()
        check(lib().bpx_batch_get_config(self._h, ctypes.byref(c)))
        return c.to_dict()

    def set_config(self, lane_voxels=2, groups=0, order=0, lds_depth=8, launches=1):
        c = nat.BatchConfig(int(lane_voxels), int(groups), int(order), 0, int(lds_depth), int(launches))
        check(lib().bpx_batch_set_config(self._h, ctypes.byref(c)))

    def execute(self, stream=None):
        check(lib().bpx_batch_execute(self._h, stream))

    def synchronize(self):
        check(lib().bpx_batch_synchronize(self._h))

    def volume(self, i):
        out = np.zeros(self.shapes[i], np.complex64 if self.complex else np.float32)
        check(lib().bpx_batch_read_volume(self._h, int(i), ptr(out)))
        return out

    def latencies(self):
        out = np.zeros(len(self.shapes), np.float64)
        check(lib().bpx_batch_latencies(self._h, ptr(out)))
        return out

    def benchmark(self, warmup=1, iterations=5):
        t = nat.BatchTiming()
        check(lib().bpx_batch_benchmark(self._h, int(warmup), int(iterations), ctypes.byref(t)))
        return dict(per_exec_ms=t.per_exec_ms, updates_per_s=t.updates_per_s,
                    updates=int(t.updates), iterations=t.iterations)

    def close(self):
        if getattr(self, "_h", None):
            lib().bpx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
