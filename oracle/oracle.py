"""ctypes wrapper around oracle/liboracle.so -- the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker.  The product library
(backprojector_amd) never imports it.

Each wrapper names the C function it calls; the C file cites the reference file:line
that function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

PAIR_DTYPE = np.dtype(
    [("camera_point", np.float32, (3,)), ("laser_point", np.float32, (3,)),
     ("camera_wall", np.float32), ("laser_wall", np.float32)]
)
assert PAIR_DTYPE.itemsize == 32


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def make_pairs(camera_point, laser_point, camera_wall, laser_wall):
    n = len(camera_wall)
    pr = np.zeros(n, dtype=PAIR_DTYPE)
    pr["camera_point"] = camera_point
    pr["laser_point"] = laser_point
    pr["camera_wall"] = camera_wall
    pr["laser_wall"] = laser_wall
    return pr


# ---------------------------------------------------------------- A.1 (GPU formula)
def a1_voxel_position(zero, inc, ix, iy, iz):
    out = np.zeros(3, np.float32)
    lib().or_a1_voxel_position(_p(_f32(zero)), _p(_f32(inc)), ctypes.c_uint32(ix),
                               ctypes.c_uint32(iy), ctypes.c_uint32(iz), _p(out))
    return out


def a1_bins(pairs, voxel_xyz, pair_idx, zero, inc, t0, dT, T):
    """or_a1_bins: (bin index or -1, rounded float bin) per (voxel, pair) query."""
    vx = _u32(voxel_xyz).reshape(-1, 3)
    pi = _u32(pair_idx).reshape(-1)
    n = len(pi)
    bins = np.zeros(n, np.int32)
    rnd = np.zeros(n, np.float32)
    lib().or_a1_bins(_p(pairs), _p(vx), _p(pi), ctypes.c_uint64(n), _p(_f32(zero)),
                     _p(_f32(inc)), ctypes.c_float(t0), ctypes.c_float(dT),
                     ctypes.c_uint32(T), _p(bins), _p(rnd))
    return bins, rnd


def a1_confocal_bin(lw, cw, s, t0, dT):
    """or_a1_confocal_bin: rounded bin of a confocal pair at squared distance s."""
    L = lib()
    L.or_a1_confocal_bin.restype = ctypes.c_float
    return float(L.or_a1_confocal_bin(ctypes.c_float(lw), ctypes.c_float(cw), ctypes.c_float(s),
                                      ctypes.c_float(t0), ctypes.c_float(dT)))


def confocal_thresholds(lw, cw, n, t0, dT):
    """or_confocal_thresholds: [P][n] smallest squared distance whose bin is >= j."""
    lw, cw = _f32(lw).reshape(-1), _f32(cw).reshape(-1)
    out = np.zeros((len(lw), int(n)), np.float32)
    lib().or_confocal_thresholds(_p(lw), _p(cw), ctypes.c_uint32(len(lw)), ctypes.c_uint32(n),
                                 ctypes.c_float(t0), ctypes.c_float(dT), _p(out))
    return out


def a1_volume(H, pairs, N, zero, inc, t0, dT, block=0):
    """or_a1_volume_f32: reference GPU-path volume in the reference summation order."""
    H = _f32(H)
    P = len(pairs)
    T = H.size // P
    N = _u32(N)
    out = np.zeros(tuple(int(n) for n in N), np.float32)
    lib().or_a1_volume_f32(_p(H), ctypes.c_uint32(T), _p(pairs), ctypes.c_uint32(P), _p(N),
                           _p(_f32(zero)), _p(_f32(inc)), ctypes.c_float(t0),
                           ctypes.c_float(dT), ctypes.c_uint32(block), _p(out))
    return out


def a1_volume_f64acc(H, pairs, N, zero, inc, t0, dT):
    """or_a1_volume_f64acc: same fp32 taps, all pairs, exact (double) accumulation."""
    H = _f32(H)
    P = len(pairs)
    T = H.size // P
    N = _u32(N)
    out = np.zeros(tuple(int(n) for n in N), np.float64)
    lib().or_a1_volume_f64acc(_p(H), ctypes.c_uint32(T), _p(pairs), ctypes.c_uint32(P), _p(N),
                              _p(_f32(zero)), _p(_f32(inc)), ctypes.c_float(t0),
                              ctypes.c_float(dT), _p(out))
    return out


def a1_voxels_f64acc(H, pairs, voxel_xyz, zero, inc, t0, dT):
    """or_a1_voxels_f64acc: exact-accumulated A.1 sums for a list of voxels [n][3]."""
    H = _f32(H)
    P = len(pairs)
    T = H.size // P
    vx = _u32(voxel_xyz).reshape(-1, 3)
    out = np.zeros(len(vx), np.float64)
    lib().or_a1_voxels_f64acc(_p(H), ctypes.c_uint32(T), _p(pairs), ctypes.c_uint32(P),
                              _p(_f32(zero)), _p(_f32(inc)), ctypes.c_float(t0),
                              ctypes.c_float(dT), _p(vx), ctypes.c_int64(len(vx)), _p(out))
    return out


def fp64_voxels(H, pairs, voxel_xyz, zero, inc, t0, dT):
    """or_fp64_voxels: FP64 oracle sums for a list of voxels [n][3] (real samples)."""
    H = _f32(H)
    P = len(pairs)
    T = H.size // P
    vx = _u32(voxel_xyz).reshape(-1, 3)
    out = np.zeros(len(vx), np.float64)
    lib().or_fp64_voxels(_p(H), ctypes.c_uint32(T), _p(pairs), ctypes.c_uint32(P),
                         _p(_f32(zero)), _p(_f32(inc)), ctypes.c_double(t0),
                         ctypes.c_double(dT), _p(vx), ctypes.c_int64(len(vx)), _p(out))
    return out


def a1_volume_c64(H, pairs, N, zero, inc, t0, dT, block=0):
    """or_a1_volume_c64: complex samples (interleaved), reference order."""
    Hc = np.ascontiguousarray(H, dtype=np.complex64)
    P = len(pairs)
    T = Hc.size // P
    N = _u32(N)
    out = np.zeros(tuple(int(n) for n in N), np.complex64)
    lib().or_a1_volume_c64(_p(Hc), ctypes.c_uint32(T), _p(pairs), ctypes.c_uint32(P), _p(N),
                           _p(_f32(zero)), _p(_f32(inc)), ctypes.c_float(t0),
                           ctypes.c_float(dT), ctypes.c_uint32(block), _p(out))
    return out


def fp64_volume(H, pairs, N, zero, inc, t0, dT):
    """or_fp64_volume: the A.1 geometry entirely in binary64 (error yardstick)."""
    cplx = np.iscomplexobj(H)
    Hf = np.ascontiguousarray(H, dtype=np.complex64 if cplx else np.float32)
    P = len(pairs)
    T = Hf.size // P
    N = _u32(N)
    out = np.zeros(tuple(int(n) for n in N), np.complex128 if cplx else np.float64)
    lib().or_fp64_volume(_p(Hf), ctypes.c_uint32(T), _p(pairs), ctypes.c_uint32(P), _p(N),
                         _p(_f32(zero)), _p(_f32(inc)), ctypes.c_double(t0),
                         ctypes.c_double(dT), ctypes.c_int(1 if cplx else 0), _p(out))
    return out


# ---------------------------------------------------------------- A.2 (CPU formula)
def a2_grid(volume_position, volume_size, N, diag_sum_in_double=True):
    zero = np.zeros(3, np.float32)
    vs = np.zeros(3, np.float32)
    w = ctypes.c_float(0)
    lib().or_a2_grid(_p(_f32(volume_position)), _p(_f32(volume_size)), _p(_u32(N)),
                     ctypes.c_int(int(diag_sum_in_double)), _p(zero), _p(vs), ctypes.byref(w))
    return zero, vs, np.float32(w.value)


def a2_volume(H, pairs, N, volume_position, volume_size, t0, dT, diag_sum_in_double=True,
              vox_begin=0, vox_end=0):
    """or_a2_volume: reference CPU path (voxel-diagonal window, double distances)."""
    cplx = np.iscomplexobj(H)
    Hf = np.ascontiguousarray(H, dtype=np.complex64 if cplx else np.float32)
    P = len(pairs)
    T = Hf.size // P
    N = _u32(N)
    out = np.zeros(tuple(int(n) for n in N), np.complex64 if cplx else np.float32)
    lib().or_a2_volume(_p(Hf), ctypes.c_uint32(T), _p(pairs), ctypes.c_uint32(P), _p(N),
                       _p(_f32(volume_position)), _p(_f32(volume_size)), ctypes.c_float(t0),
                       ctypes.c_float(dT), ctypes.c_int(1 if cplx else 0),
                       ctypes.c_int(int(diag_sum_in_double)), ctypes.c_int64(vox_begin),
                       ctypes.c_int64(vox_end), _p(out))
    return out


def a2_volume_fp64(H, pairs, N, volume_position, volume_size, t0, dT):
    H = _f32(H)
    P = len(pairs)
    T = H.size // P
    N = _u32(N)
    out = np.zeros(tuple(int(n) for n in N), np.float64)
    lib().or_a2_volume_fp64(_p(H), ctypes.c_uint32(T), _p(pairs), ctypes.c_uint32(P), _p(N),
                            _p(_f32(volume_position)), _p(_f32(volume_size)),
                            ctypes.c_double(t0), ctypes.c_double(dT), _p(out))
    return out


# ---------------------------------------------------------------- host pipeline
def precompute_pairs(laser_position, laser_grid, camera_position, camera_grid, confocal):
    lg = _f32(laser_grid)
    cg = _f32(camera_grid)
    Lx, Ly = lg.shape[:2]
    Cx, Cy = cg.shape[:2]
    n = Cx * Cy if confocal else Lx * Ly * Cx * Cy
    out = np.zeros(n, PAIR_DTYPE)
    lib().or_precompute_pairs(_p(_f32(laser_position)), _p(lg), ctypes.c_uint32(Lx),
                              ctypes.c_uint32(Ly), _p(_f32(camera_position)), _p(cg),
                              ctypes.c_uint32(Cx), ctypes.c_uint32(Cy),
                              ctypes.c_int(1 if confocal else 0), _p(out))
    return out


def basis(volume_position, volume_size, N, wall_normal=(0, 0, 0)):
    zero = np.zeros(3, np.float32)
    inc = np.zeros(9, np.float32)
    lib().or_basis(_p(_f32(volume_position)), _p(_f32(volume_size)), _p(_u32(N)),
                   _p(_f32(wall_normal)), _p(zero), _p(inc))
    return zero, inc


def time_limits(laser_position, laser_grid, camera_position, camera_grid, volume_size,
                volume_position, t0, dT):
    lg = _f32(laser_grid)
    cg = _f32(camera_grid)
    mn = ctypes.c_float(0)
    mx = ctypes.c_float(0)
    rc = lib().or_time_limits(_p(_f32(laser_position)), _p(lg), ctypes.c_uint32(lg.shape[0]),
                              ctypes.c_uint32(lg.shape[1]), _p(_f32(camera_position)), _p(cg),
                              ctypes.c_uint32(cg.shape[0]), ctypes.c_uint32(cg.shape[1]),
                              _p(_f32(volume_size)), _p(_f32(volume_position)),
                              ctypes.c_float(t0), ctypes.c_float(dT), ctypes.byref(mn),
                              ctypes.byref(mx))
    if rc != 0:
        raise ValueError("time limits need a camera grid with at least 3 rows")
    return np.float32(mn.value), np.float32(mx.value)


def transient_chunk(data, min_idx, max_idx):
    cplx = np.iscomplexobj(data)
    d = np.ascontiguousarray(data, dtype=np.complex64 if cplx else np.float32)
    T = d.shape[-1]
    rows = d.size // T
    lib().or_chunk_length.restype = ctypes.c_uint32
    Tc = lib().or_chunk_length(ctypes.c_float(min_idx), ctypes.c_float(max_idx))
    out = np.zeros(d.shape[:-1] + (Tc,), d.dtype)
    lib().or_transient_chunk(_p(d), ctypes.c_uint64(rows), ctypes.c_uint32(T),
                             ctypes.c_float(min_idx), ctypes.c_float(max_idx),
                             ctypes.c_int(1 if cplx else 0), _p(out))
    return out


def chunked_t0(t0, min_idx, dT):
    lib().or_chunked_t0.restype = ctypes.c_float
    return np.float32(lib().or_chunked_t0(ctypes.c_float(t0), ctypes.c_float(min_idx),
                                          ctypes.c_float(dT)))


def phasor_taps(wavelength, dT, times=4, max_taps=1 << 16):
    c = np.zeros(max_taps, np.float32)
    s = np.zeros(max_taps, np.float32)
    n = lib().or_phasor_taps(ctypes.c_float(wavelength), ctypes.c_float(dT),
                             ctypes.c_uint32(times), _p(c), _p(s), ctypes.c_int(max_taps))
    if n <= 0:
        raise ValueError("invalid phasor pulse")
    return c[:n].copy(), s[:n].copy()


def phasor_filter(chunk, wavelength, dT, times=4):
    c, s = phasor_taps(wavelength, dT, times)
    x = _f32(chunk)
    Tc = x.shape[-1]
    rows = x.size // Tc
    out = np.zeros(x.shape, np.complex64)
    lib().or_phasor_filter(_p(x), ctypes.c_uint64(rows), ctypes.c_uint32(Tc), _p(c), _p(s),
                           ctypes.c_int(len(c)), _p(out))
    return out


def num_threads():
    return int(lib().or_num_threads())


# ---------------------------------------------------------------- full pipelines
def backproject_pipeline(transient_data, camera_grid, laser_grid, camera_position,
                         laser_position, t0, dT, is_confocal, volume_position, volume_size,
                         voxels_per_side, wall_normal=(0, 0, 0), use_cpu=False):
    """bp::backproject (backproject.hpp:762-909), row-major inputs, GPU or CPU formula.
    Returns (volume, dict of intermediates)."""
    mn, mx = time_limits(laser_position, laser_grid, camera_position, camera_grid,
                         volume_size, volume_position, t0, dT)
    pairs = precompute_pairs(laser_position, laser_grid, camera_position, camera_grid,
                             is_confocal)
    zero, inc = basis(volume_position, volume_size, voxels_per_side, wall_normal)
    chunk = transient_chunk(transient_data, mn, mx)
    ct0 = chunked_t0(t0, mn, dT)
    Tc = chunk.shape[-1]
    H = chunk.reshape(-1, Tc)
    if use_cpu:
        vol = a2_volume(H, pairs, voxels_per_side, volume_position, volume_size, ct0, dT)
    else:
        vol = a1_volume(H, pairs, voxels_per_side, zero, inc, ct0, dT)
    return vol, dict(min_idx=mn, max_idx=mx, pairs=pairs, zero=zero, inc=inc, H=H, t0=ct0)


def phasor_pipeline(transient_data, camera_grid, laser_grid, camera_position, laser_position,
                    t0, dT, is_confocal, volume_position, volume_size, voxels_per_side,
                    wavelength, use_cpu=False):
    """bp::phasor_reconstruction (backproject.hpp:932-1039), row-major, Naive access."""
    mn, mx = time_limits(laser_position, laser_grid, camera_position, camera_grid,
                         volume_size, volume_position, t0, dT)
    imn, imx = int(mn), int(mx)  # :962-969 tuple<float,float> -> int truncation
    zero, inc = basis(volume_position, volume_size, voxels_per_side)
    chunk = transient_chunk(transient_data, np.float32(imn), np.float32(imx))
    cchunk = phasor_filter(chunk, wavelength, dT)
    pairs = precompute_pairs(laser_position, laser_grid, camera_position, camera_grid,
                             is_confocal)
    ct0 = chunked_t0(t0, np.float32(imn), dT)
    Tc = cchunk.shape[-1]
    H = cchunk.reshape(-1, Tc)
    if use_cpu:
        vol = a2_volume(H, pairs, voxels_per_side, volume_position, volume_size, ct0, dT)
    else:
        vol = a1_volume_c64(H, pairs, voxels_per_side, zero, inc, ct0, dT)
    return vol, dict(min_idx=imn, max_idx=imx, pairs=pairs, zero=zero, inc=inc, H=H, t0=ct0)
