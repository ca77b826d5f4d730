"""Drop-in for the reference's Python module `nlosbpy` (nlosbpy.cpp:81-120).

Same functions, argument names, order and defaults:

    backproject(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
                laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
                voxels_per_side, wall_normal, assume_row_major=False, use_cpu=False)
        -> float32 [N0, N1, N2]
    phasor(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
           laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
           voxels_per_side, wavelength, assume_row_major=False, use_cpu=False)
        -> complex64 [N0, N1, N2]

Every computation runs on the GPU through libbpx.so.  `use_cpu=True` selects the reference
CPU path's *semantics* (grid spacing size/N, voxel-diagonal window, double-precision
distances, backproject.hpp:138-216), still computed on the GPU.  `assume_row_major=False`
means the arrays are column-major (Fortran-order) transposes of the row-major layout; they
are transposed before the pipeline runs (the reference's column-major branch reads out of
bounds, backproject.hpp:384-385, 405).

`phasor` and `phasor_multi` also take `output=` (keyword, default "complex": the reference's
result): "magnitude" (|v|, what the reference CLI writes, main.cpp:162-166), "phase"
(atan2(im, re)) or "magnitude_phase" (a (magnitude, phase) pair), computed on the device so
that only the real volumes are copied back.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._native import check, lib, ptr

__all__ = ["backproject", "phasor", "phasor_multi"]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _shape(a):
    return np.ascontiguousarray(np.asarray(a.shape, dtype=np.int64))


def _vec3(a, name):
    v = _f32(a).reshape(-1)
    if v.size != 3:
        raise ValueError(f"{name} must have 3 elements")
    return v


def _prep(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
          laser_position, volume_position, volume_size, voxels_per_side):
    H = _f32(transient_data)
    cg = _f32(camera_grid_positions)
    lg = _f32(laser_grid_positions)
    if cg.ndim != 3 or lg.ndim != 3:
        raise ValueError("grid positions must be 3-D arrays")
    N = np.ascontiguousarray(np.asarray(voxels_per_side, dtype=np.uint32).reshape(-1))
    if N.size != 3:
        raise ValueError("voxels_per_side must have 3 elements")
    return (H, cg, lg, _vec3(camera_position, "camera_position"),
            _vec3(laser_position, "laser_position"), _vec3(volume_position, "volume_position"),
            _vec3(volume_size, "volume_size"), N)


def backproject(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
                laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
                voxels_per_side, wall_normal, assume_row_major=False, use_cpu=False):
    """bp::backproject (backproject.hpp:762-909) -> float32 volume [N0, N1, N2]."""
    H, cg, lg, cp, lp, vp, vs, N = _prep(transient_data, camera_grid_positions,
                                         laser_grid_positions, camera_position, laser_position,
                                         volume_position, volume_size, voxels_per_side)
    wn = _vec3(wall_normal, "wall_normal")
    out = np.zeros(tuple(int(x) for x in N), np.float32)
    dshape, cshape, lshape = _shape(H), _shape(cg), _shape(lg)
    check(lib().bpx_nlos_backproject(ptr(H), ptr(dshape), H.ndim, ptr(cg), ptr(cshape), ptr(lg),
                                     ptr(lshape), ptr(cp), ptr(lp), float(t0), float(deltaT),
                                     1 if is_confocal else 0, ptr(vp), ptr(vs), ptr(N), ptr(wn),
                                     1 if assume_row_major else 0, 1 if use_cpu else 0, ptr(out)))
    return out


def prepare(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
            laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
            voxels_per_side, wall_normal=(0, 0, 0), assume_row_major=True):
    """The reference pipeline's intermediates before the device call (backproject.hpp:779-847):
    dict(chunk [rows][Tc] float32, pairs, zero, inc, chunked_t0, min_index, max_index)."""
    import ctypes as C
    from ._native import PAIR_DTYPE, NlosPrepared
    H, cg, lg, cp, lp, vp, vs, N = _prep(transient_data, camera_grid_positions,
                                         laser_grid_positions, camera_position, laser_position,
                                         volume_position, volume_size, voxels_per_side)
    wn = _vec3(wall_normal, "wall_normal")
    dshape, cshape, lshape = _shape(H), _shape(cg), _shape(lg)
    out = NlosPrepared()
    check(lib().bpx_nlos_prepare(ptr(H), ptr(dshape), H.ndim, ptr(cg), ptr(cshape), ptr(lg),
                                 ptr(lshape), ptr(cp), ptr(lp), float(t0), float(deltaT),
                                 1 if is_confocal else 0, ptr(vp), ptr(vs), ptr(N), ptr(wn),
                                 1 if assume_row_major else 0, C.byref(out)))
    try:
        rows, Tc = int(out.rows), int(out.Tc)
        chunk = np.ctypeslib.as_array(out.chunk, shape=(rows * Tc,)).reshape(rows, Tc).copy()
        praw = np.ctypeslib.as_array(C.cast(out.pairs, C.POINTER(C.c_uint8)),
                                     shape=(int(out.num_pairs) * 32,)).copy()
        pairs = praw.view(PAIR_DTYPE)
        return dict(chunk=chunk, pairs=pairs, zero=np.array(out.zero[:], np.float32),
                    inc=np.array(out.inc[:], np.float32), chunked_t0=np.float32(out.chunked_t0),
                    min_index=np.float32(out.min_index), max_index=np.float32(out.max_index))
    finally:
        lib().bpx_nlos_prepared_free(C.byref(out))


_OUTPUTS = {"complex": 0, "magnitude": 1, "phase": 2, "magnitude_phase": 3}


def _phasor_volumes(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
                    laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
                    voxels_per_side, wavelengths, assume_row_major, use_cpu, output):
    """Volumes of bpx_nlos_phasor_output, shaped [count, ...] per the output kind."""
    if output not in _OUTPUTS:
        raise ValueError(f"output must be one of {sorted(_OUTPUTS)}, not {output!r}")
    H, cg, lg, cp, lp, vp, vs, N = _prep(transient_data, camera_grid_positions,
                                         laser_grid_positions, camera_position, laser_position,
                                         volume_position, volume_size, voxels_per_side)
    wl = np.ascontiguousarray(np.asarray(wavelengths, dtype=np.float32).reshape(-1))
    shape = tuple(int(x) for x in N)
    if output == "complex":
        out = np.zeros((len(wl),) + shape, np.complex64)
    elif output == "magnitude_phase":
        out = np.zeros((len(wl), 2) + shape, np.float32)
    else:
        out = np.zeros((len(wl),) + shape, np.float32)
    dshape, cshape, lshape = _shape(H), _shape(cg), _shape(lg)
    check(lib().bpx_nlos_phasor_output(ptr(H), ptr(dshape), H.ndim, ptr(cg), ptr(cshape), ptr(lg),
                                       ptr(lshape), ptr(cp), ptr(lp), float(t0), float(deltaT),
                                       1 if is_confocal else 0, ptr(vp), ptr(vs), ptr(N), ptr(wl),
                                       len(wl), 1 if assume_row_major else 0, 1 if use_cpu else 0,
                                       _OUTPUTS[output], ptr(out)))
    return out


def phasor(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
           laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
           voxels_per_side, wavelength, assume_row_major=False, use_cpu=False, *, output="complex"):
    """bp::phasor_reconstruction (backproject.hpp:932-1039) -> complex64 volume (or, with
    output=, its magnitude, phase or both as float32 volumes)."""
    if output == "complex":
        H, cg, lg, cp, lp, vp, vs, N = _prep(transient_data, camera_grid_positions,
                                             laser_grid_positions, camera_position, laser_position,
                                             volume_position, volume_size, voxels_per_side)
        out = np.zeros(tuple(int(x) for x in N), np.complex64)
        dshape, cshape, lshape = _shape(H), _shape(cg), _shape(lg)
        check(lib().bpx_nlos_phasor(ptr(H), ptr(dshape), H.ndim, ptr(cg), ptr(cshape), ptr(lg),
                                    ptr(lshape), ptr(cp), ptr(lp), float(t0), float(deltaT),
                                    1 if is_confocal else 0, ptr(vp), ptr(vs), ptr(N),
                                    float(wavelength), 1 if assume_row_major else 0,
                                    1 if use_cpu else 0, ptr(out)))
        return out
    v = _phasor_volumes(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
                        laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
                        voxels_per_side, [wavelength], assume_row_major, use_cpu, output)[0]
    return (v[0], v[1]) if output == "magnitude_phase" else v


def phasor_multi(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
                 laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
                 voxels_per_side, wavelengths, assume_row_major=False, use_cpu=False, *, output="complex"):
    """phasor() at several wavelengths of one capture -> complex64 [len(wavelengths), N0, N1, N2]
    (an extension with no counterpart in the reference module; volume i equals
    phasor(..., wavelengths[i], ...)).  The capture is copied to the GPU and cropped once.
    output= as for phasor(): float32 [len, N0, N1, N2], or [len, 2, N0, N1, N2] for
    "magnitude_phase" (magnitude first)."""
    return _phasor_volumes(transient_data, camera_grid_positions, laser_grid_positions, camera_position,
                           laser_position, t0, deltaT, is_confocal, volume_position, volume_size,
                           voxels_per_side, wavelengths, assume_row_major, use_cpu, output)
