#!/usr/bin/env python3
"""Writes tests/golden/oracle_vectors.npz: fixed inputs and the CPU oracle's outputs.

The reference holds no test vectors for this path, so these fixtures are produced by the
oracle (itself pinned by tests/test_oracle.py) and freeze its behaviour: the CPU suite
checks the oracle still reproduces them bit for bit, and the GPU suite checks the HIP path
against them.  Regenerate only on a deliberate oracle change:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

f32 = np.float32


def wall(side):
    g = np.linspace(-1, 1, side)
    X, Y = np.meshgrid(g, g, indexing="ij")
    return np.stack([X, Y, np.zeros_like(X)], -1).astype(f32)


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    o = np.array([0, -1, 0.5], f32)
    # ---- confocal 8x8 capture geometry with jittered wall points
    cg = (wall(8) + rng.uniform(-0.01, 0.01, (8, 8, 3)).astype(f32) * [1, 1, 0]).astype(f32)
    pairs = orc.precompute_pairs(o, cg, o, cg, True)
    N = np.array([9, 8, 7], np.uint32)
    zero, inc = orc.basis([0, 0, 0.8], [1, 1, 1], N)
    T, t0, dT = 1300, np.float32(2.05), np.float32(0.0025)
    H = rng.poisson(3.0, (len(pairs), T)).astype(f32)
    out.update(conf_pairs=pairs.view(np.uint8), conf_N=N, conf_zero=zero, conf_inc=inc,
               conf_T=np.uint32(T), conf_t0=t0, conf_dT=dT, conf_H=H)
    out["conf_a1_volume"] = orc.a1_volume(H, pairs, N, zero, inc, t0, dT)
    out["conf_a1_f64acc"] = orc.a1_volume_f64acc(H, pairs, N, zero, inc, t0, dT)
    out["conf_fp64"] = orc.fp64_volume(H, pairs, N, zero, inc, t0, dT)
    vp, vs = np.array([0, 0, 0.8], f32), np.array([1, 1, 1], f32)
    out["conf_a2_volume"] = orc.a2_volume(H, pairs, N, vp, vs, t0, dT)
    # ---- random bin queries (voxel, pair) incl. out-of-range
    vox = np.stack([rng.integers(0, n, 20000) for n in N], 1).astype(np.uint32)
    pid = rng.integers(0, len(pairs), 20000).astype(np.uint32)
    b, r = orc.a1_bins(pairs, vox, pid, zero, inc, t0, dT, T)
    out.update(q_vox=vox, q_pair=pid, q_bins=b, q_rounded=r)
    # ---- exhaustive 3x3 lasers x 4x4 SPADs
    lg, cg2 = wall(3), (wall(4) * f32(0.8)).astype(f32)
    ep = orc.precompute_pairs(o, lg, o, cg2, False)
    He = rng.poisson(2.0, (len(ep), 700)).astype(f32)
    Ne = np.array([6, 6, 6], np.uint32)
    ze, ie = orc.basis([0, 0, 0.8], [1, 1, 1], Ne)
    out.update(exh_pairs=ep.view(np.uint8), exh_H=He, exh_N=Ne, exh_zero=ze, exh_inc=ie,
               exh_t0=np.float32(2.3), exh_dT=np.float32(0.004))
    out["exh_a1_f64acc"] = orc.a1_volume_f64acc(He, ep, Ne, ze, ie, 2.3, 0.004)
    out["exh_a2_volume"] = orc.a2_volume(He, ep, Ne, vp, vs, 2.3, 0.004)
    # ---- phasor taps and filtered rows
    for wl in (0.004, 0.005, 0.04):
        c, s = orc.phasor_taps(wl, 0.001)
        out[f"taps_{wl}"] = (c + 1j * s).astype(np.complex64)
    x = rng.poisson(4.0, (5, 300)).astype(f32)
    out["phasor_in"] = x
    out["phasor_out_0.005"] = orc.phasor_filter(x, 0.005, 0.001)
    # ---- pipeline: time limits / chunk / chunked t0 of a 16x16 confocal capture
    g16 = wall(16)
    mn, mx = orc.time_limits(o, g16, o, g16, vs, vp, 0.0, 0.001)
    out.update(pipe_min=mn, pipe_max=mx, pipe_t0=orc.chunked_t0(0.0, mn, 0.001))
    np.savez_compressed(os.path.join(HERE, "oracle_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "oracle_vectors.npz"))


if __name__ == "__main__":
    main()
