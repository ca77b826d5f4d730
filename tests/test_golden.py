"""Golden fixtures (tests/golden/oracle_vectors.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit for bit (a regression pin on the oracle).
GPU: the HIP path matches the fixtures -- bins and exactly-accumulated volumes to fp32
summation-order tolerance, the phasor filter bit for bit.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
f32 = np.float32


@pytest.fixture(scope="module")
def g():
    z = np.load(os.path.join(HERE, "golden", "oracle_vectors.npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    from backprojector_amd import PAIR_DTYPE
    d["conf_pairs"] = d["conf_pairs"].view(PAIR_DTYPE)
    d["exh_pairs"] = d["exh_pairs"].view(PAIR_DTYPE)
    return d


def test_oracle_reproduces_golden(oracle, g):
    pr, N, zero, inc = g["conf_pairs"], g["conf_N"], g["conf_zero"], g["conf_inc"]
    H, t0, dT, T = g["conf_H"], float(g["conf_t0"]), float(g["conf_dT"]), int(g["conf_T"])
    assert np.array_equal(oracle.a1_volume(H, pr, N, zero, inc, t0, dT), g["conf_a1_volume"])
    assert np.array_equal(oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, dT), g["conf_a1_f64acc"])
    assert np.array_equal(oracle.fp64_volume(H, pr, N, zero, inc, t0, dT), g["conf_fp64"])
    vp, vs = np.array([0, 0, 0.8], f32), np.array([1, 1, 1], f32)
    assert np.array_equal(oracle.a2_volume(H, pr, N, vp, vs, t0, dT), g["conf_a2_volume"])
    b, r = oracle.a1_bins(pr, g["q_vox"], g["q_pair"], zero, inc, t0, dT, T)
    assert np.array_equal(b, g["q_bins"]) and np.array_equal(r, g["q_rounded"])
    ep = g["exh_pairs"]
    assert np.array_equal(oracle.a1_volume_f64acc(g["exh_H"], ep, g["exh_N"], g["exh_zero"],
                                                  g["exh_inc"], float(g["exh_t0"]),
                                                  float(g["exh_dT"])), g["exh_a1_f64acc"])
    for wl in (0.004, 0.005, 0.04):
        c, s = oracle.phasor_taps(wl, 0.001)
        assert np.array_equal((c + 1j * s).astype(np.complex64), g[f"taps_{wl}"])
    assert np.array_equal(oracle.phasor_filter(g["phasor_in"], 0.005, 0.001), g["phasor_out_0.005"])


def test_golden_bins_cover_ties_and_ranges(g):
    b = g["q_bins"]
    assert (b == -1).any() and (b >= 0).mean() > 0.5


@pytest.mark.gpu
def test_gpu_matches_golden(gpu, g):
    pr, N, zero, inc = g["conf_pairs"], g["conf_N"], g["conf_zero"], g["conf_inc"]
    H, t0, dT = g["conf_H"], float(g["conf_t0"]), float(g["conf_dT"])
    vol = gpu.backprojection(H, pr, N, zero, inc, t0, dT)
    ref = g["conf_a1_f64acc"]
    # integer samples: every sum is an exact integer, so equality is required
    assert np.array_equal(vol, ref.astype(f32))
    vp, vs = np.array([0, 0, 0.8], f32), np.array([1, 1, 1], f32)
    w = gpu.window_backprojection(H, pr, N, vp, vs, t0, dT)
    assert np.array_equal(w, g["conf_a2_volume"])
    ve = gpu.backprojection(g["exh_H"], g["exh_pairs"], g["exh_N"], g["exh_zero"], g["exh_inc"],
                            float(g["exh_t0"]), float(g["exh_dT"]))
    assert np.array_equal(ve, g["exh_a1_f64acc"].astype(f32))
    got = gpu.phasor_pulse(g["phasor_in"], 0.005, 0.001)
    assert np.array_equal(got, g["phasor_out_0.005"])
