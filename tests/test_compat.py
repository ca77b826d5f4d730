"""The reference-signature C++ ABI (include/backproject_cuda_compat.hpp): a C++ caller written
against the reference's declarations compiles and links against libbpx.so (CPU), and its
result equals the oracle's (GPU)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "compat", "compat_call.cpp")
EXE = os.path.join(ROOT, "tests", "compat", "compat_call")


@pytest.fixture(scope="module")
def exe(bpx):
    lib = os.path.join(ROOT, "backprojector_amd")
    cmd = ["g++", "-std=c++17", "-O1", SRC, "-I", os.path.join(ROOT, "include"), "-L", lib,
           "-lbpx", f"-Wl,-rpath,{lib}", "-o", EXE]
    subprocess.run(cmd, check=True)
    return EXE


def test_compat_links(exe):
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_compat_call_matches_oracle(exe, gpu, oracle):
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    vol = np.array([float(v) for v in out], np.float64).reshape(6, 5, 4)
    side, T = 8, 1500
    pr = np.zeros(side * side, gpu.PAIR_DTYPE)
    for x in range(side):
        for y in range(side):
            wx = np.float32(-1.0 + 2.0 * x / (side - 1))
            wy = np.float32(-1.0 + 2.0 * y / (side - 1))
            wx = np.float32(np.float32(-1.0) + np.float32(2.0 * x) / np.float32(side - 1))
            wy = np.float32(np.float32(-1.0) + np.float32(2.0 * y) / np.float32(side - 1))
            d = np.float32(np.sqrt(np.float64(wx) ** 2 + np.float64(np.float32(wy + np.float32(1))) ** 2 + 0.25))
            pr[x * side + y] = ((wx, wy, 0), (wx, wy, 0), d, d)
    b = np.arange(T)[None, :]
    p = np.arange(side * side)[:, None]
    H = ((7 * b + 13 * p) % 61 + 1).astype(np.float32)
    zero = np.array([-0.5, -0.5, 0.3], np.float32)
    inc = np.array([0.2, 0, 0, 0, 0.25, 0, 0, 0, np.float32(1.0) / np.float32(3.0)], np.float32)
    ref = oracle.a1_volume_f64acc(H, pr, (6, 5, 4), zero, inc, 2.0, 0.0025)
    assert np.array_equal(vol, ref)
