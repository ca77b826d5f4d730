"""bpx_backprojection_noalloc (no reference counterpart; for callers that cannot allocate device
memory): called through the C ABI on device buffers the test owns, its volumes must equal
bpx_backprojection's on the same capture bit for bit, on consecutive calls of different sizes
(the second smaller, so stale scratch contents would show) and on a caller stream."""
import ctypes

import numpy as np
import pytest

from backprojector_amd import scenes
from backprojector_amd._native import BpxError, InvalidArgument, check, lib, ptr

pytestmark = pytest.mark.gpu
f32 = np.float32
H2D, D2H = 1, 2
SCRATCH_BYTES = 8 << 20   # BPX_NOALLOC_SCRATCH_BYTES in include/bpx.h


class DeviceMemory:
    """hipMalloc'd buffers of the caller (the library itself allocates nothing in this call)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.held = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(int(nbytes), 4)) == 0
        self.held.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(d, a.ctypes.data, a.nbytes, H2D) == 0
        return d

    def download(self, d, shape, dtype=f32):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, d, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.held:
            assert self.hip.hipFree(p) == 0
        self.held = []


@pytest.fixture
def dev(gpu):
    m = DeviceMemory()
    yield m
    m.free()


def capture(side, T, N, seed, coded=False):
    rng = np.random.default_rng(seed)
    hidden = np.concatenate([scenes.points(rng, 3), scenes.patch(rng, size=0.2, spacing=0.01)])
    cap = scenes.render(hidden, side, T, 0.0, 4.096 / T, True, seed=seed)
    pr = scenes.wall_pairs(cap.camera_grid, cap.camera_grid, True)
    H = cap.transient.reshape(len(pr), -1).astype(f32)
    if coded:   # integer codes: every sum exact, so the oracle must match bit for bit
        b = np.arange(T)[None, :]
        p = np.arange(len(pr))[:, None]
        H = (((7 * b + 13 * p) % 61) + 1).astype(f32)
    zero = np.array([-0.5, -0.5, 0.3], f32)
    inc = np.diag([f32(1) / f32(n - 1) for n in N]).astype(f32).ravel()
    return H, pr, tuple(N), zero, inc, 0.0, 4.096 / T


def noalloc(dev, H, pr, N, zero, inc, t0, dT, stream=None):
    P, T = H.shape
    dH = dev.upload(H)
    dV = dev.alloc(4 * int(np.prod(N)))
    n = np.asarray(N, np.uint32)
    check(lib().bpx_backprojection_noalloc(dH, H.size, T, ptr(pr), P, dV, ptr(n), ptr(np.asarray(zero, f32)),
                                           ptr(np.asarray(inc, f32)), float(t0), float(dT), stream))
    return dev.download(dV, N)


def test_noalloc_equals_backprojection_on_consecutive_calls(gpu, oracle, dev, hip_streams):
    big = capture(16, 2048, (20, 26, 18), seed=11)
    small = capture(8, 512, (9, 7, 12), seed=12, coded=True)
    print(f"noalloc: first call {big[0].shape} pairs x bins -> {big[2]}", flush=True)
    v1 = noalloc(dev, *big)
    ref1 = gpu.backprojection(*big)
    assert np.abs(ref1).max() > 0 and np.array_equal(v1, ref1)
    print(f"noalloc: second call {small[0].shape} -> {small[2]} on a caller stream", flush=True)
    v2 = noalloc(dev, *small, stream=hip_streams[0])
    assert np.array_equal(v2, gpu.backprojection(*small))
    exact = oracle.a1_volume_f64acc(*small)
    assert np.array_equal(v2, exact.astype(f32))
    v1b = noalloc(dev, *big)        # back to the larger capture: nothing stale remains
    assert np.array_equal(v1b, ref1)


def test_noalloc_limits_and_errors(gpu, dev):
    H, pr, N, zero, inc, t0, dT = capture(4, 256, (4, 4, 4), seed=3)
    n = np.asarray(N, np.uint32)
    z, i = np.asarray(zero, f32), np.asarray(inc, f32)
    dV = dev.alloc(4 * 64)
    # more pair records than the fixed array holds (T = 1 keeps the device buffer small)
    many = np.resize(pr, SCRATCH_BYTES // 32 + 1)
    dH1 = dev.alloc(4 * len(many))
    with pytest.raises(BpxError, match="scratch"):
        check(lib().bpx_backprojection_noalloc(dH1, len(many), 1, ptr(many), len(many), dV, ptr(n), ptr(z), ptr(i),
                                               0.0, dT, None))
    dH = dev.upload(H)
    with pytest.raises(BpxError, match="T must be"):   # the threshold table must fit LDS
        check(lib().bpx_backprojection_noalloc(dH, 20000, 20000, ptr(pr[:1]), 1, dV, ptr(n), ptr(z), ptr(i),
                                               0.0, dT, None))
    bad = pr.copy()
    bad["laser_wall"][3] = np.nan
    with pytest.raises(InvalidArgument, match="non-finite"):
        check(lib().bpx_backprojection_noalloc(dH, H.size, H.shape[1], ptr(bad), len(bad), dV, ptr(n), ptr(z),
                                               ptr(i), t0, dT, None))
    with pytest.raises(InvalidArgument):
        check(lib().bpx_backprojection_noalloc(dH, H.size + 1, H.shape[1], ptr(pr), len(pr), dV, ptr(n), ptr(z),
                                               ptr(i), t0, dT, None))
    # an empty capture writes zeros; a valid call afterwards still matches
    dV2 = dev.upload(np.full(64, 7.0, f32))
    check(lib().bpx_backprojection_noalloc(dH, 0, H.shape[1], ptr(pr), 0, dV2, ptr(n), ptr(z), ptr(i), t0, dT, None))
    assert not dev.download(dV2, 64).any()
    assert np.array_equal(noalloc(dev, H, pr, N, zero, inc, t0, dT), gpu.backprojection(H, pr, N, zero, inc, t0, dT))
