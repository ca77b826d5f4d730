"""The per-pair bin windows over which a confocal plan stages its fused rows (the library's own
computation, bpx_row_windows) against the oracle's bins (or_a1_bins, backproject_cuda.cu:144-162
for a confocal pair): for every sampled voxel and every pair, the element a gather reads (the
oracle's raw bin + 1, clamped to the row's ends) lies at least 3 elements inside the pair's
window -- the kernels also read the neighbouring candidate and a 16-bit element's second sample
-- or the window reaches that end of the row.  Host only: no device call."""
import numpy as np
import pytest

from backprojector_amd import scenes

f32 = np.float32


def pairs_on_wall(side, jitter, seed, legs=None):
    cg = scenes.wall_grid(side)
    cg = (cg + np.random.default_rng(seed).uniform(-jitter, jitter, cg.shape).astype(f32) * [1, 1, 0]).astype(f32)
    pr = scenes.wall_pairs(cg, cg, True)
    if legs is not None:
        pr["laser_wall"] = pr["camera_wall"] = np.asarray(legs, f32)[np.arange(len(pr)) % len(legs)]
    return pr


def box(N, pos, size):
    pos, size = np.asarray(pos, f32), np.asarray(size, f32)
    return pos - size / f32(2), np.diag([size[k] / f32(N[k] - 1) for k in range(3)]).astype(f32).ravel()


def rotated(N, zero, steps, angle):
    c, s = np.cos(angle), np.sin(angle)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    return np.asarray(zero, f32), (R @ np.diag(steps)).T.astype(f32).ravel()


CASES = {
    # name: (pairs, N, zero, inc, t0, dT, Tc)
    "bins past the row end": (pairs_on_wall(16, 0.01, 16), (20, 24, 28)) + box((20, 24, 28), (0, 0, 0.8), (1, 1, 1))
                             + (2.0, 0.001, 2048),
    "bins before and past the row": (pairs_on_wall(15, 0.01, 11), (33, 35, 31))
                                    + box((33, 35, 31), (0.0, 0.05, 0.45), (0.25, 0.25, 0.25)) + (3.3, 0.001, 2000),
    "rotated basis": (pairs_on_wall(12, 0.01, 3), (17, 19, 23)) + rotated((17, 19, 23), (-0.3, -0.2, 0.5),
                                                                          (0.03, 0.025, 0.02), 0.4) + (1.9, 0.0013, 2500),
    "wall points inside the box": (pairs_on_wall(10, 0.02, 5), (21, 21, 21))
                                  + box((21, 21, 21), (0, 0, 0.0), (1.2, 1.2, 0.4)) + (0.0, 2.0 ** -9, 2600),
    "windows wholly before or past the row": (pairs_on_wall(12, 0.01, 21, legs=(0.1, 2.2, 1.5)), (21, 19, 23))
                                             + box((21, 19, 23), (0, 0, 0.6), (0.3, 0.3, 0.3)) + (4.0, 0.001, 1200),
}


@pytest.mark.parametrize("name", list(CASES))
def test_row_windows_hold_every_oracle_bin(bpx, oracle, name):
    pr, N, zero, inc, t0, dT, Tc = CASES[name]
    lde = -(-(Tc + 2) // 16) * 16
    w = bpx.row_windows(pr, N, zero, inc, t0, dT, Tc).astype(np.int64)
    assert w.shape == (len(pr), 2)
    assert np.all(w[:, 0] < w[:, 1]) and np.all(w[:, 1] <= lde)
    rng = np.random.default_rng(0)
    corners = np.array([[i, j, k] for i in (0, N[0] - 1) for j in (0, N[1] - 1) for k in (0, N[2] - 1)])
    vox = np.concatenate([corners, np.stack([rng.integers(0, N[k], 400) for k in range(3)], 1)]).astype(np.uint32)
    pi = np.repeat(np.arange(len(pr)), len(vox)).astype(np.uint32)
    _, raw = oracle.a1_bins(pr, np.tile(vox, (len(pr), 1)), pi, zero, inc, t0, dT, Tc)
    e = np.clip(raw.astype(np.int64) + 1, 0, Tc + 1)   # the element of the bin, clamped to the row
    lo, hi = w[pi, 0], w[pi, 1]
    assert np.all((lo == 0) | (e - lo >= 3)), name
    assert np.all((hi == lde) | (hi - 1 - e >= 3)), name
    # the windows are tight enough to matter: well below the whole row on average
    assert (w[:, 1] - w[:, 0]).mean() < 0.8 * lde, name


@pytest.mark.parametrize("index", [2, 7])
def test_row_windows_hold_the_batch_workload_bins(bpx, oracle, index):
    """Captures of the 64-capture uneven batch (SeedSequence(1004)), prepared by the reference
    pipeline (time limits, chunk, chunked t0): every sampled voxel-pair bin inside its window."""
    from backprojector_amd import proposals
    cap, N = scenes.batch_captures(index + 1, 1004)[index]
    chunk, pr, Nv, zero, inc, t0, dT = proposals.plan_args(cap, N)
    Tc = np.asarray(chunk).shape[1]
    zero, inc = np.asarray(zero, f32), np.asarray(inc, f32)
    lde = -(-(Tc + 2) // 16) * 16
    w = bpx.row_windows(pr, Nv, zero, inc, t0, dT, Tc).astype(np.int64)
    rng = np.random.default_rng(index)
    vox = np.stack([rng.integers(0, Nv[k], 300) for k in range(3)], 1).astype(np.uint32)
    pi = np.repeat(np.arange(len(pr)), len(vox)).astype(np.uint32)
    _, raw = oracle.a1_bins(pr, np.tile(vox, (len(pr), 1)), pi, zero, inc, t0, dT, Tc)
    e = np.clip(raw.astype(np.int64) + 1, 0, Tc + 1)
    lo, hi = w[pi, 0], w[pi, 1]
    assert np.all((lo == 0) | (e - lo >= 3)) and np.all((hi == lde) | (hi - 1 - e >= 3))
    assert (w[:, 1] - w[:, 0]).mean() < 0.6 * lde   # ~42 % of the row on this workload


def test_row_windows_rejects_invalid_arguments(bpx):
    """Empty grid axes, a non-positive bin width and malformed vectors are argument errors."""
    pr, N, zero, inc, t0, dT, Tc = CASES["bins past the row end"]
    with pytest.raises(bpx.InvalidArgument):
        bpx.row_windows(pr, (0, N[1], N[2]), zero, inc, t0, dT, Tc)
    with pytest.raises(bpx.InvalidArgument):
        bpx.row_windows(pr, N, zero, inc, t0, 0.0, Tc)
    with pytest.raises(ValueError):
        bpx.row_windows(pr, N, zero[:2], inc, t0, dT, Tc)
    assert bpx.row_windows(pr[:0], N, zero, inc, t0, dT, Tc).shape == (0, 2)
