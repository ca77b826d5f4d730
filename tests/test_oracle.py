"""Pins the CPU oracle (oracle/bp_oracle.c) with checks the mathematics fixes.

The reference ships no tests, fixtures or golden vectors for this path, and its sources cannot
be built here (empty xtensor submodules, CUDA-only GPU path).  Each check below is chosen so
that a plausible restatement mistake fails it: a dropped term (camera wall distance), a wrong
sign (t0 added), half-to-even instead of half-away rounding, a transposed pair index, bin T-1
included in the CPU window, bin 0 read by the phasor filter, the wall-normal basis sign.
"""
import numpy as np
import pytest

f32 = np.float32


def _pair(oracle, lp, cp, lw, cw):
    return oracle.make_pairs(np.array([cp], f32), np.array([lp], f32), np.array([cw], f32),
                             np.array([lw], f32))


# ----------------------------------------------------------------- A.1 bins (GPU formula)
def test_a1_exact_path_known_answer(oracle):
    # voxel straight above the wall point at h = 0.75: d1 = d2 = 0.75 exactly; walls 1.25
    # and 0.5 -> total = ((1.25 + 0.75) + 0.75) + 0.5 = 3.25; deltaT = 2^-10 divides exactly.
    pr = _pair(oracle, [0.25, -0.5, 0.0], [0.25, -0.5, 0.0], 1.25, 0.5)
    zero = np.array([0.25, -0.5, 0.75], f32)
    inc = np.eye(3, dtype=f32).ravel() * f32(0.125)
    bins, rnd = oracle.a1_bins(pr, [[0, 0, 0]], [0], zero, inc, 0.0, 2.0 ** -10, 8192)
    assert rnd[0] == 3.25 * 1024 and bins[0] == 3328
    # t0 shifts the bin down (a wrong sign would move it up)
    bins, _ = oracle.a1_bins(pr, [[0, 0, 0]], [0], zero, inc, 0.25, 2.0 ** -10, 8192)
    assert bins[0] == 3072
    # the camera wall distance counts (dropping it gives 2.75 * 1024 = 2816)
    pr2 = _pair(oracle, [0.25, -0.5, 0.0], [0.25, -0.5, 0.0], 1.25, 0.0)
    bins, _ = oracle.a1_bins(pr2, [[0, 0, 0]], [0], zero, inc, 0.0, 2.0 ** -10, 8192)
    assert bins[0] == 2816
    # out of range above T-1 -> no tap
    bins, _ = oracle.a1_bins(pr, [[0, 0, 0]], [0], zero, inc, 0.0, 2.0 ** -10, 3328)
    assert bins[0] == -1


def test_a1_pythagorean_voxels(oracle):
    # (3,4,0)-(0,0,12) style exact distances: wall point at origin, voxel at (3k,4k,12k)*s
    # gives d = 13ks exactly for small integers and s = 2^-6.
    s = 2.0 ** -6
    pr = _pair(oracle, [0, 0, 0], [0, 0, 0], 0.0, 0.0)
    zero = np.zeros(3, f32)
    inc = np.diag([3 * s, 4 * s, 12 * s]).astype(f32).ravel()
    vox = [[k, k, k] for k in range(1, 6)]
    bins, rnd = oracle.a1_bins(pr, vox, [0] * 5, zero, inc, 0.0, s, 4096)
    assert list(rnd) == [26.0 * k for k in range(1, 6)]


def test_a1_round_half_away_from_zero(oracle):
    # total - t0 = 100.5 * 2^-8 exactly -> q = 100.5 -> round half away = 101 (rint: 100)
    pr = _pair(oracle, [0, 0, 0], [0, 0, 0], 1.25, 1.25)   # total = 1.25+0.75+0.75+1.25 = 4
    zero = np.array([0, 0, 0.75], f32)
    inc = np.eye(3, dtype=f32).ravel()
    t0 = f32(4.0) - f32(100.5 / 256)
    bins, rnd = oracle.a1_bins(pr, [[0, 0, 0]], [0], zero, inc, float(t0), 2.0 ** -8, 4096)
    assert rnd[0] == 101.0 and bins[0] == 101
    t0b = f32(4.0) - f32(101.5 / 256)
    bins, rnd = oracle.a1_bins(pr, [[0, 0, 0]], [0], zero, inc, float(t0b), 2.0 ** -8, 4096)
    assert rnd[0] == 102.0            # half-to-even would give 102 here too; both ties checked
    # negative tie: q = -0.5 -> -1 -> no tap (half-to-even gives -0 -> bin 0)
    t0c = f32(4.0) + f32(0.5 / 256)
    bins, rnd = oracle.a1_bins(pr, [[0, 0, 0]], [0], zero, inc, float(t0c), 2.0 ** -8, 4096)
    assert rnd[0] == -1.0 and bins[0] == -1


def test_confocal_bin_of_squared_distance_matches_voxel_bin(oracle):
    # a confocal pair's bin from the squared distance s equals the voxel formula's bin
    rng = np.random.default_rng(3)
    for _ in range(200):
        w = rng.uniform(-1, 1, 3).astype(f32)
        w[2] = 0
        lw, cw = f32(rng.uniform(0.5, 2)), f32(rng.uniform(0.5, 2))
        pr = _pair(oracle, w, w, lw, cw)
        zero = rng.uniform(-0.5, 0.5, 3).astype(f32)
        zero[2] = f32(rng.uniform(0.3, 1.3))
        inc = (np.eye(3) * 0.01).astype(f32).ravel()
        t0, dT = float(rng.uniform(0, 2)), float(rng.uniform(0.0007, 0.003))
        _, rnd = oracle.a1_bins(pr, [[3, 5, 7]], [0], zero, inc, t0, dT, 1 << 20)
        v = oracle.a1_voxel_position(zero, inc, 3, 5, 7)
        a, b, c = w - v
        s = (a * a + b * b) + c * c
        assert oracle.a1_confocal_bin(lw, cw, s, t0, dT) == rnd[0]


def test_confocal_thresholds_are_the_bin_switch_points(oracle):
    # S[p][j] is the smallest s with bin >= j: bin(S) >= j and bin(previous float) < j
    rng = np.random.default_rng(5)
    lw = rng.uniform(0.8, 1.6, 6).astype(f32)
    cw = rng.uniform(0.8, 1.6, 6).astype(f32)
    t0, dT, n = 2.25, 0.0011, 1200
    S = oracle.confocal_thresholds(lw, cw, n, t0, dT)
    assert np.all(np.diff(S, axis=1) >= 0)
    for p in range(len(lw)):
        for j in range(0, n, 7):
            s = S[p, j]
            assert oracle.a1_confocal_bin(lw[p], cw[p], s, t0, dT) >= j
            if s > 0:
                below = np.nextafter(s, f32(0))
                assert oracle.a1_confocal_bin(lw[p], cw[p], below, t0, dT) < j
    # bins below the shortest path start at s = 0; the bin never decreases in s
    first = oracle.a1_confocal_bin(lw[0], cw[0], 0.0, t0, dT)
    assert first > 0 and np.all(S[0, :int(first) + 1] == 0)
    ss = np.sort(rng.uniform(0, 4, 400).astype(f32))
    b = [oracle.a1_confocal_bin(lw[1], cw[1], x, t0, dT) for x in ss]
    assert np.all(np.diff(b) >= 0)
    # a known answer: total = 2 + 2 sqrt(s) with dT = 2^-8 and t0 = 0 reaches bin j >= 513 at
    # 2 sqrt(s) = (j - 0.5) / 256 - 2 exactly (ties round away, every step exact)
    S2 = oracle.confocal_thresholds([1.0], [1.0], 600, 0.0, 2.0 ** -8)
    d = ((560 - 0.5) / 256 - 2) / 2
    assert oracle.a1_confocal_bin(1.0, 1.0, f32(d) * f32(d), 0.0, 2.0 ** -8) == 560
    assert S2[0, 560] <= f32(d) * f32(d)
    assert oracle.a1_confocal_bin(1.0, 1.0, np.nextafter(S2[0, 560], f32(0)), 0.0, 2.0 ** -8) == 559


def test_a1_voxel_position_order(oracle):
    # v[k] = ((zero + inc[k] x) + inc[3+k] y) + inc[6+k] z: rows of inc are index steps
    inc = np.array([1, 10, 100, 2, 20, 200, 3, 30, 300], f32) * f32(0.5)
    v = oracle.a1_voxel_position([1, 2, 3], inc, 1, 2, 3)
    assert np.array_equal(v, np.array([1 + 0.5 + 2 + 4.5, 2 + 5 + 20 + 45, 3 + 50 + 200 + 450], f32))


def test_a1_bins_agree_with_fp64_away_from_ties(oracle):
    rng = np.random.default_rng(7)
    P = 64
    wall = rng.uniform(-1, 1, size=(P, 3)).astype(f32)
    wall[:, 2] = 0
    lw = np.linalg.norm(wall - [0, -1, 0.5], axis=1).astype(f32)
    pr = oracle.make_pairs(wall, wall, lw, lw)
    zero = np.array([-0.5, -0.5, 0.3], f32)
    inc = np.diag([1 / 15, 1 / 15, 1 / 15]).astype(f32).ravel()
    vox = rng.integers(0, 16, size=(4000, 3))
    pidx = rng.integers(0, P, size=4000)
    t0, dT = 0.0, 0.001
    bins, rnd = oracle.a1_bins(pr, vox, pidx, zero, inc, t0, dT, 8192)
    v = zero.astype(np.float64) + vox @ inc.reshape(3, 3).astype(np.float64)
    d = np.linalg.norm(v - wall[pidx].astype(np.float64), axis=1)
    q = (lw[pidx] + 2 * d + lw[pidx] - t0) / dT
    far = np.abs(q - np.floor(q) - 0.5) > 0.01
    assert far.mean() > 0.95
    assert np.array_equal(rnd[far], np.floor(q[far] + 0.5))


# ----------------------------------------------------------------- A.1 volumes
def _random_capture(rng, P=32, T=512, N=(6, 5, 7), confocal=True):
    wall = rng.uniform(-1, 1, size=(P, 3)).astype(f32)
    wall[:, 2] = 0
    cam = wall if confocal else rng.uniform(-1, 1, size=(P, 3)).astype(f32) * [1, 1, 0]
    lw = np.linalg.norm(wall - [0, -1, 0.5], axis=1).astype(f32)
    cw = np.linalg.norm(cam - [0, -1, 0.5], axis=1).astype(f32)
    pr = np.zeros(P, dtype=[("camera_point", f32, 3), ("laser_point", f32, 3),
                            ("camera_wall", f32), ("laser_wall", f32)])
    pr["camera_point"], pr["laser_point"], pr["camera_wall"], pr["laser_wall"] = cam, wall, cw, lw
    H = rng.poisson(3.0, size=(P, T)).astype(f32)
    zero = np.array([-0.5, -0.5, 0.3], f32)
    inc = np.diag([1 / (N[0] - 1), 1 / (N[1] - 1), 1 / (N[2] - 1)]).astype(f32).ravel()
    return H, pr, np.array(N, np.uint32), zero, inc


def test_a1_volume_zero_and_linearity(oracle):
    rng = np.random.default_rng(1)
    H, pr, N, zero, inc = _random_capture(rng)
    t0 = 2.0
    assert not oracle.a1_volume(np.zeros_like(H), pr, N, zero, inc, t0, 0.001).any()
    H2 = rng.poisson(2.0, size=H.shape).astype(f32)
    a = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, 0.001)
    b = oracle.a1_volume_f64acc(H2, pr, N, zero, inc, t0, 0.001)
    c = oracle.a1_volume_f64acc(3 * H + H2, pr, N, zero, inc, t0, 0.001)
    assert np.allclose(c, 3 * a + b, rtol=0, atol=1e-9)
    assert a.sum() > 0


def test_a1_volume_is_sum_of_bins(oracle):
    # H[p][b] = b + 1: the volume is the sum over pairs of (bin + 1), all integers (exact)
    rng = np.random.default_rng(2)
    H, pr, N, zero, inc = _random_capture(rng, P=16, T=3000)
    H = np.tile(np.arange(1, 3001, dtype=f32), (16, 1))
    vol = oracle.a1_volume(H, pr, N, zero, inc, 0.0, 0.001)
    vox = np.stack(np.meshgrid(*[np.arange(n) for n in N], indexing="ij"), -1).reshape(-1, 3)
    expect = np.zeros(len(vox))
    for p in range(16):
        b, _ = oracle.a1_bins(pr, vox, np.full(len(vox), p), zero, inc, 0.0, 0.001, 3000)
        expect += np.where(b >= 0, b + 1, 0)
    assert np.array_equal(vol.ravel(), expect.astype(f32))


def test_a1_reference_pair_truncation(oracle):
    # P = 300 -> B = min(256, P) = 256 -> pairs 256..299 are dropped by the reference launcher
    rng = np.random.default_rng(3)
    H, pr, N, zero, inc = _random_capture(rng, P=300, T=256, N=(3, 3, 3))
    t0 = 2.1
    full = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, 0.001)
    head = oracle.a1_volume_f64acc(H[:256], pr[:256], N, zero, inc, t0, 0.001)
    ref = oracle.a1_volume(H, pr, N, zero, inc, t0, 0.001)
    assert not np.allclose(full, head)
    assert np.allclose(ref, head, rtol=1e-6)


def test_fp64_oracle_close_to_fp32_formula(oracle):
    rng = np.random.default_rng(4)
    H, pr, N, zero, inc = _random_capture(rng, P=64, T=1024, N=(8, 8, 8))
    t0 = 2.0
    a = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, 0.001)
    b = oracle.fp64_volume(H, pr, N, zero, inc, t0, 0.001)
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel < 1e-2          # only near-tie bins differ
    assert np.mean(np.isclose(a, b, rtol=1e-12)) > 0.5


# ----------------------------------------------------------------- A.2 (CPU formula)
def test_a2_window_known_answer(oracle):
    # one voxel straight above the wall point at h = 0.75 (exact double distances), grid
    # spacing size/N = 0.25 per axis -> window w = sqrt(3 * 0.0625) (float), walls 1.25
    pr = _pair(oracle, [0, 0, 0], [0, 0, 0], 1.25, 1.25)
    H = np.ones((1, 8192), f32)
    vp = np.array([0.25, 0.25, 0.875], f32)     # corner = pos - size/2 = (0, 0, 0.75)
    vs = np.array([0.5, 0.5, 0.25], f32)        # N = (2, 2, 1) -> voxel size (0.25,0.25,0.25)
    zero, vsz, w = oracle.a2_grid(vp, vs, [2, 2, 1])
    assert np.array_equal(zero, [0, 0, 0.75]) and np.array_equal(vsz, [0.25, 0.25, 0.25])
    assert w == f32(np.sqrt(3 * 0.0625))
    dT = 2.0 ** -10
    vol = oracle.a2_volume(H, pr, [2, 2, 1], vp, vs, 0.0, dT)
    total = f32(4.0)
    first = int(np.round((total - w / 2) / f32(dT)))
    last = int(np.round((total + w / 2) / f32(dT)))
    assert vol[0, 0, 0] == last - first          # bins [first, last)
    # bin T-1 is never summed: cap T just above the window start
    T = first + 3
    vol = oracle.a2_volume(np.ones((1, T), f32), pr, [2, 2, 1], vp, vs, 0.0, dT)
    assert vol[0, 0, 0] == 2                     # [first, T-1) -> 2 bins, not 3


def test_a2_linearity_and_fp64(oracle):
    rng = np.random.default_rng(5)
    H, pr, N, zero, inc = _random_capture(rng, P=48, T=1200, N=(6, 6, 6))
    vp, vs = np.array([0, 0, 0.8], f32), np.array([1, 1, 1], f32)
    a = oracle.a2_volume(H, pr, N, vp, vs, 2.0, 0.001)
    b = oracle.a2_volume(2 * H, pr, N, vp, vs, 2.0, 0.001)
    assert np.allclose(b, 2 * a, rtol=1e-6)
    c = oracle.a2_volume_fp64(H, pr, N, vp, vs, 2.0, 0.001)
    assert np.linalg.norm(a - c) / np.linalg.norm(c) < 2e-2
    assert a.sum() > 0


# ----------------------------------------------------------------- host pipeline
def test_precompute_pairs_exhaustive_index_order(oracle):
    lg = np.arange(2 * 3 * 3, dtype=f32).reshape(2, 3, 3)
    cg = 100 + np.arange(4 * 2 * 3, dtype=f32).reshape(4, 2, 3)
    lp, cp = np.array([1, 2, 3], f32), np.array([-1, -2, -3], f32)
    pr = oracle.precompute_pairs(lp, lg, cp, cg, confocal=False)
    assert len(pr) == 2 * 3 * 4 * 2
    for lx in range(2):
        for ly in range(3):
            for cx in range(4):
                for cy in range(2):
                    p = ((lx * 3 + ly) * 4 + cx) * 2 + cy
                    assert np.array_equal(pr[p]["laser_point"], lg[lx, ly])
                    assert np.array_equal(pr[p]["camera_point"], cg[cx, cy])
                    assert pr[p]["laser_wall"] == f32(np.linalg.norm(lg[lx, ly] - lp.astype(np.float64)))
                    assert pr[p]["camera_wall"] == f32(np.linalg.norm(cg[cx, cy] - cp.astype(np.float64)))


def test_precompute_pairs_confocal_uses_camera_side(oracle):
    cg = np.random.default_rng(0).uniform(-1, 1, size=(3, 4, 3)).astype(f32)
    lg_a = np.zeros_like(cg)
    lg_b = np.ones_like(cg)
    cp = np.array([0, -1, 0.5], f32)
    a = oracle.precompute_pairs([9, 9, 9], lg_a, cp, cg, confocal=True)
    b = oracle.precompute_pairs([-9, 1, 2], lg_b, cp, cg, confocal=True)
    assert a.tobytes() == b.tobytes()
    assert np.array_equal(a["laser_point"].reshape(3, 4, 3), cg)
    assert np.array_equal(a["laser_wall"], a["camera_wall"])


def test_basis_known_answers(oracle):
    zero, inc = oracle.basis([0, 0, 0.8], [1, 1, 1], [5, 3, 9])
    assert np.array_equal(zero, np.array([-0.5, -0.5, 0.3], f32))
    assert np.array_equal(inc.reshape(3, 3), np.diag(np.array([1 / 4, 1 / 2, 1 / 8], f32)))
    # wall normal (0,-1,0): y = (0,1,0), x = y cross z = (1,0,0), z = y cross x = (0,0,-1)
    zero, inc = oracle.basis([0, 0, 0.8], [1, 1, 1], [5, 3, 9], wall_normal=[0, -1, 0])
    assert np.array_equal(inc.reshape(3, 3), np.diag(np.array([1 / 4, 1 / 2, -1 / 8], f32)))


def test_time_limits_transposed_camera_corners(oracle):
    # The reference transposes the row-major camera grid to [3][Y][X] and reads the first
    # three entries of each corner row; computed here with numpy's transpose.
    g = np.linspace(-1, 1, 5, dtype=f32)
    X, Y = np.meshgrid(g, g, indexing="ij")
    grid = np.stack([X, Y, np.full_like(X, 0.1)], -1).astype(f32)
    o = np.array([0, -1, 0.5], f32)
    vp, vs = np.array([0, 0, 0.8], f32), np.array([1, 1, 1], f32)
    mn, mx = oracle.time_limits(o, grid, o, grid, vs, vp, 0.0, 0.001)
    Tg = np.transpose(grid)                                   # [3][Y][X]
    cam = [Tg[c, y, :3] for c in (0, 2) for y in (0, 4)]
    las = [grid[x, y] for x in (0, 4) for y in (0, 4)]
    vol = [vp - vs / 2 + np.array([(m >> 2) & 1, (m >> 1) & 1, m & 1]) * vs for m in range(8)]

    def mm(A, B):
        d = [f32(np.linalg.norm(np.float64(a) - np.float64(b))) for a in A for b in B]
        return min(d), max(d)
    a, b = mm([o], las)
    c, d = mm([o], cam)
    e, f_ = mm(vol, las)
    g2, h = mm(vol, cam)
    exp_min = (((a + c) + e) + g2) / f32(0.001) - f32(100)
    exp_max = (((b + d) + f_) + h) / f32(0.001) + f32(100)
    assert mn == f32(exp_min) and mx == f32(exp_max)


def test_transient_chunk_and_t0(oracle):
    data = np.tile(np.arange(1, 51, dtype=f32), (3, 1))      # H[p][t] = t + 1
    ch = oracle.transient_chunk(data, f32(-3.5), f32(40.2))   # [floor(-3.5), ceil(40.2)) = [-4, 41)
    assert ch.shape == (3, 45)
    assert np.array_equal(ch[0, :4], np.zeros(4)) and ch[0, 4] == 1 and ch[0, 44] == 41
    ch = oracle.transient_chunk(data, f32(10.7), f32(60.1))   # [10, 61): reads 10..49, pads 11
    assert ch.shape == (3, 51) and ch[1, 0] == 11 and ch[1, 39] == 50 and not ch[1, 40:].any()
    assert oracle.chunked_t0(1.0, f32(10.7), 0.5) == f32(1.0) + f32(10.7) * f32(0.5)


# ----------------------------------------------------------------- phasor filter
@pytest.mark.parametrize("wl,pulse,cycle", [(0.004, 16, 4), (0.005, 19, 4), (0.04, 159, 39)])
def test_phasor_taps_closed_form(oracle, wl, pulse, cycle):
    c, s = oracle.phasor_taps(wl, 0.001)
    assert len(c) == pulse                      # (int)(4 * wl / dT) in float
    sigma = float(f32(4 * f32(wl)) / 6)
    k = np.arange(1, pulse + 1)
    t = 0.001 * (k - pulse / 2.0)
    g = np.exp(-t * t / (2 * sigma * sigma))
    assert np.allclose(c, np.cos(2 * np.pi * k / cycle) * g, rtol=0, atol=2e-6)
    assert np.allclose(s, np.sin(2 * np.pi * k / cycle) * g, rtol=0, atol=2e-6)


def test_phasor_filter_impulse_response(oracle):
    c, s = oracle.phasor_taps(0.005, 0.001)
    P, T = len(c), 64
    x = np.zeros((2, T), f32)
    x[0, 30] = 1.0
    x[1, 0] = 5.0                               # bin 0 is never read (idx > 0)
    y = oracle.phasor_filter(x, 0.005, 0.001)
    assert not y[1].any()
    for t in range(T):
        j = 30 - t + P // 2
        want = complex(c[j], s[j]) if 0 <= j < P else 0
        assert y[0, t] == want
