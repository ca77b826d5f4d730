"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Bin exactness is tested at scale with integer-valued samples: H[p][b] is a small positive
integer that varies with b, so every voxel sum is an integer below 2^24 and any float
summation order is exact.  A single bin off by one changes a voxel's sum, so the GPU volume
must EQUAL the oracle's exactly-accumulated volume.  Real-valued captures are compared
within 1e-5 relative L2 and each states its error against the FP64 oracle.
"""
import numpy as np
import pytest

from backprojector_amd import scenes

pytestmark = pytest.mark.gpu
f32 = np.float32


def rel_l2(a, b):
    a = np.asarray(a, np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, a.dtype)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def coded_samples(P, T, seed=0):
    """Integer sample codes in [1, 61] varying along the bin axis and across pairs."""
    b = np.arange(T, dtype=np.int64)[None, :]
    p = np.arange(P, dtype=np.int64)[:, None]
    return (((7 * b + 13 * p + seed) % 61) + 1).astype(f32)


def grid_basis(N, pos=(0, 0, 0.8), size=(1, 1, 1)):
    pos, size = np.asarray(pos, f32), np.asarray(size, f32)
    zero = pos - size / f32(2)
    inc = np.diag([size[k] / f32(N[k] - 1) for k in range(3)]).astype(f32).ravel()
    return zero, inc


def confocal_pairs(side, jitter=0.0, seed=0):
    cg = scenes.wall_grid(side)
    if jitter:
        cg = cg + np.random.default_rng(seed).uniform(-jitter, jitter, cg.shape).astype(f32) * [1, 1, 0]
        cg = cg.astype(f32)
    return scenes.wall_pairs(cg, cg, True)


def sample_voxels(N, n, seed=0):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, N[k], n) for k in range(3)], 1).astype(np.uint32)


# ----------------------------------------------------------------- bit-exact bins
@pytest.mark.parametrize("side,N,t0", [(16, (20, 24, 28), 2.0), (32, (32, 32, 32), 2.3419),
                                       (8, (7, 5, 3), 1.75)])
def test_confocal_bins_bit_exact(gpu, oracle, side, N, t0):
    pr = confocal_pairs(side, jitter=0.01, seed=side)
    T = 2048
    H = coded_samples(len(pr), T)
    zero, inc = grid_basis(N)
    vol = gpu.backprojection(H, pr, N, zero, inc, t0, 0.001)
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, 0.001)
    assert np.all(ref == np.round(ref)) and ref.max() < 2 ** 24
    mism = np.count_nonzero(vol != ref.astype(f32))
    assert mism == 0, f"{mism} voxels differ (bin mismatch)"


def test_confocal_bins_bit_exact_all_variants_and_non_separable_basis(gpu, oracle):
    pr = confocal_pairs(12, jitter=0.02, seed=3)
    T = 2500
    H = coded_samples(len(pr), T, seed=5)
    N = (18, 17, 19)
    zero, inc = grid_basis(N)
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, 2.05, 0.001)
    for v in range(9):
        vol = gpu.backprojection(H, pr, N, zero, inc, 2.05, 0.001, variant=v)
        assert np.array_equal(vol, ref.astype(f32)), f"variant {v}"
    # a rotated basis (wall normal) makes the grid non-separable and the wall non-flat
    zero2, inc2 = oracle.basis([0, 0.1, 0.8], [1, 1, 1], N, wall_normal=[0.3, -1, 0.2])
    pr2 = pr.copy()
    pr2["laser_point"][:, 2] = np.linspace(0, 0.05, len(pr2)).astype(f32)
    pr2["camera_point"] = pr2["laser_point"]
    ref2 = oracle.a1_volume_f64acc(H, pr2, N, zero2, inc2, 2.05, 0.001)
    for v in (2, 5):
        vol2 = gpu.backprojection(H, pr2, N, zero2, inc2, 2.05, 0.001, variant=v)
        assert np.array_equal(vol2, ref2.astype(f32)), f"variant {v}"


@pytest.mark.parametrize("variant,kz", [(6, 16), (7, 8), (8, 8)])
@pytest.mark.parametrize("N,size,kahan", [((35, 33, 40), (0.25, 0.22, 0.3), 0), ((16, 16, 16), (0.1, 0.1, 0.1), 3),
                                          ((50, 17, 9), (0.25, 0.08, 0.03), 0), ((1, 1, 1), (0.1, 0.1, 0.1), 0)])
def test_lds_staged_kernel_bit_exact(gpu, oracle, N, size, kahan, variant, kz):
    """The LDS-staged kernels (16 x 16 x 16 / 16 x 16 x 8 tiles with 256 lanes, 16^3 tiles with
    512 lanes; each pair's bin window copied to LDS) at a
    voxel pitch whose windows fit: bins exact against the oracle, and volumes identical to the
    texture-path kernel (same element per update, same order, same Kahan blocks).  Partial
    tiles in every axis, a one-voxel volume, and bins below 0 and above Tc."""
    pr = confocal_pairs(16, jitter=0.01, seed=7)          # flat wall
    T = 3000
    H = coded_samples(len(pr), T, seed=2)
    zero, inc = grid_basis(N, pos=(0.05, -0.02, 0.7), size=size)
    if N == (1, 1, 1):
        inc = np.diag([0.004, 0.004, 0.004]).astype(f32).ravel()
    t0 = 1.2
    with gpu.Plan(H, pr, N, zero, inc, t0, 0.001, kahan_block=kahan, variant=variant) as plan:
        assert plan.info()["lane_voxels"] == kz   # the LDS-staged kernel runs
        plan.execute()
        v6 = plan.volume()
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, 0.001)
    assert np.array_equal(v6, ref.astype(f32))
    Hr = np.random.default_rng(1).gamma(0.5, 2.0, H.shape).astype(f32)
    a = gpu.backprojection(Hr, pr, N, zero, inc, t0, 0.001, kahan_block=kahan, variant=variant)
    b = gpu.backprojection(Hr, pr, N, zero, inc, t0, 0.001, kahan_block=kahan, variant=5)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("variant", [6, 7, 8])
def test_exact_fp16_rows_for_integer_samples(gpu, oracle, variant):
    """An fp32 plan on the LDS-staged kernel with 8 voxels per lane (variants 7, 8) stages
    integer samples within fp16's exact range (|x| <= 2048) as fp16 rows: row_storage says so,
    and the volume equals the fp32 rows' (keep_f32_rows) and the oracle bit for bit; the
    16-per-lane shape (variant 6) keeps fp32 rows.  One sample that fp16 does not hold exactly
    (a fraction, or beyond 2048, or NaN) keeps the whole execution on fp32 rows."""
    pr = confocal_pairs(16, jitter=0.01, seed=7)
    T, N, t0 = 3000, (35, 33, 40), 1.2
    H = np.random.default_rng(5).integers(-2048, 2049, size=(len(pr), T)).astype(f32)
    zero, inc = grid_basis(N, pos=(0.05, -0.02, 0.7), size=(0.25, 0.22, 0.3))
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, 0.001)

    def run(h, keep):
        with gpu.Plan(h, pr, N, zero, inc, t0, 0.001, variant=variant, keep_f32_rows=keep) as plan:
            plan.execute()
            return plan.volume(), plan.info()["row_storage"]

    v16, rows = run(H, False)
    v32, rows32 = run(H, True)
    assert rows == ("f32" if variant == 6 else "f16") and rows32 == "f32"
    assert np.array_equal(v16, v32) and np.array_equal(v16, ref.astype(f32))
    for bad in (0.5, 2049.0, -3000.0, np.nan):
        Hb = H.copy()
        Hb[3, 1234] = bad
        vb, rows_b = run(Hb, False)
        vk, _ = run(Hb, True)
        assert rows_b == "f32", bad
        assert np.array_equal(vb, vk, equal_nan=True), bad


def test_exact_fp16_check_body_and_tail(gpu, oracle):
    """The device check of the samples reads them 16 bytes at a time and the last P*T mod 4
    one at a time (here 225 x 2001 = 1 mod 4): one non-integer sample at the first element,
    inside, or at the very last element keeps the execution on fp32 rows."""
    pr = confocal_pairs(15, jitter=0.01, seed=11)
    T, N, t0 = 2001, (33, 35, 31), 3.3
    assert len(pr) * T % 4 == 1
    H = np.random.default_rng(8).integers(-2048, 2049, size=(len(pr), T)).astype(f32)
    zero, inc = grid_basis(N, pos=(0.0, 0.05, 0.45), size=(0.25, 0.25, 0.25))

    def run(h):
        with gpu.Plan(h, pr, N, zero, inc, t0, 0.001, variant=7) as plan:
            plan.execute()
            return plan.volume(), plan.info()["row_storage"]

    vol, rows = run(H)
    assert rows == "f16"
    assert np.array_equal(vol, oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, 0.001).astype(f32))
    for p, b in [(0, 0), (100, 1000), (len(pr) - 1, T - 1)]:
        Hb = H.copy()
        Hb[p, b] = 0.5
        _, rows_b = run(Hb)
        assert rows_b == "f32", (p, b)


@pytest.mark.parametrize("storage,cplx", [("f16", False), ("bf16", False), ("f32", True), ("f16", True),
                                          ("bf16", True)])
def test_lds_staged_kernel_formats(gpu, storage, cplx):
    """Every storage format through the LDS-staged kernel equals the texture-path kernel bit for
    bit (same element per update, same order); complex formats use 16 x 16 x 8 tiles."""
    pr = confocal_pairs(16, jitter=0.01, seed=7)
    T = 3000
    N = (35, 33, 40)
    zero, inc = grid_basis(N, pos=(0.05, -0.02, 0.7), size=(0.25, 0.22, 0.3))
    rng = np.random.default_rng(4)
    H = rng.gamma(0.5, 2.0, (len(pr), T)).astype(f32)
    if cplx:
        H = (H - 1j * rng.gamma(0.5, 2.0, H.shape)).astype(np.complex64)
    vols = {}
    for v in (6, 7, 8, 5):
        with gpu.Plan(H, pr, N, zero, inc, 1.2, 0.001, storage=storage, variant=v) as plan:
            lv = plan.info()["lane_voxels"]
            # complex 16-bit rows: 16 x 16 x 8 tiles; complex64: two real passes, real shapes
            assert lv == {6: 8 if cplx and storage != "f32" else 16, 7: 8, 8: 8, 5: 2}[v], (v, lv)
            plan.execute()
            vols[v] = plan.volume()
    assert all(np.array_equal(vols[v], vols[5]) for v in (6, 7, 8))
    assert np.abs(vols[5]).max() > 0


@pytest.mark.parametrize("variant", [6, 7, 8])
def test_lds_staged_kernel_windows_before_the_row(gpu, oracle, variant):
    """Pair count not a multiple of any stage size (15 x 15 = 225) and t0 late enough that the
    bin windows of near voxels start before bin 0 (the staging path that zero-fills elements
    below the row) while far voxels fall past Tc: bins exact against the oracle."""
    pr = confocal_pairs(15, jitter=0.01, seed=11)
    T = 2000
    H = coded_samples(len(pr), T, seed=9)
    N = (33, 35, 31)
    zero, inc = grid_basis(N, pos=(0.0, 0.05, 0.45), size=(0.25, 0.25, 0.25))
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, 3.3, 0.001)
    with gpu.Plan(H, pr, N, zero, inc, 3.3, 0.001, variant=variant) as plan:
        assert plan.info()["kernel"].startswith("confocal_lds_kernel")
        plan.execute()
        vol = plan.volume()
    assert np.array_equal(vol, ref.astype(f32))
    # the geometry does reach both ends: raw bins below 0 and past Tc for some voxel-pairs
    vox = np.array([[0, 0, 0], [16, 17, 15], [32, 34, 30], [0, 34, 0]], np.uint32)
    _, raw = oracle.a1_bins(pr, np.tile(vox, (len(pr), 1)), np.repeat(np.arange(len(pr)), len(vox)),
                            zero, inc, 3.3, 0.001, T)
    assert raw.min() < 0 and raw.max() >= T


@pytest.mark.parametrize("variant", [-1, 5, 7])
def test_row_windows_entirely_before_or_past_the_row(gpu, oracle, variant):
    """Rows are staged only over each pair's bin window (the elements a gather can read).  A
    third of the pairs see every bin before the row (short laser/camera legs), a third every
    bin past its end (long legs), the rest straddle both ends: their windows clamp to the
    row's ends.  Volumes equal the oracle bin for bin, and the threshold readback is the full
    exact table (entries outside the windows from the same search)."""
    pr = confocal_pairs(12, jitter=0.01, seed=21)
    legs = np.array([0.1, 2.2, 1.5], f32)[np.arange(len(pr)) % 3]
    pr["laser_wall"] = legs
    pr["camera_wall"] = legs
    T, t0, dT = 1200, 4.0, 0.001
    H = coded_samples(len(pr), T, seed=4)
    N = (21, 19, 23)
    zero, inc = grid_basis(N, pos=(0.0, 0.0, 0.6), size=(0.3, 0.3, 0.3))
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, dT)
    with gpu.Plan(H, pr, N, zero, inc, t0, dT, variant=variant) as plan:
        plan.execute()
        vol = plan.volume()
        got = plan.thresholds()
    assert np.array_equal(vol, ref.astype(f32))
    assert ref.max() > 0
    want = oracle.confocal_thresholds(pr["laser_wall"], pr["camera_wall"], T + 1, t0, dT)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the geometry reaches all three cases: every bin before the row, every bin past it, both
    vox = np.array([[i, j, k] for i in (0, N[0] - 1) for j in (0, N[1] - 1) for k in (0, N[2] - 1)], np.uint32)
    _, raw = oracle.a1_bins(pr, np.tile(vox, (len(pr), 1)), np.repeat(np.arange(len(pr)), len(vox)),
                            zero, inc, t0, dT, T)
    raw = raw.reshape(len(pr), len(vox))
    assert (raw.max(1) < 0).any() and (raw.min(1) >= T).any()
    assert ((raw.min(1) < 0) & (raw.max(1) >= T)).any() or ((raw.min(1) >= 0) & (raw.max(1) < T)).any()


def _tap_error_bound(pr, N, zero, inc, t0, dT, T):
    """The plan's precondition of the exact bin lookup: 16 x 2^-24 x Q with
    Q = (max total path + |t0|) / deltaT + T + 2, the total over pairs and grid corners."""
    idx = np.array([[i, j, k] for i in (0, N[0] - 1) for j in (0, N[1] - 1) for k in (0, N[2] - 1)], np.float64)
    corners = zero.astype(np.float64) + idx @ inc.reshape(3, 3).astype(np.float64)
    lp, cp = pr["laser_point"].astype(np.float64), pr["camera_point"].astype(np.float64)
    dl = np.linalg.norm(lp[:, None, :] - corners[None], axis=2).max(1)
    dc = np.linalg.norm(cp[:, None, :] - corners[None], axis=2).max(1)
    tot = (np.abs(pr["laser_wall"]) + np.abs(pr["camera_wall"]) + dl + dc).max()
    return 16.0 * 2.0 ** -24 * ((tot + abs(t0)) / dT + T + 2.0)


def test_biased_quotient_offsets_near_the_error_bound(gpu, oracle):
    """The gathers take each element's offset from the mantissa of the quotient plus a bias,
    which rounds it to 1/8 twice on top of the host-bounded error of the quotient itself.  A
    capture whose bound sits close to its 1/4 limit (20-um voxels, 64-um bins, long paths):
    every LDS-staged shape, with fp32 and exact fp16 rows, and the texture path, bit-exact
    against the oracle; bins below 0 and above Tc included."""
    pr = confocal_pairs(8, jitter=0.01, seed=11)
    N = (24, 20, 18)
    pitch = 2e-5
    zero, inc = grid_basis(N, pos=(0.05, -0.02, 0.7), size=tuple(pitch * (n - 1) for n in N))
    dT, T, t0 = 6.4e-5, 8000, 4.1
    bound = _tap_error_bound(pr, N, zero, inc, t0, dT, T)
    assert 0.15 < bound < 0.25, bound
    H = coded_samples(len(pr), T, seed=9)
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, dT)
    assert np.count_nonzero(ref) > ref.size // 4          # most voxels see in-range bins
    for variant in (5, 6, 7, 8):
        for keep in (False, True):
            with gpu.Plan(H, pr, N, zero, inc, t0, dT, variant=variant, keep_f32_rows=keep) as plan:
                plan.execute()
                vol = plan.volume()
                info = plan.info()
            assert ("lds" in info["kernel"]) == (variant != 5), info["kernel"]
            assert info["row_storage"] == ("f16" if variant in (7, 8) and not keep else "f32"), (variant, keep, info)
            assert np.array_equal(vol, ref.astype(f32)), (variant, keep)


def test_lds_staged_kernel_falls_back(gpu, oracle):
    """Variant 6 where the LDS kernel does not apply (coarse pitch: windows too long; a wall
    that is not flat) runs the fused texture-path kernel and stays exact."""
    T = 2500
    for jitter_z, N, size in [(False, (20, 18, 22), (1, 1, 1)), (True, (35, 33, 40), (0.25, 0.22, 0.3))]:
        pr = confocal_pairs(12, jitter=0.02, seed=3)
        if jitter_z:
            pr["laser_point"][:, 2] = np.linspace(0, 0.05, len(pr)).astype(f32)
            pr["camera_point"] = pr["laser_point"]
        H = coded_samples(len(pr), T, seed=5)
        zero, inc = grid_basis(N, size=size)
        with gpu.Plan(H, pr, N, zero, inc, 2.05, 0.001, variant=6) as plan:
            assert plan.info()["lane_voxels"] == 4
            plan.execute()
            vol = plan.volume()
        ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, 2.05, 0.001)
        assert np.array_equal(vol, ref.astype(f32)), (jitter_z, N)


@pytest.mark.parametrize("variant,storage,cplx", [(-1, "f32", False), (0, "f32", False), (2, "f32", False),
                                                  (5, "f16", False), (5, "bf16", False), (5, "f32", True)])
def test_threshold_table_bit_exact(gpu, oracle, variant, storage, cplx):
    """Every (pair, bin) threshold of the staged table equals the oracle's bisection over all
    non-negative floats of the reference bin formula (fused and split layouts, all formats)."""
    rng = np.random.default_rng(17)
    pr = confocal_pairs(24, jitter=0.01, seed=3)
    pr["camera_wall"] = rng.uniform(0.6, 2.2, len(pr)).astype(f32)
    pr["laser_wall"] = rng.uniform(0.6, 2.2, len(pr)).astype(f32)
    T = 1500
    H = coded_samples(len(pr), T).astype(np.complex64) * (1 + 1j) if cplx else coded_samples(len(pr), T)
    zero, inc = grid_basis((8, 8, 8))
    for t0, dT in [(2.0, 0.001), (1.1, 0.0013), (0.0, 2.0 ** -9)]:
        with gpu.Plan(H, pr, (8, 8, 8), zero, inc, t0, dT, storage=storage, variant=variant) as plan:
            plan.execute()
            got = plan.thresholds()
        ref = oracle.confocal_thresholds(pr["laser_wall"], pr["camera_wall"], T + 1, t0, dT)
        assert got.shape == ref.shape
        bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
        assert bad.size == 0, (t0, dT, bad[:5], got.ravel()[bad[:5]], ref.ravel()[bad[:5]])


def test_round_half_away_ties_on_gpu(gpu, oracle):
    # total = 4 exactly (walls 1.25, voxel 0.75 above the wall point); t0 places q at k + 0.5
    pr = np.zeros(1, gpu.PAIR_DTYPE)
    pr["laser_wall"] = pr["camera_wall"] = 1.25
    zero = np.array([0, 0, 0.75], f32)
    inc = np.eye(3, dtype=f32).ravel()
    T = 512
    H = np.arange(1, T + 1, dtype=f32)[None, :]
    for k in (100, 101, 37, 0, -1, 510, 511):
        t0 = f32(4.0) - f32((k + 0.5) / 256)
        vol = gpu.backprojection(H, pr, (1, 1, 1), zero, inc, float(t0), 2.0 ** -8)
        ref = oracle.a1_volume_f64acc(H, pr, (1, 1, 1), zero, inc, float(t0), 2.0 ** -8)
        # round half away: k + 0.5 -> k + 1 for k >= 0, -0.5 -> -1 (no tap); code = bin + 1
        want = k + 2 if (k >= 0 and k + 1 <= T - 1) else 0
        assert vol[0, 0, 0] == ref[0, 0, 0] == want, k


def test_plan_tune_keeps_the_volume(gpu, oracle):
    """bpx_plan_tune times the gather variants that apply to a confocal plan (the texture-path
    shapes 3-5, the LDS-staged shapes 6-8 whose bin windows fit -- here all of them, the 8-voxel
    ones with exact fp16 rows) and keeps the fastest: its kernel then runs, and the volume equals
    the default kernel's and the oracle's bit for bit.  A plan of an exhaustive grid keeps its
    variant."""
    pr = confocal_pairs(16, jitter=0.01, seed=7)
    T, N, t0 = 3000, (35, 33, 40), 1.2
    H = coded_samples(len(pr), T, seed=3)
    zero, inc = grid_basis(N, pos=(0.05, -0.02, 0.7), size=(0.25, 0.22, 0.3))
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, t0, 0.001).astype(f32)
    with gpu.Plan(H, pr, N, zero, inc, t0, 0.001) as plan:
        plan.execute()
        before = plan.volume()
        got = plan.tune(reps=2)
        assert 3 <= got["variant"] <= 8, got
        assert got["kernel"].startswith("confocal_lds_kernel") == (got["variant"] >= 6), got
        plan.execute()
        after = plan.volume()
        assert plan.info()["kernel"] == got["kernel"]
    assert np.array_equal(before, ref) and np.array_equal(after, ref)
    lg = scenes.wall_grid(8)
    cg = (scenes.wall_grid(6) * f32(0.9)).astype(f32)
    prg = oracle.precompute_pairs(scenes.ORIGIN, lg, scenes.ORIGIN, cg, confocal=False)
    Hg = coded_samples(len(prg), T, seed=2)
    with gpu.Plan(Hg, prg, (16, 12, 20), *grid_basis((16, 12, 20)), 2.2, 0.001) as plan:
        k0 = plan.info()["kernel"]
        assert plan.info()["path"] == "grid"
        plan.tune(reps=1)
        assert plan.info()["kernel"] == k0
        plan.execute()
        assert np.array_equal(plan.volume(), oracle.a1_volume_f64acc(Hg, prg, (16, 12, 20), *grid_basis((16, 12, 20)),
                                                                      2.2, 0.001).astype(f32))


def test_exhaustive_grid_bins_bit_exact(gpu, oracle):
    lg = scenes.wall_grid(8)
    cg = (scenes.wall_grid(6) * f32(0.9)).astype(f32)
    pr = oracle.precompute_pairs(scenes.ORIGIN, lg, scenes.ORIGIN, cg, confocal=False)
    T = 3000
    H = coded_samples(len(pr), T, seed=2)
    N = (16, 12, 20)
    zero, inc = grid_basis(N)
    vol = gpu.backprojection(H, pr, N, zero, inc, 2.2, 0.001)
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, 2.2, 0.001)
    assert np.array_equal(vol, ref.astype(f32))


def test_arbitrary_pairs_bins_bit_exact(gpu, oracle):
    rng = np.random.default_rng(11)
    P = 300
    pr = np.zeros(P, gpu.PAIR_DTYPE)
    pr["laser_point"] = rng.uniform(-1, 1, (P, 3)) * [1, 1, 0.05]
    pr["camera_point"] = rng.uniform(-1, 1, (P, 3)) * [1, 1, 0.05]
    pr["laser_wall"] = rng.uniform(0.5, 1.5, P)
    pr["camera_wall"] = rng.uniform(0.5, 1.5, P)
    T = 4000
    H = coded_samples(P, T, seed=9)
    N = (13, 11, 9)
    zero, inc = grid_basis(N)
    vol = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001)
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, 2.0, 0.001)
    assert np.array_equal(vol, ref.astype(f32))


def test_metric_size_bins_on_sampled_voxels(gpu, oracle):
    # 128 x 128 confocal pairs into 128^3 with coded samples; 2048 random voxels checked
    pr = confocal_pairs(128)
    T = 3600
    H = coded_samples(len(pr), T, seed=1)
    N = (128, 128, 128)
    zero, inc = grid_basis(N)
    vol = gpu.backprojection(H, pr, N, zero, inc, 2.1, 0.001)
    vox = sample_voxels(N, 2048, seed=4)
    ref = oracle.a1_voxels_f64acc(H, pr, vox, zero, inc, 2.1, 0.001)
    got = vol[vox[:, 0], vox[:, 1], vox[:, 2]]
    assert np.array_equal(got, ref.astype(f32))


def test_exhaustive_full_size_bins_on_sampled_voxels(gpu, oracle):
    """The 16x16 lasers x 16x16 SPADs (65,536 pairs), T = 4096, 64^3 workload at full size with
    integer sample codes (every sum exact): 1,024 random voxels equal the oracle's exactly."""
    lg = scenes.wall_grid(16)
    pr = oracle.precompute_pairs(scenes.ORIGIN, lg, scenes.ORIGIN, lg, confocal=False)
    T = 4096
    P = len(pr)
    assert P == 65536
    H = (np.add.outer(13 * np.arange(P, dtype=np.int32), 7 * np.arange(T, dtype=np.int32)) % 61 + 1).astype(f32)
    N = (64, 64, 64)
    zero, inc = grid_basis(N)
    vol = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001)
    vox = sample_voxels(N, 1024, seed=6)
    ref = oracle.a1_voxels_f64acc(H, pr, vox, zero, inc, 2.0, 0.001)
    assert np.array_equal(vol[vox[:, 0], vox[:, 1], vox[:, 2]], ref.astype(f32))
    assert np.count_nonzero(ref) > 900


def test_headline_size_linearity(gpu):
    """128x128 confocal pairs -> 128^3, T = 4096 at full size: with integer codes every voxel sum
    is exact, so the reconstruction is exactly linear (vol(a) + vol(b) == vol(a + b))."""
    pr = confocal_pairs(128)
    T = 4096
    a = coded_samples(len(pr), T, seed=3)
    b = np.roll(coded_samples(len(pr), T, seed=8), 5, axis=1)
    N = (128, 128, 128)
    zero, inc = grid_basis(N)
    va = gpu.backprojection(a, pr, N, zero, inc, 2.0, 0.001)
    vb = gpu.backprojection(b, pr, N, zero, inc, 2.0, 0.001)
    vab = gpu.backprojection(a + b, pr, N, zero, inc, 2.0, 0.001)
    assert np.abs(va).max() > 0 and np.array_equal(va + vb, vab)


# ----------------------------------------------------------------- real-valued volumes
def test_confocal_volume_parity_rendered(gpu, oracle):
    cap = scenes.render(scenes.config_scene(1), 32, 4096, t0=2.0, seed=1001)
    pr = oracle.precompute_pairs(cap.laser_position, cap.laser_grid, cap.camera_position,
                                 cap.camera_grid, True)
    H = cap.transient.reshape(len(pr), -1)
    N = (32, 32, 32)
    zero, inc = grid_basis(N)
    vol = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001)
    exact = oracle.a1_volume_f64acc(H, pr, N, zero, inc, 2.0, 0.001)
    refo = oracle.a1_volume(H, pr, N, zero, inc, 2.0, 0.001)
    fp64 = oracle.fp64_volume(H, pr, N, zero, inc, 2.0, 0.001)
    assert rel_l2(vol, exact) <= 1e-6
    assert rel_l2(vol, refo) <= 1e-5
    err64 = rel_l2(vol, fp64)
    print(f"confocal 32^2 -> 32^3: rel L2 vs reference order {rel_l2(vol, refo):.2e}, "
          f"vs FP64 oracle {err64:.2e}")
    assert err64 < 5e-2


def test_complex_bins_bit_exact(gpu, oracle):
    pr = confocal_pairs(16, jitter=0.01, seed=8)
    T = 2048
    H = coded_samples(len(pr), T) + 1j * coded_samples(len(pr), T, seed=17)
    N = (12, 14, 16)
    zero, inc = grid_basis(N)
    re = oracle.a1_volume_f64acc(H.real.astype(f32), pr, N, zero, inc, 2.0, 0.001)
    im = oracle.a1_volume_f64acc(H.imag.astype(f32), pr, N, zero, inc, 2.0, 0.001)
    for v in (2, 5):
        for storage in ("f32", "f16", "bf16"):   # codes <= 61 are exact in every format
            vol = gpu.backprojection(H.astype(np.complex64), pr, N, zero, inc, 2.0, 0.001,
                                     variant=v, storage=storage)
            assert vol.dtype == np.complex64
            assert np.array_equal(vol.real, re.astype(f32)), (v, storage)
            assert np.array_equal(vol.imag, im.astype(f32)), (v, storage)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("variant", [2, 5])
def test_half_storage_error_vs_fp64(gpu, oracle, storage, variant):
    cap = scenes.render(scenes.config_scene(1), 32, 4096, t0=2.0, seed=1001)
    pr = oracle.precompute_pairs(cap.laser_position, cap.laser_grid, cap.camera_position,
                                 cap.camera_grid, True)
    H = cap.transient.reshape(len(pr), -1)
    N = (24, 24, 24)
    zero, inc = grid_basis(N)
    vol = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001, storage=storage, variant=variant)
    v32 = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001, variant=variant)
    fp64 = oracle.fp64_volume(H, pr, N, zero, inc, 2.0, 0.001)
    e16, e32 = rel_l2(vol, fp64), rel_l2(v32, fp64)
    print(f"{storage}: rel L2 vs FP64 oracle {e16:.2e} (fp32 storage {e32:.2e}), "
          f"vs fp32 storage {rel_l2(vol, v32):.2e}")
    if storage == "f16" and H.max() <= 2048:
        assert np.array_equal(vol, v32)      # integer counts <= 2048 are exact in fp16
    assert rel_l2(vol, v32) < 4e-3


def test_phasor_precision_full_size_vs_fp64(gpu, oracle):
    """The complex 128x128 -> 128^3 workload at full size: the benchmark capture filtered with
    the 4 mm / sigma 12 mm pulse, backprojected with complex64, complex fp16 and complex bf16
    rows (per-capture scale).  On 1,024 sampled voxels: complex64 equals the fp32 formula with
    exact accumulation to fp32 rounding, and each storage states its error against the FP64
    oracle (real and imaginary parts backprojected separately: the map is linear)."""
    cap = scenes.metric_capture()
    N = (128, 128, 128)
    pre = gpu.nlosbpy.prepare(*cap.nlos_args(N), np.zeros(3), assume_row_major=True)
    c = gpu.phasor_pulse(pre["chunk"], 0.004, cap.deltaT, sigma=0.012)
    pr, zero, inc, t0 = pre["pairs"], pre["zero"], pre["inc"], float(pre["chunked_t0"])
    vox = sample_voxels(N, 1024, seed=12)
    exact = (oracle.a1_voxels_f64acc(c.real.copy(), pr, vox, zero, inc, t0, cap.deltaT)
             + 1j * oracle.a1_voxels_f64acc(c.imag.copy(), pr, vox, zero, inc, t0, cap.deltaT))
    fp64 = (oracle.fp64_voxels(c.real.copy(), pr, vox, zero, inc, t0, cap.deltaT)
            + 1j * oracle.fp64_voxels(c.imag.copy(), pr, vox, zero, inc, t0, cap.deltaT))
    got = {}
    for st in ("f32", "f16", "bf16"):
        vol = gpu.backprojection(c, pr, N, zero, inc, t0, cap.deltaT, storage=st)
        got[st] = vol[vox[:, 0], vox[:, 1], vox[:, 2]]
        print(f"complex {st}: rel L2 vs FP64 oracle {rel_l2(got[st], fp64):.2e}, "
              f"vs exact fp32 formula {rel_l2(got[st], exact):.2e}")
    assert rel_l2(got["f32"], exact) <= 1e-6
    # the fp32 path-length formula itself is ~1 % from FP64 on phasor data: near rounding
    # boundaries the two pick neighbouring bins, and the filtered signal changes sign within a
    # few bins (measured 1.1e-2 for the reference pulse on the same voxels)
    assert rel_l2(got["f32"], fp64) <= 3e-2
    for st in ("f16", "bf16"):
        assert rel_l2(got[st], got["f32"]) < 4e-3


# ----------------------------------------------------------------- CPU-path semantics
@pytest.mark.parametrize("confocal", [True, False])
def test_window_parity(gpu, oracle, confocal):
    rng = np.random.default_rng(21)
    if confocal:
        pr = confocal_pairs(16)
    else:
        pr = oracle.precompute_pairs(scenes.ORIGIN, scenes.wall_grid(6), scenes.ORIGIN,
                                     scenes.wall_grid(5), confocal=False)
    T = 2600
    H = rng.poisson(2.0, (len(pr), T)).astype(f32)
    N = (20, 18, 16)
    vp, vs = np.array([0, 0, 0.8], f32), np.array([1, 1, 1], f32)
    vol = gpu.window_backprojection(H, pr, N, vp, vs, 2.0, 0.001)
    ref = oracle.a2_volume(H, pr, N, vp, vs, 2.0, 0.001)
    fp64 = oracle.a2_volume_fp64(H, pr, N, vp, vs, 2.0, 0.001)
    print(f"window confocal={confocal}: rel L2 vs CPU path {rel_l2(vol, ref):.2e}, "
          f"vs FP64 {rel_l2(vol, fp64):.2e}")
    assert rel_l2(vol, ref) <= 1e-5


# ----------------------------------------------------------------- phasor
@pytest.mark.parametrize("wl", [0.004, 0.005, 0.04])
def test_phasor_filter_bit_exact(gpu, oracle, wl):
    rng = np.random.default_rng(3)
    x = rng.poisson(4.0, (37, 1500)).astype(f32)
    got = gpu.phasor_pulse(x, wl, 0.001)
    want = oracle.phasor_filter(x, wl, 0.001)
    assert np.array_equal(got, want)


def test_phasor_filter_explicit_sigma_vs_fp64(gpu):
    """The explicit-sigma pulse (the 4 mm / sigma 12 mm phasor filter, which the reference's
    sigma = 4 lambda / 6 cannot express): taps k = 1..2h+1, h = ceil(3 sigma / dT), Gaussian of
    t = dT (k - pulse/2) times the carrier exp(2 pi i k dT / lambda), applied like the reference
    filter (out[t] = sum_j in[t + j - pulse/2] tap_j over 0 < idx < T).  Checked against the same
    sum in float64; no reference output exists for this form (parity unpinned by design)."""
    rng = np.random.default_rng(9)
    x = rng.poisson(4.0, (23, 1200)).astype(f32)
    wl, dT, sigma = 0.004, 0.001, 0.012
    got = gpu.phasor_pulse(x, wl, dT, sigma=sigma)
    half = int(np.ceil(3 * sigma / dT))
    pulse = 2 * half + 1
    k = np.arange(1, pulse + 1, dtype=np.float64)
    t = dT * (k - pulse / 2.0)
    g = np.exp(-t * t / (2 * sigma * sigma))
    taps = (np.cos(2 * np.pi * k * dT / wl) * g).astype(f32).astype(np.float64) + \
        1j * (np.sin(2 * np.pi * k * dT / wl) * g).astype(f32).astype(np.float64)
    T = x.shape[1]
    want = np.zeros(x.shape, np.complex128)
    for j in range(pulse):
        idx = np.arange(T) + j - pulse // 2
        ok = (idx > 0) & (idx < T)
        want[:, ok] += x[:, idx[ok]].astype(np.float64) * taps[j]
    assert rel_l2(got, want) <= 2e-7
    assert abs(np.abs(got).max() - np.abs(want).max()) <= 1e-5 * np.abs(want).max()


# ----------------------------------------------------------------- reference pipeline
def test_confocal_32x32_t2048_workload_literal_and_shifted(gpu, oracle):
    """The 32x32 confocal, T = 2048, t0 = 0, 32^3 workload as specified.  Its rendered capture is
    empty: every path through the hidden points is longer than T * deltaT = 2.048 m.  Voxel bins
    start at 2028 (the voxels nearest the wall), so with non-zero samples on the raw rows only a
    few voxel-pair bins are in range and all others must be dropped: the GPU volume equals the
    oracle's exactly.  With t0 = 2.0 m the scene is inside the window and the CPU-path semantics
    match the oracle."""
    cap = scenes.render(scenes.config_scene(1), 32, 2048, t0=0.0, seed=1001)
    assert not cap.transient.any()
    N = (32, 32, 32)
    for use_cpu in (False, True):
        v = gpu.nlosbpy.backproject(*cap.nlos_args(N), np.zeros(3), assume_row_major=True, use_cpu=use_cpu)
        assert not v.any()
    pre = gpu.nlosbpy.prepare(*cap.nlos_args(N), np.zeros(3), assume_row_major=True)
    P = len(pre["pairs"])
    coded = coded_samples(P, 2048)
    v0 = gpu.backprojection(coded, pre["pairs"], N, pre["zero"], pre["inc"], 0.0, cap.deltaT)
    e0 = oracle.a1_volume_f64acc(coded, pre["pairs"], N, pre["zero"], pre["inc"], 0.0, cap.deltaT)
    assert np.array_equal(v0, e0.astype(f32))
    assert 0 < np.count_nonzero(e0) < e0.size // 100
    shifted = scenes.render(scenes.config_scene(1), 32, 2048, t0=2.0, seed=1001)
    assert shifted.transient.any()
    ref, _ = oracle.backproject_pipeline(shifted.transient, shifted.camera_grid, shifted.laser_grid,
                                         shifted.camera_position, shifted.laser_position, shifted.t0,
                                         shifted.deltaT, True, shifted.volume_position, shifted.volume_size, N,
                                         use_cpu=True)
    v = gpu.nlosbpy.backproject(*shifted.nlos_args(N), np.zeros(3), assume_row_major=True, use_cpu=True)
    assert rel_l2(v, ref) <= 1e-5 and np.abs(ref).max() > 0


def test_nlosbpy_backproject_matches_oracle_pipeline(gpu, oracle):
    cap = scenes.render(scenes.config_scene(1), 32, 2048, t0=2.0, seed=1001)
    N = (24, 24, 24)
    args = cap.nlos_args(N)
    vol = gpu.nlosbpy.backproject(*args, np.zeros(3, f32), assume_row_major=True)
    ref, inter = oracle.backproject_pipeline(cap.transient, cap.camera_grid, cap.laser_grid,
                                             cap.camera_position, cap.laser_position, cap.t0,
                                             cap.deltaT, True, cap.volume_position,
                                             cap.volume_size, N)
    exact = oracle.a1_volume_f64acc(inter["H"], inter["pairs"], N, inter["zero"], inter["inc"],
                                    float(inter["t0"]), cap.deltaT)
    assert rel_l2(vol, exact) <= 1e-6 and rel_l2(vol, ref) <= 1e-5


def test_import_nlosbpy_matches_oracle_pipeline(gpu, oracle):
    """A reference caller's `import nlosbpy` (nlosbpy.cpp:81) runs the HIP path: plain and
    phasor volumes against the oracle's bp::backproject / bp::phasor_reconstruction."""
    import nlosbpy
    cap = scenes.render(scenes.config_scene(1), 16, 2048, t0=2.0, seed=1011)
    N = (16, 12, 20)
    vol = nlosbpy.backproject(*cap.nlos_args(N), np.zeros(3, f32), assume_row_major=True)
    ref, inter = oracle.backproject_pipeline(cap.transient, cap.camera_grid, cap.laser_grid,
                                             cap.camera_position, cap.laser_position, cap.t0,
                                             cap.deltaT, True, cap.volume_position,
                                             cap.volume_size, N)
    exact = oracle.a1_volume_f64acc(inter["H"], inter["pairs"], N, inter["zero"], inter["inc"],
                                    float(inter["t0"]), cap.deltaT)
    assert vol.dtype == np.float32 and vol.shape == N and np.abs(ref).max() > 0
    assert rel_l2(vol, exact) <= 1e-6 and rel_l2(vol, ref) <= 1e-5
    cv = nlosbpy.phasor(*cap.nlos_args(N), 0.004, assume_row_major=True)
    cref, _ = oracle.phasor_pipeline(cap.transient, cap.camera_grid, cap.laser_grid,
                                     cap.camera_position, cap.laser_position, cap.t0, cap.deltaT,
                                     True, cap.volume_position, cap.volume_size, N, 0.004)
    assert cv.dtype == np.complex64 and rel_l2(cv, cref) <= 1e-5


@pytest.mark.parametrize("t0,T,nan_pair,confocal", [
    (2.0, 2048, False, True),     # window starts inside the rows and runs past the last bin
    (4.0, 2048, True, True),      # window starts before bin 0 (zero head), one NaN pair dropped
    (1.0, 8192, False, True),     # window strictly inside the rows
    (2.0, 2048, True, False),     # exhaustive pairs, dropped pairs compacted on the device
])
def test_nlosbpy_device_crop_equals_host_chunk(gpu, t0, T, nan_pair, confocal):
    """The one-shot calls crop the source rows on the device (get_transient_chunk,
    backproject.hpp:689-740); the result must equal the plan built from the host chunk that
    the same pipeline returns (nlosbpy.prepare), GPU and CPU-path semantics alike."""
    cap = scenes.render(scenes.config_scene(1), 12, T, t0=t0, confocal=confocal,
                        laser_side=None if confocal else 4, seed=11)
    N = (16, 16, 16)
    args = list(cap.nlos_args(N))
    if nan_pair:
        args[1] = args[1].copy()
        args[1][3, 5, 0] = np.nan
    pre = gpu.nlosbpy.prepare(*args, np.zeros(3), assume_row_major=True)
    lo, hi = np.floor(pre["min_index"]), np.ceil(pre["max_index"])
    assert (lo < 0) == (t0 == 4.0) and (hi > T) == (T == 2048)
    ok = np.isfinite(pre["pairs"]["camera_point"]).all(1)
    assert ok.all() != nan_pair
    vol = gpu.nlosbpy.backproject(*args, np.zeros(3), assume_row_major=True)
    ref = gpu.backprojection(pre["chunk"], pre["pairs"], N, pre["zero"], pre["inc"], pre["chunked_t0"],
                             cap.deltaT)
    assert np.array_equal(vol, ref) and np.abs(ref).max() > 0
    vc = gpu.nlosbpy.backproject(*args, np.zeros(3), assume_row_major=True, use_cpu=True)
    rc = gpu.window_backprojection(pre["chunk"], pre["pairs"], N, cap.volume_position, cap.volume_size,
                                   pre["chunked_t0"], cap.deltaT)
    assert np.array_equal(vc, rc) and np.abs(rc).max() > 0


@pytest.mark.parametrize("t0,nan_pair,use_cpu", [(2.0, False, False), (4.0, True, False), (4.0, True, True)])
def test_nlosbpy_phasor_device_rows_equal_host_route(gpu, t0, nan_pair, use_cpu):
    """nlosbpy.phasor crops and filters on the device and hands the complex rows to the plan;
    the volume must equal the host route: chunk with the int-truncated limits
    (backproject.hpp:962-969), phasor_pulse, then the complex backprojection."""
    cap = scenes.render(scenes.config_scene(1), 12, 2048, t0=t0, seed=13)
    N = (12, 12, 12)
    args = list(cap.nlos_args(N))
    if nan_pair:
        args[1] = args[1].copy()
        args[1][2, 7, 1] = np.nan
    wl = 0.005
    pre = gpu.nlosbpy.prepare(*args, np.zeros(3), assume_row_major=True)
    imin, imax = f32(int(pre["min_index"])), f32(int(pre["max_index"]))
    lo, hi = int(np.floor(imin)), int(np.ceil(imax))
    H = cap.transient.reshape(-1, cap.transient.shape[-1])
    T = H.shape[1]
    src_from, src_to, c_from = max(lo, 0), min(max(hi, 0), T), (-lo if lo < 0 else 0)
    chunk = np.zeros((H.shape[0], hi - lo), f32)
    if src_to > src_from and c_from + src_to - src_from <= hi - lo:
        chunk[:, c_from:c_from + src_to - src_from] = H[:, src_from:src_to]
    cplx = gpu.phasor_pulse(chunk, wl, cap.deltaT)
    ct0 = f32(cap.t0) + imin * f32(cap.deltaT)
    if use_cpu:
        ref = gpu.window_backprojection(cplx, pre["pairs"], N, cap.volume_position, cap.volume_size, ct0,
                                        cap.deltaT)
    else:
        ref = gpu.backprojection(cplx, pre["pairs"], N, pre["zero"], pre["inc"], ct0, cap.deltaT)
    vol = gpu.nlosbpy.phasor(*args, wl, assume_row_major=True, use_cpu=use_cpu)
    assert (lo < 0) == (t0 == 4.0)
    assert vol.dtype == np.complex64 and np.abs(ref).max() > 0
    assert np.array_equal(vol, ref)


def test_centred_wall_basis(gpu):
    """BPX_FLAG_CENTRED_WALL_BASIS: the normal (0, -1, 0) gives the identity basis (the same
    volume as no normal); a normal along -z (NaN axes in the reference construction) and a
    normal along +x give right-handed bases whose voxels are the identity grid's voxels
    permuted -- the volumes are the identity volume permuted, bit for bit."""
    cap = scenes.render(scenes.config_scene(1), 24, 4096, t0=0.0, seed=1)
    N = (32, 32, 32)

    def vol(normal, centred):
        with gpu.Plan.from_nlos(*cap.nlos_args(N), wall_normal=normal, assume_row_major=True,
                                centred_wall_basis=centred) as pl:
            pl.execute()
            return pl.volume()

    ident = vol((0, 0, 0), False)
    assert np.abs(ident).max() > 0
    assert np.array_equal(vol((0, -1, 0), True), ident)
    assert np.array_equal(vol((0, 0, 0), True), ident)
    along_z = vol((0, 0, -1), True)              # axes (0,1,0), (0,0,1), (1,0,0)
    assert np.all(np.isfinite(along_z)) and np.array_equal(along_z, np.transpose(ident, (1, 2, 0)))
    along_x = vol((1, 0, 0), True)               # axes (0,1,0), (-1,0,0), (0,0,1)
    assert np.array_equal(along_x, np.transpose(ident[::-1], (1, 0, 2)))


@pytest.mark.parametrize("use_cpu,nan_pair", [(False, False), (True, True)])
def test_nlosbpy_phasor_multi_equals_single_calls(gpu, use_cpu, nan_pair):
    """Several wavelengths of one capture (one copy and crop, one plan re-staged per
    wavelength) give exactly the single-wavelength volumes."""
    cap = scenes.render(scenes.config_scene(1), 12, 2048, t0=2.0, seed=17)
    args = list(cap.nlos_args((10, 12, 9)))
    if nan_pair:
        args[1] = args[1].copy()
        args[1][4, 2, 0] = np.nan
    wls = [0.004, 0.005, 0.04, 0.0075]
    multi = gpu.nlosbpy.phasor_multi(*args, wls, assume_row_major=True, use_cpu=use_cpu)
    assert multi.shape == (4, 10, 12, 9) and multi.dtype == np.complex64
    for i, wl in enumerate(wls):
        one = gpu.nlosbpy.phasor(*args, wl, assume_row_major=True, use_cpu=use_cpu)
        assert np.array_equal(multi[i], one), wl
    assert np.abs(multi).max() > 0


@pytest.mark.parametrize("use_cpu", [False, True])
def test_nlosbpy_phasor_magnitude_and_phase(gpu, oracle, use_cpu):
    """output= "magnitude" / "phase" / "magnitude_phase": the complex volume's |v| (what the
    reference CLI writes, std::abs, main.cpp:162-166) and atan2(im, re), computed on the device
    in double and rounded once; against the same formulas in float64 numpy on the same call's
    complex volume (within 1 ulp), numpy's float32 abs and angle (within 2: each is within about
    one ulp of the exact value), and the magnitude against |oracle phasor_reconstruction| at the
    complex volume's tolerance."""

    def exact(c):
        re, im = c.real.astype(np.float64), c.imag.astype(np.float64)
        return np.sqrt(re * re + im * im).astype(f32), np.arctan2(im, re).astype(f32)

    cap = scenes.render(scenes.config_scene(1), 12, 2048, t0=2.0, seed=23)
    N = (10, 12, 9)
    args = list(cap.nlos_args(N))
    wls = [0.004, 0.04]
    cv = gpu.nlosbpy.phasor(*args, wls[0], assume_row_major=True, use_cpu=use_cpu)
    mag = gpu.nlosbpy.phasor(*args, wls[0], assume_row_major=True, use_cpu=use_cpu, output="magnitude")
    ph = gpu.nlosbpy.phasor(*args, wls[0], assume_row_major=True, use_cpu=use_cpu, output="phase")
    m2, p2 = gpu.nlosbpy.phasor(*args, wls[0], assume_row_major=True, use_cpu=use_cpu, output="magnitude_phase")
    assert mag.dtype == np.float32 and mag.shape == N and ph.shape == N and np.abs(cv).max() > 0
    em, ep = exact(cv)
    np.testing.assert_array_max_ulp(mag, em, maxulp=1)
    np.testing.assert_array_max_ulp(ph, ep, maxulp=1)
    np.testing.assert_array_max_ulp(mag, np.abs(cv), maxulp=2)
    np.testing.assert_array_max_ulp(ph, np.angle(cv).astype(f32), maxulp=2)
    assert np.array_equal(m2, mag) and np.array_equal(p2, ph)
    multi = gpu.nlosbpy.phasor_multi(*args, wls, assume_row_major=True, use_cpu=use_cpu, output="magnitude_phase")
    cmulti = gpu.nlosbpy.phasor_multi(*args, wls, assume_row_major=True, use_cpu=use_cpu)
    assert multi.shape == (2, 2) + N
    for i in range(2):
        em, ep = exact(cmulti[i])
        np.testing.assert_array_max_ulp(multi[i, 0], em, maxulp=1)
        np.testing.assert_array_max_ulp(multi[i, 1], ep, maxulp=1)
    if not use_cpu:
        cref, _ = oracle.phasor_pipeline(cap.transient, cap.camera_grid, cap.laser_grid,
                                         cap.camera_position, cap.laser_position, cap.t0, cap.deltaT,
                                         True, cap.volume_position, cap.volume_size, N, wls[0])
        assert rel_l2(mag, np.abs(cref)) <= 1e-5
    with pytest.raises(ValueError):
        gpu.nlosbpy.phasor(*args, wls[0], assume_row_major=True, output="real")


def test_nlosbpy_column_major_equals_row_major(gpu):
    cap = scenes.render(scenes.config_scene(1), 16, 2048, t0=2.0, seed=3)
    N = (12, 12, 12)
    a = gpu.nlosbpy.backproject(*cap.nlos_args(N), np.zeros(3), assume_row_major=True)
    cm = (np.asfortranarray(cap.transient).T, cap.camera_grid.T, cap.laser_grid.T)
    args = list(cap.nlos_args(N))
    args[0], args[1], args[2] = np.ascontiguousarray(cm[0]), np.ascontiguousarray(cm[1]), np.ascontiguousarray(cm[2])
    b = gpu.nlosbpy.backproject(*args, np.zeros(3), assume_row_major=False)
    assert np.array_equal(a, b)


def test_nlosbpy_use_cpu_semantics(gpu, oracle):
    cap = scenes.render(scenes.config_scene(1), 16, 2048, t0=2.0, seed=7)
    N = (16, 16, 16)
    vol = gpu.nlosbpy.backproject(*cap.nlos_args(N), np.zeros(3), assume_row_major=True,
                                  use_cpu=True)
    ref, _ = oracle.backproject_pipeline(cap.transient, cap.camera_grid, cap.laser_grid,
                                         cap.camera_position, cap.laser_position, cap.t0,
                                         cap.deltaT, True, cap.volume_position, cap.volume_size,
                                         N, use_cpu=True)
    assert rel_l2(vol, ref) <= 1e-5


def test_nlosbpy_exhaustive(gpu, oracle):
    cap = scenes.render(scenes.config_scene(2), 6, 2048, t0=2.0, confocal=False, laser_side=5,
                        seed=1002)
    N = (16, 16, 16)
    vol = gpu.nlosbpy.backproject(*cap.nlos_args(N), np.zeros(3), assume_row_major=True)
    ref, inter = oracle.backproject_pipeline(cap.transient, cap.camera_grid, cap.laser_grid,
                                             cap.camera_position, cap.laser_position, cap.t0,
                                             cap.deltaT, False, cap.volume_position,
                                             cap.volume_size, N)
    assert rel_l2(vol, ref) <= 1e-5


@pytest.mark.parametrize("use_cpu", [False, True])
def test_nlosbpy_phasor(gpu, oracle, use_cpu):
    cap = scenes.render(scenes.config_scene(1), 16, 2048, t0=2.0, seed=5)
    N = (12, 12, 12)
    args = list(cap.nlos_args(N))
    vol = gpu.nlosbpy.phasor(*args, 0.005, assume_row_major=True, use_cpu=use_cpu)
    ref, _ = oracle.phasor_pipeline(cap.transient, cap.camera_grid, cap.laser_grid,
                                    cap.camera_position, cap.laser_position, cap.t0, cap.deltaT,
                                    True, cap.volume_position, cap.volume_size, N, 0.005,
                                    use_cpu=use_cpu)
    assert vol.dtype == np.complex64
    assert rel_l2(vol, ref) <= 1e-5


# ----------------------------------------------------------------- edge cases
def test_edge_cases(gpu, oracle):
    pr = confocal_pairs(8)
    T = 1024
    H = coded_samples(len(pr), T)
    N = (6, 6, 6)
    zero, inc = grid_basis(N)
    # all-zero samples
    assert not gpu.backprojection(np.zeros_like(H), pr, N, zero, inc, 2.0, 0.001).any()
    # volume whose paths all fall beyond the chunk -> zeros
    assert not gpu.backprojection(H, pr, N, zero + f32(50), inc, 2.0, 0.001).any()
    # non-finite pairs contribute nothing
    bad = pr.copy()
    bad["laser_point"][3] = np.nan
    bad["camera_point"][3] = np.nan
    bad["camera_wall"][5] = np.inf
    keep = np.ones(len(pr), bool)
    keep[[3, 5]] = False
    vol = gpu.backprojection(H, bad, N, zero, inc, 2.0, 0.001)
    ref = oracle.a1_volume_f64acc(H[keep], pr[keep], N, zero, inc, 2.0, 0.001)
    assert np.array_equal(vol, ref.astype(f32))
    # T = 1 and a single pair
    v1 = gpu.backprojection(H[:1, :1], pr[:1], N, zero, inc, 2.0, 0.001)
    r1 = oracle.a1_volume_f64acc(H[:1, :1], pr[:1], N, zero, inc, 2.0, 0.001)
    assert np.array_equal(v1, r1.astype(f32))
    # a 1-voxel volume
    v2 = gpu.backprojection(H, pr, (1, 1, 1), zero, inc, 2.0, 0.001)
    r2 = oracle.a1_volume_f64acc(H, pr, (1, 1, 1), zero, inc, 2.0, 0.001)
    assert np.array_equal(v2, r2.astype(f32))


def test_reference_pair_truncation_flag(gpu, oracle):
    rng = np.random.default_rng(2)
    pr = confocal_pairs(18)          # 324 pairs -> the reference uses 256
    T = 1024
    H = coded_samples(len(pr), T)
    N = (8, 8, 8)
    zero, inc = grid_basis(N)
    vol = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001, reference_pair_truncation=True)
    ref = oracle.a1_volume(H, pr, N, zero, inc, 2.0, 0.001)
    assert np.array_equal(vol, ref)
    full = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001)
    assert not np.array_equal(full, vol)


def test_determinism_and_plan_reuse(gpu):
    pr = confocal_pairs(32)
    rng = np.random.default_rng(0)
    H = rng.poisson(3.0, (len(pr), 2048)).astype(f32)
    N = (32, 32, 32)
    zero, inc = grid_basis(N)
    a = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001)
    with gpu.Plan(H, pr, N, zero, inc, 2.0, 0.001) as plan:
        plan.execute()
        b = plan.volume()
        plan.execute()
        plan.execute()
        c = plan.volume()
        info = plan.info()
        t = plan.benchmark(1, 3)
    assert np.array_equal(a, b) and np.array_equal(b, c)
    assert info["path"] == "confocal" and info["pairs_used"] == len(pr)
    assert t["gather_kernel_ms"] > 0 and t["table_kernel_ms"] > 0


def test_plan_execute_on_caller_stream(gpu, hip_streams):
    """An execution enqueued on a caller's HIP stream is ordered after the plan's earlier work
    and is waited for by the plan's read, synchronize and destroy calls."""
    pr = confocal_pairs(32)
    rng = np.random.default_rng(3)
    H = rng.poisson(3.0, (len(pr), 2048)).astype(f32)
    N = (32, 32, 32)
    zero, inc = grid_basis(N)
    ref = gpu.backprojection(H, pr, N, zero, inc, 2.0, 0.001)
    s1, s2 = hip_streams
    with gpu.Plan(H, pr, N, zero, inc, 2.0, 0.001) as plan:
        plan.execute(s1)
        plan.execute(s2)          # after the execution on s1
        assert np.array_equal(plan.volume(), ref)
        plan.execute()                        # own stream, after the one on s2
        plan.execute(s1)
        plan.synchronize()
        assert np.array_equal(plan.volume(), ref)
        plan.execute(s2)          # destroyed while possibly still running
