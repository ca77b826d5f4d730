"""Batched uneven captures (one persistent work-queue launch replayed from a hipGraph) against
each capture reconstructed alone and against the oracle."""
import numpy as np
import pytest

from backprojector_amd import scenes
from backprojector_amd._native import BpxError

pytestmark = pytest.mark.gpu
f32 = np.float32


def capture(rng, side, T, N, coded=False):
    hidden = np.concatenate([scenes.points(rng, 3), scenes.patch(rng, size=0.2, spacing=0.01)])
    cap = scenes.render(hidden, side, T, 0.0, 4.096 / T, True, seed=int(rng.integers(1 << 30)))
    pr = scenes.wall_pairs(cap.camera_grid, cap.camera_grid, True)
    H = cap.transient.reshape(len(pr), -1).astype(f32)
    if coded:  # integer sample codes: every sum is exact, so the oracle must match exactly
        b = np.arange(T)[None, :]
        p = np.arange(len(pr))[:, None]
        H = (((7 * b + 13 * p) % 61) + 1).astype(f32)
    Nv = (N, N + 3, N - 2)            # uneven, not multiples of the tile edges
    zero = np.array([-0.5, -0.5, 0.3], f32)
    inc = np.diag([f32(1) / f32(n - 1) for n in Nv]).astype(f32).ravel()
    return (H, pr, Nv, zero, inc, 0.0, 4.096 / T)


def uneven(seed=0):
    rng = np.random.default_rng(seed)
    spec = [(8, 1024, 12), (16, 2048, 20), (12, 512, 9), (24, 1024, 16), (16, 4096, 24), (8, 512, 30)]
    return [capture(rng, *s) for s in spec]


def test_batch_equals_single_captures(gpu):
    caps = uneven()
    alone = [gpu.backprojection(*c) for c in caps]
    with gpu.Batch(caps) as b:
        assert len(b) == len(caps)
        for rep in range(3):   # graph replays with eager work and synchronisation between them
            b.execute()
            b.synchronize()
            for i, ref in enumerate(alone):
                assert np.array_equal(b.volume(i), ref), (rep, i)
            gpu.backprojection(*caps[0])
        lat = b.latencies()
        assert np.all(lat > 0) and np.all(np.isfinite(lat))


def test_batch_lds_staged_queue(gpu, oracle):
    """Captures whose bin windows fit the LDS segment (4 mm bins, fine enough voxel pitch) run on
    the LDS-staged persistent queue, the others on the texture-path queue, in the same batch;
    every volume equals its capture reconstructed alone, and the coded ones the oracle."""
    caps = mixed()
    alone = [gpu.backprojection(*c) for c in caps]
    with gpu.Batch(caps) as b:
        cfg = b.config()
        assert 3 <= cfg["lds_captures"] < len(caps), cfg
        # 16 x 16 x 8 and 16^3 LDS tiles, queue orders, one launch or two concurrent ones
        for depth, order, launches in ((8, 0, 1), (16, 2, 1), (8, 2, 1), (8, 0, 2), (16, 1, 2), (8, 0, 3), (8, 1, 3)):
            b.set_config(lane_voxels=2, groups=0, order=order, lds_depth=depth, launches=launches)
            assert b.config()["launches"] == launches
            for rep in range(2):
                b.execute()
                b.synchronize()
                for i, ref in enumerate(alone):
                    assert np.array_equal(b.volume(i), ref), (depth, launches, rep, i)
            assert np.all(b.latencies() > 0)
    for i in (0, 2, 4):
        H, pr, Nv, zero, inc, t0, dT = caps[i]
        ref = oracle.a1_volume_f64acc(H, pr, Nv, zero, inc, t0, dT)
        assert np.array_equal(alone[i], ref.astype(np.float32)), i


def mixed(seed=7):
    """Captures of which some fit the LDS-staged queue and the others go to the texture path."""
    rng = np.random.default_rng(seed)
    spec = [(8, 1024, 24), (12, 1024, 30), (16, 2048, 20), (8, 1024, 40), (10, 1024, 33), (6, 512, 12)]
    return [capture(rng, *sp, coded=(i % 2 == 0)) for i, sp in enumerate(spec)]


@pytest.mark.parametrize("launches", [1, 2, 3])
def test_batch_execute_on_caller_stream(gpu, hip_streams, launches):
    """The batch graph (with two launches: its fork/join branch of two concurrent queues)
    replayed on two caller streams."""
    caps = mixed()
    alone = [gpu.backprojection(*c) for c in caps]
    s1, s2 = hip_streams
    with gpu.Batch(caps) as b:
        b.set_config(lane_voxels=2, groups=0, order=0, launches=launches)
        cfg = b.config()
        assert 1 <= cfg["lds_captures"] <= len(caps) - 1, cfg
        for rep in range(2):
            b.execute(s1)
            b.execute(s2)
            for i, ref in enumerate(alone):   # the read waits for the execution on s2
                assert np.array_equal(b.volume(i), ref), (rep, i)
            assert np.all(b.latencies() > 0)
        b.execute(s1)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_batch_16bit_storage_equals_single_captures(gpu, storage):
    """fp16 / bf16 rows with a per-capture scale: each batched volume equals the capture
    reconstructed alone with the same storage, for both lane shapes."""
    caps = uneven(2)
    alone = [gpu.backprojection(*c, storage=storage) for c in caps]
    with gpu.Batch(caps, storage=storage) as b:
        for k in (2, 4):
            b.set_config(lane_voxels=k, groups=0, order=0)
            b.execute()
            b.synchronize()
            for i, ref in enumerate(alone):
                assert np.array_equal(b.volume(i), ref), (k, i)
    f32_alone = gpu.backprojection(*caps[4])   # integer counts: exact in fp16, rounded in bf16
    assert np.linalg.norm(alone[4] - f32_alone) <= 2e-3 * np.linalg.norm(f32_alone)


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_batch_complex_captures_equal_single_captures(gpu, storage):
    """Phasor-filtered (complex) captures in one batch: complex volumes, each equal to the
    capture reconstructed alone (complex64 or complex bf16 rows)."""
    caps = []
    for c in uneven(3)[:4]:
        H = gpu.phasor_pulse(c[0], 0.04, c[6])
        caps.append((H,) + tuple(c[1:]))
    alone = [gpu.backprojection(*c, storage=storage) for c in caps]
    with gpu.Batch(caps, storage=storage) as b:
        for k in (2, 4):
            b.set_config(lane_voxels=k, groups=0, order=0)
            b.execute()
            b.synchronize()
            for i, ref in enumerate(alone):
                v = b.volume(i)
                assert v.dtype == np.complex64 and np.array_equal(v, ref), (k, i)
    with pytest.raises(ValueError):
        gpu.Batch([caps[0], uneven(3)[1]])


@pytest.mark.parametrize("cfg", [dict(lane_voxels=4, groups=7, order=1),
                                 dict(lane_voxels=2, groups=1, order=0),
                                 dict(lane_voxels=4, groups=0, order=0),
                                 dict(lane_voxels=2, groups=0, order=1, lds_depth=16),
                                 dict(lane_voxels=4, groups=3, order=0, lds_depth=16),
                                 dict(lane_voxels=2, groups=0, order=2),
                                 dict(lane_voxels=4, groups=5, order=2, lds_depth=16),
                                 dict(lane_voxels=2, groups=0, order=0, launches=2),
                                 dict(lane_voxels=4, groups=3, order=1, lds_depth=16, launches=2),
                                 dict(lane_voxels=2, groups=0, order=0, launches=3),
                                 dict(lane_voxels=4, groups=2, order=1, launches=3)])
def test_batch_configs_identical(gpu, cfg):
    caps = uneven(1)
    alone = [gpu.backprojection(*c) for c in caps]
    with gpu.Batch(caps) as b:
        b.set_config(**cfg)
        got = b.config()
        assert got["lane_voxels"] == cfg["lane_voxels"] and got["order"] == cfg["order"]
        assert got["lds_depth"] == cfg.get("lds_depth", 8) and got["launches"] == cfg.get("launches", 1)
        b.execute()
        for i, ref in enumerate(alone):
            assert np.array_equal(b.volume(i), ref), (cfg, i)


def test_batch_tune_and_benchmark(gpu):
    caps = uneven(2)
    alone = [gpu.backprojection(*c) for c in caps]
    with gpu.Batch(caps) as b:
        cfg = b.tune(reps=1)
        assert cfg["lane_voxels"] in (2, 4) and cfg["groups"] >= 1
        t = b.benchmark(1, 2)
        P = [len(c[1]) for c in caps]
        assert t["updates"] == sum(p * int(np.prod(c[2])) for p, c in zip(P, caps))
        assert t["per_exec_ms"] > 0 and t["updates_per_s"] > 0
        for i, ref in enumerate(alone):
            assert np.array_equal(b.volume(i), ref)


def test_batch_coded_matches_oracle(gpu, oracle):
    rng = np.random.default_rng(5)
    caps = [capture(rng, 8, 512, 10, coded=True), capture(rng, 6, 256, 7, coded=True)]
    with gpu.Batch(caps) as b:
        b.execute()
        for i, c in enumerate(caps):
            ref = oracle.a1_volume_f64acc(*c)
            assert np.array_equal(b.volume(i), ref.astype(f32))


def test_batch_edge_cases(gpu):
    caps = uneven(3)[:2]
    H, pr, N, zero, inc, t0, dT = caps[0]
    empty = (np.zeros(0, f32), pr[:0], N, zero, inc, t0, dT)
    with gpu.Batch([empty, caps[1]]) as b:
        b.execute()
        assert not b.volume(0).any()
        assert np.array_equal(b.volume(1), gpu.backprojection(*caps[1]))
    with gpu.Batch([]) as b:
        assert len(b) == 0
        b.execute()
        b.synchronize()
    # a non-confocal capture is rejected
    ex = pr[:4].copy()
    ex["laser_point"] = ex["camera_point"][::-1]
    with pytest.raises(BpxError):
        gpu.Batch([(H[:4], ex, N, zero, inc, t0, dT)])
    with pytest.raises(BpxError):
        with gpu.Batch(caps) as b:
            b.set_config(lane_voxels=3)


def _codes(P, T, salt):
    """Integer sample codes 1..61 for [P][T]: every voxel sum is an integer below 2^24, so any
    summation order is exact and one wrong bin changes the sum."""
    cb = (7 * np.arange(T, dtype=np.int64) + salt) % 61
    cp = (13 * np.arange(P, dtype=np.int64)) % 61
    return ((cp[:, None] + cb[None, :]) % 61 + 1).astype(f32)


@pytest.mark.timeout(900)
def test_c4_64_uneven_captures(gpu, oracle):
    """The 64 seeded uneven captures (SeedSequence(1004).spawn(64): wall side 16-128, T 1024-4096,
    volume side 16-96), i.e. one call_cuda_backprojection per capture in the reference
    (backproject_cuda.cu:370-500), as one autotuned persistent batch.  Every volume equals its
    capture's own plan exactly; on integer-coded samples of the same geometry, 256 sampled
    voxels of 8 captures (the largest included) equal the oracle's fp32 formula exactly."""
    from backprojector_amd import proposals
    caps = scenes.batch_captures(64, 1004)
    args = [proposals.plan_args(c, N) for c, N in caps]
    del caps
    print(f"C4: {len(args)} captures prepared", flush=True)
    with gpu.Batch(args) as b:
        cfg = b.tune(reps=1)
        assert 1 <= cfg["lds_captures"] < len(args), cfg
        b.execute()
        b.synchronize()
        lat = b.latencies()
        assert np.all(lat > 0) and np.all(np.isfinite(lat))
        for i, a in enumerate(args):
            with gpu.Plan(*a) as pl:
                pl.execute()
                assert np.array_equal(b.volume(i), pl.volume()), i
    print("C4: batch volumes equal single plans", flush=True)
    work = [len(a[1]) * int(np.prod(a[2])) for a in args]
    largest = int(np.argmax(work))
    assert len(args[largest][1]) == 128 * 128 and args[largest][2] == (96, 96, 96)
    check = [largest] + [i for i in np.random.default_rng(1004).permutation(len(args)) if i != largest][:7]
    coded = [(_codes(len(a[1]), a[0].shape[1], i),) + tuple(a[1:]) for i, a in enumerate(args)]
    del args
    with gpu.Batch(coded) as b:
        b.tune(reps=1)
        b.execute()
        rng = np.random.default_rng(4)
        for i in check:
            H, pr, N, zero, inc, t0, dT = coded[i]
            vox = np.stack([rng.integers(0, n, 256) for n in N], 1).astype(np.uint32)
            ref = oracle.a1_voxels_f64acc(H, pr, vox, zero, inc, t0, dT)
            got = b.volume(i)[vox[:, 0], vox[:, 1], vox[:, 2]]
            assert ref.max() > 0 and np.array_equal(got.astype(np.float64), ref), i
