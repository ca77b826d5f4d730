"""Octree refinement on the GPU against the dense single-tap reconstruction and the oracle.

Properties (samples >= 0): every voxel the octree computes equals the dense value bit for
bit; every voxel whose dense value is >= threshold * max is computed (the per-level bound is
sound); voxels not computed are 0; threshold 0 computes everything.  Integer sample codes
make the dense/oracle comparison exact."""
import numpy as np
import pytest

from backprojector_amd import scenes

pytestmark = pytest.mark.gpu
f32 = np.float32


def confocal_setup(side=24, N=32, T=4096, seed=3, coded=False):
    cap = scenes.render(scenes.config_scene(1, np.random.default_rng(seed)), side, T, t0=0.0,
                        seed=seed)
    cg = cap.camera_grid
    pr = scenes.wall_pairs(cg, cg, True)
    H = cap.transient.reshape(len(pr), -1)
    if coded:
        b = np.arange(T)[None, :]
        p = np.arange(len(pr))[:, None]
        H = (((7 * b + 13 * p) % 61) + 1).astype(f32)
    Nv = (N, N, N)
    zero = np.array([-0.5, -0.5, 0.3], f32)
    inc = np.diag([f32(1) / f32(N - 1)] * 3).astype(f32).ravel()
    return H, pr, Nv, zero, inc


@pytest.mark.parametrize("levels,tau", [(2, 0.25), (3, 0.5), (4, 0.3)])
def test_octree_exact_where_computed_and_sound(gpu, levels, tau):
    H, pr, N, zero, inc = confocal_setup(N=32 if levels < 4 else 64)
    dense = gpu.backprojection(H, pr, N, zero, inc, 0.0, 0.001)
    vol, st = gpu.octree_backprojection(H, pr, N, zero, inc, 0.0, 0.001, levels=levels,
                                        threshold=tau)
    computed = vol != 0
    assert np.array_equal(vol[computed], dense[computed])           # exact where computed
    must = dense >= tau * dense.max()
    assert np.all(computed[must])                                     # sound pruning
    assert st["active"][levels] == st["fine_updates"] // len(pr)
    assert st["active"][levels] < np.prod(N)                          # something pruned
    rm = st["running_max"]
    assert all(a <= b for a, b in zip(rm, rm[1:])) and rm[-1] <= dense.max() * (1 + 1e-6)
    assert np.allclose(st["threshold"], [tau * m for m in rm])
    print(f"levels {levels} tau {tau}: fine voxels {st['active'][levels]} of {np.prod(N)}, "
          f"active per level {st['active']}")


def test_octree_threshold_zero_is_dense(gpu, oracle):
    H, pr, N, zero, inc = confocal_setup(side=16, N=32, T=2048, coded=True)
    vol, st = gpu.octree_backprojection(H, pr, N, zero, inc, 0.0, 0.001, levels=3, threshold=0.0)
    ref = oracle.a1_volume_f64acc(H, pr, N, zero, inc, 0.0, 0.001)
    assert np.array_equal(vol, ref.astype(f32))
    assert st["active"][3] == np.prod(N)


def test_octree_large_level_two_blocks_per_lane(gpu):
    """A level of 64^3 = 262144 blocks runs the unsegmented two-blocks-per-lane bound kernel;
    threshold 0 must still reproduce the dense volume bit for bit (LDS window-max build too)."""
    H, pr, N, zero, inc = confocal_setup(side=16, N=128, T=4096)
    dense = gpu.backprojection(H, pr, N, zero, inc, 0.0, 0.001)
    vol, st = gpu.octree_backprojection(H, pr, N, zero, inc, 0.0, 0.001, levels=2, threshold=0.0)
    assert st["active"] == [32 ** 3, 64 ** 3, 128 ** 3]
    assert np.array_equal(vol, dense)
    vol2, st2 = gpu.octree_backprojection(H, pr, N, zero, inc, 0.0, 0.001, levels=2,
                                          threshold=0.25)
    computed = vol2 != 0
    assert np.array_equal(vol2[computed], dense[computed])
    assert np.all(computed[dense >= 0.25 * dense.max()])


def test_octree_plan_reuse_and_rejects(gpu):
    H, pr, N, zero, inc = confocal_setup(side=16, N=32)
    with gpu.Plan.octree(H, pr, N, zero, inc, 0.0, 0.001, levels=2, threshold=0.3) as plan:
        plan.execute()
        a = plan.volume()
        plan.execute()
        b = plan.volume()
        st = plan.octree_stats()
    assert np.array_equal(a, b) and st["levels"] == 2
    with pytest.raises(gpu.InvalidArgument):
        gpu.octree_backprojection(H, pr, (30, 32, 32), zero, inc, 0.0, 0.001, levels=2)
    ex = pr.copy()
    ex["camera_point"][:, 0] += 0.01
    with pytest.raises(gpu.BpxError):
        gpu.octree_backprojection(H, ex, N, zero, inc, 0.0, 0.001)


def test_octree_pair_stride_and_full_resolution_fallback(gpu):
    """Coarse-level pair subset and the low-confidence fallback, driven by a net whose weights are
    all zero (every child's probability is exactly 0.5): computed voxels stay bit-identical to the
    dense reconstruction whatever the options; the subset is used only with a net (the bound rule
    alone keeps every pair); with the probability band (0.1, 0.9] every child of a level-0 block
    that passes the bound takes the full-resolution fallback at once (all 512 of its fine voxels,
    exactly), so with every pair the occupied voxels are all computed, as by the sound rule."""
    H, pr, N, zero, inc = confocal_setup(side=16, N=64, T=4096, seed=5)
    dense = gpu.backprojection(H, pr, N, zero, inc, 0.0, 0.001)
    must = dense >= 0.25 * dense.max()
    net = gpu.Net(seed=0, params=np.zeros(gpu.Net.param_count(), np.float32))
    with gpu.Plan.octree(H, pr, N, zero, inc, 0.0, 0.001, levels=4, threshold=0.25) as op:
        op.set_octree_options(0.25, 0.5, coarse_pair_stride=2)      # no net: every pair
        op.execute()
        st, vol = op.octree_stats(), op.volume()
        assert st["coarse_pairs"] == len(pr) and st["fallback_voxels"] == 0
        assert np.all((vol != 0)[must]) and np.array_equal(vol[vol != 0], dense[vol != 0])
        bound_children = st["active"][1]      # children of level-0 blocks that pass the bound
        op.attach_net(net)
        # every child takes the fallback at level 0
        op.set_octree_options(0.25, 0.9, fallback_threshold=0.1, coarse_pair_stride=1)
        op.execute()
        st, vol = op.octree_stats(), op.volume()
        computed = vol != 0
        assert st["active"][1:4] == [0, 0, 0]
        assert st["fallback_voxels"] == bound_children * 8 ** 3 == st["active"][4]
        assert np.all(computed[must]) and np.array_equal(vol[computed], dense[computed])
        # every child refined (0.5 > 0.4), features from a quarter of the pairs
        op.set_octree_options(0.25, 0.4, coarse_pair_stride=2)
        op.execute()
        st, vol = op.octree_stats(), op.volume()
        computed = vol != 0
        assert st["coarse_pairs"] == 64 and st["fallback_voxels"] == 0 and st["active"][4] > 0
        assert np.array_equal(vol[computed], dense[computed])
        # every child vetoed (0.5 <= 0.6), no fallback band: nothing below level 0
        op.set_octree_options(0.25, 0.6, coarse_pair_stride=4)
        op.execute()
        st, vol = op.octree_stats(), op.volume()
        assert st["coarse_pairs"] == 16 and st["active"][1:] == [0, 0, 0, 0] and not vol.any()
    net.close()
