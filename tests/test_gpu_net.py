"""Octree proposal net on the GPU (hand-written MFMA convolution forward / backward) against the
CPU net oracle, exact layout checks with permutation / shift weights, and on-GPU training."""
import numpy as np
import pytest

from backprojector_amd import scenes
from oracle import net_oracle as no

pytestmark = pytest.mark.gpu
f32 = np.float32


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def rand_input(rng, dims):
    X = np.zeros(tuple(dims) + (16,), f32)
    X[..., :no.NIN] = rng.normal(0, 1, tuple(dims) + (no.NIN,))
    return X


def random_net(gpu, seed):
    net = gpu.Net(seed)
    p = net.params()
    rng = np.random.default_rng(seed + 100)
    W, b = no.split(np.arange(no.NPARAMS))
    p[b.ravel().astype(np.int64)] = rng.normal(0, 0.1, b.size)   # non-zero biases
    net.set_params(p)
    return net


def exact_params(taps, perms):
    """Layer k: W[k][taps[k]][o][perms[k][o]] = 1 (a shift composed with a channel permutation)."""
    p = np.zeros(no.NPARAMS, f32)
    W, _ = no.split(np.arange(no.NPARAMS))
    for k in range(4):
        for o, i in enumerate(perms[k]):
            p[int(W[k, taps[k], o, i])] = 1.0
    return p


def test_layout_identity_shift_and_permutation(gpu):
    # tap index (dx+1)*9 + (dy+1)*3 + (dz+1): layer 0 shifts x, 1 shifts y, 2 shifts z, 3 identity;
    # layer 1 permutes channels 0..3 -> out o reads in perm[o]
    taps = [22, 16, 14, 13]
    ident = list(range(16))
    perm = [2, 0, 3, 1] + list(range(4, 16))
    params = exact_params(taps, [list(range(4)), perm, ident, ident[:8]])
    net = gpu.Net(0, params)
    dims = (5, 6, 37)                      # z not a multiple of 16; every boundary case
    rng = np.random.default_rng(1)
    X = np.zeros(dims + (16,), f32)
    X[..., :4] = rng.integers(1, 64, dims + (4,)) / 8.0   # positive, exact in bfloat16
    Y = net.forward(X)
    want = np.zeros(dims + (16,), f32)
    sh = X[1:, 1:, 1:, :4]                 # out[x, y, z] = X[x+1, y+1, z+1]
    want[:-1, :-1, :-1, :4] = sh[..., perm[:4]]
    assert np.array_equal(Y, want)


def test_forward_matches_oracle(gpu):
    net = random_net(gpu, 7)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(2)
    for dims in [(4, 5, 20), (3, 3, 16), (2, 7, 33)]:
        X = rand_input(rng, dims)
        Y = net.forward(X)
        ref, _, _ = no.forward(p, X.astype(np.float64), emulate=True)
        assert np.all(Y[..., no.NOUT:] == 0)
        # same bfloat16 rounding points; only the fp32 summation order differs
        assert rel_l2(Y[..., :no.NOUT], ref[..., :no.NOUT]) < 1e-2, dims


def test_gradients_match_oracle(gpu):
    net = random_net(gpu, 11)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(3)
    dims = (4, 6, 18)
    X = rand_input(rng, dims)
    labels = (rng.random(dims + (8,)) < 0.3).astype(np.uint8)
    mask = (rng.random(dims) < 0.8).astype(np.uint8)
    loss, g = net.gradients(X, labels, mask, pos_weight=3.0)
    rloss, rg, _ = no.loss_and_grads(p, X.astype(np.float64), labels, mask, 3.0, emulate=True)
    assert abs(loss - rloss) <= 2e-3 * abs(rloss)
    W, b = no.split(np.arange(no.NPARAMS))
    for k in range(4):
        iw = W[k].ravel().astype(np.int64)
        ib = b[k].astype(np.int64)
        assert rel_l2(g[iw], rg[iw]) < 3e-2, ("weights", k, rel_l2(g[iw], rg[iw]))
        assert rel_l2(g[ib], rg[ib]) < 3e-2, ("bias", k, rel_l2(g[ib], rg[ib]))
    # the padded input channels of layer 0 and padded outputs of layer 3 get no gradient
    assert np.all(g[W[0][..., no.NIN:].ravel().astype(np.int64)] == 0)
    assert np.all(g[W[3][:, no.NOUT:, :].ravel().astype(np.int64)] == 0)


def test_gradient_descends_loss(gpu):
    """A few plain gradient steps taken with the GPU gradients lower the oracle's loss."""
    net = random_net(gpu, 5)
    rng = np.random.default_rng(4)
    dims = (4, 4, 16)
    X = rand_input(rng, dims)
    labels = (X[..., :1] > 0).repeat(8, -1).astype(np.uint8)
    p = net.params().astype(np.float64)
    l0 = no.loss_and_grads(p, X.astype(np.float64), labels, None, 1.0)[0]
    for _ in range(5):
        _, g = net.gradients(X, labels, None, 1.0)
        p = p - 0.05 * g / np.linalg.norm(g)
        net.set_params(p.astype(f32))
    l1 = no.loss_and_grads(net.params().astype(np.float64), X.astype(np.float64), labels, None,
                           1.0)[0]
    assert l1 < l0, (l0, l1)


def small_scene_plans(gpu, seed, N=64, side=16, levels=3, threshold=0.0):
    rng = np.random.default_rng(seed)
    hidden = np.concatenate([scenes.points(rng, 3), scenes.patch(rng, size=0.2, spacing=0.005)])
    cap = scenes.render(hidden, side, 4096, seed=seed)
    pr = scenes.wall_pairs(cap.camera_grid, cap.camera_grid, True)
    H = cap.transient.reshape(len(pr), -1)
    Nv = (N, N, N)
    zero = np.array([-0.5, -0.5, 0.3], f32)
    inc = np.diag([f32(1) / f32(N - 1)] * 3).astype(f32).ravel()
    args = (H, pr, Nv, zero, inc, 0.0, 0.001)
    return (args, gpu.Plan.octree(*args, levels=levels, threshold=threshold, bounds_only=True),
            gpu.Plan(*args))


def test_train_octree_lowers_loss(gpu):
    net = gpu.Net(3)
    data = [small_scene_plans(gpu, s)[1:] for s in range(3)]
    for op, dp in data:
        op.execute()
        dp.execute()
        op.synchronize()
        dp.synchronize()
        st = op.octree_stats()
        assert st["active"][:3] == [8 ** 3, 16 ** 3, 32 ** 3] and st["active"][3] == 0
    losses = []
    for it in range(60):
        op, dp = data[it % 3]
        losses.append(net.train_octree(op, dp, occupied_fraction=0.25, pos_weight=4.0, lr=3e-3))
    assert all(np.isfinite(losses))
    first, last = np.mean(losses[:6]), np.mean(losses[-6:])
    print(f"training loss {first:.4f} -> {last:.4f}")
    assert last < 0.8 * first


def test_train_on_policy_partial_levels(gpu):
    """Training from an octree that evaluated only part of each level (bound rule + the net
    itself attached): the loss covers the evaluated blocks and stays finite."""
    net = gpu.Net(4)
    _, op, dp = small_scene_plans(gpu, 5, threshold=0.25)
    op.attach_net(net)
    dp.execute()
    losses = []
    for _ in range(8):
        op.execute()
        st = op.octree_stats()
        assert st["active"][0] == 8 ** 3 and st["active"][1] <= 16 ** 3 and st["active"][3] == 0
        losses.append(net.train_octree(op, dp, occupied_fraction=0.25, pos_weight=4.0, lr=3e-3))
    assert all(np.isfinite(losses)) and all(x > 0 for x in losses)


def bias_only_net(gpu, value):
    p = np.zeros(no.NPARAMS, f32)
    _, b = no.split(np.arange(no.NPARAMS))
    p[b[3, :no.NOUT].astype(np.int64)] = value
    return gpu.Net(0, p)


def test_attached_net_narrows_and_stays_exact(gpu):
    """The net can only veto children: always-yes reproduces the bound rule, always-no stops
    after level 0, and any net computes voxels exactly."""
    args, _, dense_plan = small_scene_plans(gpu, 9)
    dense_plan.execute()
    dense = dense_plan.volume()
    with gpu.Plan.octree(*args, levels=3, threshold=0.25) as plain_oct:
        plain_oct.execute()
        v0 = plain_oct.volume()
        a0 = plain_oct.octree_stats()["active"]
    with gpu.Plan.octree(*args, levels=3, threshold=0.25, proposal_threshold=0.5) as lo:
        lo.attach_net(bias_only_net(gpu, 40.0))
        lo.execute()
        assert np.array_equal(lo.volume(), v0) and lo.octree_stats()["active"] == a0
        lo.attach_net(bias_only_net(gpu, -40.0))
        lo.execute()
        assert lo.octree_stats()["active"] == [8 ** 3, 0, 0, 0]
        assert not lo.volume().any()
        # the bound override: an always-no net with keep ratio ~0 refines every block the bound
        # rule refines; with a keep ratio above every bound it vetoes everything again
        lo.set_octree_options(0.25, 0.5, keep_bound_ratio=1e-30)
        lo.execute()
        assert np.array_equal(lo.volume(), v0) and lo.octree_stats()["active"] == a0
        lo.set_octree_options(0.25, 0.5, keep_bound_ratio=1e30)
        lo.execute()
        assert lo.octree_stats()["active"] == [8 ** 3, 0, 0, 0]
        lo.set_octree_options(0.25, 0.5)
        lo.attach_net(random_net(gpu, 21))
        lo.execute()
        v1 = lo.volume()
        lo.attach_net(None)
        lo.execute()
        assert np.array_equal(lo.volume(), v0)
    c1 = v1 != 0
    assert np.array_equal(v1[c1], dense[c1])
