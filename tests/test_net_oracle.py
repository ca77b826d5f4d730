"""Pins the CPU net oracle's backward pass with central finite differences (float64, no
bfloat16 rounding), and its convolution against a direct loop."""
import numpy as np

from oracle import net_oracle as no


def _setup(seed=0, dims=(3, 4, 5)):
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.3, no.NPARAMS)
    W, b = no.split(params)
    W[0, :, :, no.NIN:] = 0.0
    X = np.zeros(dims + (16,))
    X[..., :no.NIN] = rng.normal(0, 1, dims + (no.NIN,))
    labels = (rng.random(dims + (8,)) < 0.3).astype(np.uint8)
    mask = (rng.random(dims) < 0.8).astype(np.uint8)
    return params, X, labels, mask


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(3, 2, 4, 16))
    W = rng.normal(size=(27, 16, 16))
    Y = no.conv3(X, W)
    n0, n1, n2, _ = X.shape
    for (x, y, z) in [(0, 0, 0), (2, 1, 3), (1, 0, 2)]:
        acc = np.zeros(16)
        for t in range(27):
            dx, dy, dz = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
            xs, ys, zs = x + dx, y + dy, z + dz
            if 0 <= xs < n0 and 0 <= ys < n1 and 0 <= zs < n2:
                acc += W[t] @ X[xs, ys, zs]
        assert np.allclose(Y[x, y, z], acc)


def test_oracle_gradients_finite_differences():
    params, X, labels, mask = _setup()
    loss, g, _ = no.loss_and_grads(params, X, labels, mask, pos_weight=3.0, emulate=False)
    rng = np.random.default_rng(2)
    # probe parameters of every layer, weights and biases (skip padded zero-input columns)
    W, _ = no.split(np.arange(no.NPARAMS))
    idx = [int(W[k, t, o, i]) for k in range(4) for (t, o, i) in
           [(13, 1, 2), (0, 5, 3), (26, 7, 1)]] + [no.LAYERS * no.KW + 16 * k + 3 for k in range(4)]
    idx += list(rng.integers(no.KW, 3 * no.KW, 6))
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        l1 = no.loss_and_grads(p1, X, labels, mask, 3.0, emulate=False)[0]
        l2 = no.loss_and_grads(p2, X, labels, mask, 3.0, emulate=False)[0]
        num = (l1 - l2) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (i, num, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0])
    r = no.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie 1 + 2^-8 rounds to even (1.0)
    assert r[2] == 1.0078125                    # above the tie rounds up
    assert abs(r[3] + 3.140625) < 1e-12
