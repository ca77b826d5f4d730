"""CPU oracle of the octree proposal net -- TEST INFRASTRUCTURE ONLY (tests/ may import it).

The learned component has no reference counterpart (the reference's octree is the broken
backproject.hpp:264-335 / backproject_cuda.cu:183-367).  This restates, in float64 numpy, the
net the GPU runs: four 3x3x3 convolutions over a level grid, 16 channels, zero padding, ReLU
between layers, 8 output logits, weighted binary cross-entropy.  With emulate=True it rounds to
bfloat16 exactly where the GPU stores bfloat16 (input, weights, hidden activations, the logit
gradient and the hidden-layer gradients), so the comparison isolates summation order.
Its own backward pass is pinned by central finite differences (tests/test_net_oracle.py).

Parameter layout (same as libbpx): W[4][27][16 out][16 in] then b[4][16]; tap
t = (dx+1)*9 + (dy+1)*3 + (dz+1); grids [n0][n1][n2][16], z fastest.
"""
from __future__ import annotations

import numpy as np

LAYERS, KW, NIN, NOUT = 4, 27 * 16 * 16, 5, 8
NPARAMS = LAYERS * (KW + 16)


def bf16(x):
    """Round to bfloat16 (round to nearest even), returned as float64."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split(params):
    p = np.asarray(params, np.float64)
    W = p[:LAYERS * KW].reshape(LAYERS, 27, 16, 16)
    b = p[LAYERS * KW:].reshape(LAYERS, 16)
    return W, b


def _shifts():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def conv3(X, W):
    """Y[v][o] = sum_t sum_i X[v + off(t)][i] W[t][o][i], zero padding."""
    n0, n1, n2, _ = X.shape
    Xp = np.zeros((n0 + 2, n1 + 2, n2 + 2, 16))
    Xp[1:-1, 1:-1, 1:-1] = X
    Y = np.zeros((n0, n1, n2, 16))
    for t, dx, dy, dz in _shifts():
        Y += Xp[1 + dx:1 + dx + n0, 1 + dy:1 + dy + n1, 1 + dz:1 + dz + n2] @ W[t].T
    return Y


def conv3_dx(dZ, W):
    """dX[v][i] = sum_t sum_o dZ[v - off(t)][o] W[t][o][i]."""
    n0, n1, n2, _ = dZ.shape
    Dp = np.zeros((n0 + 2, n1 + 2, n2 + 2, 16))
    Dp[1:-1, 1:-1, 1:-1] = dZ
    dX = np.zeros((n0, n1, n2, 16))
    for t, dx, dy, dz in _shifts():
        dX += Dp[1 - dx:1 - dx + n0, 1 - dy:1 - dy + n1, 1 - dz:1 - dz + n2] @ W[t]
    return dX


def conv3_dw(dZ, X):
    """dW[t][o][i] = sum_v dZ[v][o] X[v + off(t)][i]."""
    n0, n1, n2, _ = X.shape
    Xp = np.zeros((n0 + 2, n1 + 2, n2 + 2, 16))
    Xp[1:-1, 1:-1, 1:-1] = X
    dW = np.zeros((27, 16, 16))
    D = dZ.reshape(-1, 16)
    for t, dx, dy, dz in _shifts():
        dW[t] = D.T @ Xp[1 + dx:1 + dx + n0, 1 + dy:1 + dy + n1, 1 + dz:1 + dz + n2].reshape(-1, 16)
    return dW


def forward(params, X, emulate=True):
    W, b = split(params)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    Wq = r(W)
    acts = [r(X)]
    for k in range(LAYERS):
        Z = conv3(acts[-1], Wq[k]) + b[k]
        if k < LAYERS - 1:
            acts.append(r(np.maximum(Z, 0.0)))
        else:
            Z[..., NOUT:] = 0.0
            logits = Z
    return logits, acts, Wq


def bce(logits, labels, mask, pos_weight):
    """Loss sum and dL/dlogits for logits [..][16] (channels < 8), labels [..][8]."""
    z = logits[..., :NOUT]
    y = labels.astype(np.float64)
    m = np.ones(z.shape[:-1]) if mask is None else mask.astype(np.float64)
    lsp = np.where(z >= 0, -np.log1p(np.exp(-z)), z - np.log1p(np.exp(z)))
    lsn = lsp - z
    loss = -(pos_weight * y * lsp + (1 - y) * lsn) * m[..., None]
    s = 1.0 / (1.0 + np.exp(-z))
    g = (s * (pos_weight * y + 1 - y) - pos_weight * y) * m[..., None]
    G = np.zeros(logits.shape)
    G[..., :NOUT] = g
    return float(loss.sum()), G


def loss_and_grads(params, X, labels, mask=None, pos_weight=1.0, emulate=True):
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    logits, acts, Wq = forward(params, X, emulate)
    loss, dZ = bce(logits, labels, mask, pos_weight)
    dZ = r(dZ)
    gW = np.zeros((LAYERS, 27, 16, 16))
    gb = np.zeros((LAYERS, 16))
    for k in range(LAYERS - 1, -1, -1):
        gW[k] = conv3_dw(dZ, acts[k])
        gb[k] = dZ.reshape(-1, 16).sum(0)
        if k == 0:
            break
        dX = conv3_dx(dZ, Wq[k])
        dZ = r(np.where(acts[k] > 0, dX, 0.0))
    return loss, np.concatenate([gW.ravel(), gb.ravel()]), logits
