"""float64 CPU restatement of the config-selector network -- TEST INFRASTRUCTURE ONLY.

Same parameter layout as cccl_amd/csrc/mlp.hip (W1[16][64], b1, W2[64][64], b2,
W3[64][C], b3), same loss (masked mean squared error).  `grad` is analytic backprop;
`numeric_grad` is central finite differences of `loss`, which pins `grad`.
"""

from __future__ import annotations

import numpy as np

IN, H = 16, 64


def unpack(p: np.ndarray, out: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    shapes = [(IN, H), (H,), (H, H), (H,), (H, out), (out,)]
    parts = []
    for s in shapes:
        k = int(np.prod(s))
        parts.append(p[o:o + k].reshape(s))
        o += k
    assert o == p.size
    return parts


def forward(p, x, out):
    W1, b1, W2, b2, W3, b3 = unpack(p, out)
    z1 = x @ W1 + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2 + b2
    a2 = np.maximum(z2, 0)
    y = a2 @ W3 + b3
    return y, (x, z1, a1, z2, a2)


def loss(p, x, t, m, out):
    y, _ = forward(p, np.asarray(x, np.float64), out)
    return float((m * (y - t) ** 2).sum() / m.sum())


def grad(p, x, t, m, out):
    x = np.asarray(x, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, out)
    y, (x, z1, a1, z2, a2) = forward(p, x, out)
    dy = 2.0 * m * (y - t) / m.sum()
    gW3 = a2.T @ dy
    gb3 = dy.sum(0)
    dz2 = (dy @ W3.T) * (z2 > 0)
    gW2 = a1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (z1 > 0)
    gW1 = x.T @ dz1
    gb1 = dz1.sum(0)
    return np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])


def numeric_grad(p, x, t, m, out, idx, eps=1e-6):
    p = np.asarray(p, np.float64).copy()
    g = np.zeros(len(idx))
    for k, i in enumerate(idx):
        old = p[i]
        p[i] = old + eps
        lp = loss(p, x, t, m, out)
        p[i] = old - eps
        lm = loss(p, x, t, m, out)
        p[i] = old
        g[k] = (lp - lm) / (2 * eps)
    return g
