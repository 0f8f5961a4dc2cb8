"""Pins the float64 CPU network used to check the selector's HIP kernels: analytic
backprop equals central finite differences."""

import numpy as np

from oracle import mlp_ref


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    out = 7
    P = 16 * 64 + 64 + 64 * 64 + 64 + 64 * out + out
    p = rng.standard_normal(P) * 0.2
    x = rng.standard_normal((40, 16))
    t = rng.standard_normal((40, out))
    m = (rng.random((40, out)) > 0.3).astype(np.float64)
    g = mlp_ref.grad(p, x, t, m, out)
    idx = rng.choice(P, 200, replace=False)
    gn = mlp_ref.numeric_grad(p, x, t, m, out, idx)
    assert np.allclose(g[idx], gn, rtol=1e-5, atol=1e-8)


def test_loss_is_masked_mean_square():
    out = 3
    P = 16 * 64 + 64 + 64 * 64 + 64 + 64 * out + out
    p = np.zeros(P)
    p[-out:] = [1.0, 2.0, 3.0]  # zero weights: y = b3 for every sample
    x = np.ones((2, 16))
    t = np.zeros((2, out))
    m = np.array([[1, 0, 1], [0, 0, 1]], float)
    assert abs(mlp_ref.loss(p, x, t, m, out) - (1 + 9 + 9) / 3) < 1e-12
