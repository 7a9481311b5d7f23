"""Pins the MLP oracle: its exact-FP64 gradient against central
differences, and the BF16 rounding helper against known bit patterns."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8,
                  -2.5, 3.14159265])
    r = M.bf16(x)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7
    assert r[3] == 1.0 + 2 ** -6 and r[4] == -2.5
    assert abs(r[5] - np.pi) < 2 ** -7 * 4


def test_gradient_central_differences():
    rng = np.random.default_rng(0)
    dims = [7, 16, 16, 16, 5]
    n = sum(dims[l] * dims[l + 1] + dims[l + 1] for l in range(4))
    p = rng.standard_normal(n) * 0.4
    X = rng.standard_normal((9, 7))
    T = rng.standard_normal((9, 5))
    mu, sig, tau = rng.standard_normal(7), 1 + rng.random(7), 1 + rng.random(5)
    _, g = M.loss_and_grad(p, dims, X, T, mu, sig, tau, rounding=False)
    h = 1e-6
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, dims, X, T, mu, sig, tau, False)
        lm, _ = M.loss_and_grad(p - e, dims, X, T, mu, sig, tau, False)
        assert abs((lp - lm) / (2 * h) - g[i]) < 1e-7 * max(1, abs(g[i]))
