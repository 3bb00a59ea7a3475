"""Pins the MLP oracle (oracle/mlp.py) by central finite differences. CPU only."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, h, n_out, B = 5, 7, 4, 9
    n = mlp.layout(n_in, h, n_out)["total"]
    p = rng.standard_normal(n) * 0.5
    X = rng.standard_normal((B, n_in))
    T = rng.standard_normal((B, n_out))
    M = (rng.random((B, n_out)) < 0.7).astype(float)
    g, L = mlp.grad(p, X, T, M, n_in, h, n_out)
    assert abs(L - mlp.loss(p, X, T, M, n_in, h, n_out)) < 1e-12
    eps = 1e-6
    fd = np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = eps
        fd[i] = (mlp.loss(p + e, X, T, M, n_in, h, n_out) - mlp.loss(p - e, X, T, M, n_in, h, n_out)) / (2 * eps)
    assert np.max(np.abs(fd - g)) < 1e-6 * max(1.0, np.max(np.abs(g)))


def test_masked_entries_do_not_contribute():
    rng = np.random.default_rng(1)
    n_in, h, n_out, B = 3, 4, 2, 5
    p = rng.standard_normal(mlp.layout(n_in, h, n_out)["total"])
    X = rng.standard_normal((B, n_in))
    T = rng.standard_normal((B, n_out))
    M = np.zeros((B, n_out))
    M[:, 0] = 1
    T2 = T.copy()
    T2[:, 1] += 100.0
    g1, _ = mlp.grad(p, X, T, M, n_in, h, n_out)
    g2, _ = mlp.grad(p, X, T2, M, n_in, h, n_out)
    assert np.array_equal(g1, g2)


def test_adam_first_step_is_sign_times_lr():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -0.25, 0.0])
    p1, m, v = mlp.adam(p, g, np.zeros(3), np.zeros(3), 0.1, 0.9, 0.999, 1e-12, 1)
    assert np.allclose(p1, p - 0.1 * np.sign(g))
