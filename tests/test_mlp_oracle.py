"""Pins the MLP oracle (oracle/mlp.py): the gradient by central finite differences, Adam by the constant-gradient
invariant (every bias-corrected step is lr * sign(g)) and a hand-worked three-step case. CPU only."""
import numpy as np
import pytest

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


def test_adam_constant_gradient_steps_are_sign_times_lr():
    # With a constant gradient g the moments are m_t = (1 - b1^t) g and v_t = (1 - b2^t) g^2, so the
    # bias-corrected ratio is exactly sign(g) at every step t: a wrong bias-correction exponent (t - 1, or none)
    # changes the step at t >= 2.
    g = np.array([0.5, -2.0, 3.0])
    p = np.zeros(3)
    m = np.zeros(3)
    v = np.zeros(3)
    for t in range(1, 6):
        p_new, m, v = mlp.adam(p, g, m, v, 0.01, 0.9, 0.999, 1e-12, t)
        assert np.allclose(p_new - p, -0.01 * np.sign(g), rtol=1e-9, atol=0)
        p = p_new


def test_adam_two_steps_hand_worked():
    # g1 = +1, g2 = -1, b1 = 0.9, b2 = 0.999, lr = 0.1, p0 = 0:
    #   step 1: p1 = -0.1 (sign step)
    #   step 2: m2 = 0.9*0.1 - 0.1 = -0.01, m2_hat = -0.01 / (1 - 0.81) = -1/19
    #           v2 = 0.999*0.001 + 0.001 = 0.001999, v2_hat = 0.001999 / (1 - 0.998001) = 1
    #           p2 = p1 - 0.1 * (-1/19) / 1 = -0.1 * 18/19
    #   step 3 (g3 = +1): m3 = 0.9*(-0.01) + 0.1 = 0.091, m3_hat = 0.091 / (1 - 0.729) = 0.091/0.271
    #           v3 = 0.999*0.001999 + 0.001 = 0.002997001, v3_hat = 0.002997001 / (1 - 0.997002999) = 1
    #           p3 = p2 - 0.1 * 0.091/0.271
    p = np.zeros(1)
    m = np.zeros(1)
    v = np.zeros(1)
    p, m, v = mlp.adam(p, np.array([1.0]), m, v, 0.1, 0.9, 0.999, 0.0, 1)
    assert p[0] == pytest.approx(-0.1, rel=1e-12)
    p, m, v = mlp.adam(p, np.array([-1.0]), m, v, 0.1, 0.9, 0.999, 0.0, 2)
    assert m[0] == pytest.approx(-0.01, rel=1e-12)
    assert v[0] == pytest.approx(0.001999, rel=1e-12)
    assert p[0] == pytest.approx(-0.1 * 18 / 19, rel=1e-12)
    p, m, v = mlp.adam(p, np.array([1.0]), m, v, 0.1, 0.9, 0.999, 0.0, 3)
    assert p[0] == pytest.approx(-0.1 * 18 / 19 - 0.1 * 0.091 / 0.271, rel=1e-12)
