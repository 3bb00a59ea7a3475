"""CPU oracle of the selector MLP — TEST INFRASTRUCTURE ONLY.

FP64 numpy restatement of the network in include/muxi_spmv/muxi_spmv_tune.h
(in -> h -> h -> out, ReLU on the hidden layers, parameters W1,b1,W2,b2,W3,b3
row-major in one flat array), its masked mean-squared-error loss and the
gradient by backpropagation, plus Adam. Pinned in tests/test_mlp_oracle.py:
the gradient by central finite differences of the loss, Adam by the
constant-gradient invariant (each bias-corrected step is lr * sign(g) at every
t) and a hand-worked three-step case.
"""
from __future__ import annotations

import numpy as np


def layout(n_in: int, h: int, n_out: int) -> dict:
    o = {}
    k = 0
    for name, size in [("w1", h * n_in), ("b1", h), ("w2", h * h), ("b2", h), ("w3", n_out * h), ("b3", n_out)]:
        o[name] = (k, k + size)
        k += size
    o["total"] = k
    return o


def unpack(p: np.ndarray, n_in: int, h: int, n_out: int):
    L = layout(n_in, h, n_out)
    g = lambda n: p[L[n][0]:L[n][1]]  # noqa: E731
    return (g("w1").reshape(h, n_in), g("b1"), g("w2").reshape(h, h), g("b2"), g("w3").reshape(n_out, h), g("b3"))


def forward(p, X, n_in, h, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), n_in, h, n_out)
    X = np.asarray(X, np.float64)
    H1 = np.maximum(X @ W1.T + b1, 0)
    H2 = np.maximum(H1 @ W2.T + b2, 0)
    return H2 @ W3.T + b3, (X, H1, H2)


def loss(p, X, T, M, n_in, h, n_out) -> float:
    Y, _ = forward(p, X, n_in, h, n_out)
    cnt = max(float(np.sum(M)), 1.0)
    return float(np.sum(M * (Y - T) ** 2) / cnt)


def grad(p, X, T, M, n_in, h, n_out) -> tuple[np.ndarray, float]:
    p = np.asarray(p, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, h, n_out)
    Y, (X, H1, H2) = forward(p, X, n_in, h, n_out)
    M = np.asarray(M, np.float64)
    cnt = max(float(M.sum()), 1.0)
    dY = 2.0 * M * (Y - np.asarray(T, np.float64)) / cnt
    gW3 = dY.T @ H2
    gb3 = dY.sum(0)
    dH2 = (dY @ W3) * (H2 > 0)
    gW2 = dH2.T @ H1
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ W2) * (H1 > 0)
    gW1 = dH1.T @ X
    gb1 = dH1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return g, float(np.sum(M * (Y - T) ** 2) / cnt)


def adam(p, g, m, v, lr, b1, b2, eps, t):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
