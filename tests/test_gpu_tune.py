"""GPU tests of the device generators, matrix features, sweep and the
selector MLP kernels (gradients against the FP64 oracle in oracle/mlp.py)."""
import math

import numpy as np
import pytest
import torch

from oracle import mlp as omlp
from muxi_spmv_amd import tune
from gpu_util import assert_close_fp32

pytestmark = pytest.mark.gpu


def host_features(rp, ci, C):
    R = rp.size - 1
    lens = np.diff(rp.astype(np.int64))
    rows = np.repeat(np.arange(R), lens)
    cen = rows * (C / R)
    N = ci.size
    return np.array([R, N, lens.mean(), lens.var(), lens.max(initial=0),
                     np.abs(ci - cen).sum() / N / C if N else 0, np.mean(ci // 32 == rows // 32) if N else 0,
                     np.mean(lens == 0)])


@pytest.mark.parametrize("family,p", [("uniform", (5, 15, 0)), ("powerlaw", (2.0, 1, 5000)), ("banded", (64, 16, 0)),
                                      ("blockdiag", (32, 0, 0)), ("poisson", (9.5, 0, 0))])
def test_generator_structure_and_features(family, p):
    R = 20000
    A = tune.gen_matrix(family, R, R, *p, seed=7)
    rp, ci, va = A.host()
    lens = np.diff(rp.astype(np.int64))
    assert rp[0] == 0 and rp[-1] == A.nnz == ci.size
    row = np.repeat(np.arange(R), lens)
    key = row.astype(np.int64) * R + ci
    assert np.all(np.diff(key) > 0)  # distinct and ascending within rows
    assert ci.min() >= 0 and ci.max() < R
    assert np.all(np.abs(va) <= 1.0)
    if family == "uniform":
        assert lens.min() >= 5 and lens.max() <= 15
    if family == "banded":
        assert np.all(np.abs(ci - row) <= 64) and np.all(lens == 16)
    if family == "blockdiag":
        assert np.all(ci // 32 == row // 32) and np.all(lens == 32)
    if family == "poisson":
        assert abs(lens.mean() - 9.5) < 0.2
    if family == "powerlaw":
        assert lens.min() >= 1 and lens.max() <= 5000
    f = tune.features(A)
    assert np.allclose(f, host_features(rp, ci, R), rtol=1e-9, atol=1e-12)
    # same seed, same matrix
    B = tune.gen_matrix(family, R, R, *p, seed=7)
    rp2, ci2, va2 = B.host()
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(va, va2)


def test_generator_permute_keeps_row_multisets():
    A = tune.gen_matrix("banded", 5000, 5000, 20, 8, seed=3, permute=True)
    rp, ci, va = A.host()
    assert rp[-1] == 5000 * 8 and sorted(np.diff(rp).tolist()) == [8] * 5000
    assert np.unique(ci).size > 4000
    f = tune.features(A)
    assert f[5] > 0.1  # bandwidth destroyed by the permutation


def test_sweep_times_every_feasible_candidate_and_results_are_right():
    A = tune.gen_matrix("powerlaw", 100000, 100000, 2.0, 1, 20000, seed=9)
    x = tune.gen_vector(100000, seed=1)
    y = torch.empty(100000, device="cuda")
    fl = tune.Flusher(64 << 20)
    t = tune.sweep(A, x, y, fl, reps=2)
    assert "merge:8" in t and "csr5:256" in t and "ell:0" not in t
    assert all(v > 0 for v in t.values())
    from muxi_spmv_amd import generators as G
    rp, ci, va = A.host()
    h = G.HostCSR(rp, ci, va, 100000, 100000)
    assert_close_fp32(y, h, x.cpu().numpy())  # y from the last candidate timed


def mlp_case(seed, n_in=10, h=64, n_out=14, B=37):
    rng = np.random.default_rng(seed)
    sel = tune.Selector(h, n_in, n_out, seed=seed)
    X = rng.standard_normal((B, n_in)).astype(np.float32)
    T = rng.standard_normal((B, n_out)).astype(np.float32)
    M = (rng.random((B, n_out)) < 0.8).astype(np.float32)
    return sel, X, T, M


def test_mlp_forward_matches_oracle():
    sel, X, T, M = mlp_case(0)
    Y = sel.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    Yo, _ = omlp.forward(sel.params.cpu().numpy(), X, 10, 64, 14)
    assert np.max(np.abs(Y - Yo)) < 1e-4 * max(1, np.abs(Yo).max())


def test_mlp_backward_matches_oracle():
    sel, X, T, M = mlp_case(1)
    G, L = sel.backward(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), torch.from_numpy(M).cuda())
    Go, Lo = omlp.grad(sel.params.cpu().numpy(), X, T, M, 10, 64, 14)
    assert abs(L - Lo) <= 1e-4 * abs(Lo)
    G = G.cpu().numpy()
    assert np.max(np.abs(G - Go)) <= 1e-4 * np.max(np.abs(Go))


def test_mlp_adam_matches_oracle():
    sel, X, T, M = mlp_case(2)
    from muxi_spmv_amd._lib import load
    p0 = sel.params.cpu().numpy().astype(np.float64)
    g = np.random.default_rng(3).standard_normal(p0.size).astype(np.float32)
    m = torch.zeros_like(sel.params)
    v = torch.zeros_like(sel.params)
    gd = torch.from_numpy(g).cuda()
    pe, me, ve = p0, np.zeros(p0.size), np.zeros(p0.size)
    for t in range(1, 4):
        load().muxi_mlp_adam(sel.params.data_ptr(), gd.data_ptr(), m.data_ptr(), v.data_ptr(), sel.params.numel(),
                             1e-2, 0.9, 0.999, 1e-8, t, None)
        pe, me, ve = omlp.adam(pe, g.astype(np.float64), me, ve, 1e-2, 0.9, 0.999, 1e-8, t)
    torch.cuda.synchronize()
    assert np.max(np.abs(sel.params.cpu().numpy() - pe)) < 1e-5


def test_mlp_learns_a_toy_selection_problem():
    # target: candidate 0 best when x0 > 0 else candidate 1; the trained pick must be right >= 95%
    rng = np.random.default_rng(4)
    X = rng.standard_normal((600, 10)).astype(np.float32)
    T = np.zeros((600, 14), np.float32)
    T[:, 0] = np.where(X[:, 0] > 0, 0.0, 1.0)
    T[:, 1] = np.where(X[:, 0] > 0, 1.0, 0.0)
    T[:, 2:] = 2.0
    M = np.ones_like(T)
    sel = tune.Selector(64, 10, 14, seed=5)
    losses = sel.fit(X, T, M, epochs=2000, lr=3e-3, batch=128)
    assert losses[-1] < 0.2 * losses[0]
    Xt = rng.standard_normal((200, 10)).astype(np.float32)
    pick = sel.pick(Xt)
    # errors only near the decision boundary x0 = 0
    ok = pick == np.where(Xt[:, 0] > 0, 0, 1)
    assert np.mean(ok) >= 0.9
    assert np.all(ok[np.abs(Xt[:, 0]) > 0.5])


def test_library_auto_pick_matches_python_selector():
    """The C-side pick (device features, model input, normalisation, HIP forward, feasibility mask) agrees with
    the Python selector on the same network; random weights suffice for this."""
    from muxi_spmv_amd._lib import load
    rng = np.random.default_rng(0)
    for n_in in (10, 11):  # the 11th input is the FP64 flag
        sel = tune.Selector(64, n_in=n_in, seed=3)
        sel.mu = torch.from_numpy(rng.standard_normal(n_in).astype(np.float32))
        sel.sd = torch.from_numpy((1 + rng.random(n_in)).astype(np.float32))
        sel.install()
        try:
            assert load().muxi_selector_state() == 2
            for i, (fam, p) in enumerate([("uniform", (5, 15, 0)), ("powerlaw", (2.0, 1, 5000)),
                                          ("banded", (64, 16, 0)), ("blockdiag", (32, 0, 0)), ("poisson", (30.0, 0, 0))]):
                A = tune.gen_matrix(fam, 30000, 30000, *p, seed=20 + i)
                f = tune.features(A)
                for xt in (torch.float32, torch.float64):
                    P = sel.predict(tune.feature_vector(f, xt == torch.float64)[None])[0]
                    P = np.where(tune.feasible_mask(f, xt == torch.float64) > 0, P, np.inf)
                    name = tune.library_pick(A, xt)
                    assert name in tune.CAND_NAMES
                    assert P[tune.CAND_NAMES.index(name)] <= P.min() + 1e-4 * (1 + abs(P.min()))
        finally:
            load().muxi_selector_reset()


def test_selector_clear_falls_back_to_fixed_rule():
    from muxi_spmv_amd._lib import load
    lib = load()
    lib.muxi_selector_clear()
    try:
        assert lib.muxi_selector_state() == 0
        small = tune.gen_matrix("uniform", 20000, 20000, 5, 15, seed=1)
        assert tune.library_pick(small) == "merge:4"
        wide = tune.gen_matrix("uniform", 20000, 2_000_000, 5, 15, seed=2)  # x is 8 MB: column slabs
        assert tune.library_pick(wide) == "xslab:4"
    finally:
        lib.muxi_selector_reset()
    assert lib.muxi_selector_state() in (0, 1)


def test_selector_ensemble_is_the_mean_of_its_members():
    """Selector.ensemble (stacked / block-diagonal / averaged layers) predicts the mean of the members' predictions."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((40, tune.N_INPUT)).astype(np.float32)
    members = []
    for k in range(3):
        m = tune.Selector(16, seed=10 + k)
        m.mu = torch.from_numpy(rng.standard_normal(tune.N_INPUT).astype(np.float32))
        m.sd = torch.from_numpy((1 + rng.random(tune.N_INPUT)).astype(np.float32))
        members.append(m)
    for m in members[1:]:
        m.mu, m.sd = members[0].mu.clone(), members[0].sd.clone()
    e = tune.Selector.ensemble(members)
    assert e.h == 48
    mean = np.mean([m.predict(X) for m in members], 0)
    assert np.allclose(e.predict(X), mean, rtol=1e-5, atol=1e-5)
