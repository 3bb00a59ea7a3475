"""Parity of the library's host permutations (muxi_spmv_prep.h) with the
oracle: bit-exact maps and permuted arrays. CPU only (no device calls)."""
import numpy as np
import pytest

import oracle
from muxi_spmv_amd import generators as G
from muxi_spmv_amd import prep


def zoo():
    yield "uniform", G.uniform_rows(3000, 2500, 0, 20, seed=31)
    yield "powerlaw", G.powerlaw_rows(5000, 5000, seed=32, kmax=800)
    yield "banded", G.banded(4000, 16, 8, seed=33)
    yield "blockdiag", G.block_diagonal(512, 32, seed=34)
    rng = np.random.default_rng(35)
    lens = rng.integers(0, 6, 700)
    lens[:10] = 0
    cols = G._distinct_sorted(lens, np.zeros(700, np.int64), np.full(700, 900, np.int64), rng)
    yield "sparse_cols", G._finish(lens, cols, rng.uniform(-1, 1, cols.size).astype(np.float32), 700, 900)


ZOO = dict(zoo())


@pytest.mark.parametrize("name", list(ZOO))
def test_reorder_columns_com_bit_exact(name):
    h = ZOO[name]
    fwd, nci, stats = prep.reorder_columns(h.rowPtr, h.colIdx, h.numCols)
    efwd, enci = oracle.reorder_columns_com(h.rowPtr, h.colIdx, h.numCols)
    assert np.array_equal(fwd, efwd)
    assert np.array_equal(nci, enci)
    assert stats["numReorders"] == int((fwd != np.arange(h.numCols)).sum())
    x = G.uniform_vector(h.numCols, 3)
    assert np.array_equal(prep.reorder_x(x, fwd), oracle.reorder_x(x, efwd))


@pytest.mark.parametrize("name", list(ZOO))
def test_minrow_and_frequency_bit_exact(name):
    h = ZOO[name]
    p, n = prep.reorder_columns_minrow(h.rowPtr, h.colIdx, h.numCols)
    ep, en = oracle.reorder_columns_minrow(h.rowPtr, h.colIdx, h.numCols)
    assert np.array_equal(p, ep) and np.array_equal(n, en)
    p, n = prep.reorder_columns_by_frequency(h.rowPtr, h.colIdx, h.numCols)
    ep, en = oracle.reorder_columns_freq(h.rowPtr, h.colIdx, h.numCols)
    assert np.array_equal(p, ep) and np.array_equal(n, en)


@pytest.mark.parametrize("name", list(ZOO))
def test_rows_by_length_bit_exact(name):
    h = ZOO[name]
    got = prep.reorder_rows_by_length(h.rowPtr, h.colIdx, h.values)
    exp = oracle.reorder_rows_by_length(h.rowPtr, h.colIdx, h.values)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_sort_columns_within_rows_bit_exact():
    rng = np.random.default_rng(36)
    R = 400
    lens = rng.integers(0, 80, R)
    rp = np.zeros(R + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = rng.integers(0, 50, int(rp[-1])).astype(np.int32)  # duplicate columns on purpose
    v = rng.integers(-4, 4, int(rp[-1])).astype(np.float32)
    a = prep.sort_columns_within_rows(rp, ci, v)
    b = oracle.sort_columns_within_rows(rp, ci, v)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_rcm_is_permutation_and_reduces_bandwidth():
    h = G.banded(3000, 20, 6, seed=37)
    q = np.random.default_rng(38).permutation(3000).astype(np.int32)  # symmetric scramble of the band
    pr = G.HostCSR(*prep.permute_symmetric(h.rowPtr, h.colIdx, h.values, q), 3000, 3000)
    bw0 = prep.bandwidth(pr.rowPtr, pr.colIdx)
    perm = prep.rcm(pr.rowPtr, pr.colIdx)
    assert sorted(perm.tolist()) == list(range(3000))
    nrp, nci, nv = prep.permute_symmetric(pr.rowPtr, pr.colIdx, pr.values, perm)
    bw1 = prep.bandwidth(nrp, nci)
    assert bw1 < bw0 / 10
    # P A P^T times P x equals P (A x)
    x = G.uniform_vector(3000, 39)
    y = oracle.spmv_f64(pr.rowPtr, pr.colIdx, pr.values, x)[0]
    yp = oracle.spmv_f64(nrp, nci, nv, x[perm])[0]
    assert np.allclose(yp, y[perm], rtol=0, atol=1e-12)


def test_rcm_on_path_graph_gives_band_one():
    # a path graph numbered randomly: RCM recovers bandwidth 1
    n = 200
    rng = np.random.default_rng(40)
    lab = rng.permutation(n)
    rows = [[] for _ in range(n)]
    for i in range(n - 1):
        a, b = lab[i], lab[i + 1]
        rows[a].append(b)
        rows[b].append(a)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate([sorted(r) for r in rows]).astype(np.int32)
    perm = prep.rcm(rp, ci)
    nrp, nci, _ = prep.permute_symmetric(rp, ci, np.ones(ci.size, np.float32), perm)
    assert prep.bandwidth(nrp, nci) == 1


@pytest.mark.parametrize("G", [1, 2, 3, 8, 64])
def test_shard_rows_are_merge_path_cuts(G):
    """Shard bounds are the merge-path coordinate rows of the G equal diagonals (the oracle's search, the
    reference's merge_path_search shape), cover every row once and carry about (R + N) / G path items each."""
    import oracle
    from muxi_spmv_amd import generators as Gen
    for h in (Gen.powerlaw_rows(5000, 5000, seed=2, kmax=800), Gen.uniform_rows(3000, 3000, 0, 9, seed=3)):
        rp = h.rowPtr
        R, N = rp.size - 1, int(rp[-1])
        b = prep.shard_rows(rp, G)
        L = R + N
        d = np.array([g * L // G for g in range(G + 1)], np.int32)
        rows, _ = oracle.merge_path_coords(rp, N, d)
        exp = rows.copy()
        exp[0], exp[G] = 0, R
        assert np.array_equal(b, exp)
        assert b[0] == 0 and b[-1] == R and np.all(np.diff(b) >= 0)
        # moving a cut to its row start shifts a shard by at most one row (its nonzeros and its end), plus rounding
        items = np.diff(b) + np.diff(rp[b])
        assert np.all(np.abs(items - L / G) <= np.diff(rp).max() + 2)


def test_shard_rows_edge_cases():
    rp = np.array([0, 0, 0, 5, 5], np.int32)  # empty rows around one row
    assert prep.shard_rows(rp, 1).tolist() == [0, 4]
    b = prep.shard_rows(rp, 4)
    assert b[0] == 0 and b[-1] == 4 and np.all(np.diff(b) >= 0)
    assert prep.shard_rows(np.array([0], np.int32), 3).tolist() == [0, 0, 0, 0]
