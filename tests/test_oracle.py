"""Pins the CPU oracle (oracle/) against known answers, closed forms, brute
force and independent numpy formulations. CPU only."""
import math

import numpy as np
import pytest

import oracle
from muxi_spmv_amd import generators as G

U32 = 2.0 ** -24


def rand_csr(R, C, max_len, seed, empty_frac=0.2, dtype=np.float32):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, size=R)
    lens[rng.random(R) < empty_frac] = 0
    lens = np.minimum(lens, C)
    rp = np.zeros(R + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    cols = np.concatenate([np.sort(rng.choice(C, size=k, replace=False)) for k in lens]) if lens.sum() else \
        np.zeros(0, np.int64)
    vals = rng.uniform(-1, 1, size=cols.size).astype(dtype)
    return rp, cols.astype(np.int32), vals


# ---------------------------------------------------------------- FP32 loop / FP64 oracle
def test_diagonal_known_answer():
    # tests/correctness/test_spmv.cu:14-30,63-67: diagonal n=1024, x[i]=i+1, y == x (tol 1e-5)
    n = 1024
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    v = np.ones(n, np.float32)
    x = np.arange(1, n + 1, dtype=np.float32)
    y = oracle.spmv_csr_f32(rp, ci, v, x)
    assert np.max(np.abs(y - x)) <= 1e-5
    y64, _ = oracle.spmv_f64(rp, ci, v, x)
    assert np.array_equal(y64, x.astype(np.float64))


def test_laplacian_closed_form():
    # 5-point stencil, x = 1: interior 0, edge 1, corner 2
    n = 37
    A = G.laplacian_2d(n)
    x = np.ones(n * n)
    y64, _ = oracle.spmv_f64(A.rowPtr, A.colIdx, A.values, x)
    exp = np.zeros((n, n))
    exp[0, :] += 1
    exp[-1, :] += 1
    exp[:, 0] += 1
    exp[:, -1] += 1
    assert np.array_equal(y64, exp.ravel())
    y32 = oracle.spmv_csr_f32(A.rowPtr, A.colIdx, A.values.astype(np.float32), x.astype(np.float32))
    assert np.array_equal(y32.astype(np.float64), exp.ravel())


def test_banded_harmonic_closed_form():
    # banded 1/(|i-j|+1) (src/generators/matrix_generator.cu:52-84), x = 1 -> sum of 1/(d+1) over the band
    n, bw = 200, 7
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(max(0, i - bw), min(n, i + bw + 1)):
            rows.append(i)
            cols.append(j)
            vals.append(1.0 / (abs(i - j) + 1))
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, np.array(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    y64, _ = oracle.spmv_f64(rp, np.array(cols), np.array(vals), np.ones(n))
    for i in range(n):
        exp = math.fsum(1.0 / (abs(i - j) + 1) for j in range(max(0, i - bw), min(n, i + bw + 1)))
        assert y64[i] == exp


def test_f64_dot2_matches_exact_fsum():
    rp, ci, v = rand_csr(300, 500, 60, seed=3, dtype=np.float64)
    x = np.random.default_rng(4).standard_normal(500) * 1e3
    y, ab = oracle.spmv_f64(rp, ci, v, x)
    for r in range(300):
        terms = [float(v[i]) * float(x[ci[i]]) for i in range(rp[r], rp[r + 1])]
        # products are exact in the fsum reference only if we split them; use fractions-free bound:
        exact = math.fsum(terms)
        # Dot2 error <= u*|y| + gamma_n^2 * sum|terms|, far below fsum's product rounding
        assert abs(y[r] - exact) <= 2 ** -52 * abs(exact) + 4 * 2 ** -53 * ab[r] + 1e-300


def test_f32_loop_within_summation_bound():
    rp, ci, v = rand_csr(2000, 3000, 80, seed=5)
    x = np.random.default_rng(6).uniform(-1, 1, 3000).astype(np.float32)
    y32 = oracle.spmv_csr_f32(rp, ci, v, x)
    y64, ab = oracle.spmv_f64(rp, ci, v, x)
    n = np.diff(rp)
    bound = (n + 1) * U32 * ab * 1.0001 + 1e-30
    assert np.all(np.abs(y32 - y64) <= bound)


def int_csr(R, C, max_len, seed, empty_frac=0.2):
    """Random CSR with small-integer values: every product and partial sum is an integer far below 2^53, so any
    order of FP64 summation gives the exact integer answer."""
    rp, ci, _ = rand_csr(R, C, max_len, seed, empty_frac)
    rng = np.random.default_rng(seed + 1000)
    v = rng.integers(-50, 51, size=ci.size).astype(np.float64)
    x = rng.integers(-50, 51, size=C).astype(np.float64)
    return rp, ci, v, x


def dense_of(rp, ci, v, C):
    R = rp.size - 1
    D = np.zeros((R, C), np.int64)
    for r in range(R):
        D[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]].astype(np.int64)
    return D


def test_f64_plain_exact_on_integers_and_closed_forms():
    # integer data: the plain FP64 loop must equal the exact integer product (numpy int64 on the dense matrix)
    # and the Dot2 oracle bit for bit
    rp, ci, v, x = int_csr(400, 300, 40, seed=31)
    y = oracle.spmv_csr_f64_plain(rp, ci, v, x)
    exact = dense_of(rp, ci, v, 300) @ x.astype(np.int64)
    assert np.array_equal(y, exact.astype(np.float64))
    assert np.array_equal(y, oracle.spmv_f64(rp, ci, v, x)[0])
    # Laplacian with x = 1 (interior 0, edges 1, corners 2) and the diagonal known answer
    n = 29
    A = G.laplacian_2d(n)
    y = oracle.spmv_csr_f64_plain(A.rowPtr, A.colIdx, A.values, np.ones(n * n))
    exp = np.zeros((n, n))
    exp[0, :] += 1
    exp[-1, :] += 1
    exp[:, 0] += 1
    exp[:, -1] += 1
    assert np.array_equal(y, exp.ravel())
    d = np.arange(1, 65, dtype=np.float64)
    y = oracle.spmv_csr_f64_plain(np.arange(65, dtype=np.int32), np.arange(64, dtype=np.int32), np.ones(64), d)
    assert np.array_equal(y, d)


def test_f64_plain_within_summation_bound():
    # real data: |plain - exact| <= gamma_n * sum |a x| (recursive summation), gamma_n ~ n u with u = 2^-53;
    # the row order of the loop is what a dropped or doubled term would break by far more than this
    rp, ci, v = rand_csr(500, 800, 60, seed=33, dtype=np.float64)
    x = np.random.default_rng(34).standard_normal(800)
    y = oracle.spmv_csr_f64_plain(rp, ci, v, x)
    for r in range(500):
        terms = [float(v[i]) * float(x[ci[i]]) for i in range(rp[r], rp[r + 1])]
        n = len(terms)
        bound = (n + 1) * 2.0 ** -53 * math.fsum(abs(t) for t in terms) * 1.0001 + 1e-300
        assert abs(y[r] - math.fsum(terms)) <= bound


def test_f32_mt_equals_single_thread():
    rp, ci, v = rand_csr(5000, 5000, 30, seed=7)
    x = np.random.default_rng(8).uniform(-1, 1, 5000).astype(np.float32)
    assert np.array_equal(oracle.spmv_csr_f32(rp, ci, v, x), oracle.spmv_csr_f32_mt(rp, ci, v, x, 4))


# ---------------------------------------------------------------- merge path
def brute_merge_coords(rp, k):
    """Walk the merge path item by item: nnz of row r, then its row end."""
    R = rp.size - 1
    N = int(rp[-1])
    r, n = 0, 0
    for _ in range(k):
        if r < R and n < rp[r + 1]:
            n += 1
        else:
            r += 1
    assert r <= R and n <= N
    return r, n


@pytest.mark.parametrize("seed", range(6))
def test_merge_coords_brute_force(seed):
    rp, ci, v = rand_csr(25, 40, 9, seed=seed, empty_frac=0.35)
    R, N = rp.size - 1, int(rp[-1])
    ks = np.arange(R + N + 1, dtype=np.int32)
    r, n = oracle.merge_path_coords(rp, N, ks)
    for k in range(R + N + 1):
        br, bn = brute_merge_coords(rp, k)
        assert (r[k], n[k]) == (br, bn), k
        # the invariant in closed form
        if r[k] < R:
            assert rp[r[k]] + r[k] <= k <= rp[r[k] + 1] + r[k]


def test_merge_coords_leading_trailing_empty_rows():
    rp = np.array([0, 0, 0, 3, 3, 5, 5, 5], np.int32)
    R, N = 7, 5
    r, n = oracle.merge_path_coords(rp, N, np.arange(R + N + 1, dtype=np.int32))
    for k in range(R + N + 1):
        assert (r[k], n[k]) == brute_merge_coords(rp, k)
    assert (r[-1], n[-1]) == (R, N)


def test_merge_partitions_closed_form():
    for L_R, L_N, P in [(10, 20, 3), (2_000_000, 14_700_000, 3328), (5, 0, 7), (1, 1, 1)]:
        pos = oracle.merge_partitions(L_R, L_N, P)
        L = L_R + L_N
        per = -(-L // P)
        assert pos.tolist() == [min(t * per, L) for t in range(P + 1)]


def test_merge_partition_count_reference_constants():
    # 64-wide wavefront: ePP=16, mult=8, 104 SMs, 4 waves per 256-thread block -> cap 3328
    assert oracle.merge_partition_count(2_000_000, 14_700_000) == 3328
    assert oracle.merge_partition_count(10, 5) == 1  # (15)/16 = 0 -> max(1, 0)
    assert oracle.merge_partition_count(1000, 600) == 100
    # _v2: ePP=64, mult=4 (spmv_csr5.cu:1031-1038); ultra: ePP=8, mult=16
    assert oracle.merge_partition_count(2_000_000, 14_700_000, 64, 4) == 104 * 4 * 4
    assert oracle.merge_partition_count(100_000, 1_000_000, 8, 16) == min(1_100_000 // 8, 104 * 4 * 16)


# ---------------------------------------------------------------- CSR5 tiles
@pytest.mark.parametrize("sigma", [64, 256, 512, 100, 1024])
def test_csr5_tiles_vs_searchsorted(sigma):
    rp, ci, v = rand_csr(3000, 3000, 40, seed=sigma, empty_frac=0.3)
    N = int(rp[-1])
    tr, to = oracle.csr5_tiles(rp, N, sigma)
    starts = np.arange(tr.size, dtype=np.int64) * sigma
    exp = np.searchsorted(rp[:-1], starts, side="right") - 1
    assert np.array_equal(tr, exp)
    assert np.array_equal(to, starts - rp[exp])
    # the tile's first nonzero lies in its row
    assert np.all((rp[tr] <= starts) & (starts < rp[tr + 1]))


def test_csr5_default_sigma():
    assert oracle.csr5_default_sigma(64) == 256
    assert oracle.csr5_default_sigma(32) == 128


# ---------------------------------------------------------------- ELL
def test_ell_layout_reconstructs_csr():
    rp, ci, v = rand_csr(500, 700, 12, seed=11)
    w = oracle.ell_width(rp)
    assert w == int(np.diff(rp).max())
    ec, ev = oracle.csr_to_ell(rp, ci, v, w)
    R = rp.size - 1
    dense_ell = np.zeros((R, 700))
    for i in range(w):
        for r in range(R):
            c = ec[r + i * R]
            if c >= 0:
                dense_ell[r, c] += ev[r + i * R]
            else:
                assert ev[r + i * R] == 0.0
    dense = np.zeros((R, 700))
    for r in range(R):
        dense[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    assert np.array_equal(dense, dense_ell)
    # padding count and slot order
    assert int((ec < 0).sum()) == R * w - int(rp[-1])


def test_ell_truncation():
    rp = np.array([0, 5, 6], np.int32)
    ci = np.arange(6, dtype=np.int32)
    v = np.arange(6, dtype=np.float64) + 1
    ec, ev = oracle.csr_to_ell(rp, ci, v, 3)
    assert ec.tolist() == [0, 5, 1, -1, 2, -1]
    assert ev.tolist() == [1, 6, 2, 0, 3, 0]


@pytest.mark.parametrize("extra", [0, 3])
def test_ell_spmv_exact_on_integers(extra):
    # the ELL product over the packed slots (padding skipped) equals the exact integer product of the CSR
    # matrix: width = max row length, and with `extra` more padded slots per row
    rp, ci, v, x = int_csr(300, 250, 25, seed=41)
    w = oracle.ell_width(rp) + extra
    ec, ev = oracle.csr_to_ell(rp, ci, v, w)
    y = oracle.spmv_ell_f64(ec, ev, w, x)
    exact = dense_of(rp, ci, v, 250) @ x.astype(np.int64)
    assert np.array_equal(y, exact.astype(np.float64))
    assert np.array_equal(y, oracle.spmv_csr_f64_plain(rp, ci, v, x))


def test_ell_spmv_truncated_width_keeps_first_slots():
    # a width below the longest row keeps each row's first `w` nonzeros (spmv_ellpack.cu:49): the product equals
    # the exact product of the CSR matrix cut to those entries
    rp, ci, v, x = int_csr(200, 180, 30, seed=43, empty_frac=0.0)
    w = 7
    ec, ev = oracle.csr_to_ell(rp, ci, v, w)
    y = oracle.spmv_ell_f64(ec, ev, w, x)
    R = rp.size - 1
    cut = np.array([min(rp[r + 1] - rp[r], w) for r in range(R)])
    crp = np.zeros(R + 1, np.int32)
    np.cumsum(cut, out=crp[1:])
    keep = np.concatenate([np.arange(rp[r], rp[r] + cut[r]) for r in range(R)])
    exact = dense_of(crp, ci[keep], v[keep], 180) @ x.astype(np.int64)
    assert np.array_equal(y, exact.astype(np.float64))
    assert (np.diff(rp) > w).any()  # the case really truncates


# ---------------------------------------------------------------- permutations
def test_reorder_columns_com_vs_numpy():
    rp, ci, v = rand_csr(400, 300, 20, seed=21, empty_frac=0.1)
    fwd, nc = oracle.reorder_columns_com(rp, ci, 300)
    rows = np.repeat(np.arange(400), np.diff(rp))
    s = np.zeros(300)
    np.add.at(s, ci, rows)  # exact: integers < 2^24
    cnt = np.bincount(ci, minlength=300)
    key = s.astype(np.float32)
    nz = cnt > 0
    key[nz] = key[nz] / cnt[nz].astype(np.float32)
    exp = np.lexsort((np.arange(300), key))
    assert np.array_equal(fwd, exp)
    inv = np.empty(300, np.int64)
    inv[fwd] = np.arange(300)
    assert np.array_equal(nc, inv[ci])
    xr = oracle.reorder_x(np.arange(300, dtype=np.float32), fwd)
    assert np.array_equal(xr, fwd.astype(np.float32))


def test_sort_columns_within_rows_vs_lexsort():
    rng = np.random.default_rng(22)
    R = 200
    lens = rng.integers(0, 15, R)
    rp = np.zeros(R + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = rng.integers(0, 10, int(rp[-1])).astype(np.int32)  # duplicates on purpose
    v = rng.integers(-3, 3, int(rp[-1])).astype(np.float32)
    nci, nv = oracle.sort_columns_within_rows(rp, ci, v)
    row = np.repeat(np.arange(R), lens)
    order = np.lexsort((v, ci, row))
    assert np.array_equal(nci, ci[order]) and np.array_equal(nv, v[order])


def test_minrow_and_freq_orderings():
    rp, ci, v = rand_csr(300, 250, 10, seed=23)
    perm, nc = oracle.reorder_columns_minrow(rp, ci, 250)
    assert sorted(perm.tolist()) == list(range(250))
    rows = np.repeat(np.arange(300), np.diff(rp))
    minrow = np.full(250, 300)
    np.minimum.at(minrow, ci, rows)
    assert np.all(np.diff(minrow[perm]) >= 0)
    inv = np.empty(250, np.int64)
    inv[perm] = np.arange(250)
    assert np.array_equal(nc, inv[ci])
    perm2, nc2 = oracle.reorder_columns_freq(rp, ci, 250)
    cnt = np.bincount(ci, minlength=250)
    assert sorted(perm2.tolist()) == list(range(250))
    assert np.all(np.diff(cnt[perm2]) <= 0)


def test_rows_by_length_is_stable_sort_and_restore():
    rp, ci, v = rand_csr(500, 400, 25, seed=24)
    rm, irm, nrp, nci, nv = oracle.reorder_rows_by_length(rp, ci, v)
    lens = np.diff(rp)
    assert np.array_equal(rm, np.argsort(lens, kind="stable"))
    assert np.array_equal(irm[rm], np.arange(500))
    x = np.random.default_rng(25).uniform(-1, 1, 400).astype(np.float32)
    y = oracle.spmv_csr_f32(rp, ci, v, x)
    yp = oracle.spmv_csr_f32(nrp, nci, nv, x)
    assert np.array_equal(oracle.restore_order(yp, rm), y)
