"""Pins the format-conversion oracle (oracle/formats.py) with scipy.sparse's independent conversions and a
hand-worked example. CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import formats as F


def _rand_csr(R, C, density, seed, dtype=np.float64):
    A = sp.random(R, C, density=density, random_state=seed, format="csr", dtype=dtype)
    A.sort_indices()
    return A


def test_hand_example():
    # [[1 0 2 0 0]
    #  [0 0 0 0 0]
    #  [3 4 0 0 5]]
    rp = np.array([0, 2, 2, 5], np.int32)
    ci = np.array([0, 2, 0, 1, 4], np.int32)
    v = np.array([1, 2, 3, 4, 5], np.float64)
    r, c, vv = F.csr_to_coo(rp, ci, v)
    assert r.tolist() == [0, 0, 2, 2, 2]
    trp, tci, tv = F.transpose(rp, ci, v, 5)
    assert trp.tolist() == [0, 2, 3, 4, 4, 5] and tci.tolist() == [0, 2, 2, 0, 2] and tv.tolist() == [1, 3, 4, 2, 5]
    brp, bci, bv = F.csr_to_bsr(rp, ci, v, 3, 5, 2)
    # block rows {0,1}, {2}; block cols {0,1}, {2,3}, {4}
    assert brp.tolist() == [0, 2, 4] and bci.tolist() == [0, 1, 0, 2]
    assert bv.tolist() == [1, 0, 0, 0,  2, 0, 0, 0,  3, 4, 0, 0,  5, 0, 0, 0]
    # back: explicit zeros inside numRows x numCols stay, padding (row 3, column 5) goes
    rp2, ci2, v2 = F.bsr_to_csr(brp, bci, bv, 3, 5, 2)
    assert rp2.tolist() == [0, 4, 8, 11]
    assert ci2.tolist() == [0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 4]
    assert v2.tolist() == [1, 0, 2, 0, 0, 0, 0, 0, 3, 4, 5]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_transpose_matches_scipy(seed):
    A = _rand_csr(700, 500, 0.02, seed)
    T = A.T.tocsr()
    T.sort_indices()
    rp, ci, v = F.transpose(A.indptr, A.indices, A.data, 500)
    assert np.array_equal(rp, T.indptr) and np.array_equal(ci, T.indices) and np.array_equal(v, T.data)


def test_coo_to_csr_matches_scipy_and_keeps_duplicates_in_order():
    rng = np.random.default_rng(4)
    A = _rand_csr(400, 300, 0.03, 5)
    r, c, v = F.csr_to_coo(A.indptr, A.indices, A.data)
    perm = rng.permutation(r.size)
    rp, ci, vv = F.coo_to_csr(r[perm], c[perm], v[perm], 400)
    assert np.array_equal(rp, A.indptr) and np.array_equal(ci, A.indices) and np.array_equal(vv, A.data)
    # duplicates: same (row, col) three times, out of order
    rp, ci, vv = F.coo_to_csr([1, 0, 1, 1], [2, 0, 2, 2], [10.0, 1.0, 20.0, 30.0], 2)
    assert rp.tolist() == [0, 1, 4] and ci.tolist() == [0, 2, 2, 2] and vv.tolist() == [1, 10, 20, 30]


@pytest.mark.parametrize("b", [1, 2, 4, 8])
def test_bsr_round_trip_matches_scipy(b):
    A = _rand_csr(64, 96, 0.05, 6)
    B = A.tobsr(blocksize=(b, b))
    B.sort_indices()
    brp, bci, bv = F.csr_to_bsr(A.indptr, A.indices, A.data, 64, 96, b)
    assert np.array_equal(brp, B.indptr) and np.array_equal(bci, B.indices)
    assert np.array_equal(bv, B.data.reshape(-1))
    rp, ci, v = F.bsr_to_csr(brp, bci, bv, 64, 96, b)
    Cm = B.tocsr()  # scipy keeps the stored zeros of each block
    Cm.sort_indices()
    assert np.array_equal(rp, Cm.indptr) and np.array_equal(ci, Cm.indices) and np.array_equal(v, Cm.data)
    assert np.array_equal(sp.csr_matrix((v, ci, rp), shape=(64, 96)).toarray(), A.toarray())


def test_bsr_ragged_edges_against_dense():
    A = _rand_csr(37, 29, 0.1, 7)
    for b in (3, 4, 5):
        brp, bci, bv = F.csr_to_bsr(A.indptr, A.indices, A.data, 37, 29, b)
        D = np.zeros((-(-37 // b) * b, -(-29 // b) * b))
        for br in range(brp.size - 1):
            for j in range(brp[br], brp[br + 1]):
                D[br * b:(br + 1) * b, bci[j] * b:(bci[j] + 1) * b] = bv[j * b * b:(j + 1) * b * b].reshape(b, b)
        assert np.array_equal(D[:37, :29], A.toarray()) and not D[37:].any() and not D[:, 29:].any()
        rp, ci, v = F.bsr_to_csr(brp, bci, bv, 37, 29, b)
        assert np.array_equal(sp.csr_matrix((v, ci, rp), shape=(37, 29)).toarray(), A.toarray())
        assert rp[-1] == sum(min(b, 37 - (br * b)) * sum(min(b, 29 - bci[j] * b) for j in range(brp[br], brp[br + 1]))
                             for br in range(brp.size - 1))


def test_bsr_duplicates_rejected():
    with pytest.raises(ValueError):
        F.csr_to_bsr(np.array([0, 2]), np.array([1, 1]), np.array([1.0, 2.0]), 1, 4, 2)
