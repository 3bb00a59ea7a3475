"""Matrix Market I/O (muxi_spmv_io.h): the reference's readMTXHeader / readMTXFile / writeMTXFile
(src/generators/mtx_io.cu:70-254) on hand-written files whose CSR is worked out by hand, cross-checked
with scipy.io's independent Matrix Market reader and writer, plus write->read round trips and the
error cases. Host code only: no device calls."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from muxi_spmv_amd import generators as G
from muxi_spmv_amd import mtx
from muxi_spmv_amd._lib import SpmvError


def _put(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


GENERAL = """%%MatrixMarket matrix coordinate real general
% a comment line
%

4 5 6
3 2 -1.5
1 1 2.0
1 5 0.25
3 1 7
4 4 1e-3
1 3 -4
"""


def test_general_file_by_hand(tmp_path):
    p = _put(tmp_path, "g.mtx", GENERAL)
    h = mtx.read_header(p)
    assert (h.numRows, h.numCols, h.nnz) == (4, 5, 6)
    assert not (h.isSymmetric or h.isComplex or h.isPattern or h.isInteger)
    for dt in (np.float32, np.float64):
        rp, ci, v, R, C = mtx.read(p, dt)
        assert (R, C) == (4, 5)
        # row 0: (0,0)=2 (0,2)=-4 (0,4)=.25; row 1 empty; row 2: (2,0)=7 (2,1)=-1.5; row 3: (3,3)=1e-3
        assert rp.tolist() == [0, 3, 3, 5, 6]
        assert ci.tolist() == [0, 2, 4, 0, 1, 3]
        assert v.dtype == dt
        assert np.array_equal(v, np.array([2, -4, 0.25, 7, -1.5, 1e-3], dt))


PATTERN_SYM = """%%MatrixMarket matrix coordinate pattern symmetric
3 3 4
1 1
2 1
3 2
3 3
"""


def test_pattern_symmetric_expand_and_as_stored(tmp_path):
    p = _put(tmp_path, "ps.mtx", PATTERN_SYM)
    h = mtx.read_header(p)
    assert h.isPattern and h.isSymmetric and h.nnz == 4
    rp, ci, v, R, C = mtx.read(p, np.float64, expand_symmetric=True)
    # lower triangle (0,0) (1,0) (2,1) (2,2) mirrored: the diagonal once, off-diagonals twice
    dense = np.array([[1, 1, 0], [1, 0, 1], [0, 1, 1]], np.float64)
    assert np.array_equal(sp.csr_matrix((v, ci, rp), shape=(R, C)).toarray(), dense)
    assert rp.tolist() == [0, 2, 4, 6] and ci.tolist() == [0, 1, 0, 2, 1, 2]
    rp, ci, v, R, C = mtx.read(p, np.float64, expand_symmetric=False)
    assert rp.tolist() == [0, 1, 2, 4] and ci.tolist() == [0, 0, 1, 2] and np.all(v == 1)


def test_integer_and_duplicates_keep_file_order(tmp_path):
    p = _put(tmp_path, "i.mtx", "%%MatrixMarket matrix coordinate integer general\n2 2 4\n2 2 5\n1 2 3\n2 2 -7\n1 1 1\n")
    h = mtx.read_header(p)
    assert h.isInteger
    rp, ci, v, _, _ = mtx.read(p, np.float32)
    assert rp.tolist() == [0, 2, 4] and ci.tolist() == [0, 1, 1, 1]
    assert v.tolist() == [1, 3, 5, -7]  # duplicates kept, in file order


def test_empty_matrix(tmp_path):
    p = _put(tmp_path, "e.mtx", "%%MatrixMarket matrix coordinate real general\n3 7 0\n")
    rp, ci, v, R, C = mtx.read(p)
    assert (R, C) == (3, 7) and rp.tolist() == [0, 0, 0, 0] and ci.size == 0 and v.size == 0
    q = tmp_path / "e2.mtx"
    mtx.write(q, rp, ci, v, R, C)
    rp2, ci2, v2, R2, C2 = mtx.read(q)
    assert (R2, C2) == (3, 7) and rp2.tolist() == [0, 0, 0, 0]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_round_trip_is_exact(tmp_path, dt):
    h = G.powerlaw_rows(1500, 1200, seed=5, kmax=300)
    v = np.random.default_rng(6).standard_normal(h.colIdx.size).astype(dt) * 1e3
    p = tmp_path / "rt.mtx"
    mtx.write(p, h.rowPtr, h.colIdx, v, h.numRows, h.numCols)
    rp, ci, v2, R, C = mtx.read(p, dt)
    assert (R, C) == (h.numRows, h.numCols)
    assert np.array_equal(rp, h.rowPtr) and np.array_equal(ci, h.colIdx)
    assert np.array_equal(v2, v)  # %.9g / %.17g print every bit back


def test_scipy_reads_what_we_write_and_we_read_what_scipy_writes(tmp_path):
    A = sp.random(300, 420, density=0.03, random_state=8, format="csr", dtype=np.float64)
    A.sort_indices()
    p = tmp_path / "ours.mtx"
    mtx.write(p, A.indptr, A.indices, A.data, 300, 420)
    B = scipy.io.mmread(str(p)).tocsr()
    B.sort_indices()
    assert B.shape == A.shape and np.array_equal(B.indptr, A.indptr) and np.array_equal(B.indices, A.indices)
    assert np.array_equal(B.data, A.data)
    q = tmp_path / "scipy.mtx"
    M = sp.random(250, 250, density=0.05, random_state=9, format="coo") * 10
    scipy.io.mmwrite(str(q), M, precision=17)
    rp, ci, v, R, C = mtx.read(q, np.float64)
    E = M.tocsr()
    E.sort_indices()
    assert (R, C) == E.shape
    assert np.array_equal(rp, E.indptr) and np.array_equal(ci, E.indices) and np.array_equal(v, E.data)


def test_symmetric_write_then_expand(tmp_path):
    rng = np.random.default_rng(10)
    L = sp.random(200, 200, density=0.04, random_state=11, format="csr")
    S = (L + L.T + sp.diags(rng.uniform(1, 2, 200))).tocsr()
    S.sort_indices()
    p = tmp_path / "sym.mtx"
    mtx.write(p, S.indptr, S.indices, S.data, 200, 200, symmetric=True)
    h = mtx.read_header(p)
    assert h.isSymmetric and h.nnz == (S.nnz + 200) // 2  # upper triangle with the diagonal
    rp, ci, v, _, _ = mtx.read(p, np.float64)
    assert np.array_equal(rp, S.indptr) and np.array_equal(ci, S.indices) and np.array_equal(v, S.data)
    # scipy agrees on the symmetric file
    assert np.array_equal(scipy.io.mmread(str(p)).toarray(), S.toarray())


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "INVALID_MATRIX"),    # dense layout
    ("%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n", "INVALID_MATRIX"),   # bad banner
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "UNSUPPORTED_TYPE"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "INVALID_MATRIX"),  # row out of range
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n", "INVALID_MATRIX"),  # column 0
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n2 2 1\n", "INVALID_MATRIX"),  # short
    ("%%MatrixMarket matrix coordinate real general\n% only comments\n", "INVALID_MATRIX"),  # no size line
])
def test_errors(tmp_path, text, code):
    p = _put(tmp_path, "bad.mtx", text)
    with pytest.raises(SpmvError, match=code):
        mtx.read(p)


def test_missing_file(tmp_path):
    with pytest.raises(SpmvError, match="cannot open"):
        mtx.read_header(tmp_path / "nope.mtx")
    with pytest.raises(SpmvError, match="cannot open"):
        mtx.read(tmp_path / "nope.mtx")


def test_coo_read_keeps_stored_entries_sorted(tmp_path):
    p = _put(tmp_path, "ps.mtx", PATTERN_SYM)
    ri, ci, v, R, C = mtx.read_coo(p, np.float64)  # no mirroring, like readMTXToCOO
    assert (R, C) == (3, 3)
    assert ri.tolist() == [0, 1, 2, 2] and ci.tolist() == [0, 0, 1, 2] and v.tolist() == [1, 1, 1, 1]
    ri, ci, v, _, _ = mtx.read_coo(_put(tmp_path, "g.mtx", GENERAL), np.float32)
    assert ri.tolist() == [0, 0, 0, 2, 2, 3] and ci.tolist() == [0, 2, 4, 0, 1, 3]


def test_coo_write_order_and_round_trip(tmp_path):
    ri = np.array([2, 0, 1, 0], np.int32)
    ci = np.array([1, 3, 0, 0], np.int32)
    v = np.array([0.1, -2.5, 3e-8, 7.0], np.float64)
    p = tmp_path / "c.mtx"
    mtx.write_coo(p, ri, ci, v, 3, 4)
    lines = [ln for ln in p.read_text().splitlines() if not ln.startswith("%")]
    assert lines[0] == "3 4 4" and lines[1].split()[:2] == ["3", "2"]  # given order, 1-based
    r2, c2, v2, R, C = mtx.read_coo(p, np.float64)
    order = np.lexsort((ci, ri))
    assert (R, C) == (3, 4) and np.array_equal(r2, ri[order]) and np.array_equal(c2, ci[order])
    assert np.array_equal(v2, v[order])
    with pytest.raises(SpmvError, match="INVALID_MATRIX"):
        mtx.write_coo(tmp_path / "bad.mtx", [3], [0], np.ones(1), 3, 4)


def test_huge_declared_size_is_an_error_not_a_crash(tmp_path):
    # the header claims 2e9 entries; the file holds one
    p = _put(tmp_path, "h.mtx", "%%MatrixMarket matrix coordinate real general\n100 100 2000000000\n1 1 1\n")
    with pytest.raises(SpmvError, match="INVALID_MATRIX"):
        mtx.read(p)
