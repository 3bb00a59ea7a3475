"""GPU parity for the device format conversions (muxi_spmv_formats.h): bit-exact against the numpy oracle
(oracle/formats.py, itself pinned to scipy.sparse) on the structural edge cases, FP32 / FP64 / BF16 values,
repeat calls byte-identical, error paths, a deterministic A^T x through a plan on the transpose, and
full-size round trips on the 2M x 2M power-law matrix."""
import numpy as np
import pytest
import torch

from muxi_spmv_amd import api, formats
from muxi_spmv_amd import generators as G
from muxi_spmv_amd._lib import SpmvError
from oracle import formats as F
from gpu_util import assert_close_fp32, dev_csr, matrix_zoo

pytestmark = pytest.mark.gpu

ZOO = matrix_zoo()
DTYPES = [torch.float32, torch.float64, torch.bfloat16]


def _np(t):
    return t.cpu().numpy() if t.dtype != torch.bfloat16 else t.view(torch.int16).cpu().numpy()


def _vals(h, dtype):
    """Host values as stored in `dtype` and the bit pattern the device holds."""
    t = torch.from_numpy(np.ascontiguousarray(h.values)).to(dtype)
    return t, _np(t)


def _same_csr(M, rp, ci, v):
    assert np.array_equal(M.d_rowPtr.cpu().numpy(), rp)
    assert np.array_equal(M.d_colIdx.cpu().numpy(), ci)
    assert np.array_equal(_np(M.d_values), v)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("mname", list(ZOO))
def test_csr_coo_round_trip_bit_exact(mname, dtype):
    h = ZOO[mname]
    t, bits = _vals(h, dtype)
    A = api.CSRMatrix(h.rowPtr, h.colIdx, t, h.numCols)
    coo = formats.csr_to_coo(A)
    r, c, v = F.csr_to_coo(h.rowPtr, h.colIdx, bits)
    assert np.array_equal(coo.d_rowIdx.cpu().numpy(), r) and np.array_equal(coo.d_colIdx.cpu().numpy(), c)
    assert np.array_equal(_np(coo.d_values), v)
    # shuffled COO (unsorted path) back to CSR: the oracle's stable (row, column) order
    perm = np.random.default_rng(1).permutation(r.size)
    sh = formats.COOMatrix(r[perm], c[perm], t[torch.from_numpy(perm)], h.numRows, h.numCols)
    B = formats.coo_to_csr(sh)
    _same_csr(B, *F.coo_to_csr(r[perm], c[perm], bits[perm], h.numRows))
    # sorted COO (the no-sort path) gives the original CSR back
    _same_csr(formats.coo_to_csr(coo), h.rowPtr, h.colIdx, bits)


def test_coo_duplicates_keep_input_order():
    ri = np.array([3, 0, 3, 1, 3, 0], np.int32)
    ci = np.array([2, 5, 2, 1, 2, 5], np.int32)
    v = torch.tensor([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    B = formats.coo_to_csr(formats.COOMatrix(ri, ci, v, 4, 6))
    _same_csr(B, *F.coo_to_csr(ri, ci, v.numpy(), 4))
    assert B.d_values.cpu().tolist() == [2.0, 6.0, 4.0, 1.0, 3.0, 5.0]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("mname", list(ZOO))
def test_transpose_bit_exact_and_repeatable(mname, dtype):
    h = ZOO[mname]
    t, bits = _vals(h, dtype)
    A = api.CSRMatrix(h.rowPtr, h.colIdx, t, h.numCols)
    T1 = formats.csr_transpose(A)
    assert (T1.numRows, T1.numCols, T1.nnz) == (h.numCols, h.numRows, h.nnz)
    _same_csr(T1, *F.transpose(h.rowPtr, h.colIdx, bits, h.numCols))
    T2 = formats.csr_transpose(A)
    assert torch.equal(T1.d_colIdx, T2.d_colIdx) and torch.equal(T1.d_values.view(torch.uint8),
                                                                    T2.d_values.view(torch.uint8))
    # (A^T)^T = A when columns are sorted within rows (the generators sort them)
    _same_csr(formats.csr_transpose(T1), h.rowPtr, h.colIdx, bits)


def test_transpose_unsorted_columns():
    h = G.uniform_rows(2000, 1500, 0, 30, seed=11)
    rng = np.random.default_rng(12)
    ci = h.colIdx.copy()
    for r in range(h.numRows):  # scramble columns inside each row
        a, b = h.rowPtr[r], h.rowPtr[r + 1]
        p = rng.permutation(b - a)
        ci[a:b] = ci[a:b][p]
    A = api.CSRMatrix(h.rowPtr, ci, torch.from_numpy(h.values), h.numCols)
    _same_csr(formats.csr_transpose(A), *F.transpose(h.rowPtr, ci, h.values, h.numCols))


@pytest.mark.parametrize("mname", ["uniform", "powerlaw", "empty_rows", "banded", "blockdiag", "tiny", "single_row"])
@pytest.mark.parametrize("kernel,param", [("auto", 0), ("merge", 4), ("xslab", 1028), ("adaptive", 1024)])
def test_transpose_plan_computes_At_x(mname, kernel, param):
    """y = A^T x through a plan on muxi_csr_transpose: FP32 tolerance against the FP64 oracle on A^T,
    and bitwise identical on a second run (no atomics anywhere)."""
    h = ZOO[mname]
    A = dev_csr(h)
    T = formats.csr_transpose(A)
    rp, ci, v = F.transpose(h.rowPtr, h.colIdx, h.values, h.numCols)
    hT = G.HostCSR(rp, ci, v, h.numCols, h.numRows)
    x = G.uniform_vector(h.numRows, 5)
    p = api.Plan(T, kernel, param)
    y = torch.full((h.numCols,), float("nan"), device="cuda")
    p.exec(torch.from_numpy(x).cuda(), y)
    y2 = torch.empty_like(y)
    p.exec(torch.from_numpy(x).cuda(), y2)
    torch.cuda.synchronize()
    assert_close_fp32(y, hT, x)
    assert torch.equal(y, y2)


@pytest.mark.parametrize("b", [1, 2, 3, 4, 8, 16, 64])
@pytest.mark.parametrize("mname", ["uniform", "empty_rows", "banded", "blockdiag", "tiny", "single_row", "powerlaw"])
def test_bsr_round_trip_bit_exact(mname, b):
    h = ZOO[mname]
    if b == 64 and h.nnz > 100_000:
        pytest.skip("64 x 64 blocks around scattered entries: 4096 slots per entry")
    A = dev_csr(h)
    B = formats.csr_to_bsr(A, b)
    brp, bci, bv = F.csr_to_bsr(h.rowPtr, h.colIdx, h.values, h.numRows, h.numCols, b)
    assert (B.numBlockRows, B.numBlocks, B.blockSize) == (brp.size - 1, bci.size, b)
    assert np.array_equal(B.d_rowPtr.cpu().numpy(), brp) and np.array_equal(B.d_colIdx.cpu().numpy(), bci)
    assert np.array_equal(B.d_values.cpu().numpy(), bv)
    Cm = formats.bsr_to_csr(B)
    _same_csr(Cm, *F.bsr_to_csr(brp, bci, bv, h.numRows, h.numCols, b))


@pytest.mark.parametrize("dtype", DTYPES)
def test_bsr_dtypes_and_ragged_edges(dtype):
    h = G.uniform_rows(37, 29, 0, 6, seed=13)
    t, bits = _vals(h, dtype)
    A = api.CSRMatrix(h.rowPtr, h.colIdx, t, h.numCols)
    for b in (3, 5, 7):
        B = formats.csr_to_bsr(A, b)
        brp, bci, bv = F.csr_to_bsr(h.rowPtr, h.colIdx, bits, 37, 29, b)
        assert np.array_equal(_np(B.d_values), bv)
        _same_csr(formats.bsr_to_csr(B), *F.bsr_to_csr(brp, bci, bv, 37, 29, b))


def test_bsr_to_csr_given_blocks():
    """A hand-made BSR (2 x 2 blocks, 3 x 5 matrix): padding dropped, stored zeros kept."""
    brp = np.array([0, 2, 4], np.int32)
    bci = np.array([0, 1, 0, 2], np.int32)
    bv = torch.tensor([1, 0, 0, 0, 2, 0, 0, 0, 3, 4, 0, 0, 5, 0, 0, 0], dtype=torch.float64)
    Cm = formats.bsr_to_csr(formats.BSRMatrix(brp, bci, bv, 3, 5, 2))
    assert Cm.d_rowPtr.cpu().tolist() == [0, 4, 8, 11]
    assert Cm.d_colIdx.cpu().tolist() == [0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 4]
    assert Cm.d_values.cpu().tolist() == [1, 0, 2, 0, 0, 0, 0, 0, 3, 4, 5]


def test_empty_matrices():
    A = api.CSRMatrix(np.zeros(6, np.int32), np.zeros(0, np.int32), torch.zeros(0), 4)
    T = formats.csr_transpose(A)
    assert T.d_rowPtr.cpu().tolist() == [0] * 5 and T.nnz == 0
    coo = formats.csr_to_coo(A)
    assert coo.nnz == 0
    assert formats.coo_to_csr(coo).d_rowPtr.cpu().tolist() == [0] * 6
    B = formats.csr_to_bsr(A, 2)
    assert B.numBlocks == 0 and B.d_rowPtr.cpu().tolist() == [0, 0, 0, 0]
    assert formats.bsr_to_csr(B).d_rowPtr.cpu().tolist() == [0] * 6


def test_errors():
    # rowPtr that is not monotone
    A = api.CSRMatrix(np.array([0, 3, 2, 4], np.int32), np.zeros(4, np.int32), torch.ones(4), 3)
    with pytest.raises(SpmvError, match="INVALID_MATRIX"):
        formats.csr_transpose(A)
    # column out of range
    A = api.CSRMatrix(np.array([0, 1, 2], np.int32), np.array([0, 7], np.int32), torch.ones(2), 3)
    with pytest.raises(SpmvError, match="INVALID_MATRIX"):
        formats.csr_to_coo(A)
    # COO row out of range
    with pytest.raises(SpmvError, match="INVALID_MATRIX"):
        formats.coo_to_csr(formats.COOMatrix(np.array([0, 5]), np.array([0, 1]), torch.ones(2), 3, 3))
    # duplicate entry cannot be placed in one block slot
    A = api.CSRMatrix(np.array([0, 2], np.int32), np.array([1, 1], np.int32), torch.ones(2), 4)
    with pytest.raises(SpmvError, match="duplicate"):
        formats.csr_to_bsr(A, 2)
    with pytest.raises(SpmvError, match="INVALID_MATRIX"):
        formats.csr_to_bsr(dev_csr(ZOO["tiny"]), 65)


# ---------------------------------------------------------------- full size
@pytest.fixture(scope="module")
def powerlaw_full():
    h = G.cfg_powerlaw_2m()
    return h, dev_csr(h)


def test_full_size_round_trips(powerlaw_full):
    """2M x 2M power-law (14.4M nnz): (A^T)^T = A and CSR -> COO -> CSR = A, bit for bit; A^T's row
    lengths are A's column counts."""
    h, A = powerlaw_full
    T = formats.csr_transpose(A)
    counts = np.bincount(h.colIdx, minlength=h.numCols)
    assert np.array_equal(np.diff(T.d_rowPtr.cpu().numpy().astype(np.int64)), counts)
    TT = formats.csr_transpose(T)
    assert torch.equal(TT.d_rowPtr, A.d_rowPtr) and torch.equal(TT.d_colIdx, A.d_colIdx)
    assert torch.equal(TT.d_values, A.d_values)
    del T, TT
    coo = formats.csr_to_coo(A)
    back = formats.coo_to_csr(coo)
    assert torch.equal(back.d_rowPtr, A.d_rowPtr) and torch.equal(back.d_colIdx, A.d_colIdx)
    assert torch.equal(back.d_values, A.d_values)


def test_cpp_formats_caller(tmp_path):
    """COO / BSR / Matrix Market / transpose names of the C++ shim, from a reference-style caller."""
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parent / "cpp" / "test_formats_io"
    assert exe.exists(), "build with python -m muxi_spmv_amd.build"
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
