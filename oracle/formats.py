"""CPU oracle for the format conversions — TEST INFRASTRUCTURE ONLY (numpy).

The reference declares these conversions without defining them (src/formats/sparse_formats.h:389-397:
coo_to_csr, csr_to_coo, bsr_to_csr; src/spmv/csr/spmv_csr.cuh:1107-1114: spmv_csr_transpose), so each
function here restates the textbook definition with the output order the library documents:
  * CSR -> COO: row of each entry repeated from rowPtr, entry order kept;
  * COO -> CSR: stable sort by (row, column), duplicates kept in input order;
  * transpose: stable sort of the entries by column (rows ascending inside an output row);
  * CSR -> BSR: blocks of bs x bs, block columns ascending, row-major inside a block, zero padding;
  * BSR -> CSR: every stored block entry inside numRows x numCols, explicit zeros kept, ordered by
    (row, block, column inside the block).
Pinned against scipy.sparse's independent implementations (tests/test_formats_oracle.py) and a
hand-worked example. Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np


def rows_of(rowPtr) -> np.ndarray:
    rp = np.asarray(rowPtr, np.int64)
    return np.repeat(np.arange(rp.size - 1, dtype=np.int32), np.diff(rp))


def csr_to_coo(rowPtr, colIdx, values):
    """csr_to_coo (sparse_formats.h:393-394)."""
    return rows_of(rowPtr), np.asarray(colIdx, np.int32).copy(), np.asarray(values).copy()


def _offsets(keys: np.ndarray, n: int) -> np.ndarray:
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(keys, minlength=n)[:n], out=rp[1:])
    return rp.astype(np.int32)


def coo_to_csr(rowIdx, colIdx, values, numRows: int):
    """coo_to_csr (sparse_formats.h:390-391): stable (row, column) order."""
    ri, ci, v = np.asarray(rowIdx, np.int64), np.asarray(colIdx, np.int32), np.asarray(values)
    order = np.lexsort((ci, ri))  # lexsort is stable: ties keep input order
    return _offsets(ri, numRows), ci[order], v[order]


def transpose(rowPtr, colIdx, values, numCols: int):
    """A^T in CSR (spmv_csr_transpose, spmv_csr.cuh:1107-1114)."""
    ci = np.asarray(colIdx, np.int64)
    order = np.argsort(ci, kind="stable")
    return _offsets(ci, numCols), rows_of(rowPtr)[order], np.asarray(values)[order]


def csr_to_bsr(rowPtr, colIdx, values, numRows: int, numCols: int, b: int):
    """(blockRowPtr, blockColIdx, blockValues[nb * b * b]) — the inverse of bsr_to_csr."""
    r = rows_of(rowPtr).astype(np.int64)
    c = np.asarray(colIdx, np.int64)
    v = np.asarray(values)
    nbr, nbc = -(-numRows // b), -(-numCols // b)
    key = (r // b) * nbc + c // b
    ukey = np.unique(key)
    brp = np.searchsorted(ukey, np.arange(nbr + 1, dtype=np.int64) * nbc).astype(np.int32)
    slot = np.searchsorted(ukey, key) * b * b + (r % b) * b + (c % b)
    if np.unique(slot).size != slot.size:
        raise ValueError("duplicate (row, column) entries")
    bv = np.zeros(ukey.size * b * b, v.dtype)
    bv[slot] = v
    return brp, (ukey % nbc).astype(np.int32), bv


def bsr_to_csr(blockRowPtr, blockColIdx, blockValues, numRows: int, numCols: int, b: int):
    """bsr_to_csr (sparse_formats.h:396-397)."""
    brp = np.asarray(blockRowPtr, np.int64)
    bci = np.asarray(blockColIdx, np.int64)
    bv = np.asarray(blockValues).reshape(-1)
    nb = bci.size
    j = np.repeat(np.arange(nb), b * b)
    e = np.tile(np.arange(b * b), nb)
    br = np.repeat(np.arange(brp.size - 1), np.diff(brp))[j]
    r = br * b + e // b
    c = bci[j] * b + e % b
    keep = (r < numRows) & (c < numCols)
    r, c, j, jj, val = r[keep], c[keep], j[keep], (e % b)[keep], bv[keep]
    order = np.lexsort((jj, j, r))
    return _offsets(r, numRows), c[order].astype(np.int32), val[order]
