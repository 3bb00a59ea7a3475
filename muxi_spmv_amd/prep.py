"""Preprocessing mirror: permutations of src/preprocessing/column_reorder.cuh
and the test-resident reorderings, over include/muxi_spmv/muxi_spmv_prep.h.

Host functions take numpy arrays (as the reference's take host arrays); the
*_dev functions apply maps to device tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, load


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a.size else None


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def reorder_columns(rowPtr, colIdx, numCols: int):
    """reorderColumns (column_reorder.cuh:42-141) -> (forwardMap, newColIdx, stats dict)."""
    rp, ci = _i32(rowPtr), _i32(colIdx)
    fwd = np.empty(numCols, np.int32)
    nci = np.empty(ci.size, np.int32)
    st = _lib.muxi_column_reorder_stats_t()
    check(load().muxi_reorder_columns(rp.size - 1, numCols, _p(rp), _p(ci), _p(fwd), _p(nci), C.byref(st)),
          "reorder_columns")
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    return fwd, nci, stats


def reorder_x(x, forwardMap) -> np.ndarray:
    """reorderXVector (column_reorder.cuh:150-161)."""
    xx = np.ascontiguousarray(x, np.float32)
    f = _i32(forwardMap)
    out = np.empty(f.size, np.float32)
    check(load().muxi_reorder_x_f32(_p(xx), _p(out), _p(f), f.size), "reorder_x")
    return out


def sort_columns_within_rows(rowPtr, colIdx, values):
    """sortColumnsWithinRows (column_reorder.cuh:175-201); returns new arrays."""
    rp = _i32(rowPtr)
    ci = _i32(colIdx).copy()
    v = np.ascontiguousarray(values, np.float32).copy()
    check(load().muxi_sort_columns_within_rows_f32(rp.size - 1, _p(rp), _p(ci), _p(v)), "sort_columns_within_rows")
    return ci, v


def reorder_columns_minrow(rowPtr, colIdx, numCols: int):
    """reorderColumnsRCM (tests/test_column_reordering.cu:156-197)."""
    rp, ci = _i32(rowPtr), _i32(colIdx)
    perm = np.empty(numCols, np.int32)
    nci = np.empty(ci.size, np.int32)
    check(load().muxi_reorder_columns_minrow(rp.size - 1, numCols, _p(rp), _p(ci), _p(perm), _p(nci)),
          "reorder_columns_minrow")
    return perm, nci


def reorder_columns_by_frequency(rowPtr, colIdx, numCols: int):
    """reorderColumnsByFrequency (tests/test_column_reordering.cu:116-153)."""
    rp, ci = _i32(rowPtr), _i32(colIdx)
    perm = np.empty(numCols, np.int32)
    nci = np.empty(ci.size, np.int32)
    check(load().muxi_reorder_columns_by_frequency(rp.size - 1, numCols, _p(rp), _p(ci), _p(perm), _p(nci)),
          "reorder_columns_by_frequency")
    return perm, nci


def reorder_rows_by_length(rowPtr, colIdx, values):
    """reorderMatrixByRowLength (tests/test_row_reordering.cu:99-141)."""
    rp, ci = _i32(rowPtr), _i32(colIdx)
    v = np.ascontiguousarray(values, np.float32)
    R = rp.size - 1
    rm, irm = np.empty(R, np.int32), np.empty(R, np.int32)
    nrp, nci, nv = np.empty(R + 1, np.int32), np.empty(ci.size, np.int32), np.empty(ci.size, np.float32)
    check(load().muxi_reorder_rows_by_length_f32(R, _p(rp), _p(ci), _p(v), _p(rm), _p(irm), _p(nrp), _p(nci),
                                                 _p(nv)), "reorder_rows_by_length")
    return rm, irm, nrp, nci, nv


def rcm(rowPtr, colIdx) -> np.ndarray:
    """Reverse Cuthill-McKee order (perm[new] = old) of a square pattern."""
    rp, ci = _i32(rowPtr), _i32(colIdx)
    perm = np.empty(rp.size - 1, np.int32)
    check(load().muxi_rcm(rp.size - 1, _p(rp), _p(ci), _p(perm)), "rcm")
    return perm


def permute_symmetric(rowPtr, colIdx, values, perm):
    """B = P A P^T for perm (new -> old)."""
    rp, ci, pm = _i32(rowPtr), _i32(colIdx), _i32(perm)
    v = np.ascontiguousarray(values)
    vt = {np.dtype(np.float32): _lib.SPMV_TYPE_FLOAT, np.dtype(np.float64): _lib.SPMV_TYPE_DOUBLE}[v.dtype]
    n = rp.size - 1
    nrp, nci, nv = np.empty(n + 1, np.int32), np.empty(ci.size, np.int32), np.empty_like(v)
    check(load().muxi_permute_symmetric(n, _p(rp), _p(ci), _p(v), vt, _p(pm), _p(nrp), _p(nci), _p(nv)),
          "permute_symmetric")
    return nrp, nci, nv


def shard_rows(rowPtr, G: int) -> np.ndarray:
    """Row bounds[G+1] of G shards with about equal (rows + nonzeros), no row split (muxi_shard_rows)."""
    rp = _i32(rowPtr)
    b = np.empty(G + 1, np.int32)
    check(load().muxi_shard_rows(_p(rp), rp.size - 1, G, _p(b)), "shard_rows")
    return b


def bandwidth(rowPtr, colIdx) -> int:
    rp, ci = _i32(rowPtr), _i32(colIdx)
    return int(load().muxi_bandwidth(rp.size - 1, _p(rp), _p(ci)))


# ---- device application
def _dptr(t: torch.Tensor):
    return t.data_ptr() if t.numel() else None


def _s():
    return torch.cuda.current_stream().cuda_stream or None


def remap_indices_dev(idx: torch.Tensor, mapping: torch.Tensor) -> torch.Tensor:
    out = torch.empty_like(idx)
    check(load().muxi_remap_indices_dev(_dptr(idx), _dptr(mapping), _dptr(out), idx.numel(), _s()), "remap_indices")
    return out


def gather_dev(src: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    """dst[i] = src[idx[i]] (reorderXVector on the device)."""
    dst = torch.empty(idx.numel(), dtype=src.dtype, device=src.device)
    check(load().muxi_gather_dev(_dptr(src), _dptr(idx), _dptr(dst), idx.numel(), src.element_size(), _s()),
          "gather")
    return dst


def scatter_dev(src: torch.Tensor, idx: torch.Tensor, n: int | None = None) -> torch.Tensor:
    """dst[idx[i]] = src[i] (restoreOrder)."""
    dst = torch.empty(n if n is not None else src.numel(), dtype=src.dtype, device=src.device)
    check(load().muxi_scatter_dev(_dptr(src), _dptr(idx), _dptr(dst), src.numel(), src.element_size(), _s()),
          "scatter")
    return dst


def permute_rows_dev(A, rowMap: torch.Tensor, newRowPtr: torch.Tensor):
    """Rows of A gathered by rowMap into a new CSRMatrix (newRowPtr precomputed)."""
    from .api import CSRMatrix

    nci = torch.empty_like(A.d_colIdx)
    nv = torch.empty_like(A.d_values)
    check(load().muxi_permute_rows_dev(A.numRows, _dptr(A.d_rowPtr), _dptr(A.d_colIdx), _dptr(A.d_values),
                                       A.d_values.element_size(), _dptr(rowMap), _dptr(newRowPtr), _dptr(nci),
                                       _dptr(nv), _s()), "permute_rows")
    return CSRMatrix(newRowPtr, nci, nv, A.numCols)
