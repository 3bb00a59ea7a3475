"""COO and BSR containers and the device conversions of include/muxi_spmv/muxi_spmv_formats.h (the
reference's COOMatrix / BSRMatrix, src/formats/sparse_formats.h:113-267, and its declared-only coo_to_csr,
csr_to_coo, bsr_to_csr, sparse_formats.h:389-397, and spmv_csr_transpose, spmv_csr.cuh:1107-1114).

Every conversion runs in the library on the GPU; the arrays it allocates are copied into torch tensors and
released at once, so the Python objects own ordinary torch memory."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import SpmvError, check, load
from .api import _TORCH2SPMV, CSRMatrix, _ptr, _stream

_ELEM = {torch.float32: ("<f4", torch.float32), torch.float64: ("<f8", torch.float64),
         torch.bfloat16: ("<i2", torch.int16)}


class _DevView:
    """A library-owned device array seen through __cuda_array_interface__ (aliases, never frees)."""

    def __init__(self, ptr: int, n: int, typestr: str):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False), "version": 2,
                                         "strides": None}


def _take(ptr, n: int, dtype: torch.dtype) -> torch.Tensor:
    """Copy n elements of a library device array into a new torch tensor."""
    if n == 0:
        return torch.empty(0, dtype=dtype, device="cuda")
    if dtype == torch.int32:
        return torch.as_tensor(_DevView(ptr, n, "<i4"), device="cuda").clone()
    typestr, _ = _ELEM[dtype]
    return torch.as_tensor(_DevView(ptr, n, typestr), device="cuda").clone().view(dtype)


def _csr_from_lib(c: _lib.muxi_csr_t, dtype: torch.dtype) -> CSRMatrix:
    try:
        torch.cuda.synchronize()
        rp = _take(c.d_rowPtr, c.numRows + 1, torch.int32)
        ci = _take(c.d_colIdx, c.nnz, torch.int32)
        va = _take(c.d_values, c.nnz, dtype)
        torch.cuda.synchronize()
    finally:
        load().muxi_csr_free_device(C.byref(c))
    return CSRMatrix(rp, ci, va, c.numCols)


class COOMatrix:
    """COOMatrix<T>: numRows, numCols, nnz and device arrays d_rowIdx, d_colIdx (int32[N]), d_values (T[N])."""

    def __init__(self, rowIdx, colIdx, values, numRows: int, numCols: int):
        self.d_rowIdx = torch.as_tensor(rowIdx, dtype=torch.int32).cuda().contiguous()
        self.d_colIdx = torch.as_tensor(colIdx, dtype=torch.int32).cuda().contiguous()
        self.d_values = values.cuda().contiguous() if torch.is_tensor(values) else torch.as_tensor(values).cuda()
        if self.d_values.dtype not in _TORCH2SPMV:
            raise SpmvError(-8, "COOMatrix", f"unsupported value dtype {self.d_values.dtype}")
        self.numRows, self.numCols, self.nnz = int(numRows), int(numCols), int(self.d_colIdx.numel())
        if self.d_rowIdx.numel() != self.nnz or self.d_values.numel() != self.nnz:
            raise SpmvError(-2, "COOMatrix", "inconsistent array sizes")
        self.dtype = self.d_values.dtype
        c = self.c = _lib.muxi_coo_t()
        c.numRows, c.numCols, c.nnz = self.numRows, self.numCols, self.nnz
        c.d_rowIdx, c.d_colIdx, c.d_values = _ptr(self.d_rowIdx), _ptr(self.d_colIdx), _ptr(self.d_values)


class BSRMatrix:
    """BSRMatrix<T, BlockSize>: block rowPtr (int32[numBlockRows+1]), block colIdx (int32[numBlocks]) and
    values (T[numBlocks * bs * bs], row-major inside each block), on the device."""

    def __init__(self, rowPtr, colIdx, values, numRows: int, numCols: int, blockSize: int):
        b = int(blockSize)
        self.d_rowPtr = torch.as_tensor(rowPtr, dtype=torch.int32).cuda().contiguous()
        self.d_colIdx = torch.as_tensor(colIdx, dtype=torch.int32).cuda().contiguous()
        self.d_values = values.cuda().contiguous().reshape(-1) if torch.is_tensor(values) else \
            torch.as_tensor(values).cuda().reshape(-1)
        self.numRows, self.numCols, self.blockSize = int(numRows), int(numCols), b
        self.numBlockRows = int(self.d_rowPtr.numel()) - 1
        self.numBlockCols = (self.numCols + b - 1) // b if b > 0 else 0
        self.numBlocks = int(self.d_colIdx.numel())
        self.dtype = self.d_values.dtype
        c = self.c = _lib.muxi_bsr_t()
        c.numRows, c.numCols, c.numBlockRows, c.numBlockCols = self.numRows, self.numCols, self.numBlockRows, \
            self.numBlockCols
        c.numBlocks, c.blockSize = self.numBlocks, b
        c.d_rowPtr, c.d_colIdx, c.d_values = _ptr(self.d_rowPtr), _ptr(self.d_colIdx), _ptr(self.d_values)


def csr_to_coo(A: CSRMatrix, stream=None) -> COOMatrix:
    c = _lib.muxi_coo_t()
    check(load().muxi_csr_to_coo(C.byref(A.c), A.vtype, C.byref(c), _stream(stream)), "csr_to_coo")
    try:
        torch.cuda.synchronize()
        ri = _take(c.d_rowIdx, c.nnz, torch.int32)
        ci = _take(c.d_colIdx, c.nnz, torch.int32)
        va = _take(c.d_values, c.nnz, A.dtype)
        torch.cuda.synchronize()
    finally:
        load().muxi_coo_free_device(C.byref(c))
    return COOMatrix(ri, ci, va, c.numRows, c.numCols)


def coo_to_csr(A: COOMatrix, stream=None) -> CSRMatrix:
    c = _lib.muxi_csr_t()
    check(load().muxi_coo_to_csr(C.byref(A.c), _TORCH2SPMV[A.dtype], C.byref(c), _stream(stream)), "coo_to_csr")
    return _csr_from_lib(c, A.dtype)


def csr_transpose(A: CSRMatrix, stream=None) -> CSRMatrix:
    c = _lib.muxi_csr_t()
    check(load().muxi_csr_transpose(C.byref(A.c), A.vtype, C.byref(c), _stream(stream)), "csr_transpose")
    return _csr_from_lib(c, A.dtype)


def csr_to_bsr(A: CSRMatrix, blockSize: int, stream=None) -> BSRMatrix:
    c = _lib.muxi_bsr_t()
    check(load().muxi_csr_to_bsr(C.byref(A.c), A.vtype, int(blockSize), C.byref(c), _stream(stream)), "csr_to_bsr")
    try:
        torch.cuda.synchronize()
        b = c.blockSize
        rp = _take(c.d_rowPtr, c.numBlockRows + 1, torch.int32)
        ci = _take(c.d_colIdx, c.numBlocks, torch.int32)
        va = _take(c.d_values, c.numBlocks * b * b, A.dtype)
        torch.cuda.synchronize()
    finally:
        load().muxi_bsr_free_device(C.byref(c))
    return BSRMatrix(rp, ci, va, c.numRows, c.numCols, b)


def bsr_to_csr(A: BSRMatrix, stream=None) -> CSRMatrix:
    c = _lib.muxi_csr_t()
    check(load().muxi_bsr_to_csr(C.byref(A.c), _TORCH2SPMV[A.dtype], C.byref(c), _stream(stream)), "bsr_to_csr")
    return _csr_from_lib(c, A.dtype)
