"""Matrix Market files <-> CSR arrays (the reference's readMTXHeader / readMTXFile / writeMTXFile,
src/generators/mtx_io.cu), through the library's host entry points (include/muxi_spmv/muxi_spmv_io.h)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import check, load


class MTXHeader(C.Structure):
    _fields_ = [("isSymmetric", C.c_bool), ("isComplex", C.c_bool), ("isPattern", C.c_bool),
                ("isInteger", C.c_bool), ("numRows", C.c_int), ("numCols", C.c_int), ("nnz", C.c_int)]


def read_header(path) -> MTXHeader:
    h = MTXHeader()
    check(load().muxi_mtx_read_header(str(path).encode(), C.byref(h)), "mtx_read_header")
    return h


def read(path, dtype=np.float32, expand_symmetric: bool = True):
    """(rowPtr, colIdx, values, numRows, numCols) as numpy arrays (readMTXFile<T>)."""
    c = _lib.muxi_csr_t()
    fn = load().muxi_mtx_read_csr_f64 if np.dtype(dtype) == np.float64 else load().muxi_mtx_read_csr_f32
    check(fn(str(path).encode(), int(expand_symmetric), C.byref(c)), "mtx_read")
    try:
        R, N = c.numRows, c.nnz
        rp = np.ctypeslib.as_array(C.cast(c.rowPtr, C.POINTER(C.c_int)), (R + 1,)).copy()
        if N:
            ci = np.ctypeslib.as_array(C.cast(c.colIdx, C.POINTER(C.c_int)), (N,)).copy()
            ct = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
            va = np.ctypeslib.as_array(C.cast(c.values, C.POINTER(ct)), (N,)).copy()
        else:
            ci, va = np.zeros(0, np.int32), np.zeros(0, dtype)
        return rp, ci, va, R, c.numCols
    finally:
        load().muxi_csr_free_host(C.byref(c))


def write(path, rowPtr, colIdx, values, numRows: int, numCols: int, symmetric: bool = False) -> None:
    """writeMTXFile<T> from host arrays (values FP32 or FP64)."""
    rp = np.ascontiguousarray(rowPtr, np.int32)
    ci = np.ascontiguousarray(colIdx, np.int32)
    va = np.ascontiguousarray(values)
    c = _lib.muxi_csr_t()
    c.numRows, c.numCols, c.nnz = numRows, numCols, int(ci.size)
    c.rowPtr = rp.ctypes.data
    c.colIdx = ci.ctypes.data
    c.values = va.ctypes.data
    if va.dtype not in (np.float32, np.float64):
        raise TypeError("values must be float32 or float64")
    fn = load().muxi_mtx_write_csr_f64 if va.dtype == np.float64 else load().muxi_mtx_write_csr_f32
    check(fn(str(Path(path)).encode(), C.byref(c), int(symmetric)), "mtx_write")


def read_coo(path, dtype=np.float32):
    """(rowIdx, colIdx, values, numRows, numCols): the entries as stored, (row, column) order (readMTXToCOO)."""
    c = _lib.muxi_coo_t()
    fn = load().muxi_mtx_read_coo_f64 if np.dtype(dtype) == np.float64 else load().muxi_mtx_read_coo_f32
    check(fn(str(path).encode(), C.byref(c)), "mtx_read_coo")
    try:
        N = c.nnz
        if N == 0:
            return np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, dtype), c.numRows, c.numCols
        ct = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
        ri = np.ctypeslib.as_array(C.cast(c.rowIdx, C.POINTER(C.c_int)), (N,)).copy()
        ci = np.ctypeslib.as_array(C.cast(c.colIdx, C.POINTER(C.c_int)), (N,)).copy()
        va = np.ctypeslib.as_array(C.cast(c.values, C.POINTER(ct)), (N,)).copy()
        return ri, ci, va, c.numRows, c.numCols
    finally:
        load().muxi_coo_free_host(C.byref(c))


def write_coo(path, rowIdx, colIdx, values, numRows: int, numCols: int) -> None:
    """writeCOOToMTX from host arrays, entries in the given order."""
    ri = np.ascontiguousarray(rowIdx, np.int32)
    ci = np.ascontiguousarray(colIdx, np.int32)
    va = np.ascontiguousarray(values)
    if va.dtype not in (np.float32, np.float64):
        raise TypeError("values must be float32 or float64")
    c = _lib.muxi_coo_t()
    c.numRows, c.numCols, c.nnz = numRows, numCols, int(ci.size)
    c.rowIdx, c.colIdx, c.values = ri.ctypes.data, ci.ctypes.data, va.ctypes.data
    fn = load().muxi_mtx_write_coo_f64 if va.dtype == np.float64 else load().muxi_mtx_write_coo_f32
    check(fn(str(Path(path)).encode(), C.byref(c)), "mtx_write_coo")
