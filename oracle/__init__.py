"""CPU oracle for the SpMV path — TEST INFRASTRUCTURE ONLY.

ctypes wrappers over oracle/liboracle.so (built from oracle/oracle.cpp by
oracle/Makefile). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline use this package, as the checker and the CPU baseline; the
library under muxi_spmv_amd/ never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    r = subprocess.run(["make", "-C", str(HERE), "-B", "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- SpMV
def spmv_csr_f32(rowPtr, colIdx, values, x) -> np.ndarray:
    rp, ci, v, xx = _i32(rowPtr), _i32(colIdx), _f32(values), _f32(x)
    R = rp.size - 1
    y = np.empty(R, np.float32)
    lib().oracle_spmv_csr_f32(C.c_int(R), _p(rp), _p(ci), _p(v), _p(xx), _p(y))
    return y


def spmv_csr_f32_mt(rowPtr, colIdx, values, x, threads: int) -> np.ndarray:
    rp, ci, v, xx = _i32(rowPtr), _i32(colIdx), _f32(values), _f32(x)
    R = rp.size - 1
    y = np.empty(R, np.float32)
    lib().oracle_spmv_csr_f32_mt(C.c_int(R), _p(rp), _p(ci), _p(v), _p(xx), _p(y), C.c_int(threads))
    return y


def spmv_f64(rowPtr, colIdx, values, x) -> tuple[np.ndarray, np.ndarray]:
    """FP64 oracle (Dot2-compensated) and per-row sum |a_ij x_j|."""
    rp, ci, v, xx = _i32(rowPtr), _i32(colIdx), _f64(values), _f64(x)
    R = rp.size - 1
    y = np.empty(R, np.float64)
    ab = np.empty(R, np.float64)
    lib().oracle_spmv_f64(C.c_int(R), _p(rp), _p(ci), _p(v), _p(xx), _p(y), _p(ab))
    return y, ab


def spmv_csr_f64_plain(rowPtr, colIdx, values, x) -> np.ndarray:
    rp, ci, v, xx = _i32(rowPtr), _i32(colIdx), _f64(values), _f64(x)
    R = rp.size - 1
    y = np.empty(R, np.float64)
    lib().oracle_spmv_csr_f64_plain(C.c_int(R), _p(rp), _p(ci), _p(v), _p(xx), _p(y))
    return y


# ---------------------------------------------------------------- merge path
def merge_partition_count(numRows, nnz, ePP=16, mult=8, numSMs=104, warpsPerBlock=4) -> int:
    return int(lib().oracle_merge_partition_count(numRows, nnz, ePP, mult, numSMs, warpsPerBlock))


def merge_partitions(numRows, nnz, P) -> np.ndarray:
    pos = np.empty(P + 1, np.int32)
    lib().oracle_merge_partitions(C.c_int(numRows), C.c_int(nnz), C.c_int(P), _p(pos))
    return pos


def merge_path_coords(rowPtr, nnz, pos) -> tuple[np.ndarray, np.ndarray]:
    rp, ps = _i32(rowPtr), _i32(pos)
    n = ps.size
    r = np.empty(n, np.int32)
    z = np.empty(n, np.int32)
    lib().oracle_merge_path_coords(_p(rp), C.c_int(rp.size - 1), C.c_int(nnz), _p(ps), C.c_int(n), _p(r), _p(z))
    return r, z


# ---------------------------------------------------------------- CSR5
def csr5_tiles(rowPtr, nnz, sigma) -> tuple[np.ndarray, np.ndarray]:
    rp = _i32(rowPtr)
    T = (nnz + sigma - 1) // sigma
    tr = np.empty(T, np.int32)
    to = np.empty(T, np.int32)
    lib().oracle_csr5_tiles(_p(rp), C.c_int(rp.size - 1), C.c_int(nnz), C.c_int(sigma), C.c_int(T), _p(tr), _p(to))
    return tr, to


def csr5_default_sigma(warpSize=64) -> int:
    return int(lib().oracle_csr5_default_sigma(warpSize))


# ---------------------------------------------------------------- ELL
def ell_width(rowPtr) -> int:
    rp = _i32(rowPtr)
    return int(lib().oracle_ell_width(C.c_int(rp.size - 1), _p(rp)))


def csr_to_ell(rowPtr, colIdx, values, width) -> tuple[np.ndarray, np.ndarray]:
    rp, ci, v = _i32(rowPtr), _i32(colIdx), _f64(values)
    R = rp.size - 1
    ec = np.empty(R * width, np.int32)
    ev = np.empty(R * width, np.float64)
    lib().oracle_csr_to_ell(C.c_int(R), _p(rp), _p(ci), _p(v), C.c_int(width), _p(ec), _p(ev))
    return ec, ev


def spmv_ell_f64(ecol, evals, width, x) -> np.ndarray:
    ec, ev, xx = _i32(ecol), _f64(evals), _f64(x)
    R = ec.size // width if width else 0
    y = np.empty(R, np.float64)
    lib().oracle_spmv_ell_f64(C.c_int(R), C.c_int(width), _p(ec), _p(ev), _p(xx), _p(y))
    return y


# ---------------------------------------------------------------- permutations
def reorder_columns_com(rowPtr, colIdx, numCols) -> tuple[np.ndarray, np.ndarray]:
    rp, ci = _i32(rowPtr), _i32(colIdx)
    fwd = np.empty(numCols, np.int32)
    nc = np.empty(ci.size, np.int32)
    lib().oracle_reorder_columns_com(C.c_int(rp.size - 1), C.c_int(numCols), _p(rp), _p(ci), _p(fwd), _p(nc))
    return fwd, nc


def reorder_x(x, forwardMap) -> np.ndarray:
    xx, f = _f32(x), _i32(forwardMap)
    out = np.empty(f.size, np.float32)
    lib().oracle_reorder_x(C.c_int(f.size), _p(xx), _p(f), _p(out))
    return out


def sort_columns_within_rows(rowPtr, colIdx, values) -> tuple[np.ndarray, np.ndarray]:
    rp = _i32(rowPtr)
    ci = _i32(colIdx).copy()
    v = _f32(values).copy()
    lib().oracle_sort_columns_within_rows(C.c_int(rp.size - 1), _p(rp), _p(ci), _p(v))
    return ci, v


def reorder_columns_minrow(rowPtr, colIdx, numCols) -> tuple[np.ndarray, np.ndarray]:
    rp, ci = _i32(rowPtr), _i32(colIdx)
    perm = np.empty(numCols, np.int32)
    nc = np.empty(ci.size, np.int32)
    lib().oracle_reorder_columns_minrow(C.c_int(rp.size - 1), C.c_int(numCols), _p(rp), _p(ci), _p(perm), _p(nc))
    return perm, nc


def reorder_columns_freq(rowPtr, colIdx, numCols) -> tuple[np.ndarray, np.ndarray]:
    rp, ci = _i32(rowPtr), _i32(colIdx)
    perm = np.empty(numCols, np.int32)
    nc = np.empty(ci.size, np.int32)
    lib().oracle_reorder_columns_freq(C.c_int(rp.size - 1), C.c_int(numCols), _p(rp), _p(ci), _p(perm), _p(nc))
    return perm, nc


def reorder_rows_by_length(rowPtr, colIdx, values):
    rp, ci, v = _i32(rowPtr), _i32(colIdx), _f32(values)
    R = rp.size - 1
    rm, irm = np.empty(R, np.int32), np.empty(R, np.int32)
    nrp, nci, nv = np.empty(R + 1, np.int32), np.empty(ci.size, np.int32), np.empty(ci.size, np.float32)
    lib().oracle_reorder_rows_by_length(C.c_int(R), _p(rp), _p(ci), _p(v), _p(rm), _p(irm), _p(nrp), _p(nci), _p(nv))
    return rm, irm, nrp, nci, nv


def restore_order(reordered_y, rowMap) -> np.ndarray:
    yy, rm = _f32(reordered_y), _i32(rowMap)
    out = np.empty(rm.size, np.float32)
    lib().oracle_restore_order(C.c_int(rm.size), _p(yy), _p(rm), _p(out))
    return out
