"""Python host mirror of the reference's SpMV interface over the C-ABI.

Names, argument meaning and error behaviour follow the reference's
`namespace muxi_spmv` (src/api/spmv_api.h, src/spmv/csr/spmv_csr.cuh,
src/spmv/csr5/spmv_csr5.cuh, src/spmv/ellpack/spmv_ellpack.cuh): a non-success
`spmv_status_t` raises `SpmvError` carrying the status. Device memory is
held in torch tensors (plumbing only); every computation runs in
libmuxi_spmv_amd.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import SpmvError, check, load

_TORCH2SPMV = {torch.float32: _lib.SPMV_TYPE_FLOAT, torch.float64: _lib.SPMV_TYPE_DOUBLE,
               torch.bfloat16: _lib.SPMV_TYPE_BF16}


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr() if t.numel() > 0 else None


def _stream(stream=None) -> int | None:
    if stream is None:
        return torch.cuda.current_stream().cuda_stream or None
    if isinstance(stream, torch.cuda.Stream):
        return stream.cuda_stream or None
    return stream


def default_opts(sync: int = 0, operation: int = _lib.SPMV_OP_NORMAL, stream=None) -> _lib.spmv_opts_t:
    o = load().muxi_spmv_default_opts()
    o.sync = sync
    o.operation = operation
    o.stream = _stream(stream)
    return o


class CSRMatrix:
    """CSRMatrix<T> (src/formats/sparse_formats.h:24-103): numRows, numCols,
    nnz and the device arrays d_rowPtr (int32[R+1]), d_colIdx (int32[N]),
    d_values (T[N]) as torch tensors on the current GPU."""

    def __init__(self, rowPtr, colIdx, values, numCols: int, device: str | torch.device = "cuda"):
        rp = torch.as_tensor(np.asarray(rowPtr, dtype=np.int32) if not torch.is_tensor(rowPtr) else rowPtr)
        ci = torch.as_tensor(np.asarray(colIdx, dtype=np.int32) if not torch.is_tensor(colIdx) else colIdx)
        va = values if torch.is_tensor(values) else torch.as_tensor(np.asarray(values))
        if rp.dtype != torch.int32 or ci.dtype != torch.int32:
            raise SpmvError(-2, "CSRMatrix", "rowPtr and colIdx must be int32")
        if va.dtype not in _TORCH2SPMV:
            raise SpmvError(-8, "CSRMatrix", f"unsupported value dtype {va.dtype}")
        self.numRows = int(rp.numel()) - 1
        self.numCols = int(numCols)
        self.nnz = int(ci.numel())
        if self.numRows < 0 or va.numel() != self.nnz:
            raise SpmvError(-2, "CSRMatrix", "inconsistent array sizes")
        self.d_rowPtr = rp.to(device).contiguous()
        self.d_colIdx = ci.to(device).contiguous()
        self.d_values = va.to(device).contiguous()
        self.dtype = self.d_values.dtype
        self.c = _lib.muxi_csr_t()
        self._sync_struct()

    def _sync_struct(self):
        c = self.c
        c.numRows, c.numCols, c.nnz = self.numRows, self.numCols, self.nnz
        c.d_rowPtr, c.d_colIdx, c.d_values = _ptr(self.d_rowPtr), _ptr(self.d_colIdx), _ptr(self.d_values)
        c.ownsHostMemory = False
        c.ownsDeviceMemory = False

    def with_values(self, values: torch.Tensor) -> "CSRMatrix":
        """Same pattern, other value array (e.g. a BF16 copy)."""
        m = CSRMatrix.__new__(CSRMatrix)
        m.numRows, m.numCols, m.nnz = self.numRows, self.numCols, self.nnz
        m.d_rowPtr, m.d_colIdx = self.d_rowPtr, self.d_colIdx
        m.d_values = values.contiguous()
        m.dtype = m.d_values.dtype
        m.c = _lib.muxi_csr_t()
        m._sync_struct()
        return m

    @property
    def vtype(self) -> int:
        return _TORCH2SPMV[self.dtype]


def _fn(base: str, dtype: torch.dtype):
    suffix = {torch.float32: "f32", torch.float64: "f64"}.get(dtype)
    if suffix is None:
        raise SpmvError(-8, base, f"no {base} variant for {dtype}")
    return getattr(load(), f"{base}_{suffix}")


def _check_vec(t, dtype, n: int, name: str, what: str = "vector") -> None:
    """Raises SpmvError(SPMV_ERROR_INVALID_VECTOR) unless t is a contiguous tensor of `dtype` with at least n
    elements: the kernels read and write through its raw address, so a short, strided or mistyped tensor would
    otherwise be read or written out of bounds."""
    if not torch.is_tensor(t):
        raise SpmvError(-3, name, f"{what} must be a torch tensor")
    if t.dtype != dtype:
        raise SpmvError(-3, name, f"{what} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise SpmvError(-3, name, f"{what} must be contiguous")
    if t.numel() < n:
        raise SpmvError(-3, name, f"{what} needs >= {n} elements (got {t.numel()})")


def _check_xy(x, y, dtype, n_x: int, n_y: int, name: str) -> None:
    """x and y as _check_vec requires, both on the same GPU (a host tensor's address is not a device address)."""
    _check_vec(x, dtype, n_x, name, "x")
    _check_vec(y, dtype, n_y, name, "y")
    for t, what in ((x, "x"), (y, "y")):
        if not t.is_cuda:
            raise SpmvError(-3, name, f"{what} must be on the GPU (got {t.device})")
    if x.device != y.device:
        raise SpmvError(-3, name, f"x and y on different devices ({x.device}, {y.device})")


def _dims(matrix) -> tuple[int, int]:
    """(numRows, numCols) of a CSRMatrix, CSR5Matrix or ELLPACKMatrix."""
    if isinstance(matrix, CSR5Matrix):
        return matrix.csr.numRows, matrix.csr.numCols
    if isinstance(matrix, CSRMatrix):
        return matrix.numRows, matrix.numCols
    c = matrix.c
    return int(c.numRows), int(c.numCols)


# ---------------------------------------------------------------- handle / API
class Handle:
    """spmv_handle_t (src/api/spmv_api.h:18-43)."""

    def __init__(self):
        self._h = C.POINTER(_lib.muxi_spmv_handle_t)()
        check(load().muxi_spmv_create_handle(C.byref(self._h)), "spmv_create_handle")

    @property
    def fields(self) -> _lib.muxi_spmv_handle_t:
        return self._h.contents

    def set_stream(self, stream) -> None:
        check(load().muxi_spmv_set_stream(self._h, _stream(stream)), "spmv_set_stream")

    def exec(self, fmt: int, matrix, x, y, alpha=1.0, beta=0.0, transpose=False) -> None:
        name = "muxi_spmv_exec_transpose" if transpose else "muxi_spmv_exec"
        rows, cols = _dims(matrix)
        if transpose:
            rows, cols = cols, rows
        dtype = matrix.csr.dtype if isinstance(matrix, CSR5Matrix) else matrix.dtype
        _check_xy(x, y, dtype, cols, rows, name)
        f = _fn(name, x.dtype)
        check(f(self._h, fmt, C.cast(C.pointer(matrix.c), C.c_void_p), _ptr(x), _ptr(y), alpha, beta), name)

    def destroy(self) -> None:
        if self._h:
            check(load().muxi_spmv_destroy_handle(self._h), "spmv_destroy_handle")
            self._h = C.POINTER(_lib.muxi_spmv_handle_t)()

    def __del__(self):
        try:
            if self._h:
                load().muxi_spmv_destroy_handle(self._h)
        except Exception:
            pass


def spmv_csr(A: CSRMatrix, x: torch.Tensor, y: torch.Tensor, alpha=1.0, beta=0.0, opts=None) -> None:
    """spmv_csr<T> (src/spmv/csr/spmv_csr.cu:52-249): y = alpha*A*x + beta*y."""
    if opts is not None and opts.operation == _lib.SPMV_OP_TRANSPOSE:
        _check_xy(x, y, A.dtype, A.numRows, A.numCols, "spmv_csr")
    else:
        _check_xy(x, y, A.dtype, A.numCols, A.numRows, "spmv_csr")
    o = opts if opts is not None else default_opts()
    f = _fn("muxi_spmv_csr", A.dtype)
    check(f(C.byref(A.c), _ptr(x), _ptr(y), alpha, beta, C.byref(o)), "spmv_csr")


def spmv_csr_exec(A, x, y, alpha=1.0, beta=0.0, opts=None) -> None:
    """spmv_csr_exec<T> (src/api/spmv_api.h:82-92)."""
    spmv_csr(A, x, y, alpha, beta, opts)


def spmv_merge_based(A: CSRMatrix, x, y, stream=None) -> None:
    """spmv_merge_based<T> (src/spmv/csr5/spmv_csr5.cu:930-1014): y = A*x."""
    _check_xy(x, y, A.dtype, A.numCols, A.numRows, "spmv_merge_based")
    check(_fn("muxi_spmv_merge_based", A.dtype)(C.byref(A.c), _ptr(x), _ptr(y), _stream(stream)),
          "spmv_merge_based")


def spmv_merge_based_v2(A: CSRMatrix, x, y, stream=None, useOptimized: bool = False) -> None:
    """spmv_merge_based_v2<T> (src/spmv/csr5/spmv_csr5.cu:1019-1072)."""
    _check_xy(x, y, A.dtype, A.numCols, A.numRows, "spmv_merge_based_v2")
    check(_fn("muxi_spmv_merge_based_v2", A.dtype)(C.byref(A.c), _ptr(x), _ptr(y), _stream(stream),
                                                    bool(useOptimized)), "spmv_merge_based_v2")


def spmv_sparse_optimized(A: CSRMatrix, x, y, stream=None) -> None:
    """spmv_sparse_optimized<T> (src/spmv/csr/spmv_sparse_optimized.cu:80-121)."""
    _check_xy(x, y, A.dtype, A.numCols, A.numRows, "spmv_sparse_optimized")
    check(_fn("muxi_spmv_sparse_optimized", A.dtype)(C.byref(A.c), _ptr(x), _ptr(y), _stream(stream)),
          "spmv_sparse_optimized")


# ---------------------------------------------------------------- merge plan pieces
def merge_partition_count(numRows: int, nnz: int, elementsPerPartition: int = 16, maxPartitionMultiplier: int = 8,
                          numSMs: int = 104) -> int:
    """P = max(1, min((R+N)/ePP, nSM*(256/64)*mult)) (spmv_csr5.cu:942-972)."""
    return int(load().muxi_merge_partition_count(numRows, nnz, elementsPerPartition, maxPartitionMultiplier,
                                                 numSMs))


def merge_partitions(numRows: int, nnz: int, numPartitions: int, stream=None) -> torch.Tensor:
    """compute_merge_partitions_kernel (spmv_csr5.cu:912-925): int32[P+1] diagonals."""
    pos = torch.empty(numPartitions + 1, dtype=torch.int32, device="cuda")
    check(load().muxi_merge_partitions(numRows, nnz, numPartitions, _ptr(pos), _stream(stream)),
          "merge_partitions")
    return pos


def merge_path_coords(A: CSRMatrix, pos: torch.Tensor, stream=None) -> tuple[torch.Tensor, torch.Tensor]:
    """merge_path_search (spmv_csr5.cu:497-533) at every diagonal in pos."""
    n = pos.numel()
    row = torch.empty(n, dtype=torch.int32, device="cuda")
    nz = torch.empty(n, dtype=torch.int32, device="cuda")
    check(load().muxi_merge_path_coords(_ptr(A.d_rowPtr), A.numRows, A.nnz, _ptr(pos), n, _ptr(row), _ptr(nz),
                                        _stream(stream)), "merge_path_coords")
    return row, nz


# ---------------------------------------------------------------- CSR5
class CSR5Matrix:
    """CSR5Matrix<T> (src/formats/sparse_formats.h:284-386) after csr5_preprocess."""

    def __init__(self, csr: CSRMatrix):
        self.csr = csr  # keeps the aliased CSR arrays alive
        self.c = _lib.muxi_csr5_t()

    @property
    def sigma(self) -> int:
        return self.c.sigma

    @property
    def numTiles(self) -> int:
        return self.c.numTiles

    @property
    def conversionTimeMs(self) -> float:
        return self.c.conversionTimeMs

    def tile_row_ptr(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.c.tile_row_ptr, (self.c.numTiles,)).copy() if self.c.numTiles else \
            np.zeros(0, np.int32)

    def tile_nnz_offset(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.c.tile_nnz_offset, (self.c.numTiles,)).copy() if self.c.numTiles else \
            np.zeros(0, np.int32)

    def __del__(self):
        try:
            load().muxi_csr5_destroy(C.byref(self.c))
        except Exception:
            pass


def csr5_preprocess(csr: CSRMatrix, sigma: int = 0, stream=None) -> CSR5Matrix:
    """csr5_preprocess<T> (src/spmv/csr5/spmv_csr5.cu:223-275)."""
    m = CSR5Matrix(csr)
    check(load().muxi_csr5_preprocess(C.byref(csr.c), C.byref(m.c), sigma, _stream(stream)), "csr5_preprocess")
    return m


def spmv_csr5(A: CSR5Matrix, x, y, alpha=1.0, beta=0.0, opts=None) -> None:
    """spmv_csr5<T> (src/spmv/csr5/spmv_csr5.cu:277-362)."""
    _check_xy(x, y, A.csr.dtype, A.csr.numCols, A.csr.numRows, "spmv_csr5")
    o = opts if opts is not None else default_opts()
    check(_fn("muxi_spmv_csr5", A.csr.dtype)(C.byref(A.c), _ptr(x), _ptr(y), alpha, beta, C.byref(o)), "spmv_csr5")


# ---------------------------------------------------------------- ELLPACK
class ELLPACKMatrix:
    """ELLPACKMatrix<T> (src/spmv/ellpack/spmv_ellpack.cuh:26-39)."""

    def __init__(self, dtype):
        self.c = _lib.muxi_ell_t()
        self.dtype = dtype

    @property
    def numElements(self) -> int:
        return self.c.numElements

    def arrays(self) -> tuple[torch.Tensor, torch.Tensor]:
        """Copies of the column-major slots (colIdx, values) for inspection."""
        n = self.c.numRows * self.c.numElements
        col = torch.empty(n, dtype=torch.int32, device="cuda")
        val = torch.empty(n, dtype=self.dtype, device="cuda")
        if n:
            st = _stream(None)
            check(load().muxi_device_copy(_ptr(col), self.c.d_colIdx, 4 * n, st), "device_copy")
            check(load().muxi_device_copy(_ptr(val), self.c.d_values, val.element_size() * n, st), "device_copy")
        return col, val

    def __del__(self):
        try:
            load().muxi_ellpack_destroy(C.byref(self.c))
        except Exception:
            pass


def csr_to_ellpack(csr: CSRMatrix, maxElements: int = 0) -> ELLPACKMatrix:
    """csr_to_ellpack<T> (src/spmv/ellpack/spmv_ellpack.cu:11-70)."""
    e = ELLPACKMatrix(csr.dtype)
    check(_fn("muxi_csr_to_ellpack", csr.dtype)(C.byref(csr.c), C.byref(e.c), maxElements), "csr_to_ellpack")
    return e


def spmv_ellpack(E: ELLPACKMatrix, x, y, stream=None) -> None:
    """spmv_ellpack<T> (src/spmv/ellpack/spmv_ellpack.cu:72-91)."""
    _check_xy(x, y, E.dtype, int(E.c.numCols), int(E.c.numRows), "spmv_ellpack")
    check(_fn("muxi_spmv_ellpack", E.dtype)(C.byref(E.c), _ptr(x), _ptr(y), _stream(stream)), "spmv_ellpack")


# ---------------------------------------------------------------- plans
@dataclass
class PlanInfo:
    kernel: str
    param0: int
    scratch_bytes: int


class Plan:
    """A matrix bound to one kernel and its scratch; exec allocates nothing."""

    def __init__(self, A: CSRMatrix, kernel: str = "auto", param0: int = 0, xtype=torch.float32,
                 compensated: bool = False, stream=None, row_base: int = 0, rows_full: int | None = None):
        """With row_base / rows_full, A is rows [row_base, row_base + A.numRows) of a matrix of rows_full rows
        (a row shard): muxi_plan_create_rows."""
        self.A = A
        self.xtype = xtype
        self.nrhs = 1
        self._p = C.c_void_p()
        if row_base == 0 and rows_full is None:
            check(load().muxi_plan_create(C.byref(A.c), A.vtype, _TORCH2SPMV[xtype], _lib.KERNELS[kernel], param0,
                                          int(compensated), _stream(stream), C.byref(self._p)), "plan_create")
        else:
            full = A.numRows + row_base if rows_full is None else rows_full
            check(load().muxi_plan_create_rows(C.byref(A.c), A.vtype, _TORCH2SPMV[xtype], _lib.KERNELS[kernel],
                                               param0, int(compensated), _stream(stream), int(row_base), int(full),
                                               C.byref(self._p)), "plan_create_rows")

    @classmethod
    def spmm(cls, A: CSRMatrix, nrhs: int, ipt: int = 0, xtype=torch.float32, stream=None) -> "Plan":
        """SpMM plan: Y = alpha*A*X + beta*Y, X (numCols x nrhs) and Y (numRows x nrhs) row-major; ipt 0 takes
        the items per lane group from the tuned schedule table."""
        self = cls.__new__(cls)
        self.A = A
        self.xtype = xtype
        self.nrhs = nrhs
        self._p = C.c_void_p()
        check(load().muxi_plan_create_spmm(C.byref(A.c), A.vtype, _TORCH2SPMV[xtype], nrhs, ipt, _stream(stream),
                                           C.byref(self._p)), "plan_create_spmm")
        return self

    @classmethod
    def checked(cls, A: CSRMatrix, rtol: float, compensated: bool = False, stream=None) -> "Plan":
        """Checked plan (muxi_plan_create_checked): FP32 x and y on the LDS column slabs; every row whose
        a-posteriori FP32 (or Dot2) error bound exceeds rtol * |y_r| is recomputed in FP64 in the same exec."""
        self = cls.__new__(cls)
        self.A = A
        self.xtype = torch.float32
        self.nrhs = 1
        self._p = C.c_void_p()
        check(load().muxi_plan_create_checked(C.byref(A.c), A.vtype, 1 if compensated else 0, float(rtol),
                                              _stream(stream), C.byref(self._p)), "plan_create_checked")
        return self

    def checked_rows(self, stream=None) -> np.ndarray:
        """Ids of the rows the last exec recomputed in FP64 (sorted); waits for the stream."""
        n = C.c_int()
        cap = max(int(self.A.numRows), 1)
        buf = (C.c_int * cap)()
        check(load().muxi_plan_checked_rows(self._p, C.byref(n), buf, cap, _stream(stream)), "plan_checked_rows")
        k = min(n.value, cap)
        return np.sort(np.frombuffer(buf, dtype=np.int32, count=k).copy())

    def info(self) -> PlanInfo:
        k, p0, sb = C.c_int(), C.c_int(), C.c_size_t()
        check(load().muxi_plan_info(self._p, C.byref(k), C.byref(p0), C.byref(sb)), "plan_info")
        return PlanInfo(_lib.KERNEL_NAMES[k.value], p0.value, sb.value)

    def exec(self, x: torch.Tensor, y: torch.Tensor, alpha=1.0, beta=0.0, nrhs: int | None = None,
             stream=None) -> None:
        n = self.nrhs if nrhs is None else nrhs
        _check_xy(x, y, self.xtype, self.A.numCols * n, self.A.numRows * n, "plan_exec")
        check(load().muxi_plan_exec(self._p, _ptr(x), _ptr(y), float(alpha), float(beta), n, _stream(stream)),
              "plan_exec")

    def destroy(self):
        if self._p:
            load().muxi_plan_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def oracle_f64(A: CSRMatrix, x: torch.Tensor, stream=None) -> torch.Tensor:
    """FP64 device oracle: y64 = A*x with compensated FP64 accumulation."""
    y = torch.empty(A.numRows, dtype=torch.float64, device="cuda")
    check(load().muxi_spmv_oracle_f64(C.byref(A.c), A.vtype, _TORCH2SPMV[x.dtype], _ptr(x), _ptr(y),
                                      _stream(stream)), "spmv_oracle_f64")
    return y


def convert_values(src: torch.Tensor, dtype: torch.dtype, stream=None) -> torch.Tensor:
    """Device value conversion (FP32 -> BF16 is round-to-nearest-even)."""
    dst = torch.empty(src.numel(), dtype=dtype, device=src.device)
    check(load().muxi_convert_values(_ptr(src), _TORCH2SPMV[src.dtype], _ptr(dst), _TORCH2SPMV[dtype],
                                     src.numel(), _stream(stream)), "convert_values")
    return dst
