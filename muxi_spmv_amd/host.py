"""SpMV with x and y in host memory (include/muxi_spmv/muxi_spmv_host.h): pinned buffers, row chunks with
their own device plans, PCIe copies overlapped with the kernels. The reference's end-to-end experiments are
tests/test_pinned_memory.cu:191-330 and tests/test_multi_stream.cu:121-290."""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, load
from .api import _TORCH2SPMV, CSRMatrix

_NP2SPMV = {np.dtype(np.float32): _lib.SPMV_TYPE_FLOAT, np.dtype(np.float64): _lib.SPMV_TYPE_DOUBLE}


class _Pinned:
    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        check(load().muxi_host_alloc(C.byref(self.ptr), max(int(nbytes), 1)), "host_alloc")
        weakref.finalize(self, load().muxi_host_free, self.ptr.value)


def pinned_empty(n: int, dtype=np.float32) -> np.ndarray:
    """A numpy array in pinned host memory (freed when the array is collected)."""
    dt = np.dtype(dtype)
    owner = _Pinned(n * dt.itemsize)
    buf = (C.c_char * max(n * dt.itemsize, 1)).from_address(owner.ptr.value)
    a = np.frombuffer(buf, dtype=dt, count=n)
    buf._owner = owner  # the array keeps buf alive, buf keeps the pinned allocation
    return a


@dataclass
class HostTiming:
    total_ms: float
    first_copy_in_ms: float
    kernels_ms: float
    copy_out_tail_ms: float


class HostPlan:
    """A matrix (device arrays) run on host x / y: `chunks` row ranges, each with its own device plan."""

    def __init__(self, A: CSRMatrix, chunks: int = 4, kernel: str = "auto", param0: int = 0,
                 xtype=torch.float32):
        self.A = A
        self.xtype = np.dtype(np.float64 if xtype == torch.float64 else np.float32)
        self._p = C.c_void_p()
        check(load().muxi_host_plan_create(C.byref(A.c), A.vtype, _TORCH2SPMV[xtype], int(chunks),
                                           _lib.KERNELS[kernel], int(param0), C.byref(self._p)), "host_plan_create")
        self.chunks = int(chunks)

    def exec(self, x: np.ndarray, y: np.ndarray, alpha: float = 1.0, beta: float = 0.0) -> None:
        if x.dtype != self.xtype or y.dtype != self.xtype or not x.flags.c_contiguous or not y.flags.c_contiguous:
            raise TypeError(f"x and y must be contiguous {self.xtype}")
        if x.size != self.A.numCols or y.size != self.A.numRows:
            raise ValueError("x or y has the wrong length")
        check(load().muxi_host_plan_exec(self._p, x.ctypes.data, y.ctypes.data, float(alpha), float(beta)),
              "host_plan_exec")

    def timing(self) -> HostTiming:
        t = [C.c_float() for _ in range(4)]
        check(load().muxi_host_plan_timing(self._p, *[C.byref(v) for v in t]), "host_plan_timing")
        return HostTiming(*[v.value for v in t])

    def chunk(self, k: int) -> tuple[int, int, str, int]:
        r0, r1, kk, pp = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(load().muxi_host_plan_chunk(self._p, k, C.byref(r0), C.byref(r1), C.byref(kk), C.byref(pp)),
              "host_plan_chunk")
        return r0.value, r1.value, _lib.KERNEL_NAMES.get(kk.value, str(kk.value)), pp.value

    def destroy(self) -> None:
        if self._p:
            load().muxi_host_plan_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
