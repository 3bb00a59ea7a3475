"""Batched SpMV mirror (include/muxi_spmv/muxi_spmv_batch.h)."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .api import _TORCH2SPMV, _stream
from ._lib import check, load


BATCH_SLABS = 0x10000  # ipt flag: the slab mode (MUXI_BATCH_SLABS)
TUNED_MERGE = -1  # ipt: the merge mode with the tuned schedule


class BatchPlan:
    """Binds matrices (objects with .c and .dtype) and their x / y tensors."""

    def __init__(self, mats, xs, ys, ipt: int = 0, stream=None):
        """ipt: 0 = the slab mode (the block-diagonal matrix through an LDS column-slab plan) when the batch
        qualifies, else the merge mode with the tuned schedule; TUNED_MERGE (< 0) = that tuned merge schedule;
        BATCH_SLABS = the slab mode; other values = merge items per lane (4, 8, 16) | workgroups per CU << 8."""
        n = len(mats)
        self.mats, self.xs, self.ys = mats, xs, ys  # keep alive
        arr = (_lib.muxi_csr_t * n)(*[m.c for m in mats])
        xp = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
        yp = (C.c_void_p * n)(*[y.data_ptr() for y in ys])
        dt = _TORCH2SPMV[mats[0].dtype]
        self._p = C.c_void_p()
        check(load().muxi_batch_plan_create(arr, n, dt, _TORCH2SPMV[xs[0].dtype], xp, yp, ipt, _stream(stream),
                                            C.byref(self._p)), "batch_plan_create")

    def exec(self, alpha=1.0, beta=0.0, stream=None):
        check(load().muxi_batch_exec(self._p, alpha, beta, _stream(stream)), "batch_exec")

    def exec_per_matrix(self, kernel: str = "vector", param0: int = 0, stream=None):
        check(load().muxi_batch_exec_per_matrix(self._p, _lib.KERNELS[kernel], param0, _stream(stream)),
              "batch_exec_per_matrix")

    def info(self) -> dict:
        t, g, s = C.c_int(), C.c_int(), C.c_size_t()
        check(load().muxi_batch_info(self._p, C.byref(t), C.byref(g), C.byref(s)), "batch_info")
        i, w = C.c_int(), C.c_int()
        check(load().muxi_batch_schedule(self._p, C.byref(i), C.byref(w)), "batch_schedule")
        return {"tiles": t.value, "grid": g.value, "scratch_bytes": s.value, "ipt": i.value, "wg_per_cu": w.value,
                "mode": "slabs" if i.value == 0 else "merge"}

    def graph(self, per_matrix_kernel: str | None = None, param0: int = 0, alpha=1.0, beta=0.0, stream=None):
        """A captured hipGraph of one exec (or of the per-matrix loop)."""
        return Graph(self, per_matrix_kernel, param0, alpha, beta, stream)

    def destroy(self):
        if self._p:
            load().muxi_batch_plan_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Graph:
    def __init__(self, plan: BatchPlan, per_matrix_kernel, param0, alpha, beta, stream=None):
        self.plan = plan
        s = stream if stream is not None else torch.cuda.current_stream()
        sp = _stream(s)
        if not sp:
            raise RuntimeError("graph capture needs a non-default stream")
        k = _lib.KERNELS[per_matrix_kernel] if per_matrix_kernel else 0
        self._g = C.c_void_p()
        check(load().muxi_batch_graph_create(plan._p, alpha, beta, k, param0, sp, C.byref(self._g)),
              "batch_graph_create")

    def launch(self, stream=None):
        check(load().muxi_graph_launch(self._g, _stream(stream)), "graph_launch")

    def __del__(self):
        try:
            if self._g:
                load().muxi_graph_destroy(self._g)
        except Exception:
            pass
