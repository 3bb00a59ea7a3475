"""Autotune sweep, matrix features and the learned kernel selector.

* `gen_matrix` builds seeded matrices on the device (muxi_gen_matrix).
* `features` returns the eight matrix features (muxi_matrix_features).
* `sweep` times every candidate (kernel, parameter) on one matrix, cold:
  a cache-flush kernel writes 512 MiB between repetitions, and the kernel
  time is the HIP-event time around the plan's main kernel plus its fix-up.
* `Selector` is the MLP (HIP forward/backward/Adam kernels) that predicts
  each candidate's log time from the features; its pick is the argmin.
  Trained on sweep records; the sweep stays the fallback and the measure of the pick's regret.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import statistics
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .api import Plan, _TORCH2SPMV
from ._lib import check, load

CANDIDATES: list[tuple[str, int]] = [
    ("scalar", 0), ("vector", 0),
    ("vwave", 2), ("vwave", 4), ("vwave", 8), ("vwave", 16), ("vwave", 32),
    ("merge", 4), ("merge", 8), ("merge", 16),
    ("csr5", 256), ("csr5", 512), ("csr5", 1024),
    ("adaptive", 512), ("adaptive", 1024), ("adaptive", 2048),
    ("xslab", (2 << 8) | 4), ("xslab", (4 << 8) | 4), ("xslab", (8 << 8) | 4),
    ("xslab", (2 << 8) | 8), ("xslab", (4 << 8) | 8), ("xslab", (8 << 8) | 8),
    ("ell", 0),
    ("lslab", 4),
]
CAND_NAMES = [f"{k}:{p}" for k, p in CANDIDATES]
FEATURE_NAMES = ["rows", "nnz", "mean_len", "var_len", "max_len", "bandwidth", "diag_block_fill", "empty_frac"]

GEN_FAMILIES = {"uniform": 0, "powerlaw": 1, "banded": 2, "blockdiag": 3, "poisson": 4}


class DeviceCSR:
    """A CSR matrix whose device arrays the library allocated (muxi_gen_matrix)."""

    def __init__(self, c: _lib.muxi_csr_t, dtype: torch.dtype):
        self.c = c
        self.dtype = dtype
        self.numRows, self.numCols, self.nnz = c.numRows, c.numCols, c.nnz

    @property
    def vtype(self) -> int:
        return _TORCH2SPMV[self.dtype]

    def host(self):
        """Copies to host numpy arrays (rowPtr, colIdx, values)."""
        R, N = self.numRows, self.nnz
        rp = torch.empty(R + 1, dtype=torch.int32, device="cuda")
        ci = torch.empty(max(N, 1), dtype=torch.int32, device="cuda")
        va = torch.empty(max(N, 1), dtype=self.dtype, device="cuda")
        lib = load()
        s = torch.cuda.current_stream().cuda_stream or None
        check(lib.muxi_device_copy(rp.data_ptr(), self.c.d_rowPtr, 4 * (R + 1), s), "copy")
        if N:
            check(lib.muxi_device_copy(ci.data_ptr(), self.c.d_colIdx, 4 * N, s), "copy")
            check(lib.muxi_device_copy(va.data_ptr(), self.c.d_values, va.element_size() * N, s), "copy")
        return rp.cpu().numpy(), ci[:N].cpu().numpy(), va[:N].cpu().numpy()

    def free(self):
        if self.c.d_rowPtr:
            load().muxi_gen_free(C.byref(self.c))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gen_matrix(family: str, R: int, C_: int, p0: float = 0, p1: float = 0, p2: float = 0, normal: bool = False,
               dtype=torch.float32, seed: int = 0, permute: bool = False) -> DeviceCSR:
    c = _lib.muxi_csr_t()
    s = torch.cuda.current_stream().cuda_stream or None
    check(load().muxi_gen_matrix(GEN_FAMILIES[family], R, C_, float(p0), float(p1), float(p2), int(normal),
                                 _TORCH2SPMV[dtype], seed, s, C.byref(c)), "gen_matrix")
    if permute:
        check(load().muxi_gen_permute(C.byref(c), _TORCH2SPMV[dtype], seed ^ 0x5A5A, s), "gen_permute")
    return DeviceCSR(c, dtype)


def gen_vector(n: int, dtype=torch.float32, seed: int = 0) -> torch.Tensor:
    x = torch.empty(n, dtype=dtype, device="cuda")
    check(load().muxi_gen_vector(x.data_ptr(), _TORCH2SPMV[dtype], n, seed,
                                 torch.cuda.current_stream().cuda_stream or None), "gen_vector")
    return x


def features(A) -> np.ndarray:
    out = (C.c_double * 8)()
    check(load().muxi_matrix_features(C.byref(A.c), torch.cuda.current_stream().cuda_stream or None, out),
          "matrix_features")
    return np.array(out[:], dtype=np.float64)


def feature_vector(f: np.ndarray, fp64: bool = False) -> np.ndarray:
    """Model inputs: logs of the size-like features, the fractions as they are, and 1 for FP64 x and y."""
    rows, nnz, mean, var, mx, band, diag, empty = f
    return np.array([math.log1p(rows), math.log1p(nnz), math.log1p(mean), math.log1p(math.sqrt(max(var, 0.0))),
                     math.log1p(mx), band, diag, empty, math.log1p(mx / max(mean, 1e-9)),
                     math.log1p(nnz) - math.log1p(rows), 1.0 if fp64 else 0.0], dtype=np.float32)


N_INPUT = 11


def ell_feasible(f: np.ndarray, limit_slots: float = 4.0) -> bool:
    """ELL only when its padding stays within limit_slots x nnz."""
    rows, nnz, mx = f[0], f[1], f[4]
    return nnz > 0 and rows * mx <= limit_slots * nnz and rows * mx < 2 ** 31


class Flusher:
    def __init__(self, nbytes: int = 512 << 20):
        self.buf = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")

    def __call__(self):
        check(load().muxi_flush_caches(self.buf.data_ptr(), self.buf.numel() * 4,
                                       torch.cuda.current_stream().cuda_stream or None), "flush")


def time_plan_cold(plan: Plan, x, y, reps: int, flusher: Flusher) -> float:
    """Median over reps of the event time of one cold exec (ms)."""
    lib = load()
    s = torch.cuda.current_stream().cuda_stream or None
    ev = [C.c_void_p(), C.c_void_p()]
    for e in ev:
        check(lib.muxi_event_create(C.byref(e)), "event_create")
    ts = []
    try:
        for _ in range(reps):
            flusher()
            check(lib.muxi_event_record(ev[0], s), "event_record")
            check(lib.muxi_plan_exec(plan._p, x.data_ptr(), y.data_ptr(), 1.0, 0.0, 1, s), "plan_exec")
            check(lib.muxi_event_record(ev[1], s), "event_record")
            ms = C.c_float()
            check(lib.muxi_event_elapsed_ms(ev[0], ev[1], C.byref(ms)), "elapsed")
            ts.append(ms.value)
    finally:
        for e in ev:
            lib.muxi_event_destroy(e)
    return statistics.median(ts)


LSLAB_MIN_COLS = 1 << 16  # below this x sits in every L2 and the two launches of lslab cannot pay off
LSLAB_MAX_COLS = 256 * 32768  # the selector's lslab domain, as selector.cpp (plans themselves take 512 slabs)


def feasible(kernel: str, f: np.ndarray, fp64: bool = False) -> bool:
    """Whether the sweep times this kernel on a (square) matrix with features f and FP64 x if fp64. The library's
    selector applies the same rules (selector.cpp, candidate_allowed)."""
    if kernel == "ell":
        return ell_feasible(f)
    if kernel in ("scalar", "vector"):
        return f[4] <= 20000  # one lane / one wavefront per row on 10^4+ rows: hopeless and slow to time
    if kernel == "lslab":  # FP32 x in LDS
        return not fp64 and LSLAB_MIN_COLS <= f[0] <= LSLAB_MAX_COLS
    return True


def sweep(A, x: torch.Tensor, y: torch.Tensor, flusher: Flusher, reps: int = 5,
          candidates=CANDIDATES, f: np.ndarray | None = None) -> dict[str, float]:
    """Cold time (ms) of every feasible candidate; infeasible ones are absent."""
    f = features(A) if f is None else f
    fp64 = x.dtype == torch.float64
    out = {}
    for k, p in candidates:
        if not feasible(k, f, fp64):
            continue
        plan = Plan(A, k, p, xtype=x.dtype)
        time_plan_cold(plan, x, y, 1, flusher)  # warm the code path
        out[f"{k}:{p}"] = time_plan_cold(plan, x, y, reps, flusher)
        plan.destroy()
    return out


# ---------------------------------------------------------------- the MLP
class Selector:
    """in -> hidden -> hidden -> len(CANDIDATES) MLP on the device."""

    def __init__(self, hidden: int = 64, n_in: int = N_INPUT, n_out: int = len(CANDIDATES), seed: int = 0):
        self.n_in, self.h, self.n_out = n_in, hidden, n_out
        n = int(load().muxi_mlp_param_count(n_in, hidden, n_out))
        g = torch.Generator().manual_seed(seed)
        parts = []
        for fan_in, fan_out in [(n_in, hidden), (hidden, hidden), (hidden, n_out)]:
            bound = math.sqrt(6.0 / fan_in)  # He-uniform for ReLU layers
            parts.append((torch.rand(fan_out * fan_in, generator=g) * 2 - 1) * bound)
            parts.append(torch.zeros(fan_out))
        self.params = torch.cat(parts).cuda()
        assert self.params.numel() == n
        self.mu = torch.zeros(n_in)
        self.sd = torch.ones(n_in)
        self.tmu = torch.zeros(n_out)

    def _norm(self, X: np.ndarray) -> torch.Tensor:
        X = np.asarray(X, np.float32)[:, :self.n_in]  # a 10-input model ignores the FP64 flag
        return ((torch.from_numpy(np.ascontiguousarray(X)) - self.mu) / self.sd).cuda().contiguous()

    def forward(self, Xn: torch.Tensor) -> torch.Tensor:
        B = Xn.shape[0]
        Y = torch.empty((B, self.n_out), device="cuda")
        w = torch.empty(int(load().muxi_mlp_workspace_bytes(B, self.n_in, self.h, self.n_out)) // 4 + 1,
                        device="cuda")
        check(load().muxi_mlp_forward(self.params.data_ptr(), self.n_in, self.h, self.n_out, Xn.data_ptr(), B,
                                      Y.data_ptr(), w.data_ptr(), torch.cuda.current_stream().cuda_stream or None),
              "mlp_forward")
        return Y

    def backward(self, Xn: torch.Tensor, T: torch.Tensor, M: torch.Tensor) -> tuple[torch.Tensor, float]:
        B = Xn.shape[0]
        G = torch.empty_like(self.params)
        loss = torch.empty(1, device="cuda")
        w = torch.empty(int(load().muxi_mlp_workspace_bytes(B, self.n_in, self.h, self.n_out)) // 4 + 1,
                        device="cuda")
        check(load().muxi_mlp_backward(self.params.data_ptr(), self.n_in, self.h, self.n_out, Xn.data_ptr(), B,
                                       T.data_ptr(), M.data_ptr(), G.data_ptr(), loss.data_ptr(), w.data_ptr(),
                                       torch.cuda.current_stream().cuda_stream or None), "mlp_backward")
        return G, float(loss.item())

    def fit(self, X: np.ndarray, T: np.ndarray, M: np.ndarray, epochs: int = 3000, lr: float = 3e-3,
            batch: int = 256, seed: int = 0) -> list[float]:
        """X [n, in] raw model inputs, T [n, out] log-times (relative to each
        matrix's best), M [n, out] masks. Adam on minibatches."""
        X = np.asarray(X, np.float32)[:, :self.n_in]
        self.mu = torch.from_numpy(X.mean(0))
        self.sd = torch.from_numpy(X.std(0) + 1e-6)
        Xn = self._norm(X)
        Tt = torch.from_numpy(np.asarray(T, np.float32)).cuda()
        Mt = torch.from_numpy(np.asarray(M, np.float32)).cuda()
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        rng = np.random.default_rng(seed)
        n = X.shape[0]
        losses = []
        for t in range(1, epochs + 1):
            idx = torch.from_numpy(rng.choice(n, size=min(batch, n), replace=False)).cuda()
            G, loss = self.backward(Xn[idx].contiguous(), Tt[idx].contiguous(), Mt[idx].contiguous())
            check(load().muxi_mlp_adam(self.params.data_ptr(), G.data_ptr(), m.data_ptr(), v.data_ptr(),
                                       self.params.numel(), lr, 0.9, 0.999, 1e-8, t,
                                       torch.cuda.current_stream().cuda_stream or None), "mlp_adam")
            if t % 100 == 0 or t == 1:
                losses.append(loss)
        return losses

    def predict(self, X: np.ndarray) -> np.ndarray:
        return self.forward(self._norm(np.atleast_2d(X))).cpu().numpy()

    def pick(self, X: np.ndarray, mask: np.ndarray | None = None) -> np.ndarray:
        P = self.predict(X)
        if mask is not None:
            P = np.where(mask > 0, P, np.inf)
        return P.argmin(1)

    @staticmethod
    def _layout(n_in: int, h: int, n_out: int) -> list[tuple[str, tuple]]:
        """Parameter blocks in the library's order: W1 (h x in), b1, W2 (h x h), b2, W3 (out x h), b3."""
        return [("w1", (h, n_in)), ("b1", (h,)), ("w2", (h, h)), ("b2", (h,)), ("w3", (n_out, h)), ("b3", (n_out,))]

    def _blocks(self) -> dict[str, np.ndarray]:
        p = self.params.detach().cpu().numpy().astype(np.float64)
        out, k = {}, 0
        for name, shape in self._layout(self.n_in, self.h, self.n_out):
            n = int(np.prod(shape))
            out[name] = p[k:k + n].reshape(shape)
            k += n
        return out

    @classmethod
    def ensemble(cls, members: list["Selector"]) -> "Selector":
        """One network whose output is the mean of the members' outputs: the members' first layers stacked, their
        second layers block-diagonal, their output layers side by side and scaled by 1/K. The members share the
        input normalisation (trained on the same inputs)."""
        K = len(members)
        m0 = members[0]
        n_in, h, n_out = m0.n_in, m0.h, m0.n_out
        B = [m._blocks() for m in members]
        w1 = np.concatenate([b["w1"] for b in B], 0)
        b1 = np.concatenate([b["b1"] for b in B])
        w2 = np.zeros((K * h, K * h))
        for k, b in enumerate(B):
            w2[k * h:(k + 1) * h, k * h:(k + 1) * h] = b["w2"]
        b2 = np.concatenate([b["b2"] for b in B])
        w3 = np.concatenate([b["w3"] for b in B], 1) / K
        b3 = np.mean([b["b3"] for b in B], 0)
        e = cls(K * h, n_in, n_out)
        e.params = torch.from_numpy(np.concatenate([w1.ravel(), b1, w2.ravel(), b2, w3.ravel(), b3])
                                    .astype(np.float32)).cuda()
        e.mu, e.sd = m0.mu.clone(), m0.sd.clone()
        return e

    def state(self) -> dict:
        return {"n_in": self.n_in, "hidden": self.h, "n_out": self.n_out, "candidates": CAND_NAMES,
                "params": self.params.cpu().tolist(), "mu": self.mu.tolist(), "sd": self.sd.tolist()}

    @classmethod
    def from_state(cls, st: dict) -> "Selector":
        s = cls(st["hidden"], st["n_in"], st["n_out"])
        s.params = torch.tensor(st["params"], dtype=torch.float32).cuda()
        s.mu = torch.tensor(st["mu"], dtype=torch.float32)
        s.sd = torch.tensor(st["sd"], dtype=torch.float32)
        return s

    def save(self, path: str | Path) -> None:
        Path(path).write_text(json.dumps(self.state()))

    def install(self) -> None:
        """Makes this network the library's selector for MUXI_KERNEL_AUTO plans (muxi_selector_set)."""
        ks = (C.c_int * self.n_out)(*[_lib.KERNELS[k] for k, _ in CANDIDATES])
        ps = (C.c_int * self.n_out)(*[p for _, p in CANDIDATES])
        params = self.params.detach().float().cpu().contiguous()
        mu = (C.c_float * self.n_in)(*self.mu.tolist())
        sd = (C.c_float * self.n_in)(*self.sd.tolist())
        check(load().muxi_selector_set(params.data_ptr(), self.n_in, self.h, self.n_out, mu, sd, ks, ps),
              "selector_set")

    @classmethod
    def load(cls, path: str | Path) -> "Selector":
        return cls.from_state(json.loads(Path(path).read_text()))


def library_pick(A, xtype=torch.float32) -> str:
    """The (kernel, param) an AUTO plan takes for A, as a candidate name (muxi_selector_pick)."""
    k, p = C.c_int(), C.c_int()
    check(load().muxi_selector_pick(C.byref(A.c), _TORCH2SPMV[xtype], torch.cuda.current_stream().cuda_stream or None,
                                    C.byref(k), C.byref(p)), "selector_pick")
    return f"{_lib.KERNEL_NAMES[k.value]}:{p.value}"


def feasible_mask(f: np.ndarray, fp64: bool = False) -> np.ndarray:
    return np.array([1.0 if feasible(k, f, fp64) else 0.0 for k, _ in CANDIDATES], np.float32)


def targets_from_sweep(times: dict[str, float]) -> tuple[np.ndarray, np.ndarray]:
    """log(t_c / t_best) per candidate, with a mask of measured candidates."""
    best = min(times.values())
    T = np.zeros(len(CANDIDATES), np.float32)
    M = np.zeros(len(CANDIDATES), np.float32)
    for i, name in enumerate(CAND_NAMES):
        if name in times:
            T[i] = math.log(times[name] / best)
            M[i] = 1.0
    return T, M


def reference_rule(f: np.ndarray) -> str:
    """The reference's kernel choice on a 64-wide wavefront (spmv_csr.cu:84-128):
    integer avgNnz < 32 -> merge-path, else one wavefront per row."""
    avg = int(f[1]) // max(int(f[0]), 1)
    return "merge:8" if avg < 32 else "vector:0"
