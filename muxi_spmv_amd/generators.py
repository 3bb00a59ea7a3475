"""Seeded synthetic CSR generators for the benchmark and test workloads.

Each returns host arrays (rowPtr int32[R+1], colIdx int32[N], values) with
column indices distinct and sorted within every row. Sampling uses numpy's
PCG64 generator seeded by the caller, so a (family, sizes, seed) triple
always yields the same matrix.

Families:
  uniform_rows    row length ~ U{lo..hi}, columns uniform without replacement
  powerlaw_rows   row length from p(k) ~ k^-a on [kmin, kmax] (i.i.d. per row,
                  so the row order is already shuffled)
  laplacian_2d    5-point stencil on an n x n grid, diagonal 4, neighbours -1
  banded          nnz_per_row distinct columns uniform in [i-hb, i+hb] ∩ [0, R)
  block_diagonal  dense bs x bs diagonal blocks
  permuted        any matrix with random row and column permutations
  poisson_rows    row length ~ Poisson(mean), clamped to [0, C]
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class HostCSR:
    rowPtr: np.ndarray
    colIdx: np.ndarray
    values: np.ndarray
    numRows: int
    numCols: int

    @property
    def nnz(self) -> int:
        return int(self.colIdx.size)

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.rowPtr.astype(np.int64))


def _distinct_sorted(lens: np.ndarray, lo: np.ndarray, span: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """For each row r, lens[r] distinct integers from [lo[r], lo[r]+span[r]),
    returned row-major and ascending within each row (lens[r] <= span[r])."""
    lens = lens.astype(np.int64)
    R = lens.size
    N = int(lens.sum())
    if N == 0:
        return np.zeros(0, np.int64)
    big = np.int64(1) << np.int64(32)
    dense = lens == span
    row = np.repeat(np.arange(R, dtype=np.int64), lens)
    draw = lo[row] + np.floor(rng.random(N) * span[row]).astype(np.int64)
    if dense.any():
        # rows that take their whole range: 0..span-1 in order
        starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
        offs = np.arange(N, dtype=np.int64) - np.repeat(starts, lens)
        m = dense[row]
        draw[m] = lo[row[m]] + offs[m]
    key = row * big + draw
    key.sort()
    for _ in range(10000):
        dup = np.nonzero(key[1:] == key[:-1])[0] + 1
        if dup.size == 0:
            break
        r = key[dup] // big
        key[dup] = r * big + lo[r] + np.floor(rng.random(dup.size) * span[r]).astype(np.int64)
        key.sort()
    else:  # pragma: no cover - needs ~span draws per row
        raise RuntimeError("column sampling did not converge")
    return key % big


def _finish(lens: np.ndarray, cols: np.ndarray, vals: np.ndarray, R: int, C: int) -> HostCSR:
    rp = np.zeros(R + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    if rp[-1] > np.iinfo(np.int32).max:
        raise ValueError("nnz exceeds int32")
    return HostCSR(rp.astype(np.int32), cols.astype(np.int32), vals, R, C)


def uniform_rows(R: int, C: int, lo: int, hi: int, seed: int, dtype=np.float32) -> HostCSR:
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, size=R).astype(np.int64)
    lens = np.minimum(lens, C)
    cols = _distinct_sorted(lens, np.zeros(R, np.int64), np.full(R, C, np.int64), rng)
    vals = rng.uniform(-1.0, 1.0, size=cols.size).astype(dtype)
    return _finish(lens, cols, vals, R, C)


def powerlaw_lengths(R: int, exponent: float, kmin: int, kmax: int, rng: np.random.Generator) -> np.ndarray:
    k = np.arange(kmin, kmax + 1, dtype=np.float64)
    p = k ** (-exponent)
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    u = rng.random(R)
    return (np.searchsorted(cdf, u, side="right") + kmin).astype(np.int64).clip(kmin, kmax)


def powerlaw_rows(R: int, C: int, seed: int, exponent: float = 2.0, kmin: int = 1, kmax: int = 100000,
                  dtype=np.float32) -> HostCSR:
    rng = np.random.default_rng(seed)
    lens = np.minimum(powerlaw_lengths(R, exponent, kmin, kmax, rng), C)
    cols = _distinct_sorted(lens, np.zeros(R, np.int64), np.full(R, C, np.int64), rng)
    vals = rng.uniform(-1.0, 1.0, size=cols.size).astype(dtype)
    return _finish(lens, cols, vals, R, C)


def laplacian_2d(n: int, dtype=np.float64) -> HostCSR:
    R = n * n
    r = np.arange(R, dtype=np.int64)
    i, j = r // n, r % n
    cand = np.stack([r - n, r - 1, r, r + 1, r + n], axis=1)
    ok = np.stack([i > 0, j > 0, np.ones(R, bool), j < n - 1, i < n - 1], axis=1)
    vals = np.where(cand == r[:, None], 4.0, -1.0)
    lens = ok.sum(axis=1)
    cols = cand[ok]
    v = vals[ok].astype(dtype)
    return _finish(lens, cols, v, R, R)


def banded(R: int, half_bw: int, nnz_per_row: int, seed: int, dtype=np.float32, normal: bool = True) -> HostCSR:
    rng = np.random.default_rng(seed)
    r = np.arange(R, dtype=np.int64)
    lo = np.maximum(r - half_bw, 0)
    hi = np.minimum(r + half_bw, R - 1)
    span = hi - lo + 1
    lens = np.minimum(np.full(R, nnz_per_row, np.int64), span)
    cols = _distinct_sorted(lens, lo, span, rng)
    vals = (rng.standard_normal(cols.size) if normal else rng.uniform(-1, 1, cols.size)).astype(dtype)
    return _finish(lens, cols, vals, R, R)


def block_diagonal(R: int, bs: int, seed: int, dtype=np.float32) -> HostCSR:
    rng = np.random.default_rng(seed)
    r = np.arange(R, dtype=np.int64)
    lo = (r // bs) * bs
    span = np.minimum(lo + bs, R) - lo
    lens = span.copy()
    cols = _distinct_sorted(lens, lo, span, rng)
    vals = rng.uniform(-1.0, 1.0, size=cols.size).astype(dtype)
    return _finish(lens, cols, vals, R, R)


def poisson_rows(R: int, C: int, mean: float, seed: int, dtype=np.float32) -> HostCSR:
    rng = np.random.default_rng(seed)
    lens = np.minimum(rng.poisson(mean, size=R).astype(np.int64), C)
    cols = _distinct_sorted(lens, np.zeros(R, np.int64), np.full(R, C, np.int64), rng)
    vals = rng.uniform(-1.0, 1.0, size=cols.size).astype(dtype)
    return _finish(lens, cols, vals, R, C)


def permuted(A: HostCSR, seed: int) -> HostCSR:
    """P_r A P_c^T with random row and column permutations (columns re-sorted per row)."""
    rng = np.random.default_rng(seed)
    prow = rng.permutation(A.numRows)
    pcol = rng.permutation(A.numCols).astype(np.int64)
    lens = A.row_lengths()[prow]
    rp = A.rowPtr.astype(np.int64)
    starts = rp[prow]
    idx = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens) + np.arange(int(lens.sum()))
    cols = pcol[A.colIdx[idx]]
    vals = A.values[idx]
    row = np.repeat(np.arange(A.numRows, dtype=np.int64), lens)
    order = np.lexsort((cols, row))
    return _finish(lens, cols[order], vals[order], A.numRows, A.numCols)


def uniform_vector(n: int, seed: int, dtype=np.float32) -> np.ndarray:
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=n).astype(dtype)


# ---- the named workloads ------------------------------------------------------
def cfg_uniform_100k(seed: int = 1) -> HostCSR:
    """100,000 x 100,000, row length U{5..15}, FP32."""
    return uniform_rows(100_000, 100_000, 5, 15, seed)


def cfg_powerlaw_2m(seed: int = 2) -> HostCSR:
    """2,000,000 x 2,000,000, row lengths p(k) ~ k^-2 on [1, 100000], FP32."""
    return powerlaw_rows(2_000_000, 2_000_000, seed)


def cfg_laplacian_4096() -> HostCSR:
    """5-point Laplacian on a 4096 x 4096 grid, FP64."""
    return laplacian_2d(4096)


def cfg_banded_4m(seed: int = 4) -> HostCSR:
    """4,000,000 rows, half-bandwidth 64, 16 nnz/row, values N(0,1)."""
    return banded(4_000_000, 64, 16, seed, dtype=np.float64)
