"""Shared helpers for the GPU parity tests: device upload and the tolerance
rule against the FP64 oracle."""
import numpy as np
import torch

import oracle
from muxi_spmv_amd import api
from muxi_spmv_amd import generators as G

U32 = 2.0 ** -24
U64 = 2.0 ** -53


def dev_csr(h: G.HostCSR, dtype=torch.float32) -> api.CSRMatrix:
    vals = torch.from_numpy(np.ascontiguousarray(h.values)).to(dtype)
    return api.CSRMatrix(h.rowPtr, h.colIdx, vals, h.numCols)


def fp32_bound(rowPtr, absrow, extra_ulps=2.0):
    """|y - y64| <= (n_r + extra) * u32 * sum_j |a_rj x_j|: the worst-case
    error of FP32 products plus any-order FP32 summation of n_r terms."""
    n = np.diff(np.asarray(rowPtr, np.int64)).astype(np.float64)
    return (n + extra_ulps) * U32 * absrow * 1.0001 + 1e-37


def fp64_bound(rowPtr, absrow):
    n = np.diff(np.asarray(rowPtr, np.int64)).astype(np.float64)
    return (n + 2) * U64 * absrow * 1.0001 + 1e-300


def assert_close_fp32(y, h: G.HostCSR, x, values=None, rel_l2=1e-5):
    """y (FP32 result) vs the FP64 oracle on the same stored values and x:
    the per-row worst-case bound, and a relative L2 error of at most rel_l2.
    1e-5 is the general cap: a single 60,000-term row can dominate the norm
    with a legitimately larger relative error; the benchmark workloads are
    held to 1e-6 where they are tested."""
    v = h.values if values is None else values
    y64, ab = oracle.spmv_f64(h.rowPtr, h.colIdx, np.asarray(v, np.float64), np.asarray(x, np.float64))
    yy = y.detach().cpu().double().numpy() if torch.is_tensor(y) else np.asarray(y, np.float64)
    err = np.abs(yy - y64)
    bound = fp32_bound(h.rowPtr, ab)
    bad = np.nonzero(err > bound)[0]
    assert bad.size == 0, f"{bad.size} rows over bound, first {bad[:5]} err {err[bad[:5]]} bound {bound[bad[:5]]}"
    nrm = np.linalg.norm(y64)
    if nrm > 0:
        rl2 = np.linalg.norm(yy - y64) / nrm
        assert rl2 <= rel_l2, rl2
    return y64


def assert_close_fp64(y, h: G.HostCSR, x):
    y64, ab = oracle.spmv_f64(h.rowPtr, h.colIdx, np.asarray(h.values, np.float64), np.asarray(x, np.float64))
    yy = y.detach().cpu().numpy()
    err = np.abs(yy - y64)
    assert np.all(err <= fp64_bound(h.rowPtr, ab))
    return y64


def matrix_zoo():
    """Small matrices covering the structural edge cases."""
    z = {}
    z["uniform"] = G.uniform_rows(3000, 3000, 5, 15, seed=1)
    z["powerlaw"] = G.powerlaw_rows(20000, 20000, seed=2, kmax=6000)
    # empty rows everywhere, including leading and trailing runs
    rng = np.random.default_rng(3)
    lens = rng.integers(0, 12, 4000)
    lens[rng.random(4000) < 0.4] = 0
    lens[:37] = 0
    lens[-53:] = 0
    cols = G._distinct_sorted(lens, np.zeros(4000, np.int64), np.full(4000, 5000, np.int64), rng)
    z["empty_rows"] = G._finish(lens, cols, rng.uniform(-1, 1, cols.size).astype(np.float32), 4000, 5000)
    # one very long row among short ones
    lens = np.full(3000, 3, np.int64)
    lens[1234] = 60000
    cols = G._distinct_sorted(lens, np.zeros(3000, np.int64), np.full(3000, 70000, np.int64), rng)
    z["long_row"] = G._finish(lens, cols, rng.uniform(-1, 1, cols.size).astype(np.float32), 3000, 70000)
    z["single_row"] = G.uniform_rows(1, 500, 200, 200, seed=4)
    z["tiny"] = G.uniform_rows(7, 9, 0, 3, seed=5)
    z["banded"] = G.banded(5000, 64, 16, seed=6, normal=True)
    z["blockdiag"] = G.block_diagonal(2048, 32, seed=7)
    return z
