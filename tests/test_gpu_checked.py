"""Checked plans (muxi_plan_create_checked): the per-row FP64 fallback.

A checked exec sums |a_rj x_j| per row beside y on the LDS column slabs, tests each row's a-posteriori error bound
-- (n+2)u S for FP32 sums, 2u|y| + 2((n+2)u)^2 S for Dot2 (u = 2^-24, slightly inflated for the rounding of S) --
against rtol |y_r|, and recomputes the rows that fail in FP64. The matrix here has benign rows (positive values
and x, at most 16 nonzeros: S = |y|, far inside the bound) and rows built to cancel:
  mild    x = 1 at the columns used, values (1, 1e-3, -1) in column order: y = 1e-3, S = 2.001 -- fails the
          FP32 bound at rtol 1e-5 (and a plain FP32 sum is off by 5e-5 relative), passes the Dot2 bound;
  extreme values (1, 1e-9, -1): a plain FP32 sum gives 0; fails both bounds;
  zero    values (1, -1): y = 0 with S = 2 -- fails both.
Exactly those rows must come back as recomputed, and every row of y must lie within rtol of the FP64 oracle
(recomputed rows within FP32 rounding of it).
"""
import numpy as np
import pytest
import torch

from muxi_spmv_amd import api
from muxi_spmv_amd import generators as G
from gpu_util import dev_csr

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _matrix(R=60_000, C=100_000, seed=3):
    rng = np.random.default_rng(seed)
    kind = np.zeros(R, np.int8)  # 0 benign, 1 mild, 2 extreme, 3 zero
    special = rng.choice(R, 300, replace=False)
    kind[special[:100]] = 1
    kind[special[100:200]] = 2
    kind[special[200:]] = 3
    rows, cols, vals = [], [], []
    for r in range(R):
        k = kind[r]
        if k == 0:
            n = int(rng.integers(1, 17))
            c = np.unique(rng.integers(8, C, n))
            v = rng.uniform(0.5, 1.5, c.size)
        else:
            base = int(rng.integers(0, 3))  # x = 1 at columns 0..7
            if k == 3:
                c, v = np.array([base, base + 3]), np.array([1.0, -1.0])
            else:
                c, v = np.array([base, base + 3, base + 5]), np.array([1.0, 1e-3 if k == 1 else 1e-9, -1.0])
        rows.append(np.full(c.size, r))
        cols.append(c)
        vals.append(v)
    rr = np.concatenate(rows)
    rp = np.zeros(R + 1, np.int32)
    np.add.at(rp, rr + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    h = G.HostCSR(rp, np.concatenate(cols).astype(np.int32), np.concatenate(vals).astype(np.float32), R, C)
    x = rng.uniform(0.5, 1.5, C).astype(np.float32)
    x[:8] = 1.0
    return h, x, kind


@pytest.fixture(scope="module")
def cancel():
    h, x, kind = _matrix()
    A = dev_csr(h)
    xd = torch.from_numpy(x).cuda()
    y64 = api.oracle_f64(A, xd).cpu().numpy()
    return h, A, xd, kind, y64


@pytest.mark.parametrize("comp", [False, True])
def test_checked_recomputes_exactly_the_failing_rows(cancel, comp):
    h, A, x, kind, y64 = cancel
    p = api.Plan.checked(A, RTOL, compensated=comp)
    assert p.info().kernel == "lslab"
    y = torch.full((h.numRows,), float("nan"), device="cuda")
    p.exec(x, y)
    torch.cuda.synchronize()
    got = p.checked_rows()
    want = np.nonzero(kind >= (2 if comp else 1))[0]
    assert np.array_equal(got, want), (got.size, want.size, np.setdiff1d(got, want)[:5], np.setdiff1d(want, got)[:5])
    yy = y.cpu().numpy().astype(np.float64)
    err = np.abs(yy - y64)
    assert np.all(err <= RTOL * np.abs(y64) * 1.0001), np.nonzero(err > RTOL * np.abs(y64))[0][:5]
    # recomputed rows: FP64 sums rounded once to FP32
    assert np.all(err[want] <= 2.0 ** -24 * np.abs(y64[want]) * 1.0001)
    assert np.all(yy[kind == 3] == 0.0)
    # without the fallback a plain FP32 sum loses the extreme rows entirely
    if not comp:
        q = api.Plan(A, "lslab", 4 | 0x100)
        y2 = torch.empty_like(y)
        q.exec(x, y2)
        torch.cuda.synchronize()
        e2 = np.abs(y2.cpu().numpy().astype(np.float64) - y64)
        assert np.all(e2[kind == 2] >= 0.5 * np.abs(y64[kind == 2]))


@pytest.mark.parametrize("comp", [False, True])
def test_checked_alpha_beta_and_repeat(cancel, comp):
    h, A, x, kind, y64 = cancel
    p = api.Plan.checked(A, RTOL, compensated=comp)
    y0 = torch.from_numpy(np.linspace(-1, 1, h.numRows).astype(np.float32)).cuda()
    for rep in range(2):  # the recomputed-row count starts from zero on every exec
        y = y0.clone()
        p.exec(x, y, alpha=2.0, beta=0.5)
        torch.cuda.synchronize()
        ref = 2.0 * y64 + 0.5 * y0.cpu().numpy().astype(np.float64)
        err = np.abs(y.cpu().numpy().astype(np.float64) - ref)
        scale = 2.0 * np.abs(y64) + 0.5 * np.abs(y0.cpu().numpy())
        assert np.all(err <= 2 * RTOL * scale + 1e-30), rep
        assert p.checked_rows().size == int((kind >= (2 if comp else 1)).sum()), rep


def test_checked_benign_matrix_recomputes_nothing():
    """Positive values and x, rows of 1 to 16 nonzeros: every bound is within rtol, nothing is recomputed."""
    rng = np.random.default_rng(7)
    R = C = 200_000
    lens = rng.integers(1, 17, R)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = rng.integers(0, C, int(rp[-1])).astype(np.int32)
    h = G.HostCSR(rp, ci, rng.uniform(0.5, 1.5, ci.size).astype(np.float32), R, C)
    A = dev_csr(h)
    x = torch.from_numpy(np.random.default_rng(8).uniform(0.5, 1.5, h.numCols).astype(np.float32)).cuda()
    p = api.Plan.checked(A, RTOL)
    y = torch.empty(h.numRows, device="cuda")
    p.exec(x, y)
    torch.cuda.synchronize()
    assert p.checked_rows().size == 0
    y64 = api.oracle_f64(A, x)
    rl2 = float(torch.linalg.vector_norm(y.double() - y64) / torch.linalg.vector_norm(y64))
    assert rl2 <= 1e-6


def test_checked_plan_limits():
    # 15,000,000 columns: 524 slabs of 28672 (the checked plain plan's width), over the 512 a plan takes
    h = G.HostCSR(np.array([0, 1], np.int32), np.array([14_999_999], np.int32), np.ones(1, np.float32), 1,
                  15_000_000)
    with pytest.raises(api.SpmvError, match="INVALID_MATRIX"):
        api.Plan.checked(dev_csr(h), RTOL)
    q = api.Plan(dev_csr(G.HostCSR(np.array([0, 1], np.int32), np.array([0], np.int32), np.ones(1, np.float32), 1,
                                   1)), "merge", 4)
    with pytest.raises(api.SpmvError, match="INVALID_HANDLE"):
        q.checked_rows()
