"""Parity at the named workload sizes that the other files do not cover.

* 100,000 x 100,000 uniform rows (5..15 nonzeros, FP32): every kernel against the CPU FP32 loop
  (tests/test_column_reorder.cu:87-93) under the reference's own tolerance rule |d| <= max(1e-4 |ref|, 1e-6)
  (tests/deep_analysis.cu:255), and against the FP64 oracle under the FP32 summation bound and rel L2 <= 1e-6.
* 4,000,000-row banded matrix (half-bandwidth 64, 16 nonzeros a row, N(0,1)) after RCM: values FP64 / FP32 /
  BF16 with FP32 x, plain and Dot2-compensated FP32 accumulation; rel L2 against the FP64 oracle on the stored
  values <= 1e-6, compensated no worse than plain, BF16 storage error below its unit roundoff.
"""
import numpy as np
import pytest
import torch

import oracle
from muxi_spmv_amd import api, prep, tune
from muxi_spmv_amd import generators as G
from gpu_util import assert_close_fp32, dev_csr

pytestmark = pytest.mark.gpu

KALL = [("scalar", 0), ("vector", 0), ("vwave", 8), ("adaptive", 1024), ("merge", 4), ("merge", 8), ("csr5", 256),
        ("ell", 0), ("xslab", 1028), ("xslab", 0), ("lslab", 8), ("auto", 0)]


@pytest.fixture(scope="module")
def uniform_100k():
    h = G.cfg_uniform_100k()
    return h, dev_csr(h)


@pytest.mark.parametrize("kernel,param", KALL)
def test_uniform_100k_every_kernel(uniform_100k, kernel, param):
    h, A = uniform_100k
    x = G.uniform_vector(h.numCols, 1)
    y = torch.full((h.numRows,), float("nan"), device="cuda")
    api.Plan(A, kernel, param).exec(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    yy = y.cpu().numpy()
    ref = oracle.spmv_csr_f32(h.rowPtr, h.colIdx, h.values, x)
    assert np.all(np.abs(yy - ref) <= np.maximum(1e-4 * np.abs(ref), 1e-6))
    assert_close_fp32(yy, h, x, rel_l2=1e-6)


@pytest.fixture(scope="module")
def banded_4m_rcm():
    B = tune.gen_matrix("banded", 4_000_000, 4_000_000, 64, 16, 0, normal=True, dtype=torch.float64, seed=4)
    rp, ci, va = B.host()
    B.free()
    perm = prep.rcm(rp, ci)
    nrp, nci, nva = prep.permute_symmetric(rp, ci, va, perm)
    assert prep.bandwidth(nrp, nci) <= 2 * prep.bandwidth(rp, ci)  # RCM keeps the band (64 -> 91)
    return api.CSRMatrix(nrp, nci, torch.from_numpy(nva), 4_000_000)


def _rel_l2(y, y64):
    return float(torch.linalg.vector_norm(y.double() - y64) / torch.linalg.vector_norm(y64))


@pytest.mark.parametrize("vt", [torch.float64, torch.float32, torch.bfloat16])
def test_banded_4m_precision(banded_4m_rcm, vt):
    A64 = banded_4m_rcm
    x = torch.from_numpy(G.uniform_vector(A64.numCols, 99)).cuda()
    y64 = api.oracle_f64(A64, x)  # the FP64 values
    Av = A64.with_values(api.convert_values(A64.d_values, vt))
    y64v = api.oracle_f64(Av, x)  # the stored values
    errs = {}
    for kernel, param, comp in (("merge", 4, False), ("merge", 4, True), ("xslab", 0, False), ("lslab", 8, False),
                                ("lslab", 8, True), ("auto", 0, False)):
        y = torch.full((A64.numRows,), float("nan"), device="cuda")
        api.Plan(Av, kernel, param, compensated=comp).exec(x, y)
        torch.cuda.synchronize()
        errs[(kernel, comp)] = _rel_l2(y, y64v)
        assert errs[(kernel, comp)] <= 1e-6, (kernel, comp, errs[(kernel, comp)])
        if vt == torch.bfloat16:
            assert _rel_l2(y, y64) <= 2.0 ** -8  # the storage rounding, not the kernel
    assert errs[("merge", True)] <= errs[("merge", False)] * 1.01
    assert errs[("lslab", True)] <= errs[("lslab", False)] * 1.01


@pytest.mark.parametrize("vt", [torch.float32, torch.bfloat16])
def test_banded_4m_integers_exact(banded_4m_rcm, vt):
    """The RCM-ordered banded matrix at full size with integer values in [-3, 3] (exact in BF16 too) and x in
    [-4, 4]: y equals the host's integer product row for row for ELL (16-bit offsets, four slot groups), the
    AUTO pick and merge."""
    A64 = banded_4m_rcm
    rp = A64.d_rowPtr.cpu().numpy().astype(np.int64)
    ci = A64.d_colIdx.cpu().numpy().astype(np.int64)
    i = np.arange(ci.size, dtype=np.int64)
    vi = ((i * 2654435761) >> 7) % 7 - 3
    xi = (np.arange(A64.numCols, dtype=np.int64) * 40503 >> 3) % 9 - 4
    exp = np.add.reduceat(vi * xi[ci], rp[:-1])
    exp[np.diff(rp) == 0] = 0
    A = A64.with_values(api.convert_values(torch.from_numpy(vi.astype(np.float64)).cuda(), vt))
    x = torch.from_numpy(xi.astype(np.float32)).cuda()
    for kernel in ("ell", "auto", "merge"):
        y = torch.full((A64.numRows,), float("nan"), device="cuda")
        api.Plan(A, kernel, 4 if kernel == "merge" else 0).exec(x, y)
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), exp.astype(np.float32)), kernel
