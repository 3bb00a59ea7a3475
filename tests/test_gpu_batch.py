"""GPU parity of the batched persistent SpMV and its hipGraph replays."""
import numpy as np
import pytest
import torch

from muxi_spmv_amd import api, batch
from muxi_spmv_amd import generators as G
import oracle
from gpu_util import assert_close_fp32, dev_csr, fp32_bound, matrix_zoo

pytestmark = pytest.mark.gpu


def make_batch(n=40, seed=0, dtype=torch.float32):
    rng = np.random.default_rng(seed)
    hs = []
    for i in range(n):
        R = int(np.exp(rng.uniform(np.log(50), np.log(20000))))
        mean = float(np.exp(rng.uniform(np.log(1), np.log(64))))
        hs.append(G.poisson_rows(R, R, mean, seed=1000 + i))
    zoo = matrix_zoo()
    hs += [zoo["empty_rows"], zoo["long_row"], zoo["tiny"], zoo["single_row"]]
    np_dt = np.float64 if dtype == torch.float64 else np.float32
    mats = [dev_csr(h, dtype) for h in hs]
    xs_h = [G.uniform_vector(h.numCols, 77 + i, np_dt) for i, h in enumerate(hs)]
    xs = [torch.from_numpy(x).cuda() for x in xs_h]
    ys = [torch.full((h.numRows,), float("nan"), dtype=dtype, device="cuda") for h in hs]
    return hs, mats, xs_h, xs, ys


@pytest.mark.parametrize("ipt", [4, 8, 16])
def test_batch_matches_oracle(ipt):
    hs, mats, xs_h, xs, ys = make_batch()
    bp = batch.BatchPlan(mats, xs, ys, ipt)
    bp.exec()
    torch.cuda.synchronize()
    for h, x, y in zip(hs, xs_h, ys):
        assert_close_fp32(y, h, x)


def test_batch_fp64_and_alpha_beta():
    hs, mats, xs_h, xs, ys = make_batch(12, seed=3, dtype=torch.float64)
    for y in ys:
        y.fill_(1.0)
    bp = batch.BatchPlan(mats, xs, ys)
    bp.exec(alpha=2.0, beta=0.5)
    torch.cuda.synchronize()
    from oracle import spmv_f64
    for h, x, y in zip(hs, xs_h, ys):
        # y = 2*A*x + 0.5*1: compare A*x recovered from y against the FP64 oracle
        y64, ab = spmv_f64(h.rowPtr, h.colIdx, h.values.astype(np.float64), x)
        got = (y.cpu().numpy() - 0.5) / 2.0
        n = np.diff(h.rowPtr.astype(np.int64))
        assert np.all(np.abs(got - y64) <= (n + 4) * 2.0 ** -53 * (ab + 1.0) * 2)


def test_per_matrix_paths_match():
    hs, mats, xs_h, xs, ys = make_batch(20, seed=5)
    bp = batch.BatchPlan(mats, xs, ys)
    for kern, p in [("vector", 0), ("scalar", 0), ("vwave", 8)]:
        for y in ys:
            y.fill_(float("nan"))
        bp.exec_per_matrix(kern, p)
        torch.cuda.synchronize()
        for h, x, y in zip(hs, xs_h, ys):
            assert_close_fp32(y, h, x)


MODES = {"merge": batch.TUNED_MERGE, "slabs": batch.BATCH_SLABS}


@pytest.mark.parametrize("mode", MODES)
def test_graph_replays_match_eager(mode):
    hs, mats, xs_h, xs, ys = make_batch(24, seed=7)
    bp = batch.BatchPlan(mats, xs, ys, MODES[mode])
    assert bp.info()["mode"] == mode
    bp.exec()
    torch.cuda.synchronize()
    eager = [y.clone() for y in ys]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g = bp.graph(stream=s)
        gp = bp.graph("vector", 0, stream=s)
    torch.cuda.synchronize()
    for rep in range(4):
        for y in ys:
            y.fill_(float("nan"))  # eager work between replays
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        for y, e in zip(ys, eager):
            assert torch.equal(y, e), rep  # bitwise: same kernels, same order
        for y in ys:
            y.zero_()
        gp.launch(s)
        s.synchronize()
        for h, x, y in zip(hs, xs_h, ys):
            assert_close_fp32(y, h, x)


@pytest.mark.parametrize("mode", MODES)
def test_batch_integers_exact(mode):
    """Integer values in [-3, 3] and x in [-4, 4] on the mixed batch (Poisson matrices, empty rows, a long
    row, tiny matrices): every matrix's y equals its integer product on the host, bit for bit, in both modes."""
    hs, mats, xs_h, xs, ys = make_batch(n=30, seed=8)
    mats_i, xs_i, exps = [], [], []
    for m, h in enumerate(hs):
        i = np.arange(h.nnz, dtype=np.int64)
        vi = ((i * 2654435761 + m) >> 7) % 7 - 3
        xi = (np.arange(h.numCols, dtype=np.int64) * 40503 >> 3) % 9 - 4
        e = np.zeros(h.numRows, np.int64)
        if h.nnz:
            rows = np.repeat(np.arange(h.numRows), np.diff(h.rowPtr))
            np.add.at(e, rows, vi * xi[h.colIdx])
        exps.append(e)
        mats_i.append(dev_csr(G.HostCSR(h.rowPtr, h.colIdx, vi.astype(np.float32), h.numRows, h.numCols)))
        xs_i.append(torch.from_numpy(xi.astype(np.float32)).cuda())
    bp = batch.BatchPlan(mats_i, xs_i, ys, MODES[mode])
    assert bp.info()["mode"] == mode
    bp.exec()
    torch.cuda.synchronize()
    for e, y in zip(exps, ys):
        assert np.array_equal(y.cpu().numpy(), e.astype(np.float32))


def test_batch_slabs_alpha_beta_and_nan_y():
    """The slab mode scatters y with alpha and beta: y = 2 A x + 0.5 y over y = 1, and beta = 0 over a NaN y
    (never read); members with empty rows, one long row and tiny matrices included."""
    hs, mats, xs_h, xs, ys = make_batch(18, seed=11)
    bp = batch.BatchPlan(mats, xs, ys, batch.BATCH_SLABS)
    info = bp.info()
    assert info["mode"] == "slabs" and info["tiles"] == 0 and info["scratch_bytes"] > 0
    for y in ys:
        y.fill_(1.0)
    bp.exec(alpha=2.0, beta=0.5)
    torch.cuda.synchronize()
    for h, x, y in zip(hs, xs_h, ys):
        y64, ab = oracle.spmv_f64(h.rowPtr, h.colIdx, h.values.astype(np.float64), np.asarray(x, np.float64))
        want = 2.0 * y64 + 0.5
        # twice the product's FP32 bound, then the rounding of alpha * sum and of the sum with beta * y
        tol = 2.0 * fp32_bound(h.rowPtr, ab) + 2.0 ** -23 * (np.abs(want) + 2.0 * np.abs(y64) + 1.0)
        assert np.all(np.abs(y.double().cpu().numpy() - want) <= tol)
    for y in ys:
        y.fill_(float("nan"))
    bp.exec()
    torch.cuda.synchronize()
    for h, x, y in zip(hs, xs_h, ys):
        assert_close_fp32(y, h, x)


def test_batch_slabs_needs_fp32():
    hs, mats, xs_h, xs, ys = make_batch(6, seed=12, dtype=torch.float64)
    with pytest.raises(Exception, match="UNSUPPORTED_FORMAT"):
        batch.BatchPlan(mats, xs, ys, batch.BATCH_SLABS)
    assert batch.BatchPlan(mats, xs, ys).info()["mode"] == "merge"  # the default: merge for FP64


def test_batch_slabs_degenerate_members():
    """The slab mode with members that have no rows, rows but no columns, rows and columns but no nonzeros, and
    many tiny members inside one copy workgroup's range (1,024 elements): every y equals its integer product."""
    rng = np.random.default_rng(21)
    hs = []
    for i in range(60):
        kind = i % 6
        if kind == 0:
            hs.append(G.HostCSR(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 0, 5))
        elif kind == 1:
            hs.append(G.HostCSR(np.zeros(4, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 3, 0))
        elif kind == 2:
            hs.append(G.HostCSR(np.zeros(8, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), 7, 9))
        else:
            R = int(rng.integers(1, 40))
            hs.append(G.poisson_rows(R, R, 3.0, seed=3000 + i))
    hs.append(G.poisson_rows(70_000, 70_000, 6.0, seed=77))  # one member over two slabs
    mats, xs, ys, exps = [], [], [], []
    for m, h in enumerate(hs):
        i = np.arange(h.nnz, dtype=np.int64)
        vi = ((i * 2654435761 + m) >> 7) % 7 - 3
        xi = (np.arange(h.numCols, dtype=np.int64) * 40503 + m >> 3) % 9 - 4
        e = np.zeros(h.numRows, np.int64)
        if h.nnz:
            rows = np.repeat(np.arange(h.numRows), np.diff(h.rowPtr))
            np.add.at(e, rows, vi * xi[h.colIdx])
        exps.append(e)
        mats.append(dev_csr(G.HostCSR(h.rowPtr, h.colIdx, vi.astype(np.float32), h.numRows, h.numCols)))
        xs.append(torch.from_numpy(xi.astype(np.float32)).cuda())
        ys.append(torch.full((h.numRows,), float("nan"), device="cuda"))
    bp = batch.BatchPlan(mats, xs, ys, batch.BATCH_SLABS)
    assert bp.info()["mode"] == "slabs"
    bp.exec()
    torch.cuda.synchronize()
    for m, (e, y) in enumerate(zip(exps, ys)):
        assert np.array_equal(y.cpu().numpy(), e.astype(np.float32)), m


def test_batch_slabs_rejects_a_member_whose_row_pointers_do_not_start_at_zero():
    """Row pointers from 2 to nnz + 2 would shift the member's block of the block-diagonal matrix against its
    copied columns: the slab mode refuses them."""
    good = G.poisson_rows(100, 100, 4.0, seed=1)
    rp = good.rowPtr.astype(np.int32) + 2
    ci = np.concatenate([np.zeros(2, np.int32), good.colIdx.astype(np.int32)])
    va = np.concatenate([np.zeros(2, np.float32), good.values.astype(np.float32)])
    shifted = api.CSRMatrix(torch.from_numpy(rp), torch.from_numpy(ci), torch.from_numpy(va), 100)
    shifted.c.nnz = good.nnz  # the member's nonzeros, with rowPtr[0] = 2
    x = torch.ones(100, device="cuda")
    y = torch.zeros(100, device="cuda")
    with pytest.raises(api.SpmvError, match="INVALID_MATRIX"):
        batch.BatchPlan([dev_csr(good), shifted], [x, x], [y, y.clone()], batch.BATCH_SLABS)
