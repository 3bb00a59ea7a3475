"""GPU parity for the host-memory path (muxi_spmv_host.h): y computed from host x into host y, chunked and
overlapped, against the FP64 oracle within the FP32 bound; pinned and pageable buffers; alpha / beta;
more chunks than rows; FP64; the 2M power-law matrix at full size."""
import numpy as np
import pytest
import torch

import oracle
from muxi_spmv_amd import api, host
from muxi_spmv_amd import generators as G
from muxi_spmv_amd._lib import SpmvError
from gpu_util import assert_close_fp32, assert_close_fp64, dev_csr, fp32_bound, matrix_zoo

pytestmark = pytest.mark.gpu

ZOO = matrix_zoo()


@pytest.mark.parametrize("mname", ["uniform", "powerlaw", "empty_rows", "long_row", "banded", "blockdiag", "tiny",
                                   "single_row"])
@pytest.mark.parametrize("chunks", [1, 3, 8])
@pytest.mark.parametrize("pinned", [True, False])
def test_host_plan_matches_oracle(mname, chunks, pinned):
    h = ZOO[mname]
    A = dev_csr(h)
    hp = host.HostPlan(A, chunks)
    xs = G.uniform_vector(h.numCols, 9)
    if pinned:
        x = host.pinned_empty(h.numCols)
        x[:] = xs
        y = host.pinned_empty(h.numRows)
    else:
        x, y = xs.copy(), np.empty(h.numRows, np.float32)
    y[:] = np.nan
    hp.exec(x, y)
    assert_close_fp32(y, h, xs)
    # chunks cover the rows in order
    ends = [hp.chunk(k)[:2] for k in range(chunks)]
    assert ends[0][0] == 0 and ends[-1][1] == h.numRows
    assert all(ends[k][1] == ends[k + 1][0] for k in range(chunks - 1))
    t = hp.timing()
    assert t.total_ms > 0 and t.kernels_ms >= 0 and t.total_ms + 1e-3 >= t.kernels_ms
    hp.destroy()


def test_alpha_beta_and_repeat():
    h = ZOO["powerlaw"]
    A = dev_csr(h)
    hp = host.HostPlan(A, 4)
    x = G.uniform_vector(h.numCols, 10)
    y0 = G.uniform_vector(h.numRows, 11)
    y = y0.copy()
    hp.exec(x, y, 2.0, -0.5)
    y64, ab = oracle.spmv_f64(h.rowPtr, h.colIdx, h.values.astype(np.float64), x.astype(np.float64))
    exp = 2.0 * y64 - 0.5 * y0.astype(np.float64)
    bound = 2.0 * fp32_bound(h.rowPtr, ab) + 2.0 ** -23 * (np.abs(exp) + 0.5 * np.abs(y0))
    assert np.all(np.abs(y - exp) <= bound)
    # beta = 0 never reads y; a second run gives the same bits
    y1 = np.full(h.numRows, np.nan, np.float32)
    y2 = np.full(h.numRows, np.inf, np.float32)
    hp.exec(x, y1)
    hp.exec(x, y2)
    assert np.array_equal(y1, y2)


def test_more_chunks_than_rows_and_fp64():
    h = ZOO["tiny"]
    A = dev_csr(h, torch.float64)
    hp = host.HostPlan(A, 16, xtype=torch.float64)
    x = G.uniform_vector(h.numCols, 12).astype(np.float64)
    y = np.full(h.numRows, np.nan)
    hp.exec(x, y)
    assert_close_fp64(torch.from_numpy(y), h, x)


def test_banded_sends_x_in_pieces():
    h = ZOO["banded"]
    A = dev_csr(h)
    hp = host.HostPlan(A, 8, "merge", 4)
    x = G.uniform_vector(h.numCols, 13)
    y = np.empty(h.numRows, np.float32)
    hp.exec(x, y)
    assert_close_fp32(y, h, x)
    assert [hp.chunk(k)[2] for k in range(8)] == ["merge"] * 8


def test_errors():
    h = ZOO["tiny"]
    A = dev_csr(h)
    with pytest.raises(SpmvError):
        host.HostPlan(A, 0)
    with pytest.raises(SpmvError):
        host.HostPlan(A, 65)
    hp = host.HostPlan(A, 2)
    with pytest.raises(ValueError):
        hp.exec(np.zeros(h.numCols + 1, np.float32), np.zeros(h.numRows, np.float32))
    with pytest.raises(TypeError):
        hp.exec(np.zeros(h.numCols), np.zeros(h.numRows))


def test_full_size_powerlaw_end_to_end():
    """2M x 2M power-law: host x -> host y with 4 chunks, rel L2 <= 1e-6 against the device FP64 oracle."""
    h = G.cfg_powerlaw_2m()
    A = dev_csr(h)
    hp = host.HostPlan(A, 4)
    x = host.pinned_empty(h.numCols)
    x[:] = G.uniform_vector(h.numCols, 14)
    y = host.pinned_empty(h.numRows)
    hp.exec(x, y)
    y64 = api.oracle_f64(A, torch.from_numpy(np.array(x)).cuda()).cpu().numpy()
    rl2 = np.linalg.norm(y.astype(np.float64) - y64) / np.linalg.norm(y64)
    assert rl2 <= 1e-6, rl2


@pytest.mark.parametrize("kernel", ["ell", "auto"])
@pytest.mark.parametrize("chunks", [1, 3, 4])
def test_row_blocks_of_a_banded_matrix(kernel, chunks):
    """Chunk plans are row blocks of the whole matrix: the ELL plan takes its 16-bit column offsets from the
    whole matrix's row (row + first row of the chunk), so later chunks of a banded matrix, whose columns lie
    near their rows in the whole matrix but far from their 0-based chunk rows, decode the right columns."""
    rng = np.random.default_rng(3)
    R, half = 90_000, 300
    rows = [np.unique(np.clip(r + rng.integers(-half, half + 1, size=5), 0, R - 1)) for r in range(R)]
    rp = np.zeros(R + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.int32)
    h = G.HostCSR(rp, ci, rng.uniform(-1, 1, ci.size).astype(np.float32), R, R)
    A = dev_csr(h)
    hp = host.HostPlan(A, chunks, kernel)
    xs = G.uniform_vector(R, 12)
    y = np.full(R, np.nan, np.float32)
    hp.exec(xs, y)
    assert_close_fp32(y, h, xs)
    if kernel == "ell":
        assert all(hp.chunk(k)[2] == "ell" for k in range(chunks))
    hp.destroy()


@pytest.mark.parametrize("chunks", [1, 4, 8])
def test_full_size_powerlaw_integers_exact(chunks):
    """Host x -> host y on the 2M power-law pattern with integer values in [-3, 3] and x in [-4, 4]: every chunk
    plan (a row block, AUTO) and the column pieces of x give y equal to the host's integer product."""
    h = G.cfg_powerlaw_2m()
    i = np.arange(h.nnz, dtype=np.int64)
    vi = ((i * 2654435761) >> 7) % 7 - 3
    xi = (np.arange(h.numCols, dtype=np.int64) * 40503 >> 3) % 9 - 4
    exp = np.add.reduceat(vi * xi[h.colIdx], h.rowPtr[:-1].astype(np.int64))
    exp[np.diff(h.rowPtr) == 0] = 0
    A = dev_csr(G.HostCSR(h.rowPtr, h.colIdx, vi.astype(np.float32), h.numRows, h.numCols))
    hp = host.HostPlan(A, chunks)
    x = host.pinned_empty(h.numCols)
    x[:] = xi.astype(np.float32)
    y = host.pinned_empty(h.numRows)
    y[:] = np.nan
    hp.exec(x, y)
    assert np.array_equal(np.array(y), exp.astype(np.float32))
    hp.destroy()
