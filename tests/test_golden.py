"""Golden fixtures (tests/golden/fixtures.npz, written by
tests/golden/make_golden.py): the oracle and the library's host permutations
reproduce them bit for bit. CPU only."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from muxi_spmv_amd import prep

FIX = np.load(Path(__file__).resolve().parent / "golden" / "fixtures.npz", allow_pickle=False)
CASES = sorted({k.split("_")[0] for k in FIX.files})


@pytest.mark.parametrize("c", CASES)
def test_oracle_reproduces_fixtures(c):
    f = lambda k: FIX[f"{c}_{k}"]  # noqa: E731
    rp, ci, va, C = f("rowPtr"), f("colIdx"), f("values"), int(f("numCols"))
    N = int(rp[-1])
    R = rp.size - 1
    assert np.array_equal(oracle.merge_partitions(R, N, f("merge_pos").size - 1), f("merge_pos"))
    r, n = oracle.merge_path_coords(rp, N, f("merge_pos"))
    assert np.array_equal(r, f("merge_row")) and np.array_equal(n, f("merge_nnz"))
    tr, to = oracle.csr5_tiles(rp, N, 16)
    assert np.array_equal(tr, f("csr5_row")) and np.array_equal(to, f("csr5_off"))
    ec, ev = oracle.csr_to_ell(rp, ci, va, oracle.ell_width(rp))
    assert np.array_equal(ec, f("ell_col")) and np.array_equal(ev, f("ell_val"))
    for name, fn in [("com", lambda: oracle.reorder_columns_com(rp, ci, C)),
                     ("minrow", lambda: oracle.reorder_columns_minrow(rp, ci, C)),
                     ("freq", lambda: oracle.reorder_columns_freq(rp, ci, C))]:
        a, b = fn()
        key = "fwd" if name == "com" else "perm"
        assert np.array_equal(a, f(f"{name}_{key}")) and np.array_equal(b, f(f"{name}_cols")), name
    assert np.array_equal(oracle.reorder_rows_by_length(rp, ci, va)[0], f("rowlen_map"))
    assert np.array_equal(oracle.spmv_csr_f32(rp, ci, va, f("x")), f("y_f32"))


@pytest.mark.parametrize("c", CASES)
def test_library_reproduces_fixtures(c):
    f = lambda k: FIX[f"{c}_{k}"]  # noqa: E731
    rp, ci, va, C = f("rowPtr"), f("colIdx"), f("values"), int(f("numCols"))
    fwd, nci, _ = prep.reorder_columns(rp, ci, C)
    assert np.array_equal(fwd, f("com_fwd")) and np.array_equal(nci, f("com_cols"))
    p, n = prep.reorder_columns_minrow(rp, ci, C)
    assert np.array_equal(p, f("minrow_perm")) and np.array_equal(n, f("minrow_cols"))
    p, n = prep.reorder_columns_by_frequency(rp, ci, C)
    assert np.array_equal(p, f("freq_perm")) and np.array_equal(n, f("freq_cols"))
    assert np.array_equal(prep.reorder_rows_by_length(rp, ci, va)[0], f("rowlen_map"))
