"""Writes tests/golden/fixtures.npz: small seeded inputs and the oracle's
outputs for every integer-exact function on the path.

Why fixtures: reorderColumnsRCM and reorderColumnsByFrequency in the reference
(tests/test_column_reordering.cu:116-197) order columns with an unstable
std::sort whose comparator ties on most columns; the order among ties is
whatever libstdc++'s introsort produces for that input. Nothing in the
mathematics fixes it, so these fixtures (written with g++ 11.4 / libstdc++
on x86-64) pin it, and the test suite checks that both the oracle and the
library reproduce them on the machine where the tests run. The other entries
(merge-path coordinates, CSR5 tiles, ELL slots, centre-of-mass and row-length
orders) are fixed by their definitions and are kept as regression vectors.

Run from the repository root: python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402


def small_csr(R, C, max_len, seed, empty_frac=0.25):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, size=R)
    lens[rng.random(R) < empty_frac] = 0
    lens = np.minimum(lens, C)
    rp = np.zeros(R + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    cols = np.concatenate([np.sort(rng.choice(C, size=k, replace=False)) for k in lens]).astype(np.int32)
    vals = rng.uniform(-1, 1, size=cols.size).astype(np.float32)
    return rp, cols, vals


def main():
    out = {}
    for case, (R, C, L, seed) in enumerate([(40, 30, 6, 1), (300, 200, 12, 2), (1000, 1500, 20, 3)]):
        rp, ci, va = small_csr(R, C, L, seed)
        N = int(rp[-1])
        p = f"c{case}_"
        out[p + "rowPtr"], out[p + "colIdx"], out[p + "values"], out[p + "numCols"] = rp, ci, va, np.int32(C)
        P = max(1, (R + N) // 7)
        pos = oracle.merge_partitions(R, N, P)
        out[p + "merge_pos"] = pos
        out[p + "merge_row"], out[p + "merge_nnz"] = oracle.merge_path_coords(rp, N, pos)
        out[p + "csr5_row"], out[p + "csr5_off"] = oracle.csr5_tiles(rp, N, 16)
        w = oracle.ell_width(rp)
        out[p + "ell_col"], out[p + "ell_val"] = oracle.csr_to_ell(rp, ci, va, w)
        out[p + "com_fwd"], out[p + "com_cols"] = oracle.reorder_columns_com(rp, ci, C)
        out[p + "minrow_perm"], out[p + "minrow_cols"] = oracle.reorder_columns_minrow(rp, ci, C)
        out[p + "freq_perm"], out[p + "freq_cols"] = oracle.reorder_columns_freq(rp, ci, C)
        out[p + "rowlen_map"] = oracle.reorder_rows_by_length(rp, ci, va)[0]
        x = np.random.default_rng(seed + 100).uniform(-1, 1, C).astype(np.float32)
        out[p + "x"] = x
        out[p + "y_f32"] = oracle.spmv_csr_f32(rp, ci, va, x)
    path = Path(__file__).resolve().parent / "fixtures.npz"
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({path.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
