"""The shipped schedule tables (muxi_spmv_amd/sched_tables.json) against the committed sweep DB
(muxi_spmv_amd/sched_db.json), on the CPU: the tables are what the fit gives for the DB's training workloads, the
library's compiled tables pick what the Python tables pick, and on the held-out workloads the tuned picks beat the
former fixed defaults (batched: 16 items x 8 workgroups per CU; SpMM: 8 items)."""
import ctypes as C
import json

import pytest

from muxi_spmv_amd import _lib
from muxi_spmv_amd import sched_tune as S


@pytest.fixture(scope="module")
def shipped():
    return json.loads(S.TABLES.read_text()), S.load_db()


def test_tables_are_the_fit_of_the_db(shipped):
    tables, recs = shipped
    refit = S.fit(recs)
    assert refit["batch"] == tables["batch"]
    assert refit["spmm"] == tables["spmm"]
    n_b = sum(1 for r in recs if r["kind"] == "batch")
    assert n_b == (len(S.BATCH_ROW_LEN_EDGES) + 1) * (len(S.BATCH_ROWS_EDGES) + 1) * S.BATCH_REPEATS + S.BATCH_MIXED
    assert {r["j"] for r in recs if r["kind"] == "spmm"} == set(range(len(S.SPMM_MATRICES)))


def test_held_out_regret_beats_fixed_defaults(shipped):
    tables, recs = shipped
    ev = S.evaluate(tables, recs)
    for key, st_ in tables["held_out_regret"].items():
        assert ev[key] == pytest.approx(st_), key
    bt, bf, st, sf = ev["batch_tuned"], ev["batch_fixed_16x8"], ev["spmm_tuned"], ev["spmm_fixed_8"]
    assert bt["n"] == 17 and st["n"] == 70
    assert bt["median"] <= 0.02 and bt["p95"] <= 0.15 and bt["median"] < bf["median"] and bt["p95"] < bf["p95"]
    assert st["median"] <= 0.02 and st["p95"] <= 0.10 and st["median"] < sf["median"] and st["p95"] < sf["p95"]


def test_library_tables_pick_what_the_python_tables_pick(shipped):
    tables, recs = shipped
    lib = _lib.load()
    for r in recs:
        if r["kind"] == "batch":
            i, w = C.c_int(), C.c_int()
            assert lib.muxi_batch_schedule_for(r["mean_row_len"], r["mean_rows"], C.byref(i), C.byref(w)) == 0
            assert (i.value, w.value) == S.batch_pick(tables, r["mean_row_len"], r["mean_rows"]), r["i"]
        else:
            path = r["rows"] + r["nnz"]
            assert lib.muxi_spmm_schedule_for(r["k"], float(path)) == S.spmm_pick(tables, r["k"], path)
    # k outside the tuned values: the largest tuned k not above it
    assert lib.muxi_spmm_schedule_for(48, 1e7) == S.spmm_pick(tables, 32, 1e7)
    assert lib.muxi_spmm_schedule_for(3, 1e4) == S.spmm_pick(tables, 2, 1e4)
