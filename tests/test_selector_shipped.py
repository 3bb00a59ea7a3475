"""The selector weights shipped in the package, scored on the held-out matrices of the sweep DB they were
trained beside (muxi_spmv_amd/selector_db.jsonl.gz: 2,000 training + 400 held-out matrices of 1e4 to 4e6 rows
and 300 + 60 of 4e6 to 2e7 rows, all in FP32 and again with FP64 values, x and y; every candidate timed cold on
an MI355X). Predictions come from the FP64 oracle network (oracle/mlp.py), so this runs on the CPU; the device
forward pass is checked against the same oracle in tests/test_gpu_tune.py."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import mlp as omlp

PKG = Path(__file__).resolve().parent.parent / "muxi_spmv_amd"
WEIGHTS = PKG / "selector_weights.json"
DB = PKG / "selector_db.jsonl.gz"


def _load():
    from muxi_spmv_amd import selector_train, tune
    st = json.loads(WEIGHTS.read_text())
    recs = selector_train.load_db(DB)
    return st, recs, tune


@pytest.mark.skipif(not (WEIGHTS.exists() and DB.exists()), reason="no shipped selector")
def test_shipped_selector_candidates_match_the_library():
    st, recs, tune = _load()
    assert st["candidates"] == tune.CAND_NAMES
    assert st["n_in"] == tune.N_INPUT and st["n_out"] == len(tune.CANDIDATES)
    assert sum(1 for i in recs if i < FP64) == 2400 + 360
    assert sum(1 for i in recs if i >= FP64) == 2400 + 360


FP64 = 100_000  # selector_train.FP64_BASE: index FP64 + i is matrix i with FP64 values, x and y
BIG = 50_000  # selector_train.BIG_BASE: 300 training then 60 held-out matrices of 4e6 to 2e7 rows


@pytest.mark.skipif(not (WEIGHTS.exists() and DB.exists()), reason="no shipped selector")
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_shipped_selector_regret_on_held_out_matrices(precision):
    st, recs, tune = _load()
    off = FP64 if precision == "fp64" else 0
    test = [recs[off + i] for i in list(range(2000, 2400)) + list(range(BIG + 300, BIG + 360))]
    train = [recs[off + i] for i in list(range(2000)) + list(range(BIG, BIG + 300))]
    fp64 = precision == "fp64"
    X = np.array([tune.feature_vector(np.array(r["features"]), fp64) for r in test], np.float64)
    Xn = (X - np.array(st["mu"])) / np.array(st["sd"])
    Y, _ = omlp.forward(np.array(st["params"], np.float64), Xn, st["n_in"], st["hidden"], st["n_out"])
    names = st["candidates"]

    def regret(r, name):
        best = min(r["times"].values())
        return r["times"][name] / best - 1.0 if name in r["times"] else math.inf

    learned = []
    for r, y in zip(test, Y):
        f = np.array(r["features"])
        y = np.where(tune.feasible_mask(f, fp64) > 0, y, np.inf)
        learned.append(regret(r, names[int(np.argmin(y))]))
    learned = np.array(learned)
    # the method without the learned part: the reference's fixed rule, and the best single configuration
    rule = np.array([regret(r, tune.reference_rule(np.array(r["features"]))) for r in test])
    tot = {n: sum(math.log(r["times"][n] / min(r["times"].values())) if n in r["times"] else 10.0 for r in train)
           for n in names}
    fixed_name = min(tot, key=tot.get)
    fixed = np.array([regret(r, fixed_name) for r in test])
    assert np.all(np.isfinite(learned))  # never picks a candidate the sweep would not run
    assert np.median(learned) <= 0.05
    assert np.percentile(learned, 95) <= 0.25
    assert np.median(learned) < np.median(fixed[np.isfinite(fixed)])
    assert np.median(learned) < np.median(rule[np.isfinite(rule)])
    assert np.mean(learned) < 0.5 * np.mean(fixed[np.isfinite(fixed)])


def test_retime_tag_resumes():
    """--retime times the named kernels again on every record, except records already tagged with the
    --retime-tag of the same re-timing (so a re-timing split over several calls resumes where it stopped);
    candidates a record lacks are always added."""
    from muxi_spmv_amd import selector_train as st
    from muxi_spmv_amd import tune

    f = [1e5, 1e6, 10.0, 4.0, 15.0, 0.01, 0.1, 0.0]
    cands = [("xslab", 1028), ("merge", 4), ("ell", 0)]
    times = {"xslab:1028": 0.1, "merge:4": 0.2, "ell:0": 0.3}
    rec = {"features": f, "times": dict(times)}
    assert st._todo(None, cands, ()) == cands
    assert st._todo(rec, cands, ()) == []
    assert st._todo(rec, cands, ("xslab",)) == [("xslab", 1028)]
    assert st._todo(rec, cands, ("xslab", "ell"), "t1") == [("xslab", 1028), ("ell", 0)]
    tagged = dict(rec, retimed=["t1"])
    assert st._todo(tagged, cands, ("xslab", "ell"), "t1") == []
    assert st._todo(tagged, cands, ("xslab",), "t2") == [("xslab", 1028)]
    missing = dict(tagged, times={"merge:4": 0.2})
    assert st._todo(missing, cands, ("xslab",), "t1") == [("xslab", 1028), ("ell", 0)]
    assert tune.feasible("ell", np.array(f))
