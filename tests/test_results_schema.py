"""The benchmark layer's results file (muxi_spmv_amd.benchmark): the reference's writeResultsJSON layout
(src/benchmark/performance_benchmark.cu:298-335) with its statistics (calculateStats, :171-185), written and
validated on the CPU with synthetic timings."""
import json
import math
import statistics

import pytest

from muxi_spmv_amd import benchmark as B
from muxi_spmv_amd import bench_suite


def _rec(**kw):
    args = dict(rows=1000, cols=2000, nnz=5000, times_ms=[0.5, 0.3, 0.4, 0.7, 0.2], bytes_=90_000)
    args.update(kw)
    return B.performance_result(args.pop("rows"), args.pop("cols"), args.pop("nnz"), args.pop("times_ms"),
                                args.pop("bytes_"), **args)


def test_statistics_follow_the_reference_definitions():
    ts = [0.5, 0.3, 0.4, 0.7, 0.2]
    st = B.time_stats(ts)
    assert st["timeMs"] == 0.4  # the median (sorted[n // 2])
    assert (st["minTimeMs"], st["maxTimeMs"]) == (0.2, 0.7)
    assert st["avgTimeMs"] == pytest.approx(0.42)
    assert st["stdDevMs"] == pytest.approx(math.sqrt(sum((t - 0.42) ** 2 for t in ts) / 5))  # population
    assert B.time_stats([0.3, 0.1])["timeMs"] == 0.3  # even count: the upper middle, as times[size / 2]
    with pytest.raises(ValueError):
        B.time_stats([])


def test_performance_fields_and_correctness_flag():
    r = _rec(rel_l2=3e-7, max_abs=1e-5, tolerance=1e-6, cpu={"timeMs": 2.5, "kind": "t", "threads": 1},
             workload="w", kernel="lslab:4")
    p = r["performance"]
    assert p["gflops"] == pytest.approx(2 * 5000 / 0.4 / 1e6)
    assert p["bandwidthGBs"] == pytest.approx(90_000 / 0.4 / 1e6)
    assert p["bandwidthUtilization"] == pytest.approx(100 * p["bandwidthGBs"] / B.HBM_PEAK_GBS)
    assert r["rooflineFrac"] == pytest.approx(p["bandwidthGBs"] / B.HBM_PEAK_GBS)
    assert r["matrix"] == {"rows": 1000, "cols": 2000, "nnz": 5000, "sparsity": 5000 / (1000 * 2000)}
    assert r["correctnessPassed"] is True and r["workload"] == "w" and r["kernel"] == "lslab:4"
    assert _rec(rel_l2=2e-6, tolerance=1e-6)["correctnessPassed"] is False
    assert _rec()["correctnessPassed"] is True  # no check made
    assert _rec(k=8)["performance"]["gflops"] == pytest.approx(8 * p["gflops"])


def test_written_file_round_trips_and_validates(tmp_path):
    recs = [_rec(rel_l2=1e-7, max_abs=1e-6, tolerance=1e-6), _rec(rows=7, cols=7, nnz=0, times_ms=[1.0])]
    path = tmp_path / "r.json"
    B.write_results_json(str(path), recs)
    doc = json.loads(path.read_text())
    assert list(doc) == ["results"] and len(doc["results"]) == 2
    B.validate_results(doc)
    e = doc["results"][0]
    for key in ("matrix", "device", "performance", "correctnessPassed"):  # the reference's four keys
        assert key in e
    for f in ("timeMs", "minTimeMs", "maxTimeMs", "avgTimeMs", "stdDevMs", "gflops", "bandwidthGBs",
              "bandwidthUtilization"):
        assert f in e["performance"]


@pytest.mark.parametrize("breaks", ["no_results", "missing_perf_field", "bool_as_number", "nan", "flag_type",
                                    "median_outside", "cpu_without_time"])
def test_validator_rejects(breaks):
    doc = {"results": [_rec(rel_l2=1e-7, tolerance=1e-6)]}
    r = doc["results"][0]
    if breaks == "no_results":
        doc = {"result": []}
    elif breaks == "missing_perf_field":
        del r["performance"]["stdDevMs"]
    elif breaks == "bool_as_number":
        r["matrix"]["rows"] = True
    elif breaks == "nan":
        r["performance"]["gflops"] = float("nan")
    elif breaks == "flag_type":
        r["correctnessPassed"] = 1
    elif breaks == "median_outside":
        r["performance"]["timeMs"] = 9.0
    elif breaks == "cpu_without_time":
        r["cpu"] = {"kind": "x"}
    with pytest.raises(ValueError):
        B.validate_results(doc)


def test_suite_results_file_name():
    assert bench_suite.results_path("precision") == "spmv_results_precision.json"
    assert bench_suite.results_path("spmm", "/tmp/x.json") == "/tmp/x.json"


def test_cpu_time_helper_runs_the_scipy_product():
    import numpy as np

    rp = np.array([0, 2, 2, 3], np.int32)
    ci = np.array([0, 2, 1], np.int32)
    va = np.array([1.0, 2.0, 3.0], np.float32)
    c = bench_suite.cpu_ms(rp, ci, va, np.ones(3, np.float32), reps=2)
    assert c["timeMs"] >= 0 and c["threads"] == 1
