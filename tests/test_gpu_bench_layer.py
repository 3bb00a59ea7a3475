"""GPU: the benchmark layer (muxi_spmv_bench.h) measures the reference's entry points as
measureCSRPerformance / measureCSR5Performance do (performance_benchmark.cu:56-169, 348-479): statistics of the
event-timed iterations, the reference's GFLOP/s and bandwidth at the median, the CSR5 conversion, the finite-y
check, and y itself against the oracle; and a reference-style C++ caller of the C++ names passes."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from muxi_spmv_amd import api, benchmark
from gpu_util import dev_csr, fp32_bound, fp64_bound, matrix_zoo

pytestmark = pytest.mark.gpu

ZOO = matrix_zoo()


def _relations(r, h, fb):
    assert r["iterations"] == 10
    assert 0 < r["minTimeMs"] <= r["timeMs"] <= r["maxTimeMs"]
    assert r["minTimeMs"] <= r["avgTimeMs"] <= r["maxTimeMs"] and r["stdDevMs"] >= 0
    assert r["gflops"] == pytest.approx(2.0 * h.nnz / r["timeMs"] / 1e6, rel=1e-12)
    b = h.nnz * (fb + 4) + 4 * (h.numRows + 1) + fb * (h.numCols + h.numRows)
    assert r["bandwidthGBs"] == pytest.approx(b / r["timeMs"] / 1e6, rel=1e-12)
    assert r["peakBandwidthGBs"] > 0
    assert r["bandwidthUtilization"] == pytest.approx(100.0 * r["bandwidthGBs"] / r["peakBandwidthGBs"], rel=1e-12)
    assert (r["numRows"], r["numCols"], r["nnz"]) == (h.numRows, h.numCols, h.nnz)
    assert r["correctnessPassed"]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_measure_csr_performance(dtype):
    h = ZOO["powerlaw"]
    A = dev_csr(h, dtype)
    xh = np.random.default_rng(3).uniform(-1, 1, h.numCols)
    x = torch.from_numpy(xh).to(dtype).cuda()
    y = torch.full((h.numRows,), float("nan"), dtype=dtype, device="cuda")
    r = benchmark.measure_csr_performance(A, x, y)
    _relations(r, h, 8 if dtype == torch.float64 else 4)
    assert not r["useCSR5"] and r["warpSize"] == 64
    y64, ab = oracle.spmv_f64(h.rowPtr, h.colIdx, h.values.astype(np.float64), x.double().cpu().numpy())
    bound = (fp64_bound if dtype == torch.float64 else fp32_bound)(h.rowPtr, 1.0) * ab + 1e-300
    assert np.all(np.abs(y.double().cpu().numpy() - y64) <= bound)


def test_measure_csr5_performance():
    h = ZOO["uniform"]
    A = dev_csr(h)
    x = torch.from_numpy(np.random.default_rng(4).uniform(-1, 1, h.numCols).astype(np.float32)).cuda()
    y = torch.full((h.numRows,), float("nan"), device="cuda")
    r, m = benchmark.measure_csr5_performance(A, x, y, 256)
    _relations(r, h, 4)
    assert r["useCSR5"] and r["tileSize"] == 256 == m.sigma and r["numTiles"] == m.numTiles == -(-h.nnz // 256)
    assert r["conversionTimeMs"] > 0
    assert r["totalTimeMs"] == pytest.approx(r["timeMs"] + r["conversionTimeMs"], rel=1e-12)
    assert r["conversionOverheadPercent"] == pytest.approx(100.0 * r["conversionTimeMs"] / r["timeMs"], rel=1e-12)
    y64, ab = oracle.spmv_f64(h.rowPtr, h.colIdx, h.values.astype(np.float64), x.double().cpu().numpy())
    assert np.all(np.abs(y.double().cpu().numpy() - y64) <= fp32_bound(h.rowPtr, 1.0) * ab + 1e-37)


def test_non_finite_y_fails_the_check_and_off_passes():
    h = ZOO["uniform"]
    A = dev_csr(h)
    x = torch.full((h.numCols,), float("nan"), device="cuda")
    y = torch.empty(h.numRows, device="cuda")
    assert not benchmark.measure_csr_performance(A, x, y, iterations=3)["correctnessPassed"]
    assert benchmark.measure_csr_performance(A, x, y, iterations=3, check_correctness=False)["correctnessPassed"]


def test_cpp_benchmark_caller(tmp_path):
    """measureCSRPerformance<float/double>, measureCSR5Performance<float>, printPerformanceResult and
    writeResultsJSON from a reference-style C++ caller (tests/cpp/test_benchmark_layer.cpp)."""
    exe = Path(__file__).resolve().parent / "cpp" / "test_benchmark_layer"
    assert exe.exists(), "build with python -m muxi_spmv_amd.build"
    out = tmp_path / "results.json"
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout and "=== Performance Results ===" in r.stdout
    doc = json.loads(out.read_text())
    assert len(doc["results"]) == 3 and all(e["correctnessPassed"] for e in doc["results"])


@pytest.mark.parametrize("alg", ["adaptive", "rowsplit", "lrb", "nnzsplit"])
@pytest.mark.parametrize("mname", ["powerlaw", "uniform", "empty_rows", "long_row"])
def test_rocsparse_cross_check(mname, alg):
    """The vendor library's y (rocSPARSE, every CSR algorithm) within the FP32 bound of the FP64 oracle, as ours
    is: the comparison times a correct product. y starts as NaN: the library's atomic algorithms add into y, so
    the comparison must clear it first."""
    h = ZOO[mname]
    A = dev_csr(h)
    x = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, h.numCols).astype(np.float32)).cuda()
    y = torch.full((h.numRows,), float("nan"), device="cuda")
    try:
        r = benchmark.compare_with_rocsparse(A, x, y, alg, iterations=3)
    except api.SpmvError as e:
        if e.status == -7 and "rejects algorithm" in str(e):  # the rocSPARSE loaded with torch predates it
            pytest.skip(str(e))
        raise
    assert r["timeMs"] > 0 and r["gflops"] == pytest.approx(2.0 * h.nnz / r["timeMs"] / 1e6, rel=1e-12)
    assert r["analysisMs"] >= 0
    y64, ab = oracle.spmv_f64(h.rowPtr, h.colIdx, h.values.astype(np.float64), x.double().cpu().numpy())
    yh = y.double().cpu().numpy()
    err = np.abs(yh - y64)
    bound = fp32_bound(h.rowPtr, 1.0) * ab + 1e-37
    bad = ~(err <= bound)
    assert not bad.any(), (int(bad.sum()), int(np.isnan(yh).sum()), float(np.nanmax(err / bound)),
                           np.flatnonzero(bad)[:5].tolist())


def test_rocsparse_fp64():
    h = ZOO["powerlaw"]
    A = dev_csr(h, torch.float64)
    x = torch.from_numpy(np.random.default_rng(6).uniform(-1, 1, h.numCols)).cuda()
    y = torch.full((h.numRows,), float("nan"), dtype=torch.float64, device="cuda")
    benchmark.compare_with_rocsparse(A, x, y, "adaptive", iterations=3)
    y64, ab = oracle.spmv_f64(h.rowPtr, h.colIdx, h.values.astype(np.float64), x.cpu().numpy())
    assert np.all(np.abs(y.cpu().numpy() - y64) <= fp64_bound(h.rowPtr, 1.0) * ab + 1e-300)


def test_rocsparse_timing_program(tmp_path):
    """muxi_spmv_amd/rocsparse_timing (the installation's rocSPARSE, outside a PyTorch process): every CSR
    algorithm runs, its y within 1e-5 relative L2 of a host FP64 loop, one JSON line out."""
    exe = Path(__file__).resolve().parents[1] / "muxi_spmv_amd" / "rocsparse_timing"
    if not exe.exists():
        pytest.fail("muxi_spmv_amd/rocsparse_timing is not built (python -m muxi_spmv_amd.build)")
    h = ZOO["powerlaw"]
    h.rowPtr.astype(np.int32).tofile(tmp_path / "rowPtr.i32")
    h.colIdx.astype(np.int32).tofile(tmp_path / "colIdx.i32")
    h.values.astype(np.float32).tofile(tmp_path / "values.f32")
    np.random.default_rng(7).uniform(-1, 1, h.numCols).astype(np.float32).tofile(tmp_path / "x.f32")
    r = subprocess.run([str(exe), str(tmp_path), str(h.numRows), str(h.numCols), str(h.nnz), "5"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    doc = json.loads(r.stdout.strip().splitlines()[-1])
    assert set(doc["us"]) == {"adaptive", "rowsplit", "lrb", "nnzsplit"} and all(v > 0 for v in doc["us"].values())
    assert all(v <= 1e-5 for v in doc["rel_l2_vs_fp64"].values()) and "librocsparse" in doc["library"]
