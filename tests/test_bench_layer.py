"""The benchmark layer's C ABI on a CPU host (include/muxi_spmv/muxi_spmv_bench.h): the result and config structs
laid out as the header compiles them, the reference's metric formulas (performance_benchmark.h:79-111), its
results file (performance_benchmark.cu:298-335), and the argument errors returned before any device work."""
import ctypes as C
import json
import shutil
import subprocess
from pathlib import Path

import pytest

from muxi_spmv_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def test_struct_layouts_match_the_compiled_header(tmp_path):
    """offsetof and sizeof of every field as gcc lays out muxi_spmv_bench.h, against the ctypes mirrors."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "muxi_spmv/muxi_spmv_bench.h"', 'int main(void) {']
    for st in ("muxi_perf_result_t", "muxi_bench_config_t"):
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for name, _ in getattr(_lib, st)._fields_:
            lines.append(f'printf("{st} {name} %zu\\n", offsetof({st}, {name}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        st, name, val = line.split()
        cls = getattr(_lib, st)
        got = C.sizeof(cls) if name == "size" else getattr(cls, name).offset
        assert got == int(val), (st, name, got, val)


def test_default_config_is_the_references():
    c = _lib.load().muxi_bench_default_config()
    assert (c.warmupIterations, c.measureIterations, c.checkCorrectness, c.syncAfterEach) == (3, 10, True, True)
    assert c.stream is None and c.floatType == _lib.SPMV_TYPE_FLOAT and not c.compareCusparse


def test_metric_formulas():
    lib = _lib.load()
    # 1M nonzeros, 100k x 100k, 0.01 ms: 2e6 flops -> 200 GFLOP/s; bytes 1M*(4+4) + 4*100001 + 4*(100k+100k)
    assert lib.muxi_calculate_gflops(1_000_000, 0.01) == pytest.approx(200.0, rel=1e-15)
    b = 1_000_000 * 8 + 4 * 100_001 + 4 * 200_000
    assert b == 9_200_004
    assert lib.muxi_calculate_bandwidth(1_000_000, 100_000, 100_000, 0.01, 4) == pytest.approx(b / 0.01 / 1e6,
                                                                                             rel=1e-15)
    # FP64: 12 bytes a nonzero; and a count past 2^31 bytes (the reference's int product would wrap) stays exact
    b64 = 300_000_000 * 12 + 4 * 2_000_001 + 8 * 4_000_000
    assert b64 > 2 ** 31
    assert lib.muxi_calculate_bandwidth(300_000_000, 2_000_000, 2_000_000, 1.0, 8) == pytest.approx(b64 / 1e6,
                                                                                                  rel=1e-15)


def _result(**kw):
    r = _lib.muxi_perf_result_t()
    for k, v in kw.items():
        setattr(r, k, v)
    return r


def test_results_file_has_the_references_layout(tmp_path):
    lib = _lib.load()
    rs = (_lib.muxi_perf_result_t * 2)()
    rs[0] = _result(numRows=100, numCols=200, nnz=1000, sparsity=0.05, deviceId=0, warpSize=64,
                    peakBandwidthGBs=8000.0, timeMs=0.123456, minTimeMs=0.1, maxTimeMs=0.2, avgTimeMs=0.15,
                    stdDevMs=0.01, gflops=16.2, bandwidthGBs=123.456, bandwidthUtilization=1.543,
                    correctnessPassed=True)
    rs[1] = _result(numRows=1, numCols=1, nnz=1, correctnessPassed=False)
    path = tmp_path / "r.json"
    assert lib.muxi_write_results_json(rs, 2, str(path).encode()) == 0
    text = path.read_text()
    doc = json.loads(text)
    assert list(doc) == ["results"] and len(doc["results"]) == 2
    r0 = doc["results"][0]
    assert list(r0) == ["matrix", "device", "performance", "correctnessPassed"]
    assert r0["matrix"] == {"rows": 100, "cols": 200, "nnz": 1000, "sparsity": 0.05}
    assert r0["device"] == {"id": 0, "warpSize": 64, "peakBandwidthGBs": 8000.0}
    assert list(r0["performance"]) == ["timeMs", "minTimeMs", "maxTimeMs", "avgTimeMs", "stdDevMs", "gflops",
                                       "bandwidthGBs", "bandwidthUtilization"]
    # the reference's number formats: %.4f times, %.2f rates, %.1f utilization, %.6f sparsity
    assert '"timeMs": 0.1235,' in text and '"bandwidthGBs": 123.46,' in text and '"bandwidthUtilization": 1.5' in text
    assert r0["correctnessPassed"] is True and doc["results"][1]["correctnessPassed"] is False
    # an empty list, and an unwritable path (the reference's INVALID_MATRIX)
    assert lib.muxi_write_results_json(rs, 0, str(tmp_path / "e.json").encode()) == 0
    assert json.loads((tmp_path / "e.json").read_text()) == {"results": []}
    assert lib.muxi_write_results_json(rs, 1, str(tmp_path / "no" / "dir.json").encode()) == -2


def test_measure_argument_errors_before_any_device_work():
    lib = _lib.load()
    cfg = lib.muxi_bench_default_config()
    res = _lib.muxi_perf_result_t()
    A = _lib.muxi_csr_t(numRows=4, numCols=4, nnz=0)
    dummy = C.c_void_p(1)
    assert lib.muxi_measure_csr_performance(None, dummy, dummy, C.byref(cfg), None, C.byref(res)) == -2
    assert lib.muxi_measure_csr_performance(C.byref(A), None, dummy, C.byref(cfg), None, C.byref(res)) == -3
    bad = lib.muxi_bench_default_config()
    bad.measureIterations = 0
    assert lib.muxi_measure_csr_performance(C.byref(A), dummy, dummy, C.byref(bad), None, C.byref(res)) == -3
    cus = lib.muxi_bench_default_config()
    cus.compareCusparse = True
    assert lib.muxi_measure_csr_performance(C.byref(A), dummy, dummy, C.byref(cus), None, C.byref(res)) == -7
    half = lib.muxi_bench_default_config()
    half.floatType = _lib.SPMV_TYPE_HALF
    assert lib.muxi_measure_csr_performance(C.byref(A), dummy, dummy, C.byref(half), None, C.byref(res)) == -8
    c5 = _lib.muxi_csr5_t()
    assert lib.muxi_measure_csr5_performance(C.byref(A), None, dummy, dummy, C.byref(cfg), None, C.byref(res),
                                             256) == -2
    assert lib.muxi_measure_csr5_performance(C.byref(A), C.byref(c5), dummy, None, C.byref(cfg), None,
                                             C.byref(res), 256) == -3


def test_rocsparse_comparison_argument_errors_before_any_device_work():
    """muxi_compare_with_rocsparse rejects bad arguments before it loads rocSPARSE or touches the device."""
    lib = _lib.load()
    cfg = lib.muxi_bench_default_config()
    dummy = C.c_void_p(1)
    A = _lib.muxi_csr_t(numRows=4, numCols=4, nnz=0)
    A.d_rowPtr = dummy
    t, g, a = C.c_double(), C.c_double(), C.c_double()

    def call(Am, x, y, c, alg, tp=C.byref(t)):
        return lib.muxi_compare_with_rocsparse(Am, x, y, c, alg, tp, C.byref(g), C.byref(a))

    assert call(None, dummy, dummy, C.byref(cfg), 0) == -2
    assert call(C.byref(_lib.muxi_csr_t(numRows=4, numCols=4, nnz=0)), dummy, dummy, C.byref(cfg), 0) == -2
    assert call(C.byref(A), None, dummy, C.byref(cfg), 0) == -3
    assert call(C.byref(A), dummy, dummy, C.byref(cfg), 0, tp=None) == -3
    bad = lib.muxi_bench_default_config()
    bad.measureIterations = 0
    assert call(C.byref(A), dummy, dummy, C.byref(bad), 0) == -3
    cus = lib.muxi_bench_default_config()
    cus.compareCusparse = True
    assert call(C.byref(A), dummy, dummy, C.byref(cus), 0) == -7
    half = lib.muxi_bench_default_config()
    half.floatType = _lib.SPMV_TYPE_HALF
    assert call(C.byref(A), dummy, dummy, C.byref(half), 0) == -8
    for alg in (-1, 4):
        assert call(C.byref(A), dummy, dummy, C.byref(cfg), alg) == -7
        assert b"alg must be" in lib.muxi_spmv_last_error()


def test_rocsparse_timing_rejects_a_malformed_matrix_before_any_device_work(tmp_path):
    """muxi_spmv_amd/rocsparse_timing checks the CSR arrays it is given on the host (row pointers from 0 to nnz,
    non-decreasing; columns inside the matrix) and exits with 2 before it allocates device memory."""
    import numpy as np

    exe = ROOT / "muxi_spmv_amd" / "rocsparse_timing"
    if not exe.exists():
        pytest.skip("rocsparse_timing is not built (python -m muxi_spmv_amd.build)")
    rp = np.array([0, 2, 3], np.int32)
    vals = np.ones(3, np.float32)
    x = np.ones(4, np.float32)
    for name, ci in (("column past the end", np.array([0, 4, 1], np.int32)),
                     ("negative column", np.array([0, -1, 1], np.int32))):
        rp.tofile(tmp_path / "rowPtr.i32")
        ci.tofile(tmp_path / "colIdx.i32")
        vals.tofile(tmp_path / "values.f32")
        x.tofile(tmp_path / "x.f32")
        r = subprocess.run([str(exe), str(tmp_path), "2", "4", "3", "3"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "not a CSR matrix" in r.stderr, (name, r.returncode, r.stderr)
    np.array([0, 2, 1, 3], np.int32).tofile(tmp_path / "rowPtr.i32")  # ends at nnz, decreasing in between
    np.array([0, 1, 1], np.int32).tofile(tmp_path / "colIdx.i32")
    r = subprocess.run([str(exe), str(tmp_path), "3", "4", "3", "3"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "not a CSR matrix" in r.stderr
    r = subprocess.run([str(exe), str(tmp_path / "missing"), "2", "4", "3", "3"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "cannot read" in r.stderr
