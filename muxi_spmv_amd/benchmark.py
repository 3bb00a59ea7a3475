"""Benchmark layer: byte/flop model, cold-cache timing and the results JSON.

Counterpart of src/benchmark/performance_benchmark.{h,cu}:
  calculate_gflops      calculateGFlops (performance_benchmark.h:79-85): 2*N*k / t
  algorithmic_bytes     calculateBandwidth (:96-111), generalised to value
                        and vector element sizes and k right-hand sides:
                        N*(bv + 4) + 4*(R+1) + k*C*bx + k*R*by (+ k*R*by when beta != 0)
  measure_plan          measureCSRPerformance (performance_benchmark.cu:56-169):
                        warmups, then timed launches; here every launch reads a
                        different copy of the matrix so that no launch finds its
                        data in L2 or the 256 MiB Infinity Cache ("cold").
  time_stats            calculateStats (:171-185) + the median as timeMs
  performance_result    PerformanceResult (performance_benchmark.h:21-53) as one results entry
  write_results_json    writeResultsJSON (:298-335) layout, extended and validated
                        (validate_results)
"""
from __future__ import annotations

import ctypes as C
import json
import math
import statistics
import time
from dataclasses import asdict, dataclass, field

import torch

from . import _lib
from .api import Plan, _ptr, _stream
from ._lib import check, load

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
MALL_BYTES = 256 << 20


def calculate_gflops(nnz: int, time_ms: float, k: int = 1) -> float:
    return 2.0 * nnz * k / time_ms / 1e6


def algorithmic_bytes(R: int, C: int, N: int, bv: int, bx: int, by: int | None = None, k: int = 1,
                      beta_nonzero: bool = False) -> int:
    by = bx if by is None else by
    b = N * (bv + 4) + 4 * (R + 1) + k * C * bx + k * R * by
    if beta_nonzero:
        b += k * R * by
    return int(b)


def calculate_bandwidth(bytes_: int, time_ms: float) -> float:
    return bytes_ / time_ms / 1e6


class Events:
    """A pool of library-owned HIP events."""

    def __init__(self, n: int):
        self.ev = []
        for _ in range(n):
            e = C.c_void_p()
            check(load().muxi_event_create(C.byref(e)), "event_create")
            self.ev.append(e)

    def elapsed(self, i: int, j: int) -> float:
        ms = C.c_float()
        check(load().muxi_event_elapsed_ms(self.ev[i], self.ev[j], C.byref(ms)), "event_elapsed")
        return ms.value

    def __del__(self):
        try:
            for e in self.ev:
                load().muxi_event_destroy(e)
        except Exception:
            pass


@dataclass
class Measurement:
    ms_per_step: float          # wall clock over the timed launches / steps
    kernel_ms: list = field(default_factory=list)  # main kernel per launch (HIP events)
    exec_ms: list = field(default_factory=list)    # the whole launch, every kernel (HIP events)
    steps: int = 0

    @property
    def kernel_ms_mean(self) -> float:
        return statistics.fmean(self.kernel_ms) if self.kernel_ms else float("nan")

    @property
    def kernel_ms_median(self) -> float:
        return statistics.median(self.kernel_ms) if self.kernel_ms else float("nan")

    @property
    def exec_ms_mean(self) -> float:
        return statistics.fmean(self.exec_ms) if self.exec_ms else float("nan")

    @property
    def exec_ms_median(self) -> float:
        return statistics.median(self.exec_ms) if self.exec_ms else float("nan")


def copies_for_cold(bytes_per_copy: int, floor: int = 3) -> int:
    """Enough rotating copies that the others evict a copy from the MALL."""
    need = (2 * MALL_BYTES) // max(1, bytes_per_copy) + 2
    return int(max(floor, min(64, need)))


def measure_plans(plans: list[Plan], xs: list[torch.Tensor], ys: list[torch.Tensor], steps: int, warmup: int,
                  alpha: float = 1.0, beta: float = 0.0, nrhs: int = 1, record_kernel: bool = True,
                  event_every: int = 4) -> Measurement:
    """Launches plans[i % n] on xs/ys round-robin: `warmup` untimed, then
    `steps` timed between two device synchronisations. Every `event_every`-th
    timed launch (a sample spread evenly over the timed region and over the
    rotating copies) carries two HIP events: on the first, third, ... sample
    they bracket the whole launch (the main kernel and any combine or fix-up
    kernel after it, with no marker between them: exec_ms), on the others the
    main kernel alone (kernel_ms). The other launches carry no event markers,
    which on a short kernel would add their own cost to the wall clock."""
    n = len(plans)
    s = _stream(None)
    lib = load()
    for i in range(warmup):
        check(lib.muxi_plan_exec(plans[i % n]._p, _ptr(xs[i % n]), _ptr(ys[i % n]), alpha, beta, nrhs, s),
              "plan_exec")
    every = max(1, int(event_every))
    sampled = [i for i in range(steps) if i % every == every - 1] if record_kernel else []
    ev = Events(2 * len(sampled)) if sampled else None
    slot = {i: k for k, i in enumerate(sampled)}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        j = i % n
        k = slot.get(i)
        if k is None:
            check(lib.muxi_plan_exec(plans[j]._p, _ptr(xs[j]), _ptr(ys[j]), alpha, beta, nrhs, s), "plan_exec")
        elif k % 2 == 1:  # the main kernel
            check(lib.muxi_plan_exec_ev(plans[j]._p, _ptr(xs[j]), _ptr(ys[j]), alpha, beta, nrhs, s,
                                        ev.ev[2 * k], ev.ev[2 * k + 1]), "plan_exec_ev")
        else:  # the whole launch, with no marker between its kernels
            check(lib.muxi_plan_exec_ev(plans[j]._p, _ptr(xs[j]), _ptr(ys[j]), alpha, beta, nrhs, s,
                                        ev.ev[2 * k], None), "plan_exec_ev")
            check(lib.muxi_event_record(ev.ev[2 * k + 1], s), "event_record")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    m = Measurement(ms_per_step=(t1 - t0) * 1e3 / steps, steps=steps)
    if ev is not None:
        m.exec_ms = [ev.elapsed(2 * k, 2 * k + 1) for k in range(0, len(sampled), 2)]
        m.kernel_ms = [ev.elapsed(2 * k, 2 * k + 1) for k in range(1, len(sampled), 2)]
    return m


def time_stats(times_ms: list[float]) -> dict:
    """minTimeMs / maxTimeMs / avgTimeMs / stdDevMs as calculateStats (performance_benchmark.cu:171-185; population
    standard deviation) and timeMs = the median, as measureCSRPerformance reports it (:135)."""
    ts = sorted(float(t) for t in times_ms)
    if not ts:
        raise ValueError("no timings")
    avg = statistics.fmean(ts)
    return {"timeMs": ts[len(ts) // 2], "minTimeMs": ts[0], "maxTimeMs": ts[-1], "avgTimeMs": avg,
            "stdDevMs": statistics.pstdev(ts) if len(ts) > 1 else 0.0}


def performance_result(rows: int, cols: int, nnz: int, times_ms: list[float], bytes_: int, *, k: int = 1,
                       rel_l2: float | None = None, max_abs: float | None = None, tolerance: float | None = None,
                       device: dict | None = None, cpu: dict | None = None, **extra) -> dict:
    """One entry of the results file: the reference's PerformanceResult fields in its JSON layout
    (writeResultsJSON, performance_benchmark.cu:298-335) -- matrix, device, performance, correctnessPassed --
    extended with the error against the FP64 oracle, the fraction of the HBM roofline, the CPU time of the same
    product where one was measured, and the workload's own keys (`extra`). `bytes_` is the algorithmic byte count
    of one launch; correctnessPassed is rel_l2 <= tolerance (true when no check was made, as the reference does
    when checkCorrectness is off)."""
    st = time_stats(times_ms)
    ms = st["timeMs"]
    dev = device if device is not None else {"id": 0, "warpSize": 64, "peakBandwidthGBs": HBM_PEAK_GBS}
    gbs = calculate_bandwidth(bytes_, ms)
    perf = dict(st, gflops=calculate_gflops(nnz, ms, k), bandwidthGBs=gbs,
                bandwidthUtilization=100.0 * gbs / float(dev["peakBandwidthGBs"]), iterations=len(times_ms))
    passed = True if (rel_l2 is None or tolerance is None) else bool(rel_l2 <= tolerance)
    rec = {"matrix": {"rows": int(rows), "cols": int(cols), "nnz": int(nnz),
                      "sparsity": float(nnz) / (float(rows) * float(cols)) if rows and cols else 0.0},
           "device": dev, "performance": perf, "correctnessPassed": passed,
           "errors": {"relL2VsFp64": rel_l2, "maxAbsVsFp64": max_abs, "tolerance": tolerance},
           "rooflineFrac": gbs / HBM_PEAK_GBS, "cpu": cpu}
    rec.update(extra)
    return rec


_NUM = (int, float)
_SCHEMA = {
    "matrix": {"rows": int, "cols": int, "nnz": int, "sparsity": _NUM},
    "device": {"id": int, "warpSize": int, "peakBandwidthGBs": _NUM},
    "performance": {"timeMs": _NUM, "minTimeMs": _NUM, "maxTimeMs": _NUM, "avgTimeMs": _NUM, "stdDevMs": _NUM,
                    "gflops": _NUM, "bandwidthGBs": _NUM, "bandwidthUtilization": _NUM},
    "errors": {"relL2VsFp64": (*_NUM, type(None)), "maxAbsVsFp64": (*_NUM, type(None)),
               "tolerance": (*_NUM, type(None))},
}


def validate_results(doc: dict) -> None:
    """Raises ValueError naming the first entry and field that does not follow the results-file layout."""
    if not isinstance(doc, dict) or not isinstance(doc.get("results"), list):
        raise ValueError("results file: top level must be {\"results\": [...]}")
    for i, r in enumerate(doc["results"]):
        if not isinstance(r, dict):
            raise ValueError(f"results[{i}]: not an object")
        for sec, fields in _SCHEMA.items():
            if not isinstance(r.get(sec), dict):
                raise ValueError(f"results[{i}].{sec}: missing")
            for f, t in fields.items():
                v = r[sec].get(f, KeyError)
                if v is KeyError:
                    raise ValueError(f"results[{i}].{sec}.{f}: missing")
                if isinstance(v, bool) or not isinstance(v, t):
                    raise ValueError(f"results[{i}].{sec}.{f}: {type(v).__name__}")
                if isinstance(v, float) and not math.isfinite(v):
                    raise ValueError(f"results[{i}].{sec}.{f}: not finite")
        if not isinstance(r.get("correctnessPassed"), bool):
            raise ValueError(f"results[{i}].correctnessPassed: missing or not a bool")
        if not isinstance(r.get("rooflineFrac"), _NUM) or isinstance(r.get("rooflineFrac"), bool):
            raise ValueError(f"results[{i}].rooflineFrac: missing")
        cpu = r.get("cpu")
        if cpu is not None and not (isinstance(cpu, dict) and isinstance(cpu.get("timeMs"), _NUM)):
            raise ValueError(f"results[{i}].cpu: needs timeMs")
        p = r["performance"]
        if not p["minTimeMs"] <= p["timeMs"] <= p["maxTimeMs"]:
            raise ValueError(f"results[{i}].performance: timeMs outside [minTimeMs, maxTimeMs]")


def write_results_json(path: str, results: list[dict]) -> None:
    """{results:[{matrix{rows,cols,nnz,sparsity}, device{...}, performance{...}, correctnessPassed, ...}]},
    validated before it is written."""
    doc = {"results": results}
    validate_results(doc)
    with open(path, "w") as f:
        json.dump(doc, f, indent=2)


def _perf_dict(r) -> dict:
    return {name: getattr(r, name) for name, _ in r._fields_}


def _bench_config(warmup: int, iterations: int, check_correctness: bool, sync_after_each: bool, stream,
                  float_type: int):
    c = load().muxi_bench_default_config()
    c.warmupIterations, c.measureIterations = int(warmup), int(iterations)
    c.checkCorrectness, c.syncAfterEach = bool(check_correctness), bool(sync_after_each)
    c.stream = _stream(stream)
    c.floatType = float_type
    return c


def measure_csr_performance(A, x: torch.Tensor, y: torch.Tensor, *, warmup: int = 3, iterations: int = 10,
                            check_correctness: bool = True, sync_after_each: bool = True, stream=None) -> dict:
    """measureCSRPerformance<T> (performance_benchmark.cu:56-169) through muxi_measure_csr_performance: the
    dispatching SpMV (spmv_csr) timed with HIP events, warmup then `iterations` measured calls; returns the
    PerformanceResult fields (timeMs is the median) as a dict. A is an api.CSRMatrix of float or double values."""
    from .api import SpmvError, _check_xy
    if A.dtype not in (torch.float32, torch.float64):
        raise SpmvError(-8, "measure_csr_performance", "float or double values, as the reference's template")
    _check_xy(x, y, A.dtype, A.numCols, A.numRows, "measure_csr_performance")
    ft = _lib.SPMV_TYPE_DOUBLE if A.dtype == torch.float64 else _lib.SPMV_TYPE_FLOAT
    cfg = _bench_config(warmup, iterations, check_correctness, sync_after_each, stream, ft)
    r = _lib.muxi_perf_result_t()
    check(load().muxi_measure_csr_performance(C.byref(A.c), _ptr(x), _ptr(y), C.byref(cfg), None, C.byref(r)),
          "measure_csr_performance")
    return _perf_dict(r)


def measure_csr5_performance(A, x: torch.Tensor, y: torch.Tensor, sigma: int = 0, *, warmup: int = 3,
                             iterations: int = 10, check_correctness: bool = True, sync_after_each: bool = True,
                             stream=None) -> tuple[dict, object]:
    """measureCSR5Performance<T> (performance_benchmark.cu:348-479): the CSR -> CSR5 conversion timed with HIP
    events, then the CSR5 SpMV as measure_csr_performance times spmv_csr; returns the PerformanceResult fields
    and the api.CSR5Matrix the conversion built."""
    from .api import CSR5Matrix, SpmvError, _check_xy
    if A.dtype not in (torch.float32, torch.float64):
        raise SpmvError(-8, "measure_csr5_performance", "float or double values, as the reference's template")
    _check_xy(x, y, A.dtype, A.numCols, A.numRows, "measure_csr5_performance")
    ft = _lib.SPMV_TYPE_DOUBLE if A.dtype == torch.float64 else _lib.SPMV_TYPE_FLOAT
    cfg = _bench_config(warmup, iterations, check_correctness, sync_after_each, stream, ft)
    m = CSR5Matrix(A)
    r = _lib.muxi_perf_result_t()
    check(load().muxi_measure_csr5_performance(C.byref(A.c), C.byref(m.c), _ptr(x), _ptr(y), C.byref(cfg), None,
                                               C.byref(r), int(sigma)), "measure_csr5_performance")
    return _perf_dict(r), m


ROCSPARSE_ALGS = {"adaptive": 0, "rowsplit": 1, "lrb": 2, "nnzsplit": 3}


def compare_with_rocsparse(A, x: torch.Tensor, y: torch.Tensor, alg: str = "adaptive", *, warmup: int = 3,
                           iterations: int = 10, sync_after_each: bool = True, stream=None) -> dict:
    """The vendor library's CSR SpMV (rocSPARSE rocsparse_spmv by stages, algorithm `alg`) on the same matrix
    and vectors, timed as measure_csr_performance times ours (muxi_compare_with_rocsparse, the counterpart of the
    reference's compiled-out compareWithCusparse, performance_benchmark.cu:173-270): {"timeMs", "gflops",
    "analysisMs"}; y is cleared first and holds its result. SpmvError -7 when rocSPARSE cannot be loaded."""
    from .api import SpmvError, _check_xy
    if A.dtype not in (torch.float32, torch.float64):
        raise SpmvError(-8, "compare_with_rocsparse", "float or double values")
    _check_xy(x, y, A.dtype, A.numCols, A.numRows, "compare_with_rocsparse")
    ft = _lib.SPMV_TYPE_DOUBLE if A.dtype == torch.float64 else _lib.SPMV_TYPE_FLOAT
    cfg = _bench_config(warmup, iterations, False, sync_after_each, stream, ft)
    t, g, a = C.c_double(), C.c_double(), C.c_double()
    check(load().muxi_compare_with_rocsparse(C.byref(A.c), _ptr(x), _ptr(y), C.byref(cfg), ROCSPARSE_ALGS[alg],
                                             C.byref(t), C.byref(g), C.byref(a)), "compare_with_rocsparse")
    return {"timeMs": t.value, "gflops": g.value, "analysisMs": a.value}


def device_record() -> dict:
    p = torch.cuda.get_device_properties(0)
    return {"id": 0, "name": p.name, "warpSize": 64, "computeUnits": p.multi_processor_count,
            "peakBandwidthGBs": HBM_PEAK_GBS}
