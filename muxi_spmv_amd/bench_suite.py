"""Secondary benchmark workloads (one JSON line each on stdout).

python -m muxi_spmv_amd.bench_suite precision|batched|spmm|laplacian|shard|host|kernels [--out FILE] [--results FILE]

precision  values stored FP64 / FP32 / BF16, x FP32, FP32 accumulation plain
           and compensated (Dot2), on the 2M power-law matrix and on a
           4M-row banded matrix (half-bandwidth 64, 16 nnz/row, N(0,1)) after
           RCM; rel-L2 and max-abs error of y against the FP64 oracle (the
           FP64 values and x), plus time and GB/s for each; and checked
           plans (FP32 values, rtol 1e-3, plain and Dot2): time, rows
           recomputed in FP64 and the largest relative row error.
batched    256 CSR matrices, rows log-uniform in [1e3, 2e5], mean nnz/row
           log-uniform in [4, 64], Poisson row lengths: the default batch
           plan (the slab mode here: the block-diagonal matrix through an
           LDS column-slab plan; eager and hipGraph), the merge mode at every
           schedule and at the tuned one, against one CSR-vector launch per
           matrix (eager and hipGraph).
spmm       Y = A X on the power-law matrix for k = 8 and 32 (row-major X).
laplacian  4096^2 5-point Laplacian in FP64, CSR kernels, ELLPACK and the AUTO pick.
shard      one GPU's share of an 8-GPU row split (muxi_shard_rows: equal merge-path
           diagonals moved to row starts, x replicated) of the power-law and
           Laplacian matrices: the heaviest shard run with the AUTO plan, next
           to the whole matrix on one GPU.
kernels    every configuration of the autotune sweep (and the scalar and vector kernels) on one matrix
           (--matrix powerlaw: 2M power-law FP32; laplacian: 4096^2 Laplacian FP64), --reps cold launches
           each: the process to run under rocprofv3 for every kernel's counters side by side.
host       x and y in host memory (muxi_spmv_host.h): pinned and pageable buffers,
           1 / 2 / 4 / 8 row chunks overlapping PCIe copies with the kernels, on
           the power-law matrix and the FP64 Laplacian; end-to-end time from the
           host plan's own events (median of 10) next to the device-resident plan.
All timings except host's are cold: each timed launch is preceded by a 512 MiB cache flush,
and HIP events bracket the launch.

Every run also writes a results file (--results, default spmv_results_<workload>.json) in the reference's
results layout (writeResultsJSON, src/benchmark/performance_benchmark.cu:298-335): one entry per measured
configuration with matrix, device, performance {timeMs (median), min/max/avg/stdDev, gflops, bandwidthGBs,
bandwidthUtilization} and correctnessPassed (rel L2 against the FP64 oracle within the stated tolerance),
extended with the errors, the HBM roofline fraction and a CPU time for the same product (scipy.sparse CSR, one
thread) -- see benchmark.performance_result.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import statistics
import sys
import time

import numpy as np
import torch

from . import api, host, prep, tune
from . import generators as G
from .benchmark import HBM_PEAK_GBS, algorithmic_bytes, device_record, performance_result, write_results_json
from ._lib import check, load


def log(msg):
    print(f"[suite] {msg}", file=sys.stderr, flush=True)


class Timer:
    def __init__(self):
        self.ev = [C.c_void_p(), C.c_void_p()]
        for e in self.ev:
            check(load().muxi_event_create(C.byref(e)), "event_create")
        self.fl = tune.Flusher()

    def time(self, fn, reps=10, flush=True) -> float:
        """Median ms of fn() over reps, each preceded by a cache flush; every rep's ms stays in self.last."""
        s = torch.cuda.current_stream().cuda_stream or None
        lib = load()
        fn()
        ts = []
        for _ in range(reps):
            if flush:
                self.fl()
            lib.muxi_event_record(self.ev[0], s)
            fn()
            lib.muxi_event_record(self.ev[1], s)
            ms = C.c_float()
            check(lib.muxi_event_elapsed_ms(self.ev[0], self.ev[1], C.byref(ms)), "elapsed")
            ts.append(ms.value)
        self.last = ts
        return statistics.median(ts)


def errors(y: torch.Tensor, y64: torch.Tensor) -> dict:
    d = y.double() - y64
    nrm = float(torch.linalg.vector_norm(y64))
    return {"rel_l2": float(torch.linalg.vector_norm(d)) / nrm if nrm > 0 else float(torch.linalg.vector_norm(d)),
            "max_abs": float(d.abs().max()) if d.numel() else 0.0,
            "max_abs_y": float(y64.abs().max()) if d.numel() else 0.0}


TOL_FP32 = 1e-6   # rel L2 of an FP32-accumulated y against the FP64 oracle on the same stored values
TOL_FP64 = 1e-12  # FP64 accumulation


def cpu_ms(rp, ci, va, X, reps=3) -> dict:
    """The same product on the host: scipy.sparse CSR times x (or X), one thread, median of reps."""
    import scipy.sparse as sp

    rp = np.asarray(rp)
    A = sp.csr_matrix((np.asarray(va), np.asarray(ci), rp), shape=(rp.size - 1, int(np.asarray(X).shape[0])))
    X = np.asarray(X)
    A @ X
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        A @ X
        ts.append((time.perf_counter() - t) * 1e3)
    return {"timeMs": statistics.median(ts), "kind": "scipy.sparse CSR product", "threads": 1}


def _dims(A) -> tuple[int, int, int]:
    return int(A.numRows), int(A.numCols), int(A.nnz)


def precision(args) -> list[dict]:
    tm = Timer()
    out = []
    mats = {}
    h = G.cfg_powerlaw_2m()
    mats["powerlaw_2m"] = (h.rowPtr, h.colIdx, h.values.astype(np.float64), h.numCols)
    log("banded 4M: generating")
    B = tune.gen_matrix("banded", 4_000_000, 4_000_000, 64, 16, 0, normal=True, dtype=torch.float64, seed=4)
    rp, ci, va = B.host()
    B.free()
    t = time.perf_counter()
    perm = prep.rcm(rp, ci)
    nrp, nci, nva = prep.permute_symmetric(rp, ci, va, perm)
    log(f"banded 4M: RCM + permutation {time.perf_counter() - t:.1f}s, bandwidth {prep.bandwidth(rp, ci)} -> "
        f"{prep.bandwidth(nrp, nci)}")
    mats["banded_4m_rcm"] = (nrp, nci, nva, 4_000_000)
    dev = device_record()
    for name, (rp, ci, va64, ncol) in mats.items():
        xh = G.uniform_vector(ncol, 99)
        x = torch.from_numpy(xh).cuda()
        A64 = api.CSRMatrix(rp, ci, torch.from_numpy(va64), ncol)
        y64 = api.oracle_f64(A64, x)  # FP64 values, FP32 x, exact-ish FP64 sum
        cpu = cpu_ms(rp, ci, va64.astype(np.float32), xh)
        for vt in (torch.float64, torch.float32, torch.bfloat16):
            Av = A64.with_values(api.convert_values(A64.d_values, vt))
            y64s = api.oracle_f64(Av, x)  # the oracle on the values as stored
            for kernel, param in (("merge", 4), ("xslab", 0), ("lslab", 4), ("ell", 0), ("auto", 0)):
                if kernel == "ell" and name != "banded_4m_rcm":
                    continue  # ELL pads every row of the power-law matrix to its longest
                for comp in (False, True):
                    if kernel == "xslab" and comp:
                        continue  # no Dot2 variant
                    p = api.Plan(Av, kernel, param, compensated=comp)
                    info = p.info()
                    y = torch.empty(Av.numRows, device="cuda")
                    ms = tm.time(lambda: p.exec(x, y))
                    bv = {torch.float64: 8, torch.float32: 4, torch.bfloat16: 2}[vt]
                    b = algorithmic_bytes(Av.numRows, ncol, Av.nnz, bv, 4)
                    e, es = errors(y, y64), errors(y, y64s)
                    rec = {"workload": name, "values": str(vt).split(".")[-1], "kernel": f"{kernel}:{param}",
                           "runs": f"{info.kernel}:{info.param0}", "accumulation": "fp32_dot2" if comp else "fp32",
                           "ms": round(ms, 5), "GBps": round(b / ms / 1e6, 1), **e,
                           "rel_l2_vs_stored_values": es["rel_l2"]}
                    log(json.dumps(rec))
                    out.append(rec)
                    args.results.append(performance_result(
                        *_dims(Av), tm.last, b, rel_l2=es["rel_l2"], max_abs=es["max_abs"], tolerance=TOL_FP32,
                        device=dev, cpu=cpu, workload=name, values=rec["values"], kernel=rec["runs"],
                        accumulation=rec["accumulation"], relL2VsFp64Values=e["rel_l2"]))
                    p.destroy()
            if vt == torch.float32:  # checked plans: rows whose bound fails at rtol recomputed in FP64
                for comp in (False, True):
                    rtol = 1e-3
                    p = api.Plan.checked(Av, rtol, compensated=comp)
                    y = torch.empty(Av.numRows, device="cuda")
                    ms = tm.time(lambda: p.exec(x, y))
                    nre = int(p.checked_rows().size)
                    info = p.info()
                    b = algorithmic_bytes(Av.numRows, ncol, Av.nnz, 4, 4)
                    e, es = errors(y, y64), errors(y, y64s)
                    rel_row = float(((y.double() - y64s).abs() / y64s.abs().clamp_min(1e-300)).max())
                    rec = {"workload": name, "values": "float32", "kernel": f"checked:{rtol:g}",
                           "runs": f"{info.kernel}:{info.param0}",
                           "accumulation": "fp32_dot2" if comp else "fp32", "ms": round(ms, 5),
                           "GBps": round(b / ms / 1e6, 1), **e, "rel_l2_vs_stored_values": es["rel_l2"],
                           "rows_recomputed_fp64": nre, "max_row_rel_err": rel_row}
                    log(json.dumps(rec))
                    out.append(rec)
                    args.results.append(performance_result(
                        *_dims(Av), tm.last, b, rel_l2=es["rel_l2"], max_abs=es["max_abs"], tolerance=TOL_FP32,
                        device=dev, cpu=cpu, workload=name, values="float32", kernel=f"checked:{rtol:g}",
                        accumulation=rec["accumulation"], rowsRecomputedFp64=nre, maxRowRelErr=rel_row))
                    p.destroy()
    return out


def batched(args) -> list[dict]:
    from .batch import BatchPlan

    rng = np.random.default_rng(5)
    mats, xs, ys = [], [], []
    tot_b = 0
    for i in range(256):
        R = int(math.exp(rng.uniform(math.log(1e3), math.log(2e5))))
        mean = math.exp(rng.uniform(math.log(4), math.log(64)))
        A = tune.gen_matrix("poisson", R, R, mean, seed=5000 + i)
        mats.append(A)
        xs.append(tune.gen_vector(R, seed=7000 + i))
        ys.append(torch.empty(R, device="cuda"))
        tot_b += algorithmic_bytes(R, R, A.nnz, 4, 4)
    nnz = sum(m.nnz for m in mats)
    log(f"256 matrices, {sum(m.numRows for m in mats)} rows, {nnz} nnz, {tot_b / 1e6:.0f} MB")
    tm = Timer()
    s = torch.cuda.Stream()
    out = []
    from .sched_tune import BATCH_CANDIDATES

    ipt_ms = {}
    for ipt, wg in BATCH_CANDIDATES:  # items per lane of the merge tiles x workgroups per CU
        bp = BatchPlan(mats, xs, ys, ipt | (wg << 8))
        with torch.cuda.stream(s):
            ipt_ms[f"{ipt}x{wg}"] = tm.time(lambda: bp.exec())
        torch.cuda.synchronize()
        log(f"batched persistent, {ipt} items per lane x {wg} workgroups per CU: {ipt_ms[f'{ipt}x{wg}']:.4f} ms")
        bp.destroy()
    best_ipt = min(ipt_ms, key=ipt_ms.get)
    from .batch import TUNED_MERGE

    times = {}
    bm = BatchPlan(mats, xs, ys, TUNED_MERGE)  # the merge mode with the autotuner's schedule
    tuned = f"{bm.info()['ipt']}x{bm.info()['wg_per_cu']}"
    with torch.cuda.stream(s):
        t_merge = tm.time(lambda: bm.exec())
    times["batched_merge_tuned"] = tm.last
    torch.cuda.synchronize()
    bm.exec()
    torch.cuda.synchronize()
    worst_m, worst_m_abs = 0.0, 0.0
    for A, x, y in zip(mats, xs, ys):
        e = errors(y, api.oracle_f64(A, x))
        worst_m, worst_m_abs = max(worst_m, e["rel_l2"]), max(worst_m_abs, e["max_abs"])
    bm.destroy()
    bp = BatchPlan(mats, xs, ys)  # the default: the slab mode for this batch
    log(f"default batch plan: {bp.info()}")

    def worst_err():
        w, m = 0.0, 0.0
        for A, x, y in zip(mats, xs, ys):
            e = errors(y, api.oracle_f64(A, x))
            w, m = max(w, e["rel_l2"]), max(m, e["max_abs"])
        return w, m

    with torch.cuda.stream(s):
        g_batch = bp.graph(stream=s)
        g_vec = bp.graph("vector", 0, stream=s)
        t_batch = tm.time(lambda: bp.exec())
        times["batched_persistent"] = tm.last
        t_batch_g = tm.time(lambda: g_batch.launch(s))
        times["batched_persistent_graph"] = tm.last
        # correctness of the batched result
        bp.exec()
        worst, worst_abs = worst_err()
        t_vec = tm.time(lambda: bp.exec_per_matrix("vector", 0))
        times["per_matrix_vector"] = tm.last
        t_vec_g = tm.time(lambda: g_vec.launch(s))
        times["per_matrix_vector_graph"] = tm.last
        bp.exec_per_matrix("vector", 0)
        worst_v, worst_v_abs = worst_err()
    torch.cuda.synchronize()
    cpu_t = 0.0
    for A, x in zip(mats, xs):
        rp, ci, va = A.host()
        cpu_t += cpu_ms(rp, ci, va, x.cpu().numpy(), reps=1)["timeMs"]
    cpu = {"timeMs": cpu_t, "kind": "scipy.sparse CSR product per matrix, summed", "threads": 1}
    dev = device_record()
    R_all, C_all = sum(m.numRows for m in mats), sum(m.numCols for m in mats)
    for name, ms in [("batched_persistent", t_batch), ("batched_persistent_graph", t_batch_g),
                     ("batched_merge_tuned", t_merge), ("per_matrix_vector", t_vec),
                     ("per_matrix_vector_graph", t_vec_g)]:
        out.append({"workload": "batched_256", "path": name, "ms": round(ms, 4), "GBps": round(tot_b / ms / 1e6, 1),
                    "nnz": nnz})
        w, wa = ((worst, worst_abs) if name.startswith("batched_persistent") else
                 (worst_m, worst_m_abs) if name == "batched_merge_tuned" else (worst_v, worst_v_abs))
        args.results.append(performance_result(R_all, C_all, nnz, times[name], tot_b, rel_l2=w, max_abs=wa,
                                               tolerance=TOL_FP32, device=dev, cpu=cpu, workload="batched_256",
                                               path=name, matrices=len(mats), errors_are="worst member"))
    out.append({"workload": "batched_256", "speedup_graph_vs_per_matrix_graph": round(t_vec_g / t_batch_g, 2),
                "speedup_eager": round(t_vec / t_batch, 2), "speedup_vs_merge_tuned": round(t_merge / t_batch, 2),
                "worst_rel_l2": worst, "info": bp.info(),
                "merge_schedule": tuned, "merge_sweep_best_schedule": best_ipt,
                "merge_schedule_regret": round(ipt_ms[tuned] / ipt_ms[best_ipt] - 1, 4),
                "merge_ms_by_schedule": {k: round(v, 4) for k, v in ipt_ms.items()}})
    for r in out:
        log(json.dumps(r))
    return out


def spmm(args) -> list[dict]:
    h = G.cfg_powerlaw_2m()
    A = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values), h.numCols)
    tm = Timer()
    out = []
    dev = device_record()
    for k in (8, 32):
        X = torch.rand((h.numCols, k), device="cuda") * 2 - 1
        Y = torch.empty((h.numRows, k), device="cuda")
        y64 = api.oracle_f64(A, X[:, 3].contiguous())
        cpu = cpu_ms(h.rowPtr, h.colIdx, h.values, X.cpu().numpy(), reps=1)
        best = None
        recs_k = []
        tuned = api.Plan.spmm(A, k).info().param0  # the autotuner's table
        for ipt in (4, 8, 16, 32):
            p = api.Plan.spmm(A, k, ipt)
            ms = tm.time(lambda: p.exec(X, Y))
            b = algorithmic_bytes(h.numRows, h.numCols, h.nnz, 4, 4, k=k)
            rec = {"workload": "spmm_powerlaw_2m", "k": k, "ipt": ipt, "ms": round(ms, 4),
                   "GBps": round(b / ms / 1e6, 1), "GFLOPs": round(2 * h.nnz * k / ms / 1e6, 1)}
            recs_k.append(rec)
            log(json.dumps(rec))
            e = errors(Y[:, 3], y64)
            args.results.append(performance_result(h.numRows, h.numCols, h.nnz, tm.last, b, k=k, rel_l2=e["rel_l2"],
                                                   max_abs=e["max_abs"], tolerance=TOL_FP32, device=dev, cpu=cpu,
                                                   workload="spmm_powerlaw_2m", kernel=f"spmm:{ipt}", rhs=k,
                                                   errors_are="column 3"))
            if best is None or ms < best["ms"]:
                best = rec
        best["tuned_ipt"] = tuned
        best["tuned_regret"] = round(next(r["ms"] for r in recs_k if r["ipt"] == tuned) / best["ms"] - 1, 4)
        p = api.Plan.spmm(A, k, best["ipt"])
        p.exec(X, Y)
        y64 = api.oracle_f64(A, X[:, 3].contiguous())
        best["rel_l2_col3"] = errors(Y[:, 3], y64)["rel_l2"]
        out.append(best)
    return out


def laplacian(args) -> list[dict]:
    h = G.cfg_laplacian_4096()
    A = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values), h.numCols)
    x = torch.from_numpy(G.uniform_vector(h.numCols, 3, np.float64)).cuda()
    y = torch.empty(h.numRows, dtype=torch.float64, device="cuda")
    tm = Timer()
    b = algorithmic_bytes(h.numRows, h.numCols, h.nnz, 8, 8)
    out = []
    dev = device_record()
    y64 = api.oracle_f64(A, x)
    cpu = cpu_ms(h.rowPtr, h.colIdx, h.values, x.cpu().numpy(), reps=1)
    for k, p in [("ell", 0), ("adaptive", 512), ("adaptive", 1024), ("vwave", 2), ("vwave", 4), ("csr5", 512),
                 ("merge", 4), ("xslab", 0), ("auto", 0)]:
        pl = api.Plan(A, k, p, xtype=torch.float64)
        ms = tm.time(lambda: pl.exec(x, y))
        info = pl.info()
        rec = {"workload": "laplacian_4096_fp64", "kernel": f"{k}:{p}", "runs": f"{info.kernel}:{info.param0}",
               "ms": round(ms, 4), "GBps": round(b / ms / 1e6, 1), "frac_hbm": round(b / ms / 1e6 / HBM_PEAK_GBS, 4)}
        log(json.dumps(rec))
        out.append(rec)
        e = errors(y, y64)
        args.results.append(performance_result(h.numRows, h.numCols, h.nnz, tm.last, b, rel_l2=e["rel_l2"],
                                               max_abs=e["max_abs"], tolerance=TOL_FP64, device=dev, cpu=cpu,
                                               workload="laplacian_4096_fp64", kernel=rec["runs"]))
        pl.destroy()
    return out


def kernels(args) -> list[dict]:
    """Every configuration of the autotune sweep, each its own launches, on one matrix: the 2M power-law matrix in
    FP32 (--matrix powerlaw) or the 4096^2 Laplacian in FP64 (--matrix laplacian). Each configuration runs cold
    (cache flush before each of --reps launches), so that one rocprofv3 counter pass over this process gives every
    kernel's counters side by side, told apart by kernel name. The reference's one-lane and one-wavefront-per-row
    kernels (scalar, vector) run too, where the sweep leaves them out."""
    fp64 = args.matrix == "laplacian"
    h = G.cfg_laplacian_4096() if fp64 else G.cfg_powerlaw_2m()
    dt, esz = (torch.float64, 8) if fp64 else (torch.float32, 4)
    wname = "laplacian_4096_fp64" if fp64 else "powerlaw_2m_fp32"
    A = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values).to(dt), h.numCols)
    x = torch.from_numpy(G.uniform_vector(h.numCols, 3, np.float64 if fp64 else np.float32)).cuda()
    y = torch.empty(h.numRows, dtype=dt, device="cuda")
    f = tune.features(A)
    b = algorithmic_bytes(h.numRows, h.numCols, h.nnz, esz, esz)
    tm = Timer()
    y64 = api.oracle_f64(A, x)
    dev = device_record()
    cpu = cpu_ms(h.rowPtr, h.colIdx, h.values, x.cpu().numpy(), reps=1)
    tol = TOL_FP64 if fp64 else TOL_FP32
    out = []
    for k, p in tune.CANDIDATES:
        if not (k in ("scalar", "vector") or tune.feasible(k, f, fp64)):
            continue
        pl = api.Plan(A, k, p, xtype=dt)
        ms = tm.time(lambda: pl.exec(x, y), reps=args.reps)
        info = pl.info()
        e = errors(y, y64)
        rec = {"workload": wname, "kernel": f"{k}:{p}", "runs": f"{info.kernel}:{info.param0}", "ms": round(ms, 4),
               "GBps": round(b / ms / 1e6, 1), "frac_hbm": round(b / ms / 1e6 / HBM_PEAK_GBS, 4),
               "rel_l2": e["rel_l2"]}
        log(json.dumps(rec))
        out.append(rec)
        args.results.append(performance_result(h.numRows, h.numCols, h.nnz, tm.last, b, rel_l2=e["rel_l2"],
                                               max_abs=e["max_abs"], tolerance=tol, device=dev, cpu=cpu,
                                               workload=wname, kernel=rec["runs"]))
        pl.destroy()
    return out


def shard(args) -> list[dict]:
    G_ = 8
    tm = Timer()
    out = []
    for name, h, dt in (("powerlaw_2m_fp32", G.cfg_powerlaw_2m(), torch.float32),
                        ("laplacian_4096_fp64", G.cfg_laplacian_4096(), torch.float64)):
        vdt = np.float64 if dt == torch.float64 else np.float32
        esz = 8 if dt == torch.float64 else 4
        b = prep.shard_rows(h.rowPtr, G_)
        items = np.diff(b) + np.diff(h.rowPtr[b].astype(np.int64))
        g = int(np.argmax(items))  # the shard that sets an 8-GPU step's time
        r0, r1 = int(b[g]), int(b[g + 1])
        n0, n1 = int(h.rowPtr[r0]), int(h.rowPtr[r1])
        rp = (h.rowPtr[r0:r1 + 1] - n0).astype(np.int32)
        x = torch.from_numpy(G.uniform_vector(h.numCols, 7, vdt)).cuda()
        res = {"workload": name, "gpus": G_, "shard_index": g, "shard_rows": r1 - r0, "shard_nnz": n1 - n0,
               "items_max_over_mean": round(float(items.max() / items.mean()), 4)}
        for part, (A, R, rb) in (("shard", (api.CSRMatrix(rp, h.colIdx[n0:n1],
                                                          torch.from_numpy(h.values[n0:n1]).to(dt), h.numCols),
                                            r1 - r0, r0)),
                                 ("whole", (api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values).to(dt),
                                                          h.numCols), h.numRows, 0))):
            p = api.Plan(A, "auto", 0, xtype=dt, row_base=rb, rows_full=h.numRows)  # the shard is a row block
            y = torch.empty(R, dtype=dt, device="cuda")
            ms = tm.time(lambda: p.exec(x, y))
            nb = algorithmic_bytes(R, h.numCols, A.nnz, esz, esz)
            info = p.info()
            res[part] = {"kernel": f"{info.kernel}:{info.param0}", "ms": round(ms, 4), "GBps": round(nb / ms / 1e6, 1)}
            e = errors(y, api.oracle_f64(A, x))
            if part == "shard":
                cpu = cpu_ms(rp, h.colIdx[n0:n1], h.values[n0:n1], x.cpu().numpy(), reps=1)
            else:
                cpu = cpu_ms(h.rowPtr, h.colIdx, h.values, x.cpu().numpy(), reps=1)
            args.results.append(performance_result(
                R, h.numCols, A.nnz, tm.last, nb, rel_l2=e["rel_l2"], max_abs=e["max_abs"],
                tolerance=TOL_FP64 if dt == torch.float64 else TOL_FP32, device=device_record(), cpu=cpu,
                workload=name, part=part, kernel=res[part]["kernel"], gpus=G_, shard_index=g))
            p.destroy()
        res["speedup_8_gpus_if_x_replicated"] = round(res["whole"]["ms"] / res["shard"]["ms"], 2)
        log(json.dumps(res))
        out.append(res)
    return out


def host_path(args) -> list[dict]:
    tm = Timer()
    out = []
    for name, h, dt in (("powerlaw_2m_fp32", G.cfg_powerlaw_2m(), torch.float32),
                        ("laplacian_4096_fp64", G.cfg_laplacian_4096(), torch.float64)):
        vdt = np.float64 if dt == torch.float64 else np.float32
        esz = 8 if dt == torch.float64 else 4
        A = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values).to(dt), h.numCols)
        nb = algorithmic_bytes(h.numRows, h.numCols, h.nnz, esz, esz)
        xs = G.uniform_vector(h.numCols, 3, vdt)
        p = api.Plan(A, "auto", 0, xtype=dt)
        xd = torch.from_numpy(xs).cuda()
        yd = torch.empty(h.numRows, dtype=dt, device="cuda")
        dev_ms = tm.time(lambda: p.exec(xd, yd), flush=False)
        y64 = api.oracle_f64(A, xd)
        cpu = cpu_ms(h.rowPtr, h.colIdx, h.values, xs, reps=1)
        info = p.info()
        p.destroy()
        base = {"workload": name, "device_resident_ms": round(dev_ms, 4), "device_kernel": f"{info.kernel}:{info.param0}",
                "pcie_bytes": (h.numCols + h.numRows) * esz}
        for pinned in (True, False):
            if pinned:
                x = host.pinned_empty(h.numCols, vdt)
                y = host.pinned_empty(h.numRows, vdt)
            else:
                x, y = np.empty(h.numCols, vdt), np.empty(h.numRows, vdt)
            x[:] = xs
            for chunks in (1, 2, 4, 8):
                hp = host.HostPlan(A, chunks, xtype=dt)
                hp.exec(x, y)
                ts = []
                for _ in range(10):
                    hp.exec(x, y)
                    ts.append(hp.timing())
                t = sorted(ts, key=lambda v: v.total_ms)[len(ts) // 2]
                e = errors(torch.from_numpy(np.asarray(y)).cuda(), y64)
                args.results.append(performance_result(
                    h.numRows, h.numCols, h.nnz, [v.total_ms for v in ts], nb, rel_l2=e["rel_l2"],
                    max_abs=e["max_abs"], tolerance=TOL_FP64 if dt == torch.float64 else TOL_FP32,
                    device=device_record(), cpu=cpu, workload=name, path="host_buffers", pinned=pinned,
                    chunks=chunks, timing="end to end, host x in and host y out (PCIe included)"))
                rec = dict(base, pinned=pinned, chunks=chunks, total_ms=round(t.total_ms, 4),
                           first_copy_in_ms=round(t.first_copy_in_ms, 4), kernels_ms=round(t.kernels_ms, 4),
                           copy_out_tail_ms=round(t.copy_out_tail_ms, 4),
                           end_to_end_GBps=round(nb / t.total_ms / 1e6, 1),
                           kernels=[hp.chunk(k)[2] + ":" + str(hp.chunk(k)[3]) for k in range(chunks)])
                hp.destroy()
                log(json.dumps(rec))
                out.append(rec)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["precision", "batched", "spmm", "laplacian", "shard", "host", "kernels"])
    ap.add_argument("--matrix", choices=["powerlaw", "laplacian"], default="powerlaw", help="kernels: the matrix")
    ap.add_argument("--reps", type=int, default=3, help="kernels: cold launches per configuration")
    ap.add_argument("--out", default=None, help="the JSON lines printed on stdout, also to this file")
    ap.add_argument("--results", dest="results_file", default=None,
                    help="results file (default spmv_results_<workload>.json)")
    a = ap.parse_args(argv)
    a.results = []  # performance_result entries, filled by the workload
    torch.cuda.set_device(0)
    res = {"precision": precision, "batched": batched, "spmm": spmm, "laplacian": laplacian, "shard": shard,
           "host": host_path, "kernels": kernels}[a.what](a)
    txt = "\n".join(json.dumps(r) for r in res)
    print(txt, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    path = results_path(a.what, a.results_file)
    write_results_json(path, a.results)
    log(f"{len(a.results)} results -> {path}")
    return 0


def results_path(what: str, given: str | None = None) -> str:
    """The results file of a run: the one given, else spmv_results_<workload>.json in the working directory."""
    return given or f"spmv_results_{what}.json"


if __name__ == "__main__":
    sys.exit(main())
