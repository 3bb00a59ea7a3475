"""SpMV benchmark on one MI355X: FP32 SpMV on the 2M x 2M power-law matrix.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload powerlaw|uniform|laplacian]

A step is one y = A*x over the whole matrix with A, x, y resident in HBM.
Every step reads a different copy of (A, x, y) (enough copies that the
others evict it from the 256 MiB Infinity Cache), so each step is a cold
read. By default the timed kernel is the library's own choice for an AUTO
plan (the learned selector built into the library); the autotune sweep over
the candidate kernels runs before the timed region, and the line reports the
pick's regret against the sweep's best (t_pick / t_best - 1, both timed the
same way). --kernel sweep times the sweep's best instead, --kernel k:p one
configuration.

Prints one JSON line: value = algorithmic GB/s over the K timed steps
(bytes = N*(4+4) + 4*(R+1) + 4*C + 4*R per step); roofline = those bytes
over the launch's kernel time from HIP events recorded around it in every
fourth timed step (a sample spread over the timed region; markers on every
step would add their own cost to a short step). The window holds every
kernel of the launch: for the two-kernel LDS slab path the slab kernel and
the combine after it, whose shares are reported beside it;
vendor = rocSPARSE's CSR SpMV algorithms on the same matrix, warm, beside the
timed plan timed the same way (--no-vendor skips it);
cpu_baseline = the oracle's FP32 CSR loop on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

CANDIDATES = [("merge", 4), ("merge", 8), ("merge", 16), ("csr5", 256), ("csr5", 512), ("csr5", 1024),
              ("adaptive", 512), ("adaptive", 1024), ("adaptive", 2048), ("ell", 0),
              ("xslab", (2 << 8) | 4), ("xslab", (4 << 8) | 4), ("xslab", (8 << 8) | 4),
              ("xslab", (2 << 8) | 8), ("xslab", (4 << 8) | 8), ("xslab", (8 << 8) | 8),
              ("lslab", 4), ("lslab", 8), ("vwave", 4), ("vwave", 8), ("vector", 0), ("scalar", 0)]


# HBM bytes per launch (every kernel of the launch), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# `bench.py --kernel K:P` (KiB per dispatch, FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0, the factors a 1 GiB copy
# kernel reads); the runs and commands are in profiles/lslab_spmv_powerlaw_2m_fp32.md,
# profiles/xslab_spmv_powerlaw_2m_fp32.md, profiles/ell_spmv_laplacian_4096_fp64.md and
# profiles/ell_spmv_uniform_100k_fp32.md. Configurations not measured report null.
MEASURED_TRAFFIC = {
    ("powerlaw_2m_fp32", "lslab", 4): int(((48557.1 + 18917.7) * 2.0 + (28110.1 + 9772.2) * 1.0) * 1024),
    ("powerlaw_2m_fp32", "xslab", 1028): int((66336.6 * 2.0 + 15477.4 * 1.0) * 1024),
    ("powerlaw_2m_fp32", "xslab", 1032): int((68168.4 * 2.0 + 15791.2 * 1.0) * 1024),
    ("laplacian_4096_fp64", "ell", 5): int((483392.3 * 2.0 + 131072.0 * 1.0) * 1024),
    ("uniform_100k_fp32", "ell", 15): int((7443.4 * 2.0 + 390.6 * 1.0) * 1024),
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name: str):
    from muxi_spmv_amd import generators as G
    import numpy as np

    if name == "powerlaw":
        return G.cfg_powerlaw_2m(), np.float32, "powerlaw_2m_fp32"
    if name == "uniform":
        return G.cfg_uniform_100k(), np.float32, "uniform_100k_fp32"
    if name == "laplacian":
        return G.cfg_laplacian_4096(), np.float64, "laplacian_4096_fp64"
    raise SystemExit(f"unknown workload {name}")


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(h, x, budget_s: float = 12.0) -> dict:
    """The oracle's FP32 CSR loop (the reference's CPU loop) on all host cores."""
    import oracle
    from muxi_spmv_amd.benchmark import algorithmic_bytes

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(threads, 64))
    vals = h.values.astype("float32")
    xx = x.astype("float32")
    oracle.spmv_csr_f32_mt(h.rowPtr, h.colIdx, vals, xx, threads)  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.spmv_csr_f32_mt(h.rowPtr, h.colIdx, vals, xx, threads)
        reps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    ms = (time.perf_counter() - t0) * 1e3 / reps
    b = algorithmic_bytes(h.numRows, h.numCols, h.nnz, 4, 4)
    # the reference's exact semantics: one thread, rows in order
    reps1, t0 = 0, time.perf_counter()
    while True:
        oracle.spmv_csr_f32(h.rowPtr, h.colIdx, vals, xx)
        reps1 += 1
        if time.perf_counter() - t0 >= budget_s / 4:
            break
    ms1 = (time.perf_counter() - t0) * 1e3 / reps1
    return {"value": round(b / ms / 1e6, 3), "unit": "GB/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": f"full matrix, {reps} SpMVs of the FP32 CSR loop ({ms:.2f} ms each), OpenMP rows",
            "single_thread": {"value": round(b / ms1 / 1e6, 3), "ms": round(ms1, 3), "reps": reps1}}


def vendor_leg(A0, x0, y0, plan, y64, steps: int = 200) -> dict:
    """rocSPARSE's CSR SpMV (adaptive, rowsplit, LRB: muxi_compare_with_rocsparse) on the same matrix and x,
    beside the timed plan, both warm on one copy and timed alike: every launch between two HIP events, back to
    back, the median taken. Each of rocSPARSE's y is checked against the FP64 oracle's."""
    import torch

    from muxi_spmv_amd.api import SpmvError
    from muxi_spmv_amd.benchmark import compare_with_rocsparse, measure_plans

    m = measure_plans([plan], [x0], [y0], steps, 10, event_every=1)
    out = {"timing": f"warm (one copy, {steps} launches back to back), median of HIP events around each launch",
           "ours_us": round(m.exec_ms_median * 1e3, 3), "rocsparse_us": {}, "rocsparse_analysis_ms": {},
           "rocsparse_rel_l2_vs_fp64": {}}
    for alg in ("adaptive", "rowsplit", "lrb"):
        try:
            r = compare_with_rocsparse(A0, x0, y0, alg, warmup=10, iterations=steps, sync_after_each=False)
        except SpmvError as e:
            out["rocsparse_us"][alg] = None
            log(f"rocSPARSE {alg}: {e}")
            continue
        torch.cuda.synchronize()
        out["rocsparse_us"][alg] = round(r["timeMs"] * 1e3, 3)
        out["rocsparse_analysis_ms"][alg] = round(r["analysisMs"], 3)
        out["rocsparse_rel_l2_vs_fp64"][alg] = float(torch.linalg.vector_norm(y0.double() - y64)
                                                     / torch.linalg.vector_norm(y64))
        log(f"rocSPARSE {alg}: {r['timeMs'] * 1e3:.1f} us (ours {out['ours_us']:.1f} us warm), "
            f"analysis {r['analysisMs']:.1f} ms")
    with open("/proc/self/maps") as f:
        libs = sorted({ln.split()[-1] for ln in f if "librocsparse" in ln})
    out["library"] = libs[0] if libs else None
    out["system_library"] = system_rocsparse(A0, x0, steps)
    times = list(out["rocsparse_us"].values()) + list((out["system_library"] or {}).get("us", {}).values())
    best = min((v for v in times if v), default=None)
    out["speedup_vs_best_rocsparse"] = round(best / out["ours_us"], 3) if best else None
    return out


def system_rocsparse(A0, x0, steps: int) -> dict | None:
    """The ROCm installation's rocSPARSE (all four CSR algorithms) timed the same way by
    muxi_spmv_amd/rocsparse_timing, a C++ caller run as a child process: a PyTorch process holds torch's own,
    older librocsparse under the same soname. The matrix and x go over as raw arrays in a temporary folder."""
    import shutil
    import subprocess
    import tempfile

    exe = ROOT / "muxi_spmv_amd" / "rocsparse_timing"
    if A0.dtype != __import__("torch").float32 or not exe.exists():
        return None
    d = tempfile.mkdtemp(prefix="muxi_rs_")
    try:
        A0.d_rowPtr.cpu().numpy().tofile(f"{d}/rowPtr.i32")
        A0.d_colIdx.cpu().numpy().tofile(f"{d}/colIdx.i32")
        A0.d_values.cpu().numpy().tofile(f"{d}/values.f32")
        x0.cpu().numpy().tofile(f"{d}/x.f32")
        r = subprocess.run([str(exe), d, str(A0.numRows), str(A0.numCols), str(A0.nnz), str(steps)],
                           capture_output=True, text=True, timeout=300)
    except subprocess.TimeoutExpired:
        log("system rocSPARSE: timed out")
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)
    if r.stderr.strip():
        log(f"system rocSPARSE: {r.stderr.strip()}")
    try:
        res = json.loads(r.stdout.strip().splitlines()[-1])
    except (IndexError, ValueError):
        log(f"system rocSPARSE: exit {r.returncode}, no result")
        return None
    res["exit"] = r.returncode
    log(f"system rocSPARSE ({res['library']}): {res['us']} us, analysis {res['analysis_ms']} ms")
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="powerlaw")
    ap.add_argument("--kernel", default="auto",
                    help="auto (the library's pick, sweep timed beside it), sweep (the sweep's best), or kernel:param")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocSPARSE comparison")
    ap.add_argument("--sweep-steps", type=int, default=30)
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the SpMV path here is single-GPU; "
              f"a multi-GPU run splits the rows with muxi_shard_rows and plans each share with "
              f"muxi_plan_create_rows)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from muxi_spmv_amd import api
    from muxi_spmv_amd import generators as G
    from muxi_spmv_amd.benchmark import (HBM_PEAK_GBS, algorithmic_bytes, calculate_gflops, copies_for_cold,
                                         measure_plans)

    torch.cuda.set_device(0)
    t = time.perf_counter()
    h, vdt, wname = make_workload(a.workload)
    x = G.uniform_vector(h.numCols, 12345, vdt)
    log(f"workload {wname}: R={h.numRows} nnz={h.nnz} built in {time.perf_counter() - t:.1f}s")
    tdt = torch.float64 if vdt == np.float64 else torch.float32
    esz = 8 if vdt == np.float64 else 4
    bytes_step = algorithmic_bytes(h.numRows, h.numCols, h.nnz, esz, esz)
    ncopy = copies_for_cold(bytes_step)
    A0 = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values).to(tdt), h.numCols)
    mats = [A0] + [api.CSRMatrix(A0.d_rowPtr.clone(), A0.d_colIdx.clone(), A0.d_values.clone(), h.numCols)
                   for _ in range(ncopy - 1)]
    xs = [torch.from_numpy(x).cuda() for _ in range(ncopy)]
    ys = [torch.zeros(h.numRows, dtype=tdt, device="cuda") for _ in range(ncopy)]
    log(f"{ncopy} rotating copies of {bytes_step / 1e6:.1f} MB")

    # achievable HBM bandwidth: a plain 1 GiB device-to-device copy (read + write)
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    copy_gbs = 2 * src.numel() * 4 * 10 / (time.perf_counter() - t) / 1e9
    del src, dst
    log(f"device copy: {copy_gbs:.0f} GB/s")

    sweep = {}
    if a.kernel in ("auto", "sweep"):
        maxlen = int(np.diff(h.rowPtr).max())
        cands = [c for c in CANDIDATES if not (c[0] in ("vector", "scalar") and h.nnz > 50_000_000)
                 and not (c[0] == "ell" and h.numRows * maxlen > 4 * h.nnz)]
        for k, p in cands:
            plans = [api.Plan(m, k, p, xtype=tdt) for m in mats]
            ms_k = measure_plans(plans, xs, ys, a.sweep_steps, 3, record_kernel=False).ms_per_step
            sweep[f"{k}:{p}"] = round(ms_k, 5)
            log(f"sweep {k}:{p} {ms_k * 1e3:.1f} us/step")
            del plans
        best = min(sweep, key=sweep.get)
    selector = None
    torch.cuda.synchronize()
    t_plan = time.perf_counter()
    if a.kernel == "auto":
        from muxi_spmv_amd._lib import load
        plans = [api.Plan(mats[0], "auto", 0, xtype=tdt)]
        torch.cuda.synchronize()
        t_plan = time.perf_counter() - t_plan
        plans += [api.Plan(m, "auto", 0, xtype=tdt) for m in mats[1:]]
        info = plans[0].info()
        pick = f"{info.kernel}:{info.param0}"
        key = pick
        if info.kernel == "ell" and info.param0 == maxlen and "ell:0" in sweep:
            key = "ell:0"  # the same plan: an ELL width of 0 means the longest row
        if key not in sweep:  # time the pick the way the sweep timed its candidates
            sweep[key] = round(measure_plans(plans, xs, ys, a.sweep_steps, 3, record_kernel=False).ms_per_step, 5)
        state = {0: "fixed rule", 1: "learned (built in)", 2: "learned (set by caller)"}[load().muxi_selector_state()]
        selector = {"state": state, "pick": pick, "sweep_best": best,
                    "regret": round(sweep[key] / sweep[best] - 1.0, 4)}
        log(f"selector ({state}) picks {pick}; sweep best {best}; regret {selector['regret']:.3f}")
    else:
        if a.kernel != "sweep":
            best = a.kernel
        kname, kparam = best.split(":")
        plans = [api.Plan(mats[0], kname, int(kparam), xtype=tdt)]
        torch.cuda.synchronize()
        t_plan = time.perf_counter() - t_plan
        plans += [api.Plan(m, kname, int(kparam), xtype=tdt) for m in mats[1:]]
        info = plans[0].info()

    torch.cuda.synchronize()
    meas = measure_plans(plans, xs, ys, a.steps, a.warmup)
    ms = meas.ms_per_step
    value = bytes_step / ms / 1e6
    kms = meas.kernel_ms_mean
    ems = meas.exec_ms_mean
    if not ems == ems:  # no sample of the whole launch (very few steps): the wall clock per step
        ems = ms
    if not kms == kms:
        kms = ems
    achieved = bytes_step / ems / 1e6
    log(f"{info.kernel}:{info.param0}: {ms * 1e3:.2f} us/step, launch {ems * 1e3:.2f} us, main kernel "
        f"{kms * 1e3:.2f} us (median {meas.kernel_ms_median * 1e3:.2f})")

    # correctness of the measured configuration against the device FP64 oracle
    y = torch.zeros(h.numRows, dtype=tdt, device="cuda")
    plans[0].exec(xs[0], y)
    y64 = api.oracle_f64(A0, xs[0])
    torch.cuda.synchronize()
    rel_l2 = float(torch.linalg.vector_norm(y.double() - y64) / torch.linalg.vector_norm(y64))
    max_abs = float((y.double() - y64).abs().max()) if h.numRows else 0.0

    out = {
        "metric": "SpMV GB/s (frac of HBM peak) + GFLOP/s; learned-pick regret vs sweep",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64" if vdt == np.float64 else "fp32",
        "data": "synthetic (seeded generator, cold: each step reads one of "
                f"{ncopy} rotating copies)",
        "config": {"workload": wname, "rows": h.numRows, "cols": h.numCols, "nnz": h.nnz,
                   "kernel": info.kernel, "param": info.param0, "parallelism": "single GPU"},
        "gflops": round(calculate_gflops(h.nnz, ms), 3),
        "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4),
        "frac_of_device_copy": round(value / copy_gbs, 4),
        "rel_l2_vs_fp64": rel_l2,
        "max_abs_vs_fp64": max_abs,
        "plan": {"build_ms": round(t_plan * 1e3, 2), "scratch_bytes": info.scratch_bytes},
        "hbm_copy_gbs": round(copy_gbs, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((wname, info.kernel, info.param0)),
                     "kernels": "slab kernel + combine" if info.kernel == "lslab" else f"{info.kernel} main kernel"
                     + (" + carry fix-up" if info.kernel in ("merge", "csr5", "adaptive", "xslab") else ""),
                     "launch_us_mean": round(ems * 1e3, 3), "main_kernel_us_mean": round(kms * 1e3, 3),
                     "launch_us_median": round(meas.exec_ms_median * 1e3, 3), "launch_samples": len(meas.exec_ms),
                     "bytes_per_launch": bytes_step},
        "sweep_ms": sweep,
        "selector": selector,
    }
    out["vendor"] = None
    if not a.no_vendor:
        try:  # a side measurement: whatever it meets, the bench line is still printed
            out["vendor"] = vendor_leg(A0, xs[0], ys[0], plans[0], y64)
        except Exception as e:  # noqa: BLE001
            log(f"vendor comparison skipped: {e!r}")
            out["vendor"] = {"error": repr(e)}
    out["cpu_baseline"] = None
    if not a.no_cpu:
        log("cpu baseline ...")
        try:
            out["cpu_baseline"] = cpu_baseline(h, x)
        except Exception as e:  # noqa: BLE001  (the GPU figures stand without it; say why it is missing)
            log(f"cpu baseline failed: {e!r}")
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
