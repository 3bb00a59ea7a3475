"""Autotuner for the batched launch and SpMM schedules.

python -m muxi_spmv_amd.sched_tune [--sweep] [--fit] [--db PATH] [--tables PATH] [--time-budget S]

1. Sweep (on the GPU): every candidate schedule is timed cold (a 512 MiB cache flush before each run, HIP events,
   median of 5) on seeded synthetic workloads, one JSON record per workload appended to the DB (resumable: a
   workload already in the DB is skipped).
     batched  batches of Poisson-row matrices (uniform columns, like the 256-matrix benchmark), one workload per
              (mean nonzeros per row bin) x (mean rows per matrix bin) cell and repeat, plus mixed batches of 256
              matrices drawn as the benchmark draws them (rows log-uniform in [1e3, 2e5], mean nonzeros per row
              log-uniform in [4, 64]) with other seeds; candidates: merge items per lane {4, 8, 16} x workgroups
              per CU {2, 4, 8, 16}. Repeats 0-2 and 6 mixed batches train; repeat 3 and 2 mixed batches are held
              out.
     spmm     uniform, power-law and banded matrices of 2e4 to 1e6 rows at k = 1, 2, 4, 8, 16, 32, 64 right-hand
              sides; candidates: items per lane group {4, 8, 16, 32}. Matrices with an odd index are held out.
              The SpMM table is keyed by k and the merge-path length (numRows + nnz): small matrices need short
              tiles to fill the GPU, large ones long tiles (fewer carries); the mean row length keyed it worse
              (held-out p95 regret 24 % against 9 %).
2. Fit (CPU): per table cell, the candidate with the smallest geometric mean of t / t_best over the training
   workloads in the cell (the best overall where a cell has none). The tables go to sched_tables.json, which
   build.py compiles into the library (csrc/sched.hpp): muxi_batch_plan_create and muxi_plan_create_spmm take
   them when called with ipt <= 0.
3. Evaluate: regret t_pick / t_best - 1 of the table's pick on the held-out workloads, next to the former fixed
   defaults (16 items x 8 workgroups per CU; 8 items).
"""
from __future__ import annotations

import argparse
import json
import math
import statistics
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
DB = PKG / "sched_db.json"
TABLES = PKG / "sched_tables.json"

BATCH_IPT = (4, 8, 16)
BATCH_WG = (2, 4, 8, 16)
BATCH_CANDIDATES = [(i, w) for i in BATCH_IPT for w in BATCH_WG]
BATCH_ROW_LEN_EDGES = [6.0, 12.0, 24.0, 48.0]  # mean nonzeros per row: 5 bins
BATCH_ROW_LEN_RANGE = (2.0, 128.0)
BATCH_ROWS_EDGES = [1e4, 1e5]  # mean rows per matrix: 3 bins
BATCH_ROWS_RANGE = (1e3, 5e5)
BATCH_REPEATS = 4  # the last one held out
BATCH_MIXED = 8  # batches drawn like the 256-matrix benchmark (other seeds): the first 6 train, 2 held out
BATCH_DEFAULT = (16, 8)

SPMM_IPT = (4, 8, 16, 32)
SPMM_K = (1, 2, 4, 8, 16, 32, 64)
SPMM_PATH_EDGES = [5e5, 5e6]  # merge-path length numRows + nnz: 3 bins
SPMM_DEFAULT = 8
# (family, rows, p0, p1, p2): mean row lengths from about 3 to 64
SPMM_MATRICES = [
    ("uniform", 20_000, 2, 4, 0), ("uniform", 200_000, 2, 4, 0), ("uniform", 1_000_000, 2, 4, 0),
    ("uniform", 20_000, 8, 16, 0), ("uniform", 200_000, 8, 16, 0), ("uniform", 1_000_000, 8, 16, 0),
    ("uniform", 20_000, 32, 64, 0), ("uniform", 200_000, 32, 64, 0), ("uniform", 600_000, 32, 64, 0),
    ("powerlaw", 20_000, 2.0, 1, 20_000), ("powerlaw", 200_000, 2.0, 1, 100_000), ("powerlaw", 1_000_000, 2.0, 1, 100_000),
    ("powerlaw", 50_000, 2.6, 2, 10_000), ("powerlaw", 500_000, 2.6, 2, 50_000),
    ("banded", 20_000, 64, 16, 0), ("banded", 200_000, 64, 16, 0), ("banded", 1_000_000, 64, 16, 0),
    ("banded", 200_000, 16, 4, 0), ("banded", 1_000_000, 128, 48, 0), ("uniform", 100_000, 24, 40, 0),
]


def _bin(v: float, edges) -> int:
    return int(sum(1 for e in edges if v >= e))


def _log_uniform(rng, lo, hi):
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))


def batch_spec(i: int) -> dict:
    """Workload i: cell i mod 15, repeat i // 15; past the 60 cell workloads, mixed batch i - 60."""
    n_len, n_rows = len(BATCH_ROW_LEN_EDGES) + 1, len(BATCH_ROWS_EDGES) + 1
    n_cell = n_len * n_rows * BATCH_REPEATS
    if i >= n_cell:
        m = i - n_cell
        rng = np.random.default_rng(9000 + m)
        members = []
        for q in range(256):
            R = int(math.exp(rng.uniform(math.log(1e3), math.log(2e5))))
            mu = math.exp(rng.uniform(math.log(4), math.log(64)))
            members.append((R, mu, 70_000 * (m + 1) + q))
        return {"kind": "batch", "i": i, "cell": None, "mixed": True,
                "repeat": 0 if m < BATCH_MIXED - 2 else BATCH_REPEATS - 1, "members": members}
    cell = i % (n_len * n_rows)
    a, b = divmod(cell, n_rows)
    le = [BATCH_ROW_LEN_RANGE[0]] + BATCH_ROW_LEN_EDGES + [BATCH_ROW_LEN_RANGE[1]]
    re = [BATCH_ROWS_RANGE[0]] + BATCH_ROWS_EDGES + [BATCH_ROWS_RANGE[1]]
    rng = np.random.default_rng(1000 + i)
    mean_len = _log_uniform(rng, le[a], le[a + 1])
    mean_rows = _log_uniform(rng, re[b], re[b + 1])
    count = int(np.clip(1e8 / (mean_rows * mean_len), 4, 64))
    members = []
    for m in range(count):
        R = int(np.clip(mean_rows * math.exp(rng.uniform(-0.5, 0.5)), 500, 2_000_000))
        mu = float(np.clip(mean_len * math.exp(rng.uniform(-0.35, 0.35)), 1.0, 200.0))
        members.append((R, mu, 50_000 * i + m))
    return {"kind": "batch", "i": i, "cell": [a, b], "repeat": i // (n_len * n_rows), "members": members}


def batch_features(rows: list[int], nnz: list[int]) -> tuple[float, float]:
    R, N = sum(rows), sum(nnz)
    return (N / R if R else 0.0), (R / len(rows) if rows else 0.0)


class _Timer:
    def __init__(self):
        import ctypes as C

        import torch

        from . import tune
        from ._lib import check, load

        self.C, self.torch, self.check, self.lib = C, torch, check, load()
        self.ev = [C.c_void_p(), C.c_void_p()]
        for e in self.ev:
            check(self.lib.muxi_event_create(C.byref(e)), "event_create")
        self.fl = tune.Flusher()

    def cold(self, fn, reps=5) -> float:
        """Median ms of fn() over reps, each preceded by a cache flush."""
        s = self.torch.cuda.current_stream().cuda_stream or None
        fn()
        ts = []
        for _ in range(reps):
            self.fl()
            self.lib.muxi_event_record(self.ev[0], s)
            fn()
            self.lib.muxi_event_record(self.ev[1], s)
            ms = self.C.c_float()
            self.check(self.lib.muxi_event_elapsed_ms(self.ev[0], self.ev[1], self.C.byref(ms)), "elapsed")
            ts.append(ms.value)
        return statistics.median(ts)


def sweep_batch(spec: dict, tm: _Timer) -> dict:
    import torch

    from . import tune
    from .batch import BatchPlan

    mats, xs, ys = [], [], []
    try:
        for R, mu, seed in spec["members"]:
            A = tune.gen_matrix("poisson", R, R, mu, seed=seed)
            mats.append(A)
            xs.append(tune.gen_vector(R, seed=seed + 7))
            ys.append(torch.empty(R, device="cuda"))
        rows, nnz = [m.numRows for m in mats], [m.nnz for m in mats]
        times = {}
        for ipt, wg in BATCH_CANDIDATES:
            bp = BatchPlan(mats, xs, ys, ipt | (wg << 8))
            times[f"{ipt}x{wg}"] = tm.cold(lambda: bp.exec())
            bp.destroy()
        ml, mr = batch_features(rows, nnz)
        return dict(spec, members=len(spec["members"]), rows=sum(rows), nnz=sum(nnz), mean_row_len=ml,
                    mean_rows=mr, times=times)
    finally:
        for A in mats:
            A.free()


def spmm_spec(j: int) -> dict:
    fam, R, p0, p1, p2 = SPMM_MATRICES[j]
    return {"kind": "spmm", "j": j, "family": fam, "rows": R, "p": [p0, p1, p2], "seed": 300 + j,
            "split": "test" if j % 2 else "train"}


def sweep_spmm(spec: dict, tm: _Timer) -> list[dict]:
    import torch

    from . import api, tune

    fam, R = spec["family"], spec["rows"]
    A = tune.gen_matrix(fam, R, R, *spec["p"], seed=spec["seed"])
    out = []
    try:
        for k in SPMM_K:
            if R * k * 4 > (1 << 30):
                continue
            X = torch.rand((R, k), device="cuda") * 2 - 1
            Y = torch.empty((R, k), device="cuda")
            times = {}
            for ipt in SPMM_IPT:
                p = api.Plan.spmm(A, k, ipt)
                times[str(ipt)] = tm.cold(lambda: p.exec(X, Y))
                p.destroy()
            out.append(dict(spec, k=k, nnz=int(A.nnz), mean_row_len=A.nnz / max(R, 1), path=R + int(A.nnz),
                            times=times))
            del X, Y
    finally:
        A.free()
    return out


# ---------------------------------------------------------------- fit and evaluation (CPU)
def load_db(path: Path = DB) -> list[dict]:
    return json.loads(path.read_text())["records"] if path.exists() else []


def _geo_score(recs, name) -> float:
    return math.exp(statistics.fmean(math.log(r["times"][name] / min(r["times"].values())) for r in recs))


def fit(records: list[dict]) -> dict:
    batch = [r for r in records if r["kind"] == "batch" and r["repeat"] < BATCH_REPEATS - 1]
    spmm = [r for r in records if r["kind"] == "spmm" and r["split"] == "train"]
    n_len, n_rows = len(BATCH_ROW_LEN_EDGES) + 1, len(BATCH_ROWS_EDGES) + 1
    names = [f"{i}x{w}" for i, w in BATCH_CANDIDATES]
    overall = min(names, key=lambda n: _geo_score(batch, n)) if batch else f"{BATCH_DEFAULT[0]}x{BATCH_DEFAULT[1]}"
    btab = []
    for a in range(n_len):
        for b in range(n_rows):
            cell = [r for r in batch if _bin(r["mean_row_len"], BATCH_ROW_LEN_EDGES) == a
                    and _bin(r["mean_rows"], BATCH_ROWS_EDGES) == b]
            best = min(names, key=lambda n: _geo_score(cell, n)) if cell else overall
            i, w = best.split("x")
            btab.append([int(i), int(w)])
    n_sl = len(SPMM_PATH_EDGES) + 1
    stab = []
    for k in SPMM_K:
        allk = [r for r in spmm if r["k"] == k]
        overall_k = min((str(i) for i in SPMM_IPT), key=lambda n: _geo_score(allk, n)) if allk else str(SPMM_DEFAULT)
        for a in range(n_sl):
            cell = [r for r in allk if _bin(_path(r), SPMM_PATH_EDGES) == a]
            best = min((str(i) for i in SPMM_IPT), key=lambda n: _geo_score(cell, n)) if cell else overall_k
            stab.append(int(best))
    return {"batch": {"row_len_edges": BATCH_ROW_LEN_EDGES, "rows_edges": BATCH_ROWS_EDGES, "table": btab},
            "spmm": {"k": list(SPMM_K), "path_edges": SPMM_PATH_EDGES, "table": stab}}


def batch_pick(tables: dict, mean_row_len: float, mean_rows: float) -> tuple[int, int]:
    t = tables["batch"]
    a = _bin(mean_row_len, t["row_len_edges"])
    b = _bin(mean_rows, t["rows_edges"])
    i, w = t["table"][a * (len(t["rows_edges"]) + 1) + b]
    return int(i), int(w)


def _path(r: dict) -> int:
    return int(r.get("path", r["rows"] + r["nnz"]))


def spmm_pick(tables: dict, k: int, path: float) -> int:
    t = tables["spmm"]
    ki = max(i for i, kv in enumerate(t["k"]) if kv <= k) if k >= t["k"][0] else 0
    a = _bin(path, t["path_edges"])
    return int(t["table"][ki * (len(t["path_edges"]) + 1) + a])


def evaluate(tables: dict, records: list[dict]) -> dict:
    def stats(v):
        v = sorted(v)
        return {"n": len(v), "median": v[len(v) // 2] if v else None,
                "p95": v[min(len(v) - 1, int(0.95 * len(v)))] if v else None, "max": v[-1] if v else None}

    bt = [r for r in records if r["kind"] == "batch" and r["repeat"] == BATCH_REPEATS - 1]
    st = [r for r in records if r["kind"] == "spmm" and r["split"] == "test"]
    reg_b, reg_bd, reg_s, reg_sd = [], [], [], []
    for r in bt:
        best = min(r["times"].values())
        i, w = batch_pick(tables, r["mean_row_len"], r["mean_rows"])
        reg_b.append(r["times"][f"{i}x{w}"] / best - 1)
        reg_bd.append(r["times"][f"{BATCH_DEFAULT[0]}x{BATCH_DEFAULT[1]}"] / best - 1)
    for r in st:
        best = min(r["times"].values())
        reg_s.append(r["times"][str(spmm_pick(tables, r["k"], _path(r)))] / best - 1)
        reg_sd.append(r["times"][str(SPMM_DEFAULT)] / best - 1)
    return {"batch_tuned": stats(reg_b), "batch_fixed_16x8": stats(reg_bd), "spmm_tuned": stats(reg_s),
            "spmm_fixed_8": stats(reg_sd)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--sweep", action="store_true", help="time the missing workloads on the GPU")
    ap.add_argument("--fit", action="store_true", help="fit the tables from the DB and evaluate them")
    ap.add_argument("--db", default=str(DB))
    ap.add_argument("--tables", default=str(TABLES))
    ap.add_argument("--time-budget", type=float, default=None)
    a = ap.parse_args(argv)
    db = Path(a.db)
    records = load_db(db)
    if a.sweep:
        import torch

        torch.cuda.set_device(0)
        tm = _Timer()
        have_b = {r["i"] for r in records if r["kind"] == "batch"}
        have_s = {r["j"] for r in records if r["kind"] == "spmm"}
        t0 = time.time()
        n_cells = (len(BATCH_ROW_LEN_EDGES) + 1) * (len(BATCH_ROWS_EDGES) + 1)
        todo = [("batch", i) for i in range(n_cells * BATCH_REPEATS + BATCH_MIXED) if i not in have_b] + \
               [("spmm", j) for j in range(len(SPMM_MATRICES)) if j not in have_s]
        for kind, i in todo:
            if a.time_budget is not None and time.time() - t0 > a.time_budget:
                print(f"[sched] time budget reached", flush=True)
                break
            if kind == "batch":
                rec = sweep_batch(batch_spec(i), tm)
                records.append(rec)
                print(f"[sched] batch {i} cell {rec['cell'] or 'mixed'} {rec['members']} matrices {rec['nnz']} nnz: best "
                      f"{min(rec['times'], key=rec['times'].get)} {time.time() - t0:.0f}s", flush=True)
            else:
                recs = sweep_spmm(spmm_spec(i), tm)
                records.extend(recs)
                print(f"[sched] spmm {i} {SPMM_MATRICES[i][0]} {SPMM_MATRICES[i][1]} rows: best "
                      f"{[min(r['times'], key=r['times'].get) for r in recs]} {time.time() - t0:.0f}s", flush=True)
            db.write_text(json.dumps({"records": records}))
    if a.fit:
        tables = fit(records)
        tables["held_out_regret"] = evaluate(tables, records)
        Path(a.tables).write_text(json.dumps(tables, indent=1))
        print(json.dumps(tables["held_out_regret"]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
