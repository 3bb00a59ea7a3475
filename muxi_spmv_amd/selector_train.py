"""Train and evaluate the learned kernel selector.

python -m muxi_spmv_amd.selector_train [--train 2000] [--test 400] [--db PATH] [--weights PATH]

1. Matrices i = 0 .. train+test-1 come from five seeded families (uniform
   rows, power-law rows, banded, 32x32 block-diagonal, banded with random
   row and column permutations; family = i mod 5, seed 6 + i), rows
   log-uniform in [1e4, 4e6], FP32, generated on the device. Indices
   BIG_BASE + j are further matrices of the same families with rows
   log-uniform in [4e6, 2e7] (--big-train, then --big-test of them), so the
   named workloads of up to 16.8 M rows are inside the training range. Index
   FP64_BASE + i is matrix i again with FP64 values, x and y (the first
   --fp64-train training and --fp64-test held-out matrices of both ranges),
   so the network sees what the precision changes.
2. The autotune sweep times every candidate kernel/configuration cold and
   appends one JSON record per matrix to the DB (resumable: a record is
   only topped up with the candidates it lacks; --skip leaves kernels out of
   a call, --retime times kernels again after they changed; with
   --retime-tag the records re-timed are tagged, so a later call resumes).
3. The MLP (HIP forward/backward/Adam kernels) is fit on the training
   records to predict log(t_candidate / t_best), --ensemble times from
   different seeds; the shipped network averages them (one wider MLP with
   block-diagonal hidden layers). Its pick is the argmin.
4. On the held-out matrices it reports the median and p95 regret
   t_pick / t_best - 1 against the sweep, next to the reference's fixed rule
   and the best single configuration.
"""
from __future__ import annotations

import argparse
import gzip
import json
import math
import sys
import time
from pathlib import Path

import numpy as np
import torch

from . import tune

FAMILIES = ["uniform", "powerlaw", "banded", "blockdiag", "banded_perm"]
BIG_BASE = 50_000
BIG_ROWS = (4e6, 2e7)
FP64_BASE = 100_000


def matrix_spec(i: int, rows_min: float, rows_max: float) -> dict:
    if i >= FP64_BASE:
        spec = matrix_spec(i - FP64_BASE, rows_min, rows_max)
        spec.update(i=i, dtype="fp64")
        return spec
    if i >= BIG_BASE:
        rows_min, rows_max = BIG_ROWS
    rng = np.random.default_rng(6 + i)
    fam = FAMILIES[i % 5]
    R = int(math.exp(rng.uniform(math.log(rows_min), math.log(rows_max))))
    spec = {"i": i, "family": fam, "rows": R, "seed": 6 + i}
    if fam == "uniform":
        mean = math.exp(rng.uniform(math.log(2), math.log(64)))
        spec.update(p0=max(1, int(mean * 0.5)), p1=max(1, int(mean * 1.5)), p2=0)
    elif fam == "powerlaw":
        spec.update(p0=float(rng.uniform(1.8, 3.0)), p1=int(rng.integers(1, 5)), p2=100000)
    elif fam in ("banded", "banded_perm"):
        hb = int(math.exp(rng.uniform(math.log(8), math.log(256))))
        spec.update(p0=hb, p1=int(rng.integers(4, min(2 * hb + 1, 64) + 1)), p2=0)
    else:
        spec.update(p0=32, p1=0, p2=0)
    return spec


def build(spec: dict):
    fam = spec["family"]
    gfam = "banded" if fam == "banded_perm" else fam
    return tune.gen_matrix(gfam, spec["rows"], spec["rows"], spec["p0"], spec["p1"], spec["p2"], seed=spec["seed"],
                           permute=(fam == "banded_perm"), dtype=_dtype(spec))


def _dtype(spec: dict) -> torch.dtype:
    return torch.float64 if spec.get("dtype") == "fp64" else torch.float32


def load_db(path: Path) -> dict[int, dict]:
    """Records by matrix index (a later line for an index replaces an earlier one); .gz files are read as gzip."""
    recs = {}
    if path.exists():
        text = gzip.decompress(path.read_bytes()).decode() if path.suffix == ".gz" else path.read_text()
        for line in text.splitlines():
            if line.strip():
                r = json.loads(line)
                recs[r["i"]] = r
    return recs


def _todo(rec, cands, retime, retime_tag=""):
    """Candidates still to time on a record: feasible ones it lacks, plus those named for re-timing (unless the
    record already carries retime_tag, so that a re-timing split over several calls resumes)."""
    if rec is None:
        return list(cands)
    f = np.array(rec["features"])
    if retime_tag and retime_tag in rec.get("retimed", ()):
        retime = ()
    fp64 = rec.get("dtype") == "fp64"
    return [(k, p) for k, p in cands
            if tune.feasible(k, f, fp64) and (f"{k}:{p}" not in rec["times"] or k in retime)]


def run_sweeps(indices, db_path: Path, rows_min, rows_max, reps, log_every=25, time_budget_s=None,
               skip=(), retime=(), retime_tag=""):
    """Sweeps matrices into the DB (one JSON line per matrix; a later line for the same index replaces an
    earlier one). Kernels in `skip` are not timed; kernels in `retime` are timed again on records that
    already hold them; candidates a record lacks are added to it."""
    done = load_db(db_path)
    cands = [(k, p) for k, p in tune.CANDIDATES if k not in skip]
    fl = tune.Flusher()
    t0 = time.time()
    n_new = 0
    with open(db_path, "a") as f:
        for i in indices:
            todo = _todo(done.get(i), cands, retime, retime_tag)
            if not todo:
                continue
            if time_budget_s is not None and time.time() - t0 > time_budget_s:
                print(f"[sweep] time budget reached after {n_new} matrices", flush=True)
                break
            spec = matrix_spec(i, rows_min, rows_max)
            A = build(spec)
            x = tune.gen_vector(A.numCols, dtype=_dtype(spec), seed=spec["seed"])
            y = torch.empty(A.numRows, dtype=_dtype(spec), device="cuda")
            feats = tune.features(A) if i not in done else np.array(done[i]["features"])
            times = dict(done[i]["times"]) if i in done else {}
            times.update(tune.sweep(A, x, y, fl, reps=reps, candidates=todo, f=feats))
            rec = dict(spec, nnz=int(A.nnz), features=feats.tolist(), times=times)
            tags = list(done[i].get("retimed", [])) if i in done else []
            if retime_tag and retime and retime_tag not in tags:
                tags.append(retime_tag)
            if tags:
                rec["retimed"] = tags
            f.write(json.dumps(rec) + "\n")
            f.flush()
            done[i] = rec
            n_new += 1
            A.free()
            del x, y
            if n_new % log_every == 0:
                print(f"[sweep] {n_new} new, at i={i}, {time.time() - t0:.0f}s", flush=True)
    return done


def dataset(recs):
    X, T, M = [], [], []
    for r in recs:
        X.append(tune.feature_vector(np.array(r["features"]), r.get("dtype") == "fp64"))
        t, m = tune.targets_from_sweep(r["times"])
        T.append(t)
        M.append(m)
    return np.array(X), np.array(T), np.array(M)


def regrets(recs, picks: list[str]) -> np.ndarray:
    out = []
    for r, p in zip(recs, picks):
        best = min(r["times"].values())
        t = r["times"].get(p)
        out.append(float("inf") if t is None else t / best - 1.0)
    return np.array(out)


def evaluate(sel: tune.Selector, train, test) -> dict:
    Xt, _, Mt = dataset(test)
    pick_idx = sel.pick(Xt, Mt)
    picks = [tune.CAND_NAMES[k] for k in pick_idx]
    reg = regrets(test, picks)
    ref = regrets(test, [tune.reference_rule(np.array(r["features"])) for r in test])
    # best single configuration on the training set (the method without the learned part)
    names = tune.CAND_NAMES
    tot = {n: 0.0 for n in names}
    for r in train:
        best = min(r["times"].values())
        for n in names:
            tot[n] += math.log(r["times"][n] / best) if n in r["times"] else 10.0
    fixed = min(tot, key=tot.get)
    fix = regrets(test, [fixed] * len(test))
    q = lambda a, p: float(np.percentile(a[np.isfinite(a)], p)) if np.isfinite(a).any() else None  # noqa: E731
    by = {}
    groups = {"fp32": lambda r: r.get("dtype") != "fp64", "fp64": lambda r: r.get("dtype") == "fp64",
              "rows_over_4e6": lambda r: r["i"] % FP64_BASE >= BIG_BASE}
    for name, g in groups.items():
        sel_ = np.array([g(r) for r in test])
        if sel_.any():
            by[name] = {"n": int(sel_.sum()), "median": q(reg[sel_], 50), "p95": q(reg[sel_], 95),
                        "exact_best_frac": float(np.mean(reg[sel_] == 0.0)), "reference_rule_median": q(ref[sel_], 50)}
    return {
        "n_test": len(test), "n_train": len(train),
        "learned": {"median": q(reg, 50), "p95": q(reg, 95), "mean": float(np.mean(reg)),
                    "exact_best_frac": float(np.mean(reg == 0.0))},
        "by_group": by,
        "reference_rule": {"median": q(ref, 50), "p95": q(ref, 95), "mean": float(np.mean(ref[np.isfinite(ref)]))},
        "best_fixed": {"config": fixed, "median": q(fix, 50), "p95": q(fix, 95), "mean": float(np.mean(fix))},
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=400)
    ap.add_argument("--rows-min", type=float, default=1e4)
    ap.add_argument("--rows-max", type=float, default=4e6)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--db", default="selector_db.jsonl", help="sweep records to append to (JSON lines)")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--summary", default=None)
    ap.add_argument("--sweep-only", action="store_true")
    ap.add_argument("--train-only", action="store_true")
    ap.add_argument("--epochs", type=int, default=8000)
    ap.add_argument("--ensemble", type=int, default=4, help="networks averaged into the shipped one")
    ap.add_argument("--fp64-train", type=int, default=2000, help="training matrices also swept with FP64 values")
    ap.add_argument("--fp64-test", type=int, default=400, help="held-out matrices also swept with FP64 values")
    ap.add_argument("--big-train", type=int, default=300, help="training matrices with 4e6 to 2e7 rows")
    ap.add_argument("--big-test", type=int, default=60, help="held-out matrices with 4e6 to 2e7 rows")
    ap.add_argument("--time-budget", type=float, default=None, help="seconds of sweeping in this call")
    ap.add_argument("--skip", default="", help="comma-separated kernels not to time in this call")
    ap.add_argument("--retime", default="", help="comma-separated kernels to time again on existing records")
    ap.add_argument("--retime-tag", default="", help="records re-timed under this tag are not re-timed again")
    a = ap.parse_args(argv)
    torch.cuda.set_device(0)
    db = Path(a.db)
    db.parent.mkdir(parents=True, exist_ok=True)
    n = a.train + a.test
    big_train = [BIG_BASE + j for j in range(a.big_train)]
    big_test = [BIG_BASE + a.big_train + j for j in range(a.big_test)]
    f64_train = [FP64_BASE + j for j in range(min(a.fp64_train, a.train))] + [FP64_BASE + i for i in big_train]
    f64_test = [FP64_BASE + a.train + j for j in range(min(a.fp64_test, a.test))] + [FP64_BASE + i for i in big_test]
    if not a.train_only:
        # interleave so that a partial run covers both splits
        order = list(range(a.train, n)) + list(range(a.train)) + big_test + big_train + f64_test + f64_train
        run_sweeps(order, db, a.rows_min, a.rows_max, a.reps, time_budget_s=a.time_budget,
                   skip=tuple(k for k in a.skip.split(",") if k), retime=tuple(k for k in a.retime.split(",") if k),
                   retime_tag=a.retime_tag)
        if a.sweep_only:
            return 0
    recs = load_db(db)
    train = [recs[i] for i in list(range(a.train)) + big_train + f64_train if i in recs]
    test = [recs[i] for i in list(range(a.train, n)) + big_test + f64_test if i in recs]
    print(f"[train] {len(train)} train / {len(test)} held-out records", flush=True)
    X, T, M = dataset(train)
    members = []
    for k in range(max(1, a.ensemble)):
        m = tune.Selector(64, tune.N_INPUT, len(tune.CANDIDATES), seed=k)
        losses = m.fit(X, T, M, epochs=a.epochs, lr=2e-3, batch=256, seed=k)
        print(f"[train] member {k}: loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
        members.append(m)
    sel = tune.Selector.ensemble(members) if len(members) > 1 else members[0]
    res = evaluate(sel, train, test)
    res["loss_first_last"] = [losses[0], losses[-1]]
    print(json.dumps(res), flush=True)
    if a.weights:
        sel.save(a.weights)
    if a.summary:
        Path(a.summary).write_text(json.dumps(res, indent=2))
    return 0


if __name__ == "__main__":
    sys.exit(main())
