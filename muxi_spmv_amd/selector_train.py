"""Train and evaluate the learned kernel selector.

python -m muxi_spmv_amd.selector_train [--train 2000] [--test 400] [--db PATH] [--weights PATH]

1. Matrices i = 0 .. train+test-1 come from five seeded families (uniform
   rows, power-law rows, banded, 32x32 block-diagonal, banded with random
   row and column permutations; family = i mod 5, seed 6 + i), rows
   log-uniform in [1e4, 4e6], FP32, generated on the device.
2. The autotune sweep times every candidate kernel/configuration cold and
   appends one JSON record per matrix to the DB (resumable: a record is
   only topped up with the candidates it lacks; --skip leaves kernels out of
   a call, --retime times kernels again after they changed).
3. The MLP (HIP forward/backward/Adam kernels) is fit on the training
   records to predict log(t_candidate / t_best); its pick is the argmin.
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


def matrix_spec(i: int, rows_min: float, rows_max: float) -> dict:
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
                           permute=(fam == "banded_perm"))


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


def _todo(rec, cands, retime):
    """Candidates still to time on a record: feasible ones it lacks, plus those named for re-timing."""
    if rec is None:
        return list(cands)
    f = np.array(rec["features"])
    return [(k, p) for k, p in cands
            if tune.feasible(k, f) and (f"{k}:{p}" not in rec["times"] or k in retime)]


def run_sweeps(indices, db_path: Path, rows_min, rows_max, reps, log_every=25, time_budget_s=None,
               skip=(), retime=()):
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
            todo = _todo(done.get(i), cands, retime)
            if not todo:
                continue
            if time_budget_s is not None and time.time() - t0 > time_budget_s:
                print(f"[sweep] time budget reached after {n_new} matrices", flush=True)
                break
            spec = matrix_spec(i, rows_min, rows_max)
            A = build(spec)
            x = tune.gen_vector(A.numCols, seed=spec["seed"])
            y = torch.empty(A.numRows, device="cuda")
            feats = tune.features(A) if i not in done else np.array(done[i]["features"])
            times = dict(done[i]["times"]) if i in done else {}
            times.update(tune.sweep(A, x, y, fl, reps=reps, candidates=todo, f=feats))
            rec = dict(spec, nnz=int(A.nnz), features=feats.tolist(), times=times)
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
        X.append(tune.feature_vector(np.array(r["features"])))
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
    return {
        "n_test": len(test), "n_train": len(train),
        "learned": {"median": q(reg, 50), "p95": q(reg, 95), "mean": float(np.mean(reg)),
                    "exact_best_frac": float(np.mean(reg == 0.0))},
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
    ap.add_argument("--db", default="gpurun_out/selector_db.jsonl")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--summary", default=None)
    ap.add_argument("--sweep-only", action="store_true")
    ap.add_argument("--train-only", action="store_true")
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--time-budget", type=float, default=None, help="seconds of sweeping in this call")
    ap.add_argument("--skip", default="", help="comma-separated kernels not to time in this call")
    ap.add_argument("--retime", default="", help="comma-separated kernels to time again on existing records")
    a = ap.parse_args(argv)
    torch.cuda.set_device(0)
    db = Path(a.db)
    db.parent.mkdir(parents=True, exist_ok=True)
    n = a.train + a.test
    if not a.train_only:
        # interleave so that a partial run covers both splits
        order = list(range(a.train, n)) + list(range(a.train))
        run_sweeps(order, db, a.rows_min, a.rows_max, a.reps, time_budget_s=a.time_budget,
                   skip=tuple(k for k in a.skip.split(",") if k), retime=tuple(k for k in a.retime.split(",") if k))
        if a.sweep_only:
            return 0
    recs = load_db(db)
    train = [recs[i] for i in range(a.train) if i in recs]
    test = [recs[i] for i in range(a.train, n) if i in recs]
    print(f"[train] {len(train)} train / {len(test)} held-out records", flush=True)
    X, T, M = dataset(train)
    sel = tune.Selector(64, tune.N_INPUT, len(tune.CANDIDATES), seed=0)
    losses = sel.fit(X, T, M, epochs=a.epochs, lr=2e-3, batch=256)
    print(f"[train] loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
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
