"""GPU parity on the named workloads at their full sizes.

* Batched: 256 independent CSR matrices, rows log-uniform in [1e3, 2e5], mean nonzeros per row log-uniform in
  [4, 64], Poisson row lengths, uniform sorted columns (seed 5, the batched benchmark's matrices). With integer
  values in [-3, 3] and x in [-4, 4] every member's y equals the host's integer product bit for bit, eagerly and
  from hipGraph replays; with the generated U[-1, 1] values every member is within rel L2 1e-6 of the FP64 oracle,
  in both modes (merge tiles in one persistent launch; the slab mode, the block-diagonal matrix through an LDS
  column-slab plan), the default (the slab mode) runs faster than the tuned merge schedule, and its hipGraph at
  least 2x faster than one CSR-vector launch per matrix (a graph too).
* Precision sweep, power-law leg: the 2M x 2M power-law matrix (exponent 2, lengths 1..100,000, rows shuffled)
  with values stored FP64, FP32 and BF16, x FP32, FP32 and Dot2 accumulation: rel L2 <= 1e-6 against the FP64
  oracle on the stored values, Dot2 no worse than plain, BF16 storage error within its unit roundoff; AUTO's
  kernel there, plain and Dot2; and the checked plan (per-row FP64 fallback) with every row within rtol.
* Tuned schedules: the batched plan's default schedule and SpMM's default items per lane group (the autotuner's
  tables) re-timed against every candidate on the 256-matrix batch and the power-law matrix at k = 8 and 32.
* Learned selector: a seeded handful of the held-out matrices of the committed sweep DB, regenerated on the
  device. Their features equal the recorded ones, the library's AUTO pick equals the pick the shipped network
  makes from the recorded features (the FP64 oracle network), and re-timed cold the pick is within 15 % (or 2 us)
  of the best recorded candidate re-timed beside it.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from oracle import mlp as omlp
from muxi_spmv_amd import api, batch, selector_train, tune
from muxi_spmv_amd import generators as G

pytestmark = pytest.mark.gpu

PKG = Path(__file__).resolve().parent.parent / "muxi_spmv_amd"


# ---------------------------------------------------------------- batched, 256 matrices
@pytest.fixture(scope="module")
def batch_256():
    rng = np.random.default_rng(5)
    mats, xs, ys = [], [], []
    for i in range(256):
        R = int(math.exp(rng.uniform(math.log(1e3), math.log(2e5))))
        mean = math.exp(rng.uniform(math.log(4), math.log(64)))
        mats.append(tune.gen_matrix("poisson", R, R, mean, seed=5000 + i))
        xs.append(tune.gen_vector(R, seed=7000 + i))
        ys.append(torch.full((R,), float("nan"), device="cuda"))
    print(f"batch_256: {sum(m.numRows for m in mats)} rows, {sum(m.nnz for m in mats)} nnz", flush=True)
    return mats, xs, ys


def _int_members(mats):
    """Per member: integer values as a device CSR view, integer x, and the host integer product."""
    out = []
    for m, A in enumerate(mats):
        rp, ci, _ = A.host()
        i = np.arange(ci.size, dtype=np.int64)
        vi = ((i * 2654435761 + m) >> 7) % 7 - 3
        xi = (np.arange(A.numCols, dtype=np.int64) * 40503 + m >> 3) % 9 - 4
        e = np.zeros(A.numRows, np.int64)
        if ci.size:
            rows = np.repeat(np.arange(A.numRows), np.diff(rp))
            e = np.bincount(rows, weights=(vi * xi[ci]).astype(np.float64), minlength=A.numRows).astype(np.int64)
        Ai = api.CSRMatrix(torch.from_numpy(rp), torch.from_numpy(ci), torch.from_numpy(vi.astype(np.float32)),
                           A.numCols)
        out.append((Ai, torch.from_numpy(xi.astype(np.float32)).cuda(), e))
    return out


MODES = {"merge": batch.TUNED_MERGE, "slabs": batch.BATCH_SLABS}


@pytest.mark.parametrize("mode", MODES)
def test_batch_256_integers_exact_eager_and_graph(batch_256, mode):
    mats, _, ys = batch_256
    members = _int_members(mats)
    bp = batch.BatchPlan([a for a, _, _ in members], [x for _, x, _ in members], ys, MODES[mode])
    info = bp.info()
    assert info["mode"] == mode and (info["tiles"] > 0) == (mode == "merge")
    bp.exec()
    torch.cuda.synchronize()
    for (_, _, e), y in zip(members, ys):
        assert np.array_equal(y.cpu().numpy(), e.astype(np.float32))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g = bp.graph(stream=s)
        gv = bp.graph("vector", 0, alpha=2.0, beta=0.0, stream=s)
    torch.cuda.synchronize()
    for rep in range(3):
        for y in ys:
            y.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        for (_, _, e), y in zip(members, ys):
            assert np.array_equal(y.cpu().numpy(), e.astype(np.float32)), rep
        gv.launch(s)  # the per-matrix graph honours alpha
        s.synchronize()
        for (_, _, e), y in zip(members, ys):
            assert np.array_equal(y.cpu().numpy(), (2 * e).astype(np.float32)), rep


@pytest.mark.parametrize("mode", MODES)
def test_batch_256_rel_l2_vs_fp64_oracle(batch_256, mode):
    mats, xs, ys = batch_256
    for y in ys:
        y.fill_(float("nan"))
    bp = batch.BatchPlan(mats, xs, ys, MODES[mode])
    bp.exec()
    torch.cuda.synchronize()
    worst = 0.0
    for m, (A, x, y) in enumerate(zip(mats, xs, ys)):
        y64 = api.oracle_f64(A, x)
        err = float(torch.linalg.vector_norm(y.double() - y64) / torch.linalg.vector_norm(y64))
        worst = max(worst, err)
        assert err <= 1e-6, (m, err)
    print(f"batch_256 {mode} worst rel L2 {worst:.3g}", flush=True)


def test_batch_per_matrix_graph_alpha_beta_and_bad_kernel(batch_256):
    mats, xs, ys = batch_256
    k = 24  # a slice keeps this quick
    sub_m, sub_x, sub_y = mats[:k], xs[:k], ys[:k]
    bp = batch.BatchPlan(sub_m, sub_x, sub_y)
    for y in sub_y:
        y.fill_(1.0)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g = bp.graph("vwave", 8, alpha=2.0, beta=0.5, stream=s)
    torch.cuda.synchronize()
    g.launch(s)
    s.synchronize()
    for A, x, y in zip(sub_m, sub_x, sub_y):
        y64 = api.oracle_f64(A, x)
        ref = 2.0 * y64 + 0.5
        err = float(torch.linalg.vector_norm(y.double() - ref) / torch.linalg.vector_norm(ref))
        assert err <= 1e-6, err
    with pytest.raises(api.SpmvError, match="INTERNAL"):
        with torch.cuda.stream(s):
            bp.graph("merge", 4, stream=s)
    with pytest.raises(api.SpmvError, match="INTERNAL"):
        bp.exec_per_matrix("csr5", 256)


def _interleaved_cold(tm, fns: dict, rounds: int = 3, reps: int = 5) -> dict:
    """Cold median of every candidate, timed round-robin over several rounds (min over the rounds), so a clock
    or thermal drift between one candidate's run and another's cannot decide the comparison."""
    best = {k: math.inf for k in fns}
    for _ in range(rounds):
        for k, fn in fns.items():
            best[k] = min(best[k], tm.cold(fn, reps))
    return best


def test_batch_256_tuned_schedule_within_5pct_of_sweep(batch_256):
    """The merge schedule the autotuner's table picks against every merge candidate, re-timed cold in the same
    rounds: within 5 % of the best."""
    from muxi_spmv_amd import sched_tune

    mats, xs, ys = batch_256
    tm = sched_tune._Timer()
    bp = batch.BatchPlan(mats, xs, ys, batch.TUNED_MERGE)
    info = bp.info()
    pick = (info["ipt"], info["wg_per_cu"])
    plans = {(ipt, wg): batch.BatchPlan(mats, xs, ys, ipt | (wg << 8)) for ipt, wg in sched_tune.BATCH_CANDIDATES}
    fns = {"pick": bp.exec}
    fns.update({k: b.exec for k, b in plans.items()})
    t = _interleaved_cold(tm, fns)
    for b in plans.values():
        b.destroy()
    bp.destroy()
    t_pick = t.pop("pick")
    best = min(t, key=t.get)
    print(f"batch_256 schedule: tuned {pick} {t_pick:.4f} ms, sweep best {best} {t[best]:.4f} ms", flush=True)
    assert t_pick <= 1.05 * t[best], (pick, t_pick, best, t)


def test_batch_256_graph_beats_per_matrix_vector_graph_2x(batch_256):
    """The batched throughput target: the persistent launch against one CSR-vector launch per matrix, each captured
    as a hipGraph on the same stream, cold, timed round-robin: at least 2x faster (3.4x measured). Every member's
    rel L2 against the FP64 oracle is checked in test_batch_256_rel_l2_vs_fp64_oracle."""
    from muxi_spmv_amd import sched_tune

    mats, xs, ys = batch_256
    bp = batch.BatchPlan(mats, xs, ys)  # the default: the slab mode
    assert bp.info()["mode"] == "slabs"
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        tm = sched_tune._Timer()  # its events and cache flushes on s, with the graphs
        g = bp.graph(stream=s)
        gv = bp.graph("vector", 0, stream=s)
        t = _interleaved_cold(tm, {"batched": lambda: g.launch(s), "per_matrix": lambda: gv.launch(s)}, reps=3)
    torch.cuda.synchronize()
    del g, gv  # the graphs before the plan they launch
    bp.destroy()
    print(f"batch_256 graphs: persistent {t['batched']:.4f} ms, per-matrix vector {t['per_matrix']:.4f} ms "
          f"({t['per_matrix'] / t['batched']:.2f}x)", flush=True)
    assert t["per_matrix"] >= 2.0 * t["batched"], t


def test_batch_256_slab_mode_beats_tuned_merge(batch_256):
    """Why the default takes the slab mode for this batch: it runs faster than the tuned merge schedule, both
    timed cold round-robin (2.3x measured in the launches themselves)."""
    from muxi_spmv_amd import sched_tune

    mats, xs, ys = batch_256
    tm = sched_tune._Timer()
    slabs = batch.BatchPlan(mats, xs, ys)
    merge = batch.BatchPlan(mats, xs, ys, batch.TUNED_MERGE)
    assert slabs.info()["mode"] == "slabs" and merge.info()["mode"] == "merge"
    t = _interleaved_cold(tm, {"slabs": slabs.exec, "merge": merge.exec}, reps=3)
    slabs.destroy()
    merge.destroy()
    print(f"batch_256: slab mode {t['slabs']:.4f} ms, tuned merge {t['merge']:.4f} ms", flush=True)
    assert t["slabs"] < t["merge"], t


@pytest.mark.parametrize("k", [8, 32])
def test_spmm_powerlaw_tuned_ipt_within_5pct_of_sweep(k):
    from muxi_spmv_amd import sched_tune

    h = G.cfg_powerlaw_2m()
    A = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values), h.numCols)
    X = torch.rand((h.numCols, k), device="cuda") * 2 - 1
    Y = torch.empty((h.numRows, k), device="cuda")
    tm = sched_tune._Timer()
    p = api.Plan.spmm(A, k)
    pick = p.info().param0
    plans = {ipt: api.Plan.spmm(A, k, ipt) for ipt in sched_tune.SPMM_IPT}
    fns = {"pick": lambda: p.exec(X, Y)}
    fns.update({ipt: (lambda q=q: q.exec(X, Y)) for ipt, q in plans.items()})
    t = _interleaved_cold(tm, fns)
    for q in plans.values():
        q.destroy()
    t_pick = t.pop("pick")
    best = min(t, key=t.get)
    print(f"spmm k={k}: tuned {pick} {t_pick:.4f} ms, sweep best {best} {t[best]:.4f} ms", flush=True)
    assert t_pick <= 1.05 * t[best], (pick, t_pick, best, t)


# ---------------------------------------------------------------- precision sweep, power-law 2M
@pytest.fixture(scope="module")
def powerlaw_2m():
    h = G.cfg_powerlaw_2m()
    x = G.uniform_vector(h.numCols, 2)
    A64 = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values.astype(np.float64)), h.numCols)
    return h, x, A64


def _rel_l2(y, y64):
    return float(np.linalg.norm(y - y64) / np.linalg.norm(y64))


@pytest.mark.parametrize("vt", [torch.float64, torch.float32, torch.bfloat16])
def test_powerlaw_2m_precision(powerlaw_2m, vt):
    h, x, A64 = powerlaw_2m
    xd = torch.from_numpy(x).cuda()
    Av = A64.with_values(api.convert_values(A64.d_values, vt))
    stored = Av.d_values.double().cpu().numpy()  # the stored (rounded) values, as FP64
    y64v, _ = oracle.spmv_f64(h.rowPtr, h.colIdx, stored, x.astype(np.float64))
    y64, _ = oracle.spmv_f64(h.rowPtr, h.colIdx, h.values.astype(np.float64), x.astype(np.float64))
    errs = {}
    for kernel, param, comp in (("auto", 0, False), ("auto", 0, True), ("merge", 4, False), ("merge", 4, True),
                                ("xslab", 1032, False), ("lslab", 8, False), ("lslab", 8, True)):
        y = torch.full((h.numRows,), float("nan"), device="cuda")
        api.Plan(Av, kernel, param, compensated=comp).exec(xd, y)
        torch.cuda.synchronize()
        yy = y.double().cpu().numpy()
        errs[(kernel, comp)] = e = _rel_l2(yy, y64v)
        assert e <= 1e-6, (kernel, comp, e)
        if vt == torch.bfloat16:
            assert _rel_l2(yy, y64) <= 2.0 ** -8  # BF16 storage rounding dominates
    for kernel in ("auto", "merge", "lslab"):
        assert errs[(kernel, True)] <= errs[(kernel, False)] * 1.01, (kernel, errs)
    print(f"powerlaw_2m {vt}: " + ", ".join(f"{k}{'+dot2' if c else ''} {v:.3g}" for (k, c), v in errs.items()),
          flush=True)


@pytest.mark.parametrize("comp", [False, True])
def test_powerlaw_2m_auto_plan_kernel(powerlaw_2m, comp):
    """AUTO on the power-law matrix takes the LDS slab kernel, with Dot2 accumulation too (the compensated
    variant keeps it, with half-width slabs, rather than falling back to a kernel that gathers from L2)."""
    h, _, _ = powerlaw_2m
    A = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values), h.numCols)
    info = api.Plan(A, "auto", 0, compensated=comp).info()
    assert (info.kernel, info.param0) == ("lslab", 4), info


@pytest.mark.parametrize("comp", [False, True])
def test_powerlaw_2m_checked_plan_every_row_within_rtol(powerlaw_2m, comp):
    """The checked plan on the full power-law matrix (random signs, rows up to 90,181 nonzeros): every row of y
    within rtol of the FP64 oracle -- the rows whose bound fails are the ones recomputed in FP64."""
    h, x, A64 = powerlaw_2m
    A = api.CSRMatrix(h.rowPtr, h.colIdx, torch.from_numpy(h.values), h.numCols)
    xd = torch.from_numpy(x).cuda()
    rtol = 1e-3
    p = api.Plan.checked(A, rtol, compensated=comp)
    y = torch.full((h.numRows,), float("nan"), device="cuda")
    p.exec(xd, y)
    y64 = api.oracle_f64(A, xd)
    torch.cuda.synchronize()
    rows = p.checked_rows()
    err = (y.double() - y64).abs()
    ok = err <= rtol * y64.abs() * 1.0001
    bad = torch.nonzero(~ok).flatten()[:5].tolist()
    assert bool(ok.all()), (bad, [float(err[i]) for i in bad], [float(y64[i]) for i in bad])
    lens = np.diff(h.rowPtr.astype(np.int64))
    print(f"powerlaw_2m checked {'dot2' if comp else 'fp32'} rtol {rtol}: {rows.size} rows recomputed in FP64 "
          f"({int(lens[rows].sum()) if rows.size else 0} nonzeros)", flush=True)
    assert rows.size < h.numRows // 10


# ---------------------------------------------------------------- learned selector spot-check
def _shipped():
    st = json.loads((PKG / "selector_weights.json").read_text())
    recs = selector_train.load_db(PKG / "selector_db.jsonl.gz")
    return st, recs


def _network_pick(st, rec, fp64):
    f = np.array(rec["features"])
    X = tune.feature_vector(f, fp64)[None, :].astype(np.float64)
    Xn = (X - np.array(st["mu"])) / np.array(st["sd"])
    Y, _ = omlp.forward(np.array(st["params"], np.float64), Xn, st["n_in"], st["hidden"], st["n_out"])
    y = np.where(tune.feasible_mask(f, fp64) > 0, Y[0], np.inf)
    return st["candidates"][int(np.argmin(y))]


@pytest.mark.parametrize("i", [2000, 2001, 2002, 2003, 2004, 2011, 2016, 2017])
def test_selector_held_out_spot_check(i):
    st, recs = _shipped()
    rec = recs[i]
    spec = selector_train.matrix_spec(i, 1e4, 4e6)
    assert spec["family"] == rec["family"] and spec["rows"] == rec["rows"]
    A = selector_train.build(spec)
    try:
        f = tune.features(A)
        assert A.nnz == rec["nnz"]
        assert np.allclose(f, np.array(rec["features"]), rtol=1e-9, atol=0), (f, rec["features"])
        pick = tune.library_pick(A)
        assert pick == _network_pick(st, rec, False), (pick, rec["times"])
        best = min(rec["times"], key=rec["times"].get)
        x = tune.gen_vector(A.numCols, seed=rec["seed"])
        y = torch.empty(A.numRows, device="cuda")
        fl = tune.Flusher()
        t = {}
        for name in {pick, best}:
            k, p = name.split(":")
            plan = api.Plan(A, k, int(p))
            tune.time_plan_cold(plan, x, y, 2, fl)
            t[name] = tune.time_plan_cold(plan, x, y, 7, fl)
            plan.destroy()
        regret = t[pick] / t[best] - 1.0
        print(f"held-out {i} ({rec['family']}, {rec['rows']} rows): pick {pick} best {best} "
              f"re-timed regret {regret:.3f}, recorded {rec['times'][pick] / rec['times'][best] - 1:.3f}", flush=True)
        assert regret <= 0.15 or t[pick] - t[best] <= 0.002, (pick, best, t)  # 15 %, or 2 us on small ones
    finally:
        A.free()
