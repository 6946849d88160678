"""Schedule search for large generated programs: the measured autotuner as
data source and judge, the learned cost model as the fast path.

  collect()   seeded generator DAGs of varied size over Zipf texts of varied
              length; for each, a sample of candidate schedules (partition
              knobs max_live / max_ops and tile knobs chunk_rows /
              rows_per_step / occupancy) is compiled (in parallel worker
              processes), checked bit-exact against the first candidate and
              timed with hipEvents; rows = (plan features, measured ms).
  evaluate()  on held-out programs: the full autotune over the candidate
              space (its wall time and best schedule) against the model's
              pick (plan-only features for every candidate, GPU inference,
              then compile + time of the top pick only): the pick's gap to
              the measured best, the time ratio, and the model's
              predicted-vs-measured error on every candidate.

Nothing here changes program outputs: every schedule is checked against the
others and the model only chooses which one to build.
"""

from __future__ import annotations

import itertools
import json
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import api, costmodel, dag_gen, tuning

SPACE_GRID = {"max_live": [24, 48, 72, 96], "max_ops": [250, 1000, 4000], "chunk_rows": [4, 16, 64],
              "rows_per_step": [1, 4], "occupancy": [0, 1]}


def full_space() -> list[dict]:
    keys = sorted(SPACE_GRID)
    return [dict(zip(keys, v)) for v in itertools.product(*(SPACE_GRID[k] for k in keys))]


def sample_space(n: int, seed: int) -> list[dict]:
    sp = full_space()
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = rng.choice(len(sp), size=min(n, len(sp)), replace=False)
    return [sp[i] for i in sorted(idx)]


@dataclass
class ProgramSpec:
    nodes: int
    n_bytes: int
    seed: int

    def build(self):
        return dag_gen.generate(self.nodes, seed=self.seed), dag_gen.text_case(self.n_bytes, seed=self.seed + 7)


def random_specs(n: int, seed: int, nodes=(300, 3000), log2_bytes=(20, 24)) -> list[ProgramSpec]:
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        nn = int(math.exp(rng.uniform(math.log(nodes[0]), math.log(nodes[1]))))
        nb = int(2 ** rng.uniform(*log2_bytes))
        out.append(ProgramSpec(nn, nb, seed * 1000 + i))
    return out


@dataclass
class Measured:
    options: list = field(default_factory=list)
    ms: list = field(default_factory=list)
    features: list = field(default_factory=list)
    compile_s: float = 0.0
    measure_s: float = 0.0


def measure(ctx: api.Context, source: str, case, candidates: list[dict], steps: int = 5, processes: int = 16,
            log=None, allow_empty: bool = False, shards: int = 0) -> Measured:
    """Compile every candidate (parallel processes, each plan's kernels in
    `shards` jobs), then check and time each."""
    out = Measured()
    t0 = time.perf_counter()
    tuning.precompile(source, [case], candidates, processes=processes, shards=shards, log=log)
    out.compile_s = time.perf_counter() - t0
    if log:
        log(f"measure: {len(candidates)} candidates compiled in {out.compile_s:.0f} s; checking and timing them")
    t1 = time.perf_counter()
    prog = api.Program(source)
    shape = tuning.shape_case([case])
    res = tuning.Resident(ctx, [case])
    names = [n for n in source.split("\n") if n.startswith("output stream ")]
    outs = [n.split()[-1] for n in names][:8]
    try:
        tr = tuning.autotune(ctx, prog, [res], candidates, [], outs, steps=steps, warmup=1, log=log,
                             allow_empty=allow_empty)
    finally:
        res.free()
    words = max(1, (case.bitlength + 63) // 64)
    # a candidate that partitions exactly like an earlier one runs its code
    timed = [(t.options, t.ms) for t in tr.trials]
    ms_of = {json.dumps(o, sort_keys=True): ms for o, ms in timed}
    for cand, first in tr.same_as:
        k = json.dumps(first, sort_keys=True)
        if k in ms_of:
            timed.append((cand, ms_of[k]))
    F = costmodel.features_batch(prog, shape, [o for o, _ in timed], words, 1) if timed else []
    for (o, ms), f in zip(timed, F):
        out.options.append(o)
        out.ms.append(ms)
        out.features.append(f)
    out.measure_s = time.perf_counter() - t1
    return out


def collect(ctx: api.Context, specs: list[ProgramSpec], per_program: int = 16, seed: int = 0, processes: int = 16,
            log=None, out: str = ""):
    """Training rows (features, ms, program index) from the measured tuner;
    with `out`, the rows so far are saved after every program."""
    X, y, gid = [], [], []
    for i, sp in enumerate(specs):
        p, case = sp.build()
        cands = sample_space(per_program, seed + i)
        m = measure(ctx, p.source, case, cands, processes=processes)
        X += m.features
        y += m.ms
        gid += [i] * len(m.ms)
        if out:
            np.savez(out, X=np.array(X, dtype=np.float32), ms=np.array(y), gid=np.array(gid),
                     features=np.array(costmodel.FEATURES))
        if log:
            log(f"collect {i + 1}/{len(specs)}: {sp.nodes} nodes, {sp.n_bytes} bytes, best {min(m.ms):.3f} ms, "
                f"worst {max(m.ms):.3f} ms (compile {m.compile_s:.1f} s, measure {m.measure_s:.1f} s)")
    return np.array(X, dtype=np.float32), np.array(y, dtype=np.float64), np.array(gid)


PICK_TOP = 3   # predicted-best candidates the model's pick compiles and times


def model_pick(ctx: api.Context, source: str, case, candidates: list[dict], pred: np.ndarray,
               processes: int = 16, top: int = PICK_TOP, log=None, shards: int = 0) -> "Measured":
    """The model's choice: compile and time the `top` best-predicted
    candidates that compile within their register file (a schedule the
    autotuner would accept; rejected ones are replaced by the next
    prediction) and keep the fastest.  Returns that schedule and its time,
    with the compile and measure seconds of the whole step."""
    order = [int(i) for i in np.argsort(pred)]
    got = Measured()
    pos = 0
    while pos < len(order) and len(got.ms) < top:
        batch = [candidates[i] for i in order[pos:pos + top - len(got.ms)]]
        pos += len(batch)
        m = measure(ctx, source, case, batch, processes=processes, allow_empty=True, log=log, shards=shards)
        got.options += m.options
        got.ms += m.ms
        got.compile_s += m.compile_s
        got.measure_s += m.measure_s
    if not got.ms:
        raise RuntimeError("no candidate schedule compiles within the register file")
    b = int(np.argmin(got.ms))
    return Measured([got.options[b]], [got.ms[b]], [], got.compile_s, got.measure_s)


@dataclass
class EvalRow:
    nodes: int
    n_bytes: int
    best_ms: float
    pick_ms: float
    default_ms: float
    tune_s: float
    model_s: float
    mape: float

    @property
    def gap(self) -> float:
        return self.pick_ms / self.best_ms - 1.0


def evaluate(ctx: api.Context, model: costmodel.CostModel, specs: list[ProgramSpec], candidates: list[dict],
             processes: int = 16, default: dict | None = None, log=None, pick_top: int = PICK_TOP) -> list[EvalRow]:
    """Learned pick against the full measured autotune, per held-out program."""
    default = default if default is not None else {}
    rows = []
    for sp in specs:
        p, case = sp.build()
        # the model first (so its compile finds nothing cached): plan-only
        # features of every candidate, GPU inference, then compile and time
        # only the top pick
        t1 = time.perf_counter()
        prog = api.Program(p.source)
        shape = tuning.shape_case([case])
        F = costmodel.features_batch(prog, shape, candidates, max(1, (case.bitlength + 63) // 64), 1)
        pred = model.predict_ms(F)
        feat_s = time.perf_counter() - t1
        pick = model_pick(ctx, p.source, case, candidates, pred, processes, top=pick_top)
        model_s = time.perf_counter() - t1
        # the autotuner: compile + check + time every candidate
        t0 = time.perf_counter()
        m = measure(ctx, p.source, case, candidates, processes=processes, log=log)
        tune_s = time.perf_counter() - t0
        best = min(m.ms)
        # the schedule without any search, for the ablation
        d = measure(ctx, p.source, case, [default], processes=processes, allow_empty=True)
        measured = {json.dumps(o, sort_keys=True): t for o, t in zip(m.options, m.ms)}
        pairs = [(pr, measured[json.dumps(c, sort_keys=True)]) for pr, c in zip(pred, candidates)
                 if json.dumps(c, sort_keys=True) in measured]
        mape = float(np.mean([abs(a - b) / b for a, b in pairs]))
        row = EvalRow(sp.nodes, sp.n_bytes, best, pick.ms[0], d.ms[0] if d.ms else float("nan"), tune_s, model_s,
                      mape)
        rows.append(row)
        if log:
            log(f"eval {sp.nodes} nodes / {sp.n_bytes} B: best {best:.3f} ms, pick {row.pick_ms:.3f} ms "
                f"(gap {100 * row.gap:.1f} %), default {row.default_ms:.3f} ms, tune {tune_s:.1f} s, "
                f"model {model_s:.1f} s (features + inference {feat_s:.2f} s, pick compile {pick.compile_s:.2f} s, "
                f"pick measure {pick.measure_s:.2f} s), MAPE {100 * mape:.1f} %")
    return rows


@dataclass
class MixedResult:
    n_programs: int
    bit_ops: int
    alg_bytes: int
    ms: float
    default_ms: float
    pick_s: float
    compile_s: float
    judged: list = field(default_factory=list)   # (program index, model pick ms, autotuned best ms)

    @property
    def gbitops(self) -> float:
        return self.bit_ops / (self.ms * 1e-3) / 1e9

    @property
    def gbps(self) -> float:
        return self.alg_bytes / (self.ms * 1e-3) / 1e9


def mixed_batch(ctx: api.Context, model: costmodel.CostModel, specs: list[ProgramSpec], candidates: list[dict],
                judge_frac: float = 0.05, steps: int = 10, processes: int = 16, seed: int = 0,
                log=None, pick_top: int = PICK_TOP) -> MixedResult:
    """A batch of different generated programs run as one captured graph:
    the model picks every program's schedule from plan features, all picks
    are compiled in parallel, every program's outputs are checked against the
    same program under the default schedule, the whole batch is replayed as
    one hipGraph, and the measured autotuner re-tunes a random `judge_frac`
    of the programs as judge."""
    built = [sp.build() for sp in specs]
    t0 = time.perf_counter()
    preds = []
    for (p, case) in built:
        prog = api.Program(p.source)
        F = costmodel.features_batch(prog, tuning.shape_case([case]), candidates, max(1, (case.bitlength + 63) // 64), 1)
        preds.append(model.predict_ms(F))
    pick_s = time.perf_counter() - t0
    picks = []
    for i, ((p, case), pr) in enumerate(zip(built, preds)):
        picks.append(model_pick(ctx, p.source, case, candidates, pr, processes, top=pick_top).options[0])
        if log:
            log(f"mixed: program {i + 1}/{len(built)} pick {picks[-1]}")
    t1 = time.perf_counter()
    for i, ((p, case), opt) in enumerate(zip(built, picks)):
        tuning.precompile(p.source, [case], [opt, {}], processes=processes)
        if log:
            log(f"mixed: compiled program {i + 1}/{len(built)}")
    compile_s = time.perf_counter() - t1
    residents, plans, defaults = [], [], []
    bit_ops = alg = 0
    try:
        for (p, case), opt in zip(built, picks):
            r = tuning.Resident(ctx, [case])
            residents.append(r)
            prog = api.Program(p.source)
            pl = api.Plan(ctx, prog, tuning.shape_case([case]), opt)
            r.bind(pl)
            plans.append(pl)
            d = api.Plan(ctx, prog, tuning.shape_case([case]), {})
            r.bind(d)
            defaults.append(d)
            bit_ops += p.n_dsl_ops * case.bitlength
            nw = (case.bitlength + 63) // 64
            alg += 8 * nw * (8 + len(p.outputs))
        for pl, d, (p, case) in zip(plans, defaults, built):
            pl.run()
            d.run()
            nw = (case.bitlength + 63) // 64
            for o in p.outputs:
                if not np.array_equal(pl.download_output(0, o, nw), d.download_output(0, o, nw)):
                    raise RuntimeError(f"schedule {pl.info()['kernels'][:1]} changed output {o}")
        ms = tuning.time_runs(ctx, plans, steps, 2) * len(plans)        # one batch = every program once
        default_ms = tuning.time_runs(ctx, defaults, steps, 2) * len(defaults)
    finally:
        for pl in plans + defaults:
            pl.close()
        for r in residents:
            r.free()
    res = MixedResult(len(specs), bit_ops, alg, ms, default_ms, pick_s, compile_s)
    rng = np.random.Generator(np.random.PCG64(seed))
    n_judge = max(1, int(round(judge_frac * len(specs))))
    for i in sorted(rng.choice(len(specs), size=n_judge, replace=False).tolist()):
        p, case = built[i]
        m = measure(ctx, p.source, case, candidates, processes=processes, log=log)
        k = m.options.index(picks[i])
        res.judged.append((i, float(m.ms[k]), float(min(m.ms))))
        if log:
            log(f"judge program {i} ({specs[i].nodes} nodes, {specs[i].n_bytes} B): model pick {m.ms[k]:.3f} ms, "
                f"autotuned best {min(m.ms):.3f} ms")
    return res


DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def evaluate_large(ctx: api.Context, model: costmodel.CostModel, nodes: int, log2_bytes: int,
                   candidates: list[dict], judge: int, processes: int = 16, seed: int = 0, log=None) -> dict:
    """The model on a program far larger than anything it was trained on
    (the config-4 generator DAG): its pick is the best-predicted candidate
    that compiles within its register file, found from plan features alone;
    the measured autotuner then times a subset of `judge` candidates (the
    pick, the benchmark's default schedule and random others) as judge."""
    t0 = time.perf_counter()
    p = dag_gen.generate(nodes, seed=4)
    case = dag_gen.text_case(1 << log2_bytes, seed=11)
    gen_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    prog = api.Program(p.source)
    F = costmodel.features_batch(prog, tuning.shape_case([case]), candidates, max(1, (case.bitlength + 63) // 64), 1)
    pred = model.predict_ms(F)
    predict_s = time.perf_counter() - t1
    if log:
        log(f"large: {nodes} nodes, features and predictions for {len(candidates)} candidates in {predict_s:.1f} s")
    t2 = time.perf_counter()
    # a plan of this size compiles in 64 jobs per schedule, so progress is
    # reported every few seconds
    pick = model_pick(ctx, p.source, case, candidates, pred, processes, top=1, log=log, shards=64)
    model_s = predict_s + (time.perf_counter() - t2)
    pick_opts, pick_ms = pick.options[0], float(pick.ms[0])
    if log:
        log(f"large: pick {pick_opts} {pick_ms:.1f} ms ({model_s:.1f} s with its compile and timing)")
    rng = np.random.Generator(np.random.PCG64(seed))
    default = {"warmup_rows": 4}
    others = [candidates[i] for i in rng.permutation(len(candidates)).tolist()
              if candidates[i] not in (pick_opts, default)]
    trial = [pick_opts, default] + others[:max(0, judge - 2)]
    t3 = time.perf_counter()
    m = measure(ctx, p.source, case, trial, processes=processes, log=log, allow_empty=True, shards=64)
    tune_s = time.perf_counter() - t3
    by = {json.dumps(o, sort_keys=True): ms for o, ms in zip(m.options, m.ms)}
    best = min(m.ms) if m.ms else pick_ms
    default_ms = by.get(json.dumps(default, sort_keys=True))
    return {"nodes": nodes, "dsl_ops": p.n_dsl_ops, "bytes": 1 << log2_bytes, "candidates": len(candidates),
            "pick": pick_opts, "pick_ms": round(pick_ms, 2),
            "judged": [{"options": o, "ms": round(float(ms), 2)} for o, ms in zip(m.options, m.ms)],
            "best_judged_ms": round(float(best), 2), "gap_pct": round(100 * (pick_ms / best - 1), 2),
            "default_ms": round(float(default_ms), 2) if default_ms is not None else None,
            "model_s": round(model_s, 1), "judge_s": round(tune_s, 1), "generate_s": round(gen_s, 1),
            "predicted_pick_ms": round(float(np.min(pred)), 2)}


def load_training_data():
    """Every committed autotuner measurement file (data/schedule_runtimes*.npz)."""
    import glob
    X, ms = [], []
    for f in sorted(glob.glob(os.path.join(DATA_DIR, "schedule_runtimes*.npz"))):
        d = np.load(f)
        X.append(d["X"])
        ms.append(d["ms"])
    if not X:
        raise FileNotFoundError(f"no schedule_runtimes*.npz under {DATA_DIR}")
    return np.concatenate(X), np.concatenate(ms)


def main(argv=None) -> int:
    """python -m bitstreamzoo_amd.schedsearch --out FILE.npz [--programs N] [--per-program K]
    [--eval N] [--candidates K] [--processes P]: collect training rows with the
    measured autotuner, train the cost model on the GPU, evaluate it on
    held-out programs and print one JSON summary line."""
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="")
    ap.add_argument("--data", default="", help="reuse collected rows instead of collecting")
    ap.add_argument("--programs", type=int, default=40)
    ap.add_argument("--per-program", type=int, default=16)
    ap.add_argument("--eval", type=int, default=4)
    ap.add_argument("--candidates", type=int, default=40)
    ap.add_argument("--processes", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--min-nodes", type=int, default=300)
    ap.add_argument("--max-nodes", type=int, default=3000)
    ap.add_argument("--min-log2-bytes", type=float, default=20)
    ap.add_argument("--max-log2-bytes", type=float, default=24)
    ap.add_argument("--collect-only", action="store_true")
    ap.add_argument("--mixed", type=int, default=0, help="run a mixed batch of N programs instead of --eval")
    ap.add_argument("--large", type=int, default=0,
                    help="evaluate on one generated DAG of N nodes (the config-4 generator) instead of --eval")
    ap.add_argument("--large-log2-bytes", type=int, default=27)
    ap.add_argument("--judge", type=int, default=4, help="candidates the autotuner times as judge with --large")
    ap.add_argument("--pick-top", type=int, default=PICK_TOP,
                    help="best-predicted candidates the model's pick compiles and times")
    a = ap.parse_args(argv)

    def log(m):
        print(f"[schedsearch] {m}", file=sys.stderr, flush=True)

    ctx = api.Context(0)
    if a.data == "committed":
        X, y = load_training_data()
        gid = np.zeros(len(y), dtype=np.int64)
    elif a.data:
        d = np.load(a.data)
        X, y, gid = d["X"], d["ms"], d["gid"]
    else:
        specs = random_specs(a.programs, a.seed, nodes=(a.min_nodes, a.max_nodes),
                             log2_bytes=(a.min_log2_bytes, a.max_log2_bytes))
        X, y, gid = collect(ctx, specs, a.per_program, seed=a.seed, processes=a.processes, log=log, out=a.out)
        if a.out:
            np.savez(a.out, X=X, ms=y, gid=gid, features=np.array(costmodel.FEATURES))
        if a.collect_only:
            ctx.close()
            return 0
    model = costmodel.CostModel(ctx, seed=0)
    loss = model.fit(X, y)
    log(f"trained on {len(y)} rows in {model.train_seconds:.2f} s, loss {loss:.4f}")
    cands = sample_space(a.candidates, 12345)
    if a.large:
        r = evaluate_large(ctx, model, a.large, a.large_log2_bytes, cands, a.judge, processes=a.processes,
                           seed=a.seed, log=log)
        r.update({"train_rows": int(len(y)), "train_s": round(model.train_seconds, 2)})
        print(json.dumps(r), flush=True)
        ctx.close()
        return 0
    if a.mixed:
        specs = random_specs(a.mixed, a.seed + 200, nodes=(500, 5000), log2_bytes=(16, 22))
        r = mixed_batch(ctx, model, specs, cands, processes=a.processes, seed=a.seed, log=log, pick_top=a.pick_top)
        print(json.dumps({"programs": r.n_programs, "ms_per_batch": round(r.ms, 4),
                          "default_schedules_ms": round(r.default_ms, 4), "Gbit_ops_per_s": round(r.gbitops, 1),
                          "GBps": round(r.gbps, 1), "hbm_frac": round(r.gbps / 8000.0, 4),
                          "model_pick_s": round(r.pick_s, 2), "compile_s": round(r.compile_s, 1),
                          "judged": [{"program": i, "pick_ms": round(a_, 4), "best_ms": round(b_, 4),
                                      "gap_pct": round(100 * (a_ / b_ - 1), 2)} for i, a_, b_ in r.judged]}),
              flush=True)
        ctx.close()
        return 0
    held = random_specs(a.eval, a.seed + 100, nodes=(2000, 5000), log2_bytes=(23, 25))
    rows = evaluate(ctx, model, held, cands, processes=a.processes, log=log, pick_top=a.pick_top)
    summary = {
        "train_rows": int(len(y)), "train_programs": int(len(set(gid.tolist()))), "train_s": round(model.train_seconds, 2),
        "candidates": len(cands),
        "gap_pct": [round(100 * r.gap, 2) for r in rows],
        "time_ratio_pct": [round(100 * r.model_s / r.tune_s, 2) for r in rows],
        "mape_pct": [round(100 * r.mape, 1) for r in rows],
        "pick_vs_default_speedup": [round(r.default_ms / r.pick_ms, 3) for r in rows],
        "programs": [{"nodes": r.nodes, "bytes": r.n_bytes, "best_ms": round(r.best_ms, 4), "pick_ms": round(r.pick_ms, 4),
                      "default_ms": round(r.default_ms, 4), "tune_s": round(r.tune_s, 1), "model_s": round(r.model_s, 2)}
                     for r in rows],
    }
    print(json.dumps(summary), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
