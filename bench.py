"""Benchmark of the fused MI355X bitstream executor.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload adder32|matchstar|shift_and]
                    [--no-cpu] [--no-perop] [--no-extra]

A step is one pass of the program over one batch of synthetic input already
resident in HBM.  The headline workload is the bitsliced 32-bit ripple-carry
adder DAG (65 input streams of 2^28 bits, 193 DSL ops, 33 outputs and 33
popcounts); MatchStar on 2^26 bits and the batched Shift-And of 2048 uneven
sequences are measured beside it.  Each is timed with hipEvents on the
executor's stream over K replays of a captured hipGraph, next to the plain
one-kernel-per-op executor on the same resident buffers, and its outputs are
checked bit-for-bit against an independent host computation at full size.
The CPU baseline is the AVX-512 + OpenMP CPU executor (oracle/) on a bounded
sample of the same workload.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from bitstreamzoo_amd import api, tuning, workloads  # noqa: E402
from bitstreamzoo_amd.bsdata import BsCase  # noqa: E402

HBM_PEAK_GBPS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "adder32_fused_hbm_traffic.md")


def profiled_traffic():
    """HBM bytes per launch of the fused adder kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None
METRIC = "stream Gbit-ops/s + % HBM roofline; learned vs autotuned schedule time"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def words_popcount(w: np.ndarray) -> int:
    return int(np.unpackbits(w.view(np.uint8)).sum())


def time_plans(ctx, plans, steps, warmup, executor, launches_per_run=1):
    return tuning.time_runs(ctx, plans, steps, warmup, executor, MAX_GRAPH_LAUNCHES, launches_per_run)


MAX_GRAPH_LAUNCHES = 20000


def tuned_options(ctx, args, name, program, residents, kind, int_names, out_names, default):
    """The autotuner's pick over `kind`'s candidate grid (or `default` with
    --no-tune), and the tuning record."""
    if args.schedule:
        return json.loads(args.schedule), None
    if args.no_tune:
        return default, None
    log(f"{name}: autotuning ({kind} grid)")
    space = tuning.default_space(kind)
    tr = tuning.autotune(ctx, program, residents, space, int_names, out_names, steps=max(3, args.steps // 2),
                         warmup=1, log=log)
    best = tr.best
    log(f"{name}: best {best.options} {best.ms:.4f} ms ({tr.seconds:.1f} s, {len(tr.trials)} candidates)")
    return best.options, tr.as_json()


# ── full-size host checks (independent of the executors) ──

def adder_expected(case: BsCase, width: int):
    a = [case.inputs[f"a{i}"] for i in range(width)]
    b = [case.inputs[f"b{i}"] for i in range(width)]
    c = case.inputs["cin"]
    outs = {}
    for i in range(width):
        p = a[i] ^ b[i]
        outs[f"s{i}"] = p ^ c
        c = (a[i] & b[i]) | (p & c)
    outs["cout"] = c
    return outs


def matchstar_expected(case: BsCase):
    bl = case.bitlength
    m = (1 << bl) - 1
    M = int.from_bytes(case.inputs["M"].tobytes(), "little")
    Cc = int.from_bytes(case.inputs["C"].tobytes(), "little")
    r = ((((M & Cc) + Cc) & m) ^ Cc) | M
    return np.frombuffer(r.to_bytes(len(case.inputs["M"]) * 8, "little"), dtype=np.uint64)


def shift_and_expected_popcount(case: BsCase, pattern: str) -> int:
    bl = case.bitlength
    m = (1 << bl) - 1
    acc = m
    for i, ch in enumerate(pattern):
        v = int.from_bytes(case.inputs[f"c{ch}"].tobytes(), "little")
        acc &= (v << (len(pattern) - 1 - i)) & m
    return bin(acc).count("1")


# ── workloads ──

class Bound:
    """One plan per resident input set, built with `options`."""

    def __init__(self, ctx, program, residents, options):
        self.plans = []
        for r in residents:
            p = api.Plan(ctx, program, tuning.shape_case(r.cases), options)
            r.bind(p)
            self.plans.append(p)

    def free(self):
        for p in self.plans:
            p.close()


def bench_adder(ctx, args, perop=True):
    bl = args.bitlength or (1 << 28)
    log(f"adder32: generating 65 x {bl} bits")
    w = workloads.adder32(bl, seed=2)
    case = w.cases[0]
    prog = api.Program(w.source)
    res = tuning.Resident(ctx, [case])
    outs = [f"s{i}" for i in range(32)] + ["cout"]
    opts, tune = tuned_options(ctx, args, "adder32", prog, [res], "wide", ["n_" + o for o in outs], outs,
                               {"rows_per_step": 1, "occupancy": 2})
    B = Bound(ctx, prog, [res], opts)
    plan = B.plans[0]
    plan.run()
    ctx.sync()
    log("adder32: checking all 33 outputs and popcounts against the host computation")
    exp = adder_expected(case, 32)
    nw = len(case.inputs["a0"])
    ok = True
    for name, e in exp.items():
        got = plan.download_output(0, name, nw)
        if not np.array_equal(got, e):
            ok = False
            log(f"MISMATCH in output {name}")
        if plan.read_int(0, "n_" + name) != words_popcount(e):
            ok = False
            log(f"MISMATCH in popcount n_{name}")
    if not ok:
        raise SystemExit("adder32 outputs differ from the host computation")
    ms = time_plans(ctx, [plan], args.steps, args.warmup, api.FUSED)
    log(f"adder32: fused {ms:.4f} ms/step")
    out = {"ms_per_step": ms, "bit_ops": w.bit_ops, "alg_bytes": w.alg_bytes, "plan": plan.info(),
           "options": opts, "tune": tune}
    if perop:
        out["perop_ms_per_step"] = time_plans(ctx, [plan], max(2, args.steps // 4), 1, api.PEROP)
        log(f"adder32: per-op {out['perop_ms_per_step']:.4f} ms/step")
    B.free()
    res.free()
    return w, out


def bench_matchstar(ctx, args):
    bl = 1 << 26
    rot = 12  # distinct input sets, so the 256 MB last-level cache cannot serve them
    log(f"matchstar: generating {rot} x 2 x {bl} bits")
    ws = [workloads.matchstar(bl, seed=100 + r) for r in range(rot)]
    prog = api.Program(ws[0].source)
    residents = [tuning.Resident(ctx, w.cases) for w in ws]
    opts, tune = tuned_options(ctx, args, "matchstar", prog, residents, "stateful", ["n"], ["result"],
                               {"chunk_rows": 16, "rows_per_step": 8})
    B = Bound(ctx, prog, residents, opts)
    for p in B.plans:
        p.run()
    ctx.sync()
    for p, w in zip(B.plans, ws):
        e = matchstar_expected(w.cases[0])
        got = p.download_output(0, "result", len(e))
        if not np.array_equal(got, e) or p.read_int(0, "n") != words_popcount(e):
            raise SystemExit("matchstar output differs from the host computation")
    ms = time_plans(ctx, B.plans, args.steps, args.warmup, api.FUSED)
    unproven = sum(p.unproven(0) for p in B.plans)
    log(f"matchstar: fused {ms:.4f} ms/step ({unproven} chunks left unproven by their warm-up)")
    perop = None if args.no_perop else time_plans(ctx, B.plans, max(2, args.steps // 4), 1, api.PEROP)
    if perop is not None:
        log(f"matchstar: per-op {perop:.4f} ms/step")
    info = B.plans[0].info()
    B.free()
    floor_ms, floor_opts = one_pass_floor(ctx, args, residents, ws[0].cases[0].bitlength)
    log(f"matchstar: one-launch floor (r = M & C over the same inputs) {floor_ms:.4f} ms/step {floor_opts}")
    for r in residents:
        r.free()
    out = {"ms_per_step": ms, "bit_ops": ws[0].bit_ops,
           "alg_bytes": ws[0].alg_bytes, "plan": info, "input_sets": rot, "options": opts, "tune": tune,
           "unproven_chunks": unproven, "floor_ms": floor_ms, "floor_options": floor_opts}
    if perop is not None:
        out["perop_ms_per_step"] = perop
    return ws[0], out


FLOOR_SRC = """input stream M
input stream C
output stream r
r = M & C
"""


def one_pass_floor(ctx, args, residents, bitlength):
    """ms per step of the cheapest one-launch pass that moves MatchStar's
    bytes: r = M & C over the same rotating input sets (two streams read, one
    written, no carry, no warm-up).  A program of this size cannot run
    faster than one such launch; the best of a few tile shapes, with cached
    and nontemporal loads and stores and with 16-byte lanes, is kept."""
    prog = api.Program(FLOOR_SRC)
    best = None
    for opts in ({"chunk_rows": 4, "rows_per_step": 4}, {"chunk_rows": 8, "rows_per_step": 8},
                 {"chunk_rows": 16, "rows_per_step": 8}, {"chunk_rows": 16, "rows_per_step": 4, "interleave": 1},
                 {"chunk_rows": 4, "rows_per_step": 4, "nontemporal": 2},
                 {"chunk_rows": 8, "rows_per_step": 8, "nontemporal": 2},
                 {"chunk_rows": 8, "rows_per_step": 8, "words_per_lane": 2, "nontemporal": 2}):
        plans = []
        for r in residents:
            p = api.Plan(ctx, prog, tuning.shape_case(r.cases), opts)
            r.bind(p)
            plans.append(p)
        ms = time_plans(ctx, plans, args.steps, args.warmup, api.FUSED)
        for p in plans:
            p.close()
        if best is None or ms < best[0]:
            best = (ms, opts)
    return best


def bench_shift_and(ctx, args):
    n = 2048
    log(f"shift_and: generating {n} sequences, lengths log-uniform in [2^12, 2^22]")
    w = workloads.shift_and_batch(n_seqs=n, seed=3)
    prog = api.Program(w.source)
    res = tuning.Resident(ctx, w.cases)
    opts, tune = tuned_options(ctx, args, "shift_and", prog, [res], "small", ["n"], ["match"],
                               {"chunk_rows": 16, "rows_per_step": 4})
    B = Bound(ctx, prog, [res], opts)
    plan = B.plans[0]
    log(f"shift_and: plan bound over {n} instances ({plan.info()['n_chunks']} chunks); running")
    plan.run()
    ctx.sync()
    pattern = w.source.split("the 32-mer ")[1].split(":")[0]
    rng = np.random.Generator(np.random.PCG64(7))
    check_idx = sorted(set(rng.integers(0, n, size=16).tolist()) | {0, n - 1})
    for i in check_idx:
        if plan.read_int(i, "n") != shift_and_expected_popcount(w.cases[i], pattern):
            raise SystemExit(f"shift_and popcount of sequence {i} differs from the host computation")
    log("shift_and: outputs checked; timing the fused plan")
    ms = time_plans(ctx, [plan], args.steps, args.warmup, api.FUSED)
    log(f"shift_and: fused {ms:.4f} ms/step; timing the per-op executor (eager launches)")
    ops_per_inst = plan.info()["graph_nodes"]
    perop = None
    if not args.no_perop:
        perop = time_plans(ctx, [plan], 2, 1, api.PEROP, launches_per_run=n * ops_per_inst)
        log(f"shift_and: per-op {perop:.4f} ms/step")
    info = plan.info()
    B.free()
    res.free()
    out = {"ms_per_step": ms, "bit_ops": w.bit_ops, "alg_bytes": w.alg_bytes,
           "plan": info, "instances": n, "checked_instances": len(check_idx), "options": opts, "tune": tune}
    if perop is not None:
        out["perop_ms_per_step"] = perop
    return w, out


VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9   # CUs x 64 lanes x 2.4 GHz: 32-bit VALU lane-ops per second


def bench_regex_dag(ctx, args):
    """A generated regex-style DAG (default 200k generator nodes, ~250k DSL
    ops) over the 8 basis streams of a 2^27-byte Zipf text.  Untuned default
    schedule (its kernels are compiled by 16 worker processes first).  All 64
    outputs of the fused plan are compared word for word with the per-op
    executor's on the same inputs."""
    from bitstreamzoo_amd import dag_gen
    nodes = args.dag_nodes
    nbytes = args.bitlength or (1 << 27)
    log(f"regex_dag: generating {nodes} nodes over a {nbytes}-byte text")
    p = dag_gen.generate(nodes, seed=4)
    case = dag_gen.text_case(nbytes, seed=11)
    prog = api.Program(p.source)
    # 4 warm-up rows: faster than 1 here (fewer chunks start from a wrong
    # carry, so fewer are run again after their look-back)
    opts = json.loads(args.schedule) if args.schedule else {"warmup_rows": 4}
    t0 = time.perf_counter()
    tuning.precompile(p.source, [case], [opts], processes=16, shards=64, log=lambda m: log(f"regex_dag: {m}"))
    compile_s = time.perf_counter() - t0
    res = tuning.Resident(ctx, [case])
    plan = api.Plan(ctx, prog, tuning.shape_case([case]), opts)
    res.bind(plan)
    info = plan.info()
    ks = info["kernels"]
    log(f"regex_dag: {len(ks)} kernels compiled in {compile_s:.1f} s; running the fused plan")
    nw = (case.bitlength + 63) // 64
    plan.run()
    ctx.sync()
    fused = {o: plan.download_output(0, o, nw) for o in p.outputs}
    log("regex_dag: running the per-op executor for the output check")
    plan.set_executor(api.PEROP)
    plan.run()
    ctx.sync()
    bad = [o for o in p.outputs if not np.array_equal(plan.download_output(0, o, nw), fused[o])]
    if bad:
        raise SystemExit(f"regex_dag: {len(bad)} fused outputs differ from the per-op executor's")
    del fused
    ms = time_plans(ctx, [plan], args.steps, args.warmup, api.FUSED, launches_per_run=len(ks))
    mism = [plan.mismatches(k) for k in range(len(ks))]
    log(f"regex_dag: fused {ms:.1f} ms/step; {sum(mism)} chunks ran again after their look-back "
        f"({sum(1 for m in mism if m)} kernels)")
    perop = time_plans(ctx, [plan], 1, 0, api.PEROP, launches_per_run=info["graph_nodes"])
    log(f"regex_dag: per-op {perop:.1f} ms/step")
    plan.close()
    res.free()
    planned = 8 * nw * sum(k["inputs"] + k["outputs"] for k in ks)
    alg = 8 * nw * (8 + len(p.outputs))
    w = workloads.Workload("regex_dag", p.source, [case], bit_ops=p.n_dsl_ops * case.bitlength, alg_bytes=alg)
    return w, {"ms_per_step": ms, "perop_ms_per_step": perop, "bit_ops": w.bit_ops, "alg_bytes": alg,
               "planned_bytes": planned, "lane_ops": 2 * p.n_dsl_ops * nw, "plan": info, "options": opts,
               "compile_s": compile_s, "dsl_ops": p.n_dsl_ops, "outputs_checked": len(p.outputs),
               "rerun_chunks": int(sum(mism)), "chunks_per_kernel": info["n_chunks"]}


def bench_learned(ctx, args):
    """The learned cost model against the measured autotuner on held-out
    generated DAG programs: train the GPU MLP on the committed autotuner
    measurements, then per program compare the model's pick with the full
    autotune over the same candidates (gap, wall-time ratio, prediction
    error, gain over the untuned default schedule)."""
    from bitstreamzoo_amd import costmodel, schedsearch
    X, ms = schedsearch.load_training_data()
    model = costmodel.CostModel(ctx, seed=0)
    loss = model.fit(X, ms)
    log(f"learned: {model.members} networks trained on {len(ms)} autotuner measurements in "
        f"{model.train_seconds:.2f} s (loss {loss:.4f})")
    cands = schedsearch.sample_space(64, 12345)
    held = schedsearch.random_specs(args.learned_programs, 7, nodes=(1000, 2000), log2_bytes=(21, 22))
    rows = schedsearch.evaluate(ctx, model, held, cands, processes=min(16, os.cpu_count() or 1), log=log)
    return {"train_rows": int(len(ms)), "train_s": round(model.train_seconds, 3), "candidates": len(cands),
            "programs": [f"{r.nodes} nodes x {r.n_bytes} B" for r in rows],
            "gap_pct": [round(100 * r.gap, 2) for r in rows],
            "time_ratio_pct": [round(100 * r.model_s / r.tune_s, 2) for r in rows],
            "mape_pct": [round(100 * r.mape, 1) for r in rows],
            "pick_vs_default_speedup": [round(r.default_ms / r.pick_ms, 3) for r in rows]}


def bench_mixed(ctx, args, full: bool = True):
    """A mixed batch of generated regex-style DAG programs, each over its own
    text, all run per step as one captured hipGraph of every plan's launches.
    The programs are spread over `--mixed-streams` streams (one context each,
    balanced by ops x bits), which fork from the capturing stream and join it
    again, so independent programs run concurrently; the same graph with the
    streams chained one after another is timed beside it.
    Each program's schedule is the learned model's top prediction over 64
    sampled candidates (no trial runs); every program's outputs are checked
    against the same program under the default schedule, and the measured
    autotuner re-tunes a few programs over 16 of the candidates as judge.
    Sizes are reduced from 1k-300k ops over 2^16-2^26 bytes to
    `--mixed-nodes` and `--mixed-log2-bytes`, so that compiling 256 programs
    stays within the benchmark's minutes.  `full` False (the default
    benchmark run) checks and times the default schedule on a random 5 % of
    the programs only, so it compiles about half as much."""
    from bitstreamzoo_amd import costmodel, schedsearch
    n = args.mixed_programs
    lo, hi = (int(x) for x in args.mixed_nodes.split(","))
    blo, bhi = (float(x) for x in args.mixed_log2_bytes.split(","))
    specs = schedsearch.random_specs(n, 55, nodes=(lo, hi), log2_bytes=(blo, bhi))
    built = [sp.build() for sp in specs]
    X, ms_train = schedsearch.load_training_data()
    model = costmodel.CostModel(ctx, seed=0)
    model.fit(X, ms_train)
    cands = schedsearch.sample_space(64, 12345)
    t0 = time.perf_counter()
    picks = []
    for p, case in built:
        F = costmodel.features_batch(api.Program(p.source), tuning.shape_case([case]), cands,
                                     max(1, (case.bitlength + 63) // 64), 1)
        picks.append(cands[int(np.argmin(model.predict_ms(F)))])
    pick_s = time.perf_counter() - t0
    rng = np.random.Generator(np.random.PCG64(5))
    checked = set(range(n)) if full else set(rng.choice(n, size=max(1, n // 20), replace=False).tolist())
    t1 = time.perf_counter()
    tuning.precompile_many([(p.source, [case], o) for (p, case), o in zip(built, picks)] +
                           [(built[i][0].source, [built[i][1]], {}) for i in sorted(checked)], processes=16,
                           log=lambda m: log(f"mixed: {m}"))
    compile_s = time.perf_counter() - t1
    log(f"mixed: {n} programs, picks in {pick_s:.1f} s, compiled in {compile_s:.1f} s")
    residents, plans, defaults = [], [], []
    bit_ops = alg = launches = refused = 0
    n_streams = max(1, args.mixed_streams)
    ctxs = [ctx] + [api.Context(0) for _ in range(n_streams - 1)]
    bins = tuning.balance([p.n_dsl_ops * case.bitlength for p, case in built], n_streams)
    try:
        for i, ((p, case), o) in enumerate(zip(built, picks)):
            c = ctxs[bins[i]]
            r = tuning.Resident(c, [case])
            residents.append(r)
            prog = api.Program(p.source)
            try:
                pl = api.Plan(c, prog, tuning.shape_case([case]), o)
            except api.BszError as e:
                if not tuning.refused(e):
                    raise
                # the model picked a schedule whose kernels would overflow
                refused += 1
                picks[i] = {}
                pl = api.Plan(c, prog, tuning.shape_case([case]), {})
            r.bind(pl)
            plans.append(pl)
            if i in checked:
                d = api.Plan(c, prog, tuning.shape_case([case]), {})
                r.bind(d)
                defaults.append((i, d))
            bit_ops += p.n_dsl_ops * case.bitlength
            alg += 8 * ((case.bitlength + 63) // 64) * (8 + len(p.outputs))
            launches += len(pl.info()["kernels"])
        for i, d in defaults:
            plans[i].run()
            d.run()
        for c in ctxs:
            c.sync()
        for i, d in defaults:
            p, case = built[i]
            nw = (case.bitlength + 63) // 64
            for o in p.outputs:
                if not np.array_equal(plans[i].download_output(0, o, nw), d.download_output(0, o, nw)):
                    raise SystemExit(f"mixed: output {o} of program {i} differs between two schedules")
        def groups(items):
            return [(ctxs[k], [pl for i, pl in items if bins[i] == k]) for k in range(n_streams)]
        ms = tuning.time_runs_streams(ctx, groups(list(enumerate(plans))), args.steps, args.warmup)
        serial_ms = tuning.time_runs_streams(ctx, groups(list(enumerate(plans))), args.steps, args.warmup,
                                             serial=True)
        default_ms = (tuning.time_runs_streams(ctx, groups(defaults), args.steps, args.warmup) if full else None)
    finally:
        for pl in plans + [d for _, d in defaults]:
            pl.close()
        for r in residents:
            r.free()
        for c in ctxs[1:]:
            c.close()
    log(f"mixed: batch {ms:.3f} ms/step with the model's schedules on {n_streams} streams "
        f"({serial_ms:.3f} ms with the streams one after another; {launches} launches per step); "
        f"default schedules {default_ms} ms; {len(defaults)} programs checked against their default schedule")
    judged = []
    sub = [cands[i] for i in sorted(rng.choice(len(cands), size=16, replace=False).tolist())]
    for i in sorted(rng.choice(n, size=min(n, args.mixed_judged), replace=False).tolist()):
        p, case = built[i]
        trial = [picks[i]] + [c for c in sub if c != picks[i]]
        m = schedsearch.measure(ctx, p.source, case, trial, processes=16)
        if picks[i] not in m.options:
            continue
        pick_ms = float(m.ms[m.options.index(picks[i])])
        judged.append({"program": i, "nodes": specs[i].nodes, "bytes": specs[i].n_bytes,
                       "pick_ms": round(pick_ms, 4), "autotuned_best_ms": round(float(min(m.ms)), 4),
                       "gap_pct": round(100 * (pick_ms / float(min(m.ms)) - 1), 2)})
        log(f"mixed: judged program {i}: pick {pick_ms:.3f} ms, autotuned best {min(m.ms):.3f} ms")
    return {"programs": n, "nodes": [lo, hi], "log2_bytes": [blo, bhi], "ms_per_step": ms, "refused_picks": refused,
            "streams": n_streams, "streams_in_series_ms_per_step": serial_ms,
            "checked_against_default": len(defaults),
            "default_ms_per_step": default_ms, "launches_per_step": launches, "bit_ops": bit_ops,
            "alg_bytes": alg, "pick_s": round(pick_s, 2), "compile_s": round(compile_s, 1), "judged": judged}


def copy_bandwidth(ctx) -> float:
    """Achieved GB/s of a plain device copy (read + write) of 2 GiB."""
    bl = 1 << 34  # 2 GiB per buffer
    src = ctx.alloc(bl)
    dst = ctx.alloc(bl)
    g = ctx.capture(lambda: api.lib().bsz_op_copy(ctx.h, dst.dev, src.dev, bl))
    g.launch(2)
    ctx.sync()
    ctx.timer_start()
    g.launch(10)
    ms = ctx.timer_stop() / 10
    src.free()
    dst.free()
    return 2 * (bl // 8) / (ms * 1e-3) / 1e9


CPU_CALIBRATION = os.path.join(ROOT, "profiles", "cpu_port_vs_reference.md")


def cpu_host():
    """(CPU model, physical cores of the first socket) of this host."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    if phys == "0" and core is not None:
                        cores.add(core)
                    phys = core = None
        if phys == "0" and core is not None:
            cores.add(core)
    except OSError:
        pass
    return model, len(cores) or (os.cpu_count() or 1)


def calibration_ratio(key: str):
    """oracle / reference time ratio measured by oracle/calibrate_cpu.py."""
    try:
        with open(CPU_CALIBRATION) as f:
            for line in f:
                if line.startswith(f"ratio_{key}:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(w):
    """The oracle's AVX-512 + OpenMP CPU executor on the same adder32 inputs
    as the GPU run (full size), in the bsim-faithful mode (ripple `+`, serial
    `<<`, like `bsim --backend simd_omp --add ripple`) and the best-CPU mode
    (parallel `+` and `<<`).  Threads: the physical cores of one socket,
    capped at this job's 16-CPU share."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from cpu_models import Oracle
    model, socket_cores = cpu_host()
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    threads = max(1, min(socket_cores, share, os.cpu_count() or 1))
    Oracle.lib().bszo_set_threads(threads)
    threads = Oracle.lib().bszo_get_threads()
    case = w.cases[0]
    out = {}
    for mode, variant in (("faithful", Oracle.FAITHFUL), ("best", Oracle.BEST)):
        t0 = time.perf_counter()
        o = Oracle(w.source).load(case)
        o.run(variant)
        out[mode] = (o.exec_ms(), time.perf_counter() - t0)
        del o
    ms_f, wall_f = out["faithful"]
    ms_b, _ = out["best"]
    return {"value": round(w.bit_ops / (ms_f * 1e-3) / 1e9, 1), "unit": "Gbit-ops/s", "cores": threads,
            "kind": "port",
            "sample": f"the whole adder32 workload ({case.bitlength} bits per stream, the GPU run's inputs), "
                      f"bsim-faithful mode: one AVX-512+OpenMP pass per DSL op, {ms_f:.0f} ms exec "
                      f"({wall_f:.1f} s wall)",
            "best_cpu": {"value": round(w.bit_ops / (ms_b * 1e-3) / 1e9, 1), "exec_ms": round(ms_b, 1),
                         "mode": "parallel + and << (oracle BEST)"},
            "cpu_model": model, "socket_cores": socket_cores,
            "port_over_reference_time": {"faithful": calibration_ratio("adder32_faithful"),
                                         "best": calibration_ratio("adder32_best"),
                                         "source": "profiles/cpu_port_vs_reference.md"},
            "avx512": bool(Oracle.lib().bszo_has_avx512())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="adder32", choices=["adder32", "matchstar", "shift_and", "regex_dag", "mixed"])
    ap.add_argument("--no-dag", action="store_true", help="leave the regex DAG and the mixed batch out of the default run")
    ap.add_argument("--dag-nodes", type=int, default=200_000, help="generator nodes of the regex_dag workload")
    ap.add_argument("--bitlength", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-perop", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner; use fixed default schedules")
    ap.add_argument("--schedule", default="", help="JSON options for the measured workload's plan (skips tuning)")
    ap.add_argument("--mixed-programs", type=int, default=256, help="programs of the mixed batch (0 skips it)")
    ap.add_argument("--mixed-streams", type=int, default=8, help="streams the mixed batch's programs run on")
    ap.add_argument("--mixed-nodes", default="200,2000", help="generator nodes of the mixed batch, low,high")
    ap.add_argument("--mixed-log2-bytes", default="16,20", help="text bytes (log2) of the mixed batch, low,high")
    ap.add_argument("--mixed-judged", type=int, default=2, help="mixed-batch programs re-tuned by the autotuner")
    ap.add_argument("--learned-programs", type=int, default=1,
                    help="held-out programs for the learned-vs-autotuned comparison (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the executor runs on a single MI355X and has no multi-GPU "
              "path. Run with --gpus 1.", file=sys.stderr)
        return 2

    ctx = api.Context(0)
    results = {}
    extra = {}
    if args.workload == "adder32":
        w, r = bench_adder(ctx, args, perop=not args.no_perop)
    elif args.workload == "matchstar":
        w, r = bench_matchstar(ctx, args)
    elif args.workload == "regex_dag":
        w, r = bench_regex_dag(ctx, args)
    elif args.workload == "mixed":
        r = bench_mixed(ctx, args)
        m = r["ms_per_step"]
        print(json.dumps({"metric": METRIC, "workload": "mixed", "ms_per_step": round(m, 5),
                          "value": round(r["bit_ops"] / (m * 1e-3) / 1e9, 1), "unit": "Gbit-ops/s",
                          "hbm_frac": round(r["alg_bytes"] / (m * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                          **{k: v for k, v in r.items() if k != "ms_per_step"}}), flush=True)
        ctx.close()
        return 0
    else:
        w, r = bench_shift_and(ctx, args)
    results[args.workload] = r
    if not args.no_extra and args.workload == "adder32":
        for name, fn in (("matchstar", bench_matchstar), ("shift_and", bench_shift_and)):
            try:
                _, rr = fn(ctx, args)
                extra[name] = rr
            except SystemExit:
                raise
            except Exception as e:  # report, do not hide
                extra[name] = {"error": str(e)}
        if not args.no_dag:
            # the 200k-node DAG at its default schedule, 3 timed steps
            dag_args = argparse.Namespace(**dict(vars(args), steps=3, warmup=1, schedule=""))
            for name, fn in (("regex_dag", lambda c, a: bench_regex_dag(c, a)[1]),
                             ("mixed", lambda c, a: bench_mixed(c, a, full=False))):
                if name == "mixed" and args.mixed_programs <= 0:
                    continue
                try:
                    extra[name] = fn(ctx, dag_args)
                except SystemExit:
                    raise
                except Exception as e:  # report, do not hide
                    extra[name] = {"error": str(e)}
    learned = None
    if not args.no_extra and args.workload == "adder32" and args.learned_programs > 0:
        try:
            learned = bench_learned(ctx, args)
        except SystemExit:
            raise
        except Exception as e:  # report, do not hide
            learned = {"error": str(e)}
    copy_gbps = copy_bandwidth(ctx)

    ms = r["ms_per_step"]
    gbitops = r["bit_ops"] / (ms * 1e-3) / 1e9
    achieved = r["alg_bytes"] / (ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(gbitops, 1),
        "unit": "Gbit-ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded PCG64), inputs resident in HBM",
        "config": {"workload": {
            "adder32": "bitsliced 32-bit ripple-carry adder DAG, 65 inputs x 2^28 bits, 160 ops + 33 popcounts",
            "matchstar": "MatchStar on 2 x 2^26-bit streams + popcount, 12 rotating input sets",
            "shift_and": "batched Shift-And, 2048 sequences, 32-mer, one launch",
            "regex_dag": f"random regex-style DAG, {args.dag_nodes} generator nodes "
                         f"({r.get('dsl_ops')} DSL ops), 8 basis streams of a "
                         f"{w.cases[0].bitlength}-byte Zipf(1.1) text, 64 outputs"}[args.workload],
            "executor": "fused", "bitlength": w.cases[0].bitlength if args.workload != "shift_and" else None,
            "kernels": len(r["plan"]["kernels"]), "schedule": r["options"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": profiled_traffic() if args.workload == "adder32" else None,
                     "measured_copy_GBps": round(copy_gbps, 1),
                     "kernel": r["plan"]["kernels"][0].get("entry") if len(r["plan"]["kernels"]) == 1 else None},
    }
    if args.workload == "regex_dag":
        # HBM bytes the plan moves (intermediates between kernels included)
        # and the VALU bound the workload is quoted against
        line["roofline"]["planned_bytes"] = r["planned_bytes"]
        line["roofline"]["planned_frac"] = round(r["planned_bytes"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
        line["compile_s"] = round(r["compile_s"], 1)
        line["rerun_chunks"] = r["rerun_chunks"]
        line["outputs_checked_against_perop"] = r["outputs_checked"]
    # VALU side of the roofline: 2 32-bit lane-ops per 64-bit word per DSL op
    lane_ops = r.get("lane_ops", 2 * r["bit_ops"] // 64)
    line["valu"] = {"lane_ops": lane_ops, "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
                    "bound_ms": round(lane_ops / VALU_PEAK_LANE_OPS * 1e3, 4),
                    "frac": round(lane_ops / VALU_PEAK_LANE_OPS / (ms * 1e-3), 4)}
    if r.get("tune"):
        line["autotune"] = r["tune"]
    if "perop_ms_per_step" in r:
        line["perop"] = {"ms_per_step": round(r["perop_ms_per_step"], 5),
                         "value": round(r["bit_ops"] / (r["perop_ms_per_step"] * 1e-3) / 1e9, 1),
                         "speedup_fused_over_perop": round(r["perop_ms_per_step"] / ms, 2)}
    ex = {}
    for name, rr in extra.items():
        if "error" in rr:
            ex[name] = rr
            continue
        m = rr["ms_per_step"]
        ex[name] = {"ms_per_step": round(m, 5), "value": round(rr["bit_ops"] / (m * 1e-3) / 1e9, 1),
                    "GBps": round(rr["alg_bytes"] / (m * 1e-3) / 1e9, 1),
                    "hbm_frac": round(rr["alg_bytes"] / (m * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                    "schedule": rr.get("options"), "tuning_s": rr["tune"]["tuning_s"] if rr.get("tune") else None}
        if "perop_ms_per_step" in rr:
            ex[name]["perop_ms_per_step"] = round(rr["perop_ms_per_step"], 5)
            ex[name]["speedup_fused_over_perop"] = round(rr["perop_ms_per_step"] / m, 2)
        if "unproven_chunks" in rr:
            ex[name]["unproven_chunks"] = rr["unproven_chunks"]
        if name == "regex_dag":
            ex[name]["planned_bytes"] = rr["planned_bytes"]
            ex[name]["planned_hbm_frac"] = round(rr["planned_bytes"] / (m * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
            ex[name]["valu_frac"] = round(rr["lane_ops"] / VALU_PEAK_LANE_OPS / (m * 1e-3), 4)
            ex[name]["kernels"] = len(rr["plan"]["kernels"])
            ex[name]["rerun_chunks"] = rr["rerun_chunks"]
            ex[name]["compile_s"] = round(rr["compile_s"], 1)
            ex[name]["outputs_checked_against_perop"] = rr["outputs_checked"]
        if name == "mixed":
            for k in ("programs", "nodes", "log2_bytes", "default_ms_per_step", "launches_per_step", "pick_s",
                      "compile_s", "refused_picks", "checked_against_default", "judged", "streams",
                      "streams_in_series_ms_per_step"):
                ex[name][k] = rr[k]
        if "floor_ms" in rr:
            # the same bytes in the cheapest single launch (r = M & C)
            ex[name]["one_launch_floor_ms"] = round(rr["floor_ms"], 5)
            ex[name]["one_launch_floor_hbm_frac"] = round(rr["alg_bytes"] / (rr["floor_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
            ex[name]["frac_of_floor"] = round(rr["floor_ms"] / m, 4)
            ex[name]["one_launch_floor_schedule"] = rr["floor_options"]
    if ex:
        line["other_configs"] = ex
    if learned is not None:
        line["learned_vs_autotuned"] = learned
    if not args.no_cpu and args.workload == "adder32":
        log("cpu baseline: oracle AVX-512+OpenMP on the same adder32 inputs")
        line["cpu_baseline"] = cpu_baseline(w)
    ctx.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
