"""Measured autotuner for fused plans.

A schedule is a set of FusedOptions keys (chunk_rows, rows_per_step,
waves_per_block, occupancy, words_per_lane, interleave, warmup_rows, max_ops,
max_live, nontemporal): the rows a wave walks per chunk, the rows whose loads
are issued together (words per lane in flight), the workgroup width, the
register budget, 8- or 16-byte lanes, contiguous or interleaved chunks, the
fusion depth and whether stream loads and stores bypass cache retention.  The tuner builds one plan per candidate
over the same resident inputs, skips any whose compiled kernels use AGPRs or
scratch memory (read from the code objects' metadata), checks that every
other candidate produces the same popcounts and the same output words as the
first one, and times each with hipEvents over replays of a captured hipGraph.
Schedules never change output bits; the check guards that.

The reference has no schedule search (one kernel per op, fixed 256-thread
blocks: kernels.cuh:7-36, cuda_backend.cuh:53-84).
"""

from __future__ import annotations

import itertools
import json
import time
from dataclasses import dataclass, field

import numpy as np

from . import api
from .bsdata import BsCase


def _shape_desc(cases: list[BsCase]) -> tuple:
    c0 = cases[0]
    return (dict(c0.params), sorted(c0.inputs), {k: sorted(a) for k, a in c0.input_arrays.items()})


def _shape_from_desc(desc) -> api.Case:
    params, inputs, arrays = desc
    small = BsCase(bitlength=64, params=dict(params), inputs={k: np.zeros(1, np.uint64) for k in inputs},
                   input_arrays={k: {i: np.zeros(1, np.uint64) for i in idx} for k, idx in arrays.items()})
    return api.Case.from_bscase(small)


def _compile_worker(job):
    # each worker process compiles one unit at a time: the processes are the
    # parallelism (many processes times many hipRTC threads exhaust memory)
    import os
    os.environ["BSZ_COMPILE_THREADS"] = "1"
    source, desc, options = job
    try:
        api.compile_plan_only(api.Program(source), _shape_from_desc(desc), options)
    except api.BszError as e:
        # a schedule the library refuses is an answer, not a failure of the
        # pool: the plan built from it later reports the same refusal
        if "schedule refused" in str(e):
            return str(e)
        raise RuntimeError(str(e)) from None
    return True


def _drain(futs, log, what: str, t0: float) -> None:
    """Waits for every job; on the first failure the pool is shut down and
    dropped (a broken pool left behind would hold the process at exit)."""
    import concurrent.futures as cf
    global _POOL
    every = max(1, len(futs) // 10)
    try:
        for i, f in enumerate(cf.as_completed(futs)):
            f.result()
            if log and ((i + 1) % every == 0 or i + 1 == len(futs)):
                log(f"compiled {i + 1}/{len(futs)} {what} ({time.perf_counter() - t0:.0f} s)")
    except BaseException:
        if _POOL is not None:
            _POOL[1].shutdown(wait=False, cancel_futures=True)
            _POOL = None
        raise


def precompile(source: str, cases: list[BsCase], options_list: list[dict], processes: int = 16,
               shards: int = 0, log=None) -> float:
    """Fill the code-object cache for every schedule of `options_list` with
    `processes` worker processes (hipRTC compiles one unit at a time per
    process); each plan's kernels are split into `shards` (default
    `processes`) jobs over the workers, and `log` hears of every finished job.
    Returns the wall seconds spent.  Needs no GPU."""
    import concurrent.futures as cf
    t0 = time.perf_counter()
    desc = _shape_desc(cases)
    shards = shards or processes
    jobs = []
    for opts in options_list:
        for r in range(shards):
            o = dict(opts)
            o.update(compile_shard=r, compile_shards=shards)
            jobs.append((source, desc, o))
    futs = [_pool(processes).submit(_compile_worker, j) for j in jobs]
    _drain(futs, log, "shards", t0)
    return time.perf_counter() - t0


def precompile_many(jobs: list, processes: int = 16, log=None) -> float:
    """Fill the code-object cache for many programs at once: `jobs` is a list
    of (source, cases, options); each is one compile job (all of one plan's
    kernels) on the shared worker pool.  Returns the wall seconds spent."""
    import concurrent.futures as cf
    t0 = time.perf_counter()
    futs = [_pool(processes).submit(_compile_worker, (src, _shape_desc(cases), dict(opts)))
            for src, cases, opts in jobs]
    _drain(futs, log, "programs", t0)
    return time.perf_counter() - t0


_POOL = None


def _pool(processes: int):
    """Worker processes for compilation, started once (spawned: they never
    touch the GPU) and reused by every precompile() of this process."""
    global _POOL
    import concurrent.futures as cf
    import multiprocessing as mp
    if _POOL is None or _POOL[0] != processes:
        if _POOL is not None:
            _POOL[1].shutdown()
        _POOL = (processes, cf.ProcessPoolExecutor(max_workers=processes, mp_context=mp.get_context("spawn")))
    return _POOL[1]


def shape_case(cases: list[BsCase]) -> api.Case:
    """A 64-bit case with the params and input names of `cases[0]`: enough to
    record a straight-line program."""
    c0 = cases[0]
    small = BsCase(bitlength=64, params=dict(c0.params),
                   inputs={k: np.zeros(1, np.uint64) for k in c0.inputs},
                   input_arrays={k: {i: np.zeros(1, np.uint64) for i in a} for k, a in c0.input_arrays.items()})
    return api.Case.from_bscase(small)


def time_runs_streams(ctx: api.Context, groups: list, steps: int, warmup: int, serial: bool = False) -> float:
    """ms per batch, a batch being one run of every plan of `groups`, a list
    of (context, plans) with one context per stream (`ctx` may hold a group
    of its own).  The batch is captured on `ctx`'s stream: the other streams
    fork from it at the start and join it at the end, so the groups run
    concurrently inside one hipGraph; with `serial` each group starts after
    the previous one ends (the same graph shape, one group at a time).
    Replayed and timed with hipEvents on `ctx`'s stream."""
    for c, plans in groups:
        for p in plans:
            p.set_executor(api.FUSED)
            p.prepare()
            p.run()
    for c, _ in groups:
        c.sync()
    others = [c for c, _ in groups if c is not ctx]

    def batch():
        if serial:
            prev = ctx
            for c, plans in groups:
                if c is not prev:
                    c.order_after(prev)
                for p in plans:
                    p.run()
                prev = c
        else:
            for c in others:
                c.order_after(ctx)
            for c, plans in groups:
                for p in plans:
                    p.run()
        for c in others:
            ctx.order_after(c)

    g = ctx.capture(batch)
    g.launch(warmup)
    ctx.sync()
    ctx.timer_start()
    g.launch(steps)
    ms = ctx.timer_stop() / steps
    del g
    return ms


def balance(costs: list[float], n_bins: int) -> list[int]:
    """Bin of each item: longest-processing-time first, each to the bin with
    the least cost so far."""
    load = [0.0] * n_bins
    out = [0] * len(costs)
    for i in sorted(range(len(costs)), key=lambda i: -costs[i]):
        b = min(range(n_bins), key=lambda k: load[k])
        out[i] = b
        load[b] += costs[i]
    return out


class Resident:
    """Every case's inputs uploaded once to device streams."""

    def __init__(self, ctx: api.Context, cases: list[BsCase]):
        self.ctx, self.cases = ctx, cases
        self.bufs: dict[tuple[int, str, int], api.DevStream] = {}
        for i, c in enumerate(cases):
            for n, w in c.inputs.items():
                self.bufs[(i, n, -1)] = ctx.upload(w, c.bitlength)
            for n, arr in c.input_arrays.items():
                for j, w in arr.items():
                    self.bufs[(i, n, int(j))] = ctx.upload(w, c.bitlength)

    def bind(self, plan: api.Plan):
        plan.set_instances([c.bitlength for c in self.cases])
        for (i, n, j), d in self.bufs.items():
            plan.bind(i, n, d, j)
        plan.prepare()

    def free(self):
        for d in self.bufs.values():
            d.free()
        self.bufs.clear()


def time_runs(ctx: api.Context, plans: list[api.Plan], steps: int, warmup: int,
              executor: int = api.FUSED, max_graph_launches: int = 20000,
              launches_per_run: int = 1) -> float:
    """ms per plan run: `steps` runs of every plan, captured into one hipGraph
    and replayed (or enqueued eagerly when a run holds more than
    `max_graph_launches` kernels), timed with hipEvents on the context stream."""
    for p in plans:
        p.set_executor(executor)
        p.prepare()
        p.run()
    ctx.sync()
    if launches_per_run * len(plans) > max_graph_launches:
        for _ in range(warmup):
            for p in plans:
                p.run()
        ctx.sync()
        ctx.timer_start()
        for _ in range(steps):
            for p in plans:
                p.run()
        return ctx.timer_stop() / (steps * len(plans))
    g = ctx.capture(lambda: [p.run() for p in plans])
    g.launch(warmup)
    ctx.sync()
    ctx.timer_start()
    g.launch(steps)
    ms = ctx.timer_stop() / (steps * len(plans))
    del g
    return ms


def plan_signature(plan: api.Plan, resident: Resident, int_names: list[str], out_names: list[str],
                   sample_words: int = 4096) -> tuple:
    """Popcounts of every instance plus the first and last `sample_words` words
    of each output stream of instances 0 and n-1."""
    sig = []
    n = len(resident.cases)
    for i in range(n):
        for name in int_names:
            sig.append(plan.read_int(i, name))
    for i in sorted({0, n - 1}):
        nw = max(1, (resident.cases[i].bitlength + 63) // 64)
        k = min(nw, sample_words)
        for name in out_names:
            sig.append(plan.download_output(i, name, k).tobytes())
            sig.append(plan.download_output(i, name, k, offset=nw - k).tobytes())
    return tuple(sig)


def default_space(kind: str = "auto") -> list[dict]:
    """Candidate schedules.  `kind` narrows the grid: "wide" for kernels with
    many inputs (register-bound), "stateful" for add/advance chains, "small"
    for batches of short streams."""
    grid = {
        "chunk_rows": [8, 16, 32],
        "rows_per_step": [1, 2, 4, 8],
        "occupancy": [0, 2, 4],
        "waves_per_block": [4],
    }
    if kind == "wide":
        # one kernel per program where it fits (every split into smaller
        # kernels, with or without 16-byte lanes, measured slower or equal
        # on the 65-input adder while moving 2 % more bytes); nontemporal
        # loads and stores (2) measured 5 % faster on that adder.  64-row
        # chunks leave one wave per SIMD on that adder, and their time moved
        # by 8 % between processes of the same run (0.599 against 0.646 ms),
        # so they are not candidates.
        grid = {"chunk_rows": [16, 32], "rows_per_step": [1, 2], "occupancy": [1, 2],
                "waves_per_block": [4], "interleave": [0, 1], "nontemporal": [0, 2]}
    elif kind == "stateful":
        # nontemporal loads and stores (2) measured 15 % faster on MatchStar
        # over 2^26-bit streams
        grid = {"chunk_rows": [4, 8, 16, 32], "rows_per_step": [2, 4, 8, 16], "occupancy": [0],
                "waves_per_block": [1, 4], "nontemporal": [0, 2]}
    elif kind == "small":
        grid = {"chunk_rows": [2, 4, 8], "rows_per_step": [2, 4, 8], "occupancy": [0, 4],
                "waves_per_block": [2, 4], "interleave": [0, 1]}
    keys = sorted(grid)
    return [dict(zip(keys, v)) for v in itertools.product(*(grid[k] for k in keys))]


@dataclass
class Trial:
    options: dict
    ms: float
    info: dict
    compile_s: float


def register_clean(info: dict) -> bool:
    """Every compiled kernel of the plan stays in its architectural VGPRs
    (no AGPR use, no scratch); kernels that overflow spill values on every
    row, so the tuner does not use such schedules.  A kernel whose
    counts are missing or unknown (-1) counts as unclean.  (The library
    itself refuses such plans; this is the tuner's own check of the info.)
    A plan built with {"occupancy": 1, "agprs": 1} may keep values in AGPRs,
    never in scratch."""
    allow_agprs = info.get("agprs") == 1
    for k in info.get("kernels", []):
        vg, ag, sc = k.get("vgprs", -1), k.get("agprs", -1), k.get("scratch_bytes", -1)
        if vg is None or ag is None or sc is None or vg <= 0 or ag < 0 or sc != 0:
            return False
        if ag != 0 and not allow_agprs:
            return False
    return True


def refused(e: Exception) -> bool:
    """The library refused a schedule because its compiled kernels overflow
    the architectural VGPRs (a RUNTIME error naming the refusal)."""
    return isinstance(e, api.BszError) and e.code == -3 and "schedule refused" in str(e)


@dataclass
class TuneResult:
    trials: list[Trial] = field(default_factory=list)
    skipped: list = field(default_factory=list)
    same_as: list = field(default_factory=list)   # (candidate, earlier candidate with the same kernels)
    seconds: float = 0.0

    # Schedules within this fraction of the fastest count as tied: timing
    # noise between runs is about that size.
    TIE = 0.01

    @property
    def best(self) -> Trial:
        """The fastest schedule; among schedules tied with it, the first in
        candidate order, so repeated tunings of one workload agree on the
        pick whenever the measurements only differ by noise."""
        fastest = min(t.ms for t in self.trials)
        return next(t for t in self.trials if t.ms <= fastest * (1.0 + self.TIE))

    def as_json(self) -> dict:
        b = self.best
        return {"candidates": len(self.trials), "skipped": len(self.skipped), "duplicates": len(self.same_as),
                "tuning_s": round(self.seconds, 2),
                "best": {"options": b.options, "ms": round(b.ms, 5)},
                "worst_ms": round(max(t.ms for t in self.trials), 5)}


_LAUNCH_KEYS = ("waves_per_block", "warmup_rows", "store_order", "nontemporal", "interleave")


def plan_keys(program: api.Program, shape: api.Case, options_list: list[dict]) -> list[str]:
    """A key per schedule that is equal exactly when two schedules partition
    the program into the same kernels with the same shapes and launch them
    the same way (so they compile to the same code)."""
    descs = api.describe_plans(program, shape, [dict(o, detail=1) for o in options_list])
    out = []
    for o, d in zip(options_list, descs):
        launch = {k: o.get(k) for k in _LAUNCH_KEYS}
        out.append(json.dumps([d["kernels"], d.get("kernel_detail"), d["chunk_rows"], launch], sort_keys=True))
    return out


def autotune(ctx: api.Context, program: api.Program, residents: list[Resident], space: list[dict],
             int_names: list[str], out_names: list[str], steps: int = 10, warmup: int = 2,
             base_options: dict | None = None, log=None, allow_empty: bool = False) -> TuneResult:
    """Measure every candidate of `space`: one plan per resident input set
    (rotating through several sets keeps them out of the last-level cache),
    all runs of a step captured together.  Raises if a candidate's outputs
    differ from the first candidate's."""
    t0 = time.perf_counter()
    res = TuneResult()
    shape = shape_case(residents[0].cases)
    ref_sig = None
    keys = plan_keys(program, shape, [dict(base_options or {}, **c) for c in space])
    first_of = {}
    for cand, key in zip(space, keys):
        if key in first_of:
            res.same_as.append((cand, first_of[key]))
            continue
        first_of[key] = cand
        opts = dict(base_options or {})
        opts.update(cand)
        tc = time.perf_counter()
        try:
            plans = [api.Plan(ctx, program, shape, opts) for _ in residents]
        except api.BszError as e:
            if not refused(e):
                raise
            res.skipped.append((cand, "kernels use AGPRs or scratch"))
            if log:
                log(f"tune {cand}: skipped (the library refused it: compiled kernels use AGPRs or scratch)")
            continue
        compile_s = time.perf_counter() - tc
        try:
            for p, r in zip(plans, residents):
                r.bind(p)
            info = plans[0].info()
            if not register_clean(info):
                res.skipped.append((cand, "kernels use AGPRs or scratch"))
                if log:
                    log(f"tune {cand}: skipped (compiled kernels use AGPRs or scratch)")
                continue
            for p in plans:
                p.run()
            ctx.sync()
            sig = tuple(plan_signature(p, r, int_names, out_names) for p, r in zip(plans, residents))
            if ref_sig is None:
                ref_sig = sig
            elif sig != ref_sig:
                raise RuntimeError(f"schedule {cand} changed the outputs")
            ms = time_runs(ctx, plans, steps, warmup)
            res.trials.append(Trial(cand, ms, info, compile_s))
        finally:
            for p in plans:
                p.close()
        if log:
            log(f"tune {cand}: {ms:.4f} ms")
    res.seconds = time.perf_counter() - t0
    if not res.trials and not allow_empty:
        raise RuntimeError(f"all {len(space)} candidate schedules compile to kernels that use AGPRs or scratch")
    return res
