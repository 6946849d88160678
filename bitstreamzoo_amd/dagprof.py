"""Run the fused plan of a generated regex-style DAG a few times, for a
profiler to watch, and write the plan's per-kernel shapes next to it.

    python -m bitstreamzoo_amd.dagprof --nodes 20000 --log2-bytes 27 \\
        --schedule '{"warmup_rows": 4}' --out plan.json

The JSON holds every kernel's declared input and output stream counts (the
bytes a kernel must move: 8 * words per stream), the chunk count and the
chunks each kernel ran again after its look-back, so per-dispatch counters (FETCH_SIZE,
WRITE_SIZE) can be set against the planned traffic.  Kernels are compiled by
worker processes first; the runs themselves are the only GPU work.

    python -m bitstreamzoo_amd.dagprof --analyze TRACE.csv --plan plan.json

joins a rocprofv3 kernel trace (`--kernel-trace --output-format csv`) of
such a run with the plan JSON and prints one JSON line: the kernels' time
per run, the planned bytes per kernel against its average duration
(quantiles of the rate), and the time of the kernels whose chunks ran again
after their look-back.  Needs no GPU.
"""

from __future__ import annotations

import argparse
import json
import sys

from . import api, dag_gen, tuning


def analyze(trace_csv: str, plan_json: str) -> dict:
    """Per-kernel durations of a kernel trace against the plan's planned bytes."""
    import csv
    from collections import defaultdict
    with open(plan_json) as f:
        plan = json.load(f)
    ks, rerun = plan["kernels"], plan["rerun_chunks"]
    stream_bytes = 8 * ((plan["bitlength"] + 63) // 64)
    durs = defaultdict(list)
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            if r["Kernel_Name"].startswith("bsz_fk_"):
                durs[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    # kernels with the same generated code share an entry: their dispatches
    # are averaged together
    avg = {e: sum(d) / len(d) for e, d in durs.items()}
    rows = []
    for k, n_rerun in zip(ks, rerun):
        if k["entry"] not in avg:
            raise ValueError(f"kernel {k['name']} ({k['entry']}) is not in the trace")
        rows.append((avg[k["entry"]], (k["inputs"] + k["outputs"]) * stream_bytes, n_rerun))
    total_ns = sum(t for t, _, _ in rows)
    rates = sorted(b / t for t, b, _ in rows)          # bytes per ns = GB/s

    def q(x):
        return round(rates[min(len(rates) - 1, int(x * (len(rates) - 1)))], 1)

    rr = [(t, b, n) for t, b, n in rows if n > 0]
    plain = [(t, b) for t, b, n in rows if n == 0]
    return {"kernels": len(rows), "kernel_ms_per_run": round(total_ns / 1e6, 1),
            "planned_bytes": int(sum(b for _, b, _ in rows)),
            "rate_GBps_quantiles": {"5": q(0.05), "25": q(0.25), "50": q(0.5), "75": q(0.75), "95": q(0.95)},
            "kernels_without_reruns": {"count": len(plain), "ms": round(sum(t for t, _ in plain) / 1e6, 1),
                                       "GBps": round(sum(b for _, b in plain) / sum(t for t, _ in plain), 1)},
            "kernels_with_reruns": {"count": len(rr), "ms": round(sum(t for t, _, _ in rr) / 1e6, 1),
                                    "chunks_rerun": int(sum(n for _, _, n in rr)),
                                    "largest": [{"ms": round(t / 1e6, 2), "chunks_rerun": n}
                                                for t, _, n in sorted(rr, reverse=True)[:4]]}}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--nodes", type=int, default=20_000)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--log2-bytes", type=int, default=27)
    ap.add_argument("--text-seed", type=int, default=11)
    ap.add_argument("--schedule", default="{}")
    ap.add_argument("--runs", type=int, default=2)
    ap.add_argument("--processes", type=int, default=16)
    ap.add_argument("--out", default="")
    ap.add_argument("--analyze", default="", help="kernel trace CSV to join with --plan (no run)")
    ap.add_argument("--plan", default="", help="plan JSON written by --out")
    a = ap.parse_args(argv)
    if a.analyze:
        print(json.dumps(analyze(a.analyze, a.plan)))
        return 0

    def log(m):
        print(f"[dagprof] {m}", file=sys.stderr, flush=True)

    p = dag_gen.generate(a.nodes, seed=a.seed)
    case = dag_gen.text_case(1 << a.log2_bytes, seed=a.text_seed)
    opts = json.loads(a.schedule)
    tuning.precompile(p.source, [case], [opts], processes=a.processes, shards=2 * a.processes, log=log)
    ctx = api.Context(0)
    res = tuning.Resident(ctx, [case])
    plan = api.Plan(ctx, api.Program(p.source), tuning.shape_case([case]), opts)
    res.bind(plan)
    for i in range(a.runs):
        plan.run()
        ctx.sync()
        log(f"run {i + 1}/{a.runs} done")
    info = plan.info()
    ks = info["kernels"]
    summary = {"nodes": a.nodes, "dsl_ops": p.n_dsl_ops, "bitlength": case.bitlength, "schedule": opts,
               "n_chunks": info["n_chunks"], "kernels": ks,
               "rerun_chunks": [plan.mismatches(k) for k in range(len(ks))]}
    plan.close()
    res.free()
    ctx.close()
    if a.out:
        with open(a.out, "w") as f:
            json.dump(summary, f)
    log(f"{len(ks)} kernels, {sum(summary['rerun_chunks'])} chunks redone")
    return 0


if __name__ == "__main__":
    sys.exit(main())
