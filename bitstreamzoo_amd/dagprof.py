"""Run the fused plan of a generated regex-style DAG a few times, for a
profiler to watch, and write the plan's per-kernel shapes next to it.

    python -m bitstreamzoo_amd.dagprof --nodes 20000 --log2-bytes 27 \\
        --schedule '{"warmup_rows": 4}' --out gpurun_dir/plan.json

The JSON holds every kernel's declared input and output stream counts (the
bytes a kernel must move: 8 * words per stream), the chunk count and the
chunks each kernel ran again after its look-back, so per-dispatch counters (FETCH_SIZE,
WRITE_SIZE) can be set against the planned traffic.  Kernels are compiled by
worker processes first; the runs themselves are the only GPU work.
"""

from __future__ import annotations

import argparse
import json
import sys

from . import api, dag_gen, tuning


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
    a = ap.parse_args(argv)

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
