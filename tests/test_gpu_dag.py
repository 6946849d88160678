"""Fused plans of generated random regex-style DAGs (many kernels, slots,
chained adds, MatchStar / ScanThru, advances) against the oracle, under
several schedules, and as a batch of texts of uneven length."""

import pytest

from bitstreamzoo_amd import api, dag_gen, tuning
from bitstreamzoo_amd.bsdata import words_to_int
from cpu_models import Oracle
from hip_models import ctx, plan_or_refusal
from runner import run_model

pytestmark = pytest.mark.gpu

SCHEDULES = [{}, {"chunk_rows": 2, "rows_per_step": 3, "max_live": 24, "max_ops": 300},
             {"chunk_rows": 8, "rows_per_step": 1, "max_live": 160, "occupancy": 1},
             # input words loaded at their first use, and values kept in AGPRs too
             {"chunk_rows": 4, "lazy_loads": 1},
             {"chunk_rows": 8, "rows_per_step": 1, "max_live": 160, "occupancy": 1, "agprs": 1, "lazy_loads": 1}]


def oracle_outputs(p, case):
    names = Oracle(p.source).output_names()
    o, _ = run_model(Oracle, p.source, case, Oracle.BEST, stream_names=names[0], int_names=names[1])
    return o


@pytest.mark.parametrize("seed,n_bytes", [(0, 5000), (1, 64 * 64 * 9 + 5), (2, 200_003)])
@pytest.mark.parametrize("sched", range(len(SCHEDULES)))
def test_dag_plan_matches_oracle(seed, n_bytes, sched):
    p = dag_gen.generate(1200, seed=seed)
    case = dag_gen.text_case(n_bytes, seed=seed)
    exp = oracle_outputs(p, case)
    c = ctx()
    plan = plan_or_refusal(c, api.Program(p.source), tuning.shape_case([case]), SCHEDULES[sched])
    if plan is None:
        return  # refused: its kernels would overflow the VGPRs (never run)
    res = tuning.Resident(c, [case])
    res.bind(plan)
    plan.run()
    nw = (n_bytes + 63) // 64
    for name in p.outputs:
        got = words_to_int(plan.download_output(0, name, nw))
        assert got == exp[name], f"{name} (kernels={len(plan.info()['kernels'])})"
    plan.close()
    res.free()


def test_dag_batch_of_uneven_texts():
    p = dag_gen.generate(800, seed=7)
    cases = [dag_gen.text_case(n, seed=10 + i) for i, n in enumerate((100, 4096, 64 * 64 * 3 + 1, 70_000))]
    c = ctx()
    res = tuning.Resident(c, cases)
    plan = api.Plan(c, api.Program(p.source), tuning.shape_case(cases), {"chunk_rows": 4})
    res.bind(plan)
    plan.capture()
    plan.replay(3)
    for i, case in enumerate(cases):
        exp = oracle_outputs(p, case)
        nw = (case.bitlength + 63) // 64
        for name in p.outputs[:16]:
            assert words_to_int(plan.download_output(i, name, nw)) == exp[name], (i, name)
    plan.close()
    res.free()


ROWS_SWEEP = [(16, 2), (16, 3), (16, 4), (16, 5), (16, 6), (16, 7), (4, 3), (8, 3), (12, 3)]


def sweep_options(chunk_rows, rows_per_step, occ):
    return {"chunk_rows": chunk_rows, "rows_per_step": rows_per_step, "occupancy": occ, "max_live": 160,
            "max_ops": 4000}


@pytest.fixture(scope="module")
def sweep_program():
    """The wide-partition program of the sweep, its oracle outputs, and every
    schedule of the sweep compiled up front by 16 worker processes (a
    schedule the library refuses is left to the test, which then skips the
    run)."""
    from bitstreamzoo_amd import schedsearch
    spec = schedsearch.random_specs(5, 101, nodes=(2000, 5000), log2_bytes=(23, 25))[3]
    p = dag_gen.generate(spec.nodes, seed=spec.seed)
    case = dag_gen.text_case(64 * 64 * 40 + 123, seed=spec.seed + 7)
    for occ in (1, 0):
        for cr, rps in ROWS_SWEEP:
            try:
                tuning.precompile(p.source, [case], [sweep_options(cr, rps, occ)], processes=16)
            except Exception as e:  # refused schedules are checked by plan_or_refusal below
                if "schedule refused" not in str(e):
                    raise
    return p, case, oracle_outputs(p, case)


@pytest.mark.parametrize("chunk_rows,rows_per_step", ROWS_SWEEP)
@pytest.mark.parametrize("occ", [1, 0])
def test_rows_per_step_sweep_wide_kernel(sweep_program, chunk_rows, rows_per_step, occ):
    """A wide partition (max_live 160: ~100 stateful ops per kernel) with
    every rows-per-step count, including ones that do not divide the chunk."""
    p, case, exp = sweep_program
    c = ctx()
    opts = sweep_options(chunk_rows, rows_per_step, occ)
    plan = plan_or_refusal(c, api.Program(p.source), tuning.shape_case([case]), opts)
    if plan is None:
        return  # refused: its kernels would overflow the VGPRs (never run)
    res = tuning.Resident(c, [case])
    res.bind(plan)
    plan.run()
    nw = (case.bitlength + 63) // 64
    rps = [k["rows_per_step"] for k in plan.info()["kernels"]]
    bad = [n for n in p.outputs if words_to_int(plan.download_output(0, n, nw)) != exp[n]]
    plan.close()
    res.free()
    assert not bad, f"{len(bad)} outputs differ; effective rows per step {rps[:4]}"


WIDE_REGRESSIONS = [
    # (spec list args, index, schedule): wide partitions that once gave wrong
    # outputs (the high halves of shared 64-bit ORs, the miscompile that the
    # generated code guards against) while their kernels also spilled
    (dict(n=32, seed=201, nodes=(500, 5000), log2_bytes=(16, 22)), 25,
     {"chunk_rows": 4, "max_live": 160, "max_ops": 1000, "occupancy": 1, "rows_per_step": 1}),
    (dict(n=5, seed=101, nodes=(2000, 5000), log2_bytes=(23, 25)), 3,
     {"chunk_rows": 16, "max_live": 160, "max_ops": 4000, "occupancy": 0, "rows_per_step": 4}),
]


@pytest.mark.parametrize("args,idx,sched", WIDE_REGRESSIONS)
def test_wide_schedules_match_oracle_without_spills(args, idx, sched):
    from bitstreamzoo_amd import schedsearch
    sp = schedsearch.random_specs(args["n"], args["seed"], nodes=args["nodes"], log2_bytes=args["log2_bytes"])[idx]
    p = dag_gen.generate(sp.nodes, seed=sp.seed)
    case = dag_gen.text_case(min(sp.n_bytes, 400_000), seed=sp.seed + 7)
    exp = oracle_outputs(p, case)
    c = ctx()
    plan = plan_or_refusal(c, api.Program(p.source), tuning.shape_case([case]), sched)
    if plan is None:
        return  # refused: its kernels would overflow the VGPRs (never run)
    res = tuning.Resident(c, [case])
    res.bind(plan)
    plan.run()
    nw = (case.bitlength + 63) // 64
    ks = plan.info()["kernels"]
    bad = [n for n in p.outputs if words_to_int(plan.download_output(0, n, nw)) != exp[n]]
    plan.close()
    res.free()
    assert all(k.get("scratch_bytes", 0) == 0 for k in ks), [k.get("scratch_bytes") for k in ks]
    assert not bad, f"{len(bad)} outputs differ"


def test_large_dag_matches_oracle():
    """A 20k-node generated DAG (about 25k DSL ops, hundreds of kernels and
    slots) at the default schedule."""
    p = dag_gen.generate(20_000, seed=21)
    case = dag_gen.text_case((1 << 18) + 777, seed=22)
    exp = oracle_outputs(p, case)
    tuning.precompile(p.source, [case], [{}], processes=8)
    c = ctx()
    res = tuning.Resident(c, [case])
    plan = api.Plan(c, api.Program(p.source), tuning.shape_case([case]), {})
    res.bind(plan)
    plan.run()
    nw = (case.bitlength + 63) // 64
    n_kernels = len(plan.info()["kernels"])
    bad = [n for n in p.outputs if words_to_int(plan.download_output(0, n, nw)) != exp[n]]
    plan.close()
    res.free()
    assert n_kernels > 100
    assert not bad, f"{len(bad)} of {len(p.outputs)} outputs differ"
