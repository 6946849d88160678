"""The measured autotuner on the GPU: every candidate schedule gives the same
outputs (checked inside autotune) and the pick is the fastest measured."""

import pytest

from bitstreamzoo_amd import api, tuning, workloads
from hip_models import ctx as _ctx

pytestmark = pytest.mark.gpu


def test_autotune_adder_small():
    ctx = _ctx()
    w = workloads.adder32(1 << 16, seed=11)
    res = tuning.Resident(ctx, w.cases)
    outs = [f"s{i}" for i in range(32)] + ["cout"]
    space = [{"chunk_rows": 4, "rows_per_step": 1, "occupancy": 2},
             {"chunk_rows": 8, "rows_per_step": 2, "occupancy": 1},
             {"chunk_rows": 16, "rows_per_step": 1, "occupancy": 4, "waves_per_block": 2}]
    tr = tuning.autotune(ctx, api.Program(w.source), [res], space, ["n_" + o for o in outs], outs, steps=3, warmup=1)
    # two rows of 65 inputs in flight at occupancy 1 would overflow into
    # AGPRs (the code object's metadata says so): the library regenerates
    # that kernel with one row per step, so every schedule is timed clean
    assert [t.options for t in tr.trials] == space and not tr.skipped
    assert tr.trials[1].info["kernels"][0]["rows_per_step"] == 1
    assert all(k["agprs"] == 0 and k["scratch_bytes"] == 0 for t in tr.trials for k in t.info["kernels"])
    fastest = min(t.ms for t in tr.trials)
    assert tr.best.ms <= fastest * (1 + tuning.TuneResult.TIE)
    assert tr.best is next(t for t in tr.trials if t.ms <= fastest * (1 + tuning.TuneResult.TIE))
    assert all(t.ms > 0 for t in tr.trials)
    res.free()


def test_autotune_matchstar_rotations():
    ctx = _ctx()
    ws = [workloads.matchstar(64 * 64 * 40 + 5, seed=s) for s in (1, 2)]
    residents = [tuning.Resident(ctx, w.cases) for w in ws]
    space = tuning.default_space("stateful")[:6]
    tr = tuning.autotune(ctx, api.Program(ws[0].source), residents, space, ["n"], ["result"], steps=2, warmup=1)
    # rows per step divide the chunk: 8 rows per step over 4-row chunks is
    # the 4-row schedule again, and is timed once
    assert len(tr.trials) + len(tr.same_as) == 6 and len(tr.trials) == 4
    # the chosen schedule reproduces the independent big-int result
    plan = api.Plan(ctx, api.Program(ws[1].source), tuning.shape_case(ws[1].cases), tr.best.options)
    residents[1].bind(plan)
    plan.run()
    c = ws[1].cases[0]
    m = (1 << c.bitlength) - 1
    M = int.from_bytes(c.inputs["M"].tobytes(), "little")
    C = int.from_bytes(c.inputs["C"].tobytes(), "little")
    exp = ((((M & C) + C) & m) ^ C) | M
    got = plan.download_output(0, "result", len(c.inputs["M"]))
    assert int.from_bytes(got.tobytes(), "little") == exp
    assert plan.read_int(0, "n") == bin(exp).count("1")
    plan.close()
    for r in residents:
        r.free()
