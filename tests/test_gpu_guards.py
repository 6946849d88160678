"""Guards of the product path on the GPU: schedules whose kernels overflow
the architectural VGPRs are refused by every entry (plan creation, the fused
engine behind bsz_run and the bsim_hip CLI), and graph captures larger than
the library's limit are refused before anything is captured, leaving the
stream usable."""

import os
import subprocess

import numpy as np
import pytest

from bitstreamzoo_amd import _lib, api, tuning, workloads
from bitstreamzoo_amd.bsdata import write_bsdata
from hip_models import ctx as _ctx

pytestmark = pytest.mark.gpu

# an explicit occupancy 1 and one row per step: a wide partition of this
# schedule still needs AGPRs, so nothing is left to shrink
UNCLEAN = {"chunk_rows": 8, "rows_per_step": 1, "max_live": 160, "occupancy": 1}


def _dag():
    from bitstreamzoo_amd import dag_gen
    return dag_gen.generate(1200, seed=2), dag_gen.text_case(5000, seed=2)


def test_plan_create_refuses_overflowing_schedule():
    ctx = _ctx()
    p, case = _dag()
    with pytest.raises(api.BszError) as e:
        api.Plan(ctx, api.Program(p.source), tuning.shape_case([case]), UNCLEAN)
    assert e.value.code == -3 and "schedule refused" in str(e.value)


def test_cli_refuses_overflowing_schedule(tmp_path):
    p, c = _dag()
    bs = tmp_path / "dag.bs"
    bs.write_text(p.source)
    data = tmp_path / "dag.bsdata"
    write_bsdata(str(data), [{"bitlength": c.bitlength, "inputs": dict(c.inputs),
                              "input_arrays": {k: dict(v) for k, v in c.input_arrays.items()}}])
    env = dict(os.environ, BSZ_OCCUPANCY="1", BSZ_ROWS_PER_STEP="1", BSZ_CHUNK_ROWS="8", BSZ_MAX_LIVE="160")
    r = subprocess.run([_lib.CLI_PATH, str(bs), "--input", str(data)], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 1 and "schedule refused" in r.stderr, (r.returncode, r.stderr[-500:])


def test_large_perop_capture_is_refused_and_plan_still_runs():
    """450 Shift-And instances x 63 per-op launches exceed the 20k-launch
    graph limit: bsz_plan_capture fails before capturing, and the same plan
    still runs eagerly with the right popcounts."""
    ctx = _ctx()
    w = workloads.shift_and_batch(n_seqs=450, seed=5, min_log2=8, max_log2=10)
    res = tuning.Resident(ctx, w.cases)
    plan = api.Plan(ctx, api.Program(w.source), tuning.shape_case(w.cases), {})
    res.bind(plan)
    plan.set_executor(api.PEROP)
    with pytest.raises(api.BszError) as e:
        plan.capture()
    assert e.value.code == -3 and "refused" in str(e.value)
    plan.run()
    ctx.sync()
    plan.set_executor(api.FUSED)
    fused = [plan.read_int(i, "n") for i in (0, 449)]
    plan.run()
    ctx.sync()
    plan.set_executor(api.PEROP)
    plan.run()
    ctx.sync()
    assert [plan.read_int(i, "n") for i in (0, 449)] == fused
    plan.close()
    res.free()


def test_context_capture_stops_at_the_limit():
    """Ops recorded between bsz_capture_begin and bsz_capture_end are refused
    past the limit; the capture still ends and the stream stays usable."""
    ctx = _ctx()
    L = api.lib()
    a = ctx.upload(np.arange(64, dtype=np.uint64), 4096)
    b = ctx.alloc(4096)
    _lib.check(L.bsz_capture_begin(ctx.h), ctx.h)
    rc = 0
    n = 0
    for n in range(20001):
        rc = L.bsz_op_not(ctx.h, b.dev, a.dev, 4096)
        if rc != 0:
            break
    g = api.C.c_void_p()
    end_rc = L.bsz_capture_end(ctx.h, api.C.byref(g))
    assert rc == -3 and n == 20000, (rc, n)
    assert end_rc == 0
    L.bsz_graph_destroy(g)
    assert ctx.popcount(ctx.op("not", a)) == 4096 - sum(bin(i).count("1") for i in range(64))
    a.free()
    b.free()
