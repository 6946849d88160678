"""The two generated-DAG workloads at their specified shapes, against the
oracle.

* The 200k-node random regex-style DAG (250k DSL ops, window 256, 64
  outputs) at the benchmark's schedule (4 warm-up rows), over a 2^20-byte
  Zipf text compared with the oracle, and over the full 2^27-byte text in
  the same launch (two instances of one plan).  Every op of the generator
  (and, or, xor, not, advance, add, MatchStar, ScanThru) moves information
  only toward higher bit positions, so the first 2^20 bits of every output
  of the full text equal the outputs of its 2^20-byte prefix: a size-
  independent check of the full run.
* A mixed batch of 256 generated programs (their own sizes and texts), all
  plans' launches captured into one hipGraph and replayed; every program's
  outputs checked against the oracle on every replay."""

import math

import numpy as np
import pytest

from bitstreamzoo_amd import api, dag_gen, tuning
from bitstreamzoo_amd.bsdata import words_to_int
from cpu_models import Oracle
from hip_models import ctx
from runner import run_model

pytestmark = pytest.mark.gpu


def oracle_outputs(p, case):
    names = Oracle(p.source).output_names()
    o, _ = run_model(Oracle, p.source, case, Oracle.BEST, stream_names=names[0], int_names=names[1])
    return o


def test_config4_dag_200k_nodes_matches_oracle_and_prefix():
    p = dag_gen.generate(200_000, seed=4)
    assert p.n_nodes == 200_000 and p.n_dsl_ops > 240_000
    small_bytes, big_bytes = 1 << 20, 1 << 27
    small = dag_gen.text_case(small_bytes, seed=11)
    big = dag_gen.text_case(big_bytes, seed=11)
    for j in range(8):   # the small text is the big one's prefix
        assert np.array_equal(big.inputs[f"b{j}"][: small_bytes // 64], small.inputs[f"b{j}"])
    exp = oracle_outputs(p, small)
    opts = {"warmup_rows": 4}
    tuning.precompile(p.source, [small], [opts], processes=16, shards=64)
    c = ctx()
    res = tuning.Resident(c, [small, big])
    plan = api.Plan(c, api.Program(p.source), tuning.shape_case([small]), opts)
    res.bind(plan)
    plan.run()
    c.sync()
    nw = small_bytes // 64
    bad = [o for o in p.outputs if words_to_int(plan.download_output(0, o, nw)) != exp[o]]
    assert not bad, f"{len(bad)} of {len(p.outputs)} outputs differ from the oracle on 2^20 bytes"
    pre = [o for o in p.outputs
           if not np.array_equal(plan.download_output(1, o, nw), plan.download_output(0, o, nw))]
    assert not pre, f"{len(pre)} outputs of the 2^27-byte run differ from the prefix run in their first 2^20 bits"
    assert len(plan.info()["kernels"]) > 1000
    plan.close()
    res.free()


def test_config5_mixed_batch_of_256_programs_in_one_graph():
    rng = np.random.Generator(np.random.PCG64(55))
    progs, cases = [], []
    for i in range(256):
        nodes = int(math.exp(rng.uniform(math.log(120), math.log(600))))
        nb = int(2 ** rng.uniform(12, 16))
        progs.append(dag_gen.generate(nodes, seed=5000 + i, n_outputs=8))
        cases.append(dag_gen.text_case(nb, seed=7000 + i))
    tuning.precompile_many([(p.source, [cs], {}) for p, cs in zip(progs, cases)], processes=16)
    c = ctx()
    plans, residents = [], []
    for p, cs in zip(progs, cases):
        r = tuning.Resident(c, [cs])
        pl = api.Plan(c, api.Program(p.source), tuning.shape_case([cs]), {})
        r.bind(pl)
        plans.append(pl)
        residents.append(r)
    exps = [oracle_outputs(p, cs) for p, cs in zip(progs, cases)]
    g = c.capture(lambda: [pl.run() for pl in plans])
    for replay in range(2):
        g.launch(1)
        c.sync()
        bad = []
        for i, (p, cs, pl, exp) in enumerate(zip(progs, cases, plans, exps)):
            nw = (cs.bitlength + 63) // 64
            bad += [(i, o) for o in p.outputs if words_to_int(pl.download_output(0, o, nw)) != exp[o]]
        assert not bad, f"replay {replay}: {len(bad)} outputs differ, first {bad[:4]}"
    for pl in plans:
        pl.close()
    for r in residents:
        r.free()
