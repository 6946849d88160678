"""Whole programs through the C ABI (bsz_run) against the oracle: the
benchmark workloads at small sizes and seeded random programs that use every
DSL construct.  Bit-exact on every output stream and int."""

import numpy as np
import pytest

import progfuzz
from bitstreamzoo_amd import workloads
from cpu_models import Oracle
from hip_models import EXECUTORS, assert_same, hip_run
from runner import run_model

pytestmark = pytest.mark.gpu

EXEC = list(EXECUTORS)


def check(src, case, executor, label):
    names = Oracle(src).output_names()
    ora, ora_ops = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
    hip, hip_ops = hip_run(src, case, executor, *names)
    assert_same(hip, ora, case.bitlength, set(names[1]), label)
    assert hip_ops == ora_ops, f"{label}: op_count {hip_ops} != {ora_ops}"


@pytest.mark.parametrize("executor", EXEC)
@pytest.mark.parametrize("bl", [1, 100, 4096 + 17, 65536 * 3 + 5, 1 << 20])
def test_matchstar(executor, bl):
    w = workloads.matchstar(bl, seed=bl)
    check(w.source, w.cases[0], executor, f"matchstar {bl}")


@pytest.mark.parametrize("executor", EXEC)
@pytest.mark.parametrize("bl", [33, 3000, 70001])
def test_adder32(executor, bl):
    w = workloads.adder32(bl, seed=bl)
    check(w.source, w.cases[0], executor, f"adder32 {bl}")


@pytest.mark.parametrize("executor", EXEC)
def test_shift_and_matches_scalar(executor):
    rng = np.random.Generator(np.random.PCG64(11))
    pattern = "".join(workloads.BASES[i] for i in rng.integers(0, 4, size=32))
    src = workloads.shift_and_source(pattern)
    case = workloads.dna_case(rng, 20000, "s")
    # plant three occurrences
    codes = workloads.onehot_to_codes([case.inputs[f"c{b}"] for b in workloads.BASES], 20000)
    pc = np.array([workloads.BASES.index(ch) for ch in pattern], dtype=np.uint8)
    for pos in (0, 5000, 20000 - 32):
        codes[pos:pos + 32] = pc
    for k, b in enumerate(workloads.BASES):
        case.inputs[f"c{b}"] = workloads.pack_bits(codes == k, 20000)
    check(src, case, executor, "shift_and")
    hip, _ = hip_run(src, case, executor, ["match"], ["n"])
    hits = workloads.shift_and_scalar(codes, pattern)
    m = hip["match"]
    got = [j for j in range(20000) if (m >> j) & 1]
    assert got == hits.tolist() and hip["n"] == len(hits) >= 3


@pytest.mark.parametrize("executor", EXEC)
@pytest.mark.parametrize("seed", range(40))
def test_random_programs(executor, seed):
    bl = [1, 64, 65, 200, 777, 4096, 5000][seed % 7]
    src = progfuzz.random_program(seed)
    case = progfuzz.random_case(seed, bl)
    check(src, case, executor, f"fuzz seed={seed} bl={bl}")
