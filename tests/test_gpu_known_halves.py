"""Fused kernels over values the compiler can partly evaluate.

An advance or a sum of a value the compiler proves constant (`a ^ a`,
`a & ~a`, `x + x` of such a value) has a 32-bit half the compiler knows.
After a NOT that half is known to be all ones, and the AMDGPU SelectionDAG
backend of the ROCm 7.2 compiler then drops it from a 64-bit OR that shares
the value (the OR keeps the other operand's half).  The generated kernels
hide every advance and sum behind an empty asm so no value carries a known
half; these programs, run fused under several schedules, must match the
oracle bit for bit.  The first program is the shape of the generated-DAG
program in which the wrong bit was found."""

import numpy as np
import pytest

from bitstreamzoo_amd.bsdata import BsCase
from hip_models import assert_same
from test_gpu_fused import oracle_outs, run_plan

pytestmark = pytest.mark.gpu

PROGRAMS = {
    "advance_of_xor_self": """
input stream a
input stream b
input stream c
output stream n
output stream o1
output stream o2
stream z = a ^ a
stream y = z << 3
n = ~y
stream t = b | c
o1 = t | n
o2 = c | n
""",
    "sum_of_zeros": """
input stream a
input stream b
output stream n
output stream o1
output stream o2
stream na = ~a
stream z = a & na
stream s = z + z
n = ~s
o1 = b | n
o2 = a | n
""",
    "advance_by_32_and_61": """
input stream a
input stream b
output stream o1
output stream o2
output stream o3
stream z = a ^ a
stream y = z << 61
stream x = z << 32
stream ny = ~y
stream nx = ~x
o1 = b | ny
o2 = a | ny
o3 = b | nx
""",
}

SCHEDULES = [{}, {"rows_per_step": 1}, {"chunk_rows": 4, "rows_per_step": 2, "warmup_rows": 2},
             {"occupancy": 4}]


def case_for(src: str, bitlength: int, seed: int) -> BsCase:
    rng = np.random.Generator(np.random.PCG64(seed))
    nw = (bitlength + 63) // 64
    names = [ln.split()[2] for ln in src.splitlines() if ln.startswith("input stream")]
    inputs = {}
    for n in names:
        w = rng.integers(0, 2**64, size=nw, dtype=np.uint64)
        if bitlength % 64:
            w[-1] &= np.uint64((1 << (bitlength % 64)) - 1)
        inputs[n] = w
    return BsCase(bitlength=bitlength, inputs=inputs)


@pytest.mark.parametrize("sched", SCHEDULES, ids=[str(s) for s in SCHEDULES])
@pytest.mark.parametrize("name", sorted(PROGRAMS))
@pytest.mark.parametrize("bl", [56593, 64 * 64 * 33 + 5])
def test_known_halves_match_oracle(name, sched, bl):
    src = PROGRAMS[name]
    case = case_for(src, bl, seed=bl + len(name))
    ora, names = oracle_outs(src, case)
    got, _ = run_plan(src, [case], sched, names)
    assert_same(got[0], ora, bl, set(names[1]), f"{name} {sched}")
