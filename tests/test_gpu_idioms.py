"""Program shapes the reference's applications use, written here from their
descriptions: a parity (prefix XOR) built by a `for` loop of doubling
advances, a fixpoint `while` loop whose body holds a MatchStar and an
advance, and a `for` loop over an input array with an index-dependent
advance.  Both executors against the oracle, outputs and op counts."""

import numpy as np
import pytest

from bitstreamzoo_amd.bsdata import BsCase
from cpu_models import Oracle
from hip_models import assert_same, hip_run
from runner import run_model

pytestmark = pytest.mark.gpu

PREFIX_XOR = """
input stream q
param int LOG
output stream par
output int n
stream p = q
int s = 1
for i in 0..LOG {
    stream t = p << s
    p = p ^ t
    s = s * 2
}
par = p
n = popcount(par)
"""

FIXPOINT = """
input stream marker
input stream cls
output stream reach
output int steps
output int n
stream cur = marker
stream test = marker
int k = 0
while (test) {
    stream a = cur & cls
    stream b = a + cls
    stream c = b ^ cls
    stream nxt = c | cur
    stream adv = nxt << 1
    stream grow = adv & cls
    stream more = grow | nxt
    test = more ^ cur
    cur = more
    k = k + 1
}
reach = cur
steps = k
n = popcount(reach)
"""

ROW_SCAN = """
input stream rows[6]
param int K
output stream hit
output int n
stream acc = ONES
for i in 0..K {
    stream r = rows[i]
    stream shifted = r << i
    acc = acc & shifted
}
hit = acc
n = popcount(hit)
"""


def words(rng, bl, density):
    nw = (bl + 63) // 64
    bits = rng.random(nw * 64) < density
    w = np.packbits(bits.reshape(-1, 8)[:, ::-1]).view(np.uint64).copy()
    if bl % 64:
        w[-1] &= np.uint64((1 << (bl % 64)) - 1)
    return w


def case_for(name, bl, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    if name == "prefix_xor":
        return BsCase(bitlength=bl, inputs={"q": words(rng, bl, 0.05)}, params={"LOG": 6})
    if name == "fixpoint":
        return BsCase(bitlength=bl, inputs={"marker": words(rng, bl, 0.01), "cls": words(rng, bl, 0.7)})
    return BsCase(bitlength=bl, input_arrays={"rows": {i: words(rng, bl, 0.9) for i in range(6)}},
                  params={"K": 6})


SOURCES = {"prefix_xor": PREFIX_XOR, "fixpoint": FIXPOINT, "row_scan": ROW_SCAN}


@pytest.mark.parametrize("executor", ["fused", "perop"])
@pytest.mark.parametrize("bl", [1000, 64 * 64 * 20 + 7, 300_001])
@pytest.mark.parametrize("name", sorted(SOURCES))
def test_idiom_matches_oracle(name, bl, executor):
    src = SOURCES[name]
    case = case_for(name, bl, seed=bl)
    names = Oracle(src).output_names()
    ora, ora_ops = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
    hip, hip_ops = hip_run(src, case, executor, names[0], names[1])
    assert_same(hip, ora, bl, set(names[1]), f"{name} {executor}")
    assert hip_ops == ora_ops
