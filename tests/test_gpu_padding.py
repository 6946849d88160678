"""bsim_cuda's padding semantics on the MI355X (bsz_case_set_padding,
bsim_hip --padding=bsim).

With padding 64 or 32 the fused and per-op executors give bits at or above
the bitlength the values bsim_cuda gives them with --word-bits=64|32: ~,
ONES, + and << leave the stream's padding set, >> brings it back into range,
and inputs keep the bits above the bitlength that their last word holds.  The
oracle's BSIM and BSIM32 semantics are the reference (pinned against the
reference's CPU interpreter and a big-int model on the CPU).  Outputs must be
bit-exact, including zeros above the bitlength, with equal op counts."""

import numpy as np
import pytest

import progfuzz
from bitstreamzoo_amd import api, harness
from bitstreamzoo_amd.bsdata import BsCase, int_to_words, words_to_int, write_bsdata
from cpu_models import Oracle
from hip_models import assert_same, ctx

pytestmark = pytest.mark.gpu

SEM = {64: 1, 32: 2}


def _oracle(src, case, word_bits):
    names = Oracle(src).output_names()
    m = Oracle(src).load(case)
    m.set_semantics(SEM[word_bits])
    m.run(Oracle.SCALAR)
    return names, m.outputs(*names), m.op_count()


def _hip(src, case, names, word_bits, executor):
    res = ctx().run(api.Program(src), api.Case.from_bscase(case, padding=word_bits), executor)
    out = {}
    for n in names[0]:
        s = res.stream(n)
        if s is not None:
            out[n] = int.from_bytes(s.tobytes(), "little")
        a = res.array(n)
        if a is not None:
            out[n] = {i: int.from_bytes(w.tobytes(), "little") for i, w in a.items()}
    for n in names[1]:
        v = res.int_out(n)
        if v is not None:
            out[n] = v
    return out, res.op_count


def _no_bits_above(out: dict, bl: int, ints, label):
    for k, v in out.items():
        if k in ints:
            continue
        for x in (v.values() if isinstance(v, dict) else [v]):
            assert x >> bl == 0, f"{label}: {k} has bits at or above the bitlength"


@pytest.mark.parametrize("executor", [api.FUSED, api.PEROP], ids=["fused", "perop"])
@pytest.mark.parametrize("word_bits", [64, 32])
@pytest.mark.parametrize("seed", range(10))
def test_random_programs_with_bsim_padding(seed, word_bits, executor):
    bl = [1, 63, 70, 200, 257, 777, 4100, 70000, 4096 * 64 + 5, 300000][seed]
    src = progfuzz.random_program(700 + seed, n_stmts=30, allow_control=True, big_shifts=True)
    case = progfuzz.random_case(seed, bl)
    lw = (bl + 63) // 64
    rng = np.random.Generator(np.random.PCG64(seed))
    # x0's last word holds bits above the bitlength
    x0 = words_to_int(case.inputs["x0"]) | (int(rng.integers(1, 2**63)) << (64 * lw - 63))
    case.inputs["x0"] = int_to_words(x0 & ((1 << 64 * lw) - 1), lw)
    label = f"seed {seed}, bl {bl}, word_bits {word_bits}, {executor}"
    try:
        names, want, ops = _oracle(src, case, word_bits)
    except RuntimeError as e:   # e.g. the while-iteration limit: the run fails the same way
        names = Oracle(src).output_names()
        with pytest.raises(api.BszError, match=str(e)):
            _hip(src, case, names, word_bits, executor)
        return
    got, got_ops = _hip(src, case, names, word_bits, executor)
    assert_same(got, want, bl, set(names[1]), label)
    _no_bits_above(got, bl, set(names[1]), label)
    assert got_ops == ops, label


SHR_PAD = """input stream x
output stream y
output stream z
output int n
stream t = ~x
y = t >> 200
stream u = ONES + x
z = u >> 3
n = popcount(y)
"""


@pytest.mark.parametrize("executor", [api.FUSED, api.PEROP], ids=["fused", "perop"])
def test_padding_width_follows_word_bits(executor):
    """~x >> 200 at bitlength 70: zeros when clean, 56 ones with 32-bit words
    (padding to 256 bits), 70 ones with 64-bit words (padding to 512)."""
    case = BsCase(bitlength=70, inputs={"x": int_to_words(5, 2)})
    names = Oracle(SHR_PAD).output_names()
    got = {}
    for wb in (0, 32, 64):
        res = ctx().run(api.Program(SHR_PAD), api.Case.from_bscase(case, padding=wb), executor)
        got[wb] = (words_to_int(res.stream("y")), res.int_out("n"))
    assert got[0] == (0, 0)
    assert got[32] == ((1 << 56) - 1, 56)
    assert got[64] == ((1 << 70) - 1, 70)
    for wb in (32, 64):
        _, want, _ = _oracle(SHR_PAD, case, wb)
        o, _ = _hip(SHR_PAD, case, names, wb, executor)
        assert o == want


def test_cli_padding_flag(tmp_path):
    """bsim_hip --padding=bsim --word-bits=32 reads a .bsdata file and gives
    the oracle's BSIM32 outputs."""
    src = progfuzz.random_program(901, n_stmts=30, allow_control=True, big_shifts=True)
    case = progfuzz.random_case(3, 777)
    bs = tmp_path / "p.bs"
    bs.write_text(src)
    data = tmp_path / "p.bsdata"
    write_bsdata(str(data), [{"bitlength": case.bitlength, "params": dict(case.params), "inputs": dict(case.inputs),
                              "input_arrays": {k: dict(v) for k, v in case.input_arrays.items()}}])
    for wb in (32, 64):
        names, want, ops = _oracle(src, case, wb)
        be = harness.HipBackend(word_bits=wb, padding="bsim")
        out, got_ops, _ = be.run(harness.ProgramRef(str(bs)), {}, bsdata_path=str(data))
        assert_same(out, want, case.bitlength, set(names[1]), f"cli word_bits {wb}")
        assert got_ops == ops
