"""`if` bodies kept on the device (if-conversion in the fused executor).

A program shaped like the reference's regex matchers, with hundreds of
`if (cursor) { ... }` guards nested up to three deep, runs as fused kernels
without a host decision per `if`: the walk records every body, each
assigned stream becomes a per-word selection on a device predicate, and the
host reads the predicates once at the end for the op count.  Outputs, ints,
the dynamic op count and the variable count must equal the oracle's, which
runs the bodies only where the condition holds."""

import pytest

import progfuzz
import regex_ifs
from bitstreamzoo_amd import api, tuning
from cpu_models import Oracle
from hip_models import assert_same, ctx
from runner import run_model

pytestmark = pytest.mark.gpu


def _oracle(src, case):
    names = Oracle(src).output_names()
    o, ops = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
    m = Oracle(src).load(case)
    m.run(Oracle.SCALAR)
    return names, o, ops


def _hip(src, case, names, c=None):
    res = (c or ctx()).run(api.Program(src), api.Case.from_bscase(case), api.FUSED)
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
    return out, res.op_count, res.json()


@pytest.fixture(scope="module")
def regex_program_compiled():
    """The kernels of the 200- and 60-pattern programs, compiled up front by
    16 worker processes (the plan bsz_run builds, recorded with its `if`
    bodies as device selections), so no test spends minutes in hipRTC."""
    for n_patterns, seed in ((200, 7), (60, 9)):
        src = regex_ifs.program(n_patterns, seed=seed)
        case = regex_ifs.text_case(5000, seed=3, alphabet=3)
        tuning.precompile(src, [case], [{"device_if": 1}], processes=16)


@pytest.mark.parametrize("n_bytes,alphabet", [(20000, 6), (20000, 3), (4096 * 64 + 7, 4)])
def test_regex_shaped_program_runs_fused_without_host_decisions(regex_program_compiled, n_bytes, alphabet):
    """The same 200-pattern program (its kernels compile once) over texts
    where most guards hold, few hold, and over 4096 rows."""
    src = regex_ifs.program(200, seed=7)
    case = regex_ifs.text_case(n_bytes, seed=alphabet, alphabet=alphabet)
    n_if = src.count("if (")
    names, ora, ora_ops = _oracle(src, case)
    with api.Context(0) as c:   # a fresh engine: its flush count is this run's
        hip, hip_ops, js = _hip(src, case, names, c)
    assert_same(hip, ora, case.bitlength, set(names[1]), f"regex-shaped, {n_if} ifs")
    assert hip_ops == ora_ops
    # one flush at the end: no host round trip per `if`
    assert js["engine"]["flushes"] == 1, js["engine"]
    assert n_if >= 300


def test_variable_count_follows_the_taken_branches(regex_program_compiled):
    """Streams first declared inside a body exist only where the reference ran
    the body: the reported variable count equals the per-op executor's,
    which decides every `if` on the host."""
    src = regex_ifs.program(60, seed=9)
    case = regex_ifs.text_case(5000, seed=3, alphabet=3)
    fused = ctx().run(api.Program(src), api.Case.from_bscase(case), api.FUSED).json()
    perop = ctx().run(api.Program(src), api.Case.from_bscase(case), api.PEROP).json()
    assert fused["n_variables"] == perop["n_variables"]
    assert fused["dynamic_ops"] == perop["dynamic_ops"]
    assert set(fused["outputs"]) == set(perop["outputs"])


@pytest.mark.parametrize("seed", range(60, 90))
def test_random_programs_with_ifs(seed):
    """Random programs (nested if / while / for, arrays, ints read from
    popcounts) through the fused executor: converted `if`s, and `if`s that
    fall back to the host (bodies with loops or int assignments), agree
    with the oracle on outputs and op counts."""
    bl = [64, 200, 777, 5000, 70000][seed % 5]
    src = progfuzz.random_program(seed, n_stmts=40)
    case = progfuzz.random_case(seed, bl)
    try:
        names, ora, ora_ops = _oracle(src, case)
    except RuntimeError as e:   # e.g. the while-iteration limit: the run fails the same way
        names = Oracle(src).output_names()
        with pytest.raises(api.BszError, match=str(e)):
            _hip(src, case, names)
        return
    hip, hip_ops, _ = _hip(src, case, names)
    assert_same(hip, ora, bl, set(names[1]), f"seed {seed}")
    assert hip_ops == ora_ops
