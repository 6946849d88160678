"""Fused plans (bsz_plan_*) against the oracle: tile shapes that cut streams
into many chunks, carry chains that defeat the speculative warm-up so chunks
look back at their predecessor's end state and run again, three-input
boolean folding, advances of in-kernel values across
rows and chunks, multi-kernel plans with intermediate buffers, batches of
instances of uneven lengths in one launch, and hipGraph replay."""

import numpy as np
import pytest

import progfuzz
from bitstreamzoo_amd import api, workloads
from bitstreamzoo_amd.bsdata import BsCase, words_to_int
from cpu_models import Oracle
from hip_models import assert_same, ctx
from runner import run_model

pytestmark = pytest.mark.gpu

TILES = [{"chunk_rows": 1, "rows_per_step": 1, "warmup_rows": 1},
         {"chunk_rows": 4, "rows_per_step": 2, "warmup_rows": 2},
         {"chunk_rows": 16, "rows_per_step": 4, "warmup_rows": 1},
         {"chunk_rows": 8, "rows_per_step": 4, "warmup_rows": 1, "max_ops": 3, "max_live": 4},
         # one step holds the warm-up rows and the chunk (their loads issued together)
         {"chunk_rows": 4, "rows_per_step": 8, "warmup_rows": 1, "join_warmup": 1},
         {"chunk_rows": 2, "rows_per_step": 8, "warmup_rows": 3, "join_warmup": 1}]


def run_plan(src, cases, opts, names, replay=0):
    c = ctx()
    plan = api.Plan(c, api.Program(src), api.Case.from_bscase(cases[0]), opts)
    plan.set_instances([k.bitlength for k in cases])
    keep = []
    for i, k in enumerate(cases):
        for n, w in k.inputs.items():
            d = c.upload(w, k.bitlength)
            keep.append(d)
            plan.bind(i, n, d)
        for n, arr in k.input_arrays.items():
            for j, w in arr.items():
                d = c.upload(w, k.bitlength)
                keep.append(d)
                plan.bind(i, n, d, j)
    if replay:
        plan.capture()
        plan.replay(replay)
    else:
        plan.run()
    c.sync()
    info = plan.info()
    outs = []
    for i, k in enumerate(cases):
        o = {}
        nw = max(1, (k.bitlength + 63) // 64)
        for n, idx, _kind in info["outputs"]:
            v = words_to_int(plan.download_output(i, n, nw, idx))
            if idx < 0:
                o[n] = v
            else:
                o.setdefault(n, {})[idx] = v
        for n in info["ints"]:
            o[n] = plan.read_int(i, n)
        outs.append(o)
    mism = [plan.mismatches(kk) for kk in range(len(plan.info()["kernels"]))]
    for d in keep:
        d.free()
    plan.close()
    return outs, mism


def oracle_outs(src, case):
    names = Oracle(src).output_names()
    o, _ = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
    return o, names


@pytest.mark.parametrize("tile", TILES, ids=[str(t) for t in TILES])
@pytest.mark.parametrize("bl", [64, 4096 * 3 + 7, 64 * 64 * 40 + 1])
def test_matchstar_tiles(tile, bl):
    w = workloads.matchstar(bl, seed=bl + 1)
    ora, names = oracle_outs(w.source, w.cases[0])
    got, _ = run_plan(w.source, w.cases, tile, names)
    assert_same(got[0], ora, bl, set(names[1]), f"matchstar {tile}")


CARRY_SRC = """input stream a
input stream b
input stream c
output stream s
output stream t
output stream u
output int ns
output int nu
stream x = a + b
s = x
stream y = x << 3
stream z = y + c
stream v = z << 61
stream q = v ^ x
t = q + a
stream r = ~q
u = r + ONES
ns = popcount(s)
nu = popcount(u)
"""


def carry_case(bl, kind, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    nw = (bl + 63) // 64
    ones = np.full(nw, np.uint64(2**64 - 1))
    if kind == "chain":      # one carry generated at word 0 ripples through every word
        a = ones.copy(); b = np.zeros(nw, np.uint64); b[0] = 1
        c = rng.integers(0, 2**64, nw, dtype=np.uint64)
    elif kind == "blocks":   # all-propagate stretches longer than a chunk, with carries entering them
        a = rng.integers(0, 2**64, nw, dtype=np.uint64); b = rng.integers(0, 2**64, nw, dtype=np.uint64)
        for s0 in range(3, nw, 5000):
            e = min(nw, s0 + 3000)
            a[s0:e] = ones[s0:e]; b[s0:e] = 0; a[s0 - 1] = np.uint64(2**64 - 1); b[s0 - 1] |= np.uint64(1)
        c = ones.copy()
    else:
        a = rng.integers(0, 2**64, nw, dtype=np.uint64); b = rng.integers(0, 2**64, nw, dtype=np.uint64)
        c = rng.integers(0, 2**64, nw, dtype=np.uint64)
    m = np.uint64((1 << (bl % 64)) - 1) if bl % 64 else None
    for x in (a, b, c):
        if m is not None:
            x[-1] &= m
    return BsCase(bitlength=bl, inputs={"a": a, "b": b, "c": c})


@pytest.mark.parametrize("tile", TILES, ids=[str(t) for t in TILES])
@pytest.mark.parametrize("kind", ["random", "chain", "blocks"])
def test_chained_adds_and_advances(tile, kind):
    bl = 64 * 64 * 97 + 13
    case = carry_case(bl, kind, 5)
    ora, names = oracle_outs(CARRY_SRC, case)
    got, mism = run_plan(CARRY_SRC, [case], tile, names)
    assert_same(got[0], ora, bl, set(names[1]), f"carry {kind} {tile}")
    if kind == "chain" and tile["chunk_rows"] <= 4 and tile.get("max_ops", 100) > 3:
        assert sum(mism) > 0  # the speculation was wrong and chunks ran again after their look-back


def test_lookback_wait_bound_reports_an_error():
    """A look-back wait is bounded: with a bound of one poll, a chunk whose
    predecessor has not published its end state yet gives up at once, and
    the run reports an error instead of outputs.  (One carry rippling
    through thousands of 1-row chunks: no warm-up can prove a start state,
    and the end states are published one after another.)  The same plan
    shape with the library's bound is exact."""
    bl = 64 * 64 * 6000 + 13
    case = carry_case(bl, "chain", 5)
    ora, names = oracle_outs(CARRY_SRC, case)
    tile = {"chunk_rows": 1, "rows_per_step": 1, "warmup_rows": 1}
    with pytest.raises(RuntimeError, match="look-back wait"):
        run_plan(CARRY_SRC, [case], dict(tile, spin_limit=0), names)
    got, mism = run_plan(CARRY_SRC, [case], tile, names)
    assert_same(got[0], ora, bl, set(names[1]), "carry chain, default bound")
    assert sum(mism) > 0


def test_batched_shift_and_instances():
    w = workloads.shift_and_batch(n_seqs=24, min_log2=1, max_log2=17, seed=9)
    names = Oracle(w.source).output_names()
    got, _ = run_plan(w.source, w.cases, TILES[2], names)
    for i, case in enumerate(w.cases):
        ora, _ = run_model(Oracle, w.source, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
        assert_same(got[i], ora, case.bitlength, set(names[1]), f"seq {i} L={case.bitlength}")


def test_graph_replay_is_stable():
    w = workloads.matchstar(64 * 64 * 50 + 3, seed=4)
    ora, names = oracle_outs(w.source, w.cases[0])
    got, _ = run_plan(w.source, w.cases, TILES[2], names, replay=5)
    assert_same(got[0], ora, w.cases[0].bitlength, set(names[1]), "replay")


@pytest.mark.parametrize("seed", range(24))
def test_random_straight_line_plans(seed):
    src = progfuzz.random_program(seed, n_stmts=40, allow_control=False, host_ints=False)
    case = progfuzz.random_case(seed, [777, 4096 * 2 + 5, 64 * 64 * 9 + 1][seed % 3])
    case.params["P"] = case.params["P"] % 64
    names = Oracle(src).output_names()
    try:
        api.compile_plan_only(api.Program(src), api.Case.from_bscase(case))
    except Exception:
        pytest.skip("program reads a popcount on the host")
    ora, _ = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
    got, _ = run_plan(src, [case], TILES[seed % len(TILES)], names)
    assert_same(got[0], ora, case.bitlength, set(names[1]), f"seed {seed}")


WIDE_TILES = [{"chunk_rows": 2, "rows_per_step": 1, "words_per_lane": 2},
              {"chunk_rows": 16, "rows_per_step": 2, "words_per_lane": 2},
              {"chunk_rows": 32, "rows_per_step": 8, "words_per_lane": 2, "occupancy": 2}]


@pytest.mark.parametrize("tile", WIDE_TILES, ids=[str(t) for t in WIDE_TILES])
@pytest.mark.parametrize("bl", [1, 64 * 63 + 5, 64 * 64 * 7 + 64 * 3, 64 * 64 * 33 + 1])
def test_two_words_per_lane_adder(tile, bl):
    """16-byte lanes over an odd number of rows, a last row shorter than a
    lane pair, and chunks that end mid-stream."""
    w = workloads.adder32(bl, seed=bl + 1, width=8)
    ora, names = oracle_outs(w.source, w.cases[0])
    got, _ = run_plan(w.source, w.cases, tile, names)
    assert_same(got[0], ora, bl, set(names[1]), f"adder8 {tile}")


def test_two_words_per_lane_is_used_and_batches():
    """The adder's kernels take two words per lane; a batch of uneven
    instances (each its own chunk range) matches the oracle per instance."""
    cases = [workloads.adder32(n, seed=n, width=4).cases[0] for n in (100, 64 * 64 * 2 + 1, 64 * 70, 5)]
    src = workloads.adder32_source(4)
    plan = api.Plan(ctx(), api.Program(src), api.Case.from_bscase(cases[0]), WIDE_TILES[1])
    assert all(k["words_per_lane"] == 2 for k in plan.info()["kernels"])
    plan.close()
    names = Oracle(src).output_names()
    got, _ = run_plan(src, cases, WIDE_TILES[1], names)
    for i, case in enumerate(cases):
        ora, _ = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
        assert_same(got[i], ora, case.bitlength, set(names[1]), f"instance {i}")


@pytest.mark.parametrize("seed", range(12))
def test_random_straight_line_plans_two_words_per_lane(seed):
    """Random programs (popcounts included); kernels with carries or
    advances fall back to one word per lane within the same plan."""
    src = progfuzz.random_program(100 + seed, n_stmts=40, allow_control=False, host_ints=False)
    case = progfuzz.random_case(100 + seed, [777, 4096 * 2 + 5, 64 * 64 * 9 + 1][seed % 3])
    case.params["P"] = case.params["P"] % 64
    names = Oracle(src).output_names()
    try:
        api.compile_plan_only(api.Program(src), api.Case.from_bscase(case))
    except Exception:
        pytest.skip("program reads a popcount on the host")
    ora, _ = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
    got, _ = run_plan(src, [case], WIDE_TILES[seed % len(WIDE_TILES)], names)
    assert_same(got[0], ora, case.bitlength, set(names[1]), f"seed {seed}")


INTERLEAVED = [{"chunk_rows": 4, "rows_per_step": 1, "interleave": 1},
               {"chunk_rows": 16, "rows_per_step": 4, "interleave": 1},
               {"chunk_rows": 64, "rows_per_step": 1, "interleave": 1, "occupancy": 2}]


@pytest.mark.parametrize("tile", INTERLEAVED, ids=[str(t) for t in INTERLEAVED])
@pytest.mark.parametrize("bl", [1, 64 * 63 + 5, 64 * 64 * 7 + 64 * 3, 64 * 64 * 97 + 13])
def test_interleaved_rows_adder(tile, bl):
    """Chunks that take every n-th row: a last chunk with fewer rows, rows per
    step past the stream's end, popcounts per chunk."""
    w = workloads.adder32(bl, seed=bl + 3, width=8)
    ora, names = oracle_outs(w.source, w.cases[0])
    got, _ = run_plan(w.source, w.cases, tile, names)
    assert_same(got[0], ora, bl, set(names[1]), f"adder8 {tile}")


def test_interleaved_rows_batch_and_stateful_kernels():
    """A batch of uneven instances (each interleaves over its own chunks),
    and a program whose stateful kernels keep contiguous chunks."""
    cases = [workloads.adder32(n, seed=n, width=4).cases[0] for n in (100, 64 * 64 * 2 + 1, 64 * 70, 5)]
    src = workloads.adder32_source(4)
    names = Oracle(src).output_names()
    got, _ = run_plan(src, cases, INTERLEAVED[1], names)
    for i, case in enumerate(cases):
        ora, _ = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
        assert_same(got[i], ora, case.bitlength, set(names[1]), f"instance {i}")
    bl = 64 * 64 * 97 + 13
    case = carry_case(bl, "blocks", 5)
    ora, names = oracle_outs(CARRY_SRC, case)
    got, _ = run_plan(CARRY_SRC, [case], dict(INTERLEAVED[1], max_ops=3, max_live=4), names)
    assert_same(got[0], ora, bl, set(names[1]), "carry with interleave")


@pytest.mark.parametrize("seed", range(8))
def test_random_straight_line_plans_interleaved(seed):
    src = progfuzz.random_program(200 + seed, n_stmts=40, allow_control=False, host_ints=False)
    case = progfuzz.random_case(200 + seed, [777, 4096 * 2 + 5, 64 * 64 * 9 + 1][seed % 3])
    case.params["P"] = case.params["P"] % 64
    names = Oracle(src).output_names()
    try:
        api.compile_plan_only(api.Program(src), api.Case.from_bscase(case))
    except Exception:
        pytest.skip("program reads a popcount on the host")
    ora, _ = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
    got, _ = run_plan(src, [case], dict(INTERLEAVED[seed % len(INTERLEAVED)], max_ops=8), names)
    assert_same(got[0], ora, case.bitlength, set(names[1]), f"seed {seed}")


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("bitop3", [1, 0])
def test_three_input_boolean_functions(seed, bitop3):
    """Boolean chains folded into three-input functions (v_bitop3_b32) give
    the oracle's words, and so do the same chains without the folding."""
    src = progfuzz.bool_chain_program(seed)
    bl = 64 * 64 * 5 + 17
    rng = np.random.default_rng(100 + seed)
    case = BsCase(bitlength=bl, inputs={n: workloads.random_words(rng, bl) for n in "abcd"})
    ora, names = oracle_outs(src, case)
    got, _ = run_plan(src, [case], {"bitop3": bitop3, "chunk_rows": 4}, names)
    assert_same(got[0], ora, bl, set(names[1]), f"boolean chains seed {seed} bitop3 {bitop3}")
