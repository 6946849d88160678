"""The fused executor's code generator on the CPU: plans are recorded,
partitioned and compiled by hipRTC for gfx950 (no device needed) for the
benchmark workloads, straight-line reference programs and random programs,
under several tile configurations."""

import numpy as np
import pytest

import progfuzz
import ref_fixtures
from bitstreamzoo_amd import api, workloads
from bitstreamzoo_amd._lib import BszError
from bitstreamzoo_amd.bsdata import BsCase, read_cases

OPTS = [{}, {"chunk_rows": 1, "rows_per_step": 1}, {"chunk_rows": 8, "rows_per_step": 2, "warmup_rows": 2},
        {"max_ops": 7, "max_live": 6}]


def shape_for(src, bl=64):
    rng = np.random.Generator(np.random.PCG64(0))
    return api.Case.from_bscase(progfuzz.random_case(0, bl))


@pytest.mark.parametrize("opt", OPTS, ids=[str(o) for o in OPTS])
def test_workload_plans_compile(opt):
    rng = np.random.Generator(np.random.PCG64(1))
    cases = [(workloads.MATCHSTAR_BS, workloads.matchstar(256).cases[0]),
             (workloads.adder32_source(8), workloads.adder32(256, width=8).cases[0]),
             (workloads.shift_and_source("ACGTTGCAACGT"), workloads.dna_case(rng, 256, "s"))]
    for src, case in cases:
        info = api.compile_plan_only(api.Program(src), api.Case.from_bscase(case), opt)
        assert info["kernels"], info
        if opt.get("max_ops") == 7:
            assert len(info["kernels"]) >= 2 or info["graph_nodes"] < 10


def test_matchstar_plan_shape():
    info = api.compile_plan_only(api.Program(workloads.MATCHSTAR_BS),
                                 api.Case.from_bscase(workloads.matchstar(256).cases[0]))
    k = info["kernels"]
    assert len(k) == 1 and k[0]["nodes"] == 4 and k[0]["state_words"] == 1 and k[0]["pops"] == 1
    assert info["slots"] == 0


def test_host_decisions_are_rejected_by_plan_recording():
    src = "input stream a\noutput stream o\nif (a) { o = ~a }\n"
    c = BsCase(bitlength=64, inputs={"a": np.zeros(1, np.uint64)})
    with pytest.raises(BszError) as e:
        api.compile_plan_only(api.Program(src), api.Case.from_bscase(c))
    assert e.value.code == -3


def _straight_line_fuzz():
    out = []
    for seed in range(60):
        src = progfuzz.random_program(seed, n_stmts=40, allow_control=False)
        out.append((seed, src))
    return out


@pytest.mark.parametrize("seed,src", _straight_line_fuzz()[:24], ids=lambda x: str(x)[:6])
def test_random_straight_line_plans_compile(seed, src):
    case = api.Case.from_bscase(progfuzz.random_case(seed, 200))
    try:
        info = api.compile_plan_only(api.Program(src), case, OPTS[seed % len(OPTS)])
    except BszError as e:
        assert e.code == -3, str(e)[:3000]  # reads a popcount on the host: not recordable
        return
    assert info["graph_nodes"] > 0


@pytest.mark.skipif(not ref_fixtures.available(), reason="reference tree not present")
def test_reference_straight_line_programs_compile():
    n_ok = 0
    for cid, bs, path in ref_fixtures.unit_cases()[::4]:
        src = open(bs).read()
        if "if (" in src or "while" in src or len(src) > 60000:
            continue
        case = read_cases(path)[0]
        try:
            api.compile_plan_only(api.Program(src), api.Case.from_bscase(case))
            n_ok += 1
        except BszError as e:
            assert e.code == -3, (cid, str(e)[:2000])
    assert n_ok >= 5


def test_describe_matches_compile_only_shapes():
    from bitstreamzoo_amd import dag_gen
    p = dag_gen.generate(700, seed=4)
    case = dag_gen.text_case(5000, seed=4)
    prog, shape = api.Program(p.source), api.Case.from_bscase(case)
    opts = [{"max_live": 32, "chunk_rows": 4}, {"max_ops": 100, "occupancy": 1, "rows_per_step": 4}, {}]
    desc = api.describe_plans(prog, shape, opts)
    assert len(desc) == len(opts)
    for d, o in zip(desc, opts):
        info = api.compile_plan_only(prog, shape, dict(o, compile=0))
        keys = ("nodes", "inputs", "outputs", "pops", "state_words", "rows_per_step", "occupancy")
        assert [[k[x] for x in keys] for k in d["kernels"]] == [[k[x] for x in keys] for k in info["kernels"]]


@pytest.mark.parametrize("occ", [1, 2])
def test_adder_kernel_fits_its_register_file(occ):
    """The 32-bit adder fuses into one kernel with 65 input and 33 output
    streams.  Its compiled code object (metadata read on the host) stays in
    the architectural VGPRs of the requested occupancy: no AGPRs, no scratch.
    Without the store ordering the compiler sinks all 33 stores to the end of
    the row and spills."""
    w = workloads.adder32(1 << 16, seed=1)
    info = api.compile_plan_only(api.Program(w.source), api.Case.from_bscase(w.cases[0]),
                                 {"chunk_rows": 32, "rows_per_step": 1, "occupancy": occ, "waves_per_block": 4})
    (k,) = info["kernels"]
    assert k["scratch_bytes"] == 0 and k["agprs"] == 0, k
    assert 0 < k["vgprs"] <= 512 // occ and k["vgprs"] <= 256, k


def test_two_words_per_lane_shapes():
    """Kernels without carries or advances take 16-byte lanes; stateful ones
    in the same plan stay at 8 bytes."""
    src = """input stream a
input stream b
output stream x
output stream y
x = a ^ b
stream s = a + b
y = s ^ x
"""
    case = BsCase(bitlength=64 * 64 * 4, inputs={"a": np.zeros(256, np.uint64), "b": np.zeros(256, np.uint64)})
    opts = {"chunk_rows": 8, "rows_per_step": 2, "words_per_lane": 2, "max_ops": 1}
    info = api.compile_plan_only(api.Program(src), api.Case.from_bscase(case), opts)
    wpl = {k["state_words"] > 0: k["words_per_lane"] for k in info["kernels"]}
    assert wpl == {False: 2, True: 1}, info["kernels"]


def test_plan_keys_merge_equivalent_schedules():
    """The autotuner times a partition once: schedules whose kernels and
    launches coincide share a key, others do not."""
    from bitstreamzoo_amd import tuning
    w = workloads.matchstar(64 * 64 * 4, seed=1)
    space = [{"chunk_rows": 4, "rows_per_step": 4}, {"chunk_rows": 4, "rows_per_step": 8},
             {"chunk_rows": 8, "rows_per_step": 8}, {"chunk_rows": 4, "rows_per_step": 4, "waves_per_block": 1}]
    k = tuning.plan_keys(api.Program(w.source), tuning.shape_case(w.cases), space)
    assert k[0] == k[1] and len({k[0], k[2], k[3]}) == 3


def test_string_results_come_from_one_complete_call():
    """A compile-only plan's info can grow between calls (kernels compiled by
    other processes meanwhile report their registers); the text returned is
    always one call's whole output."""
    import ctypes as C
    from bitstreamzoo_amd.api import _fill_str
    outs = iter(['{"a":1}', '{"a":"' + "x" * 70000 + '"}', '{"a":"' + "x" * 70000 + '"}'])

    def call(buf, cap):
        t = next(outs).encode()
        C.memmove(buf, t, min(len(t), cap - 1))
        return len(t)

    assert _fill_str(call, cap=4) == '{"a":"' + "x" * 70000 + '"}'


def test_overflowing_kernel_is_regenerated_smaller():
    """Two rows of the adder's 65 inputs in flight at occupancy 1 compile to a
    kernel that uses AGPRs, which spills values on every row.  The library
    never runs such a kernel: it regenerates
    it with fewer rows per step, here one, which fits the 256 VGPRs."""
    w = workloads.adder32(1 << 16, seed=11)
    prog, case = api.Program(w.source), api.Case.from_bscase(w.cases[0])
    info = api.compile_plan_only(prog, case, {"chunk_rows": 8, "rows_per_step": 2, "occupancy": 1})
    (k,) = info["kernels"]
    assert k["rows_per_step"] == 1 and k["agprs"] == 0 and k["scratch_bytes"] == 0 and 0 < k["vgprs"] <= 256, k


def test_default_schedule_never_spills():
    """The batched Shift-And at the default schedule (8 rows per step at the
    estimated occupancy) compiles to a kernel that spills to scratch; the
    regenerated kernel issues fewer rows per step and has no scratch."""
    w = workloads.shift_and_batch(n_seqs=4, seed=5, min_log2=8, max_log2=10)
    from bitstreamzoo_amd import tuning
    info = api.compile_plan_only(api.Program(w.source), tuning.shape_case(w.cases), {})
    assert all(k["agprs"] == 0 and k["scratch_bytes"] == 0 and k["vgprs"] > 0 for k in info["kernels"]), info
    assert info["kernels"][0]["rows_per_step"] < 8


def test_kernel_that_cannot_shrink_is_refused():
    """A wide random-DAG partition (max_live 160) at an explicit occupancy 1
    and one row per step still needs AGPRs: nothing is left to shrink, so
    the schedule is refused with a RUNTIME error instead of being run."""
    from bitstreamzoo_amd import dag_gen, tuning
    p = dag_gen.generate(1200, seed=2)
    case = dag_gen.text_case(5000, seed=2)
    with pytest.raises(BszError) as e:
        api.compile_plan_only(api.Program(p.source), tuning.shape_case([case]),
                              {"chunk_rows": 8, "rows_per_step": 1, "max_live": 160, "occupancy": 1})
    assert e.value.code == -3 and "schedule refused" in str(e.value) and "AGPRs" in str(e.value)


def test_register_check_fails_closed():
    """The tuner's own check treats missing or unknown register counts as
    unclean, so a kernel without metadata is never timed or picked."""
    from bitstreamzoo_amd import tuning
    good = {"vgprs": 120, "agprs": 0, "scratch_bytes": 0}
    assert tuning.register_clean({"kernels": [good]})
    for bad in ({"vgprs": 120, "agprs": 4, "scratch_bytes": 0}, {"vgprs": 120, "agprs": 0, "scratch_bytes": 16},
                {"vgprs": -1, "agprs": -1, "scratch_bytes": -1}, {"vgprs": 120}, {}):
        assert not tuning.register_clean({"kernels": [good, bad]}), bad


def test_code_object_cache_rejects_foreign_files(tmp_path):
    """Cached code objects carry a header with their key (compiler version,
    options, arch and source); a file whose header is missing or wrong is
    ignored and recompiled, never loaded."""
    import glob
    opts = {"cache_dir": str(tmp_path)}
    w = workloads.matchstar(256)
    prog, case = api.Program(w.source), api.Case.from_bscase(w.cases[0])
    api.compile_plan_only(prog, case, opts)
    files = glob.glob(str(tmp_path / "*.co"))
    assert len(files) == 1
    good = open(files[0], "rb").read()
    assert good[:8] == b"BSZCO002"
    with open(files[0], "wb") as f:
        f.write(good[16:])                      # a bare code object: no header
    info = api.compile_plan_only(prog, case, opts)
    assert info["kernels"][0]["vgprs"] > 0
    assert open(files[0], "rb").read() == good  # recompiled and rewritten


def test_device_if_plan_groups_selections_by_level():
    """A 200-pattern regex-shaped program (417 `if`s) recorded with device
    selections (the plan bsz_run builds for it) partitions by predicate
    level: far fewer kernels than `if`s, each selection reading a predicate
    finished by an earlier kernel (checked inside the planner), and no
    kernel over the state-word cap.  Such a recorded plan is never run
    through bsz_plan_create."""
    import regex_ifs
    src = regex_ifs.program(200, seed=7)
    case = regex_ifs.text_case(2000, seed=6, alphabet=6)
    info = api.compile_plan_only(api.Program(src), api.Case.from_bscase(case), {"device_if": 1, "compile": 0})
    ks = info["kernels"]
    assert len(ks) < src.count("if (") / 4, len(ks)
    assert max(k["state_words"] for k in ks) <= 32


def test_plan_create_rejects_null_arguments():
    """bsz_plan_create checks its arguments before touching a device."""
    import ctypes as C
    L = api.lib()
    h = C.c_void_p()
    w = workloads.matchstar(256)
    rc = L.bsz_plan_create(None, api.Program(w.source).h, api.Case.from_bscase(w.cases[0]).h, b'{"device_if":1}',
                           C.byref(h))
    assert rc != 0


def test_lazy_loads_and_agprs_grow_partitions_without_spilling():
    """Loading each input word at its first use lets a partition count it as
    a value only up to its last use, and AGPRs (occupancy 1) give a wave 512
    registers: a generated DAG then needs fewer, larger kernels, and the
    compiled kernels keep every value in registers (AGPRs allowed, scratch
    never)."""
    from bitstreamzoo_amd import dag_gen, tuning
    p = dag_gen.generate(3000, seed=21)
    case = dag_gen.text_case(1 << 14, seed=22)
    prog, shape = api.Program(p.source), tuning.shape_case([case])
    base = {"warmup_rows": 4}
    big = dict(base, lazy_loads=1, occupancy=1, agprs=1, max_live=160, rows_per_step=1)
    d_base, d_lazy, d_big = api.describe_plans(prog, shape, [base, dict(base, lazy_loads=1), big])
    traffic = [sum(k["inputs"] + k["outputs"] for k in d["kernels"]) for d in (d_base, d_lazy, d_big)]
    assert len(d_big["kernels"]) < len(d_lazy["kernels"]) < len(d_base["kernels"])
    assert traffic[2] < traffic[1] < traffic[0], traffic
    assert max(k["state_words"] for k in d_big["kernels"]) > 32   # more state words with the AGPRs
    info = api.compile_plan_only(prog, shape, big)
    assert info["agprs"] == 1
    assert all(k["scratch_bytes"] == 0 and k["vgprs"] > 0 for k in info["kernels"]), info["kernels"]
    assert tuning.register_clean(info)
    # a plan without the AGPR option never runs a kernel that uses them
    assert not tuning.register_clean(dict(info, agprs=0)) or all(k["agprs"] == 0 for k in info["kernels"])


def test_join_warmup_is_its_own_kernel_and_compiles_clean():
    """`join_warmup` changes the generated MatchStar kernel only where a
    step spans the warm-up rows and the chunk, and the joined kernel keeps
    every value in registers."""
    from bitstreamzoo_amd import workloads
    w = workloads.matchstar(64 * 64 * 8, seed=1)
    prog, shape = api.Program(w.source), api.Case.from_bscase(w.cases[0])

    def entry(o):
        info = api.compile_plan_only(prog, shape, o)
        (k,) = info["kernels"]
        assert k["scratch_bytes"] == 0 and k["agprs"] == 0, k
        return k["entry"]

    spans = {"chunk_rows": 4, "rows_per_step": 8}
    assert entry(spans) != entry(dict(spans, join_warmup=1))
    short = {"chunk_rows": 4, "rows_per_step": 4}   # warm-up + chunk exceed the step
    assert entry(short) == entry(dict(short, join_warmup=1))


def test_boolean_chains_fold_into_three_input_functions():
    """With `bitop3` (the default) single-use boolean links fold into
    three-input functions: the generated kernel differs from the unfolded
    one, compiles within its registers, and needs no more of them."""
    src = progfuzz.bool_chain_program(0)
    case = BsCase(bitlength=64, inputs={n: np.zeros(1, np.uint64) for n in "abcd"})
    prog, shape = api.Program(src), api.Case.from_bscase(case)
    on = api.compile_plan_only(prog, shape, {"chunk_rows": 4})["kernels"]
    off = api.compile_plan_only(prog, shape, {"chunk_rows": 4, "bitop3": 0})["kernels"]
    assert len(on) == len(off) == 1
    assert on[0]["entry"] != off[0]["entry"]
    assert on[0]["scratch_bytes"] == 0 and on[0]["agprs"] == 0
    assert on[0]["vgprs"] <= off[0]["vgprs"]


BOOL3 = [  # (DSL body over inputs a, b, c into r, the same function in Python)
    ("stream t = a & b\nr = t | c\n", lambda a, b, c: (a & b) | c),
    ("stream t = a ^ b\nr = t ^ c\n", lambda a, b, c: a ^ b ^ c),
    ("stream t = a | b\nstream u = ~t\nr = u & c\n", lambda a, b, c: (1 - (a | b)) & c),
    ("stream nb = ~b\nstream t = a & nb\nr = t ^ c\n", lambda a, b, c: (a & (1 - b)) ^ c),
    ("stream t = c | a\nr = b & t\n", lambda a, b, c: b & (c | a)),
]


@pytest.mark.parametrize("body,fn", BOOL3)
def test_three_input_truth_tables(body, fn, tmp_path, monkeypatch):
    """The output's `bop3<T>` carries the truth table of the folded
    expression: bit (a << 2 | b << 1 | c) of T, the operands passed in that
    order (the operand order of v_bitop3_b32), as an independent evaluation
    of the expression gives it.  Leaves are ordered by graph node, and an
    input becomes a node at its first use, so a dead first statement reads
    a, b, c in turn; the partition drops it."""
    import glob
    import re
    monkeypatch.setenv("BSZ_DUMP_DIR", str(tmp_path))
    src = ("input stream a\ninput stream b\ninput stream c\noutput stream r\n"
           "stream z0 = a & b\nstream z1 = z0 & c\n" + body)
    case = BsCase(bitlength=64, inputs={n: np.zeros(1, np.uint64) for n in "abc"})
    api.compile_plan_only(api.Program(src), api.Case.from_bscase(case), {"chunk_rows": 1, "rows_per_step": 1})
    (unit,) = [open(f).read() for f in glob.glob(str(tmp_path / "*.hip"))]
    (v,) = set(re.findall(r"lane_st\(urow_w\(out0 \+ ru0 \* 64ull\), off8, (v\d+) & m\)", unit))
    (call,) = set(re.findall(r"const u64 " + v + r" = bop3<(\d+)u>\((\w+?)_0, (\w+?)_0, (\w+?)_0\);", unit))
    # the operands are the three inputs' words, in the order a, b, c
    loads = {x: int(j) for x, j in re.findall(r"const u64 (x\d+)_0 = .*lane_ld\(urow\(in(\d+) ", unit)}
    assert [loads[x] for x in call[1:]] == [0, 1, 2], (call, loads)
    want = sum(1 << i for i in range(8) if fn((i >> 2) & 1, (i >> 1) & 1, i & 1))
    assert int(call[0]) == want, (call, want)


def test_refused_schedule_is_an_answer_of_the_compile_pool():
    """Compiling a schedule the library refuses in the worker pool returns
    the refusal instead of breaking the pool, and the pool keeps working."""
    from bitstreamzoo_amd import dag_gen, tuning
    p = dag_gen.generate(1200, seed=2)
    case = dag_gen.text_case(5000, seed=2)
    refused = {"chunk_rows": 8, "rows_per_step": 1, "max_live": 160, "occupancy": 1}
    tuning.precompile(p.source, [case], [refused, {}], processes=2)
    msg = tuning._compile_worker((p.source, tuning._shape_desc([case]), refused))
    assert isinstance(msg, str) and "schedule refused" in msg
    import pickle
    e = pickle.loads(pickle.dumps(BszError(-3, "schedule refused: x")))
    assert e.code == -3 and "schedule refused" in str(e)
