"""CPU-side checks of the shipped library: it loads, exports the C ABI,
its front-end accepts and rejects exactly what the oracle does, static op
counts agree, and the CLI's input decoding matches the Python reader.
No kernel runs here (no GPU in this container)."""

import json
import os
import subprocess

import numpy as np
import pytest

import progfuzz
import ref_fixtures
from bitstreamzoo_amd import _lib, api, workloads
from bitstreamzoo_amd.bsdata import read_cases, write_bsdata
from cpu_models import Oracle


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [s for s in _lib.header_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert len(_lib.header_symbols()) >= 40


def _programs():
    progs = [("matchstar", workloads.MATCHSTAR_BS), ("adder32", workloads.adder32_source()),
             ("shift_and", workloads.shift_and_source("ACGTACGTTTGACCAGTACGATCGATCGATGC"))]
    progs += [(f"fuzz{i}", progfuzz.random_program(i)) for i in range(20)]
    if ref_fixtures.available():
        progs += [(f"ref:{d}/{n}", open(p).read()) for d, n, p in ref_fixtures.programs()]
    return progs


@pytest.mark.parametrize("name,src", _programs(), ids=[p[0] for p in _programs()])
def test_frontend_static_ops_match_oracle(name, src):
    prog = api.Program(src)
    case = api.Case(64)
    o = Oracle(src)
    o.load(progfuzz.random_case(0, 64) if name.startswith("fuzz") else
           __import__("bitstreamzoo_amd.bsdata", fromlist=["BsCase"]).BsCase(bitlength=64))
    try:
        o.run(Oracle.SCALAR)
    except RuntimeError:
        pass  # runtime errors do not matter here; static ops are computed first
    if name.startswith("fuzz"):
        case = api.Case.from_bscase(progfuzz.random_case(0, 64))
    assert prog.static_ops(case) == o.static_ops()


BAD_PROGRAMS = [
    "stream x = a & b & c\n",                  # two operators
    "stream x = ~(a & b)\n",                   # operand not primary
    "input stream a\nif (a & a) { }\n",        # non-variable condition
    "int n = popcount(a & b)\n",               # non-variable popcount arg
    "stream x = a $ b\n",                      # bad character
    "stream x = \n",                           # missing expression
    "for i in 0..3 { input stream q }\n",      # declaration inside a block
    "stream x = a << 99999999999\n",           # literal beyond int32 (std::stoi)
    "output\n",                                 # truncated declaration
]


@pytest.mark.parametrize("src", BAD_PROGRAMS)
def test_parse_errors_agree_with_oracle(src):
    with pytest.raises(SyntaxError):
        Oracle(src)
    with pytest.raises(_lib.BszError) as e:
        api.Program(src)
    assert e.value.code == -2


def test_cli_parse_error_contract(tmp_path):
    bs = tmp_path / "bad.bs"
    bs.write_text("stream x = a & b & c\n")
    r = subprocess.run([_lib.CLI_PATH, str(bs)], input="{}", capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.startswith("Parse error:")


def test_cli_dump_input_matches_python_reader(tmp_path):
    rng = np.random.Generator(np.random.PCG64(4))
    bl = 777
    ins = {"a": workloads.random_words(rng, bl), "b": workloads.random_words(rng, bl)}
    arr = {"v": {0: workloads.random_words(rng, bl), 2: workloads.random_words(rng, bl)}}
    p = tmp_path / "t.bsdata"
    write_bsdata(str(p), [{"bitlength": bl, "params": {"K": 5}, "inputs": ins, "input_arrays": arr}])
    bs = tmp_path / "p.bs"
    bs.write_text("input stream a\noutput stream a\n")
    r = subprocess.run([_lib.CLI_PATH, str(bs), "--input", str(p), "--dump-input"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    c = read_cases(str(p))[0]
    assert d["bitlength"] == bl and d["params"] == {"K": 5}
    for k, w in c.inputs.items():
        assert int(d["inputs"][k], 16) == int.from_bytes(w.tobytes(), "little")
    for k, a in c.input_arrays.items():
        for i, w in a.items():
            assert int(d["input_arrays"][k][str(i)], 16) == int.from_bytes(w.tobytes(), "little")


@pytest.mark.skipif(not ref_fixtures.available(), reason="reference tree not present")
def test_cli_reads_every_reference_fixture():
    """The C++ .bsdata reader decodes every committed fixture like the Python reader."""
    bs = "/dev/null"
    n = 0
    for cid, _, path in ref_fixtures.unit_cases():
        if os.path.getsize(path) > 400_000:
            continue
        r = subprocess.run([_lib.CLI_PATH, bs, "--input", path, "--dump-input"], capture_output=True, text=True)
        assert r.returncode == 0, (cid, r.stderr)
        d = json.loads(r.stdout)
        ds = d if isinstance(d, list) else [d]
        cs = read_cases(path)
        assert len(ds) == len(cs), cid
        for dj, c in zip(ds, cs):
            assert dj["bitlength"] == c.bitlength
            for k, w in c.inputs.items():
                assert int(dj["inputs"][k], 16) == int.from_bytes(w.tobytes(), "little"), (cid, k)
            for k, a in c.input_arrays.items():
                for i, w in a.items():
                    assert int(dj["input_arrays"][k][str(i)], 16) == int.from_bytes(w.tobytes(), "little"), (cid, k, i)
        n += 1
    assert n > 50


@pytest.mark.parametrize("section,label", [
    ({"x": {"_bin": [1 << 62, 16]}}, "_bin offset near 2^64"),
    ({"x": {"_bin": [0, 1 << 40]}}, "_bin longer than the file"),
    ({"x": {"_bin": [0]}}, "_bin descriptor too short"),
    ({"x": {"_bin": [-8, 8]}}, "negative offset"),
    ({"x": {"_bin": ["0", 8]}}, "offset not an integer"),
    ({"a": {"_array_bin": [2, 0]}}, "_array_bin descriptor too short"),
    ({"a": {"_array_bin": [3, 0, 8]}}, "_array_bin past the end"),
    ({"a": {"_array_bin": [2, 8, (1 << 63) - 1]}}, "_array_bin stride overflow"),
    ({"a": {"_broadcast_pack": [1 << 20, 0, 1]}}, "_broadcast_pack past the end"),
])
def test_cli_rejects_malformed_bsdata(tmp_path, section, label):
    """Section descriptors that are short, negative, non-integer or point
    outside the binary part are parse errors (exit 1, 'Error reading bsdata'),
    never out-of-bounds reads; offsets near 2^64 cannot wrap the check."""
    header = {"bitlength": 64, "inputs": {k: v for k, v in section.items() if "_array_bin" not in v
                                          and "_broadcast_pack" not in v},
              "input_arrays": {k: v for k, v in section.items() if "_array_bin" in v or "_broadcast_pack" in v}}
    path = tmp_path / "bad.bsdata"
    path.write_bytes(json.dumps(header).encode() + b"\0" + bytes(16))
    r = subprocess.run([_lib.CLI_PATH, "/dev/null", "--input", str(path), "--dump-input"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1, (label, r.stdout[:200])
    assert "Error reading bsdata" in r.stderr and "Invalid bsdata section" in r.stderr, (label, r.stderr)


def test_cli_reads_well_formed_sections_at_the_blob_edge(tmp_path):
    """A _bin / _array_bin / _broadcast_pack whose range ends exactly at the
    end of the binary part is read in full."""
    blob = bytes(range(16)) + bytes([0b101])
    header = {"bitlength": 64, "inputs": {"x": {"_bin": [8, 8]}},
              "input_arrays": {"a": {"_array_bin": [2, 0, 8]}, "b": {"_broadcast_pack": [3, 16, 1]}}}
    path = tmp_path / "edge.bsdata"
    path.write_bytes(json.dumps(header).encode() + b"\0" + blob)
    r = subprocess.run([_lib.CLI_PATH, "/dev/null", "--input", str(path), "--dump-input"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert int(d["inputs"]["x"], 16) == int.from_bytes(blob[8:16], "little")
    assert int(d["input_arrays"]["a"]["1"], 16) == int.from_bytes(blob[8:16], "little")


def _static_ops_three_ways(src, bscase):
    from cpu_models import RefInterp
    ref = RefInterp(src).load(bscase).static_ops()
    o = Oracle(src).load(bscase)
    try:
        o.run(Oracle.SCALAR)
    except RuntimeError:
        pass  # a runtime error after the static count does not matter here
    return ref, o.static_ops(), api.Program(src).static_ops(api.Case.from_bscase(bscase))


@pytest.mark.skipif(not ref_fixtures.available(), reason="reference tree not present")
def test_static_ops_pinned_to_reference_on_every_fixture():
    """count_ops_ast of the reference itself (oracle/_ref, ast_liveness.cpp:
    629-727) equals the oracle's and the product's static op count on every
    committed fixture program under its own params (for bodies counted times
    the trip count, if/while bodies once)."""
    from cpu_models import RefInterp
    if not RefInterp.available():
        pytest.skip("oracle/_ref not built")
    n = 0
    for cid, bs, path in ref_fixtures.unit_cases():
        ref, ora, prod = _static_ops_three_ways(open(bs).read(), read_cases(path)[0])
        assert ref == ora == prod, (cid, ref, ora, prod)
        n += 1
    assert n >= 100


@pytest.mark.skipif(not ref_fixtures.available(), reason="reference tree not present")
def test_static_ops_pinned_to_reference_on_control_flow_programs():
    """The same on random programs with nested for / if / while, arrays and
    popcount assignments, where a body counted once versus per iteration
    would differ."""
    from cpu_models import RefInterp
    if not RefInterp.available():
        pytest.skip("oracle/_ref not built")
    seen_control = 0
    for seed in range(80):
        src = progfuzz.random_program(seed, n_stmts=30, allow_control=True)
        ref, ora, prod = _static_ops_three_ways(src, progfuzz.random_case(seed, 128))
        assert ref == ora == prod, (seed, ref, ora, prod, src)
        seen_control += ("if (" in src) or ("while" in src)
    assert seen_control >= 20


def test_frontend_and_planner_at_a_million_ops():
    """A generated 1M-node regex-style DAG (1.25M DSL ops, 40 MB of source,
    the size class of the reference's regex_large) goes through the
    front-end, the op counter, the recording walk and the partitioner and
    slot planner in bounded time; kernel sources are not generated for a
    description."""
    import time
    from bitstreamzoo_amd import dag_gen, tuning
    p = dag_gen.generate(1_000_000, seed=4)
    assert p.n_dsl_ops > 1_200_000
    case = dag_gen.text_case(1 << 12, seed=11)
    t0 = time.perf_counter()
    prog = api.Program(p.source)
    shape = tuning.shape_case([case])
    assert prog.static_ops(shape) == p.n_dsl_ops
    t_parse = time.perf_counter() - t0
    (d,) = api.describe_plans(prog, shape, [{"warmup_rows": 4}])
    t_plan = time.perf_counter() - t0 - t_parse
    assert len(d["kernels"]) > 10_000 and 0 < d["slots"] < 10_000
    assert sum(k["nodes"] for k in d["kernels"]) >= p.n_dsl_ops * 0.9
    assert t_parse < 20 and t_plan < 60, (t_parse, t_plan)
