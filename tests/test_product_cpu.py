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
