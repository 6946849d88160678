"""The program walker's device-side `if` (if-conversion), checked on the CPU.

tests/walker/mock_engine.cpp implements the executor's Engine interface
with host vectors and device-style predicates; it is compiled here with g++
and the address and undefined-behaviour sanitizers together with the
library's own walker, front-end and .bsdata reader.  Every program runs
twice, with predicates (bodies walked unconditionally, assignments turned
into selections, op counts and variable creation settled at the end) and
with host decisions (the reference's behaviour), and must give the same
outputs, ints, op count and variable count; outputs and op counts must also
equal the oracle's."""

import json
import os
import subprocess

import pytest

import progfuzz
import regex_ifs
from bitstreamzoo_amd.bsdata import write_bsdata
from cpu_models import Oracle
from runner import run_model

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "bitstreamzoo_amd", "csrc")


@pytest.fixture(scope="module")
def mock(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("walker") / "mock_engine")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}",
           os.path.join(HERE, "walker", "mock_engine.cpp"), os.path.join(HERE, "walker", "stubs.cpp"),
           os.path.join(CSRC, "executor.cpp"), os.path.join(CSRC, "frontend.cpp"), os.path.join(CSRC, "bsdata_io.cpp"),
           "-o", out, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("cannot build the walker harness here: " + r.stderr[-500:])
    return out


def run_mock(mock, tmp_path, src, case, preds):
    bs, data = tmp_path / "p.bs", tmp_path / "c.bsdata"
    bs.write_text(src)
    write_bsdata(str(data), [{"bitlength": case.bitlength, "params": dict(case.params), "inputs": dict(case.inputs),
                              "input_arrays": {k: dict(v) for k, v in case.input_arrays.items()}}])
    r = subprocess.run([mock, str(bs), str(data), "1" if preds else "0"], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    return r.returncode, r.stdout, r.stderr


def compare(mock, tmp_path, src, case, label):
    rc1, out1, err1 = run_mock(mock, tmp_path, src, case, True)
    rc0, out0, err0 = run_mock(mock, tmp_path, src, case, False)
    assert "Sanitizer" not in err1 and "runtime error" not in err1, (label, err1[-2000:])
    assert rc1 == rc0, (label, rc1, rc0, err1[-500:], err0[-500:])
    if rc0 != 0:
        return None  # both raised the same run-time error
    a, b = json.loads(out1), json.loads(out0)
    assert a == b, label
    return a


def test_regex_shaped_programs_predicated_equal_host_decided(mock, tmp_path):
    for seed, (n, nb, al) in enumerate([(300, 20000, 6), (200, 5000, 3), (80, 777, 2), (150, 64 * 64 * 3, 4)]):
        src = regex_ifs.program(n, seed=seed + 1)
        case = regex_ifs.text_case(nb, seed=seed, alphabet=al)
        got = compare(mock, tmp_path, src, case, f"regex {seed}")
        names = Oracle(src).output_names()
        ora, ops = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
        assert got["op_count"] == ops
        for k, v in ora.items():
            assert int(got["streams"][k], 16) == v, (seed, k)


@pytest.mark.parametrize("block", range(4))
def test_random_programs_predicated_equal_host_decided(mock, tmp_path, block):
    """Nested if / while / for, arrays, broadcast and popcount ints; bodies
    that cannot be converted (loops, int assignments) or that raise fall
    back to host decisions inside or outside converted bodies."""
    n_checked = 0
    for seed in range(block * 25, block * 25 + 25):
        bl = [1, 64, 65, 200, 777, 4096, 5000][seed % 7]
        src = progfuzz.random_program(seed, n_stmts=40)
        case = progfuzz.random_case(seed, bl)
        got = compare(mock, tmp_path, src, case, f"fuzz {seed}")
        if got is None:
            continue
        names = Oracle(src).output_names()
        ora, ops = run_model(Oracle, src, case, Oracle.SCALAR, stream_names=names[0], int_names=names[1])
        assert got["op_count"] == ops, seed
        n_checked += 1
    assert n_checked >= 10
