"""HipBackend, the benchmark-harness driver of bsim_hip (the interface of the
reference's CudaBackend, benchmark/base.py:181-254)."""

import json

import pytest

from bitstreamzoo_amd import harness, workloads
from bitstreamzoo_amd.bsdata import words_to_int, write_bsdata


def test_input_json_layout():
    d = harness.build_input_json({"M": 5, "C": -1, "big": 1 << 70}, {"k": 3}, {"a": {0: 1, 2: 4}}, 100,
                                 iterations=2, feedback={"o": "M"})
    assert d == {"inputs": {"M": 5, "C": "-0x1", "big": hex(1 << 70)}, "params": {"k": 3}, "n_vectors": 100,
                 "input_arrays": {"a": [1, 0, 4]}, "iterations": 2, "feedback": {"o": "M"}}
    json.dumps(d)


def test_parse_outputs():
    got = harness.parse_outputs({"r": "0xff", "n": 8, "arr": {"0": "0x1", "3": "0x10"}})
    assert got == {"r": 255, "n": 8, "arr": {0: 1, 3: 16}}


def test_parse_error_is_reported(tmp_path):
    bs = tmp_path / "bad.bs"
    bs.write_text("input stream x\noutput stream y\ny = x & & x\n")
    be = harness.HipBackend()
    with pytest.raises(RuntimeError, match="Parse error"):
        be.run(harness.ProgramRef(str(bs)), {"x": 1}, bitlength=8)


def test_unknown_backend_name():
    with pytest.raises(ValueError):
        harness.make_backend("cuda")


def _matchstar_expected(case):
    bl = case.bitlength
    M, C = words_to_int(case.inputs["M"]), words_to_int(case.inputs["C"])
    r = ((((M & C) + C) & ((1 << bl) - 1)) ^ C) | M
    return r, bin(r).count("1")


@pytest.mark.gpu
@pytest.mark.parametrize("executor", ["fused", "perop"])
def test_run_stdin_and_bsdata(tmp_path, executor):
    bs = tmp_path / "matchstar.bs"
    bs.write_text(workloads.MATCHSTAR_BS)
    prog = harness.ProgramRef(str(bs))
    be = harness.HipBackend(executor=executor)
    w = workloads.matchstar(5000, seed=4)
    c = w.cases[0]
    exp, n = _matchstar_expected(c)
    out, ops, ms = be.run(prog, {"M": words_to_int(c.inputs["M"]), "C": words_to_int(c.inputs["C"])},
                          bitlength=c.bitlength)
    assert out["result"] == exp and out["n"] == n
    assert ops == 5 * 1 and ms >= 0.0
    path = tmp_path / "ms.bsdata"
    write_bsdata(str(path), [{"bitlength": c.bitlength, "inputs": dict(c.inputs)}])
    out2, _, _ = be.run(prog, {}, bsdata_path=str(path))
    assert out2 == out


@pytest.mark.gpu
def test_run_batch_one_process_per_file(tmp_path):
    bs = tmp_path / "matchstar.bs"
    bs.write_text(workloads.MATCHSTAR_BS)
    prog = harness.ProgramRef(str(bs))
    paths, exps = [], []
    for s, bl in ((1, 100), (2, 70000)):
        c = workloads.matchstar(bl, seed=s).cases[0]
        p = tmp_path / f"m{s}.bsdata"
        write_bsdata(str(p), [{"bitlength": bl, "inputs": dict(c.inputs)}])
        paths.append(str(p))
        exps.append(_matchstar_expected(c))
    res = harness.HipBackend().run_batch(prog, paths)
    assert len(res) == 2
    for r, (exp, n) in zip(res, exps):
        o = harness.parse_outputs(r["outputs"])
        assert o["result"] == exp and o["n"] == n


def test_read_output_bsdata_decodes_bin_references(tmp_path):
    """The --output-bsdata file: the JSON bsim_hip prints, a NUL, then the
    words; streams and array elements referenced by offset and length."""
    blob = (5).to_bytes(8, "little") + (1 << 63).to_bytes(8, "little") + (7).to_bytes(16, "little")
    head = {"op_count": 3, "outputs": {"s": {"_bin": [0, 16]}, "a": {"2": {"_bin": [16, 16]}}, "n": 4}}
    p = tmp_path / "o.bsdata"
    p.write_bytes(json.dumps(head).encode() + b"\0" + blob)
    (c,) = harness.read_output_bsdata(str(p))
    assert c["outputs"] == {"s": 5 | (1 << 127), "a": {2: 7}, "n": 4} and c["op_count"] == 3


@pytest.mark.gpu
@pytest.mark.parametrize("executor", ["fused", "perop"])
def test_binary_outputs_equal_hex_outputs(tmp_path, executor):
    """bsim_hip --output-bsdata returns the same outputs as the hex JSON."""
    bs = tmp_path / "matchstar.bs"
    bs.write_text(workloads.MATCHSTAR_BS)
    prog = harness.ProgramRef(str(bs))
    c = workloads.matchstar(300_001, seed=9).cases[0]
    args = ({"M": words_to_int(c.inputs["M"]), "C": words_to_int(c.inputs["C"])},)
    hexo = harness.HipBackend(executor=executor).run(prog, *args, bitlength=c.bitlength)
    bino = harness.HipBackend(executor=executor, binary_outputs=True).run(prog, *args, bitlength=c.bitlength)
    assert hexo[0] == bino[0] and hexo[1] == bino[1]
    exp, n = _matchstar_expected(c)
    assert bino[0]["result"] == exp and bino[0]["n"] == n
