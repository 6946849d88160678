"""Runs of the HIP executors through the C ABI, shaped like the oracle's
results (Python ints) so parity tests compare dictionaries."""

from __future__ import annotations

import pytest

from bitstreamzoo_amd import api
from bitstreamzoo_amd.bsdata import BsCase, words_to_int

EXECUTORS = {"perop": api.PEROP, "fused": api.FUSED}

_ctx = None


def ctx() -> api.Context:
    global _ctx
    if _ctx is None:
        _ctx = api.Context(0)
    return _ctx


def hip_run(source: str, case: BsCase, executor: str, stream_names, int_names):
    prog = api.Program(source)
    res = ctx().run(prog, api.Case.from_bscase(case), EXECUTORS[executor])
    out = {}
    for n in stream_names:
        s = res.stream(n)
        if s is not None:
            out[n] = words_to_int(s)
        a = res.array(n)
        if a is not None:
            out[n] = {i: words_to_int(w) for i, w in a.items()}
    for n in int_names:
        v = res.int_out(n)
        if v is not None:
            out[n] = v
    return out, res.op_count


def assert_same(hip: dict, ora: dict, bitlength: int, int_names, label=""):
    mask = (1 << bitlength) - 1 if bitlength > 0 else -1
    assert set(hip) == set(ora), f"{label}: output sets differ: {sorted(set(hip) ^ set(ora))}"
    for k, v in ora.items():
        h = hip[k]
        if isinstance(v, dict):
            assert set(h) == set(v), f"{label}: {k} indices"
            for i in v:
                assert h[i] & mask == v[i] & mask, f"{label}: {k}[{i}]"
        elif k in int_names:
            assert h == v, f"{label}: int {k}: {h} != {v}"
        else:
            assert h & mask == v & mask, f"{label}: stream {k}"


def plan_or_refusal(c: api.Context, program: api.Program, shape: api.Case, options: dict):
    """A plan, or None when the library refuses the schedule because its
    compiled kernels would use AGPRs or scratch.  Only schedules that ask
    for occupancy 1 (the whole register file) may be refused: automatic
    occupancy keeps every kernel within 256 VGPRs."""
    from bitstreamzoo_amd import tuning
    try:
        return api.Plan(c, program, shape, options)
    except api.BszError as e:
        if tuning.refused(e) and options.get("occupancy") == 1:
            return None
        raise
