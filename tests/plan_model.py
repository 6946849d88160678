"""Host-side model of a fused plan: executes the partition exactly as the
kernels would see it (each kernel reads only its declared inputs from the
slot / segment-output / source buffers and writes only its declared outputs),
with big-int bounded semantics.  Compared with the oracle it checks the
partitioner and the slot planner independently of code generation."""

from __future__ import annotations

import json

from bitstreamzoo_amd import api
from bitstreamzoo_amd.bsdata import BsCase, words_to_int

INPUT, ZERO, ONES, NOT, AND, OR, XOR, ADD, SHL, SHR = range(10)
SOURCE, SLOT, SEGOUT = range(3)


def describe_detail(source: str, case: BsCase, options: dict) -> dict:
    prog = api.Program(source)
    shape = api.Case.from_bscase(BsCase(bitlength=64, params=dict(case.params),
                                        inputs={k: v[:1] for k, v in case.inputs.items()},
                                        input_arrays={k: {i: w[:1] for i, w in a.items()}
                                                      for k, a in case.input_arrays.items()}))
    return api.describe_plans(prog, shape, [dict(options, detail=1)])[0]


def run_plan_model(d: dict, case: BsCase) -> dict:
    bl = case.bitlength
    mask = (1 << bl) - 1 if bl > 0 else 0
    graph = d["graph"]
    sources = []
    for name, idx in d["sources"]:
        w = case.inputs[name] if idx < 0 else case.input_arrays[name][idx]
        sources.append(words_to_int(w) & mask)
    slots, segout = {}, {}

    def const(n):
        op = graph[n][0]
        return 0 if op == ZERO else mask if op == ONES else None

    for kd in d["kernel_detail"]:
        bufs = {}
        for kind, ident, node in kd["inputs"]:
            if kind == SOURCE:
                bufs[node] = sources[ident]
            elif kind == SLOT:
                if ident not in slots:
                    raise AssertionError(f"kernel reads slot {ident} before any kernel wrote it")
                bufs[node] = slots[ident]
            else:
                if node not in segout:
                    raise AssertionError(f"kernel reads segment output {node} before it was written")
                bufs[node] = segout[node]
        local = {}

        def val(a):
            if a in local:
                return local[a]
            c = const(a)
            if c is not None:
                return c
            if a not in bufs:
                raise AssertionError(f"operand {a} is neither computed in the kernel nor one of its inputs")
            return bufs[a]

        for n in kd["nodes"]:
            op, a, b, k, _ = graph[n]
            if op == ZERO or op == ONES:
                v = const(n)
            elif op == NOT:
                v = ~val(a) & mask
            elif op == AND:
                v = val(a) & val(b)
            elif op == OR:
                v = val(a) | val(b)
            elif op == XOR:
                v = val(a) ^ val(b)
            elif op == ADD:
                v = (val(a) + val(b)) & mask
            elif op == SHL:
                v = (val(a) << k) & mask
            elif op == SHR:
                v = val(a) >> k
            else:
                raise AssertionError(f"unexpected op {op} in a kernel")
            local[n] = v
        for node, slot in kd["outputs"]:
            v = local[node] if node in local else const(node)
            if slot >= 0:
                slots[slot] = v
            else:
                segout[node] = v
    out = {}
    for name, node in d["named_outputs"]:
        op = graph[node][0]
        if op == INPUT:
            out[name] = sources[graph[node][4]]
        elif op in (ZERO, ONES):
            out[name] = const(node)
        else:
            out[name] = segout[node]
    return out
