"""Case execution shared by the parity tests: iterations / feedback loop
(main.cu:84-140) on top of any model with the case interface."""

from __future__ import annotations

import copy

from bitstreamzoo_amd.bsdata import BsCase, int_to_words


def run_model(model_cls, source: str, case: BsCase, *run_args, stream_names=None, int_names=None):
    """Run `case` through `model_cls` honouring iterations/feedback; returns
    (outputs as python ints, total op_count)."""
    cur = copy.copy(case)
    cur.inputs = dict(case.inputs)
    outs, ops = {}, 0
    for _ in range(max(1, case.iterations)):
        m = model_cls(source).load(cur)
        m.run(*run_args)
        if stream_names is None:
            stream_names, int_names = m.output_names()
        outs = m.outputs(stream_names, int_names)
        ops += m.op_count()
        for out_name, in_name in case.feedback.items():
            s = m.stream(out_name)
            if s is not None:
                cur.inputs[in_name] = s
    return outs, ops
