"""Partition order (plan option `order`), checked without a GPU.

Order 1 replaces the recorded node order by a list schedule that keeps few
values live, so each kernel holds more nodes and fewer values cross kernel
boundaries through HBM.  These tests run the partition through the host plan
model (every kernel sees only its declared inputs and writes only its declared
outputs, with the slot planner's buffer reuse) and compare every output with
the oracle, under both orders; and they pin the planned-transfer reduction on
the generator's regex-style DAG.
"""

import pytest

from bitstreamzoo_amd import api, dag_gen
from bitstreamzoo_amd.bsdata import words_to_int
from cpu_models import Oracle
from plan_model import describe_detail, run_plan_model
from progfuzz import random_case, random_program


def _transfers(info: dict) -> int:
    return sum(k["inputs"] + k["outputs"] for k in info["kernels"])


def _check_against_oracle(source: str, case, options: dict) -> dict:
    d = describe_detail(source, case, options)
    got = run_plan_model(d, case)
    o = Oracle(source).load(case).run(Oracle.SCALAR)
    mask = (1 << case.bitlength) - 1
    compared = 0
    for name in got:
        want = o.stream(name)
        if want is None:  # an output array: the plan model names its elements by the array
            continue
        assert got[name] & mask == words_to_int(want) & mask, (options, name)
        compared += 1
    assert compared >= 1
    return d


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("order", [0, 1])
def test_generated_dag_partition_matches_oracle(seed, order):
    p = dag_gen.generate(3000, seed=seed)
    case = dag_gen.text_case(1000 + 37 * seed, seed=seed)
    d = _check_against_oracle(p.source, case, {"warmup_rows": 4, "order": order})
    assert len(d["kernel_detail"]) > 10


@pytest.mark.parametrize("seed", range(12))
def test_random_straight_line_programs_under_both_orders(seed):
    src = random_program(seed, n_stmts=60, allow_control=False, host_ints=False)
    case = random_case(seed, 64 * 5 + 13 * seed)
    for order in (0, 1):
        # a small value budget forces many cuts, so slots are reused often
        _check_against_oracle(src, case, {"order": order, "max_live": 8})


def test_order_is_a_permutation_of_the_recorded_nodes():
    p = dag_gen.generate(2000, seed=9)
    case = dag_gen.text_case(512, seed=9)
    d0 = describe_detail(p.source, case, {"order": 0})
    d1 = describe_detail(p.source, case, {"order": 1})
    n0 = sorted(n for k in d0["kernel_detail"] for n in k["nodes"])
    n1 = sorted(n for k in d1["kernel_detail"] for n in k["nodes"])
    assert n0 == n1 and len(set(n1)) == len(n1)
    # the recorded order keeps node ids increasing across the plan; the list
    # schedule does not (it is a different topological order)
    flat0 = [n for k in d0["kernel_detail"] for n in k["nodes"]]
    flat1 = [n for k in d1["kernel_detail"] for n in k["nodes"]]
    assert flat0 == sorted(flat0) and flat1 != sorted(flat1)


def test_list_schedule_halves_planned_transfers_of_the_regex_dag():
    # 20k generator nodes, the bench's schedule: the recorded order plans
    # 1.94 stream transfers per DSL op, the list schedule 0.99
    p = dag_gen.generate(20_000, seed=4)
    case = dag_gen.text_case(4096, seed=11)
    prog, shape = api.Program(p.source), api.Case.from_bscase(case)
    i0, i1 = api.describe_plans(prog, shape, [{"warmup_rows": 4}, {"warmup_rows": 4, "order": 1}])
    t0, t1 = _transfers(i0), _transfers(i1)
    assert t0 / p.n_dsl_ops > 1.8
    assert t1 < 0.6 * t0, (t0, t1)
    assert len(i1["kernels"]) < 0.6 * len(i0["kernels"])
    # the same register budget: no kernel of either order holds more state
    # words than the per-kernel cap
    assert max(k["state_words"] for k in i1["kernels"]) <= 32


def test_order_is_ignored_with_device_selections():
    # a program recorded with device-side `if` selections keeps the
    # predicate-level order, whatever the option says
    import regex_ifs
    src = regex_ifs.program(40, seed=3)
    case = regex_ifs.text_case(1000, seed=2, alphabet=6)
    prog, shape = api.Program(src), api.Case.from_bscase(case)
    i0 = api.compile_plan_only(prog, shape, {"device_if": 1, "compile": 0})
    i1 = api.compile_plan_only(prog, shape, {"device_if": 1, "compile": 0, "order": 1})
    assert len(i0["kernels"]) > 1
    assert [(k["nodes"], k["inputs"], k["outputs"]) for k in i0["kernels"]] == \
        [(k["nodes"], k["inputs"], k["outputs"]) for k in i1["kernels"]]
