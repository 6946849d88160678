"""The random regex-style DAG generator and its Zipf text: seeded
determinism, the op mix, the locality window, liveness of every node, the
Zipf law of the text and the basis-stream transposition; and the generated
programs' semantics (oracle against the reference interpreter)."""

import re

import numpy as np
import pytest

from bitstreamzoo_amd import dag_gen
from bitstreamzoo_amd.bsdata import words_to_int
from cpu_models import Oracle, RefInterp
from runner import run_model


def test_same_seed_same_program():
    a, b, c = dag_gen.generate(3000, seed=5), dag_gen.generate(3000, seed=5), dag_gen.generate(3000, seed=6)
    assert a.digest() == b.digest() and a.source == b.source
    assert a.digest() != c.digest()


@pytest.fixture(scope="module")
def big():
    return dag_gen.generate(200_000, seed=0)


def test_op_mix_matches_spec(big):
    n = big.n_nodes
    for kind, p in dag_gen.DEFAULT_MIX.items():
        frac = big.kind_counts[kind] / n
        assert abs(frac - p) < 0.005, (kind, frac, p)
    # MatchStar counts 4 DSL ops, ScanThru 3: ~250k DSL ops for 200k nodes
    assert 240_000 < big.n_dsl_ops < 260_000


def _defs_and_uses(src):
    defs, uses = {}, {}
    for ln, line in enumerate(src.splitlines()):
        m = re.match(r"(?:stream )?(\w+) = (.*)$", line)
        if not m:
            continue
        defs[m.group(1)] = ln
        for tok in re.findall(r"\b[a-z]\w*\b", m.group(2)):
            uses.setdefault(tok, []).append(ln)
    return defs, uses


def test_locality_window_and_liveness():
    p = dag_gen.generate(5000, seed=3, window=64)
    defs, uses = _defs_and_uses(p.source)
    node_idx = {f"v{i}": i for i in range(8, 8 + p.n_nodes)}
    for name, i in node_idx.items():
        assert name in uses, f"{name} is never used"
    # every generator node reads nodes at most `window` generator steps back
    for line in p.source.splitlines():
        m = re.match(r"stream (v\d+) = (.*)$", line)
        if not m:
            continue
        me = node_idx[m.group(1)]
        for tok in re.findall(r"\bv\d+\b", m.group(2)):
            assert me - 64 <= node_idx[tok] < me
    assert len(p.outputs) == 64
    assert p.source.count("\noutput stream o") == 64


def test_zipf_text_and_basis_streams():
    n = 1 << 20
    text = dag_gen.zipf_bytes(n, seed=1)
    counts = np.sort(np.bincount(text, minlength=256))[::-1].astype(np.float64)
    # rank-frequency ratios follow r^-1.1
    for r in (2, 4, 8):
        assert abs(counts[0] / counts[r - 1] - r ** 1.1) / r ** 1.1 < 0.05
    planes = dag_gen.basis_streams(text)
    back = np.zeros(n, dtype=np.uint8)
    for j, w in enumerate(planes):
        bits = np.unpackbits(w.view(np.uint8), bitorder="little")[:n]
        back |= (bits << j).astype(np.uint8)
    assert np.array_equal(back, text)
    assert np.array_equal(dag_gen.zipf_bytes(4096, seed=1), dag_gen.zipf_bytes(4096, seed=1))


@pytest.mark.skipif(not RefInterp.available(), reason="oracle/_ref not built or no AVX-512")
@pytest.mark.parametrize("seed", range(3))
def test_generated_programs_oracle_equals_reference(seed):
    p = dag_gen.generate(1500, seed=seed)
    case = dag_gen.text_case(3000 + 61 * seed, seed=seed)
    names = Oracle(p.source).output_names()
    o, oops = run_model(Oracle, p.source, case, Oracle.BEST, stream_names=names[0], int_names=names[1])
    r, rops = run_model(RefInterp, p.source, case, RefInterp.SIMD_OMP, 0, 1, stream_names=names[0],
                        int_names=names[1])
    mask = (1 << case.bitlength) - 1
    assert oops == rops == p.n_dsl_ops
    for k in names[0]:
        assert o[k] & mask == r[k] & mask, k
