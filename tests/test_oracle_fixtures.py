"""Pin the oracle against the reference's own committed fixtures.

Every unit fixture under benchmark/apps/*/datasets/tests (expected outputs
stored inside the .bsdata by the reference's generators) is run through the
oracle in all three kernel variants, and through the reference's own
Interpreter<CpuBackend> (oracle/_ref) where that was built.  Skipped where the
reference tree is absent.
"""

import os

import pytest

import ref_fixtures
from bitstreamzoo_amd.bsdata import read_cases
from cpu_models import Oracle, RefInterp
from runner import run_model

pytestmark = pytest.mark.skipif(not ref_fixtures.available(), reason="reference tree not present")

CASES = ref_fixtures.unit_cases() if ref_fixtures.available() else []


def test_fixture_inventory():
    # 133 .bsdata files are committed; the unit subset with programs present
    # is what GenericDomain runs under --unit.
    assert len(CASES) >= 100


@pytest.mark.parametrize("cid,bs,path", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_fixture(cid, bs, path):
    src = open(bs).read()
    names = Oracle(src).output_names()
    big = os.path.getsize(path) > 300_000
    for case in read_cases(path):
        for variant in ((Oracle.BEST,) if big else (Oracle.SCALAR, Oracle.BEST)):
            outs, _ = run_model(Oracle, src, case, variant, stream_names=names[0], int_names=names[1])
            bad = ref_fixtures.check_expected(outs, case.expected, case.bitlength, set(names[1]))
            assert not bad, f"{cid} {case.name} variant={variant}: mismatch in {bad[:5]}"


@pytest.mark.skipif(not RefInterp.available(), reason="oracle/_ref not built or no AVX-512")
@pytest.mark.parametrize("cid,bs,path", CASES, ids=[c[0] for c in CASES])
def test_oracle_equals_reference_interpreter(cid, bs, path):
    if os.path.getsize(path) > 300_000:
        pytest.skip("large fixture: covered against its expected outputs above")
    src = open(bs).read()
    names = Oracle(src).output_names()
    for case in read_cases(path):
        o, oops = run_model(Oracle, src, case, Oracle.FAITHFUL, stream_names=names[0], int_names=names[1])
        r, rops = run_model(RefInterp, src, case, RefInterp.SIMD_OMP, 0, 1, stream_names=names[0], int_names=names[1])
        mask = (1 << case.bitlength) - 1 if case.bitlength > 0 else -1
        assert set(o) == set(r), f"{cid}: output sets differ"
        for k in o:
            if isinstance(o[k], dict):
                assert set(o[k]) == set(r[k])
                for i in o[k]:
                    assert o[k][i] & mask == r[k][i] & mask, f"{cid}: {k}[{i}]"
            elif k in names[1]:
                assert o[k] == r[k], f"{cid}: int {k}"
            else:
                assert o[k] & mask == r[k] & mask, f"{cid}: {k}"
        assert oops == rops, f"{cid}: op_count {oops} != {rops}"
