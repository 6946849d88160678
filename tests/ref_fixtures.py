"""Discovery of the reference's committed unit fixtures.

Mirrors GenericDomain's lookup (benchmark/base.py:537-583): every
benchmark/apps/<domain>/src/*.bs program, its cases from
datasets/tests/tests.json (a list, or a dict keyed by program name with a
"_default" fallback), restricted to unit fixtures (no "provenance" block)
whose .bsdata file is committed.  Only used by CPU tests in a container where
/root/reference exists; GPU tests never read it.
"""

from __future__ import annotations

import glob
import json
import os

REF_ROOT = os.environ.get("BSZ_REFERENCE_ROOT", "/root/reference")
APPS = os.path.join(REF_ROOT, "benchmark", "apps")


def available() -> bool:
    return os.path.isdir(APPS)


def programs():
    out = []
    for bs in sorted(glob.glob(os.path.join(APPS, "*", "src", "*.bs"))):
        domain = bs.split(os.sep)[-3]
        out.append((domain, os.path.splitext(os.path.basename(bs))[0], bs))
    return out


def unit_cases():
    """[(id, bs_path, bsdata_path)] for every committed unit fixture."""
    res = []
    for domain, prog, bs in programs():
        tj = os.path.join(APPS, domain, "datasets", "tests", "tests.json")
        if not os.path.exists(tj):
            continue
        with open(tj) as f:
            data = json.load(f)
        cases = data if isinstance(data, list) else data.get(prog, data.get("_default", []))
        for c in cases:
            if "provenance" in c or not c.get("data_file", "").endswith(".bsdata"):
                continue
            p = os.path.join(os.path.dirname(tj), c["data_file"])
            if os.path.exists(p):
                res.append((f"{domain}/{prog}/{c['data_file']}", bs, p))
    return res


def check_expected(actual: dict, expected: dict, bitlength: int, int_outputs) -> list:
    """benchmark/base.py:437-472; returns the mismatching keys."""
    mask = (1 << bitlength) - 1 if bitlength > 0 else None
    bad = []
    for k, e in expected.items():
        if k not in actual:
            bad.append(k)
            continue
        a = actual[k]
        m = mask is not None and k not in int_outputs
        if isinstance(e, dict):
            if not isinstance(a, dict):
                bad.append(k)
                continue
            for idx, ev in e.items():
                av = a.get(idx, 0)
                if m:
                    av &= mask
                    ev &= mask
                if av != ev:
                    bad.append(f"{k}[{idx}]")
                    break
        else:
            av, ev = a, e
            if m:
                av &= mask
                ev &= mask
            if av != ev:
                bad.append(k)
    return bad
