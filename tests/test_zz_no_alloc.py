"""The allocation-free addition (bsz_op_add_noalloc) against bsz_op_add: the
same kernels with their scratch in the library's fixed device array instead
of memory bsz_op_add allocates.  Two calls in a row, the second shorter, so
whatever the first leaves in the shared array must not leak into the
second; carry chains that cross segments and rows exercise the fix-up
launch; sizes beyond the array are refused before anything runs."""

import numpy as np
import pytest

from bitstreamzoo_amd import _lib, api

pytestmark = pytest.mark.gpu


def _ctx():
    from hip_models import ctx
    return ctx()


def _operands(rng, bitlength, kind):
    n = (bitlength + 63) // 64
    if kind == "random":
        return rng.integers(0, 1 << 64, size=n, dtype=np.uint64), rng.integers(0, 1 << 64, size=n, dtype=np.uint64)
    # a + ~a with a single extra carry: the carry runs through every segment,
    # so each segment's rows below it all propagate and the fix-up walks them
    a = rng.integers(0, 1 << 64, size=n, dtype=np.uint64)
    b = ~a
    b[0] = np.uint64(int(b[0]) + 1) if int(b[0]) != (1 << 64) - 1 else b[0]
    return a, b


def _add(ctx, fn, a_dev, b_dev, bitlength):
    out = ctx.alloc(bitlength)
    _lib.check(getattr(api.lib(), fn)(ctx.h, out.dev, a_dev.dev, b_dev.dev, bitlength), ctx.h)
    ctx.sync()
    got = out.download()
    out.free()
    return got


@pytest.mark.parametrize("kind", ["random", "carry_through"])
def test_noalloc_add_matches_bsz_op_add_twice(kind):
    ctx = _ctx()
    rng = np.random.Generator(np.random.PCG64(17))
    for bitlength in ((1 << 22) + 77, (1 << 18) + 5):   # second call smaller
        a, b = _operands(rng, bitlength, kind)
        da, db = ctx.upload(a, bitlength), ctx.upload(b, bitlength)
        ref = _add(ctx, "bsz_op_add", da, db, bitlength)
        got = _add(ctx, "bsz_op_add_noalloc", da, db, bitlength)
        assert np.array_equal(got, ref), (kind, bitlength)
        exp = (int.from_bytes(a.tobytes(), "little") + int.from_bytes(b.tobytes(), "little")) & ((1 << bitlength) - 1)
        assert int.from_bytes(got[: len(a)].tobytes(), "little") == exp
        da.free()
        db.free()


def test_noalloc_add_refuses_oversized_streams():
    ctx = _ctx()
    a = ctx.alloc(64)
    rc = api.lib().bsz_op_add_noalloc(ctx.h, a.dev, a.dev, a.dev, 64)
    assert rc == -1   # aliasing: BSZ_E_INVALID
    b, d = ctx.alloc(64), ctx.alloc(64)
    rc = api.lib().bsz_op_add_noalloc(ctx.h, d.dev, a.dev, b.dev, (1 << 34) + 1)
    assert rc == -1 and "2^34" in api.lib().bsz_last_error(ctx.h).decode()
    for x in (a, b, d):
        x.free()
