"""Per-op HIP kernels (C ABI) against the oracle's word kernels and closed
forms.  Shift amounts and carry chains follow the reference's own edge cases
(test_ops.cpp:142-201, test_kernels.cu:300-382) and cross the lane (64),
wavefront row, block (256 words) and add-tile (1024 words) boundaries."""

import numpy as np
import pytest

from cpu_models import Oracle, oracle_words_op
from hip_models import ctx

pytestmark = pytest.mark.gpu

BITLENGTHS = [1, 63, 64, 65, 500, 4095, 4096, 4097, 65536 - 7, 65536, 65536 + 1, 262147, 1_000_003]
SHIFTS = [0, 1, 31, 32, 33, 63, 64, 65, 127, 128, 200, 4095, 4096, 4097, 70001]


def rand_stream(rng, bl):
    nw = (bl + 63) // 64
    w = rng.integers(0, 2**64, size=nw, dtype=np.uint64)
    if bl % 64:
        w[-1] &= np.uint64((1 << (bl % 64)) - 1)
    return w


def mask_words(w, bl):
    w = w.copy()
    if bl % 64:
        w[-1] &= np.uint64((1 << (bl % 64)) - 1)
    return w


@pytest.mark.parametrize("bl", BITLENGTHS)
def test_pointwise(bl):
    rng = np.random.Generator(np.random.PCG64(bl))
    a, b = rand_stream(rng, bl), rand_stream(rng, bl)
    c = ctx()
    da, db = c.upload(a, bl), c.upload(b, bl)
    for name, ref in [("and", a & b), ("or", a | b), ("xor", a ^ b), ("not", mask_words(~a, bl)), ("copy", a)]:
        d = c.op(name, da, db)
        got = d.download()
        assert np.array_equal(got, ref), f"{name} bl={bl}"
        full = d.download(d.alloc_words)
        assert not full[len(a):].any(), f"{name}: padding words not zero"
        d.free()
    for ones in (False, True):
        d = c.fill(bl, ones)
        exp = mask_words(np.full(len(a), np.uint64(2**64 - 1)), bl) if ones else np.zeros(len(a), np.uint64)
        assert np.array_equal(d.download(), exp)
        d.free()
    da.free(); db.free()


@pytest.mark.parametrize("bl", [65, 4097, 65536 + 1, 1_000_003])
@pytest.mark.parametrize("k", SHIFTS)
def test_shifts(bl, k):
    rng = np.random.Generator(np.random.PCG64(bl * 7 + k))
    a = rand_stream(rng, bl)
    c = ctx()
    da = c.upload(a, bl)
    big = int.from_bytes(a.tobytes(), "little")
    m = (1 << bl) - 1
    exp_l = ((big << k) & m).to_bytes(len(a) * 8, "little")
    exp_r = (big >> k).to_bytes(len(a) * 8, "little")
    dl = c.op("shl", da, k=k)
    dr = c.op("shr", da, k=k)
    assert dl.download().tobytes() == exp_l, f"shl {k}"
    assert dr.download().tobytes() == exp_r, f"shr {k}"
    # the oracle's serial kernels agree (masking applied for the bounded stream)
    assert np.array_equal(mask_words(oracle_words_op("shl", a, k=k), bl), dl.download())
    for d in (da, dl, dr):
        d.free()


def carry_cases(rng, bl):
    nw = (bl + 63) // 64
    out = [("random", rand_stream(rng, bl), rand_stream(rng, bl))]
    ones = mask_words(np.full(nw, np.uint64(2**64 - 1)), bl)
    one = np.zeros(nw, np.uint64); one[0] = 1
    out.append(("ones+1", ones, one))
    out.append(("no-carry", rand_stream(rng, bl) & np.uint64(0x5555555555555555),
                rand_stream(rng, bl) & np.uint64(0xAAAAAAAAAAAAAAAA)))
    # carry generated at word s-1 that ripples through all-ones words up to e
    for s, e in [(1, 64), (63, 65), (255, 257), (511, 513), (1023, 1025), (1, 2049), (3, nw)]:
        if e > nw or s >= e:
            continue
        a = rand_stream(rng, bl)
        b = rand_stream(rng, bl)
        a[s:e] = np.uint64(2**64 - 1)
        b[s:e] = 0
        a[s - 1] = np.uint64(2**64 - 1)
        b[s - 1] |= np.uint64(1)
        out.append((f"chain{s}-{e}", mask_words(a, bl), mask_words(b, bl)))
    return out


@pytest.mark.parametrize("bl", [64, 4097, 65536, 65536 * 4 + 3, 1_000_003, 1 << 23])
def test_add(bl):
    rng = np.random.Generator(np.random.PCG64(bl + 99))
    c = ctx()
    m = (1 << bl) - 1
    for name, a, b in carry_cases(rng, bl):
        da, db = c.upload(a, bl), c.upload(b, bl)
        ds = c.op("add", da, db)
        got = ds.download()
        exp = (int.from_bytes(a.tobytes(), "little") + int.from_bytes(b.tobytes(), "little")) & m
        assert got.tobytes() == exp.to_bytes(len(a) * 8, "little"), f"add {name} bl={bl}"
        assert np.array_equal(mask_words(oracle_words_op("add", a, b), bl), got)
        full = ds.download(ds.alloc_words)
        assert not full[len(a):].any()
        for d in (da, db, ds):
            d.free()


def test_add_repeated_launches_reuse_scan_state():
    """The lookback state resets itself: many launches in a row stay exact."""
    rng = np.random.Generator(np.random.PCG64(5))
    c = ctx()
    bl = 300_001
    m = (1 << bl) - 1
    for i in range(20):
        a, b = rand_stream(rng, bl), rand_stream(rng, bl)
        a[10:200] = np.uint64(2**64 - 1); b[10:200] = 0; b[9] = np.uint64(2**64 - 1); a[9] = np.uint64(2**64 - 1)
        da, db = c.upload(a, bl), c.upload(b, bl)
        ds = c.op("add", da, db)
        exp = (int.from_bytes(a.tobytes(), "little") + int.from_bytes(b.tobytes(), "little")) & m
        assert ds.download().tobytes() == exp.to_bytes(len(a) * 8, "little"), i
        for d in (da, db, ds):
            d.free()


@pytest.mark.parametrize("bl", [1, 63, 64, 65, 4097, 1_000_003, 1 << 24])
def test_popcount_and_nonzero(bl):
    rng = np.random.Generator(np.random.PCG64(bl + 3))
    c = ctx()
    a = rand_stream(rng, bl)
    da = c.upload(a, bl)
    assert c.popcount(da) == int(np.unpackbits(a.view(np.uint8)).sum())
    assert c.is_nonzero(da) == bool(a.any())
    z = np.zeros_like(a)
    dz = c.upload(z, bl)
    assert c.popcount(dz) == 0 and not c.is_nonzero(dz)
    top = np.zeros_like(a)
    top[-1] = np.uint64(1 << ((bl - 1) % 64))
    dt = c.upload(top, bl)
    assert c.popcount(dt) == 1 and c.is_nonzero(dt)
    ones = c.fill(bl, True)
    assert c.popcount(ones) == bl
    for d in (da, dz, dt, ones):
        d.free()
