"""Seeded synthetic workloads: the .bs program and the inputs of each
benchmark configuration, at any size.

  matchstar      MatchStar t0=M&C; t1=t0+C; t2=t1^C; result=t2|M and its
                 popcount; M ~ Bernoulli(0.02), C alternating 1-runs / 0-runs
                 of geometric lengths (means 48 / 16) so carries cross words.
  adder32        bitsliced 32-bit ripple-carry adder as a straight-line DAG
                 (per bit p=a^b, s=p^c, g=a&b, t=p&c, c'=g|t; the pattern of
                 benchmark/apps/circuit_sim/src/netlist_adder4.bs), 65 inputs,
                 33 outputs, a popcount of each output.
  shift_and      DNA exact match of a 32-mer (benchmark/apps/shift_and):
                 match = AND_i (class[p_i] << (31-i)) over 4 one-hot classes,
                 plus its popcount; batches of sequences of uneven length.

Random numbers come from numpy's PCG64 with the seed given; the same seed
gives the same bits on every machine.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .bsdata import BsCase, n_words_for


def pack_bits(bits: np.ndarray, bitlength: int) -> np.ndarray:
    """bool[bitlength] -> little-endian uint64 words (bit 0 = LSB of word 0)."""
    nw = n_words_for(bitlength)
    b = np.packbits(bits.astype(np.uint8), bitorder="little")
    out = np.zeros(nw * 8, dtype=np.uint8)
    out[: len(b)] = b
    return out.view(np.uint64)


def random_words(rng: np.random.Generator, bitlength: int) -> np.ndarray:
    """i.i.d. Bernoulli(0.5) stream, masked to the bitlength."""
    nw = n_words_for(bitlength)
    w = rng.integers(0, 2**64, size=nw, dtype=np.uint64)
    rem = bitlength % 64
    if rem:
        w[-1] &= np.uint64((1 << rem) - 1)
    return w


def bernoulli_words(rng: np.random.Generator, bitlength: int, p: float, chunk: int = 1 << 24) -> np.ndarray:
    nw = n_words_for(bitlength)
    out = np.zeros(nw, dtype=np.uint64)
    thr = np.uint32(min(2**32 - 1, int(p * 2**32)))
    for start in range(0, bitlength, chunk):
        n = min(chunk, bitlength - start)
        bits = rng.integers(0, 2**32, size=n, dtype=np.uint32) < thr
        b = np.packbits(bits.astype(np.uint8), bitorder="little")
        w0 = start // 64
        buf = np.zeros(((len(b) + 7) // 8) * 8, dtype=np.uint8)
        buf[: len(b)] = b
        out[w0: w0 + len(buf) // 8] |= buf.view(np.uint64)
    return out


def run_stream(rng: np.random.Generator, bitlength: int, mean_one: float, mean_zero: float) -> np.ndarray:
    """Alternating runs of ones and zeros with geometric lengths (first run's
    value drawn at random)."""
    first = int(rng.integers(0, 2))
    lens_parts, total = [], 0
    while total < bitlength:
        n = int(2 * (bitlength - total) / (mean_one + mean_zero)) // 2 + 1024
        l1 = rng.geometric(1.0 / mean_one, size=n)
        l0 = rng.geometric(1.0 / mean_zero, size=n)
        lens = np.empty(2 * n, dtype=np.int64)
        lens[0::2], lens[1::2] = (l1, l0) if first else (l0, l1)
        lens_parts.append(lens)
        total += int(lens.sum())
    lens = np.concatenate(lens_parts)
    vals = np.zeros(len(lens), dtype=np.uint8)
    vals[0 if first else 1::2] = 1
    bits = np.repeat(vals, lens)[:bitlength]
    return pack_bits(bits, bitlength)


def _compact_even_bits(x: np.ndarray) -> np.ndarray:
    """Bits 0,2,4,..,62 of each uint64 packed into the low 32 bits."""
    x = x & np.uint64(0x5555555555555555)
    x = (x | (x >> np.uint64(1))) & np.uint64(0x3333333333333333)
    x = (x | (x >> np.uint64(2))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    x = (x | (x >> np.uint64(4))) & np.uint64(0x00FF00FF00FF00FF)
    x = (x | (x >> np.uint64(8))) & np.uint64(0x0000FFFF0000FFFF)
    x = (x | (x >> np.uint64(16))) & np.uint64(0x00000000FFFFFFFF)
    return x


def dna_onehot(rng: np.random.Generator, length: int):
    """Uniform ACGT text of `length` bases as 4 one-hot streams, generated
    from random 2-bit codes 32 bases per 64-bit word.  Returns the four
    streams in BASES order."""
    nw = n_words_for(length)
    codes = rng.integers(0, 2**64, size=2 * nw, dtype=np.uint64)   # 32 bases per code word
    lo = codes
    hi = codes >> np.uint64(1)
    m = np.uint64(0x5555555555555555)
    planes = [(~hi) & (~lo) & m, (~hi) & lo & m, hi & (~lo) & m, hi & lo & m]
    out = []
    for pl in planes:
        half = _compact_even_bits(pl)
        w = half[0::2] | (half[1::2] << np.uint64(32))
        rem = length % 64
        if rem:
            w[-1] &= np.uint64((1 << rem) - 1)
        out.append(w)
    return out


def onehot_to_codes(streams, length: int) -> np.ndarray:
    """Inverse of dna_onehot for tests: base code (0..3) per position."""
    codes = np.zeros(length, dtype=np.uint8)
    for k, w in enumerate(streams):
        bits = np.unpackbits(w.view(np.uint8), bitorder="little")[:length].astype(bool)
        codes[bits] = k
    return codes


@dataclass
class Workload:
    name: str
    source: str
    cases: list           # [BsCase]
    bit_ops: int          # Σ executed DSL ops x bitlength (popcount counts 1)
    alg_bytes: int        # distinct inputs read once + outputs written once (+8 per popcount)


MATCHSTAR_BS = """\
// MatchStar(M, C) = (((M & C) + C) ^ C) | M, and its popcount.
input stream M
input stream C
output stream result
output int n
stream t0 = M & C
stream t1 = t0 + C
stream t2 = t1 ^ C
result = t2 | M
n = popcount(result)
"""


def matchstar(bitlength: int = 1 << 26, seed: int = 1) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    M = bernoulli_words(rng, bitlength, 0.02)
    C = run_stream(rng, bitlength, 48.0, 16.0)
    c = BsCase(bitlength=bitlength, inputs={"M": M, "C": C}, name=f"matchstar_{bitlength}")
    nw = n_words_for(bitlength)
    return Workload("matchstar", MATCHSTAR_BS, [c], bit_ops=5 * bitlength, alg_bytes=3 * 8 * nw + 8)


def adder32_source(width: int = 32) -> str:
    lines = ["// Bitsliced ripple-carry adder, one full adder per bit (5 ops each)."]
    for i in range(width):
        lines.append(f"input stream a{i}")
    for i in range(width):
        lines.append(f"input stream b{i}")
    lines.append("input stream cin")
    for i in range(width):
        lines.append(f"output stream s{i}")
    lines.append("output stream cout")
    for i in range(width):
        lines.append(f"output int n_s{i}")
    lines.append("output int n_cout")
    c = "cin"
    for i in range(width):
        lines += [
            f"stream p{i} = a{i} ^ b{i}",
            f"s{i} = p{i} ^ {c}",
            f"stream g{i} = a{i} & b{i}",
            f"stream t{i} = p{i} & {c}",
            f"stream c{i + 1} = g{i} | t{i}",
            f"n_s{i} = popcount(s{i})",
        ]
        c = f"c{i + 1}"
    lines.append(f"cout = {c}")
    lines.append("n_cout = popcount(cout)")
    return "\n".join(lines) + "\n"


def adder32(bitlength: int = 1 << 28, seed: int = 2, width: int = 32) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    inputs = {}
    for i in range(width):
        inputs[f"a{i}"] = random_words(rng, bitlength)
    for i in range(width):
        inputs[f"b{i}"] = random_words(rng, bitlength)
    inputs["cin"] = random_words(rng, bitlength)
    c = BsCase(bitlength=bitlength, inputs=inputs, name=f"adder{width}_{bitlength}")
    nw = n_words_for(bitlength)
    n_ops = 5 * width + (width + 1)  # full adders + popcounts
    return Workload(f"adder{width}", adder32_source(width), [c], bit_ops=n_ops * bitlength,
                    alg_bytes=(2 * width + 1 + width + 1) * 8 * nw + 8 * (width + 1))


BASES = "ACGT"


def shift_and_source(pattern: str) -> str:
    m = len(pattern)
    lines = [f"// Shift-And exact match of the {m}-mer {pattern}: match = AND_i (class[p_i] << ({m - 1}-i)).",
             "input stream cA", "input stream cC", "input stream cG", "input stream cT",
             "output stream match", "output int n"]
    prev = None
    for i, ch in enumerate(pattern):
        sh = m - 1 - i
        cur = f"s{i}"
        if sh > 0:
            lines.append(f"stream {cur} = c{ch} << {sh}")
        else:
            lines.append(f"stream {cur} = c{ch}")
        if prev is None:
            prev = cur
        else:
            lines.append(f"stream m{i} = {prev} & {cur}")
            prev = f"m{i}"
    lines.append(f"match = {prev}")
    lines.append("n = popcount(match)")
    return "\n".join(lines) + "\n"


def dna_case(rng: np.random.Generator, length: int, name: str) -> BsCase:
    planes = dna_onehot(rng, length)
    inputs = {f"c{b}": planes[k] for k, b in enumerate(BASES)}
    return BsCase(bitlength=length, inputs=inputs, name=name)


def shift_and_batch(n_seqs: int = 2048, min_log2: int = 12, max_log2: int = 22, seed: int = 3,
                    m: int = 32) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    pattern = "".join(BASES[i] for i in rng.integers(0, 4, size=m))
    lens = np.exp2(rng.uniform(min_log2, max_log2, size=n_seqs)).astype(np.int64)
    cases = [dna_case(rng, int(L), f"seq{i}") for i, L in enumerate(lens)]
    ops_per_bit = 2 * (m - 1) + 1   # (m-1) advances + (m-1) ANDs + popcount
    bit_ops = int(sum(ops_per_bit * int(L) for L in lens))
    alg = int(sum(5 * 8 * n_words_for(int(L)) + 8 for L in lens))
    return Workload("shift_and", shift_and_source(pattern), cases, bit_ops=bit_ops, alg_bytes=alg)


def shift_and_scalar(seq_codes: np.ndarray, pattern: str) -> np.ndarray:
    """Positions j where text[j-m+1..j] == pattern (the scalar reference of
    benchmark/apps/shift_and/src/run.py:24-47, restated with numpy)."""
    m = len(pattern)
    pc = np.array([BASES.index(ch) for ch in pattern], dtype=np.uint8)
    n = len(seq_codes)
    if n < m:
        return np.zeros(0, dtype=np.int64)
    win = np.lib.stride_tricks.sliding_window_view(seq_codes, m)
    hits = np.nonzero((win == pc).all(axis=1))[0]
    return hits + (m - 1)
