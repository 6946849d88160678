"""Seeded random regex-style DAG programs over Zipf-distributed text.

A generated program reads the 8 basis bit streams of a byte text (bit j of
byte i is bit i of stream b<j>, the transposition regex engines over
bitstreams start from) and computes a DAG of `n_nodes` generator nodes:

  40 % AND / OR    10 % XOR    10 % NOT    20 % advance by 1..8 (`<< k`)
  10 % add         10 % MatchStar / ScanThru

MatchStar(M, C) = ((M & C) + C) ^ C | M expands to 4 DSL statements and
ScanThru(M, C) = (M + C) & ~C to 3, so a 200k-node DAG is about 250k DSL ops.
Operands come from a locality window of the `window` previous nodes (and, with
a small probability, from the basis streams).  A node about to leave the
window unconsumed is forced in as an operand, so every node is live; the
nodes nobody consumed at the end are folded with ORs into `n_outputs` output
streams.  The same seed gives the same program on every machine.

Text bytes follow Zipf(s) over 256 symbols, rank r mapped to a byte by a
fixed permutation.  Nothing here runs on the GPU.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

from .bsdata import BsCase, n_words_for

KINDS = ("and", "or", "xor", "not", "adv", "add", "matchstar", "scanthru")
DEFAULT_MIX = {"and": 0.20, "or": 0.20, "xor": 0.10, "not": 0.10, "adv": 0.20, "add": 0.10,
               "matchstar": 0.05, "scanthru": 0.05}
DSL_OPS = {"and": 1, "or": 1, "xor": 1, "not": 1, "adv": 1, "add": 1, "matchstar": 4, "scanthru": 3}


@dataclass
class DagProgram:
    source: str
    n_nodes: int
    kind_counts: dict
    n_dsl_ops: int             # DSL statements (each one op)
    outputs: list = field(default_factory=list)
    seed: int = 0

    def digest(self) -> str:
        return hashlib.sha256(self.source.encode()).hexdigest()[:16]


def generate(n_nodes: int = 200_000, seed: int = 0, window: int = 256, n_outputs: int = 64,
             mix: dict | None = None, basis_prob: float = 0.05) -> DagProgram:
    mix = dict(mix or DEFAULT_MIX)
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = list(mix)
    probs = np.array([mix[k] for k in kinds], dtype=np.float64)
    probs /= probs.sum()
    kind_idx = rng.choice(len(kinds), size=n_nodes, p=probs)
    lines = ["// random regex-style DAG: %d nodes, seed %d, window %d" % (n_nodes, seed, window)]
    lines += [f"input stream b{j}" for j in range(8)]
    lines += [f"output stream o{j}" for j in range(n_outputs)]
    body: list[str] = []
    names = [f"b{j}" for j in range(8)]          # value names, basis first
    consumed = [True] * 8                         # inputs need no consumer
    counts = {k: 0 for k in KINDS}
    n_ops = 0
    tmp = 0

    def pick(i_node: int, forced: list) -> int:
        if forced:
            return forced.pop()
        if rng.random() < basis_prob:
            return int(rng.integers(0, 8))
        lo = max(0, len(names) - window)
        return int(rng.integers(lo, len(names)))

    for i in range(n_nodes):
        kind = kinds[int(kind_idx[i])]
        counts[kind] += 1
        forced = []
        old = len(names) - window
        if old >= 8 and not consumed[old]:
            forced.append(old)
        a = pick(i, forced)
        arity = 1 if kind in ("not", "adv") else 2
        b = pick(i, forced) if arity == 2 else -1
        consumed[a] = True
        if b >= 0:
            consumed[b] = True
        v = f"v{len(names)}"
        A, B = names[a], names[b] if b >= 0 else None
        if kind == "and":
            body.append(f"stream {v} = {A} & {B}")
        elif kind == "or":
            body.append(f"stream {v} = {A} | {B}")
        elif kind == "xor":
            body.append(f"stream {v} = {A} ^ {B}")
        elif kind == "not":
            body.append(f"stream {v} = ~{A}")
        elif kind == "adv":
            body.append(f"stream {v} = {A} << {int(rng.integers(1, 9))}")
        elif kind == "add":
            body.append(f"stream {v} = {A} + {B}")
        elif kind == "matchstar":     # M = A, C = B
            t = f"t{tmp}"
            tmp += 1
            body += [f"stream {t}a = {A} & {B}", f"stream {t}b = {t}a + {B}", f"stream {t}c = {t}b ^ {B}",
                     f"stream {v} = {t}c | {A}"]
        else:                         # scanthru: (M + C) & ~C
            t = f"t{tmp}"
            tmp += 1
            body += [f"stream {t}a = {A} + {B}", f"stream {t}b = ~{B}", f"stream {v} = {t}a & {t}b"]
        n_ops += DSL_OPS[kind]
        names.append(v)
        consumed.append(False)

    # fold the sinks into the outputs
    sinks = [j for j in range(8, len(names)) if not consumed[j]]
    if not sinks:
        sinks = [len(names) - 1]
    groups = [sinks[g::n_outputs] for g in range(min(n_outputs, len(sinks)))]
    while len(groups) < n_outputs:                # fewer sinks than outputs: reuse late nodes
        groups.append([len(names) - 1 - len(groups) % max(1, len(names) - 8)])
    outputs = []
    for j, grp in enumerate(groups):
        acc = names[grp[0]]
        for k, g in enumerate(grp[1:]):
            nxt = f"o{j}" if k == len(grp) - 2 else f"f{j}_{k}"
            body.append(f"{'stream ' if nxt != f'o{j}' else ''}{nxt} = {acc} | {names[g]}")
            n_ops += 1
            acc = nxt
        if len(grp) == 1:                         # a single sink: OR with itself keeps 3-address form
            body.append(f"o{j} = {acc} | {acc}")
            n_ops += 1
        outputs.append(f"o{j}")
    src = "\n".join(lines + body) + "\n"
    return DagProgram(source=src, n_nodes=n_nodes, kind_counts=counts, n_dsl_ops=n_ops, outputs=outputs, seed=seed)


_ZIPF_PERM_SEED = 0x5EED


def zipf_bytes(n: int, seed: int, s: float = 1.1, n_sym: int = 256) -> np.ndarray:
    """n bytes, rank r (1-based) drawn with probability proportional to r^-s,
    rank r written as byte perm[r-1] (a fixed permutation)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w = np.arange(1, n_sym + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w / w.sum())
    perm = np.random.Generator(np.random.PCG64(_ZIPF_PERM_SEED)).permutation(n_sym).astype(np.uint8)
    out = np.empty(n, dtype=np.uint8)
    chunk = 1 << 24
    for i in range(0, n, chunk):
        m = min(chunk, n - i)
        r = np.searchsorted(cdf, rng.random(m), side="right")
        np.minimum(r, n_sym - 1, out=r)
        out[i:i + m] = perm[r]
    return out


def basis_streams(text: np.ndarray) -> list[np.ndarray]:
    """The 8 basis bit streams of a byte text: bit i of stream j = bit j of
    byte i (little-endian words)."""
    n = len(text)
    nw = n_words_for(n)
    out = []
    for j in range(8):
        bits = (text >> np.uint8(j)) & np.uint8(1)
        b = np.packbits(bits, bitorder="little")
        buf = np.zeros(nw * 8, dtype=np.uint8)
        buf[: len(b)] = b
        out.append(buf.view(np.uint64))
    return out


def text_case(n_bytes: int, seed: int, s: float = 1.1) -> BsCase:
    text = zipf_bytes(n_bytes, seed, s)
    planes = basis_streams(text)
    return BsCase(bitlength=n_bytes, inputs={f"b{j}": planes[j] for j in range(8)}, name=f"zipf{seed}")
