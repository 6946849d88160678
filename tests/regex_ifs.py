"""Seeded programs shaped like the reference's compiled regex matchers
(benchmark/apps/regex/src/regex_medium.bs): 8 basis bit-planes b[0..7], per
pattern a chain of byte-class matches joined by advances, and `if (cursor)`
guards around the rest of the pattern, nested up to three deep, whose bodies
only assign streams; every pattern ends in an output and all are or-ed into
`any_match`.  Used to test `if` bodies that stay on the device."""

from __future__ import annotations

import numpy as np


ALPHABET = (0x41, 0x43, 0x47, 0x54, 0x0a, 0xff)   # the bytes texts are drawn from


def byte_class(rng, lines, name, planes):
    """A stream matching one byte value on the planes (bit j of the byte)."""
    v = int(ALPHABET[int(rng.integers(0, len(ALPHABET)))])
    terms = []
    for j in range(8):
        if (v >> j) & 1:
            terms.append(f"b[{j}]")
        else:
            nb = f"{name}_nb{j}"
            lines.append(f"stream {nb} = ~b[{j}]")
            terms.append(nb)
    acc = terms[0]
    for k, t in enumerate(terms[1:]):
        nxt = f"{name}_eq{k}"
        lines.append(f"stream {nxt} = {acc} & {t}")
        acc = nxt
    return acc


def program(n_patterns: int, seed: int) -> str:
    rng = np.random.Generator(np.random.PCG64(seed))
    lines = ["input stream b[8]", "output stream any_match"]
    outs = []
    body = []
    for p in range(n_patterns):
        nm = f"_p{p}"
        out = f"m{p}"
        outs.append(out)
        lines.append(f"output stream {out}")
        cls = [byte_class(rng, body, f"{nm}c{i}", 8) for i in range(int(rng.integers(2, 5)))]
        cur = cls[0]
        for i, c in enumerate(cls[1:]):
            adv = f"{nm}a{i}"
            body.append(f"stream {adv} = {cur} << 1")
            t = f"{nm}t{i}"
            body.append(f"stream {t} = {adv} & {c}")
            cur = t
        vm = f"{nm}vm"
        body.append(f"stream {vm} = ZERO")
        depth = int(rng.integers(1, 4))
        ind = ""
        guards = []
        for d in range(depth):
            body.append(f"{ind}if ({cur}) {{")
            guards.append(ind)
            ind += ""
            rep = cls[int(rng.integers(0, len(cls)))]
            for r in range(int(rng.integers(1, 4))):
                adv = f"{nm}g{d}a{r}"
                body.append(f"stream {adv} = {cur} << 1")
                t = f"{nm}g{d}t{r}"
                body.append(f"stream {t} = {rep} & {adv}")
                cur = t
            if rng.random() < 0.3:   # a ScanThru inside the guard: an addition
                s = f"{nm}g{d}s"
                body.append(f"stream {s} = {cur} + {rep}")
                nr = f"{nm}g{d}n"
                body.append(f"stream {nr} = ~{rep}")
                st = f"{nm}g{d}st"
                body.append(f"stream {st} = {s} & {nr}")
                cur = st
            body.append(f"{vm} = {cur}")
        for _ in range(depth):
            body.append("}")
        body.append(f"{out} = {vm}")
    acc = "ZERO"
    body.append("stream _acc0 = ZERO")
    acc = "_acc0"
    for i, o in enumerate(outs):
        nxt = f"_acc{i + 1}"
        body.append(f"stream {nxt} = {acc} | {o}")
        acc = nxt
    body.append(f"any_match = {acc}")
    return "\n".join(lines + body) + "\n"


def text_case(n_bytes: int, seed: int, alphabet: int = len(ALPHABET)):
    """Bytes drawn from the first `alphabet` bytes of ALPHABET (so patterns
    do match somewhere and some guards are false), transposed into the 8
    planes."""
    from bitstreamzoo_amd.bsdata import BsCase
    from bitstreamzoo_amd import workloads
    rng = np.random.Generator(np.random.PCG64(seed))
    vals = np.array(ALPHABET[:alphabet], dtype=np.uint8)
    text = vals[rng.integers(0, alphabet, size=n_bytes)]
    planes = {j: workloads.pack_bits(((text >> j) & 1).astype(bool), n_bytes) for j in range(8)}
    return BsCase(bitlength=n_bytes, input_arrays={"b": planes}, name=f"text{seed}")
