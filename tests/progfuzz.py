"""Seeded random .bs programs and inputs for differential tests.

Programs use every construct of the DSL (DSL.md:303-357): scalar and array
inputs, params, locals, every stream operator, ZERO/ONES, array reads and
writes, `for` loops with the loop variable in indices and shift amounts,
`if` and bounded `while` on streams, popcount into ints that later feed int
arithmetic, shift amounts and loop bounds, output ints and output arrays.
"""

from __future__ import annotations

import numpy as np

from bitstreamzoo_amd.bsdata import BsCase, n_words_for


class _Gen:
    def __init__(self, rng: np.random.Generator, n_stmts: int, allow_control: bool, big_shifts: bool,
                 host_ints: bool = True):
        self.r = rng
        self.lines: list[str] = []
        self.streams = ["x0", "x1", "x2"]
        self.ints = ["P"]
        self.n = 0
        self.n_stmts = n_stmts
        self.control = allow_control
        self.big_shifts = big_shifts
        self.depth = 0
        self.host_ints = host_ints

    def fresh(self, p="v"):
        self.n += 1
        return f"{p}{self.n}"

    def prim(self):
        r = self.r.random()
        if r < 0.06:
            return "ZERO"
        if r < 0.12:
            return "ONES"
        if r < 0.22:
            return f"arr[{int(self.r.integers(0, 5))}]"   # index 4 is never written: reads ZERO
        return str(self.r.choice(self.streams))

    def shift_amount(self):
        r = self.r.random()
        if r < 0.5:
            return str(int(self.r.integers(0, 64)))
        if r < 0.7 and self.big_shifts:
            return str(int(self.r.choice([64, 65, 127, 128, 200, 1000, 4095, 4096, 4097])))
        if r < 0.85:
            return str(self.r.choice(self.ints)) + " * 0 + " + str(int(self.r.integers(0, 9)))
        return "P"

    def expr(self):
        k = self.r.integers(0, 9)
        a, b = self.prim(), self.prim()
        if k == 0:
            return f"~{a}"
        if k in (1, 2, 3, 4):
            return f"{a} {['&', '|', '^', '+'][k - 1]} {b}"
        if k in (5, 6):
            return f"{a} << {self.shift_amount()}"
        if k == 7:
            return f"{a} >> {self.shift_amount()}"
        return a

    def emit(self, s):
        self.lines.append("    " * self.depth + s)

    def stmt(self):
        r = self.r.random()
        if self.control and self.depth < 2 and r < 0.05:
            n = int(self.r.integers(1, 4))
            iv = self.fresh("i")
            self.emit(f"for {iv} in 0..{n} {{")
            self.depth += 1
            acc = self.fresh("acc")
            self.emit(f"stream {acc} = arr[{iv}] ^ {self.prim()}")
            self.emit(f"arr[{iv} + 1] = {acc} << {iv}")
            self.ints.append(iv)
            for _ in range(int(self.r.integers(1, 3))):
                self.stmt()
            self.ints.remove(iv)
            self.depth -= 1
            self.emit("}")
            self.streams.append(acc)
            return
        if self.control and self.depth < 2 and r < 0.09:
            c = str(self.r.choice(self.streams))
            self.emit(f"if ({c}) {{")
            self.depth += 1
            for _ in range(int(self.r.integers(1, 3))):
                self.stmt()
            self.depth -= 1
            self.emit("}")
            return
        if self.control and self.depth < 2 and r < 0.12:
            w = self.fresh("w")
            self.emit(f"stream {w} = {self.prim()} | x1")
            self.emit(f"while ({w}) {{")
            self.depth += 1
            self.emit(f"{w} = {w} << {int(self.r.integers(5, 40))}")
            t = self.fresh("t")
            self.emit(f"stream {t} = {w} ^ {self.prim()}")
            self.emit(f"loopacc = loopacc ^ {t}")
            self.depth -= 1
            self.emit("}")
            return
        if r < 0.17 and self.host_ints:
            v = self.fresh("n")
            self.emit(f"int {v} = popcount({self.r.choice(self.streams)})")
            if self.depth == 0:
                self.ints.append(v)
            return
        if r < 0.20 and len(self.ints) > 1:
            self.emit(f"cnt = cnt + {self.r.choice(self.ints)} - 1")
            return
        if r < 0.25:
            self.emit(f"arr[{int(self.r.integers(0, 4))}] = {self.expr()}")
            return
        if r < 0.35 and self.depth == 0:
            v = str(self.r.choice(self.streams))
            if v.startswith("v") and v != "x0":
                self.emit(f"{v} = {self.expr()}")
                return
        v = self.fresh()
        self.emit(f"stream {v} = {self.expr()}")
        if self.depth == 0:
            self.streams.append(v)

    def program(self) -> str:
        hdr = ["input stream x0", "input stream x1", "input stream x2", "input stream arr[4]", "param int P",
               "output stream o0", "output stream o1", "output stream oarr[3]", "output stream loopacc",
               "output int cnt", "output int pc"]
        self.lines = []
        for _ in range(self.n_stmts):
            self.stmt()
        tail = [f"o0 = {self.streams[-1]}", f"o1 = {self.r.choice(self.streams)} ^ {self.r.choice(self.streams)}"]
        for i in range(3):
            tail.append(f"oarr[{i}] = {self.r.choice(self.streams)} + arr[{i}]")
        tail.append(f"pc = popcount({self.r.choice(self.streams)})")
        return "\n".join(hdr + self.lines + tail) + "\n"


def random_program(seed: int, n_stmts: int = 30, allow_control: bool = True, big_shifts: bool = True,
                   host_ints: bool = True) -> str:
    """host_ints=False: popcounts only feed int outputs, never int
    expressions, so the program records as one fused plan."""
    return _Gen(np.random.Generator(np.random.PCG64(seed)), n_stmts, allow_control, big_shifts,
                host_ints).program()


def random_case(seed: int, bitlength: int, density: float | None = None) -> BsCase:
    """Inputs for random_program: mixes dense, sparse, all-ones runs."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    nw = n_words_for(bitlength)

    def stream():
        kind = rng.integers(0, 4) if density is None else 0
        if kind == 0:
            w = rng.integers(0, 2**64, size=nw, dtype=np.uint64)
        elif kind == 1:  # sparse
            w = np.zeros(nw, dtype=np.uint64)
            for _ in range(max(1, bitlength // 200)):
                b = int(rng.integers(0, bitlength))
                w[b // 64] |= np.uint64(1 << (b % 64))
        elif kind == 2:  # long all-ones runs (carry chains)
            w = np.full(nw, np.uint64(2**64 - 1))
            for _ in range(int(rng.integers(0, 4))):
                b = int(rng.integers(0, bitlength))
                w[b // 64] &= ~np.uint64(1 << (b % 64))
        else:
            w = np.zeros(nw, dtype=np.uint64)
        rem = bitlength % 64
        if rem:
            w[-1] &= np.uint64((1 << rem) - 1)
        return w

    c = BsCase(bitlength=bitlength, params={"P": int(rng.integers(0, 70))})
    c.inputs = {f"x{i}": stream() for i in range(3)}
    c.input_arrays = {"arr": {i: stream() for i in range(4)}}
    return c


def bool_chain_program(seed: int, n_out: int = 24) -> str:
    """Outputs that are chains of NOT/AND/OR/XOR over four inputs; every
    temporary feeds the next link only (so the generator may fold a link into
    a three-input function), except where a chain reuses an earlier
    temporary as its other operand (so it may not)."""
    rng = np.random.default_rng(seed)
    lines = ["input stream a", "input stream b", "input stream c", "input stream d"]
    lines += [f"output stream r{i}" for i in range(n_out)] + ["output int n0"]
    t = 0
    for i in range(n_out):
        cur, made = str(rng.choice(["a", "b", "c", "d"])), []
        for _ in range(int(rng.integers(1, 6))):
            t += 1
            op = str(rng.choice(["&", "|", "^", "~"]))
            if op == "~":
                lines.append(f"stream t{t} = ~{cur}")
            else:
                pool = ["a", "b", "c", "d"] + (made[:-1] if rng.random() < 0.3 else [])
                other = str(rng.choice(pool))
                lines.append(f"stream t{t} = {cur} {op} {other}" if rng.random() < 0.5 else
                             f"stream t{t} = {other} {op} {cur}")
            made.append(f"t{t}")
            cur = f"t{t}"
        lines.append(f"r{i} = {cur}")
    lines.append("n0 = popcount(r0)")
    return "\n".join(lines) + "\n"

