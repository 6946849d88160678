"""Python view of the C ABI (include/bsz_hip.h): contexts, device streams,
programs, cases and runs.

This is plumbing for tests, benchmarks and Python callers; every operation
executes in libbsz_hip.so on the GPU.
"""

from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import _lib
from .bsdata import BsCase
from ._lib import as_u64, check, lib, ptr

PEROP, FUSED = 0, 1


class DevStream:
    """A device stream buffer owned by a Context (row-padded, clean)."""

    def __init__(self, ctx: "Context", dev, bitlength: int):
        self.ctx, self.dev, self.bitlength = ctx, dev, bitlength

    @property
    def n_words(self) -> int:
        return 1 if self.bitlength <= 0 else (self.bitlength + 63) // 64

    @property
    def alloc_words(self) -> int:
        return int(lib().bsz_stream_alloc_words(self.bitlength))

    def download(self, n: int | None = None) -> np.ndarray:
        n = self.n_words if n is None else n
        out = np.zeros(n, dtype=np.uint64)
        check(lib().bsz_d2h(self.ctx.h, ptr(out), self.dev, n), self.ctx.h)
        return out

    def free(self):
        if self.dev is not None:
            check(lib().bsz_buf_free(self.ctx.h, self.dev), self.ctx.h)
            self.dev = None


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().bsz_ctx_create(C.byref(h), device))
        self.h = h

    def close(self):
        if self.h:
            check(lib().bsz_ctx_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(lib().bsz_sync(self.h), self.h)

    @property
    def stream_handle(self) -> int:
        return lib().bsz_ctx_stream(self.h) or 0

    # ── streams and per-op kernels ──
    def alloc(self, bitlength: int) -> DevStream:
        p = _lib.u64p()
        check(lib().bsz_buf_alloc(self.h, bitlength, C.byref(p)), self.h)
        return DevStream(self, p, bitlength)

    def upload(self, words: np.ndarray, bitlength: int) -> DevStream:
        s = self.alloc(bitlength)
        w = as_u64(words)
        check(lib().bsz_h2d(self.h, s.dev, ptr(w), len(w), bitlength), self.h)
        return s

    def _new_like(self, a: DevStream) -> DevStream:
        return self.alloc(a.bitlength)

    def op(self, name: str, a: DevStream, b: DevStream | None = None, k: int = 0) -> DevStream:
        d = self._new_like(a)
        L = lib()
        if name in ("and", "or", "xor", "add"):
            check(getattr(L, "bsz_op_" + name)(self.h, d.dev, a.dev, b.dev, a.bitlength), self.h)
        elif name in ("not", "copy"):
            check(getattr(L, "bsz_op_" + name)(self.h, d.dev, a.dev, a.bitlength), self.h)
        elif name in ("shl", "shr"):
            check(getattr(L, "bsz_op_" + name)(self.h, d.dev, a.dev, k, a.bitlength), self.h)
        else:
            raise ValueError(name)
        return d

    def fill(self, bitlength: int, ones: bool) -> DevStream:
        d = self.alloc(bitlength)
        check(lib().bsz_op_fill(self.h, d.dev, 1 if ones else 0, bitlength), self.h)
        return d

    def popcount(self, a: DevStream) -> int:
        v = C.c_int64()
        check(lib().bsz_op_popcount(self.h, a.dev, a.bitlength, C.byref(v)), self.h)
        return v.value

    def is_nonzero(self, a: DevStream) -> bool:
        v = C.c_int()
        check(lib().bsz_op_is_nonzero(self.h, a.dev, a.bitlength, C.byref(v)), self.h)
        return bool(v.value)

    # ── programs ──
    def run(self, program: "Program", case: "Case", executor: int = FUSED) -> "Result":
        r = C.c_void_p()
        check(lib().bsz_run(self.h, program.h, case.h, executor, C.byref(r)), self.h)
        return Result(r, program)


class Program:
    def __init__(self, source: str):
        h = C.c_void_p()
        check(lib().bsz_program_parse(source.encode(), C.byref(h)))
        self.h = h
        self.source = source

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsz_program_free(self.h)
            self.h = None

    def static_ops(self, case: "Case") -> int:
        return int(lib().bsz_program_static_ops(self.h, case.h))


class Case:
    def __init__(self, bitlength: int = 0):
        h = C.c_void_p()
        check(lib().bsz_case_create(C.byref(h), bitlength))
        self.h = h
        self.bitlength = bitlength

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsz_case_free(self.h)
            self.h = None

    @classmethod
    def from_bscase(cls, c: BsCase) -> "Case":
        k = cls(c.bitlength)
        L = lib()
        for n, v in c.params.items():
            L.bsz_case_set_param(k.h, n.encode(), int(v))
        for n, w in c.inputs.items():
            w = as_u64(w)
            L.bsz_case_set_input(k.h, n.encode(), ptr(w), len(w))
        for n, arr in c.input_arrays.items():
            for i, w in arr.items():
                w = as_u64(w)
                L.bsz_case_set_array_elem(k.h, n.encode(), int(i), ptr(w), len(w))
        for n, (packed, count) in c.broadcast.items():
            L.bsz_case_set_broadcast(k.h, n.encode(), packed, count)
        if c.iterations > 1:
            L.bsz_case_set_iterations(k.h, c.iterations)
        for o, i in c.feedback.items():
            L.bsz_case_set_feedback(k.h, o.encode(), i.encode())
        return k


class Result:
    def __init__(self, h, program: Program):
        self.h = h
        self.program = program

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsz_result_free(self.h)
            self.h = None

    @property
    def n_words(self) -> int:
        return int(lib().bsz_result_n_words(self.h))

    @property
    def op_count(self) -> int:
        return int(lib().bsz_result_op_count(self.h))

    def stream(self, name: str):
        out = np.zeros(self.n_words, dtype=np.uint64)
        ok = lib().bsz_result_stream(self.h, name.encode(), ptr(out), len(out))
        return out if ok else None

    def array(self, name: str):
        L = lib()
        n = L.bsz_result_array_count(self.h, name.encode())
        if n == 0:
            return None
        idx = np.zeros(n, dtype=np.int64)
        L.bsz_result_array_indices(self.h, name.encode(), idx.ctypes.data_as(_lib.i64p), n)
        res = {}
        for i in idx.tolist():
            out = np.zeros(self.n_words, dtype=np.uint64)
            L.bsz_result_array_elem(self.h, name.encode(), i, ptr(out), len(out))
            res[i] = out
        return res

    def int_out(self, name: str):
        v = C.c_int64()
        return v.value if lib().bsz_result_int(self.h, name.encode(), C.byref(v)) else None

    def json(self) -> dict:
        L = lib()
        n = L.bsz_result_json(self.h, None, 0)
        buf = C.create_string_buffer(n + 1)
        L.bsz_result_json(self.h, buf, n + 1)
        return json.loads(buf.value.decode())
