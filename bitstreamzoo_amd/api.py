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
from ._lib import BszError, as_u64, check, lib, ptr  # noqa: F401  (BszError is part of the API)

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

    def write(self, words: np.ndarray):
        """Overwrite the stream with host words (truncated to the bitlength)."""
        w = as_u64(words)
        check(lib().bsz_h2d(self.ctx.h, self.dev, ptr(w), len(w), self.bitlength), self.ctx.h)

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

    def timer_start(self):
        check(lib().bsz_timer_start(self.h), self.h)

    def timer_stop(self) -> float:
        v = C.c_double()
        check(lib().bsz_timer_stop(self.h, C.byref(v)), self.h)
        return v.value

    def order_after(self, other: "Context"):
        """Work enqueued on this context from now on runs after the work
        enqueued so far on `other` (inside a capture on `other`: forks the
        capture into this context's stream; `other.order_after(self)` joins
        it back)."""
        check(lib().bsz_stream_order(other.h, self.h), self.h)

    def capture(self, fn) -> "Graph":
        """Capture the launches `fn()` enqueues on this context into a graph."""
        check(lib().bsz_capture_begin(self.h), self.h)
        try:
            fn()
        finally:
            g = C.c_void_p()
            check(lib().bsz_capture_end(self.h, C.byref(g)), self.h)
        return Graph(self, g)

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


class Graph:
    def __init__(self, ctx: "Context", h):
        self.ctx, self.h = ctx, h

    def launch(self, times: int = 1):
        check(lib().bsz_graph_launch(self.ctx.h, self.h, times), self.ctx.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsz_graph_destroy(self.h)
            self.h = None


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

    def set_padding(self, word_bits: int) -> "Case":
        """Value semantics of the bits at or above the bitlength: 0 keeps
        them clear (DSL.md:273-275); 64 or 32 gives them the values bsim_cuda
        gives them with --word-bits=64|32 (its streams are padded to a
        multiple of 8 words, cuda_backend.cuh:23-25; ~, ONES, + and << leave
        the padding bits set and >> brings them back)."""
        check(lib().bsz_case_set_padding(self.h, int(word_bits)))
        return self

    @classmethod
    def from_bscase(cls, c: BsCase, padding: int = 0) -> "Case":
        k = cls(c.bitlength)
        L = lib()
        if padding:
            k.set_padding(padding)
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


def _get_str(fn, h) -> str:
    n = fn(h, None, 0)
    buf = C.create_string_buffer(n + 1)
    fn(h, buf, n + 1)
    return buf.value.decode()


def _fill_str(call, cap: int = 1 << 16) -> str:
    """The string of a `call(buf, cap) -> full length` entry point, from one
    call whose whole output fitted.  (Two calls, one for the size and one for
    the text, can disagree: a compile-only plan reports the register use of
    code objects that other processes may finish in between.)"""
    for _ in range(8):
        buf = C.create_string_buffer(cap)
        n = int(call(buf, cap))
        if n < 0:
            check(n)
        if n < cap:
            return buf.raw[:n].decode()
        cap = n + n // 4 + 1
    raise RuntimeError("output kept growing between calls")


def compile_plan_only(program: "Program", shape: "Case", options: dict | None = None) -> dict:
    """Record + generate + hipRTC-compile a fused plan (no GPU needed)."""
    opts = json.dumps(options or {}).encode()
    L = lib()
    return json.loads(_fill_str(lambda buf, cap: L.bsz_plan_compile_only(program.h, shape.h, opts, buf, cap),
                                cap=1 << 25))


def describe_plans(program: "Program", shape: "Case", options_list: list[dict]) -> list[dict]:
    """Kernel shapes of the program under each schedule (no code, no device)."""
    js = json.dumps(options_list).encode()
    L = lib()
    return json.loads(_fill_str(lambda buf, cap: L.bsz_plan_describe(program.h, shape.h, js, buf, cap),
                                cap=1 << 22))


class Plan:
    """A fused plan over a batch of instances (bsz_plan_*)."""

    def __init__(self, ctx: Context, program: Program, shape: Case, options: dict | None = None):
        h = C.c_void_p()
        check(lib().bsz_plan_create(ctx.h, program.h, shape.h, json.dumps(options or {}).encode(), C.byref(h)),
              ctx.h)
        self.h, self.ctx, self.program = h, ctx, program

    def close(self):
        if self.h:
            lib().bsz_plan_free(self.h)
            self.h = None

    def __del__(self):
        if getattr(self, "h", None):
            self.close()

    def set_instances(self, bitlengths):
        arr = (C.c_int64 * len(bitlengths))(*[int(b) for b in bitlengths])
        check(lib().bsz_plan_set_instances(self.h, len(bitlengths), arr), self.ctx.h)

    def bind(self, inst: int, name: str, dev: DevStream, idx: int = -1):
        check(lib().bsz_plan_bind_input(self.h, inst, name.encode(), idx, dev.dev), self.ctx.h)

    def output_ptr(self, inst: int, name: str, idx: int = -1):
        p = _lib.u64p()
        check(lib().bsz_plan_output(self.h, inst, name.encode(), idx, C.byref(p)), self.ctx.h)
        return p

    def download_output(self, inst: int, name: str, n_words: int, idx: int = -1, offset: int = 0) -> np.ndarray:
        """Words [offset, offset + n_words) of an output stream."""
        p = self.output_ptr(inst, name, idx)
        if offset:
            p = C.cast(C.c_void_p(C.cast(p, C.c_void_p).value + 8 * offset), _lib.u64p)
        out = np.zeros(n_words, dtype=np.uint64)
        check(lib().bsz_d2h(self.ctx.h, ptr(out), p, n_words), self.ctx.h)
        return out

    def run(self):
        check(lib().bsz_plan_run(self.h), self.ctx.h)

    def set_executor(self, executor: int):
        check(lib().bsz_plan_set_executor(self.h, executor), self.ctx.h)

    def prepare(self):
        check(lib().bsz_plan_prepare(self.h), self.ctx.h)

    def capture(self):
        check(lib().bsz_plan_capture(self.h), self.ctx.h)

    def replay(self, times: int = 1):
        check(lib().bsz_plan_replay(self.h, times), self.ctx.h)

    def read_int(self, inst: int, name: str) -> int:
        v = C.c_int64()
        check(lib().bsz_plan_read_int(self.h, inst, name.encode(), C.byref(v)), self.ctx.h)
        return v.value

    def mismatches(self, kernel: int = 0) -> int:
        v = C.c_int64()
        check(lib().bsz_plan_mismatches(self.h, kernel, C.byref(v)), self.ctx.h)
        return v.value

    def unproven(self, kernel: int = 0) -> int:
        """Chunks of `kernel` whose warm-up did not prove their start state
        exact in the last run."""
        v = C.c_int64()
        check(lib().bsz_plan_unproven(self.h, kernel, C.byref(v)), self.ctx.h)
        return v.value

    def info(self) -> dict:
        return json.loads(_get_str(lib().bsz_plan_info, self.h))

    def source(self) -> str:
        return _get_str(lib().bsz_plan_source, self.h)
