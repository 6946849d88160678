"""ctypes access to the two CPU models used as checkers.

* ``Oracle`` — oracle/bsz_oracle.so, this repository's CPU restatement of the
  reference path (always available once ``make -C oracle`` has run).
* ``RefInterp`` — oracle/_ref/libbsref.so, the reference's own
  Interpreter<CpuBackend> compiled from its sources (only where those sources
  were present at build time, and only on AVX-512 hosts).

Both expose the same case interface, so a test can run one program through
either and compare.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from bitstreamzoo_amd.bsdata import BsCase, words_to_int

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "bsz_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libbsref.so")

u64p = C.POINTER(C.c_uint64)
i64p = C.POINTER(C.c_int64)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(u64p)


class _Model:
    prefix = ""
    so_path = ""
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(cls.so_path):
                raise FileNotFoundError(f"{cls.so_path} not built (run: make -C oracle)")
            lib = C.CDLL(cls.so_path)
            p = cls.prefix
            f = getattr(lib, p + "case_new"); f.restype = C.c_void_p; f.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t]
            f = getattr(lib, p + "case_free"); f.argtypes = [C.c_void_p]
            f = getattr(lib, p + "set_bitlength"); f.argtypes = [C.c_void_p, C.c_int64]
            f = getattr(lib, p + "set_param"); f.argtypes = [C.c_void_p, C.c_char_p, C.c_int64]
            f = getattr(lib, p + "set_input"); f.argtypes = [C.c_void_p, C.c_char_p, u64p, C.c_uint64]
            f = getattr(lib, p + "set_array_elem"); f.argtypes = [C.c_void_p, C.c_char_p, C.c_int64, u64p, C.c_uint64]
            f = getattr(lib, p + "set_broadcast"); f.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int64]
            f = getattr(lib, p + "op_count"); f.restype = C.c_int64; f.argtypes = [C.c_void_p]
            f = getattr(lib, p + "exec_ms"); f.restype = C.c_double; f.argtypes = [C.c_void_p]
            f = getattr(lib, p + "get_stream"); f.argtypes = [C.c_void_p, C.c_char_p, u64p, C.c_uint64]
            f = getattr(lib, p + "array_count"); f.restype = C.c_int64; f.argtypes = [C.c_void_p, C.c_char_p]
            f = getattr(lib, p + "array_indices"); f.argtypes = [C.c_void_p, C.c_char_p, i64p, C.c_int64]
            f = getattr(lib, p + "get_array_elem"); f.argtypes = [C.c_void_p, C.c_char_p, C.c_int64, u64p, C.c_uint64]
            f = getattr(lib, p + "get_int"); f.argtypes = [C.c_void_p, C.c_char_p, i64p]
            cls._extra(lib)
            cls._lib = lib
        return cls._lib

    @classmethod
    def _extra(cls, lib):
        pass

    @classmethod
    def available(cls) -> bool:
        try:
            cls.lib()
            return True
        except (OSError, FileNotFoundError):
            return False

    def __init__(self, source: str):
        self.L = self.lib()
        err = C.create_string_buffer(4096)
        h = getattr(self.L, self.prefix + "case_new")(source.encode(), err, 4096)
        if not h:
            raise SyntaxError(err.value.decode())
        self.h = h
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.L, self.prefix + "case_free")(self.h)
            self.h = None

    def _f(self, name):
        return getattr(self.L, self.prefix + name)

    def load(self, case: BsCase):
        self.bitlength = case.bitlength
        self._f("set_bitlength")(self.h, case.bitlength)
        for k, v in case.params.items():
            self._f("set_param")(self.h, k.encode(), int(v))
        for k, w in case.inputs.items():
            w = np.ascontiguousarray(w, dtype=np.uint64)
            self._keep.append(w)
            self._f("set_input")(self.h, k.encode(), _ptr(w), len(w))
        for k, arr in case.input_arrays.items():
            for idx, w in arr.items():
                w = np.ascontiguousarray(w, dtype=np.uint64)
                self._keep.append(w)
                self._f("set_array_elem")(self.h, k.encode(), int(idx), _ptr(w), len(w))
        for k, (packed, count) in case.broadcast.items():
            self._f("set_broadcast")(self.h, k.encode(), packed, count)
        self.n_words = case.n_words
        return self

    def run(self, *args):
        err = C.create_string_buffer(4096)
        rc = self._run(err, *args)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return self

    def op_count(self) -> int:
        return self._f("op_count")(self.h)

    def exec_ms(self) -> float:
        return self._f("exec_ms")(self.h)

    def stream(self, name: str):
        out = np.zeros(self.n_words, dtype=np.uint64)
        ok = self._f("get_stream")(self.h, name.encode(), _ptr(out), len(out))
        return out if ok else None

    def array(self, name: str):
        n = self._f("array_count")(self.h, name.encode())
        if n == 0:
            return None
        idx = np.zeros(n, dtype=np.int64)
        self._f("array_indices")(self.h, name.encode(), idx.ctypes.data_as(i64p), n)
        res = {}
        for i in idx.tolist():
            out = np.zeros(self.n_words, dtype=np.uint64)
            self._f("get_array_elem")(self.h, name.encode(), i, _ptr(out), len(out))
            res[i] = out
        return res

    def int_out(self, name: str):
        v = C.c_int64(0)
        ok = self._f("get_int")(self.h, name.encode(), C.byref(v))
        return v.value if ok else None

    def outputs(self, stream_names, int_names) -> dict:
        """Outputs as Python ints, shaped like base.py's parsed results."""
        res = {}
        for n in stream_names:
            s = self.stream(n)
            if s is not None:
                res[n] = words_to_int(s)
            a = self.array(n)
            if a is not None:
                res[n] = {i: words_to_int(w) for i, w in a.items()}
        for n in int_names:
            v = self.int_out(n)
            if v is not None:
                res[n] = v
        return res


class Oracle(_Model):
    prefix = "bszo_"
    so_path = ORACLE_SO
    SCALAR, FAITHFUL, BEST = 0, 1, 2

    @classmethod
    def _extra(cls, lib):
        lib.bszo_run.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t]
        lib.bszo_set_semantics.argtypes = [C.c_void_p, C.c_int]
        lib.bszo_output_streams.restype = C.c_char_p
        lib.bszo_output_streams.argtypes = [C.c_void_p]
        lib.bszo_output_ints.restype = C.c_char_p
        lib.bszo_output_ints.argtypes = [C.c_void_p]
        lib.bszo_static_ops.restype = C.c_int64
        lib.bszo_static_ops.argtypes = [C.c_void_p]
        lib.bszo_shl.argtypes = [u64p, u64p, C.c_uint64, C.c_uint64]
        lib.bszo_shr.argtypes = [u64p, u64p, C.c_uint64, C.c_uint64]
        lib.bszo_add.argtypes = [u64p, u64p, u64p, C.c_uint64, C.c_int]
        lib.bszo_popcount_bounded.restype = C.c_int64
        lib.bszo_popcount_bounded.argtypes = [u64p, C.c_uint64, C.c_int64]
        lib.bszo_is_nonzero_bounded.argtypes = [u64p, C.c_uint64, C.c_int64]
        lib.bszo_set_threads.argtypes = [C.c_int]
        lib.bszo_get_threads.restype = C.c_int
        lib.bszo_has_avx512.restype = C.c_int

    def _run(self, err, variant=0):
        return self.L.bszo_run(self.h, variant, err, 4096)

    def set_semantics(self, sem: int):
        self.L.bszo_set_semantics(self.h, sem)
        return self

    def output_names(self):
        s = self.L.bszo_output_streams(self.h).decode().split("\n")
        i = self.L.bszo_output_ints(self.h).decode().split("\n")
        return [x for x in s if x], [x for x in i if x]

    def static_ops(self) -> int:
        return self.L.bszo_static_ops(self.h)


class RefInterp(_Model):
    prefix = "bsref_"
    so_path = REF_SO
    SCALAR, SIMD, SIMD_OMP = 0, 1, 2

    @classmethod
    def available(cls) -> bool:
        try:
            with open("/proc/cpuinfo") as f:
                flags = f.read()
            if "avx512_vpopcntdq" not in flags and "avx512vpopcntdq" not in flags:
                return False
        except OSError:
            return False
        return super().available()

    @classmethod
    def _extra(cls, lib):
        lib.bsref_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_size_t]

    def _run(self, err, variant=2, add_mode=0, reuse_mem=1):
        return self.L.bsref_run(self.h, variant, add_mode, reuse_mem, err, 4096)

    def static_ops(self) -> int:
        """The reference's count_ops_ast under the loaded params (no run)."""
        self.L.bsref_static_ops.restype = C.c_int64
        self.L.bsref_static_ops.argtypes = [C.c_void_p]
        return self.L.bsref_static_ops(self.h)


def oracle_words_op(fn: str, *arrays, k: int = 0, variant: int = 0):
    """Call one of the oracle's exported word kernels."""
    lib = Oracle.lib()
    if fn in ("shl", "shr"):
        src = np.ascontiguousarray(arrays[0], dtype=np.uint64)
        dst = np.zeros_like(src)
        getattr(lib, "bszo_" + fn)(_ptr(dst), _ptr(src), k, len(src))
        return dst
    if fn == "add":
        a = np.ascontiguousarray(arrays[0], dtype=np.uint64)
        b = np.ascontiguousarray(arrays[1], dtype=np.uint64)
        d = np.zeros_like(a)
        lib.bszo_add(_ptr(d), _ptr(a), _ptr(b), len(a), variant)
        return d
    raise ValueError(fn)
