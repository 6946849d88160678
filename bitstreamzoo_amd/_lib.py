"""ctypes binding of include/bsz_hip.h (libbsz_hip.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbsz_hip.so")
CLI_PATH = os.path.join(PKG, "bsim_hip")
HEADER = os.path.join(os.path.dirname(PKG), "include", "bsz_hip.h")

u64p = C.POINTER(C.c_uint64)
i64p = C.POINTER(C.c_int64)
f32p = C.POINTER(C.c_float)
vp = C.c_void_p

# name -> (restype, argtypes)
_SIGS = {
    "bsz_ctx_create": (C.c_int, [C.POINTER(vp), C.c_int]),
    "bsz_ctx_destroy": (C.c_int, [vp]),
    "bsz_last_error": (C.c_char_p, [vp]),
    "bsz_sync": (C.c_int, [vp]),
    "bsz_ctx_stream": (vp, [vp]),
    "bsz_stream_alloc_words": (C.c_uint64, [C.c_int64]),
    "bsz_buf_alloc": (C.c_int, [vp, C.c_int64, C.POINTER(u64p)]),
    "bsz_buf_free": (C.c_int, [vp, u64p]),
    "bsz_h2d": (C.c_int, [vp, u64p, u64p, C.c_uint64, C.c_int64]),
    "bsz_d2h": (C.c_int, [vp, u64p, u64p, C.c_uint64]),
    "bsz_op_fill": (C.c_int, [vp, u64p, C.c_int, C.c_int64]),
    "bsz_op_copy": (C.c_int, [vp, u64p, u64p, C.c_int64]),
    "bsz_op_not": (C.c_int, [vp, u64p, u64p, C.c_int64]),
    "bsz_op_and": (C.c_int, [vp, u64p, u64p, u64p, C.c_int64]),
    "bsz_op_or": (C.c_int, [vp, u64p, u64p, u64p, C.c_int64]),
    "bsz_op_xor": (C.c_int, [vp, u64p, u64p, u64p, C.c_int64]),
    "bsz_op_add": (C.c_int, [vp, u64p, u64p, u64p, C.c_int64]),
    "bsz_op_add_noalloc": (C.c_int, [vp, u64p, u64p, u64p, C.c_int64]),
    "bsz_op_shl": (C.c_int, [vp, u64p, u64p, C.c_uint64, C.c_int64]),
    "bsz_op_shr": (C.c_int, [vp, u64p, u64p, C.c_uint64, C.c_int64]),
    "bsz_op_popcount": (C.c_int, [vp, u64p, C.c_int64, i64p]),
    "bsz_op_is_nonzero": (C.c_int, [vp, u64p, C.c_int64, C.POINTER(C.c_int)]),
    "bsz_program_parse": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
    "bsz_program_free": (C.c_int, [vp]),
    "bsz_program_static_ops": (C.c_int64, [vp, vp]),
    "bsz_case_create": (C.c_int, [C.POINTER(vp), C.c_int64]),
    "bsz_case_free": (C.c_int, [vp]),
    "bsz_case_set_param": (C.c_int, [vp, C.c_char_p, C.c_int64]),
    "bsz_case_set_input": (C.c_int, [vp, C.c_char_p, u64p, C.c_uint64]),
    "bsz_case_set_array_elem": (C.c_int, [vp, C.c_char_p, C.c_int64, u64p, C.c_uint64]),
    "bsz_case_set_broadcast": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_int64]),
    "bsz_case_set_iterations": (C.c_int, [vp, C.c_int]),
    "bsz_case_set_padding": (C.c_int, [vp, C.c_int]),
    "bsz_case_set_feedback": (C.c_int, [vp, C.c_char_p, C.c_char_p]),
    "bsz_case_read_bsdata": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(vp)), C.POINTER(C.c_int)]),
    "bsz_case_array_free": (C.c_int, [C.POINTER(vp), C.c_int]),
    "bsz_run": (C.c_int, [vp, vp, vp, C.c_int, C.POINTER(vp)]),
    "bsz_result_free": (C.c_int, [vp]),
    "bsz_result_stream": (C.c_int, [vp, C.c_char_p, u64p, C.c_uint64]),
    "bsz_result_array_count": (C.c_int64, [vp, C.c_char_p]),
    "bsz_result_array_indices": (C.c_int, [vp, C.c_char_p, i64p, C.c_int64]),
    "bsz_result_array_elem": (C.c_int, [vp, C.c_char_p, C.c_int64, u64p, C.c_uint64]),
    "bsz_result_int": (C.c_int, [vp, C.c_char_p, i64p]),
    "bsz_result_op_count": (C.c_int64, [vp]),
    "bsz_result_n_words": (C.c_uint64, [vp]),
    "bsz_result_json": (C.c_int64, [vp, C.c_char_p, C.c_uint64]),
    "bsz_plan_create": (C.c_int, [vp, vp, vp, C.c_char_p, C.POINTER(vp)]),
    "bsz_plan_compile_only": (C.c_int64, [vp, vp, C.c_char_p, C.c_char_p, C.c_uint64]),
    "bsz_plan_free": (C.c_int, [vp]),
    "bsz_plan_describe": (C.c_int64, [vp, vp, C.c_char_p, C.c_char_p, C.c_uint64]),
    "bsz_plan_set_instances": (C.c_int, [vp, C.c_int, i64p]),
    "bsz_plan_bind_input": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int64, u64p]),
    "bsz_plan_output": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int64, C.POINTER(u64p)]),
    "bsz_plan_run": (C.c_int, [vp]),
    "bsz_plan_capture": (C.c_int, [vp]),
    "bsz_plan_replay": (C.c_int, [vp, C.c_int]),
    "bsz_plan_read_int": (C.c_int, [vp, C.c_int, C.c_char_p, i64p]),
    "bsz_plan_mismatches": (C.c_int, [vp, C.c_int, i64p]),
    "bsz_plan_unproven": (C.c_int, [vp, C.c_int, i64p]),
    "bsz_plan_info": (C.c_int64, [vp, C.c_char_p, C.c_uint64]),
    "bsz_plan_source": (C.c_int64, [vp, C.c_char_p, C.c_uint64]),
    "bsz_plan_set_executor": (C.c_int, [vp, C.c_int]),
    "bsz_plan_prepare": (C.c_int, [vp]),
    "bsz_timer_start": (C.c_int, [vp]),
    "bsz_timer_stop": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "bsz_capture_begin": (C.c_int, [vp]),
    "bsz_capture_end": (C.c_int, [vp, C.POINTER(vp)]),
    "bsz_stream_order": (C.c_int, [vp, vp]),
    "bsz_graph_launch": (C.c_int, [vp, vp, C.c_int]),
    "bsz_graph_destroy": (C.c_int, [vp]),
    "bsz_mlp_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]),
    "bsz_mlp_free": (C.c_int, [vp]),
    "bsz_mlp_n_params": (C.c_int64, [vp]),
    "bsz_mlp_get_params": (C.c_int, [vp, f32p]),
    "bsz_mlp_set_params": (C.c_int, [vp, f32p]),
    "bsz_mlp_predict": (C.c_int, [vp, f32p, C.c_int, f32p]),
    "bsz_mlp_gradients": (C.c_int, [vp, f32p, f32p, C.c_int, f32p, C.POINTER(C.c_float)]),
    "bsz_mlp_train_step": (C.c_int, [vp, f32p, f32p, C.c_int, C.c_float, C.POINTER(C.c_float)]),
    "bsz_mlp_train": (C.c_int, [vp, f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64,
                                C.POINTER(C.c_float)]),
}

_lib = None


def header_symbols() -> list[str]:
    """Every function declared in include/bsz_hip.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(bsz_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m bitstreamzoo_amd.build` "
                              "(the HIP executor has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class BszError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg

    def __reduce__(self):   # picklable, so worker processes can return it
        return (BszError, (self.code, self.msg))


def check(rc: int, ctx=None):
    if rc != 0:
        msg = lib().bsz_last_error(ctx)
        raise BszError(rc, msg.decode() if msg else "")


def as_u64(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def ptr(a: np.ndarray):
    return a.ctypes.data_as(u64p)
