"""ctypes binding of the C-ABI library (include/hypertrace_hip.h).

There is no fallback: if the library is missing or a call fails, an exception
is raised.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libhypertrace_hip.so")

# Every exported entry point and its (restype, argtypes).
SIGNATURES = {
    "ht_last_error": (c_char_p, []),
    "ht_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "ht_ctx_destroy": (None, [c_void_p]),
    "ht_ctx_sync": (c_int, [c_void_p]),
    "ht_prog_build": (c_int, [c_void_p, c_char_p, c_char_p, c_int, POINTER(c_char_p), POINTER(c_char_p), c_int,
                              c_int, POINTER(c_void_p)]),
    "ht_prog_build_ex": (c_int, [c_void_p, c_char_p, c_char_p, c_int, POINTER(c_char_p), POINTER(c_char_p), c_int,
                                 c_int, c_char_p, POINTER(c_void_p)]),
    "ht_prog_destroy": (None, [c_void_p]),
    "ht_prog_cache_hit": (c_int, []),
    "ht_prog_kernel_info": (c_int, [c_void_p, c_char_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "ht_prog_precompile": (c_int, [c_char_p, c_char_p, c_int, POINTER(c_char_p), POINTER(c_char_p), c_int, c_int,
                                   c_char_p]),
    "ht_buf_alloc": (c_int, [c_void_p, c_size_t, POINTER(c_void_p)]),
    "ht_buf_free": (c_int, [c_void_p]),
    "ht_buf_size": (c_size_t, [c_void_p]),
    "ht_buf_write": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ht_buf_read": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ht_buf_zero": (c_int, [c_void_p]),
    "ht_render": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p]),
    "ht_render_passes": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_uint32]),
    "ht_render_wavefront": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_uint32]),
    "ht_render_wavefront_noalloc": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_uint32]),
    "ht_extract": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, c_void_p, c_void_p]),
    "ht_gamma": (c_int, [c_void_p, c_uint32, c_uint32, c_float, c_void_p, c_void_p]),
    "ht_pack": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p]),
    "ht_post": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, c_float, c_void_p, c_void_p]),
    "ht_mo_compose": (c_int, [c_void_p, c_void_p, c_int, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32,
                              c_void_p]),
    "ht_cache_create": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, POINTER(c_void_p)]),
    "ht_cache_destroy": (None, [c_void_p]),
    "ht_cache_set_params": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ht_cache_get_params": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ht_cache_get_grads": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ht_cache_counts": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
    "ht_mlp_forward": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p]),
    "ht_mlp_train_step": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_float, c_int, POINTER(c_float)]),
    "ht_render_cached": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_uint32,
                                 c_uint32, c_uint32]),
    "ht_cache_train": (c_int, [c_void_p, c_uint32, c_uint32, c_float]),
    "ht_batch_create": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, POINTER(c_void_p)]),
    "ht_batch_destroy": (None, [c_void_p]),
    "ht_batch_tiles": (c_uint32, [c_void_p]),
    "ht_batch_block": (c_uint32, [c_void_p]),
    "ht_batch_render": (c_int, [c_void_p, c_uint32, c_int, c_float, c_int]),
    "ht_batch_capture": (c_int, [c_void_p, c_uint32, c_int, c_float, c_int]),
    "ht_batch_replay": (c_int, [c_void_p]),
    "ht_seed_init": (c_int, [c_void_p, c_uint32, c_uint32, c_int, c_uint64]),
    "ht_seed_fill": (c_int, [POINTER(c_uint32), c_size_t, c_int, c_uint64]),
    "ht_seed_hash": (c_uint32, [c_uint64, c_uint32]),
    "ht_timer_begin": (c_int, [c_void_p]),
    "ht_timer_end": (c_int, [c_void_p, POINTER(c_float)]),
}


class HtError(RuntimeError):
    """A failed C-ABI call (processing::Error::{Io, Backend, Other})."""

    KINDS = {-1: "Io", -2: "Backend", -3: "Other"}

    def __init__(self, code: int, message: str):
        super().__init__(f"{self.KINDS.get(code, code)}: {message}")
        self.code = code


_lib = None


def lib():
    """Load the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise HtError(rc, lib().ht_last_error().decode(errors="replace"))


def cstr_array(items):
    arr = (c_char_p * max(1, len(items)))()
    for i, s in enumerate(items):
        arr[i] = s.encode() if isinstance(s, str) else s
    return arr
