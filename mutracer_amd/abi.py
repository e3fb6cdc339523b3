"""ctypes binding of include/mtx.h (the library's C ABI).

This module is the Python host's view of the same extern "C" surface the
N-API addon and native hosts call. It fails loudly when libmtx.so is absent:
there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmtx.so")
HEADER = os.path.join(HERE, "..", "include", "mtx.h")

MTX_OK = 0
MTX_ERROR = 1
MTX_INVALID_ARGUMENT = 2
MTX_INVALID_HANDLE = 3
MTX_BUSY = 4
MTX_INVALID_STATE = 5
MTX_CANCELLED = 7
MTX_BUFFER_TOO_SMALL = 8
MTX_CONTEXT_FAILED = 9
MTX_UNAVAILABLE = 10
MTX_TIMEOUT = 11

MTX_INPUT_DEVICE = 1
MTX_ACCEL_PER_INSTANCE_BOUNDS = 1
MTX_TRACE_DOUBLE_SIDED = 1
MTX_TRACE_COUNT = 2

_STATUS_NAMES = {v: k for k, v in globals().items() if k.startswith("MTX_") and isinstance(v, int)
                 and k not in ("MTX_INPUT_DEVICE", "MTX_ACCEL_PER_INSTANCE_BOUNDS", "MTX_TRACE_DOUBLE_SIDED",
                               "MTX_TRACE_COUNT")}


class MtxError(RuntimeError):
    def __init__(self, code, message=""):
        self.code = code
        super().__init__(f"{_STATUS_NAMES.get(code, code)}: {message}")


class InstanceDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("blas_index", C.c_uint32), ("material_word", C.c_uint32),
                ("double_sided", C.c_uint32), ("transform", C.c_double * 16), ("color", C.c_float * 3),
                ("reserved", C.c_float)]


_lib = None


def lib():
    """Loads libmtx.so. Raises if it was not built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MtxError(MTX_UNAVAILABLE, f"{LIB_PATH} is missing; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def header_symbols():
    """Function names declared in include/mtx.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:u?int\d+_t|void)\s+\*?(mtx_\w+)\s*\(", text, re.M)))


def _declare(L):
    u32, i32, u64, vp, sz = C.c_uint32, C.c_int32, C.c_uint64, C.c_void_p, C.c_size_t
    P = C.POINTER
    sig = {
        "mtx_abi_version": (u32, []),
        "mtx_device_info": (i32, [C.c_char_p, sz, P(i32)]),
        "mtx_blas_build": (i32, [vp, u32, vp, u32, u32, u32, P(u64), C.c_char_p, sz]),
        "mtx_blas_info": (i32, [u64, vp]),
        "mtx_blas_export": (i32, [u64, vp, vp, sz, P(sz), vp]),
        "mtx_blas_refit": (i32, [u64, vp, u32, vp, u32, u32, P(u32), C.c_char_p, sz]),
        "mtx_blas_destroy": (i32, [u64]),
        "mtx_accel_create": (i32, [vp, u32, vp, u32, u32, u32, P(u64), C.c_char_p, sz]),
        "mtx_accel_refit": (i32, [u64, vp, u32, vp, u32, C.c_char_p, sz]),
        "mtx_accel_export_tlas": (i32, [u64, vp, sz, P(sz), vp, sz, P(sz)]),
        "mtx_accel_destroy": (i32, [u64]),
        "mtx_trace_closest": (i32, [u64, vp, u32, vp, u32, vp]),
        "mtx_trace_closest_noalloc": (i32, [u64, vp, u32, vp, u32, vp]),
        "mtx_trace_any": (i32, [u64, vp, u32, vp, vp]),
        "mtx_trace_count": (i32, [u64, vp, u32, vp, u32, vp]),
    }
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
    # Later additions are declared by the modules that use them.
    from . import abi_ext
    abi_ext.declare(L)


def check(code, err=None):
    if code != MTX_OK:
        msg = err.value.decode("utf-8", "replace") if err is not None else ""
        raise MtxError(code, msg)


def errbuf():
    return C.create_string_buffer(512)
