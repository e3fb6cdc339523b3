"""ctypes signatures of the render, job, serving and denoise entry points."""
import ctypes as C


def declare(L):
    u32, i32, u64, vp, sz = C.c_uint32, C.c_int32, C.c_uint64, C.c_void_p, C.c_size_t
    P = C.POINTER
    sig = {}
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
