"""C ABI surface: libmtx.so loads and exports every symbol include/mtx.h
declares; calls that need no GPU behave as documented."""
import ctypes as C
import os

import pytest

from mutracer_amd import abi


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = abi.header_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version():
    assert abi.lib().mtx_abi_version() == 1


def test_invalid_handles_are_reported():
    L = abi.lib()
    assert L.mtx_blas_destroy(123456789) == abi.MTX_INVALID_HANDLE
    assert L.mtx_accel_destroy(987654321) == abi.MTX_INVALID_HANDLE
    info = (C.c_uint32 * 8)()
    assert L.mtx_blas_info(42424242, info) == abi.MTX_INVALID_HANDLE


def test_bad_arguments_rejected_before_any_gpu_work():
    L = abi.lib()
    err = abi.errbuf()
    out = C.c_uint64(0)
    assert L.mtx_blas_build(None, 0, None, 5, 8, 0, C.byref(out), err, 512) == abi.MTX_INVALID_ARGUMENT
    assert L.mtx_blas_build(None, 0, None, 0, 99, 0, C.byref(out), err, 512) == abi.MTX_INVALID_ARGUMENT
    assert b"bins" in err.value
