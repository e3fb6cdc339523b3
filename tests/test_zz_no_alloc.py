"""mtx_trace_closest_noalloc: the allocation-free closest-hit query.

Called through the C ABI, it must return exactly what mtx_trace_closest
returns for the same rays (hit flag, instance, triangle, t, barycentrics and
the visit counters), on two calls in a row with the second smaller than the
first (the shared static scratch keeps the last call's contents, so the
counters it needs are cleared by a kernel on every call), and reject inputs
beyond its stated maximum.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from mutracer_amd import abi, accel, workloads
    return torch, abi, accel, workloads


def _query(torch, abi, fn, acc, rays, flags):
    r = torch.from_numpy(rays).cuda()
    h = torch.full((rays.shape[0], 8), -7, dtype=torch.int32, device="cuda")
    abi.check(fn(acc.handle.value, C.c_void_p(r.data_ptr()), rays.shape[0], C.c_void_p(h.data_ptr()), flags, None))
    torch.cuda.synchronize()
    return h.cpu().numpy()


def test_noalloc_matches_trace_closest(env):
    torch, abi, accel, W = env
    L = abi.lib()
    pos, idx = W.random_triangles(200000, seed=5)
    blas = accel.Blas(pos, idx, 32)
    acc = accel.Accel([blas], [0], [W.identity_transform()], material=[0], double_sided=[1])
    big = W.pinhole_rays(640)          # 409,600 rays: more than one pass of the fixed grid
    small = W.pinhole_rays(96)[::3]    # 3,072 rays
    for flags in (abi.MTX_TRACE_DOUBLE_SIDED, abi.MTX_TRACE_DOUBLE_SIDED | abi.MTX_TRACE_COUNT, 0):
        for rays in (big, small):
            print(f"noalloc flags={flags} rays={rays.shape[0]}", flush=True)
            ref = _query(torch, abi, L.mtx_trace_closest, acc, rays, flags)
            got = _query(torch, abi, L.mtx_trace_closest_noalloc, acc, rays, flags)
            assert np.array_equal(got, ref)
            assert (ref[:, 0] == 1).mean() > 0.05


def test_noalloc_rejects_bad_input(env):
    torch, abi, accel, W = env
    L = abi.lib()
    pos, idx = W.random_triangles(1000, seed=1)
    acc = accel.Accel([accel.Blas(pos, idx, 8)], [0], [W.identity_transform()], material=[0], double_sided=[1])
    assert L.mtx_trace_closest_noalloc(acc.handle.value, None, 5, None, 0, None) == abi.MTX_INVALID_ARGUMENT
    r = torch.zeros((1, 8), dtype=torch.float32, device="cuda")
    h = torch.zeros((1, 8), dtype=torch.int32, device="cuda")
    assert L.mtx_trace_closest_noalloc(acc.handle.value, C.c_void_p(r.data_ptr()), (1 << 30) + 1,
                                       C.c_void_p(h.data_ptr()), 0, None) == abi.MTX_INVALID_ARGUMENT
    assert L.mtx_trace_closest_noalloc(12345678, C.c_void_p(r.data_ptr()), 1, C.c_void_p(h.data_ptr()), 0,
                                       None) == abi.MTX_INVALID_HANDLE
    assert L.mtx_trace_closest_noalloc(acc.handle.value, None, 0, None, 0, None) == abi.MTX_OK
