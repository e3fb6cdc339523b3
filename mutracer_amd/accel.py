"""Resident acceleration structures (GPU-built BLAS + TLAS) and ray queries.

Mirrors the reference's BLAS/TLAS preparation (src/renderer/bvhAccel.ts:
buildBLAS :200, buildWebTLAS :452, refitTLAS :579) and the raw first-hit
query of native diagnostics (querySceneSoftware, intersect.wgsl.ts:660-684),
executed by libmtx on the GPU. Device buffers are torch CUDA(HIP) tensors;
torch is only the allocator and stream provider here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

_P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


def _dev_ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        return None
    return C.c_void_p(int(stream.cuda_stream))


class Blas:
    """One triangle mesh's BVH2, built by the GPU binned-SAH builder."""

    def __init__(self, positions, indices, bins=8):
        L = abi.lib()
        self.handle = C.c_uint64(0)
        err = abi.errbuf()
        try:
            import torch
            is_dev = isinstance(positions, torch.Tensor) and positions.is_cuda
        except ImportError:  # pragma: no cover
            is_dev = False
        if is_dev:
            pos, idx = positions.contiguous(), indices.contiguous()
            self._keep = (pos, idx)
            code = L.mtx_blas_build(_dev_ptr(pos), pos.numel() // 3, _dev_ptr(idx), idx.numel() // 3, bins,
                                    abi.MTX_INPUT_DEVICE, C.byref(self.handle), err, 512)
            self._indices_host = None  # copied from the device tensor when refit or export needs it
        else:
            pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
            idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
            code = L.mtx_blas_build(_P(pos), pos.shape[0], _P(idx), idx.size // 3, bins, 0,
                                    C.byref(self.handle), err, 512)
            self._indices_host = idx
        abi.check(code, err)

    @property
    def indices_host(self):
        """The build's index buffer on the host (uint32, flat)."""
        if self._indices_host is None:
            self._indices_host = self._keep[1].cpu().numpy().astype(np.uint32).reshape(-1)
        return self._indices_host

    def refit(self, positions):
        """refitBLAS: new vertex positions (host array, same vertex count) with
        the build's index buffer. Returns True when a rebuild is advised."""
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        advised = C.c_uint32(0)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_blas_refit(self.handle.value, _P(pos), pos.shape[0], _P(self.indices_host),
                                           self.indices_host.size // 3, 0, C.byref(advised), err, 512), err)
        return bool(advised.value)

    def info(self):
        out = np.zeros(8, np.uint32)
        abi.check(abi.lib().mtx_blas_info(self.handle.value, _P(out)))
        return dict(node_count=int(out[0]), vec4_count=int(out[1]), max_depth=int(out[2]), max_leaf=int(out[3]),
                    prim_count=int(out[4]), resident_bytes=int(out[5]) | (int(out[6]) << 32),
                    interior_nodes=int(out[7]))

    def export(self):
        """Reference BVH2 vec4 words and the source triangle of each record."""
        L = abi.lib()
        n = C.c_size_t(0)
        abi.check(L.mtx_blas_export(self.handle.value, None, None, 0, C.byref(n), None))
        words = np.zeros(n.value, np.uint32)
        src = np.zeros(n.value // 4, np.uint32)
        abi.check(L.mtx_blas_export(self.handle.value, _P(self.indices_host), _P(words), words.size, C.byref(n),
                                    _P(src)))
        return words, src

    def __del__(self):
        try:
            if self.handle.value:
                abi.lib().mtx_blas_destroy(self.handle.value)
        except Exception:
            pass


def instance_descs(blas_index, transforms, material=None, double_sided=None, color=None):
    n = len(blas_index)
    arr = (abi.InstanceDesc * max(1, n))()
    tr = np.asarray(transforms, np.float64).reshape(n, 16)
    for i in range(n):
        d = arr[i]
        d.struct_size = C.sizeof(abi.InstanceDesc)
        d.blas_index = int(blas_index[i])
        d.material_word = (int(material[i]) << 8) if material is not None else 0
        d.double_sided = int(double_sided[i]) if double_sided is not None else 0
        for k in range(16):
            d.transform[k] = float(tr[i, k])
        c = color[i] if color is not None else (1.0, 1.0, 1.0)
        for k in range(3):
            d.color[k] = float(c[k])
    return arr, n


class Accel:
    """Scene acceleration: TLAS over instances of resident BLASes."""

    def __init__(self, blases, blas_index, transforms, material=None, double_sided=None, color=None, tlas_bins=64,
                 per_instance_bounds=True):
        L = abi.lib()
        self.blases = list(blases)
        hs = (C.c_uint64 * max(1, len(self.blases)))(*[b.handle.value for b in self.blases])
        self._descs, self.count = instance_descs(blas_index, transforms, material, double_sided, color)
        self._args = (blas_index, material, double_sided, color)
        self.handle = C.c_uint64(0)
        err = abi.errbuf()
        flags = abi.MTX_ACCEL_PER_INSTANCE_BOUNDS if per_instance_bounds else 0
        abi.check(L.mtx_accel_create(hs, len(self.blases), self._descs, self.count, tlas_bins, flags,
                                     C.byref(self.handle), err, 512), err)

    def refit(self, transforms, dirty=None):
        blas_index, material, double_sided, color = self._args
        self._descs, n = instance_descs(blas_index, transforms, material, double_sided, color)
        d = None if dirty is None else np.ascontiguousarray(dirty, np.uint32)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_accel_refit(self.handle.value, self._descs, n, None if d is None else _P(d),
                                            0 if d is None else d.size, err, 512), err)

    def export_tlas(self):
        L = abi.lib()
        tn, inn = C.c_size_t(0), C.c_size_t(0)
        abi.check(L.mtx_accel_export_tlas(self.handle.value, None, 0, C.byref(tn), None, 0, C.byref(inn)))
        tw = np.zeros(tn.value, np.uint32)
        iw = np.zeros(inn.value, np.uint32)
        abi.check(L.mtx_accel_export_tlas(self.handle.value, _P(tw), tw.size, C.byref(tn), _P(iw), iw.size,
                                          C.byref(inn)))
        return tw, iw

    def trace_closest(self, rays, hits, double_sided=False, count=False, stream=None):
        """rays: (N, 8) float32 CUDA tensor; hits: (N, 8) int32/uint32 CUDA tensor."""
        flags = (abi.MTX_TRACE_DOUBLE_SIDED if double_sided else 0) | (abi.MTX_TRACE_COUNT if count else 0)
        abi.check(abi.lib().mtx_trace_closest(self.handle.value, _dev_ptr(rays), rays.shape[0], _dev_ptr(hits),
                                              flags, _stream_ptr(stream)))

    def trace_any(self, rays, occluded, stream=None):
        abi.check(abi.lib().mtx_trace_any(self.handle.value, _dev_ptr(rays), rays.shape[0], _dev_ptr(occluded),
                                          _stream_ptr(stream)))

    def trace_counters(self, rays, counters, double_sided=False, stream=None):
        flags = abi.MTX_TRACE_DOUBLE_SIDED if double_sided else 0
        abi.check(abi.lib().mtx_trace_count(self.handle.value, _dev_ptr(rays), rays.shape[0], _dev_ptr(counters),
                                            flags, _stream_ptr(stream)))

    def __del__(self):
        try:
            if self.handle.value:
                abi.lib().mtx_accel_destroy(self.handle.value)
        except Exception:
            pass
