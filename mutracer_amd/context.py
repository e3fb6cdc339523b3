"""Persistent render contexts over raw scene descriptors (ctypes host).

Mirrors the reference's native context API (native/include/mutracer-context.h:
mt_context_create/commit/render/reset/wait/info/read_pixels and the one-job-
per-context submit/poll/info/cancel contract) on libmtx's mtx_context_* and
mtx_job_* entry points. Descriptors come from `descriptor.encode` or any host
that writes the v5..v8 format.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi_ext import ContextSnapshot, JobSnapshot, RenderSettings

IMAGE = dict(hdr=0, normal=1, albedo=2)
CONTEXT_STATES = {0: "empty", 1: "committed", 2: "failed", 3: "job_active"}
COMMIT_KINDS = {0: "none", 1: "full", 2: "incremental"}
JOB_STATES = {0: "queued", 1: "running", 2: "complete", 3: "cancelled", 4: "failed", 5: "cancelling"}
WAIT_INDEFINITELY = 0xFFFFFFFF


def _buf(data):
    b = np.frombuffer(bytes(data), np.uint8)
    return b, b.ctypes.data_as(C.c_void_p)


def settings_default(**overrides):
    """mt_render_settings_default, then the given fields."""
    s = RenderSettings()
    abi.check(abi.lib().mtx_render_settings_default(C.byref(s)))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def validate(data):
    """Validates and decodes a descriptor without a device; returns its summary."""
    b, p = _buf(data)
    info = np.zeros(8, np.uint32)
    err = abi.errbuf()
    abi.check(abi.lib().mtx_descriptor_validate(p, b.size, info.ctypes.data_as(C.c_void_p), err, 512), err)
    keys = ("width", "height", "blas_meshes", "instances", "materials", "lights", "env_width", "env_height")
    return dict(zip(keys, (int(v) for v in info)))


class Job:
    def __init__(self, handle):
        self.handle = handle

    def poll(self, max_new_samples=64):
        err = abi.errbuf()
        abi.check(abi.lib().mtx_job_poll(self.handle, max_new_samples, err, 512), err)

    def info(self):
        s = JobSnapshot()
        s.struct_size = C.sizeof(JobSnapshot)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_job_info(self.handle, C.byref(s), err, 512), err)
        return dict(state=JOB_STATES.get(s.state, s.state), requested=s.requested_samples,
                    submitted=s.submitted_samples, completed=s.completed_samples, base=s.base_samples)

    def cancel(self):
        err = abi.errbuf()
        abi.check(abi.lib().mtx_job_cancel(self.handle, err, 512), err)


class RenderContext:
    def __init__(self):
        self.handle = C.c_uint64(0)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_context_create(C.byref(self.handle), err, 512), err)

    def _call(self, name, *args):
        err = abi.errbuf()
        abi.check(getattr(abi.lib(), name)(self.handle.value, *args, err, 512), err)

    def commit(self, data, settings=None):
        b, p = _buf(data)
        if settings is None:
            self._call("mtx_context_commit", p, b.size)
        else:
            self._call("mtx_context_commit_with_settings", p, b.size, C.byref(settings))

    def render(self, additional_samples):
        self._call("mtx_context_render", additional_samples)

    def reset(self):
        self._call("mtx_context_reset")

    def wait(self, timeout_ms=WAIT_INDEFINITELY):
        self._call("mtx_context_wait", timeout_ms)

    def info(self):
        s = ContextSnapshot()
        s.struct_size = C.sizeof(ContextSnapshot)
        self._call("mtx_context_info", C.byref(s))
        return dict(width=s.width, height=s.height, samples=s.samples, state=CONTEXT_STATES.get(s.state, s.state),
                    last_commit=COMMIT_KINDS.get(s.last_commit, s.last_commit))

    def read_pixels(self, kind="hdr"):
        """RGBA f32 (height, width, 4), bottom row first, sums / completed samples."""
        written = C.c_size_t(0)
        self._call("mtx_context_read_pixels", IMAGE[kind], None, 0, C.byref(written))
        info = self.info()
        out = np.zeros((info["height"], info["width"], 4), np.float32)
        assert out.nbytes == written.value
        self._call("mtx_context_read_pixels", IMAGE[kind], out.ctypes.data_as(C.c_void_p), out.nbytes,
                   C.byref(written))
        return out

    def submit(self, additional_samples):
        job = C.c_uint64(0)
        self._call("mtx_context_submit", additional_samples, C.byref(job))
        return Job(job.value)

    def inject_failure(self):
        abi.check(abi.lib().mtx_context_inject_failure(self.handle.value))

    def destroy(self):
        if self.handle.value:
            self._call("mtx_context_destroy")
            self.handle = C.c_uint64(0)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and abi._lib is not None:
            abi.lib().mtx_context_destroy(h.value, None, 0)
