"""Batched render server host (continuous batching over libmtx rounds).

The reference's native job API runs one job per context
(mt_context_submit / mt_job_poll / mt_job_info / mt_job_cancel,
native/include/mutracer-context.h:146-187). `Server` keeps many jobs of
different scenes, resolutions and sample counts in flight: every
`step()` is one GPU round in which each running job renders a slice of its
samples inside one shared wavefront generation. `max_jobs_per_round=1`
gives the plain one-job-at-a-time form on the same code path.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import abi
from .abi_ext import JobInfo, TargetSettings, ServerConfig
from .render import KERNELS, PIXELS, camera_desc
from .sampling import sampling_pattern_code

_P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
JOB_STATES = {0: "queued", 1: "running", 2: "complete", 3: "cancelled", 4: "failed"}


class Server:
    def __init__(self, max_paths_per_round=1 << 24, max_jobs_per_round=0, max_samples_per_job_round=64,
                 hbm_budget_bytes=200 << 30):
        cfg = ServerConfig(C.sizeof(ServerConfig), max_paths_per_round, max_jobs_per_round,
                           max_samples_per_job_round, hbm_budget_bytes)
        self.handle = C.c_uint64(0)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_server_create(C.byref(cfg), C.byref(self.handle), err, 512), err)
        self.sizes = {}

    def submit(self, render_scene, width, height, samples, bounces=8, direct_light_candidates=1, pattern="sobol",
               sequence_seed=0, max_samples=None, sample_sequence_offset=0, clamping=0.0, camera=None):
        s = TargetSettings()
        s.struct_size = C.sizeof(TargetSettings)
        s.width, s.height, s.bounces = width, height, bounces
        s.direct_light_candidates = direct_light_candidates
        s.sampling_pattern = sampling_pattern_code(pattern) if isinstance(pattern, str) else int(pattern)
        s.sequence_seed = sequence_seed & 0xFFFFFFFF
        s.max_samples = samples if max_samples is None else max_samples
        s.sample_sequence_offset = sample_sequence_offset
        s.clamping = clamping
        s.kernel = KERNELS["wavefront"]
        cam = camera_desc(camera or render_scene.scene.camera, width, height)
        job = C.c_uint64(0)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_server_submit(self.handle.value, render_scene.handle.value, C.byref(s),
                                              C.byref(cam), samples, C.byref(job), err, 512), err)
        self.sizes[job.value] = (width, height)
        return job.value

    def step(self, stream=None):
        """One round; returns (jobs completed, idle)."""
        done, idle = C.c_uint32(0), C.c_uint32(0)
        sp = None if stream is None else C.c_void_p(int(stream.cuda_stream))
        abi.check(abi.lib().mtx_server_step(self.handle.value, sp, C.byref(done), C.byref(idle)))
        return done.value, bool(idle.value)

    def info(self, job):
        i = JobInfo()
        abi.check(abi.lib().mtx_server_job_info(self.handle.value, job, C.byref(i)))
        return dict(state=JOB_STATES.get(i.state, i.state), requested=i.requested_samples,
                    completed=i.completed_samples, submit_ms=i.submit_ms, start_ms=i.start_ms,
                    complete_ms=i.complete_ms)

    def read_pixels(self, job, kind="color"):
        w, h = self.sizes[job]
        out = np.zeros((h, w, 4), np.float32)
        written = C.c_size_t(0)
        abi.check(abi.lib().mtx_server_job_read_pixels(self.handle.value, job, PIXELS[kind], _P(out), out.size,
                                                       C.byref(written)))
        return out

    def read_accumulators(self, job):
        w, h = self.sizes[job]
        sums = np.zeros((h, w, 13), np.float64)
        counts = np.zeros((h, w), np.uint32)
        abi.check(abi.lib().mtx_server_job_read_accumulators(self.handle.value, job, _P(sums), sums.size,
                                                             _P(counts), counts.size))
        return sums, counts

    def cancel(self, job):
        abi.check(abi.lib().mtx_server_job_cancel(self.handle.value, job))

    def release(self, job):
        abi.check(abi.lib().mtx_server_job_release(self.handle.value, job))
        self.sizes.pop(job, None)

    def stats(self):
        out = np.zeros(5, np.uint64)
        abi.check(abi.lib().mtx_server_stats(self.handle.value, _P(out)))
        g = np.zeros(2, np.uint64)
        abi.check(abi.lib().mtx_server_graph_stats(self.handle.value, _P(g)))
        return dict(rounds=int(out[0]), rays=int(out[1]), running=int(out[2]), queued=int(out[3]),
                    admitted_bytes=int(out[4]), graph_replays=int(g[0]), graph_captures=int(g[1]))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and abi._lib is not None:
            abi.lib().mtx_server_destroy(h.value)


def run_stream(server, jobs, arrivals, release=True):
    """Feeds `jobs` (kwargs of Server.submit incl. render_scene) at wall-clock
    offsets `arrivals` (seconds, ascending) and drives rounds until all are
    done. Returns per-job latency (ms, arrival -> completion) and the makespan."""
    t0 = time.perf_counter()
    handles, arrive_ms = [None] * len(jobs), [0.0] * len(jobs)
    nxt, done = 0, 0
    pending = set()
    while done < len(jobs):
        now = time.perf_counter() - t0
        while nxt < len(jobs) and arrivals[nxt] <= now:
            handles[nxt] = server.submit(**jobs[nxt])
            arrive_ms[nxt] = server.info(handles[nxt])["submit_ms"]
            pending.add(nxt)
            nxt += 1
        _, idle = server.step()
        finished = [i for i in pending if server.info(handles[i])["state"] in ("complete", "cancelled", "failed")]
        for i in finished:
            pending.discard(i)
            done += 1
        if idle and nxt < len(jobs) and not pending:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
    makespan = time.perf_counter() - t0
    lat = []
    for i, h in enumerate(handles):
        inf = server.info(h)
        lat.append(inf["complete_ms"] - arrive_ms[i])
        if release:
            server.release(h)
    return np.array(lat), makespan, handles
