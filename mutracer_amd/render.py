"""Path-tracer host: resident render scene, render targets and converge().

Mirrors the reference's operator surface for the render hot path:
  - RenderScene  ~ WGPURayTracePass.setMeshes / setLights (WGPURayTracePass.ts
    :2246, :3180) and the TLAS refit of Engine.update (bvhAccel.ts:579);
  - RenderTarget ~ RenderTarget.trace(camera, maxSamples) (RenderingPipeline.ts
    :218) with setSamplingPattern (:2073), setSampleSequenceOffset (:2085),
    setBounce (:2040), restartSamples and readPixels (Render.ts:782-801);
  - Render.converge(samples) ~ Render.converge (Render.ts:827-900): batches of
    at most `samples_per_batch` samples until the budget is reached.
Every call goes to libmtx (HIP); nothing here computes pixels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi_ext import (AdaptiveOptions, CameraDesc, DisplaySettings, JnlmOptions, MeshDesc, SceneDesc, TextureDesc,
                      TargetSettings, WavefrontTuning)
from .sampling import sampling_pattern_code

_P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731

PIXELS = dict(color=0, albedo=1, normal=2, variance=3)
KERNELS = dict(wavefront=0, megakernel=1)


class RenderScene:
    """A committed scene: per-mesh BLASes, the TLAS, lights and environment,
    resident on the current HIP device."""

    def __init__(self, scene, blas_bins=8, tlas_bins=64, per_instance_bounds=True):
        L = abi.lib()
        self.scene = scene
        self._keep = []
        meshes = (MeshDesc * max(1, len(scene.meshes)))()
        for i, m in enumerate(scene.meshes):
            pos = np.ascontiguousarray(m["positions"], np.float32)
            nrm = np.ascontiguousarray(m["normals"], np.float32)
            idx = np.ascontiguousarray(m["indices"], np.uint32)
            uv = [None if m.get(k) is None else np.ascontiguousarray(m[k], np.float32) for k in ("uvs", "uvs1")]
            self._keep += [pos, nrm, idx] + uv
            meshes[i] = MeshDesc(C.sizeof(MeshDesc), pos.size // 3, idx.size // 3, 0, _P(pos).value,
                                 _P(nrm).value, _P(idx).value, *[None if u is None else _P(u).value for u in uv])
        textures = (TextureDesc * max(1, len(scene.textures)))()
        for i, t in enumerate(scene.textures):
            px = t.storage_pixels()
            self._keep.append(px)
            textures[i] = TextureDesc(C.sizeof(TextureDesc), px.shape[1], px.shape[0], t.storage, int(t.flip_y),
                                      int(t.uv_channel), t.wrap_code(t.wrap_s), t.wrap_code(t.wrap_t),
                                      (C.c_float * 2)(*t.uv_scale), (C.c_float * 2)(*t.uv_offset), _P(px).value)
        n = len(scene.mesh_index)
        inst = (abi.InstanceDesc * max(1, n))()
        for i in range(n):
            d = inst[i]
            d.struct_size = C.sizeof(abi.InstanceDesc)
            d.blas_index = int(scene.mesh_index[i])
            d.material_word = int(scene.instance_material[i]) << 8
            d.double_sided = 0
            d.transform[:] = [float(v) for v in scene.transforms[i]]
            d.color[:] = [1.0, 1.0, 1.0]
        mats = np.ascontiguousarray(scene.packed_materials(), np.float32)
        lights = np.ascontiguousarray(scene.packed_lights(), np.float32)
        env = None if scene.env is None else np.ascontiguousarray(scene.env, np.float32)
        self._keep += [mats, lights, env, meshes, inst, textures]
        desc = SceneDesc(C.sizeof(SceneDesc), len(scene.meshes), C.cast(meshes, C.c_void_p).value, n,
                         C.cast(inst, C.c_void_p).value, mats.shape[0], _P(mats).value, lights.shape[0],
                         _P(lights).value if lights.size else None,
                         0 if env is None else env.shape[1], 0 if env is None else env.shape[0],
                         None if env is None else _P(env).value, blas_bins, tlas_bins,
                         abi.MTX_ACCEL_PER_INSTANCE_BOUNDS if per_instance_bounds else 0,
                         len(scene.textures), C.cast(textures, C.c_void_p).value if scene.textures else None)
        self.handle = C.c_uint64(0)
        err = abi.errbuf()
        abi.check(L.mtx_scene_create(C.byref(desc), C.byref(self.handle), err, 512), err)
        self._keep = None

    def info(self):
        out = np.zeros(4, np.uint64)
        abi.check(abi.lib().mtx_scene_info(self.handle.value, _P(out)))
        return dict(triangles=int(out[0]), instances=int(out[1]), blas=int(out[2]), resident_mib=int(out[3]))

    def set_transforms(self, transforms, dirty=None):
        """Transform-only edit: refits the TLAS over the dirty instances."""
        t = np.ascontiguousarray(np.asarray(transforms, np.float64).reshape(-1, 16))
        d = None if dirty is None else np.ascontiguousarray(dirty, np.uint32)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_scene_set_transforms(self.handle.value, _P(t), t.shape[0],
                                                     None if d is None else _P(d), 0 if d is None else d.size,
                                                     err, 512), err)
        self.scene.transforms = [row.copy() for row in t]

    def update_mesh(self, mesh, positions, normals):
        """Vertex deformation of one mesh (BLAS refit + TLAS refit); returns
        True when a rebuild is advised."""
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        advised = C.c_uint32(0)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_scene_update_mesh(self.handle.value, mesh, _P(pos), _P(nrm), pos.shape[0],
                                                  C.byref(advised), err, 512), err)
        self.scene.meshes[mesh] = dict(self.scene.meshes[mesh], positions=pos, normals=nrm)
        return bool(advised.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and abi._lib is not None:
            abi.lib().mtx_scene_destroy(h.value)


TONE_MAPPINGS = dict(agx=0, aces=1, cineon=2, reinhard=3, linear=4)


def display_settings(tone_mapping="agx", gamma22=False, **grading):
    """mtx_display_settings from the library defaults (src/settings/display.ts:12-18)."""
    d = DisplaySettings()
    abi.check(abi.lib().mtx_display_settings_default(C.byref(d)))
    d.tone_mapping = TONE_MAPPINGS[tone_mapping]
    d.gamma22 = int(bool(gamma22))
    for k, v in grading.items():
        setattr(d, k, v)
    return d


def camera_desc(camera, width, height):
    pi, vi = camera.matrices(width, height)
    d = CameraDesc()
    d.struct_size = C.sizeof(CameraDesc)
    d.projection_inverse[:] = [float(v) for v in pi]
    d.view_inverse[:] = [float(v) for v in vi]
    d.aperture_radius = camera.aperture
    d.focus_distance = camera.focus_distance
    return d


class RenderTarget:
    """FP64 progressive accumulation of one view of a RenderScene."""

    def __init__(self, render_scene, width, height, bounces=8, direct_light_candidates=1, pattern="sobol",
                 sequence_seed=0, max_samples=1, sample_sequence_offset=0, clamping=0.0, camera=None,
                 kernel="wavefront", max_paths_in_flight=0):
        self.scene = render_scene
        self.width, self.height = width, height
        s = TargetSettings()
        s.struct_size = C.sizeof(TargetSettings)
        s.width, s.height, s.bounces = width, height, bounces
        s.direct_light_candidates = direct_light_candidates
        s.sampling_pattern = sampling_pattern_code(pattern) if isinstance(pattern, str) else int(pattern)
        s.sequence_seed = sequence_seed & 0xFFFFFFFF
        s.max_samples = max_samples
        s.sample_sequence_offset = sample_sequence_offset
        s.clamping = clamping
        s.kernel = KERNELS[kernel]
        s.max_paths_in_flight = max_paths_in_flight
        self.settings = s
        cam = camera_desc(camera or render_scene.scene.camera, width, height)
        self.handle = C.c_uint64(0)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_target_create(render_scene.handle.value, C.byref(s), C.byref(cam),
                                              C.byref(self.handle), err, 512), err)

    def render(self, samples, stream=None):
        """Adds `samples` samples to every pixel (asynchronous on `stream`)."""
        sp = None if stream is None else C.c_void_p(int(stream.cuda_stream))
        abi.check(abi.lib().mtx_target_render(self.handle.value, samples, sp))

    WORK_KEYS = ("closest_rays", "box_tests", "tri_tests", "instance_visits", "hits", "shadow_rays",
                 "shadow_box_tests", "shadow_tri_tests", "shadow_instance_visits")
    TIMER_CLASSES = ("generate_accumulate", "trace", "shade", "shadow", "megakernel", "tail")

    def render_counted(self, samples, stream=None):
        """render() plus the launch's traversal work (synchronous): totals under
        WORK_KEYS, and the same counters per kernel group under "wavefront"
        (trace/shadow kernels or the megakernel) and "tail" (tail kernel)."""
        sp = None if stream is None else C.c_void_p(int(stream.cuda_stream))
        n = len(self.WORK_KEYS)
        out = np.zeros(2 * n, np.uint64)
        abi.check(abi.lib().mtx_target_render_counted(self.handle.value, samples, _P(out), sp))
        groups = {g: {k: int(v) for k, v in zip(self.WORK_KEYS, out[i * n:(i + 1) * n])}
                  for i, g in enumerate(("wavefront", "tail"))}
        res = {k: groups["wavefront"][k] + groups["tail"][k] for k in self.WORK_KEYS}
        res.update(groups)
        return res

    def render_timed(self, samples, stream=None):
        """render() with HIP events around every launch: GPU ms per kernel class (synchronous)."""
        sp = None if stream is None else C.c_void_p(int(stream.cuda_stream))
        out = np.zeros(len(self.TIMER_CLASSES), np.float64)
        abi.check(abi.lib().mtx_target_render_timed(self.handle.value, samples, _P(out), sp))
        return dict(zip(self.TIMER_CLASSES, map(float, out)))

    def trace(self, camera=None, max_samples=1):
        """RenderTarget.trace: one more sample unless `max_samples` are done.
        Returns True while samples remain."""
        if camera is not None:
            self.set_camera(camera)
        done = self.stats()["samples"]
        if done >= max_samples:
            return False
        self.render(1)
        return done + 1 < max_samples

    def tuning(self):
        t = WavefrontTuning()
        abi.check(abi.lib().mtx_target_get_tuning(self.handle.value, C.byref(t)))
        return dict(max_paths_in_flight=t.max_paths_in_flight, trace_kernel=t.trace_kernel, lds_stack=t.lds_stack,
                    trace_oversubscription=t.trace_oversubscription, shade_oversubscription=t.shade_oversubscription,
                    shade_waves_per_simd=t.shade_waves_per_simd, use_graphs=t.use_graphs,
                    graph_replays=t.graph_replays, graph_captures=t.graph_captures, tail_paths=t.tail_paths,
                    queue_sort=t.queue_sort, lds_stack_closest=t.lds_stack_closest)

    def set_tuning(self, **kw):
        """Changes the wavefront launch shape (keys as returned by tuning())."""
        cur = self.tuning()
        for k in ("graph_replays", "graph_captures"):  # read-only counters
            cur.pop(k)
            kw.pop(k, None)
        unknown = set(kw) - set(cur)
        if unknown:
            raise KeyError(f"unknown tuning keys {sorted(unknown)}")
        cur.update(kw)
        t = WavefrontTuning(C.sizeof(WavefrontTuning), *[int(cur[k]) for k in (
            "max_paths_in_flight", "trace_kernel", "lds_stack", "trace_oversubscription", "shade_oversubscription",
            "shade_waves_per_simd", "use_graphs")])
        t.tail_paths = int(cur["tail_paths"])
        t.queue_sort = int(cur["queue_sort"])
        t.lds_stack_closest = int(cur["lds_stack_closest"])
        abi.check(abi.lib().mtx_target_set_tuning(self.handle.value, C.byref(t)))

    def synchronize(self):
        abi.check(abi.lib().mtx_target_synchronize(self.handle.value))

    def set_camera(self, camera):
        d = camera_desc(camera, self.width, self.height)
        abi.check(abi.lib().mtx_target_set_camera(self.handle.value, C.byref(d)))

    def set_sample_offset(self, offset):
        """setSampleSequenceOffset: Sobol index of the next sample of each pixel = count + offset."""
        abi.check(abi.lib().mtx_target_set_sample_offset(self.handle.value, offset))
        self.settings.sample_sequence_offset = offset

    def set_sample_targets(self, targets=None):
        """Per-pixel sample limits ((H, W) floats, bottom row first) or None."""
        if targets is None:
            abi.check(abi.lib().mtx_target_set_sample_targets(self.handle.value, None, 0))
            return
        t = np.ascontiguousarray(targets, np.float32)
        abi.check(abi.lib().mtx_target_set_sample_targets(self.handle.value, _P(t), t.size))

    def plan_adaptive(self, max_samples, min_samples=0, tile_size=0, absolute_error=0.0, relative_error=0.0,
                      confidence=0.0):
        """Plans and installs per-pixel sample targets from the accumulated
        (pilot) moments; returns (largest target, average target, tile variances)."""
        o = AdaptiveOptions(C.sizeof(AdaptiveOptions), max_samples, min_samples, tile_size, absolute_error,
                            relative_error, confidence)
        ts = tile_size or 16
        cols, rows = -(-self.width // ts), -(-self.height // ts)
        plan = np.zeros(2, np.float64)
        tv = np.zeros((rows, cols), np.float64)
        abi.check(abi.lib().mtx_target_plan_adaptive(self.handle.value, C.byref(o), _P(plan), _P(tv), tv.size))
        return int(plan[0]), float(plan[1]), tv

    def restart_samples(self):
        abi.check(abi.lib().mtx_target_reset(self.handle.value))

    def stats(self):
        out = np.zeros(2, np.uint64)
        abi.check(abi.lib().mtx_target_stats(self.handle.value, _P(out)))
        return dict(rays=int(out[0]), samples=int(out[1]))

    def read_pixels(self, kind="color"):
        """(H, W, 4) f32, top row first."""
        out = np.zeros((self.height, self.width, 4), np.float32)
        written = C.c_size_t(0)
        abi.check(abi.lib().mtx_target_read_pixels(self.handle.value, PIXELS[kind], _P(out), out.size,
                                                   C.byref(written)))
        return out

    def compose(self, tone_mapping="agx", gamma22=False, **grading):
        """Displayed RGBA8 (H, W, 4), top row first (exposure, contrast,
        saturation, temperature, tint, vignette as keywords)."""
        d = display_settings(tone_mapping, gamma22, **grading)
        out = np.zeros((self.height, self.width, 4), np.uint8)
        written = C.c_size_t(0)
        abi.check(abi.lib().mtx_target_compose(self.handle.value, C.byref(d), _P(out), out.size, C.byref(written)))
        return out

    def denoise_jnlm(self, tone_mapping="agx", gamma22=False, grading=None, **options):
        """JNLM of the composed image: (display-linear RGBA f32, RGBA8), both
        (H, W, 4) top row first. Options: strength, search_radius,
        patch_radius, albedo_bandwidth, normal_bandwidth, iterations."""
        d = display_settings(tone_mapping, gamma22, **(grading or {}))
        o = JnlmOptions()
        abi.check(abi.lib().mtx_jnlm_options_default(C.byref(o)))
        for k, v in options.items():
            setattr(o, k, v)
        lin = np.zeros((self.height, self.width, 4), np.float32)
        enc = np.zeros((self.height, self.width, 4), np.uint8)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_target_denoise_jnlm(self.handle.value, C.byref(d), C.byref(o), _P(lin), _P(enc),
                                                    self.width * self.height, err, 512), err)
        return lin, enc

    def read_accumulators(self):
        """(sums (H, W, 13) f64 bottom row first, counts (H, W) u32)."""
        sums = np.zeros((self.height, self.width, 13), np.float64)
        counts = np.zeros((self.height, self.width), np.uint32)
        abi.check(abi.lib().mtx_target_read_accumulators(self.handle.value, _P(sums), sums.size, _P(counts),
                                                         counts.size))
        return sums, counts

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and abi._lib is not None:
            abi.lib().mtx_target_destroy(h.value)


class Render:
    """Offline converge loop over one RenderTarget (Render.converge, Render.ts:827-900)."""

    def __init__(self, target, samples_per_batch=8):
        self.target = target
        self.samples_per_batch = samples_per_batch

    def _stage(self, cap):
        t = self.target
        t.restart_samples()
        left = cap
        while left > 0:
            n = min(self.samples_per_batch, left)
            t.render(n)
            left -= n
        t.synchronize()

    def converge(self, samples, adaptive=None):
        """Offline convergence. With `adaptive` (dict of plan_adaptive options
        plus pilot_samples, default 8) it runs the reference's two stages: a
        pilot pass, a per-pixel plan from its FP64 moments, then a final pass
        continuing the Sobol sequence after the pilot samples. Returns the
        resolved color (top row first) and a result dict."""
        t = self.target
        base = t.settings.sample_sequence_offset
        result = dict(samples=samples, average_samples_per_pixel=float(samples), pilot_samples=0, adaptive=False)
        if adaptive:
            opts = dict(adaptive)
            pilot = min(samples, max(2, round(opts.pop("pilot_samples", 8))))
            t.set_sample_targets(None)
            self._stage(pilot)
            mx, avg, _ = t.plan_adaptive(samples, **opts)
            t.set_sample_offset(base + pilot)
            self._stage(mx)
            t.set_sample_offset(base)
            result = dict(samples=mx, average_samples_per_pixel=avg, pilot_samples=pilot, adaptive=True)
        else:
            t.set_sample_targets(None)
            self._stage(samples)
        return t.read_pixels("color"), result
