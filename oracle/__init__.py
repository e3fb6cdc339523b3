"""CPU oracle for the render hot path -- test infrastructure only.

The C++ sources in this directory restate the reference algorithms (file:line
citations in each header). This module loads ``oracle/liborc.so`` with ctypes
and wraps it for numpy. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it; the product package
``mutracer_amd`` never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def math_eval(kind, x, y=None):
    """include/mtx_fmath.h on the host: kind in {"sin","cos","pow","atan2","acos","exp","sincos.s","sincos.c"}."""
    k = ["sin", "cos", "pow", "atan2", "acos", "exp", "sincos.s", "sincos.c"].index(kind)
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, np.float32)
    out = np.empty_like(x)
    L = lib()
    L.orc_math_eval.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
    L.orc_math_eval(k, _p(x), _p(y), _p(out), x.size)
    return out


def light_tables(lights, env=None):
    """Walker alias table and pmfs of the oracle (lights: (n,16) f32, env: (H,W,4) f32 or None)."""
    L = lib()
    lights = np.ascontiguousarray(np.asarray(lights, np.float32).reshape(-1, 16))
    n = lights.shape[0]
    envp = None if env is None else np.ascontiguousarray(env, np.float32)
    cap = n + 1
    prob = np.zeros(cap, np.float32)
    alias, typ, idx = (np.zeros(cap, np.int32) for _ in range(3))
    lpmf = np.zeros(max(n, 1), np.float32)
    epmf = np.zeros(1, np.float32)
    total = np.zeros(1, np.float64)
    L.orc_light_tables.restype = C.c_uint32
    L.orc_light_tables.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32] + [C.c_void_p] * 4 + [
        C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    m = L.orc_light_tables(_p(lights), n, _p(envp), 0 if envp is None else envp.shape[1],
                           0 if envp is None else envp.shape[0], _p(prob), _p(alias), _p(typ), _p(idx), cap,
                           _p(lpmf), _p(epmf), _p(total))
    return dict(prob=prob[:m], alias=alias[:m], type=typ[:m], index=idx[:m], light_pmf=lpmf[:n],
                env_pmf=float(epmf[0]), total=float(total[0]))


def env_cdf(env):
    """Environment CDF of the oracle: (H+1, W+2, 2) f32."""
    env = np.ascontiguousarray(env, np.float32)
    h, w = env.shape[0], env.shape[1]
    out = np.zeros((h + 1, w + 2, 2), np.float32)
    L = lib()
    L.orc_env_cdf.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
    L.orc_env_cdf(_p(env), w, h, _p(out))
    return out


def camera_rays(camera, width, height, pixels, pattern=0, seed=0, log_spp=0):
    """Primary rays of (px, py, sample index) rows: (n, 6) origin + direction,
    pinhole or thin lens as the packed camera's aperture says."""
    L = lib()
    L.orc_camera_rays.restype = None
    L.orc_camera_rays.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.c_void_p]
    cam = np.ascontiguousarray(camera, np.float32)
    pix = np.ascontiguousarray(pixels, np.uint32).reshape(-1, 3)
    out = np.zeros((pix.shape[0], 6), np.float32)
    L.orc_camera_rays(_p(cam), width, height, _p(pix), pix.shape[0], pattern, seed & 0xFFFFFFFF, log_spp, _p(out))
    return out


def threads():
    """OpenMP threads the oracle's batch entry points run on."""
    return int(lib().orc_threads())


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _declare(L):
    u32, f32, vp = C.c_uint32, C.c_float, C.c_void_p
    L.orc_threads.restype = C.c_int32
    L.orc_threads.argtypes = []
    for name in ("orc_sample_hash",):
        getattr(L, name).restype = u32
        getattr(L, name).argtypes = [u32]
    for name in ("orc_sample_seed", "orc_owen", "orc_morton2"):
        getattr(L, name).restype = u32
        getattr(L, name).argtypes = [u32, u32]
    L.orc_zsobol_index.restype = u32
    L.orc_zsobol_index.argtypes = [u32, u32, u32]
    L.orc_sobol_d1.restype = u32
    L.orc_sobol_d1.argtypes = [u32]
    L.orc_sobol2d_batch.argtypes = [vp, u32, u32, vp]
    L.orc_pcg4d.argtypes = [vp]
    L.orc_fuzz_albedo.restype = C.c_float
    L.orc_fuzz_albedo.argtypes = [C.c_float, C.c_float]
    L.orc_halton.restype = C.c_double
    L.orc_halton.argtypes = [u32, u32]
    L.orc_structured_sample.argtypes = [u32] * 8 + [vp]
    L.orc_bvh_build.restype = vp
    L.orc_bvh_build.argtypes = [vp, u32, u32]
    L.orc_bvh_info.argtypes = [vp, vp]
    L.orc_bvh_get.argtypes = [vp] * 7
    L.orc_bvh_free.argtypes = [vp]
    L.orc_blas_build.restype = vp
    L.orc_blas_build.argtypes = [vp, u32, vp, u32, u32]
    L.orc_tlas_build.restype = vp
    L.orc_tlas_build.argtypes = [vp, vp, vp, vp, vp, u32, vp, vp, u32, u32]
    L.orc_tlas_refit.argtypes = [vp, vp, vp, vp, vp, vp, u32, vp, vp, vp, u32]
    L.orc_flat_info.argtypes = [vp, vp]
    L.orc_flat_get.argtypes = [vp, vp, vp, vp]
    L.orc_flat_free.argtypes = [vp]
    L.orc_instance_world_bounds.argtypes = [vp, vp, vp, vp]
    L.orc_scene_create.restype = vp
    L.orc_scene_create.argtypes = [vp, C.c_uint64, vp, C.c_uint64, vp, C.c_uint64, vp, C.c_uint64, vp, u32, u32]
    L.orc_scene_free.argtypes = [vp]
    L.orc_trace_closest_batch.argtypes = [vp, vp, u32, u32, vp, vp]
    L.orc_trace_any_batch.argtypes = [vp, vp, u32, vp]
    L.orc_trace_brute_batch.argtypes = [vp, vp, u32, u32, vp]
    if hasattr(L, "orc_render"):
        _declare_render(L)
    if hasattr(L, "orc_bsdf_eval"):
        _declare_bsdf(L)


def _declare_render(L):
    vp, u32 = C.c_void_p, C.c_uint32
    L.orc_render.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp, vp]
    L.orc_render.restype = C.c_int


# ---------------------------------------------------------------- sampler
def sample_hash(x):
    return lib().orc_sample_hash(x & 0xFFFFFFFF)


def sample_seed(a, b):
    return lib().orc_sample_seed(a & 0xFFFFFFFF, b & 0xFFFFFFFF)


def owen(x, s):
    return lib().orc_owen(x & 0xFFFFFFFF, s & 0xFFFFFFFF)


def morton2(x, y):
    return lib().orc_morton2(x, y)


def zsobol_index(m, dim, log_spp):
    return lib().orc_zsobol_index(m & 0xFFFFFFFF, dim & 0xFFFFFFFF, log_spp)


def sobol_d1(i):
    return lib().orc_sobol_d1(i & 0xFFFFFFFF)


def sobol2d_bits(indices, seed):
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    out = np.zeros((idx.size, 2), dtype=np.uint32)
    lib().orc_sobol2d_batch(_p(idx), idx.size, seed & 0xFFFFFFFF, _p(out))
    return out


def sobol2d(indices, seed):
    return sobol2d_bits(indices, seed).astype(np.float64) / 16777216.0


def pcg4d(state):
    """One pcg4d step (rand.wgsl.ts:13-19) of a 4-word state; returns the new state."""
    v = np.array(state, dtype=np.uint32).reshape(4).copy()
    lib().orc_pcg4d(_p(v))
    return v


def fuzz_albedo(view_cosine, roughness):
    """Fitted directional albedo of the OpenPBR fuzz lobe (f32)."""
    return lib().orc_fuzz_albedo(view_cosine, roughness)


def halton(index, base):
    """Radical inverse of index in base, FP64 (halton.ts:6-15)."""
    return lib().orc_halton(index & 0xFFFFFFFF, base)


# Draw tags of OracleRenderer.draws: sample dimensions, or these flags | slot.
DRAW_HASHED = 0x80000000
DRAW_CANDIDATE = 0xC0000000


def structured_sample(pattern, px, py, width, index, dim, log_spp, seed):
    out = np.zeros(2, dtype=np.float32)
    lib().orc_structured_sample(pattern, px, py, width, index & 0xFFFFFFFF, dim & 0xFFFFFFFF,
                                log_spp, seed & 0xFFFFFFFF, _p(out))
    return out


# ---------------------------------------------------------------- BVH
def build_binned_sah(aabbs, bins=8):
    a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
    L = lib()
    h = L.orc_bvh_build(_p(a), a.shape[0], bins)
    info = np.zeros(4, dtype=np.uint32)
    L.orc_bvh_info(h, _p(info))
    n = int(info[0])
    bounds = np.zeros((n, 6), np.float32)
    lf = np.zeros(n, np.uint32)
    pc = np.zeros(n, np.uint32)
    pi = np.zeros(a.shape[0], np.uint32)
    order = np.zeros(n, np.uint32)
    offsets = np.zeros(n, np.uint32)
    L.orc_bvh_get(h, _p(bounds), _p(lf), _p(pc), _p(pi), _p(order), _p(offsets))
    L.orc_bvh_free(h)
    return dict(node_count=n, bounds=bounds, left_first=lf, prim_count=pc, prim_indices=pi,
                order=order, offsets=offsets, vec4_count=int(info[1]), max_depth=int(info[2]),
                max_leaf=int(info[3]))


def _flat(h, with_source=False):
    L = lib()
    info = np.zeros(5, dtype=np.uint64)
    L.orc_flat_info(h, _p(info))
    words = np.zeros(int(info[0]), np.uint32)
    inst = np.zeros(int(info[1]), np.uint32)
    source = np.zeros(int(info[0]) // 4, np.uint32) if with_source else None
    L.orc_flat_get(h, _p(words), _p(inst), _p(source))
    area = info[4:5].view(np.float64)[0]
    return dict(words=words, instance_words=inst, max_depth=int(info[2]), max_leaf=int(info[3]),
                build_surface_area=float(area), source_of_record=source)


def build_blas(positions, indices, bins=8):
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    L = lib()
    h = L.orc_blas_build(_p(p), p.shape[0], _p(i), i.size // 3, bins)
    out = _flat(h, with_source=True)
    L.orc_flat_free(h)
    return out


def refit_blas(words, positions):
    """refitBLAS (src/renderer/bvhAccel.ts:800-870) on BVH2 words whose leaf
    entries hold the triangle indices: new node boxes for `positions` (same
    vertex count), topology kept. Returns (words, refitted surface area)."""
    w = np.array(words, np.uint32).reshape(-1)
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    L = lib()
    L.orc_blas_refit_words.restype = C.c_double
    L.orc_blas_refit_words.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_uint32]
    area = L.orc_blas_refit_words(_p(w), w.size // 4, _p(p), p.shape[0])
    return w, float(area)


def _display6(display):
    d = dict(exposure=1.0, contrast=1.0, saturation=1.0, temperature=0.0, tint=0.0, vignette=0.0)
    d.update({k: v for k, v in (display or {}).items() if k in d})
    return np.array([d[k] for k in ("exposure", "contrast", "saturation", "temperature", "tint", "vignette")],
                    np.float32)


TONE = dict(agx=0, aces=1, cineon=2, reinhard=3, linear=4)


def compose(sums, display=None):
    """Displayed RGBA8 (H, W, 4), top row first, from FP64 accumulators
    (H, W, 13), bottom row first (orc_display.cpp)."""
    display = display or {}
    h, w, _ = sums.shape
    planes = np.ascontiguousarray(np.moveaxis(sums, 2, 0).reshape(13, -1), np.float64)
    out = np.zeros((h, w, 4), np.uint8)
    d6 = _display6(display)
    L = lib()
    L.orc_compose.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
    L.orc_compose(_p(planes), w, h, _p(d6), TONE[display.get("tone_mapping", "agx")], int(display.get("gamma22", 0)),
                  _p(out))
    return out


def jnlm(sums, display=None, strength=1.0, search_radius=5, patch_radius=1, albedo_bandwidth=0.4,
         normal_bandwidth=0.05, iterations=2):
    """JNLM of the composed image (orc_display.cpp): (filtered display-linear
    RGBA f32, RGBA8), both (H, W, 4) top row first."""
    display = display or {}
    h, w, _ = sums.shape
    planes = np.ascontiguousarray(np.moveaxis(sums, 2, 0).reshape(13, -1), np.float64)
    lin = np.zeros((h, w, 4), np.float32)
    enc = np.zeros((h, w, 4), np.uint8)
    opts = np.array([strength, albedo_bandwidth, normal_bandwidth], np.float32)
    shape = np.array([search_radius, patch_radius, iterations], np.uint32)
    L = lib()
    L.orc_jnlm.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                           C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_jnlm(_p(planes), w, h, _p(_display6(display)), TONE[display.get("tone_mapping", "agx")],
               int(display.get("gamma22", 0)), _p(opts), _p(shape), _p(lin), _p(enc))
    return lin, enc


def tonemap(rgb, mode):
    x = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    out = np.zeros_like(x)
    L = lib()
    L.orc_tonemap.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
    L.orc_tonemap(_p(x), x.shape[0], TONE[mode], _p(out))
    return out


def srgb_encode(rgb):
    x = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    out = np.zeros_like(x)
    L = lib()
    L.orc_srgb_encode.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
    L.orc_srgb_encode(_p(x), x.shape[0], _p(out))
    return out


def _tlas_args(inst):
    bi = np.ascontiguousarray(inst["blas_index"], np.uint32)
    bounds = np.ascontiguousarray(inst["bounds"], np.float32).reshape(-1, 6)
    tr = np.ascontiguousarray(inst["transform"], np.float64).reshape(-1, 16)
    mat = np.ascontiguousarray(inst["material_word"], np.uint32)
    col = np.ascontiguousarray(inst.get("color", np.ones((bi.size, 3))), np.float32).reshape(-1, 3)
    return bi, bounds, tr, mat, col


def build_tlas(inst, blas_offsets, vertex_offsets, bins=64, per_instance_bounds=True):
    bi, bounds, tr, mat, col = _tlas_args(inst)
    bo = np.ascontiguousarray(blas_offsets, np.uint32)
    vo = np.ascontiguousarray(vertex_offsets, np.uint32)
    L = lib()
    h = L.orc_tlas_build(_p(bi), _p(bounds), _p(tr), _p(mat), _p(col), bi.size, _p(bo), _p(vo),
                         bins, int(per_instance_bounds))
    return _TlasHandle(h, bo, vo)


class _TlasHandle:
    def __init__(self, h, bo, vo):
        self.h, self.bo, self.vo = h, bo, vo

    def get(self):
        return _flat(self.h)

    def refit(self, inst, dirty=None):
        bi, bounds, tr, mat, col = _tlas_args(inst)
        d = None if dirty is None else np.ascontiguousarray(dirty, np.uint32)
        lib().orc_tlas_refit(self.h, _p(bi), _p(bounds), _p(tr), _p(mat), _p(col), bi.size,
                             _p(self.bo), _p(self.vo), _p(d), 0 if d is None else d.size)

    def __del__(self):
        try:
            lib().orc_flat_free(self.h)
        except Exception:
            pass


def instance_world_bounds(transform, lo, hi):
    m = np.ascontiguousarray(transform, np.float64).reshape(16)
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    out = np.zeros(6, np.float32)
    lib().orc_instance_world_bounds(_p(m), _p(lo), _p(hi), _p(out))
    return out


# ---------------------------------------------------------------- traversal
class TraceScene:
    """Reference-layout scene (TLAS + concatenated BLAS + instance records)."""

    def __init__(self, tlas_words, blas_words, instance_words, positions, mat_double_sided,
                 per_instance_bounds=True):
        self.keep = [np.ascontiguousarray(tlas_words, np.uint32),
                     np.ascontiguousarray(blas_words, np.uint32),
                     np.ascontiguousarray(instance_words, np.uint32),
                     np.ascontiguousarray(positions, np.float32).reshape(-1, 3),
                     np.ascontiguousarray(mat_double_sided, np.uint8)]
        t, b, i, p, m = self.keep
        self.h = lib().orc_scene_create(_p(t), t.size, _p(b), b.size, _p(i), i.size, _p(p),
                                        p.shape[0], _p(m), m.size, int(per_instance_bounds))

    def closest(self, rays, double_sided=False, stats=False):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((r.shape[0], 8), np.uint32)
        st = np.zeros((r.shape[0], 3), np.uint32) if stats else None
        lib().orc_trace_closest_batch(self.h, _p(r), r.shape[0], int(double_sided), _p(hits), _p(st))
        return (hits, st) if stats else hits

    def any(self, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(r.shape[0], np.uint32)
        lib().orc_trace_any_batch(self.h, _p(r), r.shape[0], _p(occ))
        return occ

    def brute(self, rays, double_sided=False):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((r.shape[0], 8), np.uint32)
        lib().orc_trace_brute_batch(self.h, _p(r), r.shape[0], int(double_sided), _p(hits))
        return hits

    def __del__(self):
        try:
            lib().orc_scene_free(self.h)
        except Exception:
            pass


# ---------------------------------------------------------------- scene assembly
def assemble_scene(meshes, instances, blas_bins=8, tlas_bins=64, per_instance_bounds=True,
                   material_double_sided=None):
    """Reference-layout scene from meshes and instances (oracle side).

    meshes: list of dict(positions (V,3) f32, indices (T*3,) u32).
    instances: dict(mesh (N,), transform (N,16) f64 column-major, material (N,)).
    Follows WGPURayTracePass._uploadBVH (WGPURayTracePass.ts:2627-2694): BLAS
    buffers concatenated (offsets in vec4 units, bvhAccel.ts:288-304), one
    global vertex array, TLAS built over the instances' world boxes.
    """
    blas_words, blas_offsets, vertex_offsets, positions, roots = [], [], [], [], []
    source_tri = []
    off = voff = 0
    max_depth = 1
    for m in meshes:
        flat = build_blas(m["positions"], m["indices"], blas_bins)
        w = flat["words"]
        blas_offsets.append(off)
        vertex_offsets.append(voff)
        blas_words.append(w)
        src = flat["source_of_record"]
        source_tri.append(src)
        lo = w[0:3].view(np.float32) if w.size else np.zeros(3, np.float32)
        hi = w[4:7].view(np.float32) if w.size else np.zeros(3, np.float32)
        roots.append((lo.copy(), hi.copy()))
        max_depth = max(max_depth, flat["max_depth"])
        off += w.size // 4
        voff += np.asarray(m["positions"]).reshape(-1, 3).shape[0]
        positions.append(np.asarray(m["positions"], np.float32).reshape(-1, 3))
    mesh_idx = np.asarray(instances["mesh"], np.int64)
    tr = np.asarray(instances["transform"], np.float64).reshape(-1, 16)
    bounds = np.array([instance_world_bounds(tr[i], *roots[mesh_idx[i]]) for i in range(mesh_idx.size)],
                      np.float32).reshape(-1, 6)
    mats = np.asarray(instances.get("material", np.zeros(mesh_idx.size)), np.int64)
    inst = dict(blas_index=mesh_idx, bounds=bounds, transform=tr, material_word=(mats << 8).astype(np.uint32),
                color=instances.get("color", np.ones((mesh_idx.size, 3))))
    h = build_tlas(inst, blas_offsets, vertex_offsets, tlas_bins, per_instance_bounds)
    t = h.get()
    all_blas = np.concatenate(blas_words) if blas_words else np.zeros(0, np.uint32)
    all_pos = np.concatenate(positions) if positions else np.zeros((0, 3), np.float32)
    nmat = int(mats.max()) + 1 if mats.size else 1
    mds = np.zeros(nmat, np.uint8) if material_double_sided is None else np.asarray(material_double_sided, np.uint8)
    scene = TraceScene(t["words"], all_blas, t["instance_words"], all_pos, mds, per_instance_bounds)
    # identities for parity: record -> (mesh, source triangle); target -> source instance
    rec_tri = np.full(all_blas.size // 4, 0xFFFFFFFF, np.uint32)
    for k, src in enumerate(source_tri):
        rec_tri[blas_offsets[k]:blas_offsets[k] + src.size] = src
    ints = t["words"].view(np.int32)
    target_src = np.zeros(mesh_idx.size, np.int64)
    stack = [0] if t["words"].size else []
    while stack:
        node = stack.pop()
        c = ints[node * 4 + 3]
        if c >= 0:
            stack += [c, node + 2]
            continue
        for k in range(-c):
            target_src[ints[node * 4 + 8 + k * 4]] = ints[node * 4 + 8 + k * 4 + 2]
    scene.record_triangle = rec_tri
    scene.target_source = target_src
    scene.tlas_handle = h
    scene.tlas_max_depth = t["max_depth"]
    scene.blas_max_depth = max_depth
    scene.instance_bounds = bounds
    return scene


def hits_to_ids(scene, hits):
    """(hit, source instance, source triangle, t bits) per ray from oracle hit words."""
    h = np.asarray(hits, np.uint32).reshape(-1, 8)
    hit = h[:, 0].astype(bool)
    inst = np.where(hit, scene.target_source[np.where(hit, h[:, 1], 0)], -1)
    tri = np.where(hit, scene.record_triangle[np.where(hit, h[:, 2], 0)].astype(np.int64), -1)
    return hit, inst, tri, h[:, 3]


# ---------------------------------------------------------------- integrator
_LUT = None


def _energy_lut():
    global _LUT
    if _LUT is None:
        path = os.path.join(_HERE, "..", "mutracer_amd", "data", "openpbr_energy_lut.f32")
        _LUT = np.fromfile(path, "<f4")
        assert _LUT.size == 32 * 2177
        lib().orc_set_energy_lut(_p(_LUT))
    return _LUT


def _declare_render2(L):
    vp, u32 = C.c_void_p, C.c_uint32
    L.orc_set_energy_lut.argtypes = [vp]
    L.orc_render.restype = C.c_uint64
    L.orc_render.argtypes = [vp, vp, vp, u32, vp, u32, vp, u32, u32, vp, vp, C.c_float, u32, u32, u32, u32,
                             u32, u32, vp, vp, vp, u32, vp]
    L.orc_emissive_triangles.restype = u32
    L.orc_emissive_triangles.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, u32, vp, vp, vp, u32, vp]
    L.orc_path_draws.restype = u32
    L.orc_path_draws.argtypes = [vp, vp, vp, u32, vp, u32, vp, u32, u32, vp, vp, C.c_float, u32, u32, u32, vp, u32,
                                 vp, vp, u32, vp]
    L.orc_texset_create.restype = vp
    L.orc_texset_create.argtypes = [u32, vp, vp, u32, vp, vp, u32, vp]
    L.orc_texset_destroy.argtypes = [vp]
    L.orc_sample_map.argtypes = [vp, u32, u32, vp, vp, vp]


class TextureSet:
    """The oracle's copy of a scene's material textures and vertex UVs
    (orc_texset_create): textures are mutracer_amd.scene.Texture objects;
    uvs / uvs1 are per global vertex (zeros where a mesh has none);
    emissive_uv holds 12 floats per emissive triangle."""

    def __init__(self, textures, uvs=None, uvs1=None, emissive_uv=None):
        L = lib()
        _declare_render2(L)
        n = len(textures)
        desc = np.zeros((max(n, 1), 11), np.uint32)
        pixels = (C.c_void_p * max(n, 1))()
        keep = []
        for i, t in enumerate(textures):
            px = t.storage_pixels()
            keep.append(px)
            f = np.array([*t.uv_scale, *t.uv_offset], np.float32).view(np.uint32)
            desc[i] = [px.shape[1], px.shape[0], t.storage, int(t.flip_y), int(t.uv_channel),
                       t.wrap_code(t.wrap_s), t.wrap_code(t.wrap_t), *f]
            pixels[i] = px.ctypes.data
        uvs = np.ascontiguousarray(np.zeros((0, 2)) if uvs is None else uvs, np.float32)
        uvs1 = np.ascontiguousarray(np.zeros_like(uvs) if uvs1 is None else uvs1, np.float32)
        eu = np.ascontiguousarray(np.zeros((0, 12)) if emissive_uv is None else emissive_uv, np.float32)
        self.h = L.orc_texset_create(n, _p(desc), pixels, uvs.shape[0], _p(uvs), _p(uvs1), eu.shape[0], _p(eu))

    def sample(self, index, uv):
        """sampleMap of texture `index` at uv (n, 2): filtered RGBA (n, 4)
        and decoded color (n, 3)."""
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        rgba = np.zeros((uv.shape[0], 4), np.float32)
        dec = np.zeros((uv.shape[0], 3), np.float32)
        lib().orc_sample_map(self.h, index, uv.shape[0], _p(uv), _p(rgba), _p(dec))
        return rgba, dec

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_texset_destroy(self.h)
            self.h = None


def scene_uvs(scene):
    """Per-global-vertex UV0 / UV1 of a scene's meshes in mesh order (zeros
    where a mesh has none), and the emissive variant of UV1 (UV0 where a mesh
    has no UV1, lightSampling.ts:262-263)."""
    uv, uv1, uv1e = [], [], []
    for m in scene.meshes:
        nv = np.asarray(m["positions"]).reshape(-1, 3).shape[0]
        z = np.zeros((nv, 2), np.float32)
        a = z if m.get("uvs") is None else np.asarray(m["uvs"], np.float32).reshape(-1, 2)
        b = z if m.get("uvs1") is None else np.asarray(m["uvs1"], np.float32).reshape(-1, 2)
        uv.append(a)
        uv1.append(b)
        uv1e.append(a if m.get("uvs1") is None else b)
    cat = lambda x: np.concatenate(x) if x else np.zeros((0, 2), np.float32)  # noqa: E731
    return cat(uv), cat(uv1), cat(uv1e)


def _declare_bsdf(L):
    vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
    L.orc_bsdf_eval.argtypes = [vp, vp, u32, vp, vp, vp]
    L.orc_bsdf_sample.argtypes = [vp, vp, u32, vp, vp, vp]
    L.orc_bsdf_lobes.argtypes = [vp, vp, vp, vp]
    L.orc_metal_fresnel.argtypes = [f32, vp, vp, vp, vp]
    L.orc_coat_energy.argtypes = [vp, f32]
    L.orc_coat_energy.restype = f32


def _material_and_view(material, view):
    m = np.ascontiguousarray(material.pack() if hasattr(material, "pack") else material, np.float32)
    v = np.ascontiguousarray(view, np.float32)
    return m, v


def bsdf_eval(material, view, dirs):
    """BSDF * |cos| (n, 3) and pdf (n,) of unit world directions `dirs` (n, 3)
    for `material` on the outside of a surface with normal +z, seen along the
    unit vector `view` (toward the viewer)."""
    L = lib()
    _energy_lut()
    m, v = _material_and_view(material, view)
    d = np.ascontiguousarray(dirs, np.float32)
    f = np.zeros((d.shape[0], 3), np.float32)
    pdf = np.zeros(d.shape[0], np.float32)
    L.orc_bsdf_eval(_p(m), _p(v), d.shape[0], _p(d), _p(f), _p(pdf))
    return f, pdf


def bsdf_sample(material, view, u):
    """Directions (n, 3) and acceptance flags (n,) of the lobe sampler for
    numbers u (n, 3) = (rx, ry, lobe) in [0, 1)."""
    L = lib()
    _energy_lut()
    m, v = _material_and_view(material, view)
    uu = np.ascontiguousarray(u, np.float32)
    d = np.zeros((uu.shape[0], 3), np.float32)
    ok = np.zeros(uu.shape[0], np.uint8)
    L.orc_bsdf_sample(_p(m), _p(v), uu.shape[0], _p(uu), _p(d), _p(ok))
    return d, ok.astype(bool)


def bsdf_lobes(material, view):
    """Lobe probabilities (fuzz, coat, metal, specular, transmission, diffuse,
    subsurface) and the coat's view-side base scale (RGB)."""
    L = lib()
    _energy_lut()
    m, v = _material_and_view(material, view)
    p = np.zeros(7, np.float32)
    cbs = np.zeros(3, np.float32)
    L.orc_bsdf_lobes(_p(m), _p(v), _p(p), _p(cbs))
    return p, cbs


def metal_fresnel(cosine, f0, tint):
    """The oracle's F82-tint metal Fresnel at `cosine` and its hemispherical average."""
    f0 = np.ascontiguousarray(f0, np.float32)
    tint = np.ascontiguousarray(tint, np.float32)
    out = np.zeros(3, np.float32)
    avg = np.zeros(3, np.float32)
    lib().orc_metal_fresnel(float(cosine), _p(f0), _p(tint), _p(out), _p(avg))
    return out, avg


def coat_energy(material, cosine):
    _energy_lut()
    m = np.ascontiguousarray(material.pack() if hasattr(material, "pack") else material, np.float32)
    return float(lib().orc_coat_energy(_p(m), float(cosine)))


def emissive_triangles(scene, with_corners=False):
    """World-space emissive triangles of a mutracer_amd.scene.Scene: (n, 12)
    f32 records (p0, p1, p2, area, material, 0) and their f64 alias weights
    (and the global vertex index of each corner)."""
    L = lib()
    _declare_render2(L)
    meshes = scene.meshes
    pos = np.ascontiguousarray(np.concatenate([np.asarray(m["positions"], np.float32).reshape(-1, 3)
                                               for m in meshes]) if meshes else np.zeros((0, 3), np.float32))
    idx = np.ascontiguousarray(np.concatenate([np.asarray(m["indices"], np.uint32).reshape(-1)
                                               for m in meshes]) if meshes else np.zeros(0, np.uint32))
    voff = np.cumsum([0] + [np.asarray(m["positions"]).reshape(-1, 3).shape[0] for m in meshes]).astype(np.uint32)
    ioff = np.cumsum([0] + [np.asarray(m["indices"]).size for m in meshes]).astype(np.uint32)
    tcount = np.array([np.asarray(m["indices"]).size // 3 for m in meshes], np.uint32)
    inst_mesh = np.ascontiguousarray(scene.mesh_index, np.uint32)
    tr = np.ascontiguousarray(np.asarray(scene.transforms, np.float64).reshape(-1, 16))
    inst_mat = np.ascontiguousarray(scene.instance_material, np.uint32)
    mats = scene.packed_materials()
    args = [_p(pos), _p(voff), _p(idx), _p(ioff), _p(tcount), _p(inst_mesh), _p(tr), _p(inst_mat),
            inst_mesh.size, _p(mats)]
    n = L.orc_emissive_triangles(*args, None, None, 0, None)
    out = np.zeros((max(n, 1), 12), np.float32)
    w = np.zeros(max(n, 1), np.float64)
    corners = np.zeros((max(n, 1), 3), np.uint32)
    L.orc_emissive_triangles(*args, _p(out), _p(w), n, _p(corners))
    if with_corners:
        return out[:n], w[:n], corners[:n]
    return out[:n], w[:n]


class OracleRenderer:
    """CPU restatement of the megakernel over a mutracer_amd.scene.Scene."""

    def __init__(self, scene, width, height, bounces=8, candidates=1, pattern=0, seed=0, max_samples=None,
                 clamping=0.0, sample_offset=0):
        L = lib()
        _declare_render2(L)
        _energy_lut()
        self.scene = scene
        self.trace = assemble_scene(scene.meshes, dict(mesh=scene.mesh_index, transform=np.array(scene.transforms),
                                                       material=scene.instance_material),
                                    material_double_sided=[1 if m.double_sided else 0 for m in scene.materials])
        self.normals = np.ascontiguousarray(np.concatenate([m["normals"] for m in scene.meshes]), np.float32)
        self.materials = scene.packed_materials()
        self.lights = scene.packed_lights()
        self.env = None if scene.env is None else np.ascontiguousarray(scene.env, np.float32)
        self.camera = scene.camera.pack(width, height)
        from mutracer_amd.sampling import sampling_pattern_code, zsobol_log2spp
        if isinstance(pattern, str):
            pattern = sampling_pattern_code(pattern)
        log = zsobol_log2spp((max_samples or 1) - 1 if pattern == 1 else (max_samples or 1))
        self.settings = np.array([width, height, bounces, candidates, pattern, seed & 0xFFFFFFFF, log, sample_offset],
                                 np.uint32)
        self.clamping = clamping
        self.emissive, self.emissive_weight, corners = emissive_triangles(scene, with_corners=True)
        self.textures = None
        if any(m.has_maps() for m in scene.materials):
            uv, uv1, uv1e = scene_uvs(scene)
            eu = np.concatenate([uv[corners].reshape(-1, 6), uv1e[corners].reshape(-1, 6)], 1)
            self.textures = TextureSet(scene.textures, uv, uv1, eu)
        self.width, self.height = width, height
        self.rays = 0

    def draws(self, px, py, sample, cap=4096):
        """Tags of the random-pattern draws one path of (px, py, sample) makes,
        in order: the dimension of each sample1d/sample2d draw, DRAW_HASHED |
        slot, DRAW_CANDIDATE | salt."""
        out = np.zeros(cap, np.uint32)
        env = self.env
        ew, eh = (0, 0) if env is None else (env.shape[1], env.shape[0])
        n = lib().orc_path_draws(self.trace.h, _p(self.normals), _p(self.materials), self.materials.shape[0],
                                 _p(self.lights), self.lights.shape[0], None if env is None else _p(env), ew, eh,
                                 _p(self.camera), _p(self.settings), self.clamping, px, py, sample, _p(out), cap,
                                 _p(self.emissive), _p(self.emissive_weight), self.emissive.shape[0],
                                 None if self.textures is None else self.textures.h)
        assert n <= cap, "draw log overflow"
        return out[:n].copy()

    def render(self, sample_start, count, x0=0, y0=0, w=None, h=None):
        w = self.width if w is None else w
        h = self.height if h is None else h
        out = np.zeros((h, w, 13), np.float64)
        env = self.env
        ew, eh = (0, 0) if env is None else (env.shape[1], env.shape[0])
        self.rays += lib().orc_render(self.trace.h, _p(self.normals), _p(self.materials), self.materials.shape[0],
                                      _p(self.lights), self.lights.shape[0], None if env is None else _p(env),
                                      ew, eh, _p(self.camera), _p(self.settings), self.clamping, x0, y0, w, h,
                                      sample_start, count, _p(out), _p(self.emissive), _p(self.emissive_weight),
                                      self.emissive.shape[0], None if self.textures is None else self.textures.h)
        return out
