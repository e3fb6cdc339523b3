"""Raw scene descriptor (version 8) encoder for host-side Scenes.

The descriptor is the byte buffer `mtx_context_commit` takes, with the word
layout of the reference's native scene format (native/docs/scene-format.md,
records of native/include/mutracer-scene.h, word offsets of
native/include/mutracer-layout.h). `encode` writes a `scene.Scene` as one
geometry per mesh, one mesh record per instance, one material record per
material (OpenPBR parameters outside the basic record go in extension entries),
one light record per analytic light, a perspective camera, the material
textures (their table indices are the materials' slot values) and, when the
scene has one, the environment as an RGBA32F texture after them, used for
both lighting and the visible background.

Every value is stored as f32: instance transforms, light frames and the camera
are rounded once here. `f32_scene` returns the Scene those rounded values
describe, which renders identically to the committed descriptor.
"""
from __future__ import annotations

import copy
import math

import numpy as np

from . import scene as S

VERSION = 8
HEADER_WORDS = 32
NO_TEXTURE = 0xFFFFFFFF
BATCH_ORDINARY = 0xFFFFFFFF

# Header word offsets (mutracer-layout.h MT_HEADER_*).
H_VERSION, H_LENGTH, H_WIDTH, H_HEIGHT = 0, 1, 2, 3
H_GEOMETRY_COUNT, H_MESH_COUNT, H_MATERIAL_COUNT, H_LIGHT_COUNT, H_TEXTURE_COUNT = 4, 5, 6, 7, 8
H_GEOMETRY, H_MESH, H_MATERIAL, H_LIGHT, H_TEXTURE, H_CAMERA = 9, 10, 11, 12, 13, 14
H_BACKGROUND, H_ENV_TEXTURE, H_BG_TEXTURE, H_QUALITY, H_BATCH = 16, 20, 21, 22, 23
H_MESH_IDS, H_MATERIAL_IDS, H_ENV_SETTINGS, H_GROUNDED = 24, 25, 26, 27

GEOMETRY_WORDS, MESH_WORDS, MATERIAL_WORDS, LIGHT_WORDS, CAMERA_WORDS, TEXTURE_WORDS = 12, 21, 24, 32, 32, 13
BATCH_WORDS = 4

# Material extension ids (mutracer-material-properties.h) for the fields the
# basic 24-word material record does not carry.
EXT = dict(base_weight=1, base_diffuse_roughness=3, specular_weight=5, specular_color=6, specular_anisotropy=9,
           transmission_color=11, subsurface_weight=17, subsurface_color=18, fuzz_weight=22, fuzz_color=23,
           fuzz_roughness=24, coat_color=26, coat_ior=29, coat_darkening=30, double_sided=38)
_DEFAULT = S.Material()
# Texture slots carried as extension entries (the basic record holds map,
# normal, metallic-roughness and emissive), and the native enums.
EXT_MAPS = dict(alpha_map=40, roughness_map=47, metalness_map=48, transmission_map=49, coat_weight_map=64,
                coat_roughness_map=65, fuzz_color_map=66, fuzz_roughness_map=67, specular_weight_map=68,
                specular_color_map=69)
NATIVE_ALPHA_MODE = dict(auto=0, opaque=1, mask=2, blend=3)
NATIVE_MAP_ALPHA_USAGE = dict(opaque=0, alpha=1, transmission=2)

LIGHT_TYPE = {S.LIGHT_RECT: 0, S.LIGHT_SPHERE: 1, S.LIGHT_DIRECTIONAL: 2, S.LIGHT_SPOT: 3, S.LIGHT_DISK: 4,
              S.LIGHT_CYLINDER: 5}


def _f32(v):
    return float(np.float32(v))


def _bits(values):
    return np.asarray(values, np.float32).view(np.uint32)


class _Words:
    def __init__(self):
        self.w = [np.zeros(HEADER_WORDS, np.uint32)]
        self.n = HEADER_WORDS

    def append(self, arr):
        """Appends u32/f32 words; returns their word offset."""
        a = np.ascontiguousarray(arr)
        a = a.view(np.uint32) if a.dtype == np.float32 else a.astype(np.uint32)
        off = self.n
        self.w.append(a.reshape(-1))
        self.n += a.size
        return off

    def finish(self):
        out = np.concatenate(self.w)
        out[H_VERSION] = VERSION
        out[H_LENGTH] = out.size
        return out


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _matrix(x, y, z, t):
    """Column-major 4x4 with columns x, y, z, t (f32)."""
    m = np.zeros(16, np.float32)
    m[0:3], m[4:7], m[8:11], m[12:15] = x, y, z, t
    m[15] = 1.0
    return m


def _light_record(light):
    rec = np.zeros(LIGHT_WORDS, np.uint32)
    f = rec.view(np.float32)
    p = light.pack()
    kind = int(p[0])
    rec[0] = LIGHT_TYPE[kind]
    f[1:4] = light.color
    f[4] = light.intensity
    pos = getattr(light, "position", (0.0, 0.0, 0.0))
    if kind == S.LIGHT_RECT:
        u, v = np.asarray(light.u, np.float64), np.asarray(light.v, np.float64)
        x, y = _unit(u), -_unit(v)
        f[5], f[6], f[7] = np.linalg.norm(u), np.linalg.norm(v), light.spread
        f[12:28] = _matrix(x, y, np.cross(x, y), pos)
    elif kind == S.LIGHT_SPHERE:
        f[5] = light.radius
        f[12:28] = _matrix((1, 0, 0), (0, 1, 0), (0, 0, 1), pos)
    elif kind in (S.LIGHT_DIRECTIONAL, S.LIGHT_SPOT, S.LIGHT_DISK):
        z = -_unit(light.direction)
        x, y = S._frame(light.direction)
        f[12:28] = _matrix(x, y, z, pos)
        if kind == S.LIGHT_SPOT:
            f[5], f[6], f[7] = light.radius, light.inner_angle, light.outer_angle
        if kind == S.LIGHT_DISK:
            f[28] = light.radius
            rec[30] = int(light.normalize)
    elif kind == S.LIGHT_CYLINDER:
        a = _unit(light.axis)
        r = _unit(np.asarray(light.radial, np.float64) - a * np.dot(light.radial, a))
        f[12:28] = _matrix(a, r, np.cross(a, r), pos)
        f[28], f[29] = light.radius, light.length
        rec[30] = int(light.normalize)
    return rec


def _material_record(m, w):
    rec = np.zeros(MATERIAL_WORDS, np.uint32)
    f = rec.view(np.float32)
    f[0:3] = m.base_color
    f[3], f[4], f[5], f[6] = m.specular_roughness, m.base_metalness, m.transmission_weight, m.specular_ior
    em = np.asarray(m.emission, np.float64)
    if np.any(em != 0):
        if m.coat_weight > 0 and m.coat_weight_map is None:  # materialEmission multiplies the coat tint back in
            tint = 1.0 + min(m.coat_weight, 1.0) * (np.sqrt(np.maximum(np.asarray(m.coat_color, np.float64), 0)) - 1)
            em = em / tint
        f[7:10], f[10] = em, 1.0
    else:
        f[7:10], f[10] = (1.0, 1.0, 1.0), 0.0
    f[11], f[12], f[13] = m.coat_weight, m.coat_roughness, m.opacity
    rec[14] = int(bool(m.thin_walled))
    for k, name in enumerate(("map", "normal_map", "metallic_roughness_map", "emissive_map")):
        t = getattr(m, name)
        rec[15 + k] = NO_TEXTURE if t is None else int(t)
    ext = []

    def add(pid, u=None, f=None):
        e = np.zeros(4, np.uint32)
        e[0] = pid
        if u is not None:
            e[1] = u
        if f is not None:
            e.view(np.float32)[1:1 + len(f)] = f
        ext.append(e)
    for name, pid in EXT_MAPS.items():
        if getattr(m, name) is not None:
            add(pid, u=int(getattr(m, name)))
    if m.alpha_mode != "auto":
        add(41, u=NATIVE_ALPHA_MODE[m.alpha_mode])
    if m.alpha_cutoff != 0:
        add(42, f=[m.alpha_cutoff])
    if m.map_alpha_usage != "opaque":
        add(54, u=NATIVE_MAP_ALPHA_USAGE[m.map_alpha_usage])
    if tuple(m.normal_scale) != (1.0, 1.0):
        add(45, f=list(m.normal_scale))
    if tuple(m.uv_scale) != (1.0, 1.0):
        add(56, f=list(m.uv_scale))
    for name, pid in EXT.items():
        v, d = getattr(m, name), getattr(_DEFAULT, name)
        if v == d:
            continue
        e = np.zeros(4, np.uint32)
        e[0] = pid
        if isinstance(v, bool):
            e[1] = int(v)
        elif isinstance(v, tuple):
            e.view(np.float32)[1:4] = v
        else:
            e.view(np.float32)[1] = v
        ext.append(e)
    if ext:
        rec[19] = len(ext)
        rec[20] = w.append(np.concatenate(ext))
    return rec


def encode(scene, width, height):
    """The scene as a version-8 raw descriptor (bytes)."""
    w = _Words()
    geos = []
    for mesh in scene.meshes:
        g = np.zeros(GEOMETRY_WORDS, np.uint32)
        pos = np.asarray(mesh["positions"], np.float32)
        idx = np.asarray(mesh["indices"], np.uint32)
        g[0], g[1] = pos.size // 3, idx.size
        g[2] = w.append(pos)
        g[3] = w.append(idx)
        g[4] = w.append(np.asarray(mesh["normals"], np.float32))
        if mesh.get("uvs") is not None:
            g[5] = w.append(np.asarray(mesh["uvs"], np.float32))
        if mesh.get("uvs1") is not None:
            g[6] = w.append(np.asarray(mesh["uvs1"], np.float32))
        geos.append(g)
    meshes = []
    for i in range(len(scene.mesh_index)):
        r = np.zeros(MESH_WORDS, np.uint32)
        r[0], r[1] = scene.mesh_index[i], scene.instance_material[i]
        r[4:20] = _bits(np.asarray(scene.transforms[i], np.float64))
        meshes.append(r)
    mats = [_material_record(m, w) for m in scene.materials]
    lights = [_light_record(l) for l in scene.lights]
    textures = []
    for tex in scene.textures:  # material textures first: the slots index them directly
        px = tex.storage_pixels()
        t = np.zeros(TEXTURE_WORDS, np.uint32)
        t[0], t[1], t[2] = px.shape[1], px.shape[0], tex.storage
        t[3] = w.append(px.reshape(-1).view(np.uint32))
        t[4], t[5] = int(tex.flip_y), int(tex.uv_channel)
        t.view(np.float32)[6:10] = [*tex.uv_scale, *tex.uv_offset]
        t[10], t[11] = tex.wrap_code(tex.wrap_s), tex.wrap_code(tex.wrap_t)
        textures.append(t)
    if scene.env is not None:
        env = np.ascontiguousarray(scene.env, np.float32)
        t = np.zeros(TEXTURE_WORDS, np.uint32)
        t[0], t[1], t[2] = env.shape[1], env.shape[0], 1
        t[3] = w.append(env)
        t.view(np.float32)[6:8] = 1.0
        textures.append(t)
    cam = np.zeros(CAMERA_WORDS, np.uint32)
    cf = cam.view(np.float32)
    c = scene.camera
    _, vi = c.matrices(width, height)
    cf[1], cf[2], cf[3] = c.fov, c.near, c.far
    cf[4], cf[5], cf[6], cf[7], cf[8] = -1.0, 1.0, -1.0, 1.0, 1.0
    cf[10], cf[11], cf[12] = c.focus_distance, 24.0, 1.0
    if c.aperture > 0:  # cameraApertureRadius inverted for metres_per_unit 1
        focal = 0.5 * 24.0 / math.tan(_f32(c.fov) * math.pi / 360.0)
        cf[9] = focal / (2.0 * c.aperture * 1000.0)
    cf[16:32] = vi
    hdr = w.w[0]
    hdr[H_WIDTH], hdr[H_HEIGHT] = width, height
    hdr[H_GEOMETRY_COUNT], hdr[H_MESH_COUNT] = len(geos), len(meshes)
    hdr[H_MATERIAL_COUNT], hdr[H_LIGHT_COUNT], hdr[H_TEXTURE_COUNT] = len(mats), len(lights), len(textures)
    hdr[H_GEOMETRY] = w.append(np.concatenate(geos))
    hdr[H_MESH] = w.append(np.concatenate(meshes))
    hdr[H_MATERIAL] = w.append(np.concatenate(mats))
    hdr[H_LIGHT] = w.append(np.concatenate(lights)) if lights else HEADER_WORDS
    hdr[H_TEXTURE] = w.append(np.concatenate(textures)) if textures else HEADER_WORDS
    hdr[H_CAMERA] = w.append(cam)
    if scene.env is not None:
        hdr[H_ENV_TEXTURE] = hdr[H_BG_TEXTURE] = len(textures)
    return w.finish().tobytes()


def f32_scene(scene):
    """The Scene that `encode(scene)` describes: instance transforms rounded to
    f32 (lights and camera are rounded where the library packs them anyway)."""
    out = copy.copy(scene)
    out.transforms = [np.asarray(t, np.float32).astype(np.float64) for t in scene.transforms]
    return out


def header(data):
    """Decoded header words of a descriptor (for tests and tools)."""
    return np.frombuffer(data, np.uint32, HEADER_WORDS)
