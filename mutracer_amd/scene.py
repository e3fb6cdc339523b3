"""Host-side scene description and the procedural benchmark scenes.

A Scene holds plain numpy data: meshes (positions, normals, indices),
instances (mesh, column-major f64 transform, material), OpenPBR materials,
rect lights, an optional equirectangular RGBA f32 environment and a camera.
Its fields mirror the reference's scene inputs (OpenPBR parameters of
src/scene/openPbrSchema.ts, RectLight of src/scene/RectLight.ts, Camera of
src/scene/Camera.ts); the packing helpers produce the arrays the C ABI
(include/mtx.h) consumes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Packed material layout (MTX_MATERIAL_FLOATS = 64 floats), see include/mtx.h.
MATERIAL_FLOATS = 64
MAT_FIELDS = dict(base_weight=0, base_color=1, base_diffuse_roughness=4, base_metalness=5, specular_weight=6,
                  specular_color=7, specular_roughness=10, specular_ior=11, transmission_weight=12,
                  transmission_color=13, coat_weight=16, coat_color=17, coat_roughness=20, coat_ior=21,
                  coat_darkening=22, emission=23, double_sided=26, specular_anisotropy=27, fuzz_weight=28,
                  fuzz_color=29, fuzz_roughness=32, subsurface_weight=33, subsurface_color=34, thin_walled=37,
                  opacity=38, alpha_cutoff=40, normal_scale=42, uv_scale=44)
MAT_ALPHA_MODE, MAT_MAP_ALPHA_USAGE, MAT_MAPS = 39, 41, 46
# Texture slots in packing order (MTX_MAP_*; materialTextureSlots.ts:8-140).
MAP_SLOTS = ("map", "alpha_map", "transmission_map", "metallic_roughness_map", "roughness_map", "metalness_map",
             "coat_weight_map", "coat_roughness_map", "fuzz_color_map", "fuzz_roughness_map", "specular_weight_map",
             "specular_color_map", "emissive_map", "normal_map")
ALPHA_MODES = dict(opaque=0, mask=1, blend=2)               # materialConstants.ts:12-16
MAP_ALPHA_USAGE = dict(opaque=0, alpha=2, transmission=3)   # materialConstants.ts:2
WRAP_MODES = {"repeat": 0, "clamp-to-edge": 1, "mirrored-repeat": 2}


@dataclass
class Material:
    """OpenPBR parameters with the reference defaults (openPbrSchema.ts:32-69)."""
    base_weight: float = 1.0
    base_color: tuple = (0.8, 0.8, 0.8)
    base_diffuse_roughness: float = 0.0
    base_metalness: float = 0.0
    specular_weight: float = 1.0
    specular_color: tuple = (1.0, 1.0, 1.0)
    specular_roughness: float = 0.3
    specular_ior: float = 1.5
    specular_anisotropy: float = 0.0
    transmission_weight: float = 0.0
    transmission_color: tuple = (1.0, 1.0, 1.0)
    coat_weight: float = 0.0
    coat_color: tuple = (1.0, 1.0, 1.0)
    coat_roughness: float = 0.0
    coat_ior: float = 1.6
    coat_darkening: float = 1.0
    fuzz_weight: float = 0.0
    fuzz_color: tuple = (1.0, 1.0, 1.0)
    fuzz_roughness: float = 0.5
    subsurface_weight: float = 0.0
    subsurface_color: tuple = (0.8, 0.8, 0.8)
    thin_walled: bool = False
    emission: tuple = (0.0, 0.0, 0.0)   # materialEmission(): color * luminance, coat-tinted unless coat_weight_map
    double_sided: bool = False
    opacity: float = 1.0
    alpha_mode: str = "auto"             # auto, opaque, mask, blend
    alpha_cutoff: float = 0.0
    map_alpha_usage: str = "opaque"      # what the base map's alpha drives: opaque, alpha, transmission
    normal_scale: tuple = (1.0, 1.0)
    uv_scale: tuple = (1.0, 1.0)         # uvProjectionScale
    # Texture slots: scene texture indices or None.
    map: int | None = None
    alpha_map: int | None = None
    transmission_map: int | None = None
    metallic_roughness_map: int | None = None
    roughness_map: int | None = None
    metalness_map: int | None = None
    coat_weight_map: int | None = None
    coat_roughness_map: int | None = None
    fuzz_color_map: int | None = None
    fuzz_roughness_map: int | None = None
    specular_weight_map: int | None = None
    specular_color_map: int | None = None
    emissive_map: int | None = None
    normal_map: int | None = None

    def resolved_alpha_mode(self):
        """resolveMaterialAlphaMode (MaterialState.ts:251-264)."""
        if self.alpha_mode != "auto":
            return self.alpha_mode
        if self.alpha_cutoff > 0:
            return "mask"
        if self.opacity < 1 or self.map_alpha_usage == "alpha" or self.alpha_map is not None:
            return "blend"
        return "opaque"

    def pack(self):
        p = np.zeros(MATERIAL_FLOATS, np.float32)
        for k, off in MAT_FIELDS.items():
            v = getattr(self, k)
            if isinstance(v, tuple):
                p[off:off + len(v)] = v
            else:
                p[off] = float(v)
        p[MAT_ALPHA_MODE] = ALPHA_MODES[self.resolved_alpha_mode()]
        p[MAT_MAP_ALPHA_USAGE] = MAP_ALPHA_USAGE[self.map_alpha_usage]
        for k, name in enumerate(MAP_SLOTS):
            t = getattr(self, name)
            p[MAT_MAPS + k] = 0 if t is None else int(t) + 1
        return p

    def has_maps(self):
        return any(getattr(self, name) is not None for name in MAP_SLOTS)


@dataclass
class Texture:
    """A material texture (mt_texture_v4, mutracer-scene.h:87-95, without UDIM).

    pixels: (H, W, 4) uint8 (RGBA8: color slots decode it from sRGB, data
    slots read it as linear) or float32 (RGBA32F, linear); row 0 is sampled
    at v = 0 unless flip_y. A sample at uv reads uv * uv_scale + uv_offset,
    wrapped per axis (repeat, clamp-to-edge, mirrored-repeat), bilinearly.
    """
    pixels: np.ndarray
    flip_y: bool = False
    uv_channel: int = 0
    uv_scale: tuple = (1.0, 1.0)
    uv_offset: tuple = (0.0, 0.0)
    wrap_s: str = "repeat"
    wrap_t: str = "repeat"

    @property
    def storage(self):
        return 1 if np.asarray(self.pixels).dtype == np.float32 else 0

    def storage_pixels(self):
        """Row-major pixels as the C ABI takes them: one little-endian word
        per RGBA8 pixel (R in the low byte) or four floats per pixel."""
        px = np.asarray(self.pixels)
        if px.dtype == np.float32:
            return np.ascontiguousarray(px.reshape(px.shape[0], px.shape[1], 4))
        px = np.ascontiguousarray(px.astype(np.uint8).reshape(px.shape[0], px.shape[1], 4))
        return px.view("<u4").reshape(px.shape[0], px.shape[1])

    @staticmethod
    def wrap_code(mode):
        return WRAP_MODES[mode] if isinstance(mode, str) else int(mode)


@dataclass
class RectLight:
    """Rect area light: center, edge vectors u, v (normal = normalize(u x v)),
    radiance = color * intensity (RectLight.ts:5-22), spread 1 = Lambertian."""
    position: tuple
    u: tuple
    v: tuple
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    spread: float = 1.0

    def pack(self):
        p = np.zeros(16, np.float32)
        p[0] = LIGHT_RECT
        p[1:4] = self.position
        p[4:7] = self.u
        p[7:10] = self.v
        p[10:13] = np.asarray(self.color, np.float64) * self.intensity
        p[13] = self.spread
        p[14] = np.linalg.norm(np.cross(self.u, self.v))
        return p


# Analytic light types (lightConstants.ts:10-15).
LIGHT_DIRECTIONAL, LIGHT_SPHERE, LIGHT_RECT, LIGHT_SPOT, LIGHT_DISK, LIGHT_CYLINDER = 1, 2, 3, 4, 5, 6


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _frame(direction):
    """Unit x, y axes perpendicular to `direction` with x cross y = -direction
    (the reference's rigid light frame, lightsData.ts:103-116, with z = -direction)."""
    z = -_unit(direction)
    x = np.array([z[2], 0.0, -z[0]]) if abs(z[1]) < 0.99 else np.array([1.0, 0.0, 0.0])
    x = _unit(x - z * np.dot(x, z))
    return x, np.cross(z, x)


@dataclass
class SphereLight:
    """Sphere light of `radius` (0: point light with intensity in candela);
    emission = color * intensity (SphereLight, lightsData.ts:57-72)."""
    position: tuple
    radius: float = 0.0
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0

    def pack(self):
        p = np.zeros(16, np.float32)
        p[0] = LIGHT_SPHERE
        p[1:4] = self.position
        p[10:13] = np.asarray(self.color, np.float64) * self.intensity
        p[13] = self.radius
        p[14] = 4 * math.pi * self.radius ** 2
        return p


@dataclass
class SpotLight:
    """Spot light toward `direction` with a smooth cone between the inner and
    outer half-angles (degrees); radius 0 is a point emitter (lightsData.ts
    :80-92, getSpotLightRadiance :145-149)."""
    position: tuple
    direction: tuple
    radius: float = 0.0
    inner_angle: float = 30.0
    outer_angle: float = 45.0
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0

    def pack(self):
        p = np.zeros(16, np.float32)
        p[0] = LIGHT_SPOT
        p[1:4] = self.position
        p[4:7] = _unit(self.direction)
        p[7] = math.cos(math.radians(self.outer_angle))
        p[8] = math.cos(math.radians(min(self.inner_angle, self.outer_angle)))
        radiance = self.intensity if self.radius == 0 else self.intensity / (math.pi * self.radius ** 2)
        p[10:13] = np.asarray(self.color, np.float64) * radiance
        p[13] = self.radius
        p[14] = 4 * math.pi * self.radius ** 2
        return p


@dataclass
class DirectionalLight:
    """Directional light travelling along `direction`, irradiance = intensity."""
    direction: tuple
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0

    def pack(self):
        p = np.zeros(16, np.float32)
        p[0] = LIGHT_DIRECTIONAL
        p[4:7] = _unit(self.direction)
        p[10:13] = np.asarray(self.color, np.float64) * self.intensity
        return p


@dataclass
class DiskLight:
    """One-sided disk light emitting toward `direction` (lightsData.ts:73-77);
    with `normalize` the intensity is spread over the area."""
    position: tuple
    direction: tuple
    radius: float = 0.5
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    normalize: bool = False

    def pack(self):
        x, y = _frame(self.direction)
        area = math.pi * self.radius ** 2
        p = np.zeros(16, np.float32)
        p[0] = LIGHT_DISK
        p[1:4] = self.position
        p[4:7] = x * self.radius
        p[7:10] = y * -self.radius
        p[10:13] = np.asarray(self.color, np.float64) * (self.intensity / max(area, 1e-8) if self.normalize
                                                        else self.intensity)
        p[13] = self.radius
        p[14] = area
        return p


@dataclass
class CylinderLight:
    """Cylinder light along `axis` (length) with a `radial` reference
    direction (lightsData.ts:78-82); emits outward from its side."""
    position: tuple
    axis: tuple
    radial: tuple
    radius: float = 0.5
    length: float = 1.0
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    normalize: bool = False

    def pack(self):
        a = _unit(self.axis)
        r = _unit(np.asarray(self.radial, np.float64) - a * np.dot(self.radial, a))
        area = 2 * math.pi * self.radius * self.length
        p = np.zeros(16, np.float32)
        p[0] = LIGHT_CYLINDER
        p[1:4] = self.position
        p[4:7] = a * self.length
        p[7:10] = r * self.radius
        p[10:13] = np.asarray(self.color, np.float64) * (self.intensity / max(area, 1e-8) if self.normalize
                                                        else self.intensity)
        p[13] = self.radius
        p[14] = area
        return p


@dataclass
class Camera:
    eye: tuple
    target: tuple
    up: tuple = (0.0, 1.0, 0.0)
    fov: float = 45.0          # vertical, degrees (Camera.ts:43-44)
    near: float = 0.1
    far: float = 2000.0
    aperture: float = 0.0
    focus_distance: float = 1.0

    def matrices(self, width, height):
        """(projection inverse, view inverse) as column-major f32, computed in f64.

        fov, near and far enter as the f32 values a scene descriptor carries,
        and the perspective matrix (mat4.perspective: f/aspect, f,
        A = (far+near)/(near-far), B = 2 far near/(near-far), -1) is inverted
        in closed form, the same arithmetic as the library's descriptor decoder,
        so a committed descriptor and this camera give the same bits."""
        aspect = width / height
        fov, near, far = (float(np.float32(v)) for v in (self.fov, self.near, self.far))
        f = 1.0 / math.tan(fov * math.pi / 180.0 / 2.0)
        nf = 1.0 / (near - far)
        A = (far + near) * nf
        B = 2.0 * far * near * nf
        proj_inv = np.zeros((4, 4))
        proj_inv[0, 0] = aspect / f
        proj_inv[1, 1] = 1.0 / f
        proj_inv[3, 2] = 1.0 / B
        proj_inv[2, 3] = -1.0
        proj_inv[3, 3] = A / B
        eye, target, up = (np.asarray(a, np.float64) for a in (self.eye, self.target, self.up))
        z = eye - target
        z /= np.linalg.norm(z)
        x = np.cross(up, z)
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        world = np.eye(4)
        world[:3, 0], world[:3, 1], world[:3, 2], world[:3, 3] = x, y, z, eye
        return (proj_inv.T.reshape(16).astype(np.float32), world.T.reshape(16).astype(np.float32))

    def pack(self, width, height):
        pi, vi = self.matrices(width, height)
        out = np.zeros(36, np.float32)
        out[0:16] = pi
        out[16:32] = vi
        out[32] = self.aperture
        out[33] = self.focus_distance
        return out


@dataclass
class Scene:
    meshes: list = field(default_factory=list)        # dict(positions, normals, indices)
    mesh_index: list = field(default_factory=list)
    transforms: list = field(default_factory=list)    # 16 f64 column-major
    instance_material: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    lights: list = field(default_factory=list)
    env: np.ndarray | None = None                      # (H, W, 4) f32
    camera: Camera | None = None
    textures: list = field(default_factory=list)      # Texture, indexed by the materials' map slots

    def add_mesh(self, positions, indices, normals=None, uvs=None, uvs1=None):
        positions = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        indices = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        if normals is None:
            normals = vertex_normals(positions, indices)
        mesh = dict(positions=positions, indices=indices,
                    normals=np.ascontiguousarray(normals, np.float32).reshape(-1, 3))
        for key, value in (("uvs", uvs), ("uvs1", uvs1)):
            if value is not None:
                mesh[key] = np.ascontiguousarray(value, np.float32).reshape(-1, 2)
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_instance(self, mesh, transform, material):
        self.mesh_index.append(mesh)
        self.transforms.append(np.asarray(transform, np.float64).reshape(16))
        self.instance_material.append(material)

    @property
    def triangle_count(self):
        return sum(self.meshes[m]["indices"].size // 3 for m in self.mesh_index)

    def packed_materials(self):
        return np.stack([m.pack() for m in self.materials]).astype(np.float32)

    def packed_lights(self):
        return (np.stack([l.pack() for l in self.lights]).astype(np.float32) if self.lights
                else np.zeros((0, 16), np.float32))

    def double_sided(self):
        return [1 if self.materials[m].double_sided else 0 for m in self.instance_material]


def vertex_normals(positions, indices):
    p = positions.astype(np.float64)
    tri = indices.reshape(-1, 3)
    fn = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, tri[:, k], fn)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    ln[ln == 0] = 1
    return (n / ln).astype(np.float32)


def translate(t):
    m = np.eye(4)
    m[:3, 3] = t
    return m.T.reshape(16).copy()


def rot_y(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    m = np.eye(4)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, s, -s, c
    return m


# ---------------------------------------------------------------- config 1
def quad(a, b, c, d):
    """Two triangles a-b-c, a-c-d with the flat normal of a-b-c."""
    p = np.array([a, b, c, d], np.float32)
    n = np.cross(p[1] - p[0], p[2] - p[0])
    n = (n / np.linalg.norm(n)).astype(np.float32)
    return p, np.array([0, 1, 2, 0, 2, 3], np.uint32), np.tile(n, (4, 1))


def box_mesh(sx, sy, sz):
    """Axis-aligned box [−sx/2,sx/2]×[0,sy]×[−sz/2,sz/2], outward flat faces (12 tris)."""
    x, y, z = sx / 2, sy, sz / 2
    faces = [
        [(-x, 0, z), (x, 0, z), (x, y, z), (-x, y, z)],      # +z
        [(x, 0, -z), (-x, 0, -z), (-x, y, -z), (x, y, -z)],  # -z
        [(x, 0, z), (x, 0, -z), (x, y, -z), (x, y, z)],      # +x
        [(-x, 0, -z), (-x, 0, z), (-x, y, z), (-x, y, -z)],  # -x
        [(-x, y, z), (x, y, z), (x, y, -z), (-x, y, -z)],    # +y
        [(-x, 0, -z), (x, 0, -z), (x, 0, z), (-x, 0, z)],    # -y
    ]
    P, I, N = [], [], []
    for f in faces:
        p, i, n = quad(*f)
        I.append(i + 4 * len(P))
        P.append(p)
        N.append(n)
    return np.concatenate(P), np.concatenate(I), np.concatenate(N)


def cornell_box(width=256, height=256):
    """Config 1: box [-1,1]x[0,2]x[-1,1] (5 quads), two rotated boxes, one
    0.25x0.25 rect light of radiance 15 just under the ceiling. 34 triangles."""
    s = Scene()
    white = Material(base_color=(0.73, 0.73, 0.73))
    red = Material(base_color=(0.65, 0.05, 0.05))
    green = Material(base_color=(0.12, 0.45, 0.15))
    s.materials = [white, red, green]
    walls = [
        ([(-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1)], 0),     # floor (+y)
        ([(-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)], 0),     # ceiling (-y)
        ([(-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)], 0),   # back (+z)
        ([(-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1)], 1),   # left red (+x)
        ([(1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)], 2),       # right green (-x)
    ]
    for pts, mat in walls:
        p, i, n = quad(*pts)
        s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), mat)
    short = box_mesh(0.6, 0.6, 0.6)
    tall = box_mesh(0.6, 1.2, 0.6)
    m_short = rot_y(-18)
    m_short[:3, 3] = (0.33, 0.0, 0.3)
    m_tall = rot_y(20)
    m_tall[:3, 3] = (-0.33, 0.0, -0.3)
    s.add_instance(s.add_mesh(*short), m_short.T.reshape(16), 0)
    s.add_instance(s.add_mesh(*tall), m_tall.T.reshape(16), 0)
    s.lights = [RectLight(position=(0.0, 1.999, 0.0), u=(0.25, 0.0, 0.0), v=(0.0, 0.0, 0.25), intensity=15.0)]
    s.camera = Camera(eye=(0.0, 1.0, 3.94), target=(0.0, 1.0, 0.0), fov=39.3)
    return s


QUAD_UVS = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)


def procedural_textures(seed=7):
    """Small procedural images for textured scenes: an sRGB checker (RGBA8),
    metallic-roughness data (RGBA32F: g roughness ramp, b metal stripes), a
    bump normal map (RGBA32F, tangent space), an sRGB emissive pattern
    (RGBA8) and a coat-weight / transmission mask (RGBA8)."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:64, 0:64]
    checker = np.where(((x // 8) + (y // 8)) % 2 == 0, 1.0, 0.0)[..., None]
    c0 = np.array([230, 200, 120]) * checker + np.array([40, 70, 160]) * (1 - checker)
    checker_rgba = np.concatenate([c0, 255 * np.ones((64, 64, 1))], 2).astype(np.uint8)
    yy, xx = np.mgrid[0:32, 0:32]
    mr = np.zeros((32, 32, 4), np.float32)
    mr[..., 1] = 0.15 + 0.8 * xx / 31.0
    mr[..., 2] = ((yy // 4) % 2).astype(np.float32)
    mr[..., 3] = 1.0
    hy, hx = np.mgrid[0:64, 0:64] / 64.0
    h = np.sin(2 * np.pi * 3 * hx) * np.cos(2 * np.pi * 2 * hy)
    gx, gy = np.gradient(h)
    n = np.stack([-2.0 * gx * 8, -2.0 * gy * 8, np.ones_like(h)], 2)
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    normal_map = np.concatenate([n * 0.5 + 0.5, np.ones((64, 64, 1))], 2).astype(np.float32)
    em = rng.integers(0, 256, (16, 16, 4)).astype(np.uint8)
    em[..., 3] = 255
    my, mx = np.mgrid[0:24, 0:24]
    mask = np.zeros((24, 24, 4), np.uint8)
    mask[..., 0] = (255 * ((mx + my) % 6 < 3)).astype(np.uint8)
    mask[..., 1] = (40 + 8 * mx).astype(np.uint8)
    mask[..., 2] = (255 * (my > 11)).astype(np.uint8)
    mask[..., 3] = (60 + 8 * my).astype(np.uint8)
    return [Texture(checker_rgba, uv_scale=(4.0, 4.0), uv_offset=(0.125, 0.0)),
            Texture(mr, wrap_s="clamp-to-edge", wrap_t="mirrored-repeat"),
            Texture(normal_map, wrap_s="mirrored-repeat", uv_scale=(1.5, 1.0)),
            Texture(em, flip_y=True, uv_channel=1),
            Texture(mask, wrap_s="clamp-to-edge", wrap_t="clamp-to-edge", uv_scale=(1.2, 1.2),
                    uv_offset=(-0.1, -0.1))]


def textured_room(width=256, height=256):
    """The Cornell geometry with material textures on every slot kind: a
    checker base map on the floor, a bump normal map on the back wall, a
    metallic-roughness map on the tall box, a transmission / coat-weight /
    specular-color / fuzz-color mask on the short box, and a ceiling panel
    whose emission is textured (an emissive-triangle light with an emissive
    map and a coat weight map, sampled on UV1), besides the rect light."""
    s = Scene()
    s.textures = procedural_textures()
    floor = Material(base_color=(0.9, 0.9, 0.9), specular_roughness=0.6, map=0)
    wall = Material(base_color=(0.73, 0.73, 0.73), normal_map=2, normal_scale=(0.8, 1.2), uv_scale=(2.0, 1.0))
    red = Material(base_color=(0.65, 0.05, 0.05), roughness_map=1)
    green = Material(base_color=(0.12, 0.45, 0.15), specular_color_map=0, specular_weight_map=4)
    metal = Material(base_color=(0.9, 0.7, 0.5), base_metalness=1.0, specular_roughness=1.0,
                     metallic_roughness_map=1, metalness_map=1, coat_weight=0.4, coat_roughness_map=4)
    short = Material(base_color=(0.8, 0.8, 0.8), transmission_weight=0.8, transmission_map=4, specular_roughness=0.2,
                     fuzz_weight=0.5, fuzz_color_map=0, fuzz_roughness_map=4, coat_weight=0.5, coat_weight_map=4,
                     map=4, map_alpha_usage="transmission")
    panel = Material(base_color=(0.2, 0.2, 0.2), emission=(3.0, 2.5, 2.0), emissive_map=3, coat_weight=0.6,
                     coat_color=(0.9, 0.6, 0.3), coat_weight_map=4, double_sided=True)
    s.materials = [floor, wall, red, green, metal, short, panel]
    walls = [
        ([(-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1)], 0),
        ([(-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)], 1),
        ([(-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)], 1),
        ([(-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1)], 2),
        ([(1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)], 3),
    ]
    for pts, mat in walls:
        p, i, n = quad(*pts)
        s.add_instance(s.add_mesh(p, i, n, uvs=QUAD_UVS), np.eye(4).reshape(16), mat)
    box_uv = np.tile(QUAD_UVS, (6, 1))
    m_short = rot_y(-18)
    m_short[:3, 3] = (0.33, 0.0, 0.3)
    m_tall = rot_y(20)
    m_tall[:3, 3] = (-0.33, 0.0, -0.3)
    s.add_instance(s.add_mesh(*box_mesh(0.6, 0.6, 0.6), uvs=box_uv, uvs1=box_uv[:, ::-1] * 0.5),
                   m_short.T.reshape(16), 5)
    s.add_instance(s.add_mesh(*box_mesh(0.6, 1.2, 0.6), uvs=box_uv * 2.0), m_tall.T.reshape(16), 4)
    p, i, n = quad((-0.6, 1.95, 0.5), (-0.2, 1.95, 0.5), (-0.2, 1.95, 0.1), (-0.6, 1.95, 0.1))
    s.add_instance(s.add_mesh(p, i, n, uvs=QUAD_UVS, uvs1=QUAD_UVS * 1.5 - 0.25), np.eye(4).reshape(16), 6)
    s.lights = [RectLight(position=(0.3, 1.999, -0.2), u=(0.25, 0.0, 0.0), v=(0.0, 0.0, 0.25), intensity=12.0)]
    s.camera = Camera(eye=(0.0, 1.0, 3.94), target=(0.0, 1.0, 0.0), fov=39.3)
    return s


# ---------------------------------------------------------------- config 2
def _gradient_sky(w=16, h=8):
    """Small RGBA32F environment: a blue-to-white vertical gradient."""
    t = (np.arange(h, dtype=np.float32) + 0.5) / h
    env = np.zeros((h, w, 4), np.float32)
    env[..., 0] = (0.3 + 0.7 * t)[:, None]
    env[..., 1] = (0.5 + 0.5 * t)[:, None]
    env[..., 2] = 1.0
    env[..., 3] = 1.0
    return env


ALPHA_CASTERS = ("absent", "opaque", "mask_covered", "mask_open", "alpha_map_covered", "alpha_map_open",
                 "base_alpha_covered", "base_alpha_open", "checker")


def alpha_caster(kind):
    """A floor with a box under a downward rect light and a sky environment;
    a quad at y = 2 (outside the camera's view, with nothing above it but the
    light) casts the shadow. `kind` picks its material: absent, opaque, mask
    with opacity at / just below the cutoff, mask with an alpha map whose g
    is above / below the cutoff, mask with the base map's alpha as coverage
    above / below it, and a checker alpha map (half the texels covered).
    Every variant carries the same textures (constant RGBA32F images and an
    RGBA8 checker)."""
    s = Scene()
    s.textures = [Texture(np.tile(np.array([0.1, 0.9, 0.2, 0.3], np.float32), (2, 2, 1))),
                  Texture(np.tile(np.array([0.9, 0.3, 0.9, 0.9], np.float32), (2, 2, 1))),
                  Texture(np.tile(np.array([1.0, 1.0, 1.0, 0.9], np.float32), (2, 2, 1))),
                  Texture(np.tile(np.array([1.0, 1.0, 1.0, 0.3], np.float32), (2, 2, 1))),
                  Texture(np.where(((np.arange(8)[:, None] + np.arange(8)[None, :]) % 2)[..., None] == 0,
                                   np.array([255, 255, 255, 255], np.uint8), np.array([255, 0, 255, 255], np.uint8)),
                          uv_scale=(1.0, 1.0))]
    s.materials = [Material(base_color=(0.8, 0.8, 0.8)), Material(base_color=(0.7, 0.2, 0.1))]
    p, i, n = quad((-3, 0, 3), (3, 0, 3), (3, 0, -3), (-3, 0, -3))
    s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), 0)
    s.add_instance(s.add_mesh(*box_mesh(0.6, 0.5, 0.6)), translate((0.2, 0.0, -0.1)), 1)
    s.lights = [RectLight(position=(0.0, 3.0, 0.0), u=(0.5, 0.0, 0.0), v=(0.0, 0.0, 0.5), intensity=20.0)]
    s.env = _gradient_sky()
    s.camera = Camera(eye=(0.0, 1.5, 4.0), target=(0.0, 0.3, 0.0), fov=40.0)
    if kind == "absent":
        return s
    caster = dict(opaque=Material(base_color=(0.5, 0.5, 0.5)),
                  mask_covered=Material(base_color=(0.5, 0.5, 0.5), alpha_mode="mask", opacity=0.5, alpha_cutoff=0.5),
                  mask_open=Material(base_color=(0.5, 0.5, 0.5), alpha_mode="mask", opacity=0.4999,
                                     alpha_cutoff=0.5),
                  alpha_map_covered=Material(base_color=(0.5, 0.5, 0.5), alpha_mode="mask", alpha_cutoff=0.5,
                                             alpha_map=0),
                  alpha_map_open=Material(base_color=(0.5, 0.5, 0.5), alpha_mode="mask", alpha_cutoff=0.5,
                                          alpha_map=1),
                  base_alpha_covered=Material(base_color=(0.5, 0.5, 0.5), alpha_mode="mask", alpha_cutoff=0.5,
                                              map=2, map_alpha_usage="alpha"),
                  base_alpha_open=Material(base_color=(0.5, 0.5, 0.5), alpha_mode="mask", alpha_cutoff=0.5,
                                           map=3, map_alpha_usage="alpha"),
                  checker=Material(base_color=(0.5, 0.5, 0.5), alpha_mode="mask", alpha_cutoff=0.5, alpha_map=4))[kind]
    s.materials.append(caster)
    p, i, n = quad((-0.6, 2.0, 0.6), (0.6, 2.0, 0.6), (0.6, 2.0, -0.6), (-0.6, 2.0, -0.6))
    s.add_instance(s.add_mesh(p, i, n, uvs=QUAD_UVS), np.eye(4).reshape(16), len(s.materials) - 1)
    return s


def masked_panes(panes, checker=False):
    """A floor under a sky environment (no light, nothing else) seen through
    `panes` camera-facing quads of a mask material with opacity below its
    cutoff (fully passed through), or, with checker, of a checker alpha map
    (half the texels covered) in front of a box on the floor."""
    s = Scene()
    s.textures = [Texture(np.where(((np.arange(8)[:, None] + np.arange(8)[None, :]) % 2)[..., None] == 0,
                                   np.array([255, 255, 255, 255], np.uint8), np.array([255, 0, 255, 255], np.uint8)))]
    s.materials = [Material(base_color=(0.8, 0.7, 0.6)),
                   Material(base_color=(0.3, 0.6, 0.9), alpha_mode="mask", opacity=0.2, alpha_cutoff=0.5),
                   Material(base_color=(0.3, 0.6, 0.9), alpha_mode="mask", alpha_cutoff=0.5, alpha_map=0),
                   Material(base_color=(0.7, 0.2, 0.1))]
    p, i, n = quad((-4, 0, 4), (4, 0, 4), (4, 0, -4), (-4, 0, -4))
    s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), 0)
    if checker:
        s.add_instance(s.add_mesh(*box_mesh(0.8, 0.8, 0.8)), translate((0.0, 0.0, -0.5)), 3)
    for k in range(panes):
        z = 2.5 - 0.4 * k
        p, i, n = quad((-3, -1, z), (3, -1, z), (3, 4, z), (-3, 4, z))
        s.add_instance(s.add_mesh(p, i, n, uvs=QUAD_UVS), np.eye(4).reshape(16), 2 if checker else 1)
    s.env = _gradient_sky()
    s.camera = Camera(eye=(0.0, 1.2, 4.0), target=(0.0, 0.2, 0.0), fov=40.0)
    return s


def icosphere(freq):
    """Geodesic (class I) icosphere: 20 * freq^2 triangles, shared vertices."""
    t = (1 + 5 ** 0.5) / 2
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    keys, pts = {}, []
    tris = []

    def vid(p):
        k = tuple(np.round(p, 9))
        if k not in keys:
            keys[k] = len(pts)
            pts.append(p)
        return keys[k]

    for a, b, c in F:
        A, B, C = V[a], V[b], V[c]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = A + (B - A) * (i / freq) + (C - A) * (j / freq)
                grid[i, j] = vid(p / np.linalg.norm(p))
        for i in range(freq):
            for j in range(freq - i):
                tris.append([grid[i, j], grid[i + 1, j], grid[i, j + 1]])
                if i + j + 1 < freq:
                    tris.append([grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]])
    return np.array(pts, np.float64), np.array(tris, np.uint32).reshape(-1)


def _hash3(ix, iy, iz, seed):
    h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
    h = h & 0xFFFFFFFF
    h ^= h >> 13
    h = (h * 0x5BD1E995) & 0xFFFFFFFF
    h ^= h >> 15
    return h / 4294967295.0


def value_noise(p, seed):
    q = np.floor(p).astype(np.int64)
    f = p - q
    u = f * f * (3 - 2 * f)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (np.where(dx, u[:, 0], 1 - u[:, 0]) * np.where(dy, u[:, 1], 1 - u[:, 1]) *
                     np.where(dz, u[:, 2], 1 - u[:, 2]))
                out += w * _hash3(q[:, 0] + dx, q[:, 1] + dy, q[:, 2] + dz, seed)
    return out * 2 - 1


def fbm_displaced_sphere(base_pts, indices, seed, amplitude=0.1, octaves=4, lacunarity=2.0, gain=0.5, freq0=2.0):
    disp = np.zeros(base_pts.shape[0])
    amp, fr = 1.0, freq0
    for _ in range(octaves):
        disp += amp * value_noise(base_pts * fr, seed)
        amp *= gain
        fr *= lacunarity
    pos = base_pts * (1.0 + amplitude * disp)[:, None]
    pos = pos.astype(np.float32)
    return pos, indices, vertex_normals(pos, indices)


def sky_env(width=1024, height=512, sun_dir=(0.45, 0.57, -0.69), sun_radius_deg=0.5, sun_scale=1e4):
    """Sky gradient plus a sun lobe (radiance 1e4 x sky) of 0.5 degree radius."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * math.pi
    phi = u * 2 * math.pi
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    d = np.stack([-st * np.cos(phi)[None], np.broadcast_to(ct, (height, width)), -st * np.sin(phi)[None]], -1)
    up = np.clip(d[..., 1], 0, 1)
    horizon = np.array([0.9, 0.85, 0.8])
    zenith = np.array([0.25, 0.45, 0.85])
    ground = np.array([0.3, 0.28, 0.25])
    sky = np.where(d[..., 1:2] >= 0, horizon + (zenith - horizon) * up[..., None] ** 0.5, ground)
    sd = np.asarray(sun_dir, np.float64)
    sd /= np.linalg.norm(sd)
    cosang = (d * sd).sum(-1)
    sun = cosang > math.cos(math.radians(sun_radius_deg))
    rgb = sky.copy()
    rgb[sun] = sun_scale * np.array([1.0, 0.95, 0.85])
    env = np.ones((height, width, 4), np.float32)
    env[..., :3] = rgb
    return env


def instanced_field(n_instances=64, freq=28, width=1920, height=1080, seed=1, env_size=(1024, 512)):
    """Config 2: n displaced geodesic icospheres (20*freq^2 tris each, one
    BLAS per instance), random TRS over a 20x20 ground quad, OpenPBR
    metal/glass/coated-diffuse mix, sky+sun environment."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = Scene()
    base_pts, idx = icosphere(freq)
    s.materials = [Material(base_color=(0.5, 0.5, 0.5))]  # ground
    for i in range(n_instances):
        kind = i % 3
        if kind == 0:
            m = Material(base_color=(0.9, 0.6, 0.3), base_metalness=1.0, specular_color=(0.95, 0.8, 0.6),
                         specular_roughness=float(rng.uniform(0.05, 0.5)))
        elif kind == 1:
            m = Material(base_color=(1.0, 1.0, 1.0), transmission_weight=1.0, specular_ior=1.5,
                         specular_roughness=float(rng.uniform(0.05, 0.5)))
        else:
            m = Material(base_color=tuple(rng.uniform(0.2, 0.8, 3)), coat_weight=0.3, coat_roughness=0.1,
                         specular_roughness=0.5)
        s.materials.append(m)
    ground = np.array([[-10, 0, 10], [10, 0, 10], [10, 0, -10], [-10, 0, -10]], np.float32)
    gp, gi, gn = quad(*ground)
    s.add_instance(s.add_mesh(gp, gi, gn), np.eye(4).reshape(16), 0)
    for i in range(n_instances):
        pos, ind, nrm = fbm_displaced_sphere(base_pts, idx, seed=i)
        mesh = s.add_mesh(pos, ind, nrm)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        scale = float(rng.lognormal(0.0, 0.3))
        t = np.array([rng.uniform(-9, 9), scale * 1.05, rng.uniform(-9, 9)])
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        M = np.eye(4)
        M[:3, :3] = R * scale
        M[:3, 3] = t
        s.add_instance(mesh, M.T.reshape(16), 1 + i)
    s.env = sky_env(*env_size)
    s.camera = Camera(eye=(0.0, 7.0, 17.0), target=(0.0, 0.8, 0.0), fov=40.0)
    return s
