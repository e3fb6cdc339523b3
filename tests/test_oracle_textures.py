"""Oracle pins for material textures (oracle/orc_render.cpp sample_map,
get_material's texture samples, mat_normal), against mathematics that does
not reuse its code:

  - bilinear filtering between texel centers reproduces an affine image
    exactly (to f32 rounding) inside the image, and clamps at the edges;
  - repeat wraps the footprint across the edge (u = 0 averages the first and
    last texel) and is periodic; mirrored repeat is symmetric about 0 and 1;
  - RGBA8 texels read as byte / 255 at texel centers, and color slots decode
    them with IEC 61966-2-1 sRGB; flip_y reverses the rows; the texture's UV
    transform is uv * scale + offset;
  - end to end through the integrator: camera rays hitting an emissive quad
    whose emission is textured see, pixel by pixel, a numpy restatement of
    the image at the analytically intersected UV (plane hit, bilinear,
    sRGB), which pins the UV interpolation, the material UV scale and the
    emissive map;
  - a normal map in a stretched UV frame gives the cotangent-frame normal
    (the UV gradients of the quad; thetenthplanet.de/archives/1180), seen in
    the normal AOV;
  - a constant map is the same as multiplying the packed parameter: the
    render with constant base / metallic-roughness maps equals, bit for bit,
    the render of the folded material; an sRGB base map multiplies by the
    IEC-decoded byte; an emissive-triangle light with a constant emissive
    map converges to the folded emission (its NEE samples read the map).
Planted in a copy of the oracle, a missing half-texel shift, a flipped sign
in the cotangent frame, swapped metallic / roughness channels and a dropped
sRGB decode on the base map each fail at least one of these.
"""
import numpy as np
import pytest

import oracle
from mutracer_amd import scene as S


def _texset(textures):
    return oracle.TextureSet(textures)


def _affine(w, h):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    px = np.zeros((h, w, 4), np.float32)
    px[..., 0] = 0.3 + 0.05 * x + 0.11 * y
    px[..., 1] = 1.0 - 0.02 * x + 0.07 * y
    px[..., 2] = 0.5 * x - 0.25 * y
    px[..., 3] = 1.0
    return px


def _f(px, tx, ty):
    return 0.3 + 0.05 * tx + 0.11 * ty, 1.0 - 0.02 * tx + 0.07 * ty, 0.5 * tx - 0.25 * ty


def test_bilinear_reproduces_affine_image_and_clamps():
    w, h = 8, 5
    T = _texset([S.Texture(_affine(w, h), wrap_s="clamp-to-edge", wrap_t="clamp-to-edge")])
    rng = np.random.default_rng(1)
    uv = rng.uniform(0, 1, (400, 2)).astype(np.float32)
    got, _ = T.sample(0, uv)
    tx = np.clip(uv[:, 0].astype(np.float64) * w - 0.5, 0, w - 1)
    ty = np.clip(uv[:, 1].astype(np.float64) * h - 0.5, 0, h - 1)
    ref = np.stack(_f(None, tx, ty), 1)
    assert np.abs(got[:, :3] - ref).max() < 2e-6
    assert np.all(got[:, 3] == 1.0)


def test_repeat_wraps_across_the_edge_and_is_periodic():
    w, h = 8, 5
    px = _affine(w, h)
    T = _texset([S.Texture(px)])
    v = (2 + 0.5) / h  # row 2's center
    got, _ = T.sample(0, np.array([[0.0, v]], np.float32))
    assert np.allclose(got[0, :3], 0.5 * (px[2, w - 1, :3] + px[2, 0, :3]), atol=1e-6)
    k = np.arange(0, 64)
    uv = np.stack([k / 64.0, (63 - k) / 64.0], 1).astype(np.float32)
    a, _ = T.sample(0, uv)
    b, _ = T.sample(0, uv + np.float32(1.0))
    c, _ = T.sample(0, uv - np.float32(3.0))
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_mirrored_repeat_is_symmetric():
    T = _texset([S.Texture(_affine(8, 5), wrap_s="mirrored-repeat", wrap_t="mirrored-repeat")])
    k = np.arange(1, 64)
    u = (k / 64.0).astype(np.float32)
    base = np.stack([u, u[::-1]], 1)
    a, _ = T.sample(0, base)
    b, _ = T.sample(0, np.float32(2.0) - base)
    c, _ = T.sample(0, -base)
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_rgba8_texel_centers_and_srgb_decode():
    rng = np.random.default_rng(2)
    w, h = 16, 8
    px = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    T = _texset([S.Texture(px), S.Texture((px / 255.0).astype(np.float32))])
    y, x = np.mgrid[0:h, 0:w]
    uv = np.stack([(x.ravel() + 0.5) / w, (y.ravel() + 0.5) / h], 1).astype(np.float32)
    rgba, dec = T.sample(0, uv)
    assert np.array_equal(rgba, (px.reshape(-1, 4).astype(np.float32) / np.float32(255)))
    c = rgba[:, :3].astype(np.float64)
    iec = np.where(c < 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    assert np.abs(dec - iec).max() < 2e-6 * max(1.0, iec.max())
    _, dec32 = T.sample(1, uv)  # RGBA32F is linear: no decode
    assert np.array_equal(dec32, T.sample(1, uv)[0][:, :3])


def test_flip_y_and_uv_transform():
    rng = np.random.default_rng(3)
    px = rng.uniform(0, 1, (6, 9, 4)).astype(np.float32)
    T = _texset([S.Texture(px, flip_y=True), S.Texture(px[::-1].copy()),
                 S.Texture(px, uv_scale=(1.7, -0.6), uv_offset=(0.3, 2.25), wrap_t="mirrored-repeat"),
                 S.Texture(px, wrap_t="mirrored-repeat")])
    uv = rng.uniform(-2, 2, (300, 2)).astype(np.float32)
    assert np.array_equal(T.sample(0, uv)[0], T.sample(1, uv)[0])
    moved = np.stack([uv[:, 0] * np.float32(1.7) + np.float32(0.3), uv[:, 1] * np.float32(-0.6) + np.float32(2.25)], 1)
    assert np.array_equal(T.sample(2, uv)[0], T.sample(3, moved)[0])


def _np_sample(px, u, v, wrap=("repeat", "repeat")):
    """Independent f64 restatement: wrap, texel = wrapped * size - 0.5,
    bilinear with wrapped / clamped taps."""
    h, w = px.shape[:2]

    def wrapc(x, mode):
        if mode == "repeat":
            return x - np.floor(x)
        if mode == "clamp-to-edge":
            return np.clip(x, 0, 1)
        return 1 - np.abs(1 - (x - 2 * np.floor(x / 2)))

    def tap(c, n, mode):
        return np.mod(c, n) if mode == "repeat" else np.clip(c, 0, n - 1)

    tx, ty = wrapc(u, wrap[0]) * w - 0.5, wrapc(v, wrap[1]) * h - 0.5
    x0, y0 = np.floor(tx).astype(int), np.floor(ty).astype(int)
    fx, fy = (tx - x0)[:, None], (ty - y0)[:, None]
    a, b = tap(x0, w, wrap[0]), tap(x0 + 1, w, wrap[0])
    c, d = tap(y0, h, wrap[1]), tap(y0 + 1, h, wrap[1])
    return (px[c, a] * (1 - fx) + px[c, b] * fx) * (1 - fy) + (px[d, a] * (1 - fx) + px[d, b] * fx) * fy


def test_camera_rays_see_the_emissive_texture():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (12, 10, 4)).astype(np.uint8)
    sc = S.Scene()
    sc.textures = [S.Texture(img, uv_scale=(1.5, 0.75), uv_offset=(0.2, -0.1), wrap_t="mirrored-repeat")]
    sc.materials = [S.Material(base_color=(0.0, 0.0, 0.0), specular_weight=0.0, emission=(2.0, 1.5, 1.0),
                               emissive_map=0, uv_scale=(1.25, 2.0))]
    p, i, n = S.quad((-1, 0, 0), (1, 0, 0), (1, 2, 0), (-1, 2, 0))
    sc.add_instance(sc.add_mesh(p, i, n, uvs=S.QUAD_UVS), np.eye(4).reshape(16), 0)
    sc.camera = S.Camera(eye=(0.0, 1.0, 3.0), target=(0.0, 1.0, 0.0), fov=30.0)
    W = H = 24
    out = oracle.OracleRenderer(sc, W, H, bounces=2, max_samples=1).render(0, 1)
    y, x = np.mgrid[0:H, 0:W]
    rays = oracle.camera_rays(sc.camera.pack(W, H), W, H, np.stack([x.ravel(), y.ravel(), 0 * x.ravel()], 1))
    o, d = rays[:, :3].astype(np.float64), rays[:, 3:].astype(np.float64)
    t = -o[:, 2] / d[:, 2]
    hit = o + t[:, None] * d
    inside = (np.abs(hit[:, 0]) < 1) & (hit[:, 1] > 0) & (hit[:, 1] < 2)
    assert inside.mean() > 0.5
    u = (hit[:, 0] + 1) / 2 * 1.25
    v = hit[:, 1] / 2 * 2.0
    texel = _np_sample(img.astype(np.float64) / 255.0, u * 1.5 + 0.2, v * 0.75 - 0.1, ("repeat", "mirrored-repeat"))
    srgb = texel[:, :3]
    lin = np.where(srgb < 0.04045, srgb / 12.92, ((srgb + 0.055) / 1.055) ** 2.4)
    ref = lin * np.array([2.0, 1.5, 1.0])
    got = out[..., 0:3].reshape(-1, 3)
    # away from the quad's border (where the jittered ray may miss)
    edge = (np.abs(hit[:, 0]) > 0.97) | (hit[:, 1] < 0.03) | (hit[:, 1] > 1.97)
    sel = inside & ~edge
    assert np.abs(got[sel] - ref[sel]).max() < 2e-4
    assert np.all(got[~inside & ~edge] == 0)


@pytest.mark.parametrize("stretch", [1.0, 2.0])
def test_normal_map_cotangent_frame(stretch):
    """Quad in the xy plane with UV = (stretch * u0, v0): the cotangent frame
    is (du/dp, dv/dp) scaled by the larger length, i.e. x and y / stretch
    after normalization; a constant tangent-space normal (a, b, 1) with
    normal scale (sx, sy) maps to normalize((a*sx, b*sy/stretch... )) as
    derived below, seen in the primary normal AOV."""
    a, b, sx, sy = 0.4, -0.3, 0.8, 1.3
    tex = np.zeros((4, 4, 4), np.float32)
    tex[..., 0] = 0.5 + 0.5 * a
    tex[..., 1] = 0.5 + 0.5 * b
    tex[..., 2] = 1.0
    tex[..., 3] = 1.0
    sc = S.Scene()
    sc.textures = [S.Texture(tex)]
    sc.materials = [S.Material(normal_map=0, normal_scale=(sx, sy))]
    p, i, n = S.quad((-1, 0, 0), (1, 0, 0), (1, 2, 0), (-1, 2, 0))
    uvs = S.QUAD_UVS * np.array([stretch, 1.0], np.float32)
    sc.add_instance(sc.add_mesh(p, i, n, uvs=uvs), np.eye(4).reshape(16), 0)
    sc.lights = [S.RectLight(position=(0.0, 1.0, 2.0), u=(0.3, 0, 0), v=(0, -0.3, 0), intensity=5.0)]
    sc.camera = S.Camera(eye=(0.0, 1.0, 3.0), target=(0.0, 1.0, 0.0), fov=20.0)
    out = oracle.OracleRenderer(sc, 8, 8, bounces=1, max_samples=1).render(0, 1)
    nrm = out[2:6, 2:6, 8:11].reshape(-1, 3)
    # u = stretch * (x + 1) / 2, v = y / 2: grad u = (stretch / 2, 0), grad v = (0, 1/2);
    # the frame divides both by the larger length: dpdu = (1, 0, 0) * (stretch/2) / m,
    # dpdv = (0, 1, 0) * (1/2) / m with m = max(stretch, 1) / 2
    m = max(stretch, 1.0) / 2
    dpdu = np.array([stretch / 2 / m, 0, 0])
    dpdv = np.array([0, 0.5 / m, 0])
    t = np.array([a * sx, b * sy, 1.0])
    t /= np.linalg.norm(t)
    want = dpdu * t[0] + dpdv * t[1] + np.array([0, 0, 1.0]) * t[2]
    want /= np.linalg.norm(want)
    assert np.abs(nrm - want).max() < 2e-6


def _constant_map_scene(base_texture, mr_texture):
    sc = S.cornell_box()
    sc.textures = [base_texture, mr_texture]
    for k in range(len(sc.meshes)):
        nv = sc.meshes[k]["positions"].shape[0]
        sc.meshes[k]["uvs"] = np.random.default_rng(k).uniform(-1, 2, (nv, 2)).astype(np.float32)
    sc.materials[1].map = 0
    sc.materials[2].metallic_roughness_map = 1
    sc.materials[2].base_metalness = 0.9
    sc.materials[2].specular_roughness = 0.7
    return sc


def _folded(sc, base, mr):
    import copy
    folded = copy.deepcopy(sc)
    m = folded.materials
    m[1].map = None
    m[1].base_color = tuple(float(v) for v in np.float32(m[1].base_color) * np.float32(base))
    m[2].metallic_roughness_map = None
    m[2].base_metalness = float(np.float32(0.9) * np.float32(mr[2]))
    m[2].specular_roughness = float(np.float32(0.7) * np.float32(mr[1]))
    return folded


def test_constant_maps_equal_folded_parameters():
    """Constant base and metallic-roughness maps (RGBA32F powers of two, so
    that bilinear blends of equal texels are exact) on meshes with arbitrary
    UVs render bit for bit like the materials with the constants multiplied
    into their parameters: base color *= rgb, metalness *= b, roughness *= g."""
    base = np.array([0.5, 0.25, 1.0, 1.0], np.float32)
    mr = np.array([0.0, 0.5, 0.25, 1.0], np.float32)
    sc = _constant_map_scene(S.Texture(np.tile(base, (3, 5, 1))),
                             S.Texture(np.tile(mr, (4, 2, 1)), wrap_s="clamp-to-edge"))
    kw = dict(bounces=6, candidates=2, max_samples=4)
    a = oracle.OracleRenderer(sc, 24, 24, **kw).render(0, 4)
    b = oracle.OracleRenderer(_folded(sc, base[:3], mr), 24, 24, **kw).render(0, 4)
    plain = oracle.OracleRenderer(S.cornell_box(), 24, 24, **kw).render(0, 4)
    assert a[..., 0:3].mean() > 0.01
    assert np.abs(a[..., 0:3] - plain[..., 0:3]).max() > 0.01
    assert np.array_equal(a, b)


def test_constant_srgb_base_map_equals_decoded_color():
    """A constant RGBA8 base map multiplies the base color by the sRGB-decoded
    texel (IEC 61966-2-1): same image as the folded color up to the last-bit
    differences that bilinear blends of equal bytes leave."""
    rgba = np.array([200, 90, 30, 255], np.uint8)
    mr = np.array([0.0, 1.0, 1.0, 1.0], np.float32)
    sc = _constant_map_scene(S.Texture(np.tile(rgba, (3, 5, 1))), S.Texture(np.tile(mr, (1, 1, 1))))
    c = rgba[:3] / 255.0
    dec = np.where(c < 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    kw = dict(bounces=6, candidates=2, max_samples=16)
    a = oracle.OracleRenderer(sc, 24, 24, **kw).render(0, 16)[..., 0:3]
    b = oracle.OracleRenderer(_folded(sc, dec, mr), 24, 24, **kw).render(0, 16)[..., 0:3]
    assert np.abs(a - b).mean() < 1e-4 * np.abs(b).mean()


def test_emissive_map_on_a_triangle_light_converges_like_folded_emission():
    """An emissive-triangle light whose emission is a constant RGBA32F map c
    converges to the same image as the light with emission * c (the alias
    weights use the packed, untextured emission, so the two estimators differ
    sample by sample: the comparison is of converged means, NEE included)."""
    import copy
    c = np.array([0.5, 0.8, 0.3, 1.0], np.float32)
    sc = S.cornell_box()
    sc.lights = []
    sc.textures = [S.Texture(np.tile(c, (2, 2, 1)))]
    p, i, n = S.quad((-0.3, 1.9, 0.2), (0.3, 1.9, 0.2), (0.3, 1.9, -0.2), (-0.3, 1.9, -0.2))
    sc.materials.append(S.Material(base_color=(0.1, 0.1, 0.1), emission=(8.0, 6.0, 4.0), emissive_map=0))
    sc.add_instance(sc.add_mesh(p[::-1].copy(), i, -n, uvs=S.QUAD_UVS), np.eye(4).reshape(16), len(sc.materials) - 1)
    folded = copy.deepcopy(sc)
    folded.materials[-1].emissive_map = None
    folded.materials[-1].emission = tuple(float(v) for v in np.float32((8.0, 6.0, 4.0)) * c[:3])
    kw = dict(bounces=4, candidates=1, max_samples=256)
    a = oracle.OracleRenderer(sc, 16, 16, **kw).render(0, 256)[..., 0:3] / 256
    b = oracle.OracleRenderer(folded, 16, 16, **kw).render(0, 256)[..., 0:3] / 256
    assert a.mean() > 0.01
    assert abs(a.mean() / b.mean() - 1) < 0.03
    assert abs(a[..., 2].mean() / b[..., 2].mean() - 1) < 0.05
