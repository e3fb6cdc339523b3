"""Oracle analytic lights (sampleLight.wgsl.ts:10-253, intersectLight in
intersect.wgsl.ts:976-1122, power estimates in lightSampling.ts:72-107),
pinned by radiometry rather than by the oracle's own formulas.

A white-ish Lambertian plane (albedo a) is lit by one light; the camera looks
at the plane point under the light through a tiny field of view, with one
bounce, so the rendered radiance is a / pi times the irradiance there. Every
light type is compared with its closed-form (or numerically integrated)
irradiance:

- sphere of radius R, radiance L, centre at height h:  E = pi L (R / h)^2
- point light of intensity I at height h:              E = I / h^2
- spot light pointing down (point inside its cone):    E = I / h^2; outside: 0
- directional light straight down:                     E = I
- disk of radius R facing down at height h:            E = pi L R^2 / (h^2 + R^2)
- cylinder: the view-factor integral of its visible side, evaluated on a grid.
"""
import math

import numpy as np
import pytest

import oracle
from mutracer_amd import scene as S

A = 0.6


def _plane_scene(light, cam_eye=(0.0, 0.6, 3.0), target=(0.0, 0.0, 0.0)):
    s = S.Scene()
    s.materials = [S.Material(base_color=(A, A, A), specular_weight=0.0)]
    p, i, n = S.quad((-50, 0, 50), (50, 0, 50), (50, 0, -50), (-50, 0, -50))
    s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), 0)
    s.lights = [light]
    s.camera = S.Camera(eye=cam_eye, target=target, fov=0.05)
    return s


def _radiance(orc, s, spp=4096, candidates=1):
    out = orc.OracleRenderer(s, 4, 4, bounces=1, candidates=candidates, max_samples=spp).render(0, spp)
    return out[..., 0].sum() / out[..., 11].sum()


@pytest.mark.parametrize("radius,h", [(0.3, 1.0), (0.8, 2.0)])
def test_sphere_light_irradiance(orc, radius, h):
    le = 4.0
    got = _radiance(orc, _plane_scene(S.SphereLight(position=(0.0, h, 0.0), radius=radius, intensity=le)))
    want = A / math.pi * math.pi * le * (radius / h) ** 2
    assert abs(got / want - 1) < 0.01


def test_point_light_irradiance(orc):
    i, h = 7.0, 1.5
    got = _radiance(orc, _plane_scene(S.SphereLight(position=(0.0, h, 0.0), radius=0.0, intensity=i)), spp=64)
    want = A / math.pi * i / (h * h)
    assert abs(got / want - 1) < 1e-4  # a delta light: only the pixel footprint (~1e-5) varies


def test_spot_light_inside_and_outside_cone(orc):
    i, h = 5.0, 1.2
    down = S.SpotLight(position=(0.0, h, 0.0), direction=(0.0, -1.0, 0.0), inner_angle=20, outer_angle=30,
                       intensity=i)
    got = _radiance(orc, _plane_scene(down), spp=64)
    assert abs(got / (A / math.pi * i / (h * h)) - 1) < 1e-4
    # Tilted 60 degrees away: the point under the light is outside the outer cone.
    t = math.radians(60)
    away = S.SpotLight(position=(0.0, h, 0.0), direction=(math.sin(t), -math.cos(t), 0.0), inner_angle=20,
                       outer_angle=30, intensity=i)
    assert _radiance(orc, _plane_scene(away), spp=64) == 0.0


def test_spot_light_smooth_edge(orc):
    """Between the cone angles the emission is WGSL smoothstep(cos outer,
    cos inner, cos theta): tilt the spot so the point sits at 25 degrees."""
    i, h = 5.0, 1.2
    t = math.radians(25)
    spot = S.SpotLight(position=(0.0, h, 0.0), direction=(math.sin(t), -math.cos(t), 0.0), inner_angle=20,
                       outer_angle=30, intensity=i)
    lo, hi = math.cos(math.radians(30)), math.cos(math.radians(20))
    x = min(max((math.cos(t) - lo) / (hi - lo), 0.0), 1.0)
    want = A / math.pi * i / (h * h) * x * x * (3 - 2 * x)
    got = _radiance(orc, _plane_scene(spot), spp=64)
    assert abs(got / want - 1) < 1e-3  # footprint spread on the cone edge; linear instead of smoothstep is 4% off


def test_directional_light_irradiance(orc):
    i = 3.0
    got = _radiance(orc, _plane_scene(S.DirectionalLight(direction=(0.0, -1.0, 0.0), intensity=i)), spp=64)
    assert abs(got / (A / math.pi * i) - 1) < 1e-5


@pytest.mark.parametrize("radius,h", [(0.4, 1.0), (1.5, 0.8)])
def test_disk_light_irradiance(orc, radius, h):
    le = 3.0
    disk = S.DiskLight(position=(0.0, h, 0.0), direction=(0.0, -1.0, 0.0), radius=radius, intensity=le)
    got = _radiance(orc, _plane_scene(disk))
    want = A / math.pi * math.pi * le * radius ** 2 / (h * h + radius * radius)
    assert abs(got / want - 1) < 0.01


def test_disk_light_is_one_sided(orc):
    disk = S.DiskLight(position=(0.0, 1.0, 0.0), direction=(0.0, 1.0, 0.0), radius=0.5, intensity=3.0)
    assert _radiance(orc, _plane_scene(disk), spp=256) == 0.0


def _cylinder_irradiance(le, radius, length, h, n=1200):
    """Irradiance at the origin (normal +y) from the side of a Lambertian
    cylinder with axis along x at height h: midpoint rule over the axial
    coordinate and the angle around the axis, visible side only."""
    x = (np.arange(n) + 0.5) / n * length - length / 2
    phi = (np.arange(n) + 0.5) / n * 2 * math.pi
    X, P = np.meshgrid(x, phi, indexing="ij")
    ny, nz = np.cos(P), np.sin(P)
    px, py, pz = X, h + radius * ny, radius * nz
    d2 = px * px + py * py + pz * pz
    d = np.sqrt(d2)
    cos_recv = py / d                     # receiver normal (0, 1, 0)
    cos_emit = -(py * ny + pz * nz) / d   # surface normal (0, cos phi, sin phi)
    w = np.where((cos_recv > 0) & (cos_emit > 0), le * cos_recv * cos_emit / d2, 0.0)
    return w.sum() * (length / n) * (2 * math.pi * radius / n)


def test_cylinder_light_irradiance(orc):
    le, radius, length, h = 2.0, 0.25, 1.5, 1.0
    cyl = S.CylinderLight(position=(0.0, h, 0.0), axis=(1.0, 0.0, 0.0), radial=(0.0, 1.0, 0.0), radius=radius,
                          length=length, intensity=le)
    got = _radiance(orc, _plane_scene(cyl))
    want = A / math.pi * _cylinder_irradiance(le, radius, length, h)
    assert abs(got / want - 1) < 0.01


def test_bsdf_hits_of_area_lights_agree_with_nee(orc):
    """With four NEE candidates the estimator mixes light samples and BSDF
    hits of the sphere and disk lights (MIS); the result is the same
    irradiance."""
    for light, e in (
            (S.SphereLight(position=(0.0, 1.0, 0.0), radius=0.5, intensity=2.0), math.pi * 2.0 * 0.25),
            (S.DiskLight(position=(0.0, 1.0, 0.0), direction=(0.0, -1.0, 0.0), radius=0.7, intensity=2.0),
             math.pi * 2.0 * 0.49 / 1.49)):
        got = _radiance(orc, _plane_scene(light), candidates=4)
        assert abs(got / (A / math.pi * e) - 1) < 0.01


def test_analytic_light_powers():
    """estimateAnalyticLightPower for each type (unit-luminance colour)."""
    f = 0.5 * (1 - 0.5 * (math.cos(math.radians(20)) + math.cos(math.radians(35))))
    cases = [
        (S.SphereLight(position=(0, 0, 0), radius=0.0, intensity=2.0), 2.0 * 4 * math.pi),
        (S.SphereLight(position=(0, 0, 0), radius=0.5, intensity=3.0), 3.0 * math.pi * 4 * math.pi * 0.25),
        (S.SpotLight(position=(0, 0, 0), direction=(0, -1, 0), inner_angle=20, outer_angle=35, intensity=2.0),
         2.0 * 4 * math.pi * f),
        (S.SpotLight(position=(0, 0, 0), direction=(0, -1, 0), radius=0.2, inner_angle=20, outer_angle=35,
                     intensity=2.0), 2.0 / (math.pi * 0.04) * math.pi * 4 * math.pi * 0.04 * f),
        (S.DirectionalLight(direction=(0, -1, 0), intensity=1.5), 1.5 * 4 * math.pi),
        (S.DiskLight(position=(0, 0, 0), direction=(0, -1, 0), radius=1.0, intensity=2.0), 2.0 * math.pi * math.pi),
        (S.DiskLight(position=(0, 0, 0), direction=(0, -1, 0), radius=1.0, intensity=2.0, normalize=True),
         2.0 * math.pi),
        (S.CylinderLight(position=(0, 0, 0), axis=(1, 0, 0), radial=(0, 1, 0), radius=0.5, length=2.0,
                         intensity=1.0), math.pi * 2 * math.pi * 0.5 * 2.0),
    ]
    for light, want in cases:
        t = oracle.light_tables(light.pack()[None])
        assert abs(t["total"] / want - 1) < 1e-6, (light, t["total"], want)


def test_mixed_light_table():
    lights = np.stack([S.RectLight(position=(0, 0, 0), u=(1, 0, 0), v=(0, 0, 1), intensity=1.0).pack(),
                       S.SphereLight(position=(0, 0, 0), radius=0.0, intensity=1.0 / 4).pack(),
                       S.DirectionalLight(direction=(0, -1, 0), intensity=0.5).pack()])
    t = oracle.light_tables(lights)
    np.testing.assert_allclose(t["light_pmf"], np.array([1.0, 1.0, 2.0]) / 4.0, rtol=1e-6)


# ---------------------------------------------------------------- emissive triangles
def _emissive_quad_scene(le, half, h, face_down=True, double_sided=False, candidates_cam=(0.0, 0.6, 3.0)):
    """Diffuse plane under an emissive square (two triangles) of radiance le at
    height h; the square faces down (its geometric normal -y) unless
    face_down is False."""
    s = S.Scene()
    s.materials = [S.Material(base_color=(A, A, A), specular_weight=0.0),
                   S.Material(base_color=(0.0, 0.0, 0.0), specular_weight=0.0, emission=(le, le, le),
                              double_sided=double_sided)]
    p, i, n = S.quad((-50, 0, 50), (50, 0, 50), (50, 0, -50), (-50, 0, -50))
    s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), 0)
    pts = np.array([[-half, h, -half], [half, h, -half], [half, h, half], [-half, h, half]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32) if face_down else np.array([0, 2, 1, 0, 3, 2], np.uint32)
    nrm = np.tile(np.array([[0.0, -1.0 if face_down else 1.0, 0.0]], np.float32), (4, 1))
    s.add_instance(s.add_mesh(pts, idx, nrm), np.eye(4).reshape(16), 1)
    s.camera = S.Camera(eye=candidates_cam, target=(0.0, 0.0, 0.0), fov=0.05)
    return s


def test_emissive_triangle_list_and_weights():
    import oracle as O
    s = _emissive_quad_scene(2.0, 0.5, 1.0)
    tris, w = O.emissive_triangles(s)
    assert tris.shape == (2, 12)
    np.testing.assert_allclose(tris[:, 9], 0.5, rtol=1e-6)        # two halves of a unit square
    np.testing.assert_allclose(w, 0.5 * 2.0, rtol=1e-6)           # area * luminance
    assert np.all(tris[:, 10] == 1)
    n = np.cross(tris[0, 3:6] - tris[0, 0:3], tris[0, 6:9] - tris[0, 0:3])
    assert n[1] < 0  # faces the plane


@pytest.mark.parametrize("candidates", [1, 4])
def test_emissive_quad_matches_rect_light(orc, candidates):
    """An emissive square of radiance L lights the plane like a rect light of
    the same radiance: light samples of its triangles plus MIS-weighted BSDF
    hits give the view-factor irradiance."""
    import test_oracle_render as R
    le, half, h = 3.0, 0.4, 1.0
    got = _radiance(orc, _emissive_quad_scene(le, half, h), candidates=candidates)
    want = A / math.pi * R._parallel_rect_irradiance(le, half, half, h)
    assert abs(got / want - 1) < 0.01


def test_emissive_seen_directly(orc):
    """Camera rays take the emission with weight 1."""
    le = 2.5
    s = _emissive_quad_scene(le, 0.5, 1.0, candidates_cam=(0.0, 0.2, 0.0))
    s.camera = S.Camera(eye=(0.0, 0.2, 0.0), target=(0.0, 1.0, 0.0), up=(0.0, 0.0, 1.0), fov=0.05)
    out = orc.OracleRenderer(s, 4, 4, bounces=1, candidates=1, max_samples=4).render(0, 4)
    np.testing.assert_allclose(out[..., 0:3].sum(axis=(0, 1)) / out[..., 11].sum(), le, rtol=1e-6)


def test_emissive_back_face_only_when_double_sided(orc):
    """A square facing away from the plane lights it only when double-sided;
    then only BSDF hits reach its back (weight 1: no light sample covers it)."""
    le, half, h = 3.0, 0.4, 1.0
    assert _radiance(orc, _emissive_quad_scene(le, half, h, face_down=False), spp=512) == 0.0
    import test_oracle_render as R
    got = _radiance(orc, _emissive_quad_scene(le, half, h, face_down=False, double_sided=True), spp=16384)
    want = A / math.pi * R._parallel_rect_irradiance(le, half, half, h)
    assert abs(got / want - 1) < 0.02
