"""Oracle thin-walled surfaces and the subsurface boundary lobe
(openPbrBsdf.wgsl.ts:697-750, eval :1137-1312, sampling :1599-1640;
sampleMaterial.wgsl.ts:328-345), pinned by slab optics and by where the
light goes, not by the oracle's formulas.

The scenes are an infinite horizontal plane (normal +y) lit by an
environment that is E_UP above the horizon and E_DOWN below it, seen from
above at a known angle through a tiny field of view, with two bounces, so
every path ends in the environment after at most one scattering:

- smooth thin-walled glass (a window pane): the ray continues undeviated
  below with the slab transmittance and reflects above with the slab
  reflectance R = 2F / (1 + F), F the unpolarized Fresnel reflectance of
  the interface (textbook Fresnel equations here): radiance = (1 - R) T E_DOWN
  + R E_UP, T = 1 for a white tint and tint^(1 / cos theta_t) otherwise;
- the subsurface boundary of a solid (specular weight 0, so the interface is
  index-matched): a diffuse transmission into the object, radiance =
  E_DOWN;
- a thin-walled subsurface sheet: half diffuse reflection, half diffuse
  transmission, each tinted by the subsurface color c:
  radiance = c (E_UP + E_DOWN) / 2;
- a glass sphere made thin-walled conserves energy in a white furnace.
"""
import math

import numpy as np
import pytest

import oracle
from mutracer_amd import scene as S

E_UP, E_DOWN = 0.3, 1.0


def _env(steep=None):
    """E_UP above the horizon, E_DOWN below; `steep` replaces E_DOWN in the
    rows more than 45 degrees below the horizon (v > 3/4)."""
    e = np.zeros((32, 64, 4), np.float32)
    e[:16, :, 0:3] = E_UP      # rows of the upper hemisphere (v = acos(y) / pi < 1/2)
    e[16:, :, 0:3] = E_DOWN
    if steep is not None:
        e[24:, :, 0:3] = steep
    e[..., 3] = 1
    return e


def _plane_scene(material, cos_view, steep=None):
    s = S.Scene()
    s.materials = [material]
    p, i, n = S.quad((-500, 0, 500), (500, 0, 500), (500, 0, -500), (-500, 0, -500))
    s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), 0)
    h = 3.0 * cos_view
    d = 3.0 * math.sqrt(1.0 - cos_view * cos_view)
    s.camera = S.Camera(eye=(0.0, h, d), target=(0.0, 0.0, 0.0), fov=0.05)
    s.env = _env(steep)
    return s


def _radiance(orc, sc, spp=1024, candidates=1, bounces=2):
    out = orc.OracleRenderer(sc, 4, 4, bounces=bounces, candidates=candidates, max_samples=spp).render(0, spp)
    return out[..., 0:3].sum(axis=(0, 1)) / out[..., 11].sum()


def _fresnel(cos_i, n):
    """Unpolarized Fresnel reflectance of an interface of relative index n."""
    sin_t = math.sqrt(max(0.0, 1.0 - cos_i * cos_i)) / n
    cos_t = math.sqrt(1.0 - sin_t * sin_t)
    rs = (cos_i - n * cos_t) / (cos_i + n * cos_t)
    rp = (n * cos_i - cos_t) / (n * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp), cos_t


def test_environment_split():
    s = S.Scene()
    s.materials = [S.Material()]
    p, i, n = S.quad((1000, 1000, 1000), (1001, 1000, 1000), (1001, 1000, 1001), (1000, 1000, 1001))
    s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), 0)
    s.env = _env()
    for target, want in (((0.0, 5.0, 0.0), E_UP), ((0.0, -5.0, 0.0), E_DOWN)):
        s.camera = S.Camera(eye=(0.0, 0.0, 0.0), target=target, up=(0.0, 0.0, 1.0), fov=0.05)
        out = oracle.OracleRenderer(s, 2, 2, bounces=1, max_samples=4).render(0, 4)
        np.testing.assert_allclose(out[..., 0] / out[..., 11], want, rtol=1e-6)


@pytest.mark.parametrize("candidates", [1, 3])
@pytest.mark.parametrize("cos_view", [0.98, 0.7, 0.35])
def test_thin_window_pane(orc, candidates, cos_view):
    m = S.Material(transmission_weight=1.0, specular_roughness=0.0, specular_ior=1.5, thin_walled=True)
    got = _radiance(orc, _plane_scene(m, cos_view), spp=256, candidates=candidates)
    f, _ = _fresnel(cos_view, 1.5)
    r = 2 * f / (1 + f)
    want = (1 - r) * E_DOWN + r * E_UP
    np.testing.assert_allclose(got, want, rtol=2e-3)


def test_thin_window_tint_path_length(orc):
    """The ray leaves the pane undeviated, 30 degrees below the horizon
    (E_DOWN); refracted into glass it would run 55 degrees below, where the
    environment is darker."""
    cos_view, tint = 0.5, np.array([0.9, 0.6, 0.3])
    m = S.Material(transmission_weight=1.0, transmission_color=tuple(tint), specular_roughness=0.0,
                   specular_ior=1.5, thin_walled=True)
    got = _radiance(orc, _plane_scene(m, cos_view, steep=0.1), spp=256)
    f, cos_t = _fresnel(cos_view, 1.5)
    r = 2 * f / (1 + f)
    want = (1 - r) * tint ** (1 / cos_t) * E_DOWN + r * E_UP
    np.testing.assert_allclose(got, want, rtol=2e-3)


@pytest.mark.parametrize("cos_view", [0.9, 0.4])
def test_subsurface_boundary_transmits_diffusely(orc, cos_view):
    m = S.Material(subsurface_weight=1.0, specular_weight=0.0)
    got = _radiance(orc, _plane_scene(m, cos_view), spp=2048)
    np.testing.assert_allclose(got, E_DOWN, rtol=5e-3)


@pytest.mark.parametrize("candidates", [1, 2])
def test_thin_subsurface_sheet_splits_light(orc, candidates):
    c = np.array([0.9, 0.5, 0.2])
    m = S.Material(subsurface_weight=1.0, subsurface_color=tuple(c), specular_weight=0.0, thin_walled=True)
    got = _radiance(orc, _plane_scene(m, 0.6), spp=4096, candidates=candidates)
    np.testing.assert_allclose(got, c * (E_UP + E_DOWN) / 2, rtol=0.01)


def test_thin_glass_sphere_furnace(orc):
    import test_oracle_render as R
    m = S.Material(base_color=(1.0, 1.0, 1.0), transmission_weight=1.0, specular_ior=1.5, specular_roughness=0.0,
                   thin_walled=True)
    sc = R._furnace_scene(m, geometry="sphere")
    out = orc.OracleRenderer(sc, 16, 16, bounces=20, max_samples=32).render(0, 32)
    rgb = out[..., 0:3].sum(axis=(0, 1)) / out[..., 11].sum()
    np.testing.assert_allclose(rgb, 1.0, rtol=0.01)
