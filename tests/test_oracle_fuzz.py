"""Oracle fuzz lobe (openPbrBsdf.wgsl.ts:595-625, layered at :1400-1418,
sampled with the diffuse cosine lobe at :1482-1497, probability at
:993-998), pinned by what its layering fixes rather than by its formula:

- a fuzz-only surface (no base, no specular) under a uniform white
  environment reflects exactly weight * color * A(view cosine), where A is
  the lobe's fitted directional albedo: BSDF samples, environment samples
  and their MIS weights must add up to it, with 1 and 4 NEE candidates;
- a fuzz layer over a base scales the base by (1 - weight * A) and adds its
  own weight * color * A: the rendered value is that affine combination of
  the base-only render;
- the fit lies in [0, 1], grows toward grazing views and with roughness at
  normal incidence (the shape of a sheen albedo).

The fit's constants themselves are data of the reference (no published
table to check them against here): parity of those is unpinned.
"""
import math

import numpy as np
import pytest

import oracle
from mutracer_amd import scene as S


def _plane(material, cos_view, env=1.0):
    s = S.Scene()
    s.materials = [material]
    p, i, n = S.quad((-500, 0, 500), (500, 0, 500), (500, 0, -500), (-500, 0, -500))
    s.add_instance(s.add_mesh(p, i, n), np.eye(4).reshape(16), 0)
    h = 3.0 * cos_view
    d = 3.0 * math.sqrt(1.0 - cos_view * cos_view)
    s.camera = S.Camera(eye=(0.0, h, d), target=(0.0, 0.0, 0.0), fov=0.05)
    e = np.full((32, 64, 4), env, np.float32)
    e[..., 3] = 1
    s.env = e
    return s


def _radiance(orc, sc, spp, candidates=1):
    out = orc.OracleRenderer(sc, 4, 4, bounces=1, candidates=candidates, max_samples=spp).render(0, spp)
    return out[..., 0:3].sum(axis=(0, 1)) / out[..., 11].sum()


def test_fuzz_albedo_shape():
    cs = np.linspace(0.0, 1.0, 41)
    for r in (0.3, 0.5, 0.8, 1.0):
        a = np.array([oracle.fuzz_albedo(float(c), r) for c in cs])
        assert a.min() >= 0.0 and a.max() <= 1.0
        assert np.all(np.diff(a) <= 1e-7), r  # more fuzz reflection toward grazing
    normal = [oracle.fuzz_albedo(1.0, r) for r in (0.2, 0.4, 0.6, 0.8, 1.0)]
    assert np.all(np.diff(normal) > 0)


@pytest.mark.parametrize("candidates", [1, 4])
@pytest.mark.parametrize("cos_view,rough", [(0.9, 0.8), (0.5, 0.5), (0.25, 1.0)])
def test_fuzz_only_reflects_its_albedo(orc, candidates, cos_view, rough):
    w, color = 0.9, (1.0, 0.6, 0.3)
    m = S.Material(base_weight=0.0, specular_weight=0.0, fuzz_weight=w, fuzz_color=color, fuzz_roughness=rough)
    got = _radiance(orc, _plane(m, cos_view), 2048, candidates)
    want = w * np.array(color) * oracle.fuzz_albedo(cos_view, rough)
    np.testing.assert_allclose(got, want, rtol=0.01)


@pytest.mark.parametrize("cos_view", [0.3, 0.8])
def test_fuzz_layer_over_base(orc, cos_view):
    w, rough, color = 0.7, 0.6, (0.5, 0.9, 0.7)
    base = dict(base_color=(0.6, 0.4, 0.2), specular_roughness=0.5)
    plain = _radiance(orc, _plane(S.Material(**base), cos_view), 4096)
    fuzzy = _radiance(orc, _plane(S.Material(fuzz_weight=w, fuzz_color=color, fuzz_roughness=rough, **base),
                                  cos_view), 4096)
    a = oracle.fuzz_albedo(cos_view, rough)
    want = w * a * np.array(color) + (1.0 - w * a) * plain
    np.testing.assert_allclose(fuzzy, want, rtol=0.015)
