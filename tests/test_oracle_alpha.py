"""Oracle pins for mask alpha coverage (oracle/orc_render.cpp get_material's
alpha resolution, alpha_covers, and the integrator's pass-through vertex),
against consequences of the reference's rules that do not reuse its code:

  - a caster whose coverage stays below the cutoff, with nothing behind it
    but the light and the sky, renders bit for bit like the scene without
    it: shadow rays pass it (intersectionAlpha.wgsl.ts:80-87) and a ray that
    hits it is not alpha-occluded, so the light and the environment are
    reached exactly as by a ray that misses (rayTrace.compute.ts:533-552);
  - a caster whose coverage reaches the cutoff (opacity equal to it, or an
    alpha map / base-map alpha above it) renders bit for bit like the same
    caster with an opaque material; an alpha map reads its g channel and the
    base map its a channel (constant textures whose other channels fall on
    the other side of the cutoff), and the caster does shadow the floor;
  - panes of a fully masked-out material in front of the camera are passed
    through without consuming a depth or a coverage draw (:440-458, :527):
    the floor behind any number of panes is shaded with the same sample
    dimensions and random stream as without panes, so images agree to the
    rounding of the moved ray origin on every sampling pattern, while the
    albedo guide is still the floor's (the pane's where the sky is seen);
  - over an empty sky the same panes leave the color planes bit for bit
    unchanged and set the albedo guide to the pane's color (the primary
    hit's guides, :750-756, are not overwritten by a miss).
Planted in a copy of the oracle, a `>` for `>=` at the cutoff, the alpha
map read from its a channel, a depth increment at a passed-through vertex,
and a coverage draw at the completeVertex that follows it each fail one of
these.
"""
import numpy as np
import pytest

import oracle
from mutracer_amd import scene as S

KW = dict(bounces=3, candidates=2, max_samples=4)


def _render(sc, w=24, h=20, spp=4, **kw):
    args = dict(KW)
    args.update(kw)
    return oracle.OracleRenderer(sc, w, h, **args).render(0, spp)


def test_open_caster_equals_absent_and_covered_equals_opaque():
    absent = _render(S.alpha_caster("absent"))
    opaque = _render(S.alpha_caster("opaque"))
    lit = absent[..., 0:3].mean()
    assert lit > 0.05
    # the opaque caster shadows part of the floor
    assert np.abs(absent[..., 0:3] - opaque[..., 0:3]).max() > 0.05 * lit
    for kind in ("mask_open", "alpha_map_open", "base_alpha_open"):
        assert np.array_equal(_render(S.alpha_caster(kind)), absent), kind
    for kind in ("mask_covered", "alpha_map_covered", "base_alpha_covered"):
        assert np.array_equal(_render(S.alpha_caster(kind)), opaque), kind


def test_checker_caster_lies_between():
    absent = _render(S.alpha_caster("absent"), spp=16)[..., 0:3].sum()
    opaque = _render(S.alpha_caster("opaque"), spp=16)[..., 0:3].sum()
    checker = _render(S.alpha_caster("checker"), spp=16)[..., 0:3].sum()
    assert opaque < checker < absent


@pytest.mark.parametrize("pattern", ["sobol", "zsobol", "random"])
def test_panes_are_passed_through_without_a_depth(pattern):
    plain = _render(S.masked_panes(0), w=20, h=16, spp=4, pattern=pattern, bounces=4)
    assert plain[..., 0:3].mean() > 0.05
    for panes in (1, 3):
        got = _render(S.masked_panes(panes), w=20, h=16, spp=4, pattern=pattern, bounces=4)
        rel = np.abs(got[..., 0:3] - plain[..., 0:3]) / np.maximum(np.abs(plain[..., 0:3]), 1e-3)
        assert rel.max() < 1e-3, (panes, rel.max())
        floor = plain[..., 4] == plain[..., 4].max()  # every sample on the floor
        sky = plain[..., 4] == 0                       # every sample on the sky: the pane's guides
        assert floor.mean() > 0.3 and sky.mean() > 0.05
        assert np.array_equal(got[floor][:, 4:8], plain[floor][:, 4:8])
        pane = np.float32(S.masked_panes(1).materials[1].base_color)
        assert np.allclose(got[sky][:, 4:7] / 4, pane, atol=0)


def test_panes_over_the_sky_keep_the_color_and_take_the_guides():
    sc = S.masked_panes(2)
    sky = S.masked_panes(0)
    for s in (sc, sky):  # drop the floor: the panes are all the camera sees
        s.mesh_index, s.transforms, s.instance_material = s.mesh_index[1:], s.transforms[1:], s.instance_material[1:]
    sky.mesh_index, sky.transforms, sky.instance_material = [], [], []
    got = _render(sc, w=16, h=12, spp=2)
    ref = _render(sky, w=16, h=12, spp=2)
    assert np.array_equal(got[..., 0:4], ref[..., 0:4])
    pane = np.float32(S.masked_panes(1).materials[1].base_color)
    assert np.allclose(got[..., 4:7] / 2, pane, atol=0)
    assert np.all(ref[..., 4:7] == 0)
