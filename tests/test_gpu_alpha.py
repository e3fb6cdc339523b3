"""Mask alpha coverage on the GPU vs the oracle (row f4: alpha maps and
coverage; a10 material alpha resolution; the shadow any-hit coverage test
of every trace kernel).

Scenes (mutracer_amd.scene.alpha_caster / masked_panes): a shadow caster
over a floor, opaque, masked out or covered by its opacity, alpha map or
base-map alpha, or covered per texel by a checker alpha map; and panes of a
masked-out or checker-masked material between the camera and the floor, so
that camera and secondary rays pass through them. Both sides restate the
same rules (coverage resolution, pass-through vertex without a depth,
shadow-ray coverage test), so the FP64 accumulators agree bit for bit on
both kernel forms, every sampling pattern and every wavefront launch form
(stack or persistent walker traversal, tail kernel, queue sorting).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from mutracer_amd import abi, render, scene
    return abi, render, scene


def _gpu(render, sc, w, h, spp, tuning=None, **kw):
    tgt = render.RenderTarget(render.RenderScene(sc), w, h, max_samples=spp, **kw)
    if tuning:
        tgt.set_tuning(**tuning)
    tgt.render(spp)
    tgt.synchronize()
    sums, counts = tgt.read_accumulators()
    assert np.all(counts == spp)
    return sums


@pytest.mark.parametrize("kernel", ["wavefront", "megakernel"])
@pytest.mark.parametrize("kind", ["opaque", "mask_open", "mask_covered", "alpha_map_covered", "base_alpha_open",
                                  "checker"])
def test_casters_match_oracle(env, orc, kernel, kind):
    abi, render, scene = env
    sc = scene.alpha_caster(kind)
    w, h, spp = 40, 32, 4
    got = _gpu(render, sc, w, h, spp, bounces=3, direct_light_candidates=2, kernel=kernel)
    ref = orc.OracleRenderer(sc, w, h, bounces=3, candidates=2, max_samples=spp).render(0, spp)
    assert ref[..., 0:3].mean() > 0.01
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kernel", ["wavefront", "megakernel"])
@pytest.mark.parametrize("pattern", ["sobol", "zsobol", "random"])
@pytest.mark.parametrize("panes", ["1", "3", "checker"])
def test_panes_match_oracle(env, orc, kernel, pattern, panes):
    abi, render, scene = env
    sc = scene.masked_panes(2, checker=True) if panes == "checker" else scene.masked_panes(int(panes))
    w, h, spp = 36, 28, 4
    got = _gpu(render, sc, w, h, spp, bounces=5, direct_light_candidates=2, pattern=pattern, kernel=kernel)
    ref = orc.OracleRenderer(sc, w, h, bounces=5, candidates=2, pattern=pattern, max_samples=spp).render(0, spp)
    assert ref[..., 0:3].mean() > 0.01
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("tuning", [
    dict(trace_kernel=0, tail_paths=0),
    dict(trace_kernel=1, tail_paths=0),
    dict(trace_kernel=1, tail_paths=600),
    dict(tail_paths=1 << 30),
    dict(queue_sort=1, tail_paths=0),
])
def test_wavefront_forms_match_oracle(env, orc, tuning):
    """The checker caster's shadow rays and the checker panes' pass-through
    events through every wavefront launch form."""
    abi, render, scene = env
    for sc in (scene.alpha_caster("checker"), scene.masked_panes(2, checker=True)):
        got = _gpu(render, sc, 48, 36, 4, tuning=tuning, bounces=6, direct_light_candidates=2)
        ref = orc.OracleRenderer(sc, 48, 36, bounces=6, candidates=2, max_samples=4).render(0, 4)
        assert np.array_equal(got, ref)
