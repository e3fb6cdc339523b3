"""Display composition and the JNLM denoiser on the GPU vs the oracle.

Both sides read the same FP64 accumulators and evaluate the same f32
operations in the WGSL's order with the shared deterministic pow/exp/log2, so
the composed RGBA8 image and the JNLM's filtered display-linear RGBA agree
bit for bit with oracle/orc_display.cpp (compose.frag.wgsl.ts:100-142,
jnlm.compute.ts:47-318), for every tone curve, with grading and vignette, and
for non-default filter shapes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def target():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from mutracer_amd import render, scene
    sc = scene.instanced_field(n_instances=9, freq=10, env_size=(128, 64))
    t = render.RenderTarget(render.RenderScene(sc), 72, 40, bounces=8, direct_light_candidates=2, max_samples=8)
    t.render(8)
    return t


@pytest.mark.parametrize("tone", ["agx", "aces", "cineon", "reinhard", "linear"])
def test_compose_matches_oracle(target, orc, tone):
    sums, _ = target.read_accumulators()
    got = target.compose(tone)
    ref = orc.compose(sums, dict(tone_mapping=tone))
    assert np.array_equal(got, ref)
    grading = dict(exposure=1.7, contrast=1.2, saturation=0.8, temperature=0.3, tint=-0.2, vignette=0.5)
    got = target.compose(tone, gamma22=True, **grading)
    ref = orc.compose(sums, dict(tone_mapping=tone, gamma22=1, **grading))
    assert np.array_equal(got, ref)
    assert got.std() > 1.0


@pytest.mark.parametrize("opts", [dict(), dict(strength=2.5, search_radius=3, patch_radius=2, iterations=3),
                                  dict(patch_radius=0, albedo_bandwidth=0.2, normal_bandwidth=0.1, iterations=1),
                                  dict(patch_radius=3, search_radius=7, iterations=4)])
def test_jnlm_matches_oracle(target, orc, opts):
    sums, _ = target.read_accumulators()
    lin, enc = target.denoise_jnlm("agx", **opts)
    ref_lin, ref_enc = orc.jnlm(sums, dict(tone_mapping="agx"), **opts)
    assert np.array_equal(lin, ref_lin)
    assert np.array_equal(enc, ref_enc)
    raw = target.compose("agx")
    # the filter changes the noisy image where it has surfaces
    assert np.abs(enc.astype(int) - raw.astype(int)).max() > 0
