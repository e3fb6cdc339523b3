"""Material textures on the GPU vs the oracle (row f4: textures, sRGB
decode, UV transforms; a7 UVs and normal maps; a10 texture samples in the
material resolve).

The textured room (mutracer_amd.scene.textured_room) puts every texture
slot kind in one scene: RGBA8 (sRGB color, linear data) and RGBA32F images,
repeat / clamp / mirrored wrap per axis, UV scale and offset, flip_y, UV1,
a normal map with a stretched UV frame, a metallic-roughness map, fuzz,
coat and transmission masks, and an emissive-triangle light whose emission
is textured (emissive map plus the coat-weight-map tint) so that the NEE
samples and the BSDF hits both read it. Both sides run the same restatement
of the reference's sampleMap / getMaterial / getMatNormal, so the FP64
accumulators must agree bit for bit, on both kernel forms and every sampling
pattern. A scene whose material needs blend coverage (stochastic
transparency) is refused with MTX_UNAVAILABLE.
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


def _render(render, sc, w, h, spp, **kw):
    tgt = render.RenderTarget(render.RenderScene(sc), w, h, **kw)
    tgt.render(spp)
    tgt.synchronize()
    return tgt.read_accumulators()


@pytest.mark.parametrize("kernel", ["wavefront", "megakernel"])
@pytest.mark.parametrize("pattern", ["sobol", "zsobol", "random"])
def test_textured_room_matches_oracle(env, orc, kernel, pattern):
    abi, render, scene = env
    sc = scene.textured_room()
    w, h, spp = 56, 40, 6
    sums, counts = _render(render, sc, w, h, spp, bounces=8, direct_light_candidates=2, pattern=pattern,
                           max_samples=spp, kernel=kernel)
    ref = orc.OracleRenderer(sc, w, h, bounces=8, candidates=2, pattern=pattern, max_samples=spp).render(0, spp)
    assert np.all(counts == spp)
    assert ref[..., 0:3].mean() > 0.01
    assert np.array_equal(sums, ref)


def test_textures_change_the_image(env, orc):
    """The maps are not inert: removing them changes the rendered image
    (checked on the oracle, whose equality with the GPU the test above
    establishes)."""
    abi, render, scene = env
    sc = scene.textured_room()
    plain = scene.textured_room()
    for m in plain.materials:
        for name in scene.MAP_SLOTS:
            setattr(m, name, None)
    a = orc.OracleRenderer(sc, 32, 24, bounces=4, max_samples=2).render(0, 2)
    b = orc.OracleRenderer(plain, 32, 24, bounces=4, max_samples=2).render(0, 2)
    assert np.abs(a[..., 0:3] - b[..., 0:3]).max() > 0.05
    assert np.abs(a[..., 8:11] - b[..., 8:11]).max() > 0.01  # the normal map moves the normal AOV


@pytest.mark.parametrize("edit", ["blend", "opacity", "alpha map"])
def test_blend_coverage_is_refused(env, edit):
    """Blend coverage (explicit, or resolved from opacity < 1 or an alpha
    map without a cutoff) is refused; mask materials are rendered (their
    parity is covered with the alpha-coverage tests)."""
    abi, render, scene = env
    sc = scene.cornell_box()
    if edit == "blend":
        sc.materials[1].alpha_mode = "blend"
    elif edit == "opacity":
        sc.materials[1].opacity = 0.5
    else:
        sc.textures = [scene.Texture(np.ones((2, 2, 4), np.float32))]
        sc.materials[1].alpha_map = 0
    with pytest.raises(abi.MtxError) as e:
        render.RenderScene(sc)
    assert e.value.code == abi.MTX_UNAVAILABLE
    assert "alpha coverage" in str(e.value)


def test_texture_validation(env):
    abi, render, scene = env
    sc = scene.textured_room()
    sc.materials[0].map = len(sc.textures)  # one past the table
    with pytest.raises(abi.MtxError) as e:
        render.RenderScene(sc)
    assert e.value.code == abi.MTX_INVALID_ARGUMENT
    sc = scene.textured_room()
    sc.textures[1].pixels = sc.textures[1].pixels.copy()
    sc.textures[1].pixels[3, 4, 1] = np.nan
    with pytest.raises(abi.MtxError) as e:
        render.RenderScene(sc)
    assert e.value.code == abi.MTX_INVALID_ARGUMENT
