"""U-Net denoiser on MFMA (FP16 / BF16 storage, FP32 accumulation).

  - against an emulated-precision oracle (every conv input, weight and output
    rounded to the storage type, f64 accumulation): agreement to a few units
    of the storage precision pins the kernels' indexing and MFMA layouts;
  - against the FP64 oracle: the error each precision actually has, stated
    per layer (printed) and bounded for the output;
  - mt_denoiser_run semantics: any size (padded to 16), alpha preserved,
    finite output;
  - the other network shapes: the bundled small network and UNetLarge
    networks (synthetic He-scaled weights of the large models' exact shapes,
    9- and 3-channel) against the same oracles;
  - the whole run (autoexposure, transfers, padding, prefiltered guides,
    image tiles) against the oracle's restatement of it, and tiles against
    untiled runs of the same regions, bit for bit.
"""
import os
import numpy as np
import pytest

from oracle import unet

pytestmark = pytest.mark.gpu

# max |device - reference| / max |reference| per layer, about three times the
# largest error measured over these tests' networks and inputs (emulated:
# fp16 1.5e-3, bf16 1.3e-2; FP64: fp16 1.2e-3, bf16 2.2e-2).
EMU_TOL = {"fp16": 5e-3, "bf16": 4e-2}
FP64_OUT_TOL = {"fp16": 5e-3, "bf16": 7e-2}


@pytest.fixture(scope="module")
def net():
    import torch
    assert torch.cuda.is_available()
    from mutracer_amd import denoise
    return denoise, unet.read_tza(denoise.WEIGHTS)


def _input(h, w, seed=0):
    rng = np.random.default_rng(seed)
    x = np.zeros((h, w, 9), np.float32)
    base = rng.uniform(0.1, 0.6, (1, 1, 3))
    x[..., 0:3] = np.clip(base + rng.normal(0, 0.08, (h, w, 3)), 0, 1)
    x[..., 3:6] = rng.uniform(0, 1, (h, w, 3))
    n = rng.normal(size=(h, w, 3))
    x[..., 6:9] = n / np.linalg.norm(n, axis=2, keepdims=True)
    return x


def _images(h, w, seed):
    rng = np.random.default_rng(seed)
    color = np.concatenate([rng.gamma(2.0, 0.6, (h, w, 3)), rng.uniform(0, 1, (h, w, 1))], axis=2).astype(np.float32)
    albedo = np.concatenate([rng.uniform(0, 1, (h, w, 3)), np.ones((h, w, 1))], axis=2).astype(np.float32)
    n = rng.normal(size=(h, w, 3))
    normal = np.concatenate([n / np.linalg.norm(n, axis=2, keepdims=True), np.ones((h, w, 1))],
                            axis=2).astype(np.float32)
    return color, albedo, normal


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-6))


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_forward_vs_oracles(net, precision):
    denoise, w = net
    d = denoise.Denoiser(precision)
    x = _input(48, 80)
    out, layers = d.forward(x, layers=True)
    q = unet.to_fp16 if precision == "fp16" else unet.to_bf16
    emu_out, emu = unet.forward(w, x.astype(np.float64), quantize=q)
    ref_out, ref = unet.forward(w, x.astype(np.float64))
    for name, _, _ in denoise.LAYERS:
        e_emu, e_64 = _rel(layers[name], emu[name]), _rel(layers[name], ref[name])
        print(f"{precision} {name:11s} vs emulated {e_emu:.2e}  vs fp64 {e_64:.2e}")
        assert e_emu <= EMU_TOL[precision], f"{name} differs from the emulated {precision} network"
    assert _rel(out, ref_out) <= FP64_OUT_TOL[precision]
    assert np.all(np.isfinite(out))


def test_run_semantics(net):
    denoise, _ = net
    d = denoise.Denoiser("fp16")
    rng = np.random.default_rng(4)
    h, w = 37, 53
    color = np.concatenate([rng.gamma(2.0, 0.3, (h, w, 3)), rng.uniform(0, 1, (h, w, 1))], axis=2).astype(np.float32)
    albedo = np.concatenate([rng.uniform(0, 1, (h, w, 3)), np.ones((h, w, 1))], axis=2).astype(np.float32)
    n = rng.normal(size=(h, w, 3))
    normal = np.concatenate([n / np.linalg.norm(n, axis=2, keepdims=True), np.ones((h, w, 1))], axis=2).astype(np.float32)
    out = d.run(color, normal, albedo)
    assert out.shape == color.shape and np.all(np.isfinite(out))
    assert np.array_equal(out[..., 3], color[..., 3])
    assert not np.allclose(out[..., :3], color[..., :3])
    # A smooth (constant) image stays close to itself.
    flat = np.full((h, w, 4), 0.5, np.float32)
    nflat = np.zeros((h, w, 4), np.float32)
    nflat[..., 2] = 1
    o2 = d.run(flat, nflat, flat)
    assert abs(float(o2[8:-8, 8:-8, :3].mean()) - 0.5) < 0.1


STAGED = 16  # MTX_DENOISER_STAGED


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
@pytest.mark.parametrize("h,w", [(64, 96), (48, 80)])
def test_tiling_shapes_give_identical_outputs(net, precision, h, w):
    """Every launch shape of the direct and the LDS-staged kernel accumulates
    each output in the same (channel step, tap) order, so all shapes give the
    same bits, also where a block's four rows or a wave's pixel tiles run
    past the image (48 x 80: 3 x 5 pixels at the deepest level)."""
    denoise, _ = net
    d = denoise.Denoiser(precision)
    x = _input(h, w, seed=5)
    ref, ref_layers = d.forward(x, layers=True)
    for mt, nt in ((1, 1), (8, 7), (2, 3), (4, 2), (STAGED | 1, 1), (STAGED | 8, 7), (STAGED | 2, 3),
                   (STAGED | 4, 4), (STAGED | 4, 2)):
        for layer in range(16):
            d.set_tiling(layer, mt, nt)
            assert d.tiling(layer, w, h)[0] == mt
        out, layers = d.forward(x, layers=True)
        assert np.array_equal(out, ref), (mt, nt)
        assert all(np.array_equal(layers[k], ref_layers[k]) for k in ref_layers), (mt, nt)
    for bad in ((3, 1), (STAGED | 3, 1), (STAGED, 1), (2, 8)):
        with pytest.raises(Exception):
            d.set_tiling(0, *bad)


def test_denoiser_autotune(net):
    denoise, _ = net
    from mutracer_amd import autotune
    d = denoise.Denoiser("fp16")
    x = _input(64, 96, seed=5)
    ref = d.forward(x)
    best, before, after = autotune.tune_denoiser(d, 96, 64, iterations=1)
    assert len(best) == 16 and before > 0 and after > 0
    assert np.array_equal(d.forward(x), ref)


def test_create_backend_destroy_contract(net):
    """mt_denoiser_create/backend/destroy semantics (mutracer-denoise.h:9-53):
    NULL weights load the standard network shipped beside the library; AUTO
    and HIP run it; explicit other backends are unavailable (never fall back);
    destroyed handles are invalid; the result equals the weights-bytes form."""
    import ctypes as C
    from mutracer_amd import abi
    denoise, _ = net
    L = abi.lib()
    err = abi.errbuf()
    h = C.c_uint64(0)
    abi.check(L.mtx_denoiser_create(None, 0, C.byref(h), err, 512), err)
    b = C.c_uint32(99)
    abi.check(L.mtx_denoiser_backend(h.value, C.byref(b), err, 512), err)
    assert b.value == 4
    for backend, code in ((1, abi.MTX_UNAVAILABLE), (2, abi.MTX_UNAVAILABLE), (3, abi.MTX_UNAVAILABLE),
                          (5, abi.MTX_INVALID_ARGUMENT)):
        other = C.c_uint64(0)
        assert L.mtx_denoiser_create(None, backend, C.byref(other), err, 512) == code
        assert other.value == 0
    assert L.mtx_denoiser_create(b"/nonexistent.tza", 0, C.byref(C.c_uint64(0)), err, 512) == abi.MTX_UNAVAILABLE
    rng = np.random.default_rng(3)
    H, W = 24, 40
    color = rng.uniform(0, 2, (H, W, 4)).astype(np.float32)
    albedo = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    normal = rng.normal(size=(H, W, 4)).astype(np.float32)
    out = np.zeros_like(color)
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    abi.check(L.mtx_denoiser_run(h.value, W, H, P(color), P(normal), P(albedo), P(out), out.size, err, 512), err)
    ref = denoise.Denoiser("fp16").run(color, normal, albedo)
    assert np.array_equal(out, ref)
    abi.check(L.mtx_denoiser_destroy(h.value, err, 512), err)
    assert L.mtx_denoiser_destroy(h.value, err, 512) == abi.MTX_INVALID_HANDLE
    assert L.mtx_denoiser_run(h.value, W, H, P(color), P(normal), P(albedo), P(out), out.size, err,
                              512) == abi.MTX_INVALID_HANDLE


def _check_layers(d, w, x, precision):
    out, layers = d.forward(x, layers=True)
    q = unet.to_fp16 if precision == "fp16" else unet.to_bf16
    emu_out, emu = unet.forward(w, x.astype(np.float64), quantize=q)
    ref_out, ref = unet.forward(w, x.astype(np.float64))
    for name, _, _ in d.layers:
        e_emu, e_64 = _rel(layers[name], emu[name]), _rel(layers[name], ref[name])
        print(f"{precision} {name:11s} vs emulated {e_emu:.2e}  vs fp64 {e_64:.2e}")
        assert e_emu <= EMU_TOL[precision], f"{name} differs from the emulated {precision} network"
    assert _rel(out, ref_out) <= FP64_OUT_TOL[precision]
    assert np.all(np.isfinite(out))


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_small_network_vs_oracles(net, precision):
    denoise, _ = net
    path = os.path.join(denoise.WEIGHTS_DIR, "rt_hdr_alb_nrm_small.tza")
    d = denoise.Denoiser(precision, path)
    assert d.info()[:3] == (16, 9, 0)
    _check_layers(d, unet.read_tza(path), _input(48, 80, seed=6), precision)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
@pytest.mark.parametrize("in_channels", [9, 3])
def test_large_network_vs_oracles(net, precision, in_channels):
    denoise, _ = net
    w = unet.synthetic_weights(unet.network_shapes(unet.WIDTHS["large"], in_channels, True), seed=in_channels)
    d = denoise.Denoiser(precision, unet.write_tza(w))
    assert d.info() == (19, in_channels, 1, 1 if precision == "fp16" else 2)
    x = _input(32, 48, seed=7)[..., :in_channels]
    _check_layers(d, w, x, precision)


# max |device - oracle| / max |oracle| of the denoised color (measured: fp16
# 3.6e-3, 6.5e-3 with prefilters; bf16 1.9e-2)
RUN_TOL = {"fp16": 2e-2, "bf16": 6e-2}


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_run_vs_oracle(net, precision):
    """The whole mt_denoiser_run at a size that is no multiple of 16, with
    autoexposure and with a fixed input scale, against the oracle's run
    (storage precision emulated)."""
    denoise, w = net
    d = denoise.Denoiser(precision)
    q = unet.to_fp16 if precision == "fp16" else unet.to_bf16
    color, albedo, normal = _images(37, 53, seed=8)
    for scale in (None, 0.7):
        d.set_input_scale(scale)
        out = d.run(color, normal, albedo)
        ref = unet.run(w, color, albedo, normal, input_scale=scale, quantize=q)
        err = _rel(out[..., :3], ref[..., :3])
        print(f"{precision} run scale={scale}: {err:.2e}")
        assert err <= RUN_TOL[precision]
        assert np.array_equal(out[..., 3], color[..., 3])
    # Autoexposure is on by default and equals the oracle's scale as a fixed one.
    d.set_input_scale(unet.autoexposure(color))
    fixed = d.run(color, normal, albedo)
    d.set_input_scale(None)
    assert _rel(d.run(color, normal, albedo)[..., :3], fixed[..., :3]) < 1e-2


def test_prefiltered_guides_vs_oracle(net):
    """Clean-auxiliary pipeline: albedo and normal through 3-channel UNetLarge
    prefilters (synthetic weights), then the beauty network (the bundled
    small one) on the decoded guides."""
    denoise, _ = net
    small = os.path.join(denoise.WEIGHTS_DIR, "rt_hdr_alb_nrm_small.tza")
    wa = unet.synthetic_weights(unet.network_shapes(unet.WIDTHS["large"], 3, True), seed=11)
    wn = unet.synthetic_weights(unet.network_shapes(unet.WIDTHS["large"], 3, True), seed=12)
    d = denoise.Denoiser("fp16", small)
    fa, fn = denoise.Denoiser("fp16", unet.write_tza(wa)), denoise.Denoiser("fp16", unet.write_tza(wn))
    color, albedo, normal = _images(40, 52, seed=9)
    plain = d.run(color, normal, albedo)
    d.set_auxiliary(fa, fn)
    out = d.run(color, normal, albedo)
    ref = unet.run(unet.read_tza(small), color, albedo, normal, aux=(wa, wn), quantize=unet.to_fp16)
    err = _rel(out[..., :3], ref[..., :3])
    print(f"prefiltered run: {err:.2e}")
    assert err <= RUN_TOL["fp16"]
    assert _rel(out[..., :3], plain[..., :3]) > 10 * err  # the prefilters changed the guides
    d.set_auxiliary(None, None)
    assert np.array_equal(d.run(color, normal, albedo), plain)


def test_tiles(net):
    """A tile covering the image gives the untiled bits; 32-pixel tiles with
    16 pixels of overlap on a 96 x 64 image (regions of 48 or 64 pixels)
    equal untiled runs of those regions, stitched; tiles on an unaligned
    image follow the oracle's tiled run."""
    denoise, w = net
    d = denoise.Denoiser("fp16")
    d.set_input_scale(0.8)  # tiles of one image share its scale; crops would not
    color, albedo, normal = _images(64, 96, seed=10)
    base = d.run(color, normal, albedo)
    d.set_tiles(128, 0)
    assert np.array_equal(d.run(color, normal, albedo), base)
    d.set_tiles(32, 16)
    tiled = d.run(color, normal, albedo)
    d.set_tiles(0, 0)
    expect = np.zeros_like(tiled)
    for ty in range(0, 64, 32):
        for tx in range(0, 96, 32):
            y0, y1, x0, x1 = max(0, ty - 16), min(64, ty + 48), max(0, tx - 16), min(96, tx + 48)
            crop = [np.ascontiguousarray(a[y0:y1, x0:x1]) for a in (color, normal, albedo)]
            r = d.run(*crop)
            expect[ty:ty + 32, tx:tx + 32] = r[ty - y0:ty - y0 + 32, tx - x0:tx - x0 + 32]
    assert np.array_equal(tiled, expect)
    assert not np.array_equal(tiled, base)
    c2, a2, n2 = _images(45, 70, seed=12)
    d.set_tiles(32, 16)
    out = d.run(c2, n2, a2)
    ref = unet.run(w, c2, a2, n2, input_scale=0.8, tile=32, overlap=16, quantize=unet.to_fp16)
    assert _rel(out[..., :3], ref[..., :3]) <= RUN_TOL["fp16"]


def test_denoiser_validation(net):
    from mutracer_amd import abi
    denoise, _ = net
    d = denoise.Denoiser("fp16")
    for bad in ((17, 0), (32, 8)):
        with pytest.raises(abi.MtxError):
            d.set_tiles(*bad)
    for bad in (-1.0, 0.0, float("inf")):
        with pytest.raises(abi.MtxError):
            d.set_input_scale(bad)
    three = denoise.Denoiser("fp16", unet.write_tza(unet.synthetic_weights(
        unet.network_shapes([16] * 15 + [3], 3, False))))
    nine = denoise.Denoiser("fp16", os.path.join(denoise.WEIGHTS_DIR, "rt_hdr_alb_nrm_small.tza"))
    bf = denoise.Denoiser("bf16", unet.write_tza(unet.synthetic_weights(
        unet.network_shapes([16] * 15 + [3], 3, False))))
    with pytest.raises(abi.MtxError):
        d.set_auxiliary(nine, three)  # a 9-channel prefilter
    with pytest.raises(abi.MtxError):
        d.set_auxiliary(three, bf)  # mixed precision
    with pytest.raises(abi.MtxError):
        d.set_auxiliary(three, three)  # the same network twice
    color, albedo, normal = _images(16, 16, seed=1)
    with pytest.raises(abi.MtxError):
        three.run(color, normal, albedo)  # a prefilter network is no beauty network
    with pytest.raises(ValueError):
        three.forward(np.zeros((16, 16, 9), np.float32))


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_rendered_cornell_run_vs_oracle(net, precision):
    """mtx_denoiser_run on a rendered Cornell box (80 x 60 at 4 spp: the
    color, albedo and normal of read_pixels, normals decoded from their
    display encoding to [-1, 1]) against the oracle's run of the same
    images; the denoised image is also smoother than the noisy one."""
    denoise, w = net
    from mutracer_amd import render, scene
    t = render.RenderTarget(render.RenderScene(scene.cornell_box()), 80, 60, bounces=8, max_samples=4)
    t.render(4)
    t.synchronize()
    color, albedo, enc = (t.read_pixels(k) for k in ("color", "albedo", "normal"))
    n = np.where(enc[..., 3:4] > 0, enc[..., :3] * 2.0 - 1.0, 0.0)
    normal = np.concatenate([n, np.ones_like(enc[..., 3:4])], axis=2).astype(np.float32)
    d = denoise.Denoiser(precision)
    out = d.run(color, normal, albedo)
    q = unet.to_fp16 if precision == "fp16" else unet.to_bf16
    ref = unet.run(w, color, albedo, normal, quantize=q)
    err = _rel(out[..., :3], ref[..., :3])
    print(f"{precision} rendered Cornell run: {err:.2e}")
    assert err <= RUN_TOL[precision]
    # Neighbour differences inside the back wall drop (the noise is removed).
    def rough(img):
        c = img[10:40, 25:55, :3]
        return float(np.abs(np.diff(c, axis=1)).mean())
    print(f"{precision} neighbour difference: noisy {rough(color):.4f}, denoised {rough(out):.4f}")
    assert rough(out) < 0.5 * rough(color)
