"""U-Net denoiser on MFMA (FP16 / BF16 storage, FP32 accumulation).

  - against an emulated-precision oracle (every conv input, weight and output
    rounded to the storage type, f64 accumulation): agreement to a few units
    of the storage precision pins the kernels' indexing and MFMA layouts;
  - against the FP64 oracle: the error each precision actually has, stated
    per layer (printed) and bounded for the output;
  - mt_denoiser_run semantics: any size (padded to 16), alpha preserved,
    finite output.
"""
import numpy as np
import pytest

from oracle import unet

pytestmark = pytest.mark.gpu

# max |device - reference| / max |reference| per layer
EMU_TOL = {"fp16": 2e-2, "bf16": 8e-2}
FP64_OUT_TOL = {"fp16": 2e-2, "bf16": 1e-1}


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
