"""Pins the FP64 U-Net oracle: the weights files, the networks' shapes and
cost, the convolution against a direct per-pixel sum, the transfer curves
by their closed-form properties, the autoexposure against a brute-force sum
over explicit bins, and the tiling against the untiled run."""
import hashlib
import os

import numpy as np
import pytest

from oracle import unet

DATA = os.path.join(os.path.dirname(__file__), "..", "mutracer_amd", "data")
WEIGHTS = os.path.join(DATA, "rt_hdr_alb_nrm.tza")
SMALL = os.path.join(DATA, "rt_hdr_alb_nrm_small.tza")
REF_WEIGHTS = "/root/reference/oidn/weights"
# oidn/weights/README.md lists these SHA-256 digests.
DIGESTS = {"rt_hdr_alb_nrm.tza": "e586ef2ff48d7fbb7611986405220ed8fc5c13ca79bfc40be4dc742fbf959e1a",
           "rt_hdr_alb_nrm_small.tza": "980ea307e7825cae23ed8b1addf8de62ece84b61094795a5631ebf1c91b2e1a7",
           "rt_hdr_calb_cnrm_large.tza": "08130cdc4f33b7a513d0073757f089ac830f900613bb5bad72ca42c52b934465",
           "rt_alb_large.tza": "a8efbe6eed511f41c8f5ff9145923214cefb0703cf91a3b00a1dc131ade1600a",
           "rt_nrm_large.tza": "8ac7705ec5ac303116b40ad2f0880cd741ed1f4280828bf2f18f7635c3377257"}


def test_weights_digest_and_table():
    data = open(WEIGHTS, "rb").read()
    # oidn/weights/README.md lists this SHA-256 for rt_hdr_alb_nrm.tza.
    assert hashlib.sha256(data).hexdigest() == "e586ef2ff48d7fbb7611986405220ed8fc5c13ca79bfc40be4dc742fbf959e1a"
    w = unet.read_tza(data)
    assert len(w) == 32 and sum(a.size for a, _ in w.values()) == 919811
    assert w["enc_conv0.weight"][0].shape == (32, 9, 3, 3) and w["enc_conv0.weight"][1] == "oihw"
    assert w["dec_conv1a.weight"][0].shape == (64, 73, 3, 3)
    assert w["dec_conv0.bias"][0].shape == (3,)


def test_cost_per_1080p_pass():
    """2 x multiply-adds of the 16 convolutions at 1920 x 1088, summed here
    from the weight tensors' shapes and each layer's resolution level (the
    encoder halves it after enc_conv1..4, the decoder doubles it back):
    523.94 GFLOP, and the oracle's own count agrees."""
    w = unet.read_tza(open(WEIGHTS, "rb").read())
    level = {"enc_conv0": 0, "enc_conv1": 0, "enc_conv2": 1, "enc_conv3": 2, "enc_conv4": 3, "enc_conv5a": 4,
             "enc_conv5b": 4, "dec_conv4a": 3, "dec_conv4b": 3, "dec_conv3a": 2, "dec_conv3b": 2, "dec_conv2a": 1,
             "dec_conv2b": 1, "dec_conv1a": 0, "dec_conv1b": 0, "dec_conv0": 0}
    total = 0
    for name, lv in level.items():
        cout, cin, kh, kw = w[name + ".weight"][0].shape
        total += 2 * (1088 >> lv) * (1920 >> lv) * kh * kw * cin * cout
    assert abs(total / 1e9 - 523.94) < 0.01
    assert abs(unet.flops(1088, 1920) - total) < 1.0


def test_conv_equals_direct_sum():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(5, 7, 3))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    y = unet.conv3x3(x, w, b, relu=False)
    for yy in range(5):
        for xx in range(7):
            for o in range(4):
                s = b[o]
                for dy in range(3):
                    for dx in range(3):
                        iy, ix = yy + dy - 1, xx + dx - 1
                        if 0 <= iy < 5 and 0 <= ix < 7:
                            s += np.dot(x[iy, ix], w[o, :, dy, dx])
                assert abs(y[yy, xx, o] - s) < 1e-12


def test_forward_shapes_and_skip_order():
    w = unet.read_tza(WEIGHTS)
    x = np.random.default_rng(0).uniform(0, 1, (32, 48, 9))
    out, layers = unet.forward(w, x)
    assert out.shape == (32, 48, 3)
    assert layers["enc_conv5b"].shape == (2, 3, 96) and layers["dec_conv1a"].shape == (32, 48, 64)
    # The first decoder conv sees the upsampled tensor first: changing only the
    # skip input (pool3 path) must change dec_conv4a through its last 64 channels.
    wt = w["dec_conv4a.weight"][0].astype(np.float64)
    e5 = layers["enc_conv5b"]
    p3 = unet.pool2(layers["enc_conv3"])
    cat = np.concatenate([unet.up2(e5), p3], axis=2)
    ref = unet.conv3x3(cat, wt, w["dec_conv4a.bias"][0].astype(np.float64))
    assert np.allclose(ref, layers["dec_conv4a"])


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0])
    r = unet.to_bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12


def test_bundled_small_network():
    data = open(SMALL, "rb").read()
    assert hashlib.sha256(data).hexdigest() == DIGESTS["rt_hdr_alb_nrm_small.tza"]
    w = unet.read_tza(data)
    assert unet.plan(w) is unet.UNET
    shapes = unet.network_shapes(unet.WIDTHS["small"], 9, False)
    assert {k: a.shape for k, (a, _) in w.items() if k.endswith(".weight")} == {k + ".weight": v for k, v in
                                                                                  shapes.items()}


@pytest.mark.skipif(not os.path.isdir(REF_WEIGHTS), reason="the reference's weights directory is absent")
@pytest.mark.parametrize("name,in_channels", [("rt_hdr_calb_cnrm_large.tza", 9), ("rt_alb_large.tza", 3),
                                              ("rt_nrm_large.tza", 3)])
def test_large_network_shapes(name, in_channels):
    """The UNetLarge plan's channel bookkeeping (each decoder level's skip is
    the pooled output of the encoder level above it, the input at level 0)
    reproduces every weight shape of the three large networks."""
    data = open(os.path.join(REF_WEIGHTS, name), "rb").read()
    assert hashlib.sha256(data).hexdigest() == DIGESTS[name]
    w = unet.read_tza(data)
    assert unet.plan(w) is unet.UNET_LARGE and len(w) == 38
    shapes = unet.network_shapes(unet.WIDTHS["large"], in_channels, True)
    for layer, shape in shapes.items():
        assert w[layer + ".weight"][0].shape == shape, layer
        assert w[layer + ".bias"][0].shape == (shape[0],)


def test_synthetic_archive_round_trip_and_large_forward():
    shapes = unet.network_shapes(unet.WIDTHS["large"], 3, True)
    w = unet.synthetic_weights(shapes, seed=1)
    r = unet.read_tza(unet.write_tza(w))
    assert all(np.array_equal(r[k][0], w[k][0]) and r[k][1] == w[k][1] for k in w)
    x = np.random.default_rng(1).uniform(0, 1, (32, 16, 3))
    out, layers = unet.forward(r, x)
    assert out.shape == (32, 16, 3) and len(layers) == 19
    assert layers["enc_conv5b"].shape == (2, 1, 256) and layers["dec_conv1b"].shape == (32, 16, 64)
    # dec_conv3a reads up(dec_conv4b) then the pooled enc_conv2b (level 2).
    cat = np.concatenate([unet.up2(layers["dec_conv4b"]), unet.pool2(layers["enc_conv2b"])], axis=2)
    ref = unet.conv3x3(cat, r["dec_conv3a.weight"][0].astype(np.float64), r["dec_conv3a.bias"][0].astype(np.float64))
    assert np.allclose(ref, layers["dec_conv3a"])
    # dec_conv1a reads up(dec_conv2b) then the network input.
    cat = np.concatenate([unet.up2(layers["dec_conv2b"]), x], axis=2)
    ref = unet.conv3x3(cat, r["dec_conv1a.weight"][0].astype(np.float64), r["dec_conv1a.bias"][0].astype(np.float64))
    assert np.allclose(ref, layers["dec_conv1a"])
    assert unet.flops(32, 16, r) == sum(2 * 9 * s[0] * s[1] * (32 >> lv) * (16 >> lv)
                                        for (n, lv, _), s in zip(unet.UNET_LARGE, shapes.values()))


def test_pu_curve_closed_forms():
    """The three pieces meet at the breakpoints (X0 = pu(Y0), X1 = pu(Y1)),
    the curve increases, the inverse undoes it, and the normalization maps
    65504 to 1."""
    assert abs(unet.pu_forward(unet.PU_Y0) - unet.PU_X0) < 1e-9
    assert abs(unet.pu_forward(unet.PU_Y1) - unet.PU_X1) < 1e-7
    eps = 1e-12
    for y in (unet.PU_Y0, unet.PU_Y1):
        assert abs(unet.pu_forward(y * (1 + eps)) - unet.pu_forward(y)) < 1e-7
    y = np.logspace(-9, 4.8, 2000)
    f = unet.pu_forward(y)
    assert np.all(np.diff(f) > 0)
    assert np.allclose(unet.pu_inverse(f), y, rtol=1e-6, atol=1e-12)
    assert abs(unet.pu_forward(65504.0) * unet.PU_NORM - 1.0) < 1e-15
    assert unet.pu_forward(0.0) == 0.0


def test_srgb_closed_forms():
    v = np.linspace(0, 1, 1001)
    assert np.allclose(unet.srgb_decode(unet.srgb_encode(v)), v, atol=1e-12)
    assert abs(unet.srgb_encode(0.5) - 0.7353569830524495) < 1e-12  # 1.055 * 0.5^(1/2.4) - 0.055
    assert abs(unet.srgb_encode(2.0) - 1.0) < 1e-15 and unet.srgb_encode(-1.0) == 0.0
    # The two pieces of each curve meet at their thresholds.
    assert abs(12.92 * 0.0031308 - (1.055 * 0.0031308 ** (1 / 2.4) - 0.055)) < 1e-6
    assert abs(0.04045 / 12.92 - ((0.04045 + 0.055) / 1.055) ** 2.4) < 1e-6


def test_autoexposure_brute_force():
    rng = np.random.default_rng(7)
    c = rng.gamma(1.5, 0.4, (37, 21, 4))
    c[:12, :] = 0  # the first row of bins is dark and left out of the mean
    lum = 0.212671 * c[..., 0] + 0.715160 * c[..., 1] + 0.072169 * c[..., 2]
    logs = []
    for by in range(3):  # ceil(37 / 16) x ceil(21 / 16) bins
        for bx in range(2):
            y0, y1 = by * 37 // 3, (by + 1) * 37 // 3
            x0, x1 = bx * 21 // 2, (bx + 1) * 21 // 2
            m = lum[y0:y1, x0:x1].sum() / ((y1 - y0) * (x1 - x0))
            if m > 1e-8:
                logs.append(np.log2(m))
    assert len(logs) == 4
    assert abs(unet.autoexposure(c) - 0.18 / 2 ** (sum(logs) / len(logs))) < 1e-12
    assert abs(unet.autoexposure(np.full((40, 40, 4), 2.0)) - 0.09) < 1e-12
    assert unet.autoexposure(np.zeros((8, 8, 4))) == 1.0
    bad = np.full((16, 16, 4), 1.0)
    bad[0, 0, 0] = np.nan
    assert np.isfinite(unet.autoexposure(bad))


def test_run_tiling_and_transfers():
    """A tile covering the image equals the untiled run; the prefilter
    guides replace the raw ones; a fixed input scale equal to the
    autoexposure result gives the automatic output."""
    w = unet.synthetic_weights(unet.network_shapes(unet.WIDTHS["small"], 9, False), seed=2)
    aux = tuple(unet.synthetic_weights(unet.network_shapes([16] * 15 + [3], 3, False), seed=s) for s in (3, 4))
    rng = np.random.default_rng(5)
    h, wd = 20, 36
    color = np.concatenate([rng.gamma(2.0, 0.3, (h, wd, 3)), rng.uniform(0, 1, (h, wd, 1))], axis=2)
    albedo = rng.uniform(0, 1, (h, wd, 4))
    n = rng.normal(size=(h, wd, 3))
    normal = np.concatenate([n / np.linalg.norm(n, axis=2, keepdims=True), np.ones((h, wd, 1))], axis=2)
    base = unet.run(w, color, albedo, normal)
    assert np.array_equal(base[..., 3], color[..., 3])
    assert np.array_equal(unet.run(w, color, albedo, normal, tile=64), base)
    assert np.array_equal(unet.run(w, color, albedo, normal, input_scale=unet.autoexposure(color)), base)
    assert not np.allclose(unet.run(w, color, albedo, normal, input_scale=1.0), base)
    tiled = unet.run(w, color, albedo, normal, tile=16, overlap=16)
    assert not np.array_equal(tiled, base) and np.all(np.isfinite(tiled))
    pre = unet.run(w, color, albedo, normal, aux=aux)
    assert not np.allclose(pre, base)
