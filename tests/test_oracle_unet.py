"""Pins the FP64 U-Net oracle: the weights file, the network's shape and
cost, and the convolution against a direct per-pixel sum."""
import hashlib
import os

import numpy as np

from oracle import unet

WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "mutracer_amd", "data", "rt_hdr_alb_nrm.tza")


def test_weights_digest_and_table():
    data = open(WEIGHTS, "rb").read()
    # oidn/weights/README.md lists this SHA-256 for rt_hdr_alb_nrm.tza.
    assert hashlib.sha256(data).hexdigest() == "e586ef2ff48d7fbb7611986405220ed8fc5c13ca79bfc40be4dc742fbf959e1a"
    w = unet.read_tza(data)
    assert len(w) == 32 and sum(a.size for a, _ in w.values()) == 919811
    assert w["enc_conv0.weight"][0].shape == (32, 9, 3, 3) and w["enc_conv0.weight"][1] == "oihw"
    assert w["dec_conv1a.weight"][0].shape == (64, 73, 3, 3)
    assert w["dec_conv0.bias"][0].shape == (3,)


def test_cost_matches_survey_figure():
    assert abs(unet.flops(1088, 1920) / 1e9 - 523.94) < 0.01


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
