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
