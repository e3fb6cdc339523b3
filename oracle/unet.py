"""FP64 oracle of the OIDN RT U-Net -- test infrastructure only.

Restates the denoiser network the reference runs through oidn-webgpu
(oidnNetwork.ts:105-141, weights oidn/weights/rt_hdr_alb_nrm.tza, OIDN 2.3):

  enc_conv0 9->32, enc_conv1 32->32, pool1; enc_conv2 ->48, pool2;
  enc_conv3 ->64, pool3; enc_conv4 ->80, pool4; enc_conv5a ->96, enc_conv5b ->96;
  up(x2, nearest) + concat(pool3) -> dec_conv4a 160->112, dec_conv4b 112->112;
  up + concat(pool2) -> dec_conv3a 160->96, dec_conv3b 96->96;
  up + concat(pool1) -> dec_conv2a 128->64, dec_conv2b 64->64;
  up + concat(input) -> dec_conv1a 73->64, dec_conv1b 64->32; dec_conv0 32->3.

Every convolution is 3x3, stride 1, zero padding 1, followed by ReLU except
dec_conv0; pools are 2x2 max; concatenation puts the upsampled tensor first.
Tensors are HWC float64; weights are the TZA's fp16 values widened exactly.
The network code itself (npm oidn-webgpu 0.6.1) is not in the reference
checkout, so the layer order above is OIDN's published U-Net; parity with the
browser filter is unpinned, the oracle pins the GPU kernels' arithmetic.
"""
from __future__ import annotations

import struct

import numpy as np

LAYERS = ["enc_conv0", "enc_conv1", "enc_conv2", "enc_conv3", "enc_conv4", "enc_conv5a", "enc_conv5b",
          "dec_conv4a", "dec_conv4b", "dec_conv3a", "dec_conv3b", "dec_conv2a", "dec_conv2b",
          "dec_conv1a", "dec_conv1b", "dec_conv0"]


def read_tza(path_or_bytes):
    """Tensor archive: u16 magic 0x41D7, u8 major 2, u8 minor, u64 table offset;
    table = u32 count, then per tensor: u16 name length, name, u8 rank,
    u32 dims, rank layout chars, u8 dtype ('h' fp16, 'f' fp32), u64 data offset."""
    d = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    magic, major = struct.unpack_from("<HB", d, 0)
    if magic != 0x41D7 or major != 2:
        raise ValueError("not a version-2 tensor archive")
    (table,) = struct.unpack_from("<Q", d, 4)
    (n,) = struct.unpack_from("<I", d, table)
    p = table + 4
    out = {}
    for _ in range(n):
        (nl,) = struct.unpack_from("<H", d, p)
        p += 2
        name = d[p:p + nl].decode()
        p += nl
        rank = d[p]
        p += 1
        shape = struct.unpack_from("<" + "I" * rank, d, p)
        p += 4 * rank
        layout = d[p:p + rank].decode()
        p += rank
        dtype = chr(d[p])
        p += 1
        (off,) = struct.unpack_from("<Q", d, p)
        p += 8
        count = int(np.prod(shape))
        np_t = {"h": "<f2", "f": "<f4"}[dtype]
        arr = np.frombuffer(d, dtype=np_t, count=count, offset=off).reshape(shape)
        out[name] = (arr, layout)
    return out


def conv3x3(x, w, b, relu=True):
    """x: (H, W, Cin) f64, w: (Cout, Cin, 3, 3), b: (Cout,). Zero padding 1."""
    H, W, _ = x.shape
    xp = np.zeros((H + 2, W + 2, x.shape[2]), np.float64)
    xp[1:-1, 1:-1] = x
    y = np.zeros((H, W, w.shape[0]), np.float64)
    for dy in range(3):
        for dx in range(3):
            tap = w[:, :, dy, dx].astype(np.float64).T  # (Cin, Cout)
            y += xp[dy:dy + H, dx:dx + W].reshape(-1, x.shape[2]) .dot(tap).reshape(H, W, -1)
    y += b.astype(np.float64)
    return np.maximum(y, 0.0) if relu else y


def pool2(x):
    H, W, C = x.shape
    return x.reshape(H // 2, 2, W // 2, 2, C).max(axis=(1, 3))


def up2(x):
    return x.repeat(2, axis=0).repeat(2, axis=1)


def forward(weights, x, quantize=None):
    """Runs the U-Net on x (H, W, 9), H and W multiples of 16. Returns
    (output (H, W, 3), {layer: output}). `quantize`, if given, maps every
    conv input and weight to the storage precision of an emulated device run
    (accumulation stays f64)."""
    q = quantize or (lambda a: a)
    layers = {}

    def conv(name, t, relu=True):
        w, _ = weights[name + ".weight"]
        b, _ = weights[name + ".bias"]
        y = conv3x3(q(t), q(w.astype(np.float64)), b.astype(np.float64), relu)
        y = q(y) if quantize else y
        layers[name] = y
        return y

    x = np.asarray(x, np.float64)
    e0 = conv("enc_conv0", x)
    p1 = pool2(conv("enc_conv1", e0))
    p2 = pool2(conv("enc_conv2", p1))
    p3 = pool2(conv("enc_conv3", p2))
    p4 = pool2(conv("enc_conv4", p3))
    e5 = conv("enc_conv5b", conv("enc_conv5a", p4))
    d4 = conv("dec_conv4b", conv("dec_conv4a", np.concatenate([up2(e5), p3], axis=2)))
    d3 = conv("dec_conv3b", conv("dec_conv3a", np.concatenate([up2(d4), p2], axis=2)))
    d2 = conv("dec_conv2b", conv("dec_conv2a", np.concatenate([up2(d3), p1], axis=2)))
    d1 = conv("dec_conv1b", conv("dec_conv1a", np.concatenate([up2(d2), q(x)], axis=2)))
    out = conv("dec_conv0", d1, relu=False)
    return out, layers


def to_fp16(a):
    return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)


def to_bf16(a):
    """Round-to-nearest-even to bfloat16 (as float64)."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def flops(h, w):
    """2 x multiply-adds of the 16 convolutions at an h x w input."""
    shapes = [(0, 9, 32), (0, 32, 32), (1, 32, 48), (2, 48, 64), (3, 64, 80), (4, 80, 96), (4, 96, 96),
              (3, 160, 112), (3, 112, 112), (2, 160, 96), (2, 96, 96), (1, 128, 64), (1, 64, 64),
              (0, 73, 64), (0, 64, 32), (0, 32, 3)]
    return sum(2 * 9 * ci * co * (h >> lv) * (w >> lv) for lv, ci, co in shapes)
