"""FP64 oracle of the OIDN RT U-Nets and mt_denoiser_run -- test
infrastructure only.

Restates the denoiser networks the reference runs through oidn-webgpu
(oidnNetwork.ts:94-141; weights oidn/weights/*.tza, OIDN 2.3) and through
native OIDN (native/host/src/denoise.rs:140-215, the RT filter with hdr on):

"UNet" (rt_hdr_alb_nrm, rt_hdr_alb_nrm_small):
  enc_conv0, enc_conv1, pool1; enc_conv2, pool2; enc_conv3, pool3;
  enc_conv4, pool4; enc_conv5a, enc_conv5b;
  up(x2, nearest) + concat(pool3) -> dec_conv4a, dec_conv4b;
  up + concat(pool2) -> dec_conv3a, dec_conv3b;
  up + concat(pool1) -> dec_conv2a, dec_conv2b;
  up + concat(input) -> dec_conv1a, dec_conv1b; dec_conv0.
"UNetLarge" (rt_hdr_calb_cnrm_large, rt_alb_large, rt_nrm_large): two
  convolutions per encoder level (enc_convNa, enc_convNb, the second pooled
  for N = 1..4), enc_conv5a/b, the same decoder, and dec_conv1a, dec_conv1b,
  dec_conv1c at full resolution.

Every convolution is 3x3, stride 1, zero padding 1, followed by ReLU except
the last; pools are 2x2 max; concatenation puts the upsampled tensor first.
Channel counts come from the weights. Tensors are HWC float64; weights are
the TZA's fp16 values widened exactly.

run() restates mt_denoiser_run: OIDN's RT input processing (HDR color times
the input scale -- OIDN's autoexposure unless given -- through the 'pu'
transfer normalized so that 65504 maps to 1; albedo clamped to [0, 1];
normals clamped to [-1, 1] and mapped to [0, 1]), zero padding to a multiple
of 16, the network, and the inverse transfer; with prefilters, the albedo
(sRGB-encoded, oidnAuxiliary.wgsl.ts:2-16) and the normal (mapped to [0, 1])
first run through their 3-channel networks and are decoded back; with tiles,
each tile runs on its region grown by the overlap (oidn-webgpu's
maxTileSize / tileOverlap, WGPURenderingPipeline.ts:1330-1342).

The network code itself (npm oidn-webgpu 0.6.1) and OIDN's sources are not
in the reference checkout, so the layer order, the pu curve and the
autoexposure above are OIDN's published ones; parity with those libraries is
unpinned -- the oracle pins the GPU kernels' arithmetic, and its transfer
functions are pinned by their own closed-form properties.
"""
from __future__ import annotations

import struct

import numpy as np

# (name, level, pooled buffer or None) in network order.
UNET = [("enc_conv0", 0, None), ("enc_conv1", 0, 0), ("enc_conv2", 1, 1), ("enc_conv3", 2, 2), ("enc_conv4", 3, 3),
        ("enc_conv5a", 4, None), ("enc_conv5b", 4, None), ("dec_conv4a", 3, None), ("dec_conv4b", 3, None),
        ("dec_conv3a", 2, None), ("dec_conv3b", 2, None), ("dec_conv2a", 1, None), ("dec_conv2b", 1, None),
        ("dec_conv1a", 0, None), ("dec_conv1b", 0, None), ("dec_conv0", 0, None)]
UNET_LARGE = [("enc_conv1a", 0, None), ("enc_conv1b", 0, 0), ("enc_conv2a", 1, None), ("enc_conv2b", 1, 1),
              ("enc_conv3a", 2, None), ("enc_conv3b", 2, 2), ("enc_conv4a", 3, None), ("enc_conv4b", 3, 3),
              ("enc_conv5a", 4, None), ("enc_conv5b", 4, None), ("dec_conv4a", 3, None), ("dec_conv4b", 3, None),
              ("dec_conv3a", 2, None), ("dec_conv3b", 2, None), ("dec_conv2a", 1, None), ("dec_conv2b", 1, None),
              ("dec_conv1a", 0, None), ("dec_conv1b", 0, None), ("dec_conv1c", 0, None)]
LAYERS = [n for n, _, _ in UNET]


def plan(weights):
    """The layer plan a weight set selects by its names."""
    return UNET_LARGE if "enc_conv1a.weight" in weights else UNET


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
    """Runs the U-Net on x (H, W, C), H and W multiples of 16, C the first
    layer's input channels. Returns (output (H, W, 3), {layer: output}).
    `quantize`, if given, maps every conv input and weight to the storage
    precision of an emulated device run (accumulation stays f64)."""
    q = quantize or (lambda a: a)
    layers = {}
    steps = plan(weights)
    x = np.asarray(x, np.float64)
    pools = {}
    t = x
    for i, (name, level, pool) in enumerate(steps):
        if name.startswith("dec_") and name.endswith("a"):
            skip = q(x) if level == 0 else pools[level - 1]
            t = np.concatenate([up2(t), skip], axis=2)
        w, _ = weights[name + ".weight"]
        b, _ = weights[name + ".bias"]
        y = conv3x3(q(t), q(w.astype(np.float64)), b.astype(np.float64), relu=i + 1 < len(steps))
        t = q(y) if quantize else y
        layers[name] = t
        if pool is not None:
            t = pools[pool] = pool2(t)
    return t, layers


def to_fp16(a):
    return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)


def to_bf16(a):
    """Round-to-nearest-even to bfloat16 (as float64)."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def flops(h, w, weights=None):
    """2 x multiply-adds of the convolutions at an h x w input (the standard
    network's 16 unless `weights` gives another)."""
    if weights is None:
        shapes = [(0, 9, 32), (0, 32, 32), (1, 32, 48), (2, 48, 64), (3, 64, 80), (4, 80, 96), (4, 96, 96),
                  (3, 160, 112), (3, 112, 112), (2, 160, 96), (2, 96, 96), (1, 128, 64), (1, 64, 64),
                  (0, 73, 64), (0, 64, 32), (0, 32, 3)]
    else:
        shapes = [(lv, weights[n + ".weight"][0].shape[1], weights[n + ".weight"][0].shape[0])
                  for n, lv, _ in plan(weights)]
    return sum(2 * 9 * ci * co * (h >> lv) * (w >> lv) for lv, ci, co in shapes)


# ------------------------------------------------------------ transfers
# OIDN's perceptually uniform ('pu') HDR curve: y <= Y0 linear, y <= Y1 a
# power law, above a logarithm; X0 = pu(Y0), X1 = pu(Y1).
PU_A, PU_B, PU_C, PU_D = 1.41283765e+03, 1.64593172e+00, 4.31384981e-01, -2.94139609e-03
PU_E, PU_F, PU_G = 1.92653254e-01, 6.26026094e-03, 9.98620152e-01
PU_Y0, PU_Y1, PU_X0, PU_X1 = 1.57945760e-06, 3.22087631e-02, 2.23151711e-03, 3.70974749e-01


def pu_forward(y):
    y = np.asarray(y, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(y <= PU_Y0, PU_A * y,
                        np.where(y <= PU_Y1, PU_B * np.power(np.maximum(y, 1e-300), PU_C) + PU_D,
                                 PU_E * np.log(y + PU_F) + PU_G))


def pu_inverse(x):
    x = np.asarray(x, np.float64)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return np.where(x <= PU_X0, x / PU_A,
                        np.where(x <= PU_X1, np.power(np.maximum((x - PU_D) / PU_B, 0.0), 1.0 / PU_C),
                                 np.exp((x - PU_G) / PU_E) - PU_F))


PU_NORM = 1.0 / float(pu_forward(65504.0))
FLT_MAX = float(np.finfo(np.float32).max)


def srgb_encode(v):
    """Linear -> sRGB of values clamped to [0, 1] (oidnEncodeAlbedo)."""
    lin = np.clip(np.asarray(v, np.float64), 0.0, 1.0)
    return np.where(lin <= 0.0031308, 12.92 * lin, 1.055 * np.power(lin, 1.0 / 2.4) - 0.055)


def srgb_decode(v):
    """sRGB -> linear of values clamped to [0, 1] (oidnDecodeAlbedo)."""
    e = np.clip(np.asarray(v, np.float64), 0.0, 1.0)
    return np.where(e <= 0.04045, e / 12.92, np.power((e + 0.055) / 1.055, 2.4))


def autoexposure(color):
    """OIDN's autoexposure: ceil(h/16) x ceil(w/16) bins with bounds
    b * h // nbh, each bin's mean luminance (0.212671 r + 0.715160 g +
    0.072169 b of the color clamped to [0, FLT_MAX], NaN as 0); 0.18 over
    2^(mean log2 of the bins above 1e-8), 1 if there is none."""
    c = np.nan_to_num(np.asarray(color, np.float64)[..., :3], nan=0.0, posinf=FLT_MAX)
    c = np.clip(c, 0.0, FLT_MAX)
    lum = 0.212671 * c[..., 0] + 0.715160 * c[..., 1] + 0.072169 * c[..., 2]
    h, w = lum.shape
    nbh, nbw = (h + 15) // 16, (w + 15) // 16
    logs = []
    for bh in range(nbh):
        y0, y1 = bh * h // nbh, (bh + 1) * h // nbh
        for bw in range(nbw):
            x0, x1 = bw * w // nbw, (bw + 1) * w // nbw
            L = lum[y0:y1, x0:x1].mean()
            if L > 1e-8:
                logs.append(np.log2(L))
    return 0.18 / 2.0 ** np.mean(logs) if logs else 1.0


def _pad16(t):
    h, w = t.shape[:2]
    out = np.zeros(((h + 15) // 16 * 16, (w + 15) // 16 * 16, t.shape[2]), np.float64)
    out[:h, :w] = t
    return out


def _net_image(weights, x, quantize):
    """Network output (h, w, 3) of an (h, w, C) input tensor, zero-padded to
    a multiple of 16."""
    h, w = x.shape[:2]
    q = quantize or (lambda a: a)
    out, _ = forward(weights, q(_pad16(x)), quantize)
    return out[:h, :w]


def _tiled(weights, x, tile, overlap, quantize):
    h, w = x.shape[:2]
    if not tile:
        return _net_image(weights, x, quantize)
    out = np.zeros((h, w, 3))
    for ty in range(0, h, tile):
        for tx in range(0, w, tile):
            ty1, tx1 = min(ty + tile, h), min(tx + tile, w)
            y0, y1 = max(0, ty - overlap), min(h, ty1 + overlap)
            x0, x1 = max(0, tx - overlap), min(w, tx1 + overlap)
            # The region is padded to a multiple of 16 with the image's own
            # pixels where they exist, zeros past the image.
            ry1, rx1 = min(h, y0 + (y1 - y0 + 15) // 16 * 16), min(w, x0 + (x1 - x0 + 15) // 16 * 16)
            r = _net_image(weights, x[y0:ry1, x0:rx1], quantize)
            out[ty:ty1, tx:tx1] = r[ty - y0:ty1 - y0, tx - x0:tx1 - x0]
    return out


def run(weights, color, albedo, normal, input_scale=None, aux=None, tile=0, overlap=0, quantize=None):
    """mt_denoiser_run on (h, w, 4) images: denoised (h, w, 4), alpha from
    color. `aux` = (albedo weights, normal weights) prefilters the guides."""
    color, albedo, normal = (np.asarray(a, np.float64) for a in (color, albedo, normal))
    if aux is not None:
        fa = _tiled(aux[0], srgb_encode(np.nan_to_num(albedo[..., :3])), tile, overlap, quantize)
        fn = _tiled(aux[1], np.clip(np.nan_to_num(normal[..., :3]) * 0.5 + 0.5, 0.0, 1.0), tile, overlap, quantize)
        albedo = np.concatenate([srgb_decode(fa), albedo[..., 3:]], axis=2)
        normal = np.concatenate([np.clip(fn, 0.0, 1.0) * 2.0 - 1.0, normal[..., 3:]], axis=2)
    scale = autoexposure(color) if input_scale is None else input_scale
    c = np.clip(np.nan_to_num(color[..., :3] * scale, nan=0.0, posinf=FLT_MAX), 0.0, FLT_MAX)
    x = np.concatenate([pu_forward(c) * PU_NORM, np.clip(np.nan_to_num(albedo[..., :3]), 0.0, 1.0),
                        np.clip(np.nan_to_num(normal[..., :3]), -1.0, 1.0) * 0.5 + 0.5], axis=2)
    y = _tiled(weights, x, tile, overlap, quantize)
    out = pu_inverse(np.maximum(y, 0.0) / PU_NORM) / scale
    return np.concatenate([out, color[..., 3:4]], axis=2)


# ------------------------------------------------------ synthetic networks
# Output channels per layer of the shipped networks, in plan order (read
# from the tensor archives' tables; the tests compare them with the files).
WIDTHS = {
    "standard": [32, 32, 48, 64, 80, 96, 96, 112, 112, 96, 96, 64, 64, 64, 32, 3],
    "small": [32, 32, 32, 32, 32, 32, 32, 64, 64, 64, 64, 64, 32, 32, 32, 3],
    "large": [64, 64, 96, 96, 128, 128, 192, 192, 256, 256, 192, 192, 128, 128, 96, 96, 64, 64, 3],
}


def network_shapes(widths, in_channels, large):
    """{layer: (cout, cin, 3, 3)} of a U-Net with the given output widths:
    each layer takes the previous tensor's channels, a decoder level's first
    layer also the skip (the pooled encoder output one level up, or the
    network input at level 0)."""
    steps = UNET_LARGE if large else UNET
    shapes, pools, prev = {}, {}, in_channels
    for (name, level, pool), cout in zip(steps, widths):
        cin = prev
        if name.startswith("dec_") and name.endswith("a"):
            cin += in_channels if level == 0 else pools[level - 1]
        shapes[name] = (cout, cin, 3, 3)
        if pool is not None:
            pools[pool] = cout
        prev = cout
    return shapes


def synthetic_weights(shapes, seed=0):
    """He-scaled random fp16 weights and small biases for `shapes`, as
    {name: (array, layout)} like read_tza."""
    rng = np.random.default_rng(seed)
    out = {}
    for name, (co, ci, kh, kw) in shapes.items():
        w = rng.normal(0.0, np.sqrt(2.0 / (ci * kh * kw)), (co, ci, kh, kw)).astype(np.float16)
        out[name + ".weight"] = (w, "oihw")
        out[name + ".bias"] = (rng.normal(0.0, 0.02, co).astype(np.float16), "x")
    return out


def write_tza(tensors):
    """Bytes of a version-2 tensor archive holding `tensors` ({name: (array,
    layout)}, fp16 or fp32 arrays): header, data (64-byte aligned), table."""
    blob = bytearray(struct.pack("<HBBQ", 0x41D7, 2, 0, 0))
    entries = []
    for name, (arr, layout) in tensors.items():
        arr = np.ascontiguousarray(arr)
        code = {"float16": b"h", "float32": b"f"}[arr.dtype.name]
        while len(blob) % 64:
            blob.append(0)
        entries.append((name, arr.shape, layout, code, len(blob)))
        blob += arr.astype(arr.dtype.newbyteorder("<")).tobytes()
    table = len(blob)
    blob += struct.pack("<I", len(entries))
    for name, shape, layout, code, off in entries:
        nb = name.encode()
        blob += struct.pack("<H", len(nb)) + nb + struct.pack("<B", len(shape))
        blob += struct.pack("<" + "I" * len(shape), *shape) + layout.encode() + code + struct.pack("<Q", off)
    struct.pack_into("<Q", blob, 4, table)
    return bytes(blob)
