"""Denoiser host: the OIDN RT U-Nets on MFMA (libmtx).

`Denoiser(precision, weights)` loads a TZA network (path or bytes; default
the reference's oidn/weights/rt_hdr_alb_nrm.tza, shipped here as
mutracer_amd/data together with rt_hdr_alb_nrm_small.tza) and mirrors
mt_denoiser_run (native/include/mutracer-denoise.h:41-53) with
`run(color, normal, albedo)`; `forward(x)` runs the network alone and can
return every convolution's output for precision studies. `Denoiser.load`
mirrors the reference's model loader (src/oidn.ts:22-29,
src/renderer/oidnNetwork.ts:94-141, 181-215): the 'standard' and 'small'
models from the bundled files, 'cleanAuxLarge' with its albedo and normal
prefilter networks from a weights directory (the oidn-weights 2.3.0-1
files, which the reference fetches rather than bundles).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import struct

import numpy as np

from . import abi

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS = os.path.join(WEIGHTS_DIR, "rt_hdr_alb_nrm.tza")
# src/oidn.ts:22-29
MODEL_FILENAMES = {"small": "rt_hdr_alb_nrm_small.tza", "standard": "rt_hdr_alb_nrm.tza",
                   "cleanAuxLarge": "rt_hdr_calb_cnrm_large.tza"}
AUXILIARY_MODEL_FILENAMES = {"albedo": "rt_alb_large.tza", "normal": "rt_nrm_large.tza"}
BUNDLED_MODELS = ("small", "standard")
PRECISIONS = {"fp16": 1, "bf16": 2}
# Layer names and resolution levels of OIDN's two RT network shapes, in
# network order: "UNet" (standard, small) and "UNetLarge" (large models).
UNET = [("enc_conv0", 0), ("enc_conv1", 0), ("enc_conv2", 1), ("enc_conv3", 2), ("enc_conv4", 3),
        ("enc_conv5a", 4), ("enc_conv5b", 4), ("dec_conv4a", 3), ("dec_conv4b", 3), ("dec_conv3a", 2),
        ("dec_conv3b", 2), ("dec_conv2a", 1), ("dec_conv2b", 1), ("dec_conv1a", 0), ("dec_conv1b", 0),
        ("dec_conv0", 0)]
UNET_LARGE = [("enc_conv1a", 0), ("enc_conv1b", 0), ("enc_conv2a", 1), ("enc_conv2b", 1), ("enc_conv3a", 2),
              ("enc_conv3b", 2), ("enc_conv4a", 3), ("enc_conv4b", 3), ("enc_conv5a", 4), ("enc_conv5b", 4),
              ("dec_conv4a", 3), ("dec_conv4b", 3), ("dec_conv3a", 2), ("dec_conv3b", 2), ("dec_conv2a", 1),
              ("dec_conv2b", 1), ("dec_conv1a", 0), ("dec_conv1b", 0), ("dec_conv1c", 0)]
LAYERS = [  # (name, level, output channels) of the standard network
    ("enc_conv0", 0, 32), ("enc_conv1", 0, 32), ("enc_conv2", 1, 48), ("enc_conv3", 2, 64), ("enc_conv4", 3, 80),
    ("enc_conv5a", 4, 96), ("enc_conv5b", 4, 96), ("dec_conv4a", 3, 112), ("dec_conv4b", 3, 112),
    ("dec_conv3a", 2, 96), ("dec_conv3b", 2, 96), ("dec_conv2a", 1, 64), ("dec_conv2b", 1, 64),
    ("dec_conv1a", 0, 64), ("dec_conv1b", 0, 32), ("dec_conv0", 0, 3)]
_P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


def tza_shapes(data):
    """{tensor name: shape} from the table of a version-2 tensor archive
    (u16 magic 0x41D7, u8 version 2, u8 minor, u64 table offset; table: u32
    count, then name length u16, name, rank u8, rank u32 dims, rank layout
    bytes, dtype byte, u64 data offset)."""
    magic, major = struct.unpack_from("<HB", data, 0)
    if magic != 0x41D7 or major != 2:
        raise ValueError("not a version-2 tensor archive")
    (table,) = struct.unpack_from("<Q", data, 4)
    (n,) = struct.unpack_from("<I", data, table)
    p, out = table + 4, {}
    for _ in range(n):
        (nl,) = struct.unpack_from("<H", data, p)
        name = bytes(data[p + 2:p + 2 + nl]).decode()
        p += 2 + nl
        rank = data[p]
        out[name] = struct.unpack_from("<" + "I" * rank, data, p + 1)
        p += 1 + 5 * rank + 1 + 8
    return out


class Denoiser:
    def __init__(self, precision="fp16", weights=WEIGHTS):
        data = bytes(weights) if isinstance(weights, (bytes, bytearray, memoryview)) else open(weights, "rb").read()
        buf = (C.c_ubyte * len(data)).from_buffer_copy(data)
        self.handle = C.c_uint64(0)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_denoiser_create_from_weights(C.cast(buf, C.c_void_p), len(data), PRECISIONS[precision],
                                                C.byref(self.handle), err, 512), err)
        self.precision = precision
        info = self.info()
        self.large = bool(info[2])
        self.input_channels = info[1]
        shapes = tza_shapes(data)
        plan = UNET_LARGE if self.large else UNET
        assert len(plan) == info[0]
        self.layers = [(name, lv, shapes[name + ".weight"][0]) for name, lv in plan]
        self._aux = None

    @classmethod
    def load(cls, model="standard", weights_dir=None, auxiliary=None, precision="fp16", max_tile_size=0,
             tile_overlap=0):
        """A packaged model by name (src/oidn.ts:22-29): 'standard' and
        'small' from the bundled weights unless weights_dir is given;
        'cleanAuxLarge' needs weights_dir and loads its albedo and normal
        prefilters from there too, unless `auxiliary` = {'albedo': path,
        'normal': path} overrides them (oidnNetwork.ts:181-215). A missing
        file raises; nothing falls back to another network."""
        if model not in MODEL_FILENAMES:
            raise ValueError(f"unknown model {model!r}; expected one of {sorted(MODEL_FILENAMES)}")
        if weights_dir is None and model not in BUNDLED_MODELS:
            raise FileNotFoundError(f"the {model} weights are not bundled: pass weights_dir, a directory holding "
                                    f"the oidn-weights 2.3.0-1 files {MODEL_FILENAMES[model]} and "
                                    f"{', '.join(AUXILIARY_MODEL_FILENAMES.values())}")
        base = WEIGHTS_DIR if weights_dir is None else weights_dir
        d = cls(precision, os.path.join(base, MODEL_FILENAMES[model]))
        if auxiliary is None and model == "cleanAuxLarge":
            auxiliary = {k: os.path.join(base, f) for k, f in AUXILIARY_MODEL_FILENAMES.items()}
        if auxiliary is not None:
            d.set_auxiliary(cls(precision, auxiliary["albedo"]), cls(precision, auxiliary["normal"]))
        d.set_tiles(max_tile_size, tile_overlap)
        return d

    def info(self):
        """(convolution layers, input channels, 1 if UNetLarge, precision code)."""
        out = (C.c_uint32 * 4)()
        abi.check(abi.lib().mtx_denoiser_info(self.handle.value, out))
        return tuple(out)

    def set_auxiliary(self, albedo=None, normal=None):
        """Prefilter networks of the albedo and normal guides (3-channel
        denoisers of this precision), or None, None to remove them."""
        err = abi.errbuf()
        abi.check(abi.lib().mtx_denoiser_set_auxiliary(self.handle.value, albedo.handle.value if albedo else 0,
                                                       normal.handle.value if normal else 0, err, 512), err)
        self._aux = (albedo, normal) if albedo else None

    @property
    def auxiliary(self):
        return self._aux

    def set_tiles(self, max_tile_size=0, overlap=0):
        """Image tiles of run(): max_tile_size pixels per tile (0 = whole
        image), each run with `overlap` pixels of context per side; both
        multiples of 16."""
        abi.check(abi.lib().mtx_denoiser_set_tiles(self.handle.value, max_tile_size, overlap))

    def set_input_scale(self, scale=None):
        """Fixed HDR input scale, or None for per-image autoexposure."""
        abi.check(abi.lib().mtx_denoiser_set_input_scale(self.handle.value, math.nan if scale is None else scale))

    def run(self, color, normal, albedo):
        """RGBA f32 (H, W, 4) images -> denoised RGBA f32 (alpha of color kept)."""
        color, normal, albedo = (np.ascontiguousarray(a, np.float32) for a in (color, normal, albedo))
        h, w = color.shape[:2]
        out = np.empty_like(color)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_denoiser_run(self.handle.value, w, h, _P(color), _P(normal), _P(albedo), _P(out),
                                             out.size, err, 512), err)
        return out

    def forward(self, x, layers=False):
        """x: (H, W, C) f32 with C the network's input channels (9 or 3), H
        and W multiples of 16 -> (H, W, 3) [, per-layer outputs]."""
        x = np.ascontiguousarray(x, np.float32)
        h, w = x.shape[:2]
        if x.ndim != 3 or x.shape[2] != self.input_channels:
            raise ValueError(f"input must be (H, W, {self.input_channels})")
        out = np.empty((h, w, 3), np.float32)
        bufs = None
        ptrs = None
        if layers:
            bufs = [np.empty((h >> lv, w >> lv, c), np.float32) for _, lv, c in self.layers]
            ptrs = (C.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
        err = abi.errbuf()
        abi.check(abi.lib().mtx_denoiser_forward(self.handle.value, w, h, _P(x), _P(out),
                                                 None if ptrs is None else C.cast(ptrs, C.c_void_p), err, 512), err)
        if layers:
            return out, {name: b for (name, _, _), b in zip(self.layers, bufs)}
        return out

    STAGED = 16  # MTX_DENOISER_STAGED

    def set_tiling(self, layer, pixel_tiles=0, channel_tiles=0):
        """Launch shape of one layer's convolution (0, 0 = automatic);
        pixel_tiles | STAGED selects the LDS-staged kernel."""
        abi.check(abi.lib().mtx_denoiser_set_tiling(self.handle.value, layer, pixel_tiles, channel_tiles))

    def tiling(self, layer, width, height):
        """(pixel_tiles, channel_tiles) the layer uses at width x height."""
        mt, nt = C.c_uint32(0), C.c_uint32(0)
        abi.check(abi.lib().mtx_denoiser_get_tiling(self.handle.value, layer, width, height, C.byref(mt),
                                                    C.byref(nt)))
        return mt.value, nt.value

    def benchmark(self, width, height, iterations=10):
        ms = C.c_double(0)
        abi.check(abi.lib().mtx_denoiser_benchmark(self.handle.value, width, height, iterations, C.byref(ms)))
        return ms.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and abi._lib is not None:
            abi.lib().mtx_denoiser_destroy(h.value, None, 0)


def _main(argv=None):
    """Times network passes at a given size (HIP events), after untimed
    passes: python3 -m mutracer_amd.denoise [--width W] [--height H]
    [--precision fp16|bf16] [--iterations N] [--warmup N] [--model M]."""
    import argparse
    ap = argparse.ArgumentParser(prog="python3 -m mutracer_amd.denoise")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1088)
    ap.add_argument("--precision", default="fp16", choices=("fp16", "bf16"))
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=100, help="untimed passes first")
    ap.add_argument("--model", default="standard", choices=sorted(MODEL_FILENAMES))
    ap.add_argument("--weights-dir", default=None)
    args = ap.parse_args(argv)
    d = Denoiser.load(args.model, args.weights_dir, precision=args.precision)
    if args.warmup:
        d.benchmark(args.width, args.height, args.warmup)
    ms = d.benchmark(args.width, args.height, args.iterations)
    print(f"denoiser {args.model} {args.precision} {args.width}x{args.height}: {ms:.3f} ms per pass", flush=True)


if __name__ == "__main__":
    _main()
