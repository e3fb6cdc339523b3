"""Denoiser host: the OIDN RT U-Net on MFMA (libmtx).

`Denoiser(precision)` loads the TZA weights (the reference's
oidn/weights/rt_hdr_alb_nrm.tza, shipped here as mutracer_amd/data) and
mirrors mt_denoiser_run (native/include/mutracer-denoise.h:41-53) with
`run(color, normal, albedo)`; `forward(x)` runs the network alone and can
return every convolution's output for precision studies.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "rt_hdr_alb_nrm.tza")
PRECISIONS = {"fp16": 1, "bf16": 2}
LAYERS = [  # (name, level, output channels)
    ("enc_conv0", 0, 32), ("enc_conv1", 0, 32), ("enc_conv2", 1, 48), ("enc_conv3", 2, 64), ("enc_conv4", 3, 80),
    ("enc_conv5a", 4, 96), ("enc_conv5b", 4, 96), ("dec_conv4a", 3, 112), ("dec_conv4b", 3, 112),
    ("dec_conv3a", 2, 96), ("dec_conv3b", 2, 96), ("dec_conv2a", 1, 64), ("dec_conv2b", 1, 64),
    ("dec_conv1a", 0, 64), ("dec_conv1b", 0, 32), ("dec_conv0", 0, 3)]
_P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


class Denoiser:
    def __init__(self, precision="fp16", weights=WEIGHTS):
        data = open(weights, "rb").read()
        buf = (C.c_ubyte * len(data)).from_buffer_copy(data)
        self.handle = C.c_uint64(0)
        err = abi.errbuf()
        abi.check(abi.lib().mtx_denoiser_create_from_weights(C.cast(buf, C.c_void_p), len(data), PRECISIONS[precision],
                                                C.byref(self.handle), err, 512), err)
        self.precision = precision

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
        """x: (H, W, 9) f32, H and W multiples of 16 -> (H, W, 3) [, per-layer outputs]."""
        x = np.ascontiguousarray(x, np.float32)
        h, w = x.shape[:2]
        out = np.empty((h, w, 3), np.float32)
        bufs = None
        ptrs = None
        if layers:
            bufs = [np.empty((h >> lv, w >> lv, c), np.float32) for _, lv, c in LAYERS]
            ptrs = (C.c_void_p * 16)(*[b.ctypes.data for b in bufs])
        err = abi.errbuf()
        abi.check(abi.lib().mtx_denoiser_forward(self.handle.value, w, h, _P(x), _P(out),
                                                 None if ptrs is None else C.cast(ptrs, C.c_void_p), err, 512), err)
        if layers:
            return out, {name: b for (name, _, _), b in zip(LAYERS, bufs)}
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
    """Times network passes at a given size (HIP events), after 100 untimed
    passes: python3 -m mutracer_amd.denoise [--width W] [--height H]
    [--precision fp16|bf16] [--iterations N]."""
    import argparse
    ap = argparse.ArgumentParser(prog="python3 -m mutracer_amd.denoise")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1088)
    ap.add_argument("--precision", default="fp16", choices=("fp16", "bf16"))
    ap.add_argument("--iterations", type=int, default=20)
    args = ap.parse_args(argv)
    d = Denoiser(args.precision)
    d.benchmark(args.width, args.height, 100)
    ms = d.benchmark(args.width, args.height, args.iterations)
    print(f"denoiser {args.precision} {args.width}x{args.height}: {ms:.3f} ms per pass", flush=True)


if __name__ == "__main__":
    _main()
