"""Decodes the OpenPBR energy-compensation tables into openpbr_energy_lut.f32.

The tables are Adobe's openpbr-bsdf data (Apache-2.0) as re-quantized by the
reference (src/renderer/openPbrEnergyLut.ts). This script reads the encoded
strings from a reference checkout and decodes them with the row-delta format
documented in src/renderer/decodeOpenPbrEnergyLut.ts:1-18: directional rows
hold little-endian Uint16 min/max followed by 32 Uint8 samples, each stored as
a difference from the previous row (mod 65536 / mod 256), value =
(min + (max - min) * sample / 255) / 65535; metal and dielectric averages are
exact Uint16 UNORM; each metal average row is then replaced by the exact
integral of 2*mu times its piecewise-linear directional row.

Output: 2177 rows x 32 float32, little-endian (rows as in the reference's
OPENPBR_*_ROW constants). Run once where the reference is checked out:
    python mutracer_amd/data/make_energy_lut.py /root/reference
"""
import base64
import os
import re
import sys

import numpy as np

WIDTH, HEIGHT = 32, 2177
METAL_ROW, METAL_AVG_ROW = 1056, 1088
DIEL_ROW, DIEL_AVG_ROW, DIEL_RATIO_ROW = 1089, 2113, 2145


def _constant(text, name):
    start = text.index(f"const {name} =")
    end = text.index(";", start)
    return "".join(re.findall(r'"([A-Za-z0-9+/=]*)"', text[start:end]))


def decode(text):
    enc = b"".join(base64.b64decode(_constant(text, n)) for n in (
        "OPENPBR_ENERGY_LUT_ROW_DELTA_BASE64", "OPENPBR_METAL_ENERGY_LUT_ROW_DELTA_BASE64",
        "OPENPBR_DIELECTRIC_ENERGY_LUT_ROW_DELTA_BASE64"))
    avg = b"".join(base64.b64decode(_constant(text, n)) for n in (
        "OPENPBR_METAL_AVERAGE_U16_BASE64", "OPENPBR_DIELECTRIC_AVERAGE_U16_BASE64"))
    data = np.zeros(WIDTH * HEIGHT, np.float32)
    lo = hi = 0
    samples = np.zeros(WIDTH, np.int64)
    row_in = 0
    for row in range(HEIGHT):
        if row == METAL_AVG_ROW or DIEL_AVG_ROW <= row < DIEL_RATIO_ROW:
            continue
        o = row_in * (WIDTH + 4)
        row_in += 1
        lo = (lo + (enc[o] | enc[o + 1] << 8)) & 0xFFFF
        hi = (hi + (enc[o + 2] | enc[o + 3] << 8)) & 0xFFFF
        samples = (samples + np.frombuffer(enc[o + 4:o + 4 + WIDTH], np.uint8)) & 0xFF
        for c in range(WIDTH):  # f64 arithmetic, one rounding to f32
            data[row * WIDTH + c] = (lo + (hi - lo) * int(samples[c]) / 255) / 65535
    for i in range(32 + 32 * 32):
        v = (avg[2 * i] | avg[2 * i + 1] << 8) / 65535
        off = METAL_AVG_ROW * 32 + i if i < 32 else DIEL_AVG_ROW * 32 + i - 32
        data[off] = v
    h = 1 / (WIDTH - 1)
    for r in range(32):
        off = (METAL_ROW + r) * WIDTH
        acc = 0.0
        for i in range(WIDTH - 1):
            a, b = float(data[off + i]), float(data[off + i + 1])
            acc += h * (i * h * (a + b) + h * (a + 2 * b) / 3)
        data[METAL_AVG_ROW * WIDTH + r] = acc
    return data


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, "src/renderer/openPbrEnergyLut.ts")).read()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "openpbr_energy_lut.f32")
    decode(text).astype("<f4").tofile(out)
    print("wrote", out)
