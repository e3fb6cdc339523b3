"""Builds libmtx.so (HIP for gfx950) in-tree from mutracer_amd/csrc.

Every translation unit is compiled with:
  -ffp-contract=off            a*b+c stays unfused (bit-reproducible f32/f64)
  -fno-fast-math               IEEE division/sqrt, NaN/inf semantics kept
  -target-feature -packed-fp32-ops   no v_pk_*_f32 forms
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmtx.so")
ARCH = os.environ.get("MTX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-command-line-argument", "-Wno-pass-failed",
    f"-I{os.path.join(HERE, '..', 'include')}",
]


LUT = os.path.join(HERE, "data", "openpbr_energy_lut.f32")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _lut_source():
    """Writes the OpenPBR energy tables (2177 rows x 32 f32) as a C++ array."""
    import numpy as np
    bits = np.fromfile(LUT, dtype="<u4")
    if bits.size != 32 * 2177:
        raise RuntimeError(f"{LUT}: expected {32 * 2177} floats, found {bits.size}")
    path = os.path.join(BUILD, "mtx_energy_lut.cpp")
    rows = ",\n".join(",".join(f"0x{int(b):08x}u" for b in bits[i:i + 16]) for i in range(0, bits.size, 16))
    text = ("#include <stdint.h>\nnamespace mtx {\nextern const uint32_t kEnergyLutBits[69664];\n"
            "alignas(16) const uint32_t kEnergyLutBits[69664] = {\n" + rows + "};\n}\n")
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources() + [_lut_source()]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
