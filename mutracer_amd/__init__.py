"""mutracer_amd: MI355X-native render hot path for MuTracer.

The product is libmtx.so (HIP for gfx950, C ABI in include/mtx.h). This
package is the Python host over that ABI: resident acceleration structures,
ray queries, the path tracer's Engine/Render surface, serving and the
denoiser. Every entry point calls the HIP library; if it was not built the
import of a submodule that needs it raises (there is no CPU fallback).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
