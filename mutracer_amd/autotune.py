"""Autotuner for the wavefront launch shape.

The reference's only tuner is the tile-size predictor of TileRender.ts
(:302-490), which fits tile area to a time budget from GPU timestamps. Here
the knobs are the wavefront kernels' launch shape (mtx_wavefront_tuning):
traversal form, LDS stack depth, grid oversubscription, the shade kernel's
occupancy target, the queue length below which the tail kernel finishes a
generation, and the number of paths in flight. The tuner measures each
candidate with HIP events (RenderTarget.render_timed) on a short probe
render of the real target and keeps the fastest, searching one dimension at
a time (coordinate descent, two passes). Every candidate renders the same
samples, so tuning never changes an image; probes restart the target's
accumulation, which the caller resets anyway before a real render.

Results are cached per (device, scene size, resolution, bounces, candidates)
in a JSON file so a server or a repeated render tunes once.
"""
from __future__ import annotations

import json
import os

SPACE = {
    "trace_kernel": [0, 1],
    "lds_stack": [16, 24, 32],
    "trace_oversubscription": [1, 2, 4],
    "shade_waves_per_simd": [1, 2, 4],
    "shade_oversubscription": [1, 2],
    "max_paths_in_flight": [1 << 23, 1 << 24, 1 << 25],
    "tail_paths": [0, 16384, 65536, 262144],
    "queue_sort": [0, 1],
    "lds_stack_closest": [0, 16, 24, 32],  # 0: the any-hit depth (lds_stack)
}
ORDER = ["max_paths_in_flight", "tail_paths", "queue_sort", "trace_kernel", "trace_oversubscription",
         "shade_waves_per_simd", "shade_oversubscription", "lds_stack", "lds_stack_closest"]


def _key(target):
    from . import abi
    import ctypes as C
    name = C.create_string_buffer(256)
    cus = C.c_int32(0)
    abi.lib().mtx_device_info(name, 256, C.byref(cus))
    s = target.settings
    tris = target.scene.info()["triangles"]
    bucket = 1 << max(0, int(tris).bit_length() - 1)
    return f"{name.value.decode()}|{cus.value}|tris~{bucket}|{s.width}x{s.height}|b{s.bounces}|c{s.direct_light_candidates}"


def probe(target, tuning, samples):
    """GPU milliseconds of one probe render with `tuning` (after one warm-up)."""
    target.set_tuning(**tuning)
    target.restart_samples()
    target.render_timed(1)
    target.restart_samples()
    t = target.render_timed(samples)
    return t["generate_accumulate"] + t["trace"] + t["shade"] + t["shadow"] + t["tail"]


def tune(target, samples=2, cache_path=None, log=None):
    """Picks the fastest launch shape for `target`; applies and returns it."""
    cache_path = cache_path or os.path.join(os.path.expanduser("~"), ".cache", "mutracer_amd", "tuning.json")
    key = _key(target)
    cache = {}
    if os.path.exists(cache_path):
        try:
            cache = json.load(open(cache_path))
        except (OSError, ValueError):
            cache = {}
    if key in cache:
        target.set_tuning(**cache[key]["tuning"])
        return cache[key]["tuning"]
    best = target.tuning()
    best_ms = probe(target, best, samples)
    trials = [(dict(best), best_ms)]
    for _ in range(2):
        improved = False
        for dim in ORDER:
            for v in SPACE[dim]:
                if v == best[dim]:
                    continue
                cand = dict(best, **{dim: v})
                ms = probe(target, cand, samples)
                trials.append((cand, ms))
                if log:
                    log(f"tune {dim}={v}: {ms:.2f} ms (best {best_ms:.2f})")
                if ms < best_ms * 0.98:
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    target.set_tuning(**best)
    target.restart_samples()
    cache[key] = dict(tuning=best, ms=best_ms, probe_samples=samples, trials=len(trials))
    try:
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        json.dump(cache, open(cache_path, "w"), indent=1)
    except OSError:
        pass
    return best


# ---------------------------------------------------------------- denoiser
DENOISER_STAGED = 16  # MTX_DENOISER_STAGED: the LDS-staged convolution kernel
DENOISER_SHAPES = [(flag | mt, nt) for flag in (0, DENOISER_STAGED) for mt in (1, 2, 4, 8) for nt in (1, 2, 3, 4, 7)]


def tune_denoiser(denoiser, width, height, iterations=3, log=None):
    """Picks each U-Net layer's convolution launch shape (pixel tiles x
    output-channel tiles per wave, direct or LDS-staged kernel) by
    measuring whole forward passes with one layer changed at a time (HIP
    events, mtx_denoiser_benchmark), and applies the fastest. Returns {layer: (pixel_tiles, channel_tiles)} and the
    pass time before and after."""
    before = min(denoiser.benchmark(width, height, iterations) for _ in range(2))
    best = {}
    for layer in range(len(denoiser.layers)):
        cur = denoiser.tiling(layer, width, height)
        trials = []
        for mt, nt in DENOISER_SHAPES:
            denoiser.set_tiling(layer, mt, nt)
            trials.append((min(denoiser.benchmark(width, height, iterations) for _ in range(2)), mt, nt))
        trials.sort()
        ms, mt, nt = trials[0]
        best[layer] = (mt, nt)
        denoiser.set_tiling(layer, mt, nt)
        if log:
            log(f"denoiser layer {layer}: {cur} -> {(mt, nt)} ({ms:.3f} ms per pass)")
    after = min(denoiser.benchmark(width, height, iterations) for _ in range(2))
    return best, before, after

