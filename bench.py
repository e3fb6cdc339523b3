"""Throughput of the path-tracing hot path on one MI355X.

Workload (one step): one full image of the 1M-triangle instanced scene — 64
fBm-displaced geodesic icospheres (15,680 triangles each, one BLAS each) plus
a 2-triangle ground quad, F82 metal / dielectric / coated-diffuse OpenPBR
materials, 1024x512 sky+sun environment — at 1920x1080, 16 samples per pixel,
depth 12, 4 NEE candidates, Owen-scrambled Sobol seed 0. The scene (BVHs,
materials, environment, FP64 accumulators) is resident in HBM before timing;
each step clears the accumulators and renders the 16 spp in one megakernel
launch.

Prints one JSON line: Msamples/s (value), Mrays/s, the megakernel's roofline
(algorithmic bytes per launch from the instrumented launch's work totals over
the launch duration from HIP events on the launch stream) and the CPU
baseline: the oracle (C++ restatement of the same algorithm, OpenMP) timed on
this host on 32x32-pixel tiles of the same image, whose results are also
compared with the GPU's accumulators for those pixels.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at 1080p 1M-tri (plain megakernel)"
HBM_PEAK_GBS = 8000.0
WIDTH, HEIGHT, SPP, DEPTH, CANDIDATES = 1920, 1080, 16, 12, 4

# Algorithmic bytes per unit of megakernel work (rays stay in registers):
BOX_B = 32          # one child AABB slab test (2 x vec3 f32 + padding)
TRI_B = 48          # one triangle test (3 vertices, 12 B each, + 12 B index)
INST_B = 80         # one instance entry (w2l 4x4 f32 + 16 B record)
HIT_B = 128 + 48    # per closest hit: packed material (32 f32) + 3 normals and the index triple
SAMPLE_B = 13 * 16 + 8  # per sample: read+write of 13 FP64 accumulator planes + the u32 count


KERNEL_NAMES = {"wavefront": "wavefront launch sequence (generate, trace, shade, shadow, accumulate)",
                "megakernel": "k_render (megakernel)"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profile():
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", "megakernel_c2_1080p_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(sc, gpu_sums, seconds):
    """Oracle on 32x32 tiles spread over the frame until `seconds` elapse."""
    import oracle
    r = oracle.OracleRenderer(sc, WIDTH, HEIGHT, bounces=DEPTH, candidates=CANDIDATES, pattern=0, seed=0,
                              max_samples=SPP)
    tiles_x, tiles_y = WIDTH // 32, HEIGHT // 32
    n_tiles = tiles_x * tiles_y
    done, t0, errs = 0, time.perf_counter(), []
    k = 0
    while time.perf_counter() - t0 < seconds and k < n_tiles:
        tile = (k * 211) % n_tiles  # 211 is prime to n_tiles: a spread-out visiting order
        x0, y0 = (tile % tiles_x) * 32, (tile // tiles_x) * 32
        out = r.render(0, SPP, x0, y0, 32, 32)
        done += 32 * 32 * SPP
        g = gpu_sums[y0:y0 + 32, x0:x0 + 32, 0:3] / SPP
        ref = out[..., 0:3] / SPP
        errs.append(np.mean((g - ref) ** 2 / (ref * ref + 1e-2)))
        k += 1
    dt = time.perf_counter() - t0
    return dict(value=done / dt / 1e6, unit="Msamples/s", cores=oracle.threads(), kind="port",
                sample=f"{k} tiles of 32x32 px x {SPP} spp ({done} samples) of the same 1080p image, "
                       f"{dt:.1f} s"), float(np.mean(errs)) if errs else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel", choices=["wavefront", "megakernel"], default="wavefront")
    ap.add_argument("--max-paths", type=int, default=0, help="wavefront paths in flight (0 = library default)")
    ap.add_argument("--tuning", default="", help="wavefront tuning as key=value,... (see RenderTarget.tuning)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the path shards by "
              f"independent sample ranges; multi-GPU runs are out of scope).", file=sys.stderr)
        sys.exit(2)

    import torch
    from mutracer_amd import render, scene
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)

    t = time.perf_counter()
    sc = scene.instanced_field()
    log(f"scene generated: {sc.triangle_count} triangles in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    rs = render.RenderScene(sc)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t) * 1e3
    info = rs.info()
    log(f"scene committed on GPU (BVH build + upload): {build_ms:.0f} ms, {info}")
    tgt = render.RenderTarget(rs, WIDTH, HEIGHT, bounces=DEPTH, direct_light_candidates=CANDIDATES,
                              pattern="sobol", sequence_seed=0, max_samples=SPP, kernel=args.kernel,
                              max_paths_in_flight=args.max_paths)
    if args.tuning:
        tgt.set_tuning(**{k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))})
    stream = torch.cuda.current_stream()

    tgt.restart_samples()
    work = tgt.render_counted(SPP, stream)
    rays = tgt.stats()["rays"]
    log(f"instrumented launch: {work}, rays {rays}")
    samples = WIDTH * HEIGHT * SPP
    alg_bytes = (BOX_B * work["box_tests"] + TRI_B * work["tri_tests"] + INST_B * work["instance_visits"]
                 + HIT_B * work["hits"] + SAMPLE_B * samples)

    for _ in range(args.warmup):
        tgt.restart_samples()
        tgt.render(SPP, stream)
    torch.cuda.synchronize()
    log("warmup done")

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        tgt.restart_samples()
        starts[k].record(stream)
        tgt.render(SPP, stream)
        ends[k].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, launch {kernel_ms:.2f} ms")
    st = tgt.stats()
    assert st["samples"] == SPP
    rays_per_step = st["rays"]
    value = samples / (ms_per_step / 1e3) / 1e6
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    sums, counts = tgt.read_accumulators()
    assert np.all(counts == SPP)

    cpu, parity = None, None
    if not args.no_cpu_baseline and args.cpu_seconds > 0:
        log("timing the CPU baseline (oracle)")
        cpu, parity = cpu_baseline(sc, sums, args.cpu_seconds)

    traffic = traffic_from_profile()
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1M-tri instanced field: 64 fBm-displaced icospheres x 15,680 tris + ground quad, "
                               "1024x512 sky+sun env, OpenPBR metal/glass/coat mix",
                   "width": WIDTH, "height": HEIGHT, "spp": SPP, "depth": DEPTH,
                   "direct_light_candidates": CANDIDATES, "sampling": "sobol seed 0",
                   "triangles": info["triangles"], "accumulation": "fp64", "kernel": args.kernel,
                   "tuning": tgt.tuning() if args.kernel == "wavefront" else None},
        "mrays_per_s": round(rays_per_step / (ms_per_step / 1e3) / 1e6, 3),
        "rays_per_step": rays_per_step,
        "scene_commit_ms": round(build_ms, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": KERNEL_NAMES[args.kernel], "kernel_ms": round(kernel_ms, 3),
                     "algorithmic_bytes_per_launch": int(alg_bytes), "work_per_launch": work},
        "cpu_baseline": cpu,
        "parity_relmse_cpu_tiles": parity,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
