"""Throughput of the path-tracing hot path on one MI355X.

Workload (one step): one full image of the 1M-triangle instanced scene — 64
fBm-displaced geodesic icospheres (15,680 triangles each, one BLAS each) plus
a 2-triangle ground quad, F82 metal / dielectric / coated-diffuse OpenPBR
materials, 1024x512 sky+sun environment — at 1920x1080, 16 samples per pixel,
depth 12, 4 NEE candidates, Owen-scrambled Sobol seed 0. The scene (BVHs,
materials, environment, FP64 accumulators) is resident in HBM before timing;
each step clears the accumulators and renders the 16 spp with the wavefront
kernels (generate, then per event closest-hit trace, shade, shadow trace, then
accumulate), replayed from the step's captured hipGraph (the product path),
or, with --kernel megakernel, in one megakernel launch.

Prints one JSON line: Msamples/s (value, wall clock over the timed steps),
Mrays/s, the roofline of the dominant kernel (the closest-hit trace; its
algorithmic bytes from the instrumented launch's work totals, its GPU time
from HIP events around each of its launches on the launch stream, in a
separate instrumented pass), the CPU baseline (the oracle, a C++ restatement
of the same algorithm with OpenMP, timed on this host on 32x32-pixel tiles of
the same image, whose results are also compared with the GPU's accumulators
for those pixels), the 8M-triangle build / 16M-ray traversal microbench with
the build's roofline, the U-Net denoiser at 1080p, and the serving stream
(batched vs one job at a time).
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

METRIC = "Msamples/s at 1080p 1M-tri (plain: one job at a time)"
HBM_PEAK_GBS = 8000.0
WIDTH, HEIGHT, SPP, DEPTH, CANDIDATES = 1920, 1080, 16, 12, 4

# Algorithmic bytes per unit of megakernel work (rays stay in registers):
BOX_B = 32          # one child AABB slab test (2 x vec3 f32 + padding)
TRI_B = 48          # one triangle test (3 vertices, 12 B each, + 12 B index)
INST_B = 80         # one instance entry (w2l 4x4 f32 + 16 B record)
HIT_B = 128 + 48    # per closest hit: packed material (32 f32) + 3 normals and the index triple
RAY_B = 32 + 32     # wavefront closest-hit query: ray in (origin, direction), hit record out
SAMPLE_B = 13 * 16 + 8  # per sample: read+write of 13 FP64 accumulator planes + the u32 count


KERNEL_NAMES = {"wavefront": "k_wf_trace (closest-hit traversal of the wavefront sequence)",
                "megakernel": "k_render (megakernel)"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profile_value(name, key):
    """A `key: value` line of a committed rocprofv3 summary under profiles/, if any."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def profile_summary(db_path):
    """Per-kernel totals from a rocprofv3 results database (kernel trace and,
    when collected, counters summed per kernel)."""
    import sqlite3
    con = sqlite3.connect(db_path)
    out = ["kernel | calls | total ms | avg us"]
    for name, calls, total, avg in con.execute(
            "select name, count(*), sum(duration), avg(duration) from kernels group by name "
            "order by sum(duration) desc limit 25"):
        out.append(f"{name.split('(')[0]} | {calls} | {total / 1e6:.3f} | {avg / 1e3:.1f}")
    try:
        rows = con.execute("select k.name, i.name, sum(e.value), count(*) from rocpd_pmc_event e "
                           "join rocpd_info_pmc i on e.pmc_id = i.id "
                           "join rocpd_kernel_dispatch d on d.event_id = e.event_id join kernels k on k.id = d.id "
                           "group by k.name, i.name order by sum(e.value) desc").fetchall()
    except sqlite3.Error:
        rows = []
    if rows:
        out.append("kernel | counter | sum over dispatches | dispatches")
        for kname, cname, total, n in rows:
            out.append(f"{kname.split('(')[0]} | {cname} | {total:.1f} | {n}")
    return "\n".join(out)


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB (reads and writes mib MiB each):
    the reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}), flush=True)


SERVE_RES = [(512, 512), (1024, 768), (1920, 1080), (3840, 2160)]


def serving_benchmark(n_jobs=40, rate=40.0, per_sample=False):
    """The serving stream: 8 resident instanced-field scenes with 10k-4M
    triangles (log-uniform, PCG64 seed 5), each with a 1024x512 sky+sun
    environment; Poisson arrivals at `rate` jobs/s (seed 6) truncated to the
    first `n_jobs` jobs; resolutions uniform over SERVE_RES; 4-256 spp
    log-uniform; depth 8, 4 NEE candidates; admission by a 200 GB HBM budget.
    The same stream runs through the batched server (continuous batching: every
    running job gets a slice of each round's 2^24-path generation) and through
    one job at a time with all of its samples in one generation
    (plain_whole_job, the plain GPU path); `speedup` is served / plain_whole_job.
    With per_sample, also the reference's dispatch model (one job, one sample
    per launch sequence) as `plain_per_sample`."""
    import math
    from mutracer_amd import render, scene, serve
    rng = np.random.Generator(np.random.PCG64(5))
    tris = np.exp(rng.uniform(math.log(1e4), math.log(4e6), 8))
    scenes = []
    for k, t in enumerate(tris):
        f = max(2, int(round(math.sqrt(t / 1280))))
        scenes.append(render.RenderScene(scene.instanced_field(n_instances=64, freq=f, env_size=(1024, 512), seed=k)))
    rng = np.random.Generator(np.random.PCG64(6))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    jobs, samples = [], 0
    for i in range(n_jobs):
        w, h = SERVE_RES[rng.integers(0, len(SERVE_RES))]
        spp = int(round(math.exp(rng.uniform(math.log(4), math.log(256)))))
        jobs.append(dict(render_scene=scenes[rng.integers(0, 8)], width=w, height=h, samples=spp, bounces=8,
                         direct_light_candidates=4))
        samples += w * h * spp
    out = {}
    modes = [("served", dict(max_jobs_per_round=0, max_samples_per_job_round=16)),
             ("plain_whole_job", dict(max_jobs_per_round=1, max_samples_per_job_round=1 << 16))]
    if per_sample:
        modes.append(("plain_per_sample", dict(max_jobs_per_round=1, max_samples_per_job_round=1)))
    for mode, kw in modes:
        server = serve.Server(**kw)
        server.submit(**dict(jobs[0], samples=1))  # warm the kernels and buffers
        while not server.step()[1]:
            pass
        lat, makespan, _ = serve.run_stream(server, jobs, arrivals)
        st = server.stats()
        out[mode] = dict(msamples_per_s=round(samples / makespan / 1e6, 3), makespan_s=round(makespan, 3),
                         p50_ms=round(float(np.percentile(lat, 50)), 2), p99_ms=round(float(np.percentile(lat, 99)), 2),
                         rounds=st["rounds"], graph_replays=st["graph_replays"], graph_captures=st["graph_captures"])
        log(f"serving {mode}: {out[mode]}")
    out["speedup"] = round(out["served"]["msamples_per_s"] / out["plain_whole_job"]["msamples_per_s"], 3)
    if per_sample:
        out["speedup_vs_per_sample"] = round(out["served"]["msamples_per_s"] /
                                             out["plain_per_sample"]["msamples_per_s"], 3)
    out["stream"] = (f"first {n_jobs} jobs of the Poisson {rate}/s stream (seed 6), 8 scenes "
                     f"{int(tris.min())}-{int(tris.max())} tris with 1024x512 env, res {SERVE_RES}, 4-256 spp, "
                     f"depth 8, 4 candidates, {samples} samples")
    return out


def split_work(words):
    """Sum over split (interior) nodes of their primitive counts, from BVH2
    words (node rows [min.xyz, right or -count], [max.xyz, 0]; a leaf's
    primitive rows follow it; the left child follows its parent's two rows):
    every primitive is counted once per interior node on its root path."""
    w = np.asarray(words, np.uint32).reshape(-1, 4)
    code = w[:, 3].view(np.int32)
    frontier, depth, total = np.array([0], np.int64), 0, 0
    while frontier.size:
        c = code[frontier]
        leaf = c < 0
        total += int((-c[leaf]).astype(np.int64).sum()) * depth
        inner = frontier[~leaf]
        frontier = np.concatenate([inner + 2, code[inner].astype(np.int64)])
        depth += 1
    return total


def microbench(verify_build=False, subset=1 << 20):
    """Traversal/build microbench: 8M independent triangles (seed 3, lognormal
    edges), binned SAH with 32 bins built on the GPU; 7,997,584 pinhole
    primaries at 2828^2 plus 8M cosine-hemisphere bounces from their hits.
    Closest-hit IDs are checked bit for bit against the oracle's
    reference-order traversal on a random subset of each ray set (and, with
    verify_build, the BVH2 words against the oracle's own build)."""
    import torch
    import oracle
    from mutracer_amd import accel, workloads as W
    t = time.perf_counter()
    pos, idx = W.random_triangles(8_000_000, seed=3)
    log(f"triangles generated in {time.perf_counter() - t:.1f} s")
    dpos, didx = torch.from_numpy(pos).cuda(), torch.from_numpy(idx).cuda()
    accel.Blas(dpos, didx, 32)  # warm-up build (kernels, allocator)
    torch.cuda.synchronize()
    t = time.perf_counter()
    blas = accel.Blas(dpos, didx, 32)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t) * 1e3
    acc = accel.Accel([blas], [0], [W.identity_transform()], material=[0], double_sided=[1])
    torch.cuda.synchronize()
    words, src = blas.export()
    info = blas.info()
    split = split_work(words)
    # B_build: per split node 40 B per primitive for binning (AABB 24 +
    # centroid 12 + index 4) and 8 B per primitive for the partition, 64 B per
    # node written, 16 B per input triangle.
    nodes = info["node_count"]
    b_build = 48 * split + 64 * nodes + 16 * 8_000_000
    out = {"triangles": 8_000_000, "bins": 32, "build_ms": round(build_ms, 2),
           "build_mtris_per_s": round(8.0 / (build_ms / 1e3), 2),
           "build_roofline": {"bound": "hbm", "algorithmic_bytes": int(b_build), "split_node_primitives": split,
                              "nodes": nodes, "achieved": round(b_build / (build_ms / 1e3) / 1e9, 2),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(b_build / (build_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}}
    log(f"GPU binned-SAH build of 8M triangles: {build_ms:.1f} ms, B_build {b_build / 1e9:.2f} GB")

    def trace(rays, label):
        r = torch.from_numpy(rays).cuda()
        h = torch.zeros((rays.shape[0], 8), dtype=torch.int32, device="cuda")
        acc.trace_closest(r, h, double_sided=True)  # warm
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        acc.trace_closest(r, h, double_sided=True)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        out[label] = {"rays": rays.shape[0], "ms": round(ms, 3), "mrays_per_s": round(rays.shape[0] / ms / 1e3, 1)}
        log(f"{label}: {out[label]}")
        return h.cpu().numpy().view(np.uint32)

    prim = W.pinhole_rays(2828)
    hp = trace(prim, "primaries")
    hit = hp[:, 0].astype(bool)
    tt = hp[hit, 3].view(np.float32)
    p = prim[hit, 0:3] + prim[hit, 4:7] * tt[:, None]
    v = pos[idx.reshape(-1, 3)[hp[hit, 2].astype(np.int64)]]
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n *= np.where((n * prim[hit, 4:7]).sum(1, keepdims=True) > 0, -1, 1)
    bounce = W.hemisphere_bounces(p, n, 8_000_000, seed=0)
    hb = trace(bounce, "bounces")
    total_ms = out["primaries"]["ms"] + out["bounces"]["ms"]
    out["mrays_per_s"] = round((prim.shape[0] + bounce.shape[0]) / total_ms / 1e3, 1)

    # Oracle check on the GPU-built BVH (reference-order traversal).
    tw, iw = acc.export_tlas()
    scene = oracle.TraceScene(tw, words, iw, pos, np.ones(1, np.uint8))
    rng = np.random.default_rng(11)
    mism = 0
    for rays, hits in ((prim, hp), (bounce, hb)):
        sel = rng.choice(rays.shape[0], min(subset, rays.shape[0]), replace=False)
        ref = scene.closest(rays[sel], double_sided=True)
        g = hits[sel]
        hit = ref[:, 0].astype(bool)
        rec = np.where(hit, ref[:, 2], 0)
        same = (g[:, 0] == ref[:, 0]) & (~hit | ((src[rec] == g[:, 2]) & (g[:, 3] == ref[:, 3])))
        mism += int((~same).sum())
    out["oracle_subset_rays"] = 2 * subset
    out["oracle_mismatches"] = mism
    if verify_build:
        t = time.perf_counter()
        ref_words = oracle.build_blas(pos, idx, 32)["words"]
        out["oracle_build_s"] = round(time.perf_counter() - t, 1)
        out["build_words_identical"] = bool(np.array_equal(ref_words, words))
    return out


def denoiser_benchmark(full_precision=False, iterations=20):
    """U-Net at 1920x1080 (padded to 1088) in FP16 and BF16: GPU time per
    pass (HIP events; the median of three runs of `iterations` back-to-back
    passes, after 100 untimed passes that bring the GPU to its steady clocks
    from the idle of the CPU oracle before), TFLOP/s against the dense MFMA
    peak, and the error
    against the FP64 oracle per layer on inputs rendered from the Cornell
    box at 4 spp (a 256x256 centre crop by default, the full frame with
    full_precision)."""
    from oracle import unet
    from mutracer_amd import denoise, render, scene
    sc = scene.cornell_box()
    rs = render.RenderScene(sc)
    t = render.RenderTarget(rs, 1920, 1080, bounces=8, max_samples=4)
    t.render(4)
    color, albedo, normal = (t.read_pixels(k) for k in ("color", "albedo", "normal"))
    norm = 1.0 / _pu_forward(65504.0)
    x = np.zeros((1088, 1920, 9), np.float32)
    scale = unet.autoexposure(color)
    x[:1080, :, 0:3] = _pu_forward(np.maximum(color[..., :3].astype(np.float64) * scale, 0)) * norm
    x[:1080, :, 3:6] = np.clip(albedo[..., :3], 0, 1)
    # readPixels('normal') is display-encoded (n * 0.5 + 0.5, A = 1 where valid),
    # the network's normal input is the same mapping of n (0.5 where n = 0).
    x[:1080, :, 6:9] = np.where(normal[..., 3:4] > 0, np.clip(normal[..., :3], 0, 1), 0.5)
    gflop = unet.flops(1088, 1920) / 1e9
    weights = unet.read_tza(denoise.WEIGHTS)
    crop = x if full_precision else x[416:672, 832:1088]
    ref_out, ref = unet.forward(weights, crop.astype(np.float64))
    out = {"input": "Cornell box 1920x1080 at 4 spp (color through pu, albedo, normal), padded to 1920x1088",
           "gflop_per_pass": round(gflop, 2),
           "precision_region": "full frame" if full_precision else "256x256 centre crop"}
    for prec in ("fp16", "bf16"):
        d = denoise.Denoiser(prec)
        d.benchmark(1920, 1088, 100)
        ms = sorted(d.benchmark(1920, 1088, iterations) for _ in range(3))[1]
        got, layers = d.forward(crop, layers=True)
        errs = {name: round(float(np.abs(layers[name] - ref[name]).max() / max(np.abs(ref[name]).max(), 1e-9)), 6)
                for name, _, _ in denoise.LAYERS}
        diff = np.abs(got - ref_out)
        out[prec] = {"ms_per_pass": round(ms, 3), "tflops": round(gflop / ms, 2),
                     "mfma_frac": round(gflop / ms / 2500.0, 4),
                     "layer_max_rel_error_vs_fp64": errs,
                     "output_max_abs_error": float(diff.max()), "output_mean_abs_error": float(diff.mean()),
                     "output_p99_abs_error": float(np.percentile(diff, 99))}
        log(f"denoiser {prec}: {ms:.2f} ms, {gflop / ms:.1f} TFLOP/s, output max err {diff.max():.2e}")
    # The other network shapes at the same size, FP16: the bundled small
    # model and the large clean-auxiliary model with its two prefilters
    # (weights of the large models' exact shapes, He-scaled random values:
    # they are not bundled, and the time does not depend on the values).
    small = unet.read_tza(os.path.join(denoise.WEIGHTS_DIR, "rt_hdr_alb_nrm_small.tza"))
    large9 = unet.synthetic_weights(unet.network_shapes(unet.WIDTHS["large"], 9, True))
    large3 = unet.synthetic_weights(unet.network_shapes(unet.WIDTHS["large"], 3, True))
    out["models_fp16"] = {}
    for model, w, src in (("small", small, os.path.join(denoise.WEIGHTS_DIR, "rt_hdr_alb_nrm_small.tza")),
                          ("cleanAuxLarge", large9, unet.write_tza(large9)),
                          ("prefilter_large", large3, unet.write_tza(large3))):
        d = denoise.Denoiser("fp16", src)
        d.benchmark(1920, 1088, 50)
        ms = sorted(d.benchmark(1920, 1088, iterations) for _ in range(3))[1]
        g = unet.flops(1088, 1920, w) / 1e9
        out["models_fp16"][model] = {"ms_per_pass": round(ms, 3), "gflop_per_pass": round(g, 2),
                                     "tflops": round(g / ms, 2), "mfma_frac": round(g / ms / 2500.0, 4)}
        log(f"denoiser {model} fp16: {ms:.2f} ms, {g / ms:.1f} TFLOP/s")
    return out


def _pu_forward(y):
    """OIDN's pu transfer (as the library's input stage)."""
    y = np.asarray(y, np.float64)
    A, B, C, D, E, F, G = 1.41283765e+03, 1.64593172e+00, 4.31384981e-01, -2.94139609e-03, 1.92653254e-01, \
        6.26026094e-03, 9.98620152e-01
    Y0, Y1 = 1.57945760e-06, 3.22087631e-02
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(y <= Y0, A * y, np.where(y <= Y1, B * np.power(np.maximum(y, 1e-30), C) + D,
                                                 E * np.log(y + F) + G))


def alpha_benchmark(opaque_sums, steps=3):
    """Cost of mask coverage at the headline workload: every material of the
    instanced field as a mask material whose coverage (opacity 1, cutoff
    0.5) covers everywhere. The image must equal the opaque render bit for
    bit; the difference in time is what the coverage-testing shadow, walker
    and tail kernels and the per-hit coverage resolution cost."""
    import torch
    from mutracer_amd import render, scene
    sc = scene.instanced_field()
    for m in sc.materials:
        m.alpha_mode, m.opacity, m.alpha_cutoff = "mask", 1.0, 0.5
    tgt = render.RenderTarget(render.RenderScene(sc), WIDTH, HEIGHT, bounces=DEPTH,
                              direct_light_candidates=CANDIDATES, pattern="sobol", sequence_seed=0,
                              max_samples=SPP)
    stream = torch.cuda.current_stream()
    tgt.restart_samples()
    tgt.render(SPP, stream)
    torch.cuda.synchronize()
    sums, counts = tgt.read_accumulators()
    t0 = time.perf_counter()
    for _ in range(steps):
        tgt.restart_samples()
        tgt.render(SPP, stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return {"workload": "instanced field, every material mask (covered everywhere)", "ms_per_step": round(ms, 3),
            "msamples_per_s": round(WIDTH * HEIGHT * SPP / (ms / 1e3) / 1e6, 3),
            "equal_to_opaque": bool(np.array_equal(sums, opaque_sums))}


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
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-denoiser", action="store_true")
    ap.add_argument("--denoise-full-precision", action="store_true", help="FP64 oracle on the full 1080p frame")
    ap.add_argument("--serving-jobs", type=int, default=40, help="jobs of the serving stream (its first N arrivals)")
    ap.add_argument("--serving-per-sample", action="store_true",
                    help="also run the stream one sample per launch sequence (the reference's dispatch model)")
    ap.add_argument("--no-microbench", action="store_true")
    ap.add_argument("--microbench-subset", type=int, default=1 << 18, help="rays per set checked against the oracle")
    ap.add_argument("--kernel", choices=["wavefront", "megakernel"], default="wavefront")
    ap.add_argument("--max-paths", type=int, default=0, help="wavefront paths in flight (0 = library default)")
    ap.add_argument("--tuning", default="", help="wavefront tuning as key=value,... (see RenderTarget.tuning)")
    ap.add_argument("--profile-summary", default="", help="print a summary of a rocprofv3 results db and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, help="run one copy of N MiB and exit")
    ap.add_argument("--microbench", action="store_true", help="8M-triangle build + 16M-ray traversal microbench")
    ap.add_argument("--verify-build", action="store_true", help="with --microbench: compare with the oracle build")
    args = ap.parse_args()
    if args.profile_summary:
        print(profile_summary(args.profile_summary))
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.microbench:
        print(json.dumps({"microbench": microbench(args.verify_build, args.microbench_subset)}), flush=True)
        return
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
    if args.kernel == "wavefront":
        # Dominant kernel: the closest-hit trace k_wf_trace (ray in + hit out +
        # node/triangle/instance fetches), counted over its own launches only
        # (the tail kernel's traversal work is the "tail" group).
        wf = work["wavefront"]
        alg_bytes = (RAY_B * wf["closest_rays"] + BOX_B * wf["box_tests"] + TRI_B * wf["tri_tests"]
                     + INST_B * wf["instance_visits"])
    else:
        alg_bytes = (BOX_B * (work["box_tests"] + work["shadow_box_tests"])
                     + TRI_B * (work["tri_tests"] + work["shadow_tri_tests"])
                     + INST_B * (work["instance_visits"] + work["shadow_instance_visits"])
                     + HIT_B * work["hits"] + SAMPLE_B * samples)

    for _ in range(args.warmup):
        tgt.restart_samples()
        tgt.render(SPP, stream)
    torch.cuda.synchronize()
    log("warmup done")

    # Timed region: the product path (each step's launch sequence replayed
    # from its hipGraph), bracketed by synchronizations and the wall clock.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        tgt.restart_samples()
        tgt.render(SPP, stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    st = tgt.stats()
    assert st["samples"] == SPP
    rays_per_step = st["rays"]
    value = samples / (ms_per_step / 1e3) / 1e6
    sums, counts = tgt.read_accumulators()
    assert np.all(counts == SPP)
    tuning = tgt.tuning() if args.kernel == "wavefront" else None
    log(f"timed {args.steps} steps (graph path): {ms_per_step:.2f} ms/step, tuning {tuning}")
    # Instrumented pass (not the headline): HIP events around every launch on
    # the launch stream give each kernel class's GPU time.
    classes = []
    for k in range(max(2, min(args.steps, 5))):
        tgt.restart_samples()
        classes.append(tgt.render_timed(SPP, stream))
    class_ms = {k: float(np.mean([c[k] for c in classes])) for k in classes[0]}
    kernel_ms = class_ms["trace"] if args.kernel == "wavefront" else class_ms["megakernel"]
    eventful_ms = float(np.mean([sum(c.values()) for c in classes]))
    log(f"instrumented steps: per kernel class {class_ms} (sum {eventful_ms:.2f} ms)")
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9

    cpu, parity = None, None
    if not args.no_cpu_baseline and args.cpu_seconds > 0:
        log("timing the CPU baseline (oracle)")
        cpu, parity = cpu_baseline(sc, sums, args.cpu_seconds)

    denoiser = None
    if not args.no_denoiser:
        log("denoiser: U-Net at 1080p, FP16 and BF16")
        denoiser = denoiser_benchmark(args.denoise_full_precision)

    micro = None
    if not args.no_microbench:
        log("traversal/build microbench: 8M triangles, 16M rays")
        micro = microbench(subset=args.microbench_subset)

    serving = None
    log("mask coverage at the headline workload")
    alpha = alpha_benchmark(sums)
    log(f"mask coverage: {alpha}")
    if not args.no_serving:
        log("serving stream: batched vs one job at a time")
        serving = serving_benchmark(n_jobs=args.serving_jobs, per_sample=args.serving_per_sample)

    traffic = profile_value("wf_trace_c2_1080p_traffic.txt", "traffic_bytes_per_step")
    valu = profile_value("wf_c2_1080p_valu_issue.txt", "k_wf_trace_valu_issue_frac")
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1M-tri instanced field: 64 fBm-displaced icospheres x 15,680 tris + ground quad, "
                               "1024x512 sky+sun env, OpenPBR metal/glass/coat mix",
                   "width": WIDTH, "height": HEIGHT, "spp": SPP, "depth": DEPTH,
                   "direct_light_candidates": CANDIDATES, "sampling": "sobol seed 0",
                   "triangles": info["triangles"], "accumulation": "fp64", "kernel": args.kernel,
                   "tuning": tuning, "timed_path": "hipGraph replay"},
        "mrays_per_s": round(rays_per_step / (ms_per_step / 1e3) / 1e6, 3),
        "rays_per_step": rays_per_step,
        "scene_commit_ms": round(build_ms, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": KERNEL_NAMES[args.kernel], "kernel_ms": round(kernel_ms, 3),
                     "algorithmic_bytes_per_step": int(alg_bytes),
                     # The roof that binds this kernel is vector instruction issue, not HBM: the
                     # algorithmic bytes mostly hit in L2; measured traffic and issue rate follow.
                     "binding_roof": "valu_issue",
                     "valu_issue_frac": valu,
                     "hbm_traffic_frac": (round(traffic / (kernel_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                                          if traffic else None),
                     "work_per_step": work,
                     "kernel_class_ms_per_step": {k: round(v, 3) for k, v in class_ms.items()},
                     "instrumented_step_ms": round(eventful_ms, 3)},
        "cpu_baseline": cpu,
        "microbench": micro,
        "serving": serving,
        "mask_coverage": alpha,
        "denoiser": denoiser,
        "parity_relmse_cpu_tiles": parity,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
