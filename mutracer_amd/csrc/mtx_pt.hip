// Path-tracing megakernel for gfx950.
//
// One lane traces one path sample of one pixel: camera ray with the
// Owen-scrambled Sobol jitter, the event loop of rayTrace.compute.ts
// integrator (:891-946) with completeVertex (:514-589), shadeVertex
// (:343-512), SampleMaterial / sampleScattering / finishDirectLight
// (sampleMaterial.wgsl.ts) and the OpenPBR lobes of openPbrBsdf.wgsl.ts
// used by the scenes (coat, F82 metal + multiple scattering, dielectric
// specular, solid transmission + dielectric multiple scattering, EON
// diffuse). Traversal is the shared two-level traverse() of mtx_accel.h with
// the per-lane LDS stack. Results are accumulated in FP64 per pixel
// (running sums of color, alpha, albedo, normal and luminance squared), the
// per-pixel sample counter giving each pixel's next Sobol sample index.
#include "mtx_registry.h"
#include "mtx_shade.h"

namespace mtx {

// kCount adds per-launch totals of the traversal work to acc.work (first
// group of mtx_target_render_counted: closest rays, closest box/triangle/
// instance tests, hits, shadow rays, shadow box/triangle/instance tests).
template <bool kCount, bool kAlpha>
__global__ void __launch_bounds__(kBlock)
k_render(const GpuScene S, const RenderParams P, Accumulators acc, int32_t *__restrict__ spill) {
  __shared__ int32_t lds[kLdsStack * kBlock];
  LaneStack st{lds + threadIdx.x, kBlock,
               spill + ((size_t)blockIdx.x * kBlock + threadIdx.x) * (kMaxStack - kLdsStack)};
  const AccelView av = *S.accel;
  const Lut L{S.lut};
  uint64_t rays = 0;
  TraceCounters tc{0u, 0u, 0u}, ts{0u, 0u, 0u};
  uint64_t nClosest = 0, nHits = 0, nShadow = 0;
  const uint32_t nBlocks = (P.pixelCount + kBlock - 1) / kBlock;
  // XCD-aware ordering: blocks resident on one XCD (blockIdx % 8) walk one
  // contiguous eighth of the pixel range so neighbouring pixels share an L2.
  const uint32_t xcd = blockIdx.x & 7u, rank = blockIdx.x >> 3, perXcd = (gridDim.x + 7u) >> 3;
  const uint32_t span = (nBlocks + 7u) >> 3;
  const uint32_t bEnd = min(nBlocks, (xcd + 1u) * span);
  const int eventLimit = min((int)P.bounces * 3, 50);
  for (uint32_t pb = xcd * span + rank; pb < bEnd; pb += perXcd) {
    const uint32_t local = pb * kBlock + threadIdx.x;
    if (local >= P.pixelCount) continue;
    const uint32_t pix = P.pixelBase + local;
    const uint32_t px = pix % P.width, py = pix / P.width;
    double sums[13];
#pragma unroll
    for (int c = 0; c < 13; c++) sums[c] = acc.planes[(size_t)c * acc.pixels + pix];
    const uint32_t count0 = acc.count[pix];
    uint32_t done = 0;
    for (uint32_t k = 0; k < P.samples; k++) {
      if (acc.targets && (double)(count0 + k) >= (double)acc.targets[pix]) break;
      done++;
      uint4 rngState;  // pcg4d state of the random pattern
      Smp smp = path_sampler(P, px, py, count0 + k + P.sampleOffset, &rngState);
      PathRegs pr;
      init_path(P, smp, pr);
      for (int event = 0; event <= eventLimit; event++) {
        TraceHit h;
        rays++;
        if (kCount) nClosest++;
        traverse<false, kCount>(av, pr.ro, pr.rd, kRayMaxDistance, event != 0, path_depth(pr.flags), pr.curTri,
                                pr.curSlot, st, h, &tc);
        if (kCount) nHits += h.hit ? 1u : 0u;
        ShadowReq sr;
        const bool alive = shade_event(S, P, L, pr, h, event, smp, sr);
        if (sr.valid) {
          TraceHit sh;
          rays++;
          if (kCount) nShadow++;
          if (!traverse<true, kCount>(av, sr.ro, sr.rd, sr.tMax, true, 1, pr.curTri, pr.curSlot, st, sh, &ts,
                                      coverage<kAlpha>(&S)))
            pr.li = pr.li + sr.payload;
        }
        if (!alive) break;
      }
      add_sample(sums, finish_path(P, pr));
    }
#pragma unroll
    for (int c = 0; c < 13; c++) acc.planes[(size_t)c * acc.pixels + pix] = sums[c];
    acc.count[pix] = count0 + done;
  }
  // One atomic per wave for the ray counter.
  for (int off = 32; off > 0; off >>= 1) rays += __shfl_down(rays, off);
  if ((threadIdx.x & 63) == 0 && rays) atomicAdd((unsigned long long *)acc.rays, (unsigned long long)rays);
  if (kCount) {
    uint64_t v[kWorkGroup] = {nClosest, tc.boxes, tc.tris, tc.instances, nHits, nShadow, ts.boxes, ts.tris, ts.instances};
#pragma unroll
    for (int k = 0; k < kWorkGroup; k++) {
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off);
      if ((threadIdx.x & 63) == 0 && v[k]) atomicAdd((unsigned long long *)acc.work + k, (unsigned long long)v[k]);
    }
  }
}

void LaunchTimer::begin(int cls) {
  hipEvent_t e;
  MTX_HIP(hipEventCreate(&e));
  MTX_HIP(hipEventRecord(e, stream));
  events.push_back(e);
  classes.push_back(cls);
}
void LaunchTimer::end() {
  hipEvent_t e;
  MTX_HIP(hipEventCreate(&e));
  MTX_HIP(hipEventRecord(e, stream));
  events.push_back(e);
}
void LaunchTimer::collect(double ms[kTimerClasses]) {
  for (int k = 0; k < kTimerClasses; k++) ms[k] = 0.0;
  if (!events.empty()) MTX_HIP(hipEventSynchronize(events.back()));
  for (size_t i = 0; i + 1 < events.size(); i += 2) {
    float t = 0.0f;
    MTX_HIP(hipEventElapsedTime(&t, events[i], events[i + 1]));
    ms[classes[i / 2]] += (double)t;
  }
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  events.clear();
  classes.clear();
}
LaunchTimer::~LaunchTimer() {
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
}

void launch_render(const GpuScene &scene, const RenderParams &p, const Accumulators &acc, hipStream_t s,
                   LaunchTimer *timer) {
  if (!p.pixelCount || !p.samples) return;
  uint32_t grid = trace_grid();
  uint32_t need = ((p.pixelCount + kBlock - 1) / kBlock + 7u) / 8u * 8u;
  grid = std::min(grid, need);
  int32_t *spill = spill_for(s, grid, kBlock);
  if (timer) timer->begin(4);
  // Shadow rays test coverage only in scenes with mask materials.
  const bool masked = (scene.features & kSceneMasked) != 0u;
  if (acc.work && masked)
    hipLaunchKernelGGL((k_render<true, true>), dim3(grid), dim3(kBlock), 0, s, scene, p, acc, spill);
  else if (acc.work)
    hipLaunchKernelGGL((k_render<true, false>), dim3(grid), dim3(kBlock), 0, s, scene, p, acc, spill);
  else if (masked)
    hipLaunchKernelGGL((k_render<false, true>), dim3(grid), dim3(kBlock), 0, s, scene, p, acc, spill);
  else
    hipLaunchKernelGGL((k_render<false, false>), dim3(grid), dim3(kBlock), 0, s, scene, p, acc, spill);
  if (timer) timer->end();
  MTX_HIP(hipGetLastError());
}

}  // namespace mtx
