// Wavefront path tracer for gfx950.
//
// The integrator of the megakernel (rayTrace.compute.ts integrator, :891-946)
// split at its two ray queries, so that traversal runs in small,
// high-occupancy kernels and shading in its own kernel:
//
//   generate  camera rays for every path of the generation
//   per event e = 0 .. min(3 * bounces, 50):
//     trace   closest hit for the paths in queue e          (LDS short stack)
//     shade   shade_event(): emission, NEE candidate choice, BSDF sample,
//             Russian roulette; appends the path to queue e+1 when it
//             continues and its shadow ray to the shadow queue of event e
//     shadow  any-hit for the shadow queue; adds the payload of each
//             unoccluded ray to its path's radiance
//   accumulate each pixel's samples, in sample order, into the FP64 running
//             sums and advance the per-pixel sample counter
//
// A generation holds the paths of one job (a render target) or of several
// jobs at once (the batched server): kMulti kernels look up each path's job
// slot, so the paths of different scenes, resolutions and sample counts share
// the same queues and launches. Queues are compacted with one atomic per wave
// (ballot + popcount). Every kernel reads its queue length from device
// memory, so a launch sequence needs no host round trip. Each path's
// arithmetic is exactly the megakernel's, in the same order, so every form
// produces identical accumulators.
#include "mtx_fetch.h"
#include "mtx_registry.h"
#include "mtx_shade.h"
#include "mtx_walk.h"

namespace mtx {

constexpr int kWfTraceBlock = 128;
constexpr int kWfShadeBlock = 128;

// Reserves one slot per lane with `pred` in the queue behind `counter`; must
// be reached by all lanes of the wave.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
  const uint64_t mask = __ballot(pred);
  const uint32_t total = (uint32_t)__popcll(mask);
  uint32_t base = 0;
  if (total) {
    if (lane_id() == 0) base = atomicAdd(counter, total);
    base = __shfl(base, 0);
  }
  return base + (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
}

__device__ __forceinline__ uint32_t *queue_of(const WfBuffers &W, int event) { return (event & 1) ? W.qB : W.qA; }

template <bool kMulti>
__device__ __forceinline__ const JobSlot &job_of(const WfLaunch &L, const WfBuffers &W, uint32_t p) {
  return kMulti ? L.jobs[W.item[p]] : L.one;
}

// Paths of a job are numbered pixel-major: path pathBase + local * samples + s
// is sample s of pixel local, so a wave holds a few pixels' samples and a
// contiguous queue range covers a few image rows (coherent rays, a small
// geometry working set per window of the queue).
__device__ __forceinline__ void path_pixel(const JobSlot &J, uint32_t p, uint32_t &local, uint32_t &s) {
  const uint32_t rel = p - J.pathBase;
  local = rel / J.params.samples;
  s = rel - local * J.params.samples;
}

// Sampler of path p: pixel and sample index within its job; the random
// pattern's pcg4d state lives in W.rng[p] between kernels.
__device__ __forceinline__ Smp path_smp(const JobSlot &J, const WfBuffers &W, uint32_t p) {
  const RenderParams &P = J.params;
  uint32_t local, s;
  path_pixel(J, p, local, s);
  const uint32_t pix = P.pixelBase + local;
  return path_sampler(P, pix % P.width, pix / P.width, J.acc.count[pix] + P.sampleOffset + s, W.rng + p);
}

__device__ __forceinline__ void store_path(const WfBuffers &W, uint32_t p, const PathRegs &pr) {
  W.o[p] = make_float4(pr.ro.x, pr.ro.y, pr.ro.z, pr.scatPdf);
  W.d[p] = make_float4(pr.rd.x, pr.rd.y, pr.rd.z, pr.alpha);
  W.li[p] = make_float4(pr.li.x, pr.li.y, pr.li.z, intf((int32_t)pr.flags));
  W.beta[p] = make_float4(pr.beta.x, pr.beta.y, pr.beta.z, intf(pr.curTri));
  // The next closest-hit query skips the last hit triangle from depth 1 on
  // (traverse's bounces > 0); a path at depth 0 stores slot -1, which no
  // instance has, so the trace kernels pass bounces = 1 for every path.
  W.alb[p] = make_float4(pr.auxAlbedo.x, pr.auxAlbedo.y, pr.auxAlbedo.z,
                         intf(path_depth(pr.flags) > 0 ? pr.curSlot : -1));
  W.nrm[p] = make_float4(pr.auxNormal.x, pr.auxNormal.y, pr.auxNormal.z, 0.0f);
}

// Queue lengths of a generation start at zero; queue 0 holds every path
// unless per-pixel sample targets select them.
__global__ void k_wf_reset(WfBuffers W, uint32_t initial) {
  for (int k = threadIdx.x; k < kWfCounters; k += blockDim.x) W.counts[k] = k == 0 ? initial : 0u;
}

// Whether sample s of path p's pixel is still below the pixel's target.
__device__ __forceinline__ bool below_target(const JobSlot &J, uint32_t p) {
  if (!J.acc.targets) return true;
  uint32_t local, s;
  path_pixel(J, p, local, s);
  const uint32_t pix = J.params.pixelBase + local;
  return (double)(J.acc.count[pix] + s) < (double)J.acc.targets[pix];
}

template <bool kMulti>
__global__ void __launch_bounds__(256) k_wf_generate(const WfLaunch L, WfBuffers W, int select) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nWaves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t w = wave; w * 64u < W.nPaths; w += nWaves) {
    const uint32_t p = w * 64u + lane_id();
    bool alive = p < W.nPaths;
    if (alive) {
      uint32_t j = 0;
      if (kMulti) {  // last slot with pathBase <= p
        uint32_t lo = 0, hi = L.nJobs;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (L.jobs[mid].pathBase <= p) lo = mid; else hi = mid;
        }
        j = lo;
        W.item[p] = (uint16_t)j;
      }
      const JobSlot &J = kMulti ? L.jobs[j] : L.one;
      if (select) alive = below_target(J, p);
      if (alive) {
        Smp smp = path_smp(J, W, p);
        PathRegs pr;
        init_path(J.params, smp, pr);
        store_path(W, p, pr);
      }
    }
    if (select) {
      const uint32_t slot = wave_append(W.counts, alive);
      if (alive) W.qA[slot] = p;
    } else if (alive) {
      W.qA[p] = p;
    }
  }
}

template <int N>
__device__ __forceinline__ void flush_work(uint64_t *work, const uint64_t (&v0)[N], const int (&slots)[N]) {
#pragma unroll
  for (int k = 0; k < N; k++) {
    uint64_t v = v0[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (lane_id() == 0 && v) atomicAdd((unsigned long long *)work + slots[k], (unsigned long long)v);
  }
}

template <bool kMulti, bool kCount, int kLds>
__global__ void __launch_bounds__(kWfTraceBlock, kTraceWaves)
k_wf_trace(const WfLaunch L, WfBuffers W, int event, int32_t *__restrict__ spill) {
  __shared__ int32_t lds[kLds * kWfTraceBlock];
  const uint32_t n = W.counts[event];
  if (n == 0u || W.counts[kWfTailFlag]) return;
  LaneStackT<kLds> st{lds + threadIdx.x, kWfTraceBlock,
                      spill + ((size_t)blockIdx.x * kWfTraceBlock + threadIdx.x) * (kMaxStack - kLds)};
  const uint32_t *q = queue_of(W, event);
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long *)L.rays, (unsigned long long)n);
  TraceCounters tc{0u, 0u, 0u};
  uint32_t hits = 0;
  AccelView one{};
  if (!kMulti) one = *L.one.scene.accel;
  auto ray = [&](uint32_t i) {
    const uint32_t p = q[i];
    const float4 o = W.o[p], d = W.d[p];
    const int32_t curTri = fint(W.beta[p].w), curSlot = fint(W.alb[p].w);
    TraceHit h;
    if (kMulti) {
      const AccelView av = *job_of<true>(L, W, p).scene.accel;
      traverse<false, kCount>(av, f3of(o), f3of(d), kRayMaxDistance, event != 0, 1, curTri, curSlot, st, h, &tc);
    } else {
      traverse<false, kCount>(one, f3of(o), f3of(d), kRayMaxDistance, event != 0, 1, curTri, curSlot, st, h, &tc);
    }
    W.hitA[p] = make_float4(h.t, h.bary.x, h.bary.y, h.bary.z);
    W.hitB[p] = make_int2(h.hit ? h.slot : -1, h.tri);
    if (kCount) hits += h.hit ? 1u : 0u;
  };
  for_each_ray<kTraceFetch>(n, W.counts + kWfXcdBase + (2 * event) * kWfXcds, ray);
  if (kCount) {
    const uint64_t v[5] = {(blockIdx.x == 0 && threadIdx.x == 0) ? (uint64_t)n : 0u, tc.boxes, tc.tris, tc.instances,
                           hits};
    const int slots[5] = {kWorkClosestRays, kWorkBoxes, kWorkTris, kWorkInstances, kWorkHits};
    flush_work(L.work, v, slots);
  }
}

// kWaves: occupancy the register allocator must reach (waves per SIMD); the
// allocator spills to reach it, so the best value is measured, not assumed.
// kFull: compiled with the fuzz, subsurface and thin-wall lobes and the
// random pattern (the launcher passes false when no job of the generation
// uses them).
template <bool kMulti, int kWaves, bool kFull>
__global__ void __launch_bounds__(kWfShadeBlock, kWaves) k_wf_shade(const WfLaunch L, WfBuffers W, int event) {
  const uint32_t n = W.counts[event];
  if (n == 0u || W.counts[kWfTailFlag]) return;
  const uint32_t *q = queue_of(W, event);
  uint32_t *qNext = queue_of(W, event + 1);
  uint32_t *nNext = W.counts + event + 1;
  uint32_t *nShadow = W.counts + kWfShadowBase + event;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nWaves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t w = wave; w * 64u < n; w += nWaves) {
    const uint32_t i = w * 64u + lane_id();
    const bool active = i < n;
    bool alive = false;
    ShadowReq sr;
    sr.valid = false;
    uint32_t p = 0;
    PathRegs pr;
    if (active) {
      p = q[i];
      const JobSlot &J = job_of<kMulti>(L, W, p);
      Smp smp = path_smp(J, W, p);
      const GpuScene &S = J.scene;
      const RenderParams &P = J.params;
      const float4 o = W.o[p], d = W.d[p], l = W.li[p], b = W.beta[p], al = W.alb[p], nr = W.nrm[p];
      pr.ro = f3of(o); pr.scatPdf = o.w;
      pr.rd = f3of(d); pr.alpha = d.w;
      pr.li = f3of(l); pr.flags = (uint32_t)fint(l.w);
      pr.beta = f3of(b); pr.curTri = fint(b.w);
      pr.auxAlbedo = f3of(al); pr.curSlot = fint(al.w);
      pr.auxNormal = f3of(nr);
      TraceHit h;
      const int2 hb = W.hitB[p];
      h.hit = hb.x >= 0;
      h.slot = hb.x;
      h.tri = hb.y;
      if (h.hit) {
        const float4 ha = W.hitA[p];
        h.t = ha.x;
        h.bary = f3{ha.y, ha.z, ha.w};
        // The traversal's object-space hit data, recomputed from the same inputs.
        const InstSlot &is = S.accel->slots[h.slot];
        const TriRec r = is.tris[h.tri];
        const f3 p0 = f3of(r.v0), p1 = f3of(r.v1), p2 = f3of(r.v2);
        h.normal = normalize(cross(p1 - p0, p2 - p0));
        const Affine w2l{f3of(is.c0), f3of(is.c1), f3of(is.c2), f3of(is.c3)};
        const f3 oo = xform(w2l, pr.ro, 1.0f), od = xform(w2l, pr.rd, 0.0f);
        h.position = oo + h.t * od;
      }
      const Lut lut{S.lut};
      alive = shade_event<kFull>(S, P, lut, pr, h, event, smp, sr);
      store_path(W, p, pr);
    }
    if (W.qSort) {
      const uint32_t oct = active ? (pr.rd.x < 0.0f ? 1u : 0u) | (pr.rd.y < 0.0f ? 2u : 0u) |
                                        (pr.rd.z < 0.0f ? 4u : 0u)
                                  : 0u;
      uint32_t *bins = W.counts + kWfSortBase + kWfSortBins * (event + 1);
      for (uint32_t o = 0; o < (uint32_t)kWfSortBins; o++) {
        const bool mine = active && alive && oct == o;
        const uint32_t slot = wave_append(bins + o, mine);
        if (mine) W.qSort[(size_t)o * W.nPaths + slot] = p;
      }
    } else {
      const uint32_t slot = wave_append(nNext, active && alive);
      if (active && alive) qNext[slot] = p;
    }
    const uint32_t sslot = wave_append(nShadow, active && sr.valid);
    if (active && sr.valid) {
      W.s0[sslot] = make_float4(sr.ro.x, sr.ro.y, sr.ro.z, sr.tMax);
      W.s1[sslot] = make_float4(sr.rd.x, sr.rd.y, sr.rd.z, intf((int32_t)p));
      W.s2[sslot] = make_float4(sr.payload.x, sr.payload.y, sr.payload.z, intf(pr.curTri));
      W.s3[sslot] = pr.curSlot;
    }
  }
}

// Concatenates the octant bins of queue `event` into the queue in bin order
// and sets the queue's length.
__global__ void __launch_bounds__(256) k_wf_gather(WfBuffers W, int event) {
  if (W.counts[kWfTailFlag]) return;
  const uint32_t *bins = W.counts + kWfSortBase + kWfSortBins * event;
  uint32_t start[kWfSortBins + 1];
  start[0] = 0;
#pragma unroll
  for (int o = 0; o < kWfSortBins; o++) start[o + 1] = start[o] + bins[o];
  const uint32_t total = start[kWfSortBins];
  if (blockIdx.x == 0 && threadIdx.x == 0) W.counts[event] = total;
  uint32_t *q = queue_of(W, event);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    uint32_t o = 0;
    while (i >= start[o + 1]) o++;
    q[i] = W.qSort[(size_t)o * W.nPaths + (i - start[o])];
  }
}

__device__ __forceinline__ void add_payload(const WfBuffers &W, uint32_t p, float4 c) {
  float4 l = W.li[p];
  l.x = l.x + c.x;
  l.y = l.y + c.y;
  l.z = l.z + c.z;
  W.li[p] = l;
}

template <bool kMulti, bool kCount, int kLds, bool kAlpha>
__global__ void __launch_bounds__(kWfTraceBlock, kShadowWaves)
k_wf_shadow(const WfLaunch L, WfBuffers W, int event, int32_t *__restrict__ spill) {
  __shared__ int32_t lds[kLds * kWfTraceBlock];
  const uint32_t n = W.counts[kWfShadowBase + event];
  if (n == 0u || W.counts[kWfTailFlag]) return;
  LaneStackT<kLds> st{lds + threadIdx.x, kWfTraceBlock,
                      spill + ((size_t)blockIdx.x * kWfTraceBlock + threadIdx.x) * (kMaxStack - kLds)};
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long *)L.rays, (unsigned long long)n);
  TraceCounters tc{0u, 0u, 0u};
  AccelView one{};
  if (!kMulti) one = *L.one.scene.accel;
  auto ray = [&](uint32_t i) {
    const float4 a = W.s0[i], b = W.s1[i], c = W.s2[i];
    const uint32_t p = (uint32_t)fint(b.w);
    TraceHit h;
    bool occluded;
    if (kMulti) {
      const GpuScene &S = job_of<true>(L, W, p).scene;
      const AccelView av = *S.accel;
      occluded = traverse<true, kCount>(av, f3of(a), f3of(b), a.w, true, 1, fint(c.w), W.s3[i], st, h, &tc,
                                        coverage<kAlpha>(&S));
    } else {
      occluded = traverse<true, kCount>(one, f3of(a), f3of(b), a.w, true, 1, fint(c.w), W.s3[i], st, h, &tc,
                                        coverage<kAlpha>(&L.one.scene));
    }
    if (!occluded) add_payload(W, p, c);
  };
  for_each_ray<kShadowFetch>(n, W.counts + kWfXcdBase + (2 * event + 1) * kWfXcds, ray);
  if (kCount) {
    const uint64_t v[4] = {tc.boxes, tc.tris, tc.instances, (blockIdx.x == 0 && threadIdx.x == 0) ? (uint64_t)n : 0u};
    const int slots[4] = {kWorkShadowBoxes, kWorkShadowTris, kWorkShadowInstances, kWorkShadowRays};
    flush_work(L.work, v, slots);
  }
}

// Tail of a generation: once event `event`'s queue holds at most `tailPaths`
// paths, one lane takes each of them through all its remaining events
// (closest hit, shade_event, shadow any-hit, in the order the wavefront
// kernels apply them), so the few long paths left (glass, deep bounces) no
// longer pay three launches and a grid-wide pass per event. The queue length
// was written by earlier kernels of the stream, so every block takes the same
// decision; the flag it raises makes the remaining wavefront launches of the
// generation return at once. Per-path arithmetic is that of the wavefront
// kernels, so results do not depend on where the tail starts.
template <bool kMulti, bool kCount, bool kFull, bool kAlpha>
__global__ void __launch_bounds__(kWfShadeBlock)
k_wf_tail(const WfLaunch L, WfBuffers W, int event, uint32_t tailPaths, int32_t *__restrict__ spill) {
  __shared__ int32_t lds[kMinLdsStack * kWfShadeBlock];
  const uint32_t n = W.counts[event];
  const uint32_t flag = W.counts[kWfTailFlag];
  if (n == 0u || n > tailPaths || (flag != 0u && flag != (uint32_t)event + 1u)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) W.counts[kWfTailFlag] = (uint32_t)event + 1u;
  LaneStackT<kMinLdsStack> st{lds + threadIdx.x, kWfShadeBlock,
                              spill + ((size_t)blockIdx.x * kWfShadeBlock + threadIdx.x) * (kMaxStack - kMinLdsStack)};
  const uint32_t *q = queue_of(W, event);
  uint64_t rays = 0;
  TraceCounters tc{0u, 0u, 0u}, ts{0u, 0u, 0u};
  uint64_t nClosest = 0, nHits = 0, nShadow = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t p = q[i];
    const JobSlot &J = job_of<kMulti>(L, W, p);
    const GpuScene &S = J.scene;
    const RenderParams &P = J.params;
    const AccelView av = *S.accel;
    const Lut lut{S.lut};
    Smp smp = path_smp(J, W, p);
    PathRegs pr;
    const float4 o = W.o[p], d = W.d[p], l = W.li[p], b = W.beta[p], al = W.alb[p], nr = W.nrm[p];
    pr.ro = f3of(o); pr.scatPdf = o.w;
    pr.rd = f3of(d); pr.alpha = d.w;
    pr.li = f3of(l); pr.flags = (uint32_t)fint(l.w);
    pr.beta = f3of(b); pr.curTri = fint(b.w);
    pr.auxAlbedo = f3of(al); pr.curSlot = fint(al.w);
    pr.auxNormal = f3of(nr);
    for (int e = event; e < kWfMaxEvents; e++) {
      TraceHit h;
      rays++;
      if (kCount) nClosest++;
      traverse<false, kCount>(av, pr.ro, pr.rd, kRayMaxDistance, e != 0, path_depth(pr.flags), pr.curTri, pr.curSlot,
                              st, h, &tc);
      if (kCount) nHits += h.hit ? 1u : 0u;
      ShadowReq sr;
      const bool alive = shade_event<kFull>(S, P, lut, pr, h, e, smp, sr);
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
    store_path(W, p, pr);
  }
  for (int off = 32; off > 0; off >>= 1) rays += __shfl_down(rays, off);
  if (lane_id() == 0 && rays) atomicAdd((unsigned long long *)L.rays, (unsigned long long)rays);
  if (kCount) {
    const uint64_t v[9] = {nClosest, tc.boxes, tc.tris, tc.instances, nHits, nShadow, ts.boxes, ts.tris, ts.instances};
    int slots[9];
#pragma unroll
    for (int k = 0; k < 9; k++) slots[k] = kWorkTailBase + k;
    flush_work(L.work, v, slots);
  }
}

// Persistent closest-hit / any-hit kernels over Walker: each lane takes a
// new queue entry (wave-aggregated atomic on counts[kWfFetchBase + 2e (+1)])
// as soon as its previous ray has finished.
template <bool kMulti, bool kShadow, bool kCount, int kLds, bool kAlpha>
__global__ void __launch_bounds__(kWfTraceBlock)
k_wf_walk(const WfLaunch L, WfBuffers W, int event, int32_t *__restrict__ spill) {
  __shared__ int32_t lds[kLds * kWfTraceBlock];
  const uint32_t n = W.counts[kShadow ? kWfShadowBase + event : event];
  if (n == 0u || W.counts[kWfTailFlag]) return;
  uint32_t *fetch = W.counts + kWfFetchBase + 2 * event + (kShadow ? 1 : 0);
  WalkStack<kLds> st{lds + threadIdx.x, kWfTraceBlock,
                     spill + ((size_t)blockIdx.x * kWfTraceBlock + threadIdx.x) * (kMaxStack - kLds)};
  const uint32_t *q = queue_of(W, event);
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long *)L.rays, (unsigned long long)n);
  AccelView av{};
  const GpuScene *scene = &L.one.scene;  // the scene of the lane's ray (coverage of masked instances)
  if (!kMulti) av = *L.one.scene.accel;
  Walker<kShadow, kCount> w;
  bool have = false, exhausted = false;
  uint32_t entry = 0, path = 0;
  uint64_t totals[4] = {0, 0, 0, 0};
  const bool doubleSided = kShadow ? true : event != 0;
  const int bounces = 1;  // closest hit: depth-0 paths store slot -1 (store_path)
  while (true) {
    const bool need = !have && !exhausted;
    const uint32_t idx = wave_append(fetch, need);
    if (need) {
      if (idx < n) {
        entry = idx;
        if (kShadow) {
          const float4 a = W.s0[idx], b = W.s1[idx], c = W.s2[idx];
          path = (uint32_t)fint(b.w);
          if (kMulti) {
            scene = &job_of<true>(L, W, path).scene;
            av = *scene->accel;
          }
          w.init(av, f3of(a), f3of(b), a.w, doubleSided, bounces, fint(c.w), W.s3[idx], st);
        } else {
          path = q[idx];
          if (kMulti) av = *job_of<true>(L, W, path).scene.accel;
          const float4 o = W.o[path], d = W.d[path];
          w.init(av, f3of(o), f3of(d), kRayMaxDistance, doubleSided, bounces, fint(W.beta[path].w),
                 fint(W.alb[path].w), st);
        }
        have = true;
      } else {
        exhausted = true;
      }
    }
    if (!__any(have)) break;
    if (have && (w.done || w.step(av, st, coverage<kAlpha>(scene)))) {
      have = false;
      if (kShadow) {
        if (!w.hit) add_payload(W, path, W.s2[entry]);
      } else {
        W.hitA[path] = make_float4(w.t, w.bary.x, w.bary.y, w.bary.z);
        W.hitB[path] = make_int2(w.hit ? w.hitSlot : -1, w.hitTri);
      }
      if (kCount) {
        totals[0] += w.cnt.boxes;
        totals[1] += w.cnt.tris;
        totals[2] += w.cnt.instances;
        totals[3] += (!kShadow && w.hit) ? 1u : 0u;
      }
    }
  }
  if (kCount) {
    const uint64_t first = (blockIdx.x == 0 && threadIdx.x == 0) ? (uint64_t)n : 0u;
    if (kShadow) {
      const uint64_t v[4] = {totals[0], totals[1], totals[2], first};
      const int slots[4] = {kWorkShadowBoxes, kWorkShadowTris, kWorkShadowInstances, kWorkShadowRays};
      flush_work(L.work, v, slots);
    } else {
      const uint64_t v[5] = {first, totals[0], totals[1], totals[2], totals[3]};
      const int slots[5] = {kWorkClosestRays, kWorkBoxes, kWorkTris, kWorkInstances, kWorkHits};
      flush_work(L.work, v, slots);
    }
  }
}

// Samples of the pixel a lane handles that the accumulation takes: all of
// them, or (adaptive sampling) those still below the pixel's target.
__device__ __forceinline__ uint32_t samples_taken(const Accumulators &acc, uint32_t pix, uint32_t count0,
                                                  uint32_t samples) {
  if (!acc.targets) return samples;
  const double target = (double)acc.targets[pix];
  uint32_t done = 0;
  while (done < samples && (double)(count0 + done) < target) done++;
  return done;
}

// One job with kS samples per pixel, kS dividing 64: a wave takes 64 / kS
// pixels with one lane per path (a pixel's paths are adjacent, so the path
// loads are coalesced), puts each sample's 13 values in LDS, and then lane
// (pixel, channel) adds the pixel's samples in sample order to its FP64 sum,
// the same additions in the same order as k_wf_accumulate.
template <int kS>
__global__ void __launch_bounds__(256) k_wf_accumulate_lanes(const WfLaunch L, WfBuffers W) {
  constexpr int kPx = 64 / kS;  // pixels per wave
  __shared__ double vals[4][64][13];
  __shared__ uint32_t count0s[4][kPx], dones[4][kPx];  // per pixel of a wave: count before, samples taken
  const JobSlot &J = L.one;
  const RenderParams &P = J.params;
  const Accumulators &acc = J.acc;
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint32_t total = P.pixelCount;
  for (uint32_t first = blockIdx.x * 4u * kPx; first < total; first += gridDim.x * 4u * kPx) {  // block-uniform
    const uint32_t base = first + wv * kPx;
    {
      const uint32_t local = base + lane / kS, s = lane % kS;
      if (local < total) {
        const uint32_t pix = P.pixelBase + local;
        const uint32_t count0 = acc.count[pix], done = samples_taken(acc, pix, count0, kS);
        if (s == 0u) {
          count0s[wv][lane / kS] = count0;
          dones[wv][lane / kS] = done;
        }
        if (s < done) {
          const uint32_t p = J.pathBase + local * kS + s;
          const float4 l = W.li[p], d = W.d[p], al = W.alb[p], nr = W.nrm[p];
          PathRegs pr;
          pr.li = f3of(l);
          pr.flags = (uint32_t)fint(l.w);
          pr.alpha = d.w;
          pr.auxAlbedo = f3of(al);
          pr.auxNormal = f3of(nr);
          double v[13];
          sample_values(finish_path(P, pr), v);
#pragma unroll
          for (int c = 0; c < 13; c++) vals[wv][lane][c] = v[c];
        }
      }
    }
    __syncthreads();
    for (uint32_t k = lane; k < (uint32_t)kPx * 13u; k += 64u) {
      const uint32_t slot = k / 13u, c = k - slot * 13u, local = base + slot;
      if (local >= total) continue;
      const uint32_t pix = P.pixelBase + local, done = dones[wv][slot];
      double sum = acc.planes[(size_t)c * acc.pixels + pix];
      for (uint32_t s = 0; s < done; s++) sum += vals[wv][slot * kS + s][c];
      acc.planes[(size_t)c * acc.pixels + pix] = sum;
      if (c == 0u) acc.count[pix] = count0s[wv][slot] + done;
    }
    __syncthreads();  // the next pixels reuse vals
  }
}

template <bool kMulti>
__global__ void __launch_bounds__(256) k_wf_accumulate(const WfLaunch L, WfBuffers W) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t total = kMulti ? L.totalPixels : L.one.params.pixelCount;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    uint32_t j = 0;
    if (kMulti) {  // last slot with pixelPrefix <= g
      uint32_t lo = 0, hi = L.nJobs;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (L.jobs[mid].pixelPrefix <= g) lo = mid; else hi = mid;
      }
      j = lo;
    }
    const JobSlot &J = kMulti ? L.jobs[j] : L.one;
    const RenderParams &P = J.params;
    const Accumulators &acc = J.acc;
    const uint32_t local = g - J.pixelPrefix;
    const uint32_t pix = P.pixelBase + local;
    double sums[13];
#pragma unroll
    for (int c = 0; c < 13; c++) sums[c] = acc.planes[(size_t)c * acc.pixels + pix];
    const uint32_t count0 = acc.count[pix];
    const uint32_t done = samples_taken(acc, pix, count0, P.samples);
    // Unrolled so that the loads of several samples are in flight at once;
    // the samples are still added in order.
#pragma unroll 4
    for (uint32_t s = 0; s < done; s++) {
      const uint32_t p = J.pathBase + local * P.samples + s;
      const float4 l = W.li[p], d = W.d[p], al = W.alb[p], nr = W.nrm[p];
      PathRegs pr;
      pr.li = f3of(l);
      pr.flags = (uint32_t)fint(l.w);
      pr.alpha = d.w;
      pr.auxAlbedo = f3of(al);
      pr.auxNormal = f3of(nr);
      add_sample(sums, finish_path(P, pr));
    }
#pragma unroll
    for (int c = 0; c < 13; c++) acc.planes[(size_t)c * acc.pixels + pix] = sums[c];
    acc.count[pix] = count0 + done;
  }
}

// ------------------------------------------------------------------ host
static int blocks_per_cu(const void *kernel, int block, size_t lds) {
  int n = 0;
  MTX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, lds));
  return std::max(1, n);
}
static int cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    MTX_HIP(hipGetDevice(&dev));
    MTX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cus = std::max(cus, 1);
  }
  return cus;
}

void WfBuffers::Owner::ensure(uint32_t paths, bool sort) {
  if (sort && sortCapacity < paths) {
    qSort.alloc((size_t)paths * kWfSortBins);
    sortCapacity = paths;
  }
  if (paths <= capacity) return;
  o.alloc(paths); d.alloc(paths); li.alloc(paths); beta.alloc(paths); alb.alloc(paths); nrm.alloc(paths);
  hitA.alloc(paths); hitB.alloc(paths); qA.alloc(paths); qB.alloc(paths);
  s0.alloc(paths); s1.alloc(paths); s2.alloc(paths); s3.alloc(paths); item.alloc(paths); rng.alloc(paths);
  if (!counts.ptr) counts.alloc(kWfCounters);
  capacity = paths;
}
size_t WfBuffers::Owner::bytes() const {
  return o.bytes() * 6 + hitA.bytes() + hitB.bytes() + qA.bytes() * 2 + s0.bytes() * 3 + s3.bytes() + item.bytes() +
         rng.bytes() + qSort.bytes() + counts.bytes();
}
WfBuffers WfBuffers::Owner::view(uint32_t nPaths, bool sort) const {
  WfBuffers W;
  W.o = o.ptr; W.d = d.ptr; W.li = li.ptr; W.beta = beta.ptr; W.alb = alb.ptr; W.nrm = nrm.ptr;
  W.hitA = hitA.ptr; W.hitB = hitB.ptr; W.qA = qA.ptr; W.qB = qB.ptr;
  W.s0 = s0.ptr; W.s1 = s1.ptr; W.s2 = s2.ptr; W.s3 = s3.ptr; W.item = item.ptr; W.rng = rng.ptr;
  W.qSort = sort ? qSort.ptr : nullptr;
  W.counts = counts.ptr;
  W.nPaths = nPaths;
  return W;
}

template <bool kMulti, bool kShadow, int kLds>
static void launch_walk_lds(uint32_t grid, hipStream_t s, const WfLaunch &L, const WfBuffers &W, int e,
                            int32_t *spill) {
  if (kShadow && L.masked) {
    if (L.work)
      hipLaunchKernelGGL((k_wf_walk<kMulti, kShadow, true, kLds, true>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W, e,
                         spill);
    else
      hipLaunchKernelGGL((k_wf_walk<kMulti, kShadow, false, kLds, true>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W,
                         e, spill);
  } else if (L.work) {
    hipLaunchKernelGGL((k_wf_walk<kMulti, kShadow, true, kLds, false>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W, e,
                       spill);
  } else {
    hipLaunchKernelGGL((k_wf_walk<kMulti, kShadow, false, kLds, false>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W,
                       e, spill);
  }
}
template <bool kMulti, bool kShadow>
static void launch_walk(const WfConfig &cfg, uint32_t grid, hipStream_t s, const WfLaunch &L, const WfBuffers &W,
                        int e, int32_t *spill) {
  if (cfg.ldsStack <= 16) launch_walk_lds<kMulti, kShadow, 16>(grid, s, L, W, e, spill);
  else if (cfg.ldsStack <= 24) launch_walk_lds<kMulti, kShadow, 24>(grid, s, L, W, e, spill);
  else launch_walk_lds<kMulti, kShadow, 32>(grid, s, L, W, e, spill);
}

template <bool kMulti, bool kShadow, int kLds>
static void launch_trace_lds(uint32_t grid, hipStream_t s, const WfLaunch &L, const WfBuffers &W, int e,
                             int32_t *spill) {
  if (kShadow) {
    if (L.masked) {
      if (L.work)
        hipLaunchKernelGGL((k_wf_shadow<kMulti, true, kLds, true>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W, e,
                           spill);
      else
        hipLaunchKernelGGL((k_wf_shadow<kMulti, false, kLds, true>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W, e,
                           spill);
    } else if (L.work) {
      hipLaunchKernelGGL((k_wf_shadow<kMulti, true, kLds, false>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W, e,
                         spill);
    } else {
      hipLaunchKernelGGL((k_wf_shadow<kMulti, false, kLds, false>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W, e,
                         spill);
    }
  } else {
    if (L.work)
      hipLaunchKernelGGL((k_wf_trace<kMulti, true, kLds>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W, e, spill);
    else
      hipLaunchKernelGGL((k_wf_trace<kMulti, false, kLds>), dim3(grid), dim3(kWfTraceBlock), 0, s, L, W, e, spill);
  }
}
// Closest-hit (kShadow false) or any-hit launch with the configured LDS stack depth.
template <bool kMulti, bool kShadow>
static void launch_trace(const WfConfig &cfg, uint32_t grid, hipStream_t s, const WfLaunch &L, const WfBuffers &W,
                         int e, int32_t *spill) {
  const int depth = (!kShadow && cfg.ldsStackClosest) ? cfg.ldsStackClosest : cfg.ldsStack;
  if (depth <= 16) launch_trace_lds<kMulti, kShadow, 16>(grid, s, L, W, e, spill);
  else if (depth <= 24) launch_trace_lds<kMulti, kShadow, 24>(grid, s, L, W, e, spill);
  else launch_trace_lds<kMulti, kShadow, 32>(grid, s, L, W, e, spill);
}

struct Grids {
  uint32_t trace, shadow, walk, shade, tail, cus;
};
static Grids grids_for(const WfConfig &cfg) {
  static int closestBpc[3] = {0, 0, 0}, shadowBpc[3] = {0, 0, 0}, shadeBpc[5] = {0, 0, 0, 0, 0},
             walkBpc[3] = {0, 0, 0}, tailBpc = 0;
  const int cus = cu_count();
  auto level = [](int d) { return d <= 16 ? 0 : (d <= 24 ? 1 : 2); };
  const int li = level(cfg.ldsStack), lc = level(cfg.ldsStackClosest ? cfg.ldsStackClosest : cfg.ldsStack);
  if (!closestBpc[lc]) {
    const void *t = lc == 0 ? (const void *)k_wf_trace<false, false, 16>
                  : lc == 1 ? (const void *)k_wf_trace<false, false, 24> : (const void *)k_wf_trace<false, false, 32>;
    closestBpc[lc] = blocks_per_cu(t, kWfTraceBlock, 0);
  }
  if (!shadowBpc[li]) {
    const void *a = li == 0   ? (const void *)k_wf_shadow<false, false, 16, false>
                    : li == 1 ? (const void *)k_wf_shadow<false, false, 24, false>
                              : (const void *)k_wf_shadow<false, false, 32, false>;
    shadowBpc[li] = blocks_per_cu(a, kWfTraceBlock, 0);
  }

  const int sw = std::min(4, std::max(1, cfg.shadeWaves));
  if (!shadeBpc[sw]) {
    const void *k = sw == 1 ? (const void *)k_wf_shade<false, 1, true> : sw == 2 ? (const void *)k_wf_shade<false, 2, true>
                  : sw == 3 ? (const void *)k_wf_shade<false, 3, true> : (const void *)k_wf_shade<false, 4, true>;
    shadeBpc[sw] = blocks_per_cu(k, kWfShadeBlock, 0);
  }
  if (!walkBpc[li]) {
    const void *k = li == 0 ? (const void *)k_wf_walk<false, false, false, 16, false>
                  : li == 1 ? (const void *)k_wf_walk<false, false, false, 24, false>
                            : (const void *)k_wf_walk<false, false, false, 32, false>;
    walkBpc[li] = blocks_per_cu(k, kWfTraceBlock, 0);
  }
  if (!tailBpc) tailBpc = blocks_per_cu((const void *)k_wf_tail<false, false, true, false>, kWfShadeBlock, 0);
  Grids g;
  g.cus = (uint32_t)cus;
  g.tail = (uint32_t)(cus * tailBpc);
  g.trace = (uint32_t)(cus * closestBpc[lc] * std::max(1, cfg.traceOversub));
  g.shadow = (uint32_t)(cus * shadowBpc[li] * std::max(1, cfg.traceOversub));
  g.walk = (uint32_t)(cus * walkBpc[li] * std::max(1, cfg.traceOversub));
  g.shade = (uint32_t)(cus * shadeBpc[sw] * std::max(1, cfg.shadeOversub));
  return g;
}

template <bool kMulti, bool kFull>
static void launch_shade(const WfConfig &cfg, const Grids &g, hipStream_t s, const WfLaunch &L, const WfBuffers &W,
                         int e) {
  switch (cfg.shadeWaves) {
    case 2: hipLaunchKernelGGL((k_wf_shade<kMulti, 2, kFull>), dim3(g.shade), dim3(kWfShadeBlock), 0, s, L, W, e); break;
    case 3: hipLaunchKernelGGL((k_wf_shade<kMulti, 3, kFull>), dim3(g.shade), dim3(kWfShadeBlock), 0, s, L, W, e); break;
    case 4: hipLaunchKernelGGL((k_wf_shade<kMulti, 4, kFull>), dim3(g.shade), dim3(kWfShadeBlock), 0, s, L, W, e); break;
    default: hipLaunchKernelGGL((k_wf_shade<kMulti, 1, kFull>), dim3(g.shade), dim3(kWfShadeBlock), 0, s, L, W, e); break;
  }
}
template <bool kMulti, bool kFull>
static void launch_tail(const WfConfig &cfg, const Grids &g, hipStream_t s, const WfLaunch &L, const WfBuffers &W,
                        int e, int32_t *spill) {
  if (L.masked) {
    if (L.work)
      hipLaunchKernelGGL((k_wf_tail<kMulti, true, kFull, true>), dim3(g.tail), dim3(kWfShadeBlock), 0, s, L, W, e,
                         cfg.tailPaths, spill);
    else
      hipLaunchKernelGGL((k_wf_tail<kMulti, false, kFull, true>), dim3(g.tail), dim3(kWfShadeBlock), 0, s, L, W, e,
                         cfg.tailPaths, spill);
  } else if (L.work) {
    hipLaunchKernelGGL((k_wf_tail<kMulti, true, kFull, false>), dim3(g.tail), dim3(kWfShadeBlock), 0, s, L, W, e,
                       cfg.tailPaths, spill);
  } else {
    hipLaunchKernelGGL((k_wf_tail<kMulti, false, kFull, false>), dim3(g.tail), dim3(kWfShadeBlock), 0, s, L, W, e,
                       cfg.tailPaths, spill);
  }
}

// Launches one generation: generate, the event loop, accumulate. `full`:
// some job uses fuzz, subsurface or thin-walled materials or the random
// pattern (else the shading kernels compiled without them run).
template <bool kMulti>
static void run_generation(const WfLaunch &L, const WfBuffers &W, int eventLimit, uint32_t pixels, const WfConfig &cfg,
                           const Grids &g, int32_t *spill, hipStream_t s, LaunchTimer *timer, bool full) {
  auto tb = [&](int c) { if (timer) timer->begin(c); };
  auto te = [&]() { if (timer) timer->end(); };
  const uint32_t genGrid = std::min<uint32_t>((W.nPaths + 255) / 256, g.cus * 16u);
  const bool select = !kMulti && L.one.acc.targets != nullptr;  // server jobs have no targets
  tb(0);
  hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(256), 0, s, W, select ? 0u : W.nPaths);
  hipLaunchKernelGGL(k_wf_generate<kMulti>, dim3(genGrid), dim3(256), 0, s, L, W, select ? 1 : 0);
  te();
  for (int e = 0; e <= eventLimit; e++) {
    if (cfg.tailPaths) {
      tb(5);
      if (full) launch_tail<kMulti, true>(cfg, g, s, L, W, e, spill);
      else launch_tail<kMulti, false>(cfg, g, s, L, W, e, spill);
      te();
    }
    tb(1);
    if (cfg.walker)
      launch_walk<kMulti, false>(cfg, g.walk, s, L, W, e, spill);
    else
      launch_trace<kMulti, false>(cfg, g.trace, s, L, W, e, spill);
    te();
    tb(2);
    if (full) launch_shade<kMulti, true>(cfg, g, s, L, W, e);
    else launch_shade<kMulti, false>(cfg, g, s, L, W, e);
    if (W.qSort && e < eventLimit) hipLaunchKernelGGL(k_wf_gather, dim3(genGrid), dim3(256), 0, s, W, e + 1);
    te();
    tb(3);
    if (cfg.walker)
      launch_walk<kMulti, true>(cfg, g.walk, s, L, W, e, spill);
    else
      launch_trace<kMulti, true>(cfg, g.shadow, s, L, W, e, spill);
    te();
  }
  const uint32_t accGrid = std::min<uint32_t>((pixels + 255) / 256, g.cus * 16u);
  tb(0);
  const uint32_t spp = L.one.params.samples;
  if (!kMulti && spp != 0u && spp <= 64u && 64u % spp == 0u) {
    const uint32_t blocks = std::min<uint32_t>((pixels + 4u * (64u / spp) - 1u) / (4u * (64u / spp)), g.cus * 16u);
    switch (spp) {
      case 1: hipLaunchKernelGGL(k_wf_accumulate_lanes<1>, dim3(blocks), dim3(256), 0, s, L, W); break;
      case 2: hipLaunchKernelGGL(k_wf_accumulate_lanes<2>, dim3(blocks), dim3(256), 0, s, L, W); break;
      case 4: hipLaunchKernelGGL(k_wf_accumulate_lanes<4>, dim3(blocks), dim3(256), 0, s, L, W); break;
      case 8: hipLaunchKernelGGL(k_wf_accumulate_lanes<8>, dim3(blocks), dim3(256), 0, s, L, W); break;
      case 16: hipLaunchKernelGGL(k_wf_accumulate_lanes<16>, dim3(blocks), dim3(256), 0, s, L, W); break;
      case 32: hipLaunchKernelGGL(k_wf_accumulate_lanes<32>, dim3(blocks), dim3(256), 0, s, L, W); break;
      default: hipLaunchKernelGGL(k_wf_accumulate_lanes<64>, dim3(blocks), dim3(256), 0, s, L, W); break;
    }
  } else {
    hipLaunchKernelGGL(k_wf_accumulate<kMulti>, dim3(accGrid), dim3(256), 0, s, L, W);
  }
  te();
  MTX_HIP(hipGetLastError());
}

void WfGraph::reset() {
  if (exec) (void)hipGraphExecDestroy(exec);
  exec = nullptr;
  key = 0;
}
WfGraph::~WfGraph() {
  reset();
  if (own) (void)hipStreamDestroy(own);
}

static uint64_t fnv(uint64_t h, const void *data, size_t n) {
  const unsigned char *b = (const unsigned char *)data;
  for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

void launch_render_wavefront(const GpuScene &scene, const RenderParams &p0, const Accumulators &acc,
                             WfBuffers::Owner &buf, const WfConfig &cfg, hipStream_t s, LaunchTimer *timer,
                             WfGraph *graph) {
  if (!p0.pixelCount || !p0.samples) return;
  const bool full = needs_full_shading(scene, p0);
  const uint32_t maxPaths = std::max<uint32_t>(cfg.maxPaths, p0.pixelCount);
  const uint32_t genMax = std::max<uint32_t>(1u, maxPaths / p0.pixelCount);
  const Grids g = grids_for(cfg);
  int32_t *spill = spill_for(s, std::max(std::max(std::max(g.trace, g.shadow), g.walk), g.tail), kWfTraceBlock);
  const int eventLimit = std::min((int)p0.bounces * 3, 50);
  uint32_t done = 0;
  while (done < p0.samples) {
    const uint32_t gen = std::min(genMax, p0.samples - done);
    const uint32_t nPaths = gen * p0.pixelCount;
    buf.ensure(nPaths, cfg.queueSort);
    const WfBuffers W = buf.view(nPaths, cfg.queueSort);
    WfLaunch L;
    std::memset((void *)&L, 0, sizeof L);  // padding too: the bytes key the graph cache
    L.one.scene = scene;
    L.one.params = p0;
    L.one.params.samples = gen;
    L.one.acc = acc;
    L.one.pathBase = 0;
    L.one.pixelPrefix = 0;
    L.jobs = nullptr;
    L.nJobs = 1;
    L.totalPixels = p0.pixelCount;
    L.rays = acc.rays;
    L.work = acc.work;
    L.masked = (scene.features & kSceneMasked) != 0u ? 1u : 0u;
    if (graph && cfg.graphs && !timer) {
      // The sequence has no host-side dependence on device values, so one
      // capture replays for as long as its arguments are unchanged.
      hipStream_t gs = s;
      if (!gs) {
        if (!graph->own) MTX_HIP(hipStreamCreate(&graph->own));
        gs = graph->own;  // blocking stream: ordered with the legacy null stream
      }
      uint64_t key = 1469598103934665603ull;
      static_assert(sizeof(GpuScene) == 152 && sizeof(RenderParams) % 4 == 0, "key layouts");
      key = fnv(key, &L.one.scene, sizeof L.one.scene);
      key = fnv(key, &L.one.params, sizeof L.one.params);
      const void *ptrs[8] = {L.one.acc.planes, L.one.acc.count, L.one.acc.rays, L.one.acc.work, W.o, W.counts,
                             L.one.acc.targets, W.rng};
      key = fnv(key, ptrs, sizeof ptrs);
      key = fnv(key, &L.one.acc.pixels, sizeof L.one.acc.pixels);
      key = fnv(key, &W.nPaths, sizeof W.nPaths);
      const int64_t cf[10] = {cfg.maxPaths, cfg.traceOversub, cfg.shadeOversub, cfg.walker, cfg.ldsStack,
                              cfg.shadeWaves, cfg.graphs, cfg.tailPaths, cfg.queueSort, cfg.ldsStackClosest};
      key = fnv(key, cf, sizeof cf);
      key = fnv(key, &spill, sizeof spill);
      key = fnv(key, &eventLimit, sizeof eventLimit);
      key = fnv(key, &gs, sizeof gs);
      if (!graph->exec || graph->key != key) {
        graph->reset();
        hipGraph_t gr = nullptr;
        MTX_HIP(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
        run_generation<false>(L, W, eventLimit, p0.pixelCount, cfg, g, spill, gs, nullptr, full);
        MTX_HIP(hipStreamEndCapture(gs, &gr));
        hipError_t e = hipGraphInstantiate(&graph->exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        MTX_HIP(e);
        graph->key = key;
        graph->captures++;
      }
      MTX_HIP(hipGraphLaunch(graph->exec, gs));
      graph->replays++;
    } else {
      run_generation<false>(L, W, eventLimit, p0.pixelCount, cfg, g, spill, s, timer, full);
    }
    done += gen;
  }
}

void launch_generation_multi(const JobSlot *devSlots, uint32_t nJobs, uint32_t nPaths, uint32_t totalPixels,
                             int eventLimit, uint64_t *rays, WfBuffers::Owner &buf, const WfConfig &cfg,
                             hipStream_t s, bool full, bool masked, LaunchTimer *timer, WfGraph *graph) {
  if (!nJobs || !nPaths) return;
  MTX_REQUIRE(nJobs <= kMaxJobsPerRound, MTX_INVALID_ARGUMENT, "too many jobs in one round");
  const Grids g = grids_for(cfg);
  int32_t *spill = spill_for(s, std::max(std::max(std::max(g.trace, g.shadow), g.walk), g.tail), kWfTraceBlock);
  buf.ensure(nPaths, cfg.queueSort);
  const WfBuffers W = buf.view(nPaths, cfg.queueSort);
  WfLaunch L;
  std::memset((void *)&L, 0, sizeof L);
  L.jobs = devSlots;
  L.nJobs = nJobs;
  L.totalPixels = totalPixels;
  L.rays = rays;
  L.work = nullptr;
  L.masked = masked ? 1u : 0u;
  const int events = std::min(eventLimit, 50);
  if (graph && cfg.graphs && !timer) {
    // The caller's slot upload precedes the launch on the same stream (the
    // server's own blocking stream is ordered with the null stream).
    hipStream_t gs = s;
    if (!gs) {
      if (!graph->own) MTX_HIP(hipStreamCreate(&graph->own));
      gs = graph->own;
    }
    uint64_t key = 1469598103934665603ull;
    key = fnv(key, &L, sizeof L);
    const void *ptrs[4] = {W.o, W.counts, W.rng, spill};
    key = fnv(key, ptrs, sizeof ptrs);
    const int64_t cf[13] = {cfg.maxPaths, cfg.traceOversub, cfg.shadeOversub, cfg.walker, cfg.ldsStack,
                            cfg.shadeWaves, cfg.graphs, cfg.tailPaths, cfg.queueSort, cfg.ldsStackClosest,
                            (int64_t)nPaths, (int64_t)events, (int64_t)full};
    key = fnv(key, cf, sizeof cf);
    key = fnv(key, &gs, sizeof gs);
    if (!graph->exec || graph->key != key) {
      graph->reset();
      hipGraph_t gr = nullptr;
      MTX_HIP(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
      run_generation<true>(L, W, events, totalPixels, cfg, g, spill, gs, nullptr, full);
      MTX_HIP(hipStreamEndCapture(gs, &gr));
      hipError_t e = hipGraphInstantiate(&graph->exec, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      MTX_HIP(e);
      graph->key = key;
      graph->captures++;
    }
    MTX_HIP(hipGraphLaunch(graph->exec, gs));
    graph->replays++;
    return;
  }
  run_generation<true>(L, W, events, totalPixels, cfg, g, spill, s, timer, full);
}

}  // namespace mtx
