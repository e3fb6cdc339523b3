// Wavefront path tracer for gfx950.
//
// The integrator of the megakernel (rayTrace.compute.ts integrator, :891-946)
// split at its two ray queries, so that traversal runs in small,
// high-occupancy kernels and shading in its own kernel:
//
//   generate  camera rays for G samples of every pixel (G * pixels paths)
//   per event e = 0 .. min(3 * bounces, 50):
//     trace   closest hit for the paths in queue e          (LDS short stack)
//     shade   shade_event(): emission, NEE candidate choice, BSDF sample,
//             Russian roulette; appends the path to queue e+1 when it
//             continues and its shadow ray to the shadow queue of event e
//     shadow  any-hit for the shadow queue; adds the payload of each
//             unoccluded ray to its path's radiance
//   accumulate the G samples of each pixel, in sample order, into the FP64
//             running sums and advance the per-pixel sample counter
//
// Queues are compacted with one atomic per wave (ballot + popcount). Every
// kernel reads its queue length from device memory, so a launch sequence
// needs no host round trip and can be captured into a graph. Each path's
// arithmetic is exactly the megakernel's, in the same order, so both produce
// identical accumulators.
#include "mtx_registry.h"
#include "mtx_shade.h"
#include "mtx_walk.h"

namespace mtx {

constexpr int kWfTraceBlock = 128;
constexpr int kWfShadeBlock = 128;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

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

__device__ __forceinline__ const uint32_t *queue_of(const WfBuffers &W, int event) {
  return (event & 1) ? W.qB : W.qA;
}
__device__ __forceinline__ uint32_t *queue_of_mut(const WfBuffers &W, int event) {
  return (event & 1) ? W.qB : W.qA;
}

__global__ void __launch_bounds__(256) k_wf_generate(const RenderParams P, Accumulators acc, WfBuffers W) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0) {
    W.counts[0] = W.nPaths;
    for (int k = 1; k < kWfCounters; k++) W.counts[k] = 0u;
  }
  for (uint32_t p = tid; p < W.nPaths; p += stride) {
    const uint32_t local = p % P.pixelCount, s = p / P.pixelCount;
    const uint32_t pix = P.pixelBase + local;
    const uint32_t px = pix % P.width, py = pix / P.width;
    Smp smp = path_sampler(P, px, py, acc.count[pix] + P.sampleOffset + s);
    PathRegs pr;
    init_path(P, smp, pr);
    W.o[p] = make_float4(pr.ro.x, pr.ro.y, pr.ro.z, pr.scatPdf);
    W.d[p] = make_float4(pr.rd.x, pr.rd.y, pr.rd.z, pr.alpha);
    W.li[p] = make_float4(pr.li.x, pr.li.y, pr.li.z, intf((int32_t)pr.flags));
    W.beta[p] = make_float4(pr.beta.x, pr.beta.y, pr.beta.z, intf(pr.curTri));
    W.alb[p] = make_float4(pr.auxAlbedo.x, pr.auxAlbedo.y, pr.auxAlbedo.z, intf(pr.curSlot));
    W.nrm[p] = make_float4(pr.auxNormal.x, pr.auxNormal.y, pr.auxNormal.z, 0.0f);
    W.qA[p] = p;
  }
}

template <bool kCount>
__global__ void __launch_bounds__(kWfTraceBlock)
k_wf_trace(const AccelView *__restrict__ accel, WfBuffers W, int event, int32_t *__restrict__ spill,
           Accumulators acc) {
  __shared__ int32_t lds[kLdsStack * kWfTraceBlock];
  const uint32_t n = W.counts[event];
  if (n == 0u) return;
  LaneStack st{lds + threadIdx.x, kWfTraceBlock,
               spill + ((size_t)blockIdx.x * kWfTraceBlock + threadIdx.x) * (kMaxStack - kLdsStack)};
  const AccelView av = *accel;
  const uint32_t *q = queue_of(W, event);
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  if (tid == 0) atomicAdd((unsigned long long *)acc.rays, (unsigned long long)n);
  TraceCounters tc{0u, 0u, 0u};
  uint32_t hits = 0;
  for (uint32_t i = tid; i < n; i += stride) {
    const uint32_t p = q[i];
    const float4 o = W.o[p], d = W.d[p];
    const int32_t curTri = fint(W.beta[p].w), curSlot = fint(W.alb[p].w);
    TraceHit h;
    traverse<false, kCount>(av, f3of(o), f3of(d), kRayMaxDistance, event != 0, trace_depth(event), curTri, curSlot,
                            st, h, &tc);
    W.hitA[p] = make_float4(h.t, h.bary.x, h.bary.y, h.bary.z);
    W.hitB[p] = make_int2(h.hit ? h.slot : -1, h.tri);
    if (kCount) hits += h.hit ? 1u : 0u;
  }
  if (kCount) {
    uint64_t v[4] = {tc.boxes, tc.tris, tc.instances, hits};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off);
      if (lane_id() == 0 && v[k]) atomicAdd((unsigned long long *)acc.work + k, (unsigned long long)v[k]);
    }
  }
}

__global__ void __launch_bounds__(kWfShadeBlock)
k_wf_shade(const GpuScene S, const RenderParams P, Accumulators acc, WfBuffers W, int event) {
  const uint32_t n = W.counts[event];
  if (n == 0u) return;
  const Lut L{S.lut};
  const uint32_t *q = queue_of(W, event);
  uint32_t *qNext = queue_of_mut(W, event + 1);
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
      const uint32_t local = p % P.pixelCount, s = p / P.pixelCount;
      const uint32_t pix = P.pixelBase + local;
      Smp smp = path_sampler(P, pix % P.width, pix / P.width, acc.count[pix] + P.sampleOffset + s);
      alive = shade_event(S, P, L, pr, h, event, smp, sr);
      W.o[p] = make_float4(pr.ro.x, pr.ro.y, pr.ro.z, pr.scatPdf);
      W.d[p] = make_float4(pr.rd.x, pr.rd.y, pr.rd.z, pr.alpha);
      W.li[p] = make_float4(pr.li.x, pr.li.y, pr.li.z, intf((int32_t)pr.flags));
      W.beta[p] = make_float4(pr.beta.x, pr.beta.y, pr.beta.z, intf(pr.curTri));
      W.alb[p] = make_float4(pr.auxAlbedo.x, pr.auxAlbedo.y, pr.auxAlbedo.z, intf(pr.curSlot));
      W.nrm[p] = make_float4(pr.auxNormal.x, pr.auxNormal.y, pr.auxNormal.z, 0.0f);
    }
    const uint32_t slot = wave_append(nNext, active && alive);
    if (active && alive) qNext[slot] = p;
    const uint32_t sslot = wave_append(nShadow, active && sr.valid);
    if (active && sr.valid) {
      W.s0[sslot] = make_float4(sr.ro.x, sr.ro.y, sr.ro.z, sr.tMax);
      W.s1[sslot] = make_float4(sr.rd.x, sr.rd.y, sr.rd.z, intf((int32_t)p));
      W.s2[sslot] = make_float4(sr.payload.x, sr.payload.y, sr.payload.z, intf(pr.curTri));
      W.s3[sslot] = pr.curSlot;
    }
  }
}

template <bool kCount>
__global__ void __launch_bounds__(kWfTraceBlock)
k_wf_shadow(const AccelView *__restrict__ accel, WfBuffers W, int event, int32_t *__restrict__ spill,
            Accumulators acc) {
  __shared__ int32_t lds[kLdsStack * kWfTraceBlock];
  const uint32_t n = W.counts[kWfShadowBase + event];
  if (n == 0u) return;
  LaneStack st{lds + threadIdx.x, kWfTraceBlock,
               spill + ((size_t)blockIdx.x * kWfTraceBlock + threadIdx.x) * (kMaxStack - kLdsStack)};
  const AccelView av = *accel;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  if (tid == 0) atomicAdd((unsigned long long *)acc.rays, (unsigned long long)n);
  TraceCounters tc{0u, 0u, 0u};
  for (uint32_t i = tid; i < n; i += stride) {
    const float4 a = W.s0[i], b = W.s1[i], c = W.s2[i];
    TraceHit h;
    if (!traverse<true, kCount>(av, f3of(a), f3of(b), a.w, true, 1, fint(c.w), W.s3[i], st, h, &tc)) {
      const uint32_t p = (uint32_t)fint(b.w);
      float4 l = W.li[p];
      l.x = l.x + c.x;
      l.y = l.y + c.y;
      l.z = l.z + c.z;
      W.li[p] = l;
    }
  }
  if (kCount) {
    uint64_t v[4] = {tc.boxes, tc.tris, tc.instances, 0};
    v[3] = (blockIdx.x == 0 && threadIdx.x == 0) ? n : 0u;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off);
      if (lane_id() == 0 && v[k]) atomicAdd((unsigned long long *)acc.work + (k < 3 ? 5 + k : 4), (unsigned long long)v[k]);
    }
  }
}

// Persistent closest-hit / any-hit kernels over Walker: each lane takes a
// new queue entry (wave-aggregated atomic on counts[kWfFetchBase + 2e (+1)])
// as soon as its previous ray has finished.
template <bool kShadow, bool kCount, int kLds>
__global__ void __launch_bounds__(kWfTraceBlock)
k_wf_walk(const AccelView *__restrict__ accel, WfBuffers W, int event, int32_t *__restrict__ spill, Accumulators acc) {
  __shared__ int32_t lds[kLds * kWfTraceBlock];
  const uint32_t n = W.counts[kShadow ? kWfShadowBase + event : event];
  if (n == 0u) return;
  uint32_t *fetch = W.counts + kWfFetchBase + 2 * event + (kShadow ? 1 : 0);
  WalkStack<kLds> st{lds + threadIdx.x, kWfTraceBlock,
                     spill + ((size_t)blockIdx.x * kWfTraceBlock + threadIdx.x) * (kMaxStack - kLds)};
  const AccelView av = *accel;
  const uint32_t *q = queue_of(W, event);
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long *)acc.rays, (unsigned long long)n);
  Walker<kShadow, kCount> w;
  bool have = false, exhausted = false;
  uint32_t entry = 0;
  uint64_t totals[4] = {0, 0, 0, 0};
  const bool doubleSided = kShadow ? true : event != 0;
  const int bounces = kShadow ? 1 : trace_depth(event);
  while (true) {
    const bool need = !have && !exhausted;
    const uint32_t idx = wave_append(fetch, need);
    if (need) {
      if (idx < n) {
        entry = idx;
        if (kShadow) {
          const float4 a = W.s0[idx], b = W.s1[idx], c = W.s2[idx];
          w.init(av, f3of(a), f3of(b), a.w, doubleSided, bounces, fint(c.w), W.s3[idx], st);
        } else {
          const uint32_t p = q[idx];
          const float4 o = W.o[p], d = W.d[p];
          w.init(av, f3of(o), f3of(d), kRayMaxDistance, doubleSided, bounces, fint(W.beta[p].w), fint(W.alb[p].w),
                 st);
        }
        have = true;
      } else {
        exhausted = true;
      }
    }
    if (!__any(have)) break;
    if (have && (w.done || w.step(av, st))) {
      have = false;
      if (kShadow) {
        if (!w.hit) {
          const uint32_t p = (uint32_t)fint(W.s1[entry].w);
          const float4 c = W.s2[entry];
          float4 l = W.li[p];
          l.x = l.x + c.x;
          l.y = l.y + c.y;
          l.z = l.z + c.z;
          W.li[p] = l;
        }
      } else {
        const uint32_t p = q[entry];
        W.hitA[p] = make_float4(w.t, w.bary.x, w.bary.y, w.bary.z);
        W.hitB[p] = make_int2(w.hit ? w.hitSlot : -1, w.hitTri);
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
    if (kShadow && blockIdx.x == 0 && threadIdx.x == 0) totals[3] = n;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint64_t v = totals[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
      const int slot = k < 3 ? (kShadow ? 5 + k : k) : (kShadow ? 4 : 3);
      if (lane_id() == 0 && v) atomicAdd((unsigned long long *)acc.work + slot, (unsigned long long)v);
    }
  }
}

__global__ void __launch_bounds__(256) k_wf_accumulate(const RenderParams P, Accumulators acc, WfBuffers W) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t local = blockIdx.x * blockDim.x + threadIdx.x; local < P.pixelCount; local += stride) {
    const uint32_t pix = P.pixelBase + local;
    double sums[13];
#pragma unroll
    for (int c = 0; c < 13; c++) sums[c] = acc.planes[(size_t)c * acc.pixels + pix];
    for (uint32_t s = 0; s < W.gen; s++) {
      const uint32_t p = s * P.pixelCount + local;
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
    acc.count[pix] += W.gen;
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

void WfBuffers::Owner::ensure(uint32_t paths) {
  if (paths <= capacity) return;
  o.alloc(paths); d.alloc(paths); li.alloc(paths); beta.alloc(paths); alb.alloc(paths); nrm.alloc(paths);
  hitA.alloc(paths); hitB.alloc(paths); qA.alloc(paths); qB.alloc(paths);
  s0.alloc(paths); s1.alloc(paths); s2.alloc(paths); s3.alloc(paths);
  if (!counts.ptr) counts.alloc(kWfCounters);
  capacity = paths;
}
size_t WfBuffers::Owner::bytes() const {
  return o.bytes() * 6 + hitA.bytes() + hitB.bytes() + qA.bytes() * 2 + s0.bytes() * 3 + s3.bytes() + counts.bytes();
}
WfBuffers WfBuffers::Owner::view(uint32_t nPaths, uint32_t gen) const {
  WfBuffers W;
  W.o = o.ptr; W.d = d.ptr; W.li = li.ptr; W.beta = beta.ptr; W.alb = alb.ptr; W.nrm = nrm.ptr;
  W.hitA = hitA.ptr; W.hitB = hitB.ptr; W.qA = qA.ptr; W.qB = qB.ptr;
  W.s0 = s0.ptr; W.s1 = s1.ptr; W.s2 = s2.ptr; W.s3 = s3.ptr;
  W.counts = counts.ptr;
  W.nPaths = nPaths;
  W.gen = gen;
  return W;
}

template <bool kShadow, int kLds>
static void launch_walk_lds(uint32_t grid, hipStream_t s, const AccelView *accel, const WfBuffers &W, int e,
                            int32_t *spill, const Accumulators &acc) {
  if (acc.work)
    hipLaunchKernelGGL((k_wf_walk<kShadow, true, kLds>), dim3(grid), dim3(kWfTraceBlock), 0, s, accel, W, e, spill, acc);
  else
    hipLaunchKernelGGL((k_wf_walk<kShadow, false, kLds>), dim3(grid), dim3(kWfTraceBlock), 0, s, accel, W, e, spill, acc);
}
template <bool kShadow>
static void launch_walk(const WfConfig &cfg, uint32_t grid, hipStream_t s, const AccelView *accel, const WfBuffers &W,
                        int e, int32_t *spill, const Accumulators &acc) {
  if (cfg.ldsStack <= 16) launch_walk_lds<kShadow, 16>(grid, s, accel, W, e, spill, acc);
  else if (cfg.ldsStack <= 24) launch_walk_lds<kShadow, 24>(grid, s, accel, W, e, spill, acc);
  else launch_walk_lds<kShadow, 32>(grid, s, accel, W, e, spill, acc);
}
static const void *walk_kernel(int lds) {
  if (lds <= 16) return (const void *)k_wf_walk<false, false, 16>;
  if (lds <= 24) return (const void *)k_wf_walk<false, false, 24>;
  return (const void *)k_wf_walk<false, false, 32>;
}

void launch_render_wavefront(const GpuScene &scene, const RenderParams &p0, const Accumulators &acc,
                             WfBuffers::Owner &buf, const WfConfig &cfg, hipStream_t s, LaunchTimer *timer) {
  auto tb = [&](int c) { if (timer) timer->begin(c); };
  auto te = [&]() { if (timer) timer->end(); };
  if (!p0.pixelCount || !p0.samples) return;
  const uint32_t maxPaths = std::max<uint32_t>(cfg.maxPaths, p0.pixelCount);
  const uint32_t genMax = std::max<uint32_t>(1u, maxPaths / p0.pixelCount);
  const int cus = cu_count();
  static int traceBpc = 0, shadowBpc = 0, shadeBpc = 0;
  if (!traceBpc) {
    traceBpc = blocks_per_cu((const void *)k_wf_trace<false>, kWfTraceBlock, 0);
    shadowBpc = blocks_per_cu((const void *)k_wf_shadow<false>, kWfTraceBlock, 0);
    shadeBpc = blocks_per_cu((const void *)k_wf_shade, kWfShadeBlock, 0);
  }
  const uint32_t traceGrid = (uint32_t)(cus * std::max(traceBpc, shadowBpc) * std::max(1, cfg.traceOversub));
  const uint32_t shadeGrid = (uint32_t)(cus * shadeBpc * std::max(1, cfg.shadeOversub));
  static int walkBpc[3] = {0, 0, 0};
  const int li = cfg.ldsStack <= 16 ? 0 : (cfg.ldsStack <= 24 ? 1 : 2);
  if (!walkBpc[li]) walkBpc[li] = blocks_per_cu(walk_kernel(cfg.ldsStack), kWfTraceBlock, 0);
  const uint32_t walkGrid = (uint32_t)(cus * walkBpc[li] * std::max(1, cfg.traceOversub));
  int32_t *spill = spill_for(s, std::max(traceGrid, walkGrid), kWfTraceBlock);
  const int eventLimit = std::min((int)p0.bounces * 3, 50);
  uint32_t done = 0;
  while (done < p0.samples) {
    const uint32_t gen = std::min(genMax, p0.samples - done);
    const uint32_t nPaths = gen * p0.pixelCount;
    buf.ensure(nPaths);
    WfBuffers W = buf.view(nPaths, gen);
    RenderParams P = p0;
    P.samples = gen;
    const uint32_t genGrid = std::min<uint32_t>((nPaths + 255) / 256, (uint32_t)cus * 16u);
    tb(0);
    hipLaunchKernelGGL(k_wf_generate, dim3(genGrid), dim3(256), 0, s, P, acc, W);
    te();
    for (int e = 0; e <= eventLimit; e++) {
      tb(1);
      if (cfg.walker) {
        launch_walk<false>(cfg, walkGrid, s, scene.accel, W, e, spill, acc);
      } else if (acc.work) {
        hipLaunchKernelGGL(k_wf_trace<true>, dim3(traceGrid), dim3(kWfTraceBlock), 0, s, scene.accel, W, e, spill, acc);
      } else {
        hipLaunchKernelGGL(k_wf_trace<false>, dim3(traceGrid), dim3(kWfTraceBlock), 0, s, scene.accel, W, e, spill, acc);
      }
      te();
      tb(2);
      hipLaunchKernelGGL(k_wf_shade, dim3(shadeGrid), dim3(kWfShadeBlock), 0, s, scene, P, acc, W, e);
      te();
      tb(3);
      if (cfg.walker) {
        launch_walk<true>(cfg, walkGrid, s, scene.accel, W, e, spill, acc);
      } else if (acc.work) {
        hipLaunchKernelGGL(k_wf_shadow<true>, dim3(traceGrid), dim3(kWfTraceBlock), 0, s, scene.accel, W, e, spill, acc);
      } else {
        hipLaunchKernelGGL(k_wf_shadow<false>, dim3(traceGrid), dim3(kWfTraceBlock), 0, s, scene.accel, W, e, spill, acc);
      }
      te();
    }
    const uint32_t accGrid = std::min<uint32_t>((p0.pixelCount + 255) / 256, (uint32_t)cus * 16u);
    tb(0);
    hipLaunchKernelGGL(k_wf_accumulate, dim3(accGrid), dim3(256), 0, s, P, acc, W);
    te();
    MTX_HIP(hipGetLastError());
    done += gen;
  }
}

}  // namespace mtx
