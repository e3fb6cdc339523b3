// Work distribution and occupancy targets of the traversal kernels (gfx950:
// 64-lane waves, 8 XCDs each with its own L2).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mtx {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

constexpr int kXcds = 8;

// Occupancy the register allocator must reach for the traversal kernels
// (waves per SIMD). Measured on the 1M-triangle 1080p workload: closest-hit
// is fastest at 6 (80 VGPRs, a few spilled; 5 unconstrained, 8 spills too
// much), any-hit at 8 (64 VGPRs); re-measured with the unified closest-hit
// node step, 5 and 7 are slightly slower. The 16-entry LDS stack keeps LDS
// from capping these.
constexpr int kTraceWaves = 6;
constexpr int kShadowWaves = 8;


// Work distribution of the traversal kernels. The queue of n entries is cut
// into kXcds contiguous segments, one per XCD (workgroups are dispatched
// to the XCDs round-robin, so block b runs on XCD b % 8). Neighbouring queue
// entries are spatially coherent rays (neighbouring pixels of pixel-major
// paths, or a caller's ordered batch), so each XCD's L2 serves a small part
// of the scene at a time.
//   mode 0: grid-stride loop over the whole queue (no segments)
//   mode 1: a wave takes the next kFetchChunk entries of its XCD's segment
//           with one atomic and helps the other segments once its own is
//           drained
//   mode 2: static segments, each XCD's blocks stride over their own
// Measured on the 1M-triangle 1080p workload: closest-hit is fastest with
// mode 1 (its rays are long enough to amortize the atomic), any-hit with
// mode 0; mode 2 loses to load imbalance, chunks above 64 serialize the
// short queues of late events.
constexpr int kTraceFetch = 1;
constexpr int kShadowFetch = 0;
constexpr uint32_t kFetchChunk = 64;
__device__ __forceinline__ bool fetch_chunk(uint32_t *ctr, uint32_t n, uint32_t &seg, uint32_t &base,
                                            uint32_t &limit) {
  uint32_t b0 = 0, l0 = 0, s0 = seg;
  if (lane_id() == 0) {
    for (uint32_t k = 0; k < (uint32_t)kXcds; k++) {
      const uint32_t x = (seg + k) % (uint32_t)kXcds;
      const uint32_t b = (uint32_t)(((uint64_t)n * x) / kXcds);
      const uint32_t e = (uint32_t)(((uint64_t)n * (x + 1)) / kXcds);
      if (b >= e) continue;
      const uint32_t c = atomicAdd(ctr + x, 1u);
      if (c < (e - b + kFetchChunk - 1u) / kFetchChunk) {
        b0 = b + c * kFetchChunk;
        l0 = min(e, b0 + kFetchChunk);
        s0 = x;
        break;
      }
    }
  }
  base = __shfl(b0, 0);
  limit = __shfl(l0, 0);
  seg = __shfl(s0, 0);
  return base < limit;
}

// Calls ray(i) once for every queue entry i < n, spread over the grid by kFetchMode.
template <int kFetchMode, class F>
__device__ __forceinline__ void for_each_ray(uint32_t n, uint32_t *ctr, F &&ray) {
  if (kFetchMode == 1) {
    // Waves beyond one per chunk could only probe drained segments: they
    // leave at once (the first ceil(n / 64) waves keep fetching until every
    // segment is drained, so no chunk is left behind).
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= (n + kFetchChunk - 1u) / kFetchChunk) return;
    uint32_t seg = blockIdx.x % (uint32_t)kXcds, base, limit;
    while (fetch_chunk(ctr, n, seg, base, limit))
      for (uint32_t i = base + lane_id(); i < limit; i += 64u) ray(i);
  } else if (kFetchMode == 2 && gridDim.x % (uint32_t)kXcds == 0u) {
    const uint32_t x = blockIdx.x % (uint32_t)kXcds, nb = gridDim.x / (uint32_t)kXcds;
    const uint32_t b = (uint32_t)(((uint64_t)n * x) / kXcds), e = (uint32_t)(((uint64_t)n * (x + 1)) / kXcds);
    for (uint32_t i = b + (blockIdx.x / kXcds) * blockDim.x + threadIdx.x; i < e; i += nb * blockDim.x) ray(i);
  } else {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ray(i);
  }
}

}  // namespace mtx
