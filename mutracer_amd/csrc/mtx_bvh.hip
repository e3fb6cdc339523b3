// GPU binned-SAH BVH2 builder (gfx950) producing the reference builder's tree.
//
// The reference (src/renderer/binnedSahBvh.ts:70-236) builds depth-first on
// one CPU thread: per node it bins centroids on all three axes, sweeps the
// bins for the lowest SAH cost in f64, and partitions the node's primitive
// range in place with a swap loop. Every per-node decision depends only on
// the node's own range, so this builder processes nodes breadth-first and in
// parallel and still reproduces the tree exactly:
//   - bin boxes are min/max of f32 values (exact, order-free; kept as
//     order-preserving uint keys for atomics), counts are integer sums;
//   - the sweep runs the reference's f64 expression sequence on one lane;
//   - the in-place swap loop's final permutation is computed in parallel:
//     with x_k the k-th right-side element from the front and z_k the k-th
//     left-side element from the back, the loop performs P = #{k : x_k < z_k}
//     exchanges; left elements among the first P from the back move to x_r,
//     the k-th front right element moves to z_{k-1}-1 (z_0 = n), the (P+1)-th
//     moves to z_P-1 when x_{P+1} < z_P, and every other right element moves
//     down by one slot (the loop's back stream shifts by one).
// Node numbering differs from the reference's (children are allocated as
// adjacent pairs with atomics), but the depth-first layout, which is what the
// traversal and the exported BVH2 buffer use, depends only on the topology.
//
// Phases: nodes above kSmall primitives are split level by level with many
// workgroups per node (kChunk elements each); a node of at most kSmall
// primitives is finished by one workgroup that builds its whole subtree in
// LDS. A final pass orders interior nodes depth-first and emits the
// traversal layout of mtx_accel.h.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "mtx_scene.h"

namespace mtx {

constexpr int kChunk = 2048;
constexpr int kBuildThreads = 256;
constexpr double kFar = 1e30;

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int32_t js_to_int32(double v) {
  if (!isfinite(v)) return 0;
  double t = trunc(v);
  if (t >= -2147483648.0 && t < 2147483648.0) return (int32_t)t;
  double m = fmod(t, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return (int32_t)(uint32_t)m;
}

// binnedSahBvh.ts:46-49
__device__ __forceinline__ int32_t bin_of(float center, double nodeMin, double scale, int32_t bins) {
  int32_t b = js_to_int32(((double)center - nodeMin) * scale);
  return b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
}

struct NodeFrame {
  double nMin[3], extent[3], scale[3], area;
};

// Component i of a three-element local array through constant indices: a
// runtime index into a local array places the array (and the struct that
// holds it) in scratch memory.
__device__ __forceinline__ double pick3(const double *a, int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

__device__ __forceinline__ NodeFrame node_frame(const float *lo, const float *hi, int32_t bins,
                                                const double *minDim) {
  NodeFrame f;
  for (int a = 0; a < 3; a++) {
    f.nMin[a] = (double)lo[a];
    f.extent[a] = (double)hi[a] - f.nMin[a];
  }
  f.area = f.extent[0] * f.extent[1] + f.extent[1] * f.extent[2] + f.extent[2] * f.extent[0];
  for (int a = 0; a < 3; a++) f.scale[a] = f.extent[a] > minDim[a] ? (double)bins / f.extent[a] : 0.0;
  return f;
}

// A node's split: axis, last bin of the left side, the two children's boxes
// (the SAH sweep of binnedSahBvh.ts:153-224, sah_split_wave, over bins
// stored as keys: kmin/kmax[(axis*bins + i)*3 + k], cnt[axis*bins + i]).
struct Split {
  int32_t axis, pos;
  float lo[2][3], hi[2][3];
};


// ------------------------------------------------------------ input kernels
__global__ void k_tri_aabbs(const float *__restrict__ pos, const uint32_t *__restrict__ idx,
                            uint32_t triCount, uint32_t vertexCount, float4 *__restrict__ lo,
                            float4 *__restrict__ hi, uint32_t *__restrict__ valid) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= triCount) return;
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool ok = true;
  for (int c = 0; c < 3; c++) {
    uint32_t v = idx[t * 3 + c];
    if (v >= vertexCount) { ok = false; break; }
    for (int k = 0; k < 3; k++) {
      float x = pos[(size_t)v * 3 + k];
      mn[k] = (isnan(x) || isnan(mn[k])) ? NAN : jminf(mn[k], x);
      mx[k] = (isnan(x) || isnan(mx[k])) ? NAN : jmaxf(mx[k], x);
    }
  }
  for (int k = 0; k < 3; k++) ok = ok && isfinite(mn[k]) && isfinite(mx[k]);
  lo[t] = make_float4(mn[0], mn[1], mn[2], 0.0f);
  hi[t] = make_float4(mx[0], mx[1], mx[2], 0.0f);
  valid[t] = ok ? 1u : 0u;
}

__global__ void k_compact(const float4 *__restrict__ lo, const float4 *__restrict__ hi,
                          const uint32_t *__restrict__ valid, const uint32_t *__restrict__ offs,
                          uint32_t n, float4 *__restrict__ plo, float4 *__restrict__ phi,
                          float4 *__restrict__ cen, uint32_t *__restrict__ source) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || !valid[t]) return;
  uint32_t p = offs[t];
  float4 a = lo[t], b = hi[t];
  plo[p] = a;
  phi[p] = b;
  cen[p] = make_float4((float)(((double)a.x + (double)b.x) * 0.5), (float)(((double)a.y + (double)b.y) * 0.5),
                       (float)(((double)a.z + (double)b.z) * 0.5), 0.0f);
  source[p] = t;
}

__global__ void k_centers(const float4 *__restrict__ plo, const float4 *__restrict__ phi, uint32_t n,
                          float4 *__restrict__ cen, uint32_t *__restrict__ source) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float4 a = plo[p], b = phi[p];
  cen[p] = make_float4((float)(((double)a.x + (double)b.x) * 0.5), (float)(((double)a.y + (double)b.y) * 0.5),
                       (float)(((double)a.z + (double)b.z) * 0.5), 0.0f);
  if (source) source[p] = p;
}

// Root box with the reference's strict-compare scan: equal values (+0/-0)
// keep the earliest primitive's value. 64-bit keys: (value key, index).
__device__ __forceinline__ uint32_t canon_key(float v) { return fkey(v == 0.0f ? 0.0f : v); }

// Grid-stride: each thread reduces its primitives, then its wave, then the
// block (LDS), and the block issues six atomics.
__global__ void __launch_bounds__(256) k_root_reduce(const float4 *__restrict__ plo, const float4 *__restrict__ phi,
                                                     uint32_t n, unsigned long long *__restrict__ keys,
                                                     uint32_t *__restrict__ prim) {
  __shared__ unsigned long long part[4][6];
  unsigned long long k[6];
  for (int i = 0; i < 6; i++) k[i] = (i < 3) ? ~0ull : 0ull;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    prim[p] = p;
    const float4 a = plo[p], b = phi[p];
    const float mn[3] = {a.x, a.y, a.z}, mx[3] = {b.x, b.y, b.z};
    for (int i = 0; i < 3; i++) {
      const unsigned long long lo = ((unsigned long long)canon_key(mn[i]) << 32) | p;
      const unsigned long long hi = ((unsigned long long)canon_key(mx[i]) << 32) | (0xffffffffu - p);
      k[i] = lo < k[i] ? lo : k[i];
      k[3 + i] = hi > k[3 + i] ? hi : k[3 + i];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int i = 0; i < 6; i++) {
      unsigned long long o = __shfl_down(k[i], off);
      k[i] = i < 3 ? (o < k[i] ? o : k[i]) : (o > k[i] ? o : k[i]);
    }
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int i = 0; i < 6; i++) part[wave][i] = k[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++)
      for (int i = 0; i < 6; i++)
        k[i] = i < 3 ? (part[w][i] < k[i] ? part[w][i] : k[i]) : (part[w][i] > k[i] ? part[w][i] : k[i]);
    for (int i = 0; i < 3; i++) atomicMin(&keys[i], k[i]);
    for (int i = 3; i < 6; i++) atomicMax(&keys[i], k[i]);
  }
}

__global__ void k_root_init(const unsigned long long *__restrict__ keys, const float4 *__restrict__ plo,
                            const float4 *__restrict__ phi, uint32_t n, BNode *__restrict__ nodes,
                            uint32_t *__restrict__ counter) {
  if (threadIdx.x != 0) return;
  BNode r{};
  for (int i = 0; i < 3; i++) {
    uint32_t pmin = (uint32_t)(keys[i] & 0xffffffffu);
    uint32_t pmax = 0xffffffffu - (uint32_t)(keys[3 + i] & 0xffffffffu);
    float4 a = plo[pmin], b = phi[pmax];
    r.lo[i] = i == 0 ? a.x : (i == 1 ? a.y : a.z);
    r.hi[i] = i == 0 ? b.x : (i == 1 ? b.y : b.z);
  }
  r.leftFirst = 0;
  r.primCount = n;
  r.depth = 1;
  r.parent = -1;
  nodes[0] = r;
  *counter = 1;
}

// -------------------------------------------------------- big-node phase
struct Chunk {
  uint32_t big;    // index into the level's big-node list
  uint32_t begin;  // element range relative to the node's first
  uint32_t end;
  uint32_t pad;
};

struct BigState {
  uint32_t node;
  uint32_t split;   // 1 when the node splits this level
  int32_t axis, pos;
  uint32_t totalL, totalR, P, pad;
  float clo[2][3], chi[2][3];
};

// The level's bins (nb * 3 * bins of them; nb read on the device).
__global__ void k_bins_clear(uint32_t *__restrict__ kmin, uint32_t *__restrict__ kmax,
                             uint32_t *__restrict__ cnt, const uint32_t *__restrict__ nbp, int32_t bins) {
  const uint32_t nBins = *nbp * 3u * (uint32_t)bins;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nBins * 3; i += gridDim.x * blockDim.x) {
    if (i < nBins) cnt[i] = 0;
    kmin[i] = 0xffffffffu;
    kmax[i] = 0u;
  }
}

__global__ void __launch_bounds__(kBuildThreads)
k_bin_chunks(const Chunk *__restrict__ chunks, const BigState *__restrict__ big,
             const BNode *__restrict__ nodes, const uint32_t *__restrict__ prim,
             const float4 *__restrict__ plo, const float4 *__restrict__ phi,
             const float4 *__restrict__ cen, int32_t bins, double md0, double md1, double md2,
             uint32_t *__restrict__ gkmin, uint32_t *__restrict__ gkmax, uint32_t *__restrict__ gcnt,
             const uint32_t *__restrict__ ncp) {
  __shared__ uint32_t kmin[3 * kMaxBins * 3], kmax[3 * kMaxBins * 3], cnt[3 * kMaxBins];
  if (blockIdx.x >= *ncp) return;
  const Chunk ch = chunks[blockIdx.x];
  const BNode nd = nodes[big[ch.big].node];
  const double minDim[3] = {md0, md1, md2};
  NodeFrame f = node_frame(nd.lo, nd.hi, bins, minDim);
  for (int i = threadIdx.x; i < 3 * bins * 3; i += blockDim.x) { kmin[i] = 0xffffffffu; kmax[i] = 0u; }
  for (int i = threadIdx.x; i < 3 * bins; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += blockDim.x) {
    uint32_t p = prim[nd.leftFirst + e];
    float4 c = cen[p], a = plo[p], b = phi[p];
    float cc[3] = {c.x, c.y, c.z};
    float mn[3] = {a.x, a.y, a.z}, mx[3] = {b.x, b.y, b.z};
    for (int ax = 0; ax < 3; ax++) {
      int bin = ax * bins + bin_of(cc[ax], f.nMin[ax], f.scale[ax], bins);
      atomicAdd(&cnt[bin], 1u);
      for (int k = 0; k < 3; k++) {
        atomicMin(&kmin[bin * 3 + k], fkey(mn[k]));
        atomicMax(&kmax[bin * 3 + k], fkey(mx[k]));
      }
    }
  }
  __syncthreads();
  uint32_t base = ch.big * 3 * bins;
  for (int i = threadIdx.x; i < 3 * bins; i += blockDim.x) {
    if (!cnt[i]) continue;
    atomicAdd(&gcnt[base + i], cnt[i]);
    for (int k = 0; k < 3; k++) {
      atomicMin(&gkmin[(base + i) * 3 + k], kmin[i * 3 + k]);
      atomicMax(&gkmax[(base + i) * 3 + k], kmax[i * 3 + k]);
    }
  }
}

__device__ __forceinline__ bool goes_right(const float4 *cen, uint32_t p, const BNode &nd, const BigState &s,
                                           int32_t bins, const double *minDim) {
  NodeFrame f = node_frame(nd.lo, nd.hi, bins, minDim);
  float4 c = cen[p];
  float v = s.axis == 0 ? c.x : (s.axis == 1 ? c.y : c.z);
  return bin_of(v, pick3(f.nMin, s.axis), pick3(f.scale, s.axis), bins) > s.pos;
}

__global__ void __launch_bounds__(kBuildThreads)
k_classify(const Chunk *__restrict__ chunks, const BigState *__restrict__ big,
           const BNode *__restrict__ nodes, const uint32_t *__restrict__ prim,
           const float4 *__restrict__ cen, int32_t bins, double md0, double md1, double md2,
           uint8_t *__restrict__ isR, uint32_t *__restrict__ chunkR, const uint32_t *__restrict__ ncp) {
  if (blockIdx.x >= *ncp) return;
  const Chunk ch = chunks[blockIdx.x];
  const BigState s = big[ch.big];
  __shared__ uint32_t total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  if (!s.split) { if (threadIdx.x == 0) chunkR[blockIdx.x] = 0; return; }
  const BNode nd = nodes[s.node];
  const double minDim[3] = {md0, md1, md2};
  uint32_t local = 0;
  for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += blockDim.x) {
    bool r = goes_right(cen, prim[nd.leftFirst + e], nd, s, bins, minDim);
    isR[nd.leftFirst + e] = r ? 1 : 0;
    local += r ? 1u : 0u;
  }
  atomicAdd(&total, local);
  __syncthreads();
  if (threadIdx.x == 0) chunkR[blockIdx.x] = total;
}

// One thread per big node: exclusive prefix of right counts over its chunks.
__global__ void k_chunk_prefix(BigState *__restrict__ big, const uint32_t *__restrict__ nbp,
                               const BNode *__restrict__ nodes, const uint32_t *__restrict__ firstChunk,
                               const uint32_t *__restrict__ chunkR, uint32_t *__restrict__ chunkRBefore) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= *nbp) return;
  uint32_t c0 = firstChunk[b], c1 = firstChunk[b + 1];
  uint32_t acc = 0;
  for (uint32_t c = c0; c < c1; c++) { chunkRBefore[c] = acc; acc += chunkR[c]; }
  BigState s = big[b];
  s.totalR = acc;
  s.totalL = nodes[s.node].primCount - acc;
  s.P = 0;
  big[b] = s;
}

constexpr int kItems = kChunk / kBuildThreads;  // 8 elements per thread

// Ranks within the node; scatter x_k (right elements from the front) and
// z_k (left elements from the back). rank[] keeps each element's rank.
__global__ void __launch_bounds__(kBuildThreads)
k_xz(const Chunk *__restrict__ chunks, const BigState *__restrict__ big, const BNode *__restrict__ nodes,
     const uint8_t *__restrict__ isR, const uint32_t *__restrict__ chunkRBefore,
     uint32_t *__restrict__ X, uint32_t *__restrict__ Z, uint32_t *__restrict__ rank,
     const uint32_t *__restrict__ ncp) {
  typedef hipcub::BlockScan<uint32_t, kBuildThreads> Scan;
  __shared__ typename Scan::TempStorage tmp;
  if (blockIdx.x >= *ncp) return;
  const Chunk ch = chunks[blockIdx.x];
  const BigState s = big[ch.big];
  if (!s.split) return;
  const uint32_t first = nodes[s.node].leftFirst;
  uint32_t v[kItems];
  for (int i = 0; i < kItems; i++) {
    uint32_t e = ch.begin + threadIdx.x * kItems + i;
    v[i] = (e < ch.end) ? isR[first + e] : 0u;
  }
  uint32_t ex[kItems];
  Scan(tmp).ExclusiveSum(v, ex);
  for (int i = 0; i < kItems; i++) {
    uint32_t e = ch.begin + threadIdx.x * kItems + i;
    if (e >= ch.end) continue;
    uint32_t rR = chunkRBefore[blockIdx.x] + ex[i];  // right elements before e in the node
    if (v[i]) {
      X[first + rR] = e;
      rank[first + e] = rR;
    } else {
      uint32_t lBefore = e - rR;
      uint32_t rFromBack = s.totalL - lBefore - 1;
      Z[first + rFromBack] = e;
      rank[first + e] = rFromBack;
    }
  }
}

__global__ void __launch_bounds__(kBuildThreads)
k_count_p(const Chunk *__restrict__ chunks, BigState *__restrict__ big, const BNode *__restrict__ nodes,
          const uint8_t *__restrict__ isR, const uint32_t *__restrict__ Z, const uint32_t *__restrict__ rank,
          const uint32_t *__restrict__ ncp) {
  if (blockIdx.x >= *ncp) return;
  const Chunk ch = chunks[blockIdx.x];
  const BigState s = big[ch.big];
  if (!s.split) return;
  const uint32_t first = nodes[s.node].leftFirst;
  uint32_t local = 0;
  for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += blockDim.x) {
    if (!isR[first + e]) continue;
    uint32_t k = rank[first + e];
    if (k < s.totalL && e < Z[first + k]) local++;
  }
  for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(&big[ch.big].P, local);
}

// Destination of element e (relative) under the swap loop's permutation.
__device__ __forceinline__ uint32_t partition_dest(uint32_t e, bool right, uint32_t k, uint32_t P,
                                                   uint32_t n, uint32_t totalR, const uint32_t *X,
                                                   const uint32_t *Z) {
  if (!right) return (k < P) ? X[k] : e;  // k-th left from the back (0-based)
  // k-th right from the front (0-based); z_j = Z[j-1], z_0 = n.
  if (k < P) return (k == 0 ? n : Z[k - 1]) - 1;
  uint32_t zP = P == 0 ? n : Z[P - 1];
  if (k == P && P < totalR && X[P] < zP) return zP - 1;
  return e - 1;
}

__global__ void __launch_bounds__(kBuildThreads)
k_scatter(const Chunk *__restrict__ chunks, const BigState *__restrict__ big, const BNode *__restrict__ nodes,
          const uint8_t *__restrict__ isR, const uint32_t *__restrict__ X, const uint32_t *__restrict__ Z,
          const uint32_t *__restrict__ rank, const uint32_t *__restrict__ prim, uint32_t *__restrict__ tmp,
          const uint32_t *__restrict__ ncp) {
  if (blockIdx.x >= *ncp) return;
  const Chunk ch = chunks[blockIdx.x];
  const BigState s = big[ch.big];
  if (!s.split) return;
  const BNode nd = nodes[s.node];
  const uint32_t first = nd.leftFirst;
  for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += blockDim.x) {
    bool r = isR[first + e] != 0;
    uint32_t d = partition_dest(e, r, rank[first + e], s.P, nd.primCount, s.totalR, X + first, Z + first);
    tmp[first + d] = prim[first + e];
  }
}

__global__ void k_copyback(const Chunk *__restrict__ chunks, const BigState *__restrict__ big,
                           const BNode *__restrict__ nodes, const uint32_t *__restrict__ tmp,
                           uint32_t *__restrict__ prim, const uint32_t *__restrict__ ncp) {
  if (blockIdx.x >= *ncp) return;
  const Chunk ch = chunks[blockIdx.x];
  const BigState s = big[ch.big];
  if (!s.split) return;
  const uint32_t first = nodes[s.node].leftFirst;
  for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += blockDim.x) prim[first + e] = tmp[first + e];
}

// The level's big nodes: their records (for the chunk list) and fresh
// per-node states.
__global__ void k_level_nodes(const BNode *__restrict__ nodes, const uint32_t *__restrict__ ids,
                              const uint32_t *__restrict__ nbp, BNode *__restrict__ out, BigState *__restrict__ big) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= *nbp) return;
  out[b] = nodes[ids[b]];
  BigState st{};
  st.node = ids[b];
  big[b] = st;
}

// The level's chunk list on the device: the first chunk of each big node
// (exclusive scan of ceil(primCount / kChunk) in node order) and the count.
__global__ void __launch_bounds__(1024) k_chunk_first(const BNode *__restrict__ level,
                                                      const uint32_t *__restrict__ nbp,
                                                      uint32_t *__restrict__ first, uint32_t *__restrict__ ncp) {
  typedef hipcub::BlockScan<uint32_t, 1024> Scan;
  __shared__ typename Scan::TempStorage tmp;
  __shared__ uint32_t carry;
  const uint32_t nb = *nbp;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nb; base += 1024) {
    const uint32_t b = base + threadIdx.x;
    const uint32_t v = b < nb ? (level[b].primCount + kChunk - 1) / kChunk : 0u;
    uint32_t ex, agg;
    Scan(tmp).ExclusiveSum(v, ex, agg);
    if (b < nb) first[b] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += agg;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    first[nb] = carry;
    *ncp = carry;
  }
}
// Chunk c: kChunk elements of the node whose range of chunks holds c (every
// big node has at least one chunk, so the first-chunk list increases).
__global__ void k_chunk_fill(const BNode *__restrict__ level, const uint32_t *__restrict__ nbp,
                             const uint32_t *__restrict__ first, const uint32_t *__restrict__ ncp,
                             Chunk *__restrict__ chunks) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= *ncp) return;
  uint32_t lo = 0, hi = *nbp;  // first[lo] <= c < first[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (first[mid] <= c) lo = mid; else hi = mid;
  }
  const uint32_t e = (c - first[lo]) * (uint32_t)kChunk;
  chunks[c] = Chunk{lo, e, min(e + (uint32_t)kChunk, level[lo].primCount), 0u};
}

// Children of split big nodes; route them to the next big level or the
// small-subtree list.
__global__ void k_children(const BigState *__restrict__ big, const uint32_t *__restrict__ nbp,
                           BNode *__restrict__ nodes, uint32_t *__restrict__ counter, uint32_t *__restrict__ nextBig,
                           uint32_t *__restrict__ nextBigCount, uint32_t *__restrict__ small,
                           uint32_t *__restrict__ smallCount) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= *nbp) return;
  const BigState s = big[b];
  if (!s.split) return;
  BNode nd = nodes[s.node];
  uint32_t c = atomicAdd(counter, 2u);
  for (int side = 0; side < 2; side++) {
    BNode ch{};
    for (int k = 0; k < 3; k++) { ch.lo[k] = s.clo[side][k]; ch.hi[k] = s.chi[side][k]; }
    ch.leftFirst = side == 0 ? nd.leftFirst : nd.leftFirst + s.totalL;
    ch.primCount = side == 0 ? s.totalL : s.totalR;
    ch.depth = nd.depth + 1;
    ch.parent = (int32_t)s.node;
    nodes[c + side] = ch;
    if (ch.primCount > (uint32_t)kSmall) nextBig[atomicAdd(nextBigCount, 1u)] = c + side;
    else small[atomicAdd(smallCount, 1u)] = c + side;
  }
  nd.leftFirst = c;
  nd.primCount = 0;
  nodes[s.node] = nd;
}

// ------------------------------------------------------ small-subtree phase
constexpr int kSmallItems = kSmall / kBuildThreads;  // 4

// One subtree of at most kSmall primitives: its root node and the arrays of
// the build it belongs to (the subtrees of every mesh of a scene go out in
// one launch, so small meshes still fill the GPU).
struct SmallTask {
  BNode *nodes;
  uint32_t *prim;
  const float4 *plo, *phi, *cen;
  uint32_t *counter;
  double md[3];
  uint32_t root, pad;
};

// Nodes of at most kWaveNode primitives are split by one wave each, the
// primitives one per lane: no workgroup barriers, and the four waves of a
// workgroup work on four such subtrees at once.
constexpr int kWaveNode = 64;
constexpr int kBuildWaves = kBuildThreads / 64;

// Ordering point for LDS between the lanes of one wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct SmallBlockArrays {
  uint32_t ltmp[kSmall], X[kSmall], Z[kSmall], rk[kSmall];
  uint32_t kmin[3 * kMaxBins * 3], kmax[3 * kMaxBins * 3], cnt[3 * kMaxBins];
};
struct SmallWaveArrays {
  uint32_t kmin[3 * kMaxBins * 3], kmax[3 * kMaxBins * 3], cnt[3 * kMaxBins];
  uint32_t X[kWaveNode], Z[kWaveNode], tmp[kWaveNode];
  uint32_t stack[kWaveNode + 2];  // pending node ids and their records
  BNode snode[kWaveNode + 2];
};

// Cross-lane steps through DPP (gfx9 row shifts and row broadcasts, no LDS
// traffic); a lane without a source keeps `old`.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_u(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, kCtrl, kRowMask, 0xf, false);
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_f(float old, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                               __builtin_bit_cast(int, v), kCtrl, kRowMask, 0xf,
                                                               false));
}
// Inclusive prefix scans over segments of 32 (or 64) lanes: row_shr 1, 2, 4,
// 8 within rows of 16, then row 15's total into the next row (rows 1 and 3)
// and, for 64, lane 31's into rows 2 and 3. The min / max keep the earlier
// lane's value on ties.
template <bool kMax>
__device__ __forceinline__ float scan_minmax(float v, bool seg64) {
  const float nt = kMax ? -INFINITY : INFINITY;
  float e;
  e = dpp_f<0x111, 0xf>(nt, v); v = kMax ? (v > e ? v : e) : (v < e ? v : e);
  e = dpp_f<0x112, 0xf>(nt, v); v = kMax ? (v > e ? v : e) : (v < e ? v : e);
  e = dpp_f<0x114, 0xf>(nt, v); v = kMax ? (v > e ? v : e) : (v < e ? v : e);
  e = dpp_f<0x118, 0xf>(nt, v); v = kMax ? (v > e ? v : e) : (v < e ? v : e);
  e = dpp_f<0x142, 0xa>(nt, v); v = kMax ? (v > e ? v : e) : (v < e ? v : e);
  if (seg64) { e = dpp_f<0x143, 0xc>(nt, v); v = kMax ? (v > e ? v : e) : (v < e ? v : e); }
  return v;
}
__device__ __forceinline__ uint32_t scan_add(uint32_t v, bool seg64) {
  v += dpp_u<0x111, 0xf>(0u, v);
  v += dpp_u<0x112, 0xf>(0u, v);
  v += dpp_u<0x114, 0xf>(0u, v);
  v += dpp_u<0x118, 0xf>(0u, v);
  v += dpp_u<0x142, 0xa>(0u, v);
  if (seg64) v += dpp_u<0x143, 0xc>(0u, v);
  return v;
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// The sweep of sah_split for one wave: lane b holds bin b of an axis (with
// at most 32 bins, lanes 32..63 hold the next axis: two axes per pass, the
// scans segmented at 32 lanes); an inclusive prefix scan keeps the leftmost
// of equal minima / maxima (the left fold's strict comparisons); the right
// fold runs from the last bin down and keeps the rightmost, computed as the
// same prefix scan over the bins loaded in reverse order and moved to the
// lane of its split position; lane i then prices split position i with the
// sweep's f64 expressions, and the wave takes the lowest cost, ties to the
// earliest (axis, position). Child boxes are each side's fold over its bins
// in ascending order, as the sweep's epilogue does: a prefix scan over the
// side's bins (leftmost on ties, as the fold's strict comparison keeps the
// first) read at the side's last bin, the fold's start value kept when no
// bin beats it. Returns false for a leaf; every lane returns the same split.
__device__ __forceinline__ bool sah_split_wave(const uint32_t *kmin, const uint32_t *kmax, const uint32_t *cnt,
                                               int32_t bins, const NodeFrame &f, const double *minDim,
                                               uint32_t nodePrims, Split &out) {
  const int lane = threadIdx.x & 63;
  const bool halves = bins <= 32;
  const int b = halves ? (lane & 31) : lane, passes = halves ? 2 : 3;
  double bestCost = INFINITY;
  int bestIdx = 0x7fffffff;
  for (int pass = 0; pass < passes; pass++) {
    const int axis = halves ? 2 * pass + (lane >> 5) : pass;
    const int ax = axis < 3 ? axis : 2;
    const bool ok = axis < 3 && pick3(f.extent, ax) > pick3(minDim, ax);
    // Skip a pass whose axes are all degenerate (f is the same on every lane).
    const int axLo = halves ? 2 * pass : pass;
    const bool okLo = pick3(f.extent, axLo) > pick3(minDim, axLo);
    const bool okHi = halves && pass == 0 && f.extent[1] > minDim[1];
    if (!okLo && !okHi) continue;
    const bool inb = ok && b < bins;
    const int fb = ax * bins + (inb ? b : 0), rb = ax * bins + (inb ? bins - 1 - b : 0);
    const bool hasL = inb && cnt[fb] != 0, hasR = inb && cnt[rb] != 0;
    float lmn[3], lmx[3], rmn[3], rmx[3];
    for (int k = 0; k < 3; k++) {
      lmn[k] = hasL ? unkey(kmin[fb * 3 + k]) : INFINITY;
      lmx[k] = hasL ? unkey(kmax[fb * 3 + k]) : -INFINITY;
      rmn[k] = hasR ? unkey(kmin[rb * 3 + k]) : INFINITY;
      rmx[k] = hasR ? unkey(kmax[rb * 3 + k]) : -INFINITY;
    }
    const uint32_t lc = scan_add(inb ? cnt[fb] : 0u, !halves);
    const uint32_t rcr = scan_add(inb ? cnt[rb] : 0u, !halves);
    for (int k = 0; k < 3; k++) {
      lmn[k] = scan_minmax<false>(lmn[k], !halves);
      lmx[k] = scan_minmax<true>(lmx[k], !halves);
      rmn[k] = scan_minmax<false>(rmn[k], !halves);
      rmx[k] = scan_minmax<true>(rmx[k], !halves);
    }
    // right side of position i = the suffix from bin i + 1 = the reversed
    // scan at lane bins - 2 - i of the segment
    const int src = (halves ? (lane & 32) : 0) + (b < bins - 1 ? bins - 2 - b : 0);
    const uint32_t rightN = __shfl(rcr, src);
    float sr[6];
    for (int k = 0; k < 3; k++) { sr[k] = __shfl(rmn[k], src); sr[3 + k] = __shfl(rmx[k], src); }
    if (ok && b < bins - 1) {
      // the folds start from +-kFar: values beyond it leave the start value
      double L[6], R[6];
      for (int k = 0; k < 3; k++) {
        L[k] = (double)lmn[k] < kFar ? (double)lmn[k] : kFar;
        L[3 + k] = (double)lmx[k] > -kFar ? (double)lmx[k] : -kFar;
        R[k] = (double)sr[k] < kFar ? (double)sr[k] : kFar;
        R[3 + k] = (double)sr[3 + k] > -kFar ? (double)sr[3 + k] : -kFar;
      }
      const double lx = L[3] - L[0], ly = L[4] - L[1], lz = L[5] - L[2];
      const double rx = R[3] - R[0], ry = R[4] - R[1], rz = R[5] - R[2];
      const double lcost = lc == 0 ? kFar : (lx * ly + ly * lz + lz * lx) * (double)lc;
      const double rcost = rightN == 0 ? kFar : (rx * ry + ry * rz + rz * rx) * (double)rightN;
      const double cost = (lcost + rcost) + 0.0;
      if (cost < kFar && cost < bestCost) { bestCost = cost; bestIdx = axis * 64 + b; }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double oc = __shfl_xor(bestCost, off);
    const int oi = __shfl_xor(bestIdx, off);
    if (oc < bestCost || (oc == bestCost && oi < bestIdx)) { bestCost = oc; bestIdx = oi; }
  }
  const double splitCost = 1.0 + 1.0 * (bestIdx != 0x7fffffff ? bestCost : kFar) / f.area;
  if (!(splitCost < (double)nodePrims * 1.0)) return false;
  const int bestAxis = bestIdx >> 6, bestPos = bestIdx & 63;
  out.axis = bestAxis;
  out.pos = bestPos;
  const float farF = (float)kFar;
  const int bb = bestAxis * bins + (lane < bins ? lane : 0);
  const bool filled = lane < bins && cnt[bb] != 0;
  for (int sd = 0; sd < 2; sd++) {
    const bool mine = filled && (sd == 0 ? lane <= bestPos : lane > bestPos);
    const int at = sd == 0 ? bestPos : bins - 1;
    for (int k = 0; k < 3; k++) {
      const float vmn = readlane_f(scan_minmax<false>(mine ? unkey(kmin[bb * 3 + k]) : INFINITY, true), at);
      const float vmx = readlane_f(scan_minmax<true>(mine ? unkey(kmax[bb * 3 + k]) : -INFINITY, true), at);
      out.lo[sd][k] = (double)vmn < (double)farF ? vmn : farF;
      out.hi[sd][k] = (double)vmx > (double)-farF ? vmx : -farF;
    }
  }
  return true;
}

// One wave per big node of the level: its split from the level's bins (the
// wave sweep, with the sequential fold's choice and tie rules).
__global__ void __launch_bounds__(256) k_sweep(BigState *__restrict__ big, const uint32_t *__restrict__ nbp,
                                               const BNode *__restrict__ nodes, const uint32_t *__restrict__ gkmin,
                                               const uint32_t *__restrict__ gkmax, const uint32_t *__restrict__ gcnt,
                                               int32_t bins, double md0, double md1, double md2) {
  const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (b >= *nbp) return;  // the same for every lane of the wave
  BigState s = big[b];
  const BNode nd = nodes[s.node];
  const double minDim[3] = {md0, md1, md2};
  const NodeFrame f = node_frame(nd.lo, nd.hi, bins, minDim);
  s.split = 0;
  s.totalL = s.totalR = s.P = 0;
  if (f.area != 0.0 && nd.primCount != 1) {
    Split sp;
    const uint32_t base = b * 3u * (uint32_t)bins;
    if (sah_split_wave(gkmin + base * 3, gkmax + base * 3, gcnt + base, bins, f, minDim, nd.primCount, sp)) {
      s.split = 1;
      s.axis = sp.axis;
      s.pos = sp.pos;
      for (int c = 0; c < 2; c++)
        for (int k = 0; k < 3; k++) { s.clo[c][k] = sp.lo[c][k]; s.chi[c][k] = sp.hi[c][k]; }
    }
  }
  if ((threadIdx.x & 63) == 0) big[b] = s;
}

// Writes the two children of `parent` (split at totalL), returns the first
// child's index and, when `out` is given, the two records; one lane only.
__device__ uint32_t emit_children(BNode *__restrict__ nodes, uint32_t *__restrict__ counter, uint32_t parentId,
                                  const BNode &parent, const Split &split, uint32_t totalL, uint32_t totalR,
                                  BNode *out = nullptr) {
  const uint32_t c = atomicAdd(counter, 2u);
  for (int side = 0; side < 2; side++) {
    BNode ch{};
    for (int k = 0; k < 3; k++) { ch.lo[k] = split.lo[side][k]; ch.hi[k] = split.hi[side][k]; }
    ch.leftFirst = side == 0 ? parent.leftFirst : parent.leftFirst + totalL;
    ch.primCount = side == 0 ? totalL : totalR;
    ch.depth = parent.depth + 1;
    ch.parent = (int32_t)parentId;
    nodes[c + side] = ch;
    if (out) out[side] = ch;
  }
  BNode nd = parent;
  nd.leftFirst = c;
  nd.primCount = 0;
  nodes[parentId] = nd;
  return c;
}

__global__ void __launch_bounds__(kBuildThreads) k_small_build(const SmallTask *__restrict__ tasks, int32_t bins) {
  typedef hipcub::BlockScan<uint32_t, kBuildThreads> Scan;
  __shared__ typename Scan::TempStorage scanTmp;
  __shared__ uint32_t lp[kSmall];
  __shared__ uint8_t rflag[kSmall];
  __shared__ union {
    SmallBlockArrays a;
    SmallWaveArrays w[kBuildWaves];
  } sh;
  __shared__ uint32_t stack[kSmall + 1];   // pending nodes above kWaveNode primitives
  __shared__ uint32_t queue[kSmall];       // pending nodes of at most kWaveNode primitives
  __shared__ int32_t sp;
  __shared__ uint32_t qCount, qHead;
  __shared__ int32_t popped;
  __shared__ uint32_t curNode, P;
  __shared__ BNode cur;
  __shared__ Split split;
  __shared__ int32_t doSplit;
  __shared__ uint32_t totalR;

  const SmallTask &tk = tasks[blockIdx.x];
  BNode *__restrict__ nodes = tk.nodes;
  uint32_t *__restrict__ prim = tk.prim;
  const float4 *__restrict__ plo = tk.plo;
  const float4 *__restrict__ phi = tk.phi;
  const float4 *__restrict__ cen = tk.cen;
  uint32_t *__restrict__ counter = tk.counter;
  const double minDim[3] = {tk.md[0], tk.md[1], tk.md[2]};
  const uint32_t root = tk.root;
  const BNode rn = nodes[root];
  const uint32_t base = rn.leftFirst, total = rn.primCount;
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) lp[e] = prim[base + e];
  if (threadIdx.x == 0) {
    sp = 0;
    qCount = 0;
    qHead = 0;
    if (total > (uint32_t)kWaveNode) stack[sp++] = root;
    else queue[qCount++] = root;
  }
  __syncthreads();
  uint32_t *const kmin = sh.a.kmin, *const kmax = sh.a.kmax, *const cnt = sh.a.cnt;
  uint32_t *const ltmp = sh.a.ltmp, *const X = sh.a.X, *const Z = sh.a.Z, *const rk = sh.a.rk;
  // Phase 1: nodes above kWaveNode primitives, split by the whole workgroup.
  while (true) {
    __syncthreads();  // every lane is done with the previous node's shared state
    if (threadIdx.x == 0) {
      popped = sp > 0;
      if (popped) {
        curNode = stack[--sp];
        cur = nodes[curNode];
      }
    }
    __syncthreads();
    if (!popped) break;
    const uint32_t n = cur.primCount, first = cur.leftFirst - base;
    NodeFrame f = node_frame(cur.lo, cur.hi, bins, minDim);
    bool leaf = (f.area == 0.0) || (n == 1);
    if (!leaf) {
      for (int i = threadIdx.x; i < 3 * bins * 3; i += blockDim.x) { kmin[i] = 0xffffffffu; kmax[i] = 0u; }
      for (int i = threadIdx.x; i < 3 * bins; i += blockDim.x) cnt[i] = 0;
      __syncthreads();
      for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
        uint32_t p = lp[first + e];
        float4 c = cen[p], a = plo[p], b = phi[p];
        float cc[3] = {c.x, c.y, c.z}, mn[3] = {a.x, a.y, a.z}, mx[3] = {b.x, b.y, b.z};
        for (int ax = 0; ax < 3; ax++) {
          int bin = ax * bins + bin_of(cc[ax], f.nMin[ax], f.scale[ax], bins);
          atomicAdd(&cnt[bin], 1u);
          for (int k = 0; k < 3; k++) {
            atomicMin(&kmin[bin * 3 + k], fkey(mn[k]));
            atomicMax(&kmax[bin * 3 + k], fkey(mx[k]));
          }
        }
      }
      __syncthreads();
      // The split: the wave sweep (the sequential fold's choice and tie
      // rules) on wave 0, broadcast through shared memory.
      if (threadIdx.x < 64) {
        Split s1;
        const bool ok = sah_split_wave(kmin, kmax, cnt, bins, f, minDim, n, s1);
        if (threadIdx.x == 0) {
          doSplit = ok ? 1 : 0;
          if (ok) split = s1;
        }
      }
      __syncthreads();
      leaf = doSplit == 0;
    }
    if (!leaf) {
      // Classify and rank (elements held blocked: thread t owns 4t..4t+3).
      uint32_t v[kSmallItems], ex[kSmallItems];
      for (int i = 0; i < kSmallItems; i++) {
        uint32_t e = threadIdx.x * kSmallItems + i;
        bool r = false;
        if (e < n) {
          float4 c = cen[lp[first + e]];
          float val = split.axis == 0 ? c.x : (split.axis == 1 ? c.y : c.z);
          r = bin_of(val, pick3(f.nMin, split.axis), pick3(f.scale, split.axis), bins) > split.pos;
          rflag[e] = r ? 1 : 0;
        }
        v[i] = r ? 1u : 0u;
      }
      uint32_t aggregate;
      Scan(scanTmp).ExclusiveSum(v, ex, aggregate);
      if (threadIdx.x == 0) { totalR = aggregate; P = 0; }
      __syncthreads();
      const uint32_t totalL = n - totalR;
      for (int i = 0; i < kSmallItems; i++) {
        uint32_t e = threadIdx.x * kSmallItems + i;
        if (e >= n) continue;
        if (v[i]) { X[ex[i]] = e; rk[e] = ex[i]; }
        else { uint32_t k = totalL - (e - ex[i]) - 1; Z[k] = e; rk[e] = k; }
      }
      __syncthreads();
      uint32_t local = 0;
      for (uint32_t e = threadIdx.x; e < n; e += blockDim.x)
        if (rflag[e] && rk[e] < totalL && e < Z[rk[e]]) local++;
      if (local) atomicAdd(&P, local);
      __syncthreads();
      for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
        uint32_t d = partition_dest(e, rflag[e] != 0, rk[e], P, n, totalR, X, Z);
        ltmp[d] = lp[first + e];
      }
      __syncthreads();
      for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) lp[first + e] = ltmp[e];
      if (threadIdx.x == 0) {
        const uint32_t c = emit_children(nodes, counter, curNode, cur, split, totalL, totalR);
        const uint32_t count[2] = {totalL, totalR};
        for (int side = 1; side >= 0; side--) {
          if (count[side] > (uint32_t)kWaveNode) stack[sp++] = c + side;
          else queue[qCount++] = c + side;
        }
      }
    }
  }
  // Phase 2: each wave takes queued nodes and splits them and their
  // descendants on its own.
  __syncthreads();
  __threadfence_block();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  SmallWaveArrays &W = sh.w[wv];
  while (true) {
    uint32_t item = 0;
    if (lane == 0) item = atomicAdd(&qHead, 1u);
    item = __shfl(item, 0);
    if (item >= qCount) break;
    if (lane == 0) {
      W.stack[0] = queue[item];
      W.snode[0] = nodes[queue[item]];
    }
    int wsp = 1;
    wave_sync();
    while (wsp > 0) {
      --wsp;
      const uint32_t nodeId = W.stack[wsp];
      const BNode nd = W.snode[wsp];
      const uint32_t n = nd.primCount, first = nd.leftFirst - base;
      const NodeFrame f = node_frame(nd.lo, nd.hi, bins, minDim);
      if ((f.area == 0.0) || (n == 1)) continue;
      for (int i = lane; i < 3 * bins * 3; i += 64) { W.kmin[i] = 0xffffffffu; W.kmax[i] = 0u; }
      for (int i = lane; i < 3 * bins; i += 64) W.cnt[i] = 0;
      wave_sync();
      const bool live = (uint32_t)lane < n;
      const uint32_t p = live ? lp[first + lane] : 0u;
      float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (live) {
        c = cen[p];
        const float4 a = plo[p], b = phi[p];
        const float cc[3] = {c.x, c.y, c.z}, mn[3] = {a.x, a.y, a.z}, mx[3] = {b.x, b.y, b.z};
        for (int ax = 0; ax < 3; ax++) {
          const int bin = ax * bins + bin_of(cc[ax], f.nMin[ax], f.scale[ax], bins);
          atomicAdd(&W.cnt[bin], 1u);
          for (int k = 0; k < 3; k++) {
            atomicMin(&W.kmin[bin * 3 + k], fkey(mn[k]));
            atomicMax(&W.kmax[bin * 3 + k], fkey(mx[k]));
          }
        }
      }
      wave_sync();
      Split sp2;
      if (!sah_split_wave(W.kmin, W.kmax, W.cnt, bins, f, minDim, n, sp2)) continue;
      // The swap loop's permutation (see k_small_build's workgroup form) from
      // wave ballots: X[k] = k-th right element from the front, Z[k] = k-th
      // left element from the back.
      bool r = false;
      if (live) {
        const float val = sp2.axis == 0 ? c.x : (sp2.axis == 1 ? c.y : c.z);
        r = bin_of(val, pick3(f.nMin, sp2.axis), pick3(f.scale, sp2.axis), bins) > sp2.pos;
      }
      const unsigned long long rmask = __ballot(r);
      const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
      const uint32_t ex = (uint32_t)__popcll(rmask & below), nR = (uint32_t)__popcll(rmask);
      const uint32_t nL = n - nR;
      uint32_t rank = 0;
      if (live) {
        if (r) { W.X[ex] = lane; rank = ex; }
        else { rank = nL - ((uint32_t)lane - ex) - 1; W.Z[rank] = lane; }
      }
      wave_sync();
      const uint32_t Pw = (uint32_t)__popcll(__ballot(live && r && rank < nL && (uint32_t)lane < W.Z[rank]));
      if (live) W.tmp[partition_dest(lane, r, rank, Pw, n, nR, W.X, W.Z)] = p;
      wave_sync();
      if (live) lp[first + lane] = W.tmp[lane];
      if (lane == 0) {
        BNode ch[2];
        const uint32_t c0 = emit_children(nodes, counter, nodeId, nd, sp2, nL, nR, ch);
        W.stack[wsp] = c0 + 1;
        W.snode[wsp] = ch[1];
        W.stack[wsp + 1] = c0;
        W.snode[wsp + 1] = ch[0];
      }
      wsp += 2;
      wave_sync();
    }
  }
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) prim[base + e] = lp[e];
}

// --------------------------------------------------------------- layout
// Node counts per depth: a block counts its grid-stride share in LDS and
// adds each nonzero count to the global histogram once.
constexpr uint32_t kDepthBins = 256;
__global__ void k_depth_hist(const BNode *__restrict__ nodes, uint32_t n, uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[kDepthBins];
  for (uint32_t d = threadIdx.x; d < kDepthBins; d += blockDim.x) h[d] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&h[min(nodes[i].depth, kDepthBins - 1)], 1u);
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kDepthBins; d += blockDim.x)
    if (h[d]) atomicAdd(&hist[d], h[d]);
}

// Node ids grouped by depth (any order within a depth): a block counts its
// share per depth in LDS, reserves one range per depth with a single global
// atomic, then writes its nodes into the ranges.
__global__ void k_depth_scatter(const BNode *__restrict__ nodes, uint32_t n, uint32_t *__restrict__ cursor,
                                uint32_t *__restrict__ byDepth) {
  __shared__ uint32_t cnt[kDepthBins], base[kDepthBins];
  for (uint32_t d = threadIdx.x; d < kDepthBins; d += blockDim.x) cnt[d] = 0;
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    atomicAdd(&cnt[min(nodes[i].depth, kDepthBins - 1)], 1u);
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kDepthBins; d += blockDim.x) {
    base[d] = cnt[d] ? atomicAdd(&cursor[d], cnt[d]) : 0u;
    cnt[d] = 0;
  }
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t d = min(nodes[i].depth, kDepthBins - 1);
    byDepth[base[d] + atomicAdd(&cnt[d], 1u)] = i;
  }
}

// Interior-node subtree sizes, one depth level at a time (deepest first).
__global__ void k_subtree(const BNode *__restrict__ nodes, const uint32_t *__restrict__ level, uint32_t cnt,
                          uint32_t *__restrict__ size) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  uint32_t v = level[i];
  const BNode nd = nodes[v];
  if (nd.primCount) { size[v] = 0; return; }
  size[v] = 1 + size[nd.leftFirst] + size[nd.leftFirst + 1];
}

// Depth-first index of interior nodes, top-down: left child follows its
// parent, the right child follows the left subtree.
__global__ void k_dfs(const BNode *__restrict__ nodes, const uint32_t *__restrict__ level, uint32_t cnt,
                      const uint32_t *__restrict__ size, uint32_t *__restrict__ dfs) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  uint32_t v = level[i];
  const BNode nd = nodes[v];
  if (nd.primCount) return;
  uint32_t me = dfs[v];
  dfs[nd.leftFirst] = me + 1;
  dfs[nd.leftFirst + 1] = me + 1 + size[nd.leftFirst];
}

__device__ __forceinline__ int32_t child_ref(const BNode &c, const uint32_t *dfs, uint32_t id) {
  return c.primCount ? (int32_t)(0x80000000u | c.leftFirst) : (int32_t)dfs[id];
}

__global__ void k_emit_nodes(const BNode *__restrict__ nodes, uint32_t n, const uint32_t *__restrict__ dfs,
                             TNode *__restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BNode nd = nodes[i];
  if (nd.primCount) return;
  const BNode L = nodes[nd.leftFirst], R = nodes[nd.leftFirst + 1];
  TNode t;
  t.a = make_float4(L.lo[0], L.lo[1], L.lo[2], intf(child_ref(L, dfs, nd.leftFirst)));
  t.b = make_float4(L.hi[0], L.hi[1], L.hi[2], intf(child_ref(R, dfs, nd.leftFirst + 1)));
  t.c = make_float4(R.lo[0], R.lo[1], R.lo[2], 0.0f);
  t.d = make_float4(R.hi[0], R.hi[1], R.hi[2], 0.0f);
  out[dfs[i]] = t;
}

__global__ void k_leaf_counts(const BNode *__restrict__ nodes, uint32_t n, uint32_t *__restrict__ leafCount) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BNode nd = nodes[i];
  if (!nd.primCount) return;
  for (uint32_t k = 0; k < nd.primCount; k++) leafCount[nd.leftFirst + k] = nd.primCount;
}

// Largest leaf primitive count of a tree, max-reduced into out[0], which must
// hold no more than it: the caller passes the leaf-count array, whose first
// entry is the count of slot 0's leaf (k_emit_tris has read it by then).
__global__ void k_max_leaf(const BNode *__restrict__ nodes, uint32_t n, uint32_t *__restrict__ out) {
  uint32_t v = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v = max(v, nodes[i].primCount);
  for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0 && v) atomicMax(out, v);
}

__global__ void k_emit_tris(const uint32_t *__restrict__ prim, const uint32_t *__restrict__ source,
                            const uint32_t *__restrict__ leafCount, uint32_t n, const float *__restrict__ pos,
                            const uint32_t *__restrict__ idx, TriRec *__restrict__ out,
                            uint32_t *__restrict__ srcOut) {
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  uint32_t t = source[prim[q]];
  uint32_t i0 = idx[t * 3], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
  TriRec r;
  r.v0 = make_float4(pos[(size_t)i0 * 3], pos[(size_t)i0 * 3 + 1], pos[(size_t)i0 * 3 + 2], intf((int32_t)t));
  r.v1 = make_float4(pos[(size_t)i1 * 3], pos[(size_t)i1 * 3 + 1], pos[(size_t)i1 * 3 + 2], intf((int32_t)leafCount[q]));
  r.v2 = make_float4(pos[(size_t)i2 * 3], pos[(size_t)i2 * 3 + 1], pos[(size_t)i2 * 3 + 2], 0.0f);
  out[q] = r;
  srcOut[q] = t;
}

// --------------------------------------------------------- BLAS refit kernels
// refitBLAS (bvhAccel.ts:800-870): node boxes recomputed bottom-up from new
// vertex positions, topology kept. A leaf's box spans the finite vertices of
// its triangles, an interior node's the union of its children's; min/max are
// exact, so the boxes do not depend on the visiting order.
__global__ void k_refit_level(BNode *__restrict__ nodes, const uint32_t *__restrict__ level, uint32_t cnt,
                              const uint32_t *__restrict__ prim, const uint32_t *__restrict__ source,
                              const float *__restrict__ pos, const uint32_t *__restrict__ idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t v = level[i];
  const BNode nd = nodes[v];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (nd.primCount) {
    for (uint32_t q = nd.leftFirst; q < nd.leftFirst + nd.primCount; q++) {
      const uint32_t t = source[prim[q]];
      for (int c = 0; c < 3; c++) {
        const float *p = pos + (size_t)idx[(size_t)t * 3 + c] * 3;
        const float x = p[0], y = p[1], z = p[2];
        if (!isfinite(x) || !isfinite(y) || !isfinite(z)) continue;
        lo[0] = jminf(lo[0], x); lo[1] = jminf(lo[1], y); lo[2] = jminf(lo[2], z);
        hi[0] = jmaxf(hi[0], x); hi[1] = jmaxf(hi[1], y); hi[2] = jmaxf(hi[2], z);
      }
    }
  } else {
    const BNode L = nodes[nd.leftFirst], R = nodes[nd.leftFirst + 1];
    for (int k = 0; k < 3; k++) { lo[k] = jminf(L.lo[k], R.lo[k]); hi[k] = jmaxf(L.hi[k], R.hi[k]); }
  }
  for (int k = 0; k < 3; k++) { nodes[v].lo[k] = lo[k]; nodes[v].hi[k] = hi[k]; }
}

// Sum over nodes of 2 (xz + xy + zy): the surface-area measure refitBLAS
// compares with the build's to advise a rebuild.
__global__ void k_node_area(const BNode *__restrict__ nodes, uint32_t n, double *__restrict__ area) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0;
  if (i < n) {
    const BNode nd = nodes[i];
    const double x = (double)nd.hi[0] - nd.lo[0], y = (double)nd.hi[1] - nd.lo[1], z = (double)nd.hi[2] - nd.lo[2];
    a = 2.0 * (x * z + x * y + z * y);
    if (!(a == a) || isinf(a)) a = 0.0;
  }
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o);
  if ((threadIdx.x & 63u) == 0u && a != 0.0) atomicAdd(area, a);
}

// --------------------------------------------------------- TLAS refit kernels
__global__ void k_instance_update(const double *__restrict__ transforms, const float *__restrict__ blasBoxes,
                                  const uint32_t *__restrict__ dirty, uint32_t nDirty,
                                  const uint32_t *__restrict__ slotOfSource, InstSlot *__restrict__ slots) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nDirty) return;
  uint32_t src = dirty[i];
  const double *m = transforms + (size_t)src * 16;
  double inv[16];
  if (!invert4(m, inv)) for (int k = 0; k < 16; k++) inv[k] = 0.0;
  float box[6];
  world_bounds(m, blasBoxes + (size_t)src * 6, blasBoxes + (size_t)src * 6 + 3, box);
  InstSlot &s = slots[slotOfSource[src]];
  s.c0 = make_float4((float)inv[0], (float)inv[1], (float)inv[2], s.c0.w);
  s.c1 = make_float4((float)inv[4], (float)inv[5], (float)inv[6], s.c1.w);
  s.c2 = make_float4((float)inv[8], (float)inv[9], (float)inv[10], s.c2.w);
  s.c3 = make_float4((float)inv[12], (float)inv[13], (float)inv[14], s.c3.w);
  s.bmin = make_float4(box[0], box[1], box[2], s.bmin.w);
  s.bmax = make_float4(box[3], box[4], box[5], s.bmax.w);
}

__device__ __forceinline__ void child_box(int32_t ref, const TNode *nodes, const InstSlot *slots, float lo[3],
                                          float hi[3]) {
  for (int k = 0; k < 3; k++) { lo[k] = INFINITY; hi[k] = -INFINITY; }
  if (ref >= 0) {
    const TNode n = nodes[ref];
    float a[3] = {n.a.x, n.a.y, n.a.z}, b[3] = {n.b.x, n.b.y, n.b.z};
    float c[3] = {n.c.x, n.c.y, n.c.z}, d[3] = {n.d.x, n.d.y, n.d.z};
    for (int k = 0; k < 3; k++) { lo[k] = jminf(a[k], c[k]); hi[k] = jmaxf(b[k], d[k]); }
    return;
  }
  int32_t first = ref & 0x7fffffff;
  int32_t cnt = fint(slots[first].c1.w);
  for (int32_t q = first; q < first + cnt; q++) {
    const InstSlot &s = slots[q];
    float a[3] = {s.bmin.x, s.bmin.y, s.bmin.z}, b[3] = {s.bmax.x, s.bmax.y, s.bmax.z};
    for (int k = 0; k < 3; k++) { lo[k] = jminf(lo[k], a[k]); hi[k] = jmaxf(hi[k], b[k]); }
  }
}

// Bottom-up refit of the TLAS child boxes (refitTLAS refitNode,
// bvhAccel.ts:658-687), one workgroup walking the levels deepest first.
__global__ void k_tlas_refit(TNode *__restrict__ nodes, const InstSlot *__restrict__ slots,
                             const uint32_t *__restrict__ levelNodes, const uint32_t *__restrict__ levelStart,
                             uint32_t nLevels, AccelView *__restrict__ view) {
  for (uint32_t L = 0; L < nLevels; L++) {
    for (uint32_t i = levelStart[L] + threadIdx.x; i < levelStart[L + 1]; i += blockDim.x) {
      uint32_t v = levelNodes[i];
      TNode n = nodes[v];
      float lo[3], hi[3];
      child_box(fint(n.a.w), nodes, slots, lo, hi);
      n.a = make_float4(lo[0], lo[1], lo[2], n.a.w);
      n.b = make_float4(hi[0], hi[1], hi[2], n.b.w);
      child_box(fint(n.b.w), nodes, slots, lo, hi);
      n.c = make_float4(lo[0], lo[1], lo[2], n.c.w);
      n.d = make_float4(hi[0], hi[1], hi[2], n.d.w);
      nodes[v] = n;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float lo[3], hi[3];
    child_box(view->tlasRootRef, nodes, slots, lo, hi);
    for (int k = 0; k < 3; k++) { view->rootMin[k] = lo[k]; view->rootMax[k] = hi[k]; }
  }
}

// ------------------------------------------------------------- host side
static inline uint32_t blocks(size_t n, int b = 256) { return (uint32_t)((n + b - 1) / b); }

// Builds the tree over T.primCount boxes in T.plo/T.phi with centers T.cen.
// Splits every node above kSmall primitives level by level and appends the
// remaining subtree roots to `small` (launched by run_small_builds).
static void build_topology(Topology &T, uint32_t bins, hipStream_t s, std::vector<SmallTask> &small) {
  const uint32_t n = T.primCount;
  MTX_REQUIRE(bins >= 2 && bins <= (uint32_t)kMaxBins, MTX_INVALID_ARGUMENT, "bins must be in [2, 64]");
  T.nodes.alloc(std::max<uint32_t>(1, 2 * n));
  T.prim.alloc(std::max<uint32_t>(1, n));
  T.nodeCount = 0;
  if (n == 0) return;
  T.counter.alloc(1);
  DevBuf<uint32_t> &counter = T.counter;
  DevBuf<unsigned long long> keys(6);
  unsigned long long init[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
  keys.upload(init, 6, s);
  hipLaunchKernelGGL(k_root_reduce, dim3(std::min<uint32_t>(blocks(n), 2048)), dim3(256), 0, s, T.plo.ptr, T.phi.ptr,
                     n, keys.ptr, T.prim.ptr);
  hipLaunchKernelGGL(k_root_init, dim3(1), dim3(64), 0, s, keys.ptr, T.plo.ptr, T.phi.ptr, n, T.nodes.ptr, counter.ptr);
  BNode root;
  T.nodes.download(&root, 1, s);
  MTX_HIP(hipStreamSynchronize(s));
  const double minDim[3] = {((double)root.hi[0] - (double)root.lo[0]) * 1e-20,
                            ((double)root.hi[1] - (double)root.lo[1]) * 1e-20,
                            ((double)root.hi[2] - (double)root.lo[2]) * 1e-20};
  std::vector<uint32_t> smallIds;
  const uint32_t nb0 = n > (uint32_t)kSmall ? 1u : 0u;
  if (!nb0) smallIds.push_back(0);
  // Level buffers, sized once: a big node holds more than kSmall primitives,
  // so a level has at most n / kSmall of them and n / kChunk + nb chunks;
  // the small-subtree roots are disjoint, at most n.
  const uint32_t maxBig = n / (uint32_t)kSmall + 1u, maxChunks = n / (uint32_t)kChunk + maxBig + 1u;
  if (nb0) {
    DevBuf<uint8_t> isR(n);
    DevBuf<uint32_t> X(n), Z(n), rank(n), tmp(n), dNext(2 * maxBig + 4), dSmall(n + 2), dNb(2), dNc(1),
        dSmallCount(1), dFirst(maxBig + 1), chunkR(maxChunks), chunkRB(maxChunks);
    DevBuf<uint32_t> gkmin((size_t)maxBig * 3 * bins * 3), gkmax((size_t)maxBig * 3 * bins * 3),
        gcnt((size_t)maxBig * 3 * bins);
    DevBuf<BNode> dLevel(maxBig);
    DevBuf<BigState> dBig(maxBig);
    DevBuf<Chunk> dChunks(maxChunks);
    const uint32_t init[4] = {0u, 1u, 0u, 0u};  // dNext[0] = root; level counts {1, 0}; no small roots yet
    dNext.upload(init, 1, s);
    dNb.upload(init + 1, 2, s);
    dSmallCount.upload(init + 2, 1, s);
    // Levels run back to back with their counts on the device: every launch
    // is sized for the level's bound and its threads beyond the level's
    // nodes or chunks return at once; the host reads the next level's size
    // once per kLevelsPerCheck levels (a finished build runs empty levels).
    constexpr int kLevelsPerCheck = 4;
    const uint32_t gNodes = blocks(maxBig, 64), gChunks = maxChunks;
    for (uint32_t level = 0;;) {
      for (int k = 0; k < kLevelsPerCheck; k++, level++) {
        const uint32_t *nb = dNb.ptr + (level & 1u);
        uint32_t *nbNext = dNb.ptr + ((level + 1u) & 1u);
        hipLaunchKernelGGL(k_level_nodes, dim3(gNodes), dim3(64), 0, s, T.nodes.ptr, dNext.ptr, nb, dLevel.ptr,
                           dBig.ptr);
        hipLaunchKernelGGL(k_chunk_first, dim3(1), dim3(1024), 0, s, dLevel.ptr, nb, dFirst.ptr, dNc.ptr);
        hipLaunchKernelGGL(k_chunk_fill, dim3(blocks(maxChunks)), dim3(256), 0, s, dLevel.ptr, nb, dFirst.ptr,
                           dNc.ptr, dChunks.ptr);
        hipLaunchKernelGGL(k_bins_clear, dim3(std::min<uint32_t>(blocks((size_t)maxBig * 9 * bins), 1024)),
                           dim3(256), 0, s, gkmin.ptr, gkmax.ptr, gcnt.ptr, nb, (int32_t)bins);
        hipLaunchKernelGGL(k_bin_chunks, dim3(gChunks), dim3(kBuildThreads), 0, s, dChunks.ptr, dBig.ptr,
                           T.nodes.ptr, T.prim.ptr, T.plo.ptr, T.phi.ptr, T.cen.ptr, (int32_t)bins, minDim[0],
                           minDim[1], minDim[2], gkmin.ptr, gkmax.ptr, gcnt.ptr, dNc.ptr);
        hipLaunchKernelGGL(k_sweep, dim3(blocks((size_t)maxBig * 64)), dim3(256), 0, s, dBig.ptr, nb, T.nodes.ptr,
                           gkmin.ptr, gkmax.ptr, gcnt.ptr, (int32_t)bins, minDim[0], minDim[1], minDim[2]);
        hipLaunchKernelGGL(k_classify, dim3(gChunks), dim3(kBuildThreads), 0, s, dChunks.ptr, dBig.ptr,
                           T.nodes.ptr, T.prim.ptr, T.cen.ptr, (int32_t)bins, minDim[0], minDim[1], minDim[2],
                           isR.ptr, chunkR.ptr, dNc.ptr);
        hipLaunchKernelGGL(k_chunk_prefix, dim3(gNodes), dim3(64), 0, s, dBig.ptr, nb, T.nodes.ptr, dFirst.ptr,
                           chunkR.ptr, chunkRB.ptr);
        hipLaunchKernelGGL(k_xz, dim3(gChunks), dim3(kBuildThreads), 0, s, dChunks.ptr, dBig.ptr, T.nodes.ptr,
                           isR.ptr, chunkRB.ptr, X.ptr, Z.ptr, rank.ptr, dNc.ptr);
        hipLaunchKernelGGL(k_count_p, dim3(gChunks), dim3(kBuildThreads), 0, s, dChunks.ptr, dBig.ptr, T.nodes.ptr,
                           isR.ptr, Z.ptr, rank.ptr, dNc.ptr);
        hipLaunchKernelGGL(k_scatter, dim3(gChunks), dim3(kBuildThreads), 0, s, dChunks.ptr, dBig.ptr, T.nodes.ptr,
                           isR.ptr, X.ptr, Z.ptr, rank.ptr, T.prim.ptr, tmp.ptr, dNc.ptr);
        hipLaunchKernelGGL(k_copyback, dim3(gChunks), dim3(kBuildThreads), 0, s, dChunks.ptr, dBig.ptr,
                           T.nodes.ptr, tmp.ptr, T.prim.ptr, dNc.ptr);
        MTX_HIP(hipMemsetAsync(nbNext, 0, sizeof(uint32_t), s));
        hipLaunchKernelGGL(k_children, dim3(gNodes), dim3(64), 0, s, dBig.ptr, nb, T.nodes.ptr, counter.ptr,
                           dNext.ptr, nbNext, dSmall.ptr, dSmallCount.ptr);
        MTX_HIP(hipGetLastError());
      }
      uint32_t next = 0;
      MTX_HIP(hipMemcpyAsync(&next, dNb.ptr + (level & 1u), sizeof next, hipMemcpyDeviceToHost, s));
      MTX_HIP(hipStreamSynchronize(s));
      MTX_REQUIRE(next <= maxBig, MTX_ERROR, "BVH builder: big-node level larger than its bound");
      if (!next) break;
    }
    uint32_t nSmall = 0;
    dSmallCount.download(&nSmall, 1, s);
    MTX_HIP(hipStreamSynchronize(s));
    MTX_REQUIRE(nSmall <= n, MTX_ERROR, "BVH builder: small-subtree list larger than its bound");
    smallIds.resize(nSmall);
    dSmall.download(smallIds.data(), nSmall, s);
    MTX_HIP(hipStreamSynchronize(s));  // the level buffers are freed on return
  }
  for (uint32_t r : smallIds)
    small.push_back(SmallTask{T.nodes.ptr, T.prim.ptr, T.plo.ptr, T.phi.ptr, T.cen.ptr, counter.ptr,
                              {minDim[0], minDim[1], minDim[2]}, r, 0u});
}

// Builds every pending small subtree (one workgroup each) in one launch.
static void run_small_builds(const std::vector<SmallTask> &small, uint32_t bins, hipStream_t s) {
  if (small.empty()) return;
  DevBuf<SmallTask> dTasks;
  dTasks.upload(small.data(), small.size(), s);
  hipLaunchKernelGGL(k_small_build, dim3((uint32_t)small.size()), dim3(kBuildThreads), 0, s, dTasks.ptr,
                     (int32_t)bins);
  MTX_HIP(hipGetLastError());
  MTX_HIP(hipStreamSynchronize(s));
}

// Node count once the small subtrees are built.
static void finish_topology(Topology &T, hipStream_t s) {
  if (T.primCount == 0) return;
  uint32_t nodeCount = 0;
  T.counter.download(&nodeCount, 1, s);
  MTX_HIP(hipStreamSynchronize(s));
  T.nodeCount = nodeCount;
}

static void layout_topology(Topology &T, Layout &L, hipStream_t s) {
  const uint32_t nn = T.nodeCount;
  if (!nn) return;
  constexpr uint32_t kLevels = 256;
  DevBuf<uint32_t> hist(kLevels), cursor(kLevels), byDepth(nn), size(nn), dfs(nn);
  std::vector<uint32_t> zeros(kLevels, 0);
  hist.upload(zeros.data(), kLevels, s);
  hipLaunchKernelGGL(k_depth_hist, dim3(std::min<uint32_t>(blocks(nn), 2048)), dim3(256), 0, s, T.nodes.ptr, nn,
                     hist.ptr);
  std::vector<uint32_t> h = hist.to_host(kLevels, s);
  uint32_t maxDepth = 0;
  for (uint32_t d = 0; d < kLevels; d++) if (h[d]) maxDepth = d;
  MTX_REQUIRE(maxDepth < kLevels - 1, MTX_ERROR, "BVH deeper than 254 levels");
  std::vector<uint32_t> start(kLevels + 1, 0);
  for (uint32_t d = 0; d < kLevels; d++) start[d + 1] = start[d] + h[d];
  cursor.upload(start.data(), kLevels, s);
  hipLaunchKernelGGL(k_depth_scatter, dim3(std::min<uint32_t>(blocks(nn), 2048)), dim3(256), 0, s, T.nodes.ptr, nn,
                     cursor.ptr, byDepth.ptr);
  for (uint32_t d = maxDepth; d >= 1; d--)
    if (h[d]) hipLaunchKernelGGL(k_subtree, dim3(blocks(h[d])), dim3(256), 0, s, T.nodes.ptr, byDepth.ptr + start[d], h[d], size.ptr);
  uint32_t zero = 0;
  dfs.upload(&zero, 1, s);
  for (uint32_t d = 1; d <= maxDepth; d++)
    if (h[d]) hipLaunchKernelGGL(k_dfs, dim3(blocks(h[d])), dim3(256), 0, s, T.nodes.ptr, byDepth.ptr + start[d], h[d], size.ptr, dfs.ptr);
  uint32_t rootSize = 0;
  size.download(&rootSize, 1, s);
  BNode root;
  T.nodes.download(&root, 1, s);
  MTX_HIP(hipStreamSynchronize(s));
  L.interiorCount = root.primCount ? 0 : rootSize;
  L.tnodes.alloc(std::max<uint32_t>(1, L.interiorCount));
  if (L.interiorCount)
    hipLaunchKernelGGL(k_emit_nodes, dim3(blocks(nn)), dim3(256), 0, s, T.nodes.ptr, nn, dfs.ptr, L.tnodes.ptr);
  MTX_HIP(hipGetLastError());
  L.rootRef = root.primCount ? (int32_t)0x80000000u : 0;
  for (int k = 0; k < 3; k++) { L.rootLo[k] = root.lo[k]; L.rootHi[k] = root.hi[k]; }
  T.maxDepth = maxDepth;
}

// Builds the BLAS of every mesh: per mesh the primitive boxes and the wide
// top levels, then all meshes' small subtrees in one launch, then per mesh
// the depth-first layout and the leaf-ordered triangle records.
std::vector<std::shared_ptr<Blas>> build_blases(const MeshInput *meshes, size_t count, uint32_t bins, hipStream_t s) {
  struct Work {
    std::shared_ptr<Blas> B;
    DevBuf<float> hpos;
    DevBuf<uint32_t> hidx;
    const float *dpos = nullptr;
    const uint32_t *didx = nullptr;
  };
  std::vector<Work> W(count);
  std::vector<SmallTask> small;
  for (size_t m = 0; m < count; m++) {
    const MeshInput &in = meshes[m];
    Work &w = W[m];
    w.B = std::make_shared<Blas>();
    Blas &B = *w.B;
    B.triangleInput = in.tcount;
    B.vertexInput = in.vcount;
    w.dpos = in.pos;
    w.didx = in.idx;
    if (!in.deviceInput) {
      w.hpos.upload(in.pos, (size_t)in.vcount * 3, s);
      w.hidx.upload(in.idx, (size_t)in.tcount * 3, s);
      w.dpos = w.hpos.ptr;
      w.didx = w.hidx.ptr;
    }
    Topology &T = B.topo;
    uint32_t n = 0;
    const uint32_t tcount = in.tcount;
    if (tcount) {
      DevBuf<float4> lo(tcount), hi(tcount);
      DevBuf<uint32_t> valid(tcount), offs(tcount);
      hipLaunchKernelGGL(k_tri_aabbs, dim3(blocks(tcount)), dim3(256), 0, s, w.dpos, w.didx, tcount, in.vcount,
                         lo.ptr, hi.ptr, valid.ptr);
      size_t tmpBytes = 0;
      MTX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmpBytes, valid.ptr, offs.ptr, tcount, s));
      DevBuf<uint8_t> tmpScan(std::max<size_t>(1, tmpBytes));
      MTX_HIP(hipcub::DeviceScan::ExclusiveSum(tmpScan.ptr, tmpBytes, valid.ptr, offs.ptr, tcount, s));
      uint32_t lastOff = 0, lastValid = 0;
      MTX_HIP(hipMemcpyAsync(&lastOff, offs.ptr + tcount - 1, 4, hipMemcpyDeviceToHost, s));
      MTX_HIP(hipMemcpyAsync(&lastValid, valid.ptr + tcount - 1, 4, hipMemcpyDeviceToHost, s));
      MTX_HIP(hipStreamSynchronize(s));
      n = lastOff + lastValid;
      T.primCount = n;
      T.plo.alloc(std::max<uint32_t>(1, n));
      T.phi.alloc(std::max<uint32_t>(1, n));
      T.cen.alloc(std::max<uint32_t>(1, n));
      T.source.alloc(std::max<uint32_t>(1, n));
      hipLaunchKernelGGL(k_compact, dim3(blocks(tcount)), dim3(256), 0, s, lo.ptr, hi.ptr, valid.ptr, offs.ptr,
                         tcount, T.plo.ptr, T.phi.ptr, T.cen.ptr, T.source.ptr);
      MTX_HIP(hipGetLastError());
    }
    if (n) build_topology(T, bins, s, small);
  }
  run_small_builds(small, bins, s);
  std::vector<std::shared_ptr<Blas>> out(count);
  for (size_t m = 0; m < count; m++) {
    Work &w = W[m];
    Blas &B = *w.B;
    Topology &T = B.topo;
    out[m] = w.B;
    const uint32_t n = T.primCount;
    if (n == 0) continue;
    finish_topology(T, s);
    layout_topology(T, B.layout, s);
    DevBuf<uint32_t> leafCount(n);
    hipLaunchKernelGGL(k_leaf_counts, dim3(blocks(T.nodeCount)), dim3(256), 0, s, T.nodes.ptr, T.nodeCount,
                       leafCount.ptr);
    B.tris.alloc(n);
    B.triSource.alloc(n);
    hipLaunchKernelGGL(k_emit_tris, dim3(blocks(n)), dim3(256), 0, s, T.prim.ptr, T.source.ptr, leafCount.ptr, n,
                       w.dpos, w.didx, B.tris.ptr, B.triSource.ptr);
    MTX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_max_leaf, dim3(std::min<uint32_t>(blocks(T.nodeCount), 1024)), dim3(256), 0, s, T.nodes.ptr,
                       T.nodeCount, leafCount.ptr);
    MTX_HIP(hipGetLastError());
    B.maxLeaf = leafCount.to_host(1, s)[0];
    // The build scratch (centers, input boxes) is not needed on the hot path.
    T.cen.release();
    T.plo.release();
    T.phi.release();
    T.counter.release();
  }
  return out;
}

std::shared_ptr<Blas> build_blas(const float *pos, uint32_t vcount, const uint32_t *idx, uint32_t tcount,
                                 uint32_t bins, bool deviceInput, hipStream_t s) {
  const MeshInput in{pos, vcount, idx, tcount, deviceInput};
  return build_blases(&in, 1, bins, s)[0];
}

size_t export_blas(const Blas &B, const uint32_t *indices, uint32_t *words, size_t capWords,
                   uint32_t *sourceOfRecord, hipStream_t s) {
  const Topology &T = B.topo;
  if (T.nodeCount == 0) return 0;
  std::vector<BNode> nodes = T.nodes.to_host(T.nodeCount, s);
  size_t vec4 = 0;
  for (auto &nd : nodes) vec4 += 2 + nd.primCount;
  if (!words || capWords < vec4 * 4) return vec4 * 4;
  std::vector<uint32_t> prim = T.prim.to_host(T.primCount, s);
  std::vector<uint32_t> source = T.source.to_host(T.primCount, s);
  std::vector<uint32_t> offsets(T.nodeCount, 0);
  {
    std::vector<uint32_t> stack = {0};
    uint32_t off = 0;
    while (!stack.empty()) {
      uint32_t v = stack.back();
      stack.pop_back();
      offsets[v] = off;
      off += 2 + nodes[v].primCount;
      if (!nodes[v].primCount) { stack.push_back(nodes[v].leftFirst + 1); stack.push_back(nodes[v].leftFirst); }
    }
  }
  std::memset(words, 0, vec4 * 16);
  if (sourceOfRecord) for (size_t i = 0; i < vec4; i++) sourceOfRecord[i] = 0xffffffffu;
  for (uint32_t v = 0; v < T.nodeCount; v++) {
    const BNode &nd = nodes[v];
    size_t o = (size_t)offsets[v] * 4;
    for (int k = 0; k < 3; k++) { words[o + k] = fbits(nd.lo[k]); words[o + 4 + k] = fbits(nd.hi[k]); }
    if (!nd.primCount) { words[o + 3] = offsets[nd.leftFirst + 1]; continue; }
    words[o + 3] = (uint32_t)(-(int32_t)nd.primCount);
    for (uint32_t i = 0; i < nd.primCount; i++) {
      uint32_t t = source[prim[nd.leftFirst + i]];
      size_t e = o + 8 + (size_t)i * 4;
      if (indices) { words[e] = indices[t * 3]; words[e + 1] = indices[t * 3 + 1]; words[e + 2] = indices[t * 3 + 2]; }
      if (sourceOfRecord) sourceOfRecord[e / 4] = t;
    }
  }
  return vec4 * 4;
}

// Sum of node surface areas (k_node_area) of a topology.
static double node_area(const Topology &T, hipStream_t s) {
  DevBuf<double> acc(1);
  const double zero = 0.0;
  acc.upload(&zero, 1, s);
  hipLaunchKernelGGL(k_node_area, dim3(blocks(T.nodeCount)), dim3(256), 0, s, T.nodes.ptr, T.nodeCount, acc.ptr);
  MTX_HIP(hipGetLastError());
  return acc.to_host(1, s)[0];
}

// Node indices grouped by depth, the depth-first TNode index of every node
// and the build's surface area: computed once, on the first refit.
static void ensure_refit_data(Blas &B, hipStream_t s) {
  Topology &T = B.topo;
  const uint32_t nn = T.nodeCount;
  if (B.refit.ready || !nn) return;
  constexpr uint32_t kLevels = 256;
  DevBuf<uint32_t> hist(kLevels), cursor(kLevels), size(nn);
  std::vector<uint32_t> zeros(kLevels, 0);
  hist.upload(zeros.data(), kLevels, s);
  hipLaunchKernelGGL(k_depth_hist, dim3(std::min<uint32_t>(blocks(nn), 2048)), dim3(256), 0, s, T.nodes.ptr, nn,
                     hist.ptr);
  std::vector<uint32_t> h = hist.to_host(kLevels, s);
  std::vector<uint32_t> start(kLevels + 1, 0);
  for (uint32_t d = 0; d < kLevels; d++) start[d + 1] = start[d] + h[d];
  B.refit.byDepth.alloc(nn);
  B.refit.dfs.alloc(nn);
  cursor.upload(start.data(), kLevels, s);
  hipLaunchKernelGGL(k_depth_scatter, dim3(std::min<uint32_t>(blocks(nn), 2048)), dim3(256), 0, s, T.nodes.ptr, nn,
                     cursor.ptr, B.refit.byDepth.ptr);
  for (uint32_t d = T.maxDepth; d >= 1; d--)
    if (h[d])
      hipLaunchKernelGGL(k_subtree, dim3(blocks(h[d])), dim3(256), 0, s, T.nodes.ptr, B.refit.byDepth.ptr + start[d],
                         h[d], size.ptr);
  const uint32_t zero = 0;
  B.refit.dfs.upload(&zero, 1, s);
  for (uint32_t d = 1; d <= T.maxDepth; d++)
    if (h[d])
      hipLaunchKernelGGL(k_dfs, dim3(blocks(h[d])), dim3(256), 0, s, T.nodes.ptr, B.refit.byDepth.ptr + start[d],
                         h[d], size.ptr, B.refit.dfs.ptr);
  MTX_HIP(hipGetLastError());
  B.refit.levelStart.assign(start.begin(), start.end());
  B.refit.levelCount.assign(h.begin(), h.end());
  B.refit.buildArea = node_area(T, s);
  B.refit.ready = true;
}

bool refit_blas(Blas &B, const float *pos, const uint32_t *idx, hipStream_t s) {
  Topology &T = B.topo;
  if (!T.nodeCount) return false;
  ensure_refit_data(B, s);
  const auto &start = B.refit.levelStart;
  const auto &cnt = B.refit.levelCount;
  for (uint32_t d = T.maxDepth; d >= 1; d--)
    if (cnt[d])
      hipLaunchKernelGGL(k_refit_level, dim3(blocks(cnt[d])), dim3(256), 0, s, T.nodes.ptr,
                         B.refit.byDepth.ptr + start[d], cnt[d], T.prim.ptr, T.source.ptr, pos, idx);
  if (B.layout.interiorCount)
    hipLaunchKernelGGL(k_emit_nodes, dim3(blocks(T.nodeCount)), dim3(256), 0, s, T.nodes.ptr, T.nodeCount,
                       B.refit.dfs.ptr, B.layout.tnodes.ptr);
  DevBuf<uint32_t> leafCount(T.primCount);
  hipLaunchKernelGGL(k_leaf_counts, dim3(blocks(T.nodeCount)), dim3(256), 0, s, T.nodes.ptr, T.nodeCount,
                     leafCount.ptr);
  hipLaunchKernelGGL(k_emit_tris, dim3(blocks(T.primCount)), dim3(256), 0, s, T.prim.ptr, T.source.ptr,
                     leafCount.ptr, T.primCount, pos, idx, B.tris.ptr, B.triSource.ptr);
  MTX_HIP(hipGetLastError());
  BNode root;
  T.nodes.download(&root, 1, s);
  const double area = node_area(T, s);  // synchronizes
  for (int k = 0; k < 3; k++) { B.layout.rootLo[k] = root.lo[k]; B.layout.rootHi[k] = root.hi[k]; }
  return area > B.refit.buildArea * kDynamicBlasRebuildRatio;
}

// New BLAS root boxes for the instances of `blas`, then the TLAS refit over them.
void refit_accel_blas(Accel &A, uint32_t blas, hipStream_t s) {
  const Blas &B = *A.blases.at(blas);
  std::vector<uint32_t> list;
  for (uint32_t i = 0; i < A.instanceCount; i++)
    if (A.blasIndex[i] == blas) list.push_back(i);
  if (list.empty()) return;
  std::vector<float> boxes = A.blasBoxes.to_host((size_t)A.instanceCount * 6, s);
  for (uint32_t i : list)
    for (int k = 0; k < 3; k++) { boxes[(size_t)i * 6 + k] = B.layout.rootLo[k]; boxes[(size_t)i * 6 + 3 + k] = B.layout.rootHi[k]; }
  A.blasBoxes.upload(boxes.data(), boxes.size(), s);
  DevBuf<uint32_t> dl;
  dl.upload(list.data(), list.size(), s);
  hipLaunchKernelGGL(k_instance_update, dim3(blocks(list.size(), 64)), dim3(64), 0, s, A.transforms.ptr,
                     A.blasBoxes.ptr, dl.ptr, (uint32_t)list.size(), A.slotOfSource.ptr, A.slots.ptr);
  DevBuf<uint32_t> dStart;
  dStart.upload(A.levelStart.data(), A.levelStart.size(), s);
  hipLaunchKernelGGL(k_tlas_refit, dim3(1), dim3(256), 0, s, A.layout.tnodes.ptr, A.slots.ptr, A.levelNodes.ptr,
                     dStart.ptr, (uint32_t)(A.levelStart.size() - 1), A.view.ptr);
  MTX_HIP(hipGetLastError());
  MTX_HIP(hipStreamSynchronize(s));
}

static void fill_slot(InstSlot &sl, const double *m, const float box[6]) {
  double inv[16];
  if (!invert4(m, inv)) for (int k = 0; k < 16; k++) inv[k] = 0.0;
  sl.c0 = make_float4((float)inv[0], (float)inv[1], (float)inv[2], sl.c0.w);
  sl.c1 = make_float4((float)inv[4], (float)inv[5], (float)inv[6], sl.c1.w);
  sl.c2 = make_float4((float)inv[8], (float)inv[9], (float)inv[10], sl.c2.w);
  sl.c3 = make_float4((float)inv[12], (float)inv[13], (float)inv[14], sl.c3.w);
  sl.bmin = make_float4(box[0], box[1], box[2], sl.bmin.w);
  sl.bmax = make_float4(box[3], box[4], box[5], sl.bmax.w);
}

std::shared_ptr<Accel> build_accel(const std::vector<std::shared_ptr<Blas>> &blases, const mtx_instance_desc *inst,
                                   uint32_t count, uint32_t bins, uint32_t flags, hipStream_t s) {
  auto A = std::make_shared<Accel>();
  A->blases = blases;
  A->instanceCount = count;
  A->bins = bins;
  A->flags = flags;
  A->transformsHost.resize((size_t)count * 16);
  std::vector<float> boxes((size_t)count * 6), blasBoxes((size_t)count * 6);
  for (uint32_t i = 0; i < count; i++) {
    MTX_REQUIRE(inst[i].blas_index < blases.size(), MTX_INVALID_ARGUMENT, "instance blas_index out of range");
    const Blas &B = *blases[inst[i].blas_index];
    MTX_REQUIRE(B.topo.nodeCount > 0, MTX_INVALID_ARGUMENT, "instance refers to an empty BLAS");
    std::memcpy(&A->transformsHost[(size_t)i * 16], inst[i].transform, 128);
    for (int k = 0; k < 3; k++) { blasBoxes[i * 6 + k] = B.layout.rootLo[k]; blasBoxes[i * 6 + 3 + k] = B.layout.rootHi[k]; }
    world_bounds(inst[i].transform, B.layout.rootLo, B.layout.rootHi, &boxes[(size_t)i * 6]);
    A->blasIndex.push_back(inst[i].blas_index);
    A->material.push_back(inst[i].material_word);
    A->doubleSided.push_back(inst[i].double_sided ? 1 : 0);
    for (int k = 0; k < 3; k++) A->color.push_back(inst[i].color[k]);
    A->blasMaxDepth = std::max(A->blasMaxDepth, B.topo.maxDepth);
  }
  Topology &T = A->topo;
  T.primCount = count;
  if (count) {
    std::vector<float4> lo(count), hi(count);
    for (uint32_t i = 0; i < count; i++) {
      lo[i] = make_float4(boxes[i * 6], boxes[i * 6 + 1], boxes[i * 6 + 2], 0.0f);
      hi[i] = make_float4(boxes[i * 6 + 3], boxes[i * 6 + 4], boxes[i * 6 + 5], 0.0f);
    }
    T.plo.upload(lo.data(), count, s);
    T.phi.upload(hi.data(), count, s);
    T.cen.alloc(count);
    T.source.alloc(count);
    hipLaunchKernelGGL(k_centers, dim3(blocks(count)), dim3(256), 0, s, T.plo.ptr, T.phi.ptr, count, T.cen.ptr,
                       T.source.ptr);
    std::vector<SmallTask> small;
    build_topology(T, bins, s, small);
    run_small_builds(small, bins, s);
    finish_topology(T, s);
    layout_topology(T, A->layout, s);
  }
  A->maxDepth = T.maxDepth;
  // Instance slots in leaf ("target") order.
  std::vector<uint32_t> prim = count ? T.prim.to_host(count, s) : std::vector<uint32_t>();
  std::vector<BNode> nodes = count ? T.nodes.to_host(T.nodeCount, s) : std::vector<BNode>();
  std::vector<uint32_t> leafCount(count, 0);
  for (auto &nd : nodes)
    if (nd.primCount) for (uint32_t k = 0; k < nd.primCount; k++) leafCount[nd.leftFirst + k] = nd.primCount;
  std::vector<InstSlot> slots(std::max<uint32_t>(1, count));
  std::vector<uint32_t> slotOfSource(std::max<uint32_t>(1, count), 0);
  A->slotSource.assign(count, 0);
  for (uint32_t q = 0; q < count; q++) {
    uint32_t src = prim[q];
    const Blas &B = *blases[inst[src].blas_index];
    InstSlot sl{};
    sl.c0.w = intf(B.layout.rootRef);
    sl.c1.w = intf((int32_t)leafCount[q]);
    sl.c2.w = intf((int32_t)src);
    sl.c3.w = bitsf((inst[src].material_word & ~1u) | (inst[src].double_sided ? 1u : 0u));
    sl.bmin.w = intf((int32_t)B.topo.maxDepth);
    fill_slot(sl, inst[src].transform, &boxes[(size_t)src * 6]);
    sl.nodes = B.layout.tnodes.ptr;
    sl.tris = B.tris.ptr;
    slots[q] = sl;
    slotOfSource[src] = q;
    A->slotSource[q] = src;
  }
  A->slots.upload_min1(slots.data(), count, s);
  A->slotOfSource.upload_min1(slotOfSource.data(), count, s);
  A->transforms.upload_min1(A->transformsHost.data(), A->transformsHost.size(), s);
  A->blasBoxes.upload_min1(blasBoxes.data(), blasBoxes.size(), s);
  // TLAS node levels (deepest first) for the GPU refit.
  std::vector<std::vector<uint32_t>> levels;
  if (A->layout.interiorCount) {
    std::vector<TNode> tn = A->layout.tnodes.to_host(A->layout.interiorCount, s);
    std::vector<std::pair<uint32_t, uint32_t>> st = {{0, 0}};
    while (!st.empty()) {
      auto [v, d] = st.back();
      st.pop_back();
      if (levels.size() <= d) levels.resize(d + 1);
      levels[d].push_back(v);
      int32_t l = fint(tn[v].a.w), r = fint(tn[v].b.w);
      if (l >= 0) st.push_back({(uint32_t)l, d + 1});
      if (r >= 0) st.push_back({(uint32_t)r, d + 1});
    }
  }
  std::vector<uint32_t> flat, starts = {0};
  for (int d = (int)levels.size() - 1; d >= 0; d--) {
    flat.insert(flat.end(), levels[d].begin(), levels[d].end());
    starts.push_back((uint32_t)flat.size());
  }
  A->levelStart = starts;
  A->levelNodes.upload_min1(flat.data(), flat.size(), s);
  AccelView v{};
  v.tlasNodes = A->layout.tnodes.ptr;
  v.slots = A->slots.ptr;
  v.tlasRootRef = A->layout.rootRef;
  v.perInstanceBounds = (flags & MTX_ACCEL_PER_INSTANCE_BOUNDS) ? 1u : 0u;
  v.empty = count == 0 ? 1u : 0u;
  for (int k = 0; k < 3; k++) { v.rootMin[k] = A->layout.rootLo[k]; v.rootMax[k] = A->layout.rootHi[k]; }
  A->hostView = v;
  A->view.upload(&v, 1, s);
  MTX_HIP(hipStreamSynchronize(s));
  return A;
}

void refit_accel(Accel &A, const mtx_instance_desc *inst, uint32_t count, const uint32_t *dirty, uint32_t dirtyCount,
                 hipStream_t s) {
  MTX_REQUIRE(count == A.instanceCount, MTX_INVALID_ARGUMENT, "refit needs the same instance count");
  std::vector<uint32_t> list;
  if (dirty) {
    for (uint32_t i = 0; i < dirtyCount; i++) {
      MTX_REQUIRE(dirty[i] < count, MTX_INVALID_ARGUMENT, "dirty index out of range");
      list.push_back(dirty[i]);
    }
  } else {
    for (uint32_t i = 0; i < count; i++) list.push_back(i);
  }
  for (uint32_t i : list) {
    MTX_REQUIRE(inst[i].blas_index == A.blasIndex[i], MTX_INVALID_ARGUMENT, "refit cannot change an instance's BLAS");
    std::memcpy(&A.transformsHost[(size_t)i * 16], inst[i].transform, 128);
  }
  if (!count || list.empty()) return;
  A.transforms.upload(A.transformsHost.data(), A.transformsHost.size(), s);
  DevBuf<uint32_t> dl;
  dl.upload(list.data(), list.size(), s);
  hipLaunchKernelGGL(k_instance_update, dim3(blocks(list.size(), 64)), dim3(64), 0, s, A.transforms.ptr,
                     A.blasBoxes.ptr, dl.ptr, (uint32_t)list.size(), A.slotOfSource.ptr, A.slots.ptr);
  DevBuf<uint32_t> dStart;
  dStart.upload(A.levelStart.data(), A.levelStart.size(), s);
  hipLaunchKernelGGL(k_tlas_refit, dim3(1), dim3(256), 0, s, A.layout.tnodes.ptr, A.slots.ptr, A.levelNodes.ptr,
                     dStart.ptr, (uint32_t)(A.levelStart.size() - 1), A.view.ptr);
  MTX_HIP(hipGetLastError());
  MTX_HIP(hipStreamSynchronize(s));
}

// Reference-layout TLAS (bvhAccel.ts:385-449, web leaf boxes) and instance
// records (bvhAccel.ts:323-353) from the resident structure.
void export_tlas(const Accel &A, std::vector<uint32_t> &tw, std::vector<uint32_t> &iw, hipStream_t s) {
  tw.clear();
  iw.clear();
  const uint32_t count = A.instanceCount;
  if (!count) return;
  std::vector<InstSlot> slots = A.slots.to_host(count, s);
  std::vector<TNode> tn = A.layout.interiorCount ? A.layout.tnodes.to_host(A.layout.interiorCount, s) : std::vector<TNode>();
  AccelView v;
  A.view.download(&v, 1, s);
  MTX_HIP(hipStreamSynchronize(s));
  // Reconstruct nodes (bounds, leaf ranges) in depth-first order.
  struct N { float lo[3], hi[3]; int32_t ref; };
  std::vector<N> order;
  std::vector<std::pair<N, int>> st;
  N root{{v.rootMin[0], v.rootMin[1], v.rootMin[2]}, {v.rootMax[0], v.rootMax[1], v.rootMax[2]}, v.tlasRootRef};
  st.push_back({root, 0});
  std::vector<uint32_t> rightPatch;  // (index into order) needing right pointer
  std::vector<int32_t> parentOf;
  std::vector<uint32_t> offsets;
  uint32_t vec4 = 0;
  // Iterative DFS tracking offsets.
  struct Item { N n; int32_t parentOrder; };
  std::vector<Item> stack = {{root, -1}};
  std::vector<int32_t> parents;
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    uint32_t me = (uint32_t)order.size();
    order.push_back(it.n);
    parents.push_back(it.parentOrder);
    offsets.push_back(vec4);
    if (it.n.ref >= 0) {
      vec4 += 2;
      const TNode &t = tn[it.n.ref];
      N l{{t.a.x, t.a.y, t.a.z}, {t.b.x, t.b.y, t.b.z}, fint(t.a.w)};
      N r{{t.c.x, t.c.y, t.c.z}, {t.d.x, t.d.y, t.d.z}, fint(t.b.w)};
      stack.push_back({r, (int32_t)me});
      stack.push_back({l, (int32_t)me});
    } else {
      int32_t first = it.n.ref & 0x7fffffff;
      vec4 += 2 + (uint32_t)fint(slots[first].c1.w);
    }
  }
  const bool perInst = (A.flags & MTX_ACCEL_PER_INSTANCE_BOUNDS) != 0;
  tw.assign((size_t)(vec4 + (perInst ? count * 2 : 0)) * 4, 0u);
  // Right-child pointer: the second child pushed of each interior node.
  std::vector<int32_t> secondChild(order.size(), -1);
  for (size_t i = 0; i < order.size(); i++) {
    int32_t p = parents[i];
    if (p >= 0 && (size_t)p + 1 != i) secondChild[p] = (int32_t)i;
  }
  for (size_t i = 0; i < order.size(); i++) {
    const N &n = order[i];
    size_t o = (size_t)offsets[i] * 4;
    for (int k = 0; k < 3; k++) { tw[o + k] = fbits(n.lo[k]); tw[o + 4 + k] = fbits(n.hi[k]); }
    if (n.ref >= 0) { tw[o + 3] = offsets[secondChild[i]]; continue; }
    int32_t first = n.ref & 0x7fffffff, c = fint(slots[first].c1.w);
    tw[o + 3] = (uint32_t)(-c);
    for (int32_t k = 0; k < c; k++) {
      uint32_t target = (uint32_t)(first + k);
      uint32_t src = (uint32_t)fint(slots[target].c2.w);
      size_t e = o + 8 + (size_t)k * 4;
      tw[e] = target;
      tw[e + 1] = 0;  // BLAS offsets are buffer addresses here; filled by the caller if needed
      tw[e + 2] = src;
      if (perInst) {
        uint32_t box = vec4 + target * 2;
        tw[e + 3] = box;
        const InstSlot &sl = slots[target];
        tw[(size_t)box * 4] = fbits(sl.bmin.x); tw[(size_t)box * 4 + 1] = fbits(sl.bmin.y); tw[(size_t)box * 4 + 2] = fbits(sl.bmin.z);
        tw[(size_t)box * 4 + 4] = fbits(sl.bmax.x); tw[(size_t)box * 4 + 5] = fbits(sl.bmax.y); tw[(size_t)box * 4 + 6] = fbits(sl.bmax.z);
      }
    }
  }
  iw.assign((size_t)count * 40, 0u);
  for (uint32_t q = 0; q < count; q++) {
    const InstSlot &sl = slots[q];
    uint32_t src = (uint32_t)fint(sl.c2.w);
    uint32_t *r = &iw[(size_t)q * 40];
    r[0] = 0;
    r[1] = 0;
    r[2] = A.material[src];
    r[3] = fbits(0.0f);
    double inv[16];
    if (!invert4(&A.transformsHost[(size_t)src * 16], inv)) std::memset(inv, 0, sizeof inv);
    for (int k = 0; k < 16; k++) {
      r[4 + k] = fbits((float)inv[k]);
      r[20 + k] = fbits((float)A.transformsHost[(size_t)src * 16 + k]);
    }
    for (int k = 0; k < 3; k++) r[36 + k] = fbits(A.color[src * 3 + k]);
    r[39] = fbits(0.0f);
  }
}

}  // namespace mtx
