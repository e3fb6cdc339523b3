// Resident acceleration structure layout and the two-level traversal used by
// every ray query and by the path tracer (gfx950).
//
// Layout in HBM (all 16-byte aligned, read through the L2):
//   TNode   64 B  one interior node holding BOTH children's boxes, so an
//                 interior visit is one 64-byte fetch instead of the three
//                 scattered 16-byte reads of the vec4 BVH2 layout:
//                 a = (left.min, leftRef)  b = (left.max, rightRef)
//                 c = (right.min, 0)       d = (right.max, 0)
//                 ref >= 0: interior node index in the same array;
//                 ref <  0: leaf, first triangle slot = ref & 0x7fffffff.
//   TriRec  48 B  the triangle's three object-space vertices, gathered in
//                 leaf order (no index indirection on the hot path):
//                 v0.w = source triangle id, v1.w = leaf primitive count.
//   InstSlot 128 B per TLAS leaf entry in leaf order (the reference's
//                 "target" instance record): world-to-local columns,
//                 BLAS root ref, the web per-instance world box and the
//                 BLAS buffer addresses.
// Node topology, child visit order, box/triangle arithmetic and tie rules are
// those of the reference BVH2 traversal (intersect.wgsl.ts:213-627), so hit
// IDs and t match it bit-for-bit for identical f32 rays.
#pragma once
#include "mtx_math.h"

namespace mtx {

struct TNode {
  float4 a, b, c, d;
};
struct TriRec {
  float4 v0, v1, v2;
};
struct InstSlot {
  float4 c0;    // w2l column 0, w = BLAS root ref
  float4 c1;    // w2l column 1, w = primitive count of the TLAS leaf holding this slot
  float4 c2;    // w2l column 2, w = source instance index
  float4 c3;    // w2l column 3, w = (packed material word) | double-sided bit 0
  float4 bmin;  // world box (web per-instance bounds), w = BLAS max depth
  float4 bmax;  // w = reserved
  const TNode *nodes;
  const TriRec *tris;
  uint32_t pad[4];
};
static_assert(sizeof(TNode) == 64, "TNode");
static_assert(sizeof(TriRec) == 48, "TriRec");
static_assert(sizeof(InstSlot) == 128, "InstSlot");

// Traversal reads nodes, triangles and instance slots through global
// address-space pointers: loads through generic (flat) pointers also count
// against the LDS counter, so every wait for a node would wait for the LDS
// stack's operations too.
#if defined(__HIP_DEVICE_COMPILE__)
#define MTX_AS_GLOBAL __attribute__((address_space(1)))
#else
#define MTX_AS_GLOBAL
#endif
template <class T>
__device__ __forceinline__ const MTX_AS_GLOBAL T *as_global(const T *p) {
  return (const MTX_AS_GLOBAL T *)p;
}


struct AccelView {
  const TNode *tlasNodes;
  const InstSlot *slots;
  int32_t tlasRootRef;
  uint32_t perInstanceBounds;
  uint32_t empty;
  uint32_t pad;
  float rootMin[3], rootMax[3];
};

// intersect.wgsl.ts:1512-1521 (DEPTH_QUERY_MODE off: no padding).
MTX_HD float intersect_box(f3 o, f3 bmin, f3 bmax, f3 invD, float tMax) {
  f3 tBot = (bmin - o) * invD;
  f3 tTop = (bmax - o) * invD;
#if defined(__HIP_DEVICE_COMPILE__)
  // The hardware min/max (IEEE minNum/maxNum, min3/max3) return the non-NaN
  // operand as mnf/mxf do and differ from them only in the sign of a zero
  // result; the returned distance is only ever compared (> 0, with tMax and
  // with the sibling's), where -0 == +0, so every decision is the same.
  const float t0 = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(tBot.x, tTop.x), __builtin_fminf(tBot.y, tTop.y)),
                                   __builtin_fminf(tBot.z, tTop.z));
  const float t1 = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(tBot.x, tTop.x), __builtin_fmaxf(tBot.y, tTop.y)),
                                   __builtin_fmaxf(tBot.z, tTop.z));
#else
  f3 tNear = vmin(tBot, tTop);
  f3 tFar = vmax(tBot, tTop);
  float t0 = mxf(tNear.x, mxf(tNear.y, tNear.z));
  float t1 = mnf(tFar.x, mnf(tFar.y, tFar.z));
#endif
  return (t0 > t1 || t0 > tMax) ? -1.0f : (t0 > 0.0f ? t0 : t1);
}

MTX_HD int max_dimension(f3 v) { return v.x > v.y ? (v.x > v.z ? 0 : 2) : (v.y > v.z ? 1 : 2); }

// The third component, 1 / d[maxDim], is the same IEEE division as the
// ray's invD component, which callers pass in.
MTX_HD f3 get_shear(f3 d, f3 invD, int maxDim) {
  if (maxDim == 0) return f3{-d.y / d.x, -d.z / d.x, invD.x};
  if (maxDim == 1) return f3{-d.z / d.y, -d.x / d.y, invD.y};
  return f3{-d.x / d.z, -d.y / d.z, invD.z};
}

MTX_HD f3 permute_axes(f3 p, int maxDim) {
  return maxDim == 0 ? f3{p.y, p.z, p.x} : (maxDim == 1 ? f3{p.z, p.x, p.y} : p);
}

// Watertight ray/triangle test, intersect.wgsl.ts:1430-1493. t == 0 is a miss.
MTX_HD float intersect_triangle(f3 o, f3 p0, f3 p1, f3 p2, int maxDim, f3 shear, float tMax,
                                f3 &bary) {
  f3 a = permute_axes(p0 - o, maxDim), b = permute_axes(p1 - o, maxDim), c = permute_axes(p2 - o, maxDim);
  a = f3{a.x + shear.x * a.z, a.y + shear.y * a.z, a.z};
  b = f3{b.x + shear.x * b.z, b.y + shear.y * b.z, b.z};
  c = f3{c.x + shear.x * c.z, c.y + shear.y * c.z, c.z};
  f3 e{b.x * c.y - b.y * c.x, c.x * a.y - c.y * a.x, a.x * b.y - a.y * b.x};
  bool anyNeg = (e.x < 0.0f) | (e.y < 0.0f) | (e.z < 0.0f);
  bool anyPos = (e.x > 0.0f) | (e.y > 0.0f) | (e.z > 0.0f);
  if (anyNeg && anyPos) return 0.0f;
  float det = (e.x + e.y) + e.z;
  float z0 = a.z * shear.z, z1 = b.z * shear.z, z2 = c.z * shear.z;
  float tScaled = (e.x * z0 + e.y * z1) + e.z * z2;
  if (fabsf(tScaled) > fabsf(tMax * det)) return 0.0f;
  float invDet = 1.0f / det;
  bary = e * invDet;
  return tScaled * invDet;
}

struct TraceHit {
  float t;
  f3 bary;
  f3 normal;    // object-space geometric normal (normalize(cross(e1, e2)))
  f3 position;  // object-space o + t*d
  int32_t slot; // instance target slot, -1 on miss
  int32_t tri;  // BLAS-local triangle slot
  bool hit;
};

struct TraceCounters {
  uint32_t boxes, tris, instances;
};

// Per-lane traversal stack: the first kLds entries live in LDS
// (lds[i * ldsStride]), deeper entries spill to a per-lane global slice of
// kMaxStack - kLds entries.
constexpr int kLdsStack = 32;
constexpr int kMinLdsStack = 16;  // smallest kLds of any kernel; sizes the spill slices
constexpr int kMaxStack = 160;

template <int kLds>
struct LaneStackT {
  static_assert(kLds >= kMinLdsStack && kLds <= kMaxStack, "LDS stack depth");
  int32_t *lds;
  int32_t ldsStride;
  int32_t *spill;
  __device__ __forceinline__ void put(int i, int32_t v) {
    if (i < kLds) lds[i * ldsStride] = v; else spill[i - kLds] = v;
  }
  __device__ __forceinline__ int32_t get(int i) const {
    return i < kLds ? lds[i * ldsStride] : spill[i - kLds];
  }
};
using LaneStack = LaneStackT<kLdsStack>;

// Bit 1 of an instance slot's c3.w word: the instance's material is
// alpha-masked, so an any-hit query accepts its triangles only where the
// coverage test passes (intersectSceneQuery, intersect.wgsl.ts:643-646).
constexpr uint32_t kSlotMasked = 2u;

// Coverage test of any-hit queries that have no material data: every
// triangle occludes.
struct NoCoverage {
  __device__ __forceinline__ bool covers(int32_t, int32_t, f3) const { return true; }
};

// Two-level traversal: TLAS near-first stack, leaf instances in reverse
// order (optionally rejected by their world box), ray to object space with
// the instance's w2l, BLAS near-first stack with the root pushed untested,
// then the triangle loop with the self-hit skip and back-face culling of
// getBVHTraverseCode. anyHit returns at the first accepted triangle; on a
// masked instance a triangle is accepted only when cov.covers(slot,
// triangle, barycentrics) holds, and a rejected one leaves tMax unchanged.
template <bool kAnyHit, bool kCount, class Stack, class Cov = NoCoverage>
__device__ __forceinline__ bool traverse(const AccelView &acc, f3 ro, f3 rd, float tMax,
                                         bool doubleSided, int bounces, int32_t curTri,
                                         int32_t curSlot, Stack &st, TraceHit &h,
                                         TraceCounters *cnt, const Cov &cov = Cov()) {
  h.hit = false;
  h.slot = -1;
  h.tri = -1;
  h.t = kRayMaxDistance;
  if (acc.empty) return false;
  const f3 worldInvD = splat(1.0f) / rd;
  f3 o = ro, d = rd, invD = worldInvD, shear = splat(0.0f);
  int maxDim = 2;
  int sp = 0;
  int blasBase = -1;
  int leafFirst = 0, leafRemaining = 0;
  bool matDouble = false, masked = false;
  int32_t slotIdx = -1;
  const MTX_AS_GLOBAL TNode *bnodes = nullptr;
  const MTX_AS_GLOBAL TriRec *btris = nullptr;
  const MTX_AS_GLOBAL TNode *tnodes = as_global(acc.tlasNodes);
  const MTX_AS_GLOBAL InstSlot *slots = as_global(acc.slots);
  if (kCount) cnt->boxes++;
  if (intersect_box(o, f3{acc.rootMin[0], acc.rootMin[1], acc.rootMin[2]},
                    f3{acc.rootMax[0], acc.rootMax[1], acc.rootMax[2]}, invD, tMax) > 0.0f) {
    st.put(sp++, acc.tlasRootRef);
  }
  // Closest hit: one node step serves both levels (the node array and the
  // ray are picked per lane), so a wave whose lanes sit in the TLAS and in
  // BLASes runs the box-pair code once per iteration. Any hit keeps a
  // separate step per level, which measures faster for shadow rays. Either
  // way each lane pops, tests and pushes in the same order.
  if constexpr (!kAnyHit) {
    while (true) {
      const bool inBlas = (blasBase >= 0) & (sp > blasBase);
      if (!inBlas) {
        if (blasBase >= 0) {  // BLAS finished: back to the world ray
          blasBase = -1;
          o = ro;
          d = rd;
          invD = worldInvD;
        }
        if (leafRemaining > 0) {
          leafRemaining--;
          slotIdx = leafFirst + leafRemaining;
          const MTX_AS_GLOBAL InstSlot &s = slots[slotIdx];
          if (acc.perInstanceBounds) {
            if (kCount) cnt->boxes++;
            if (intersect_box(ro, f3of(s.bmin), f3of(s.bmax), worldInvD, tMax) <= 0.0f) continue;
          }
          if (kCount) cnt->instances++;
          float4 c0 = s.c0, c1 = s.c1, c2 = s.c2, c3 = s.c3;
          Affine w2l{f3of(c0), f3of(c1), f3of(c2), f3of(c3)};
          o = xform(w2l, ro, 1.0f);
          d = xform(w2l, rd, 0.0f);
          invD = splat(1.0f) / d;
          maxDim = max_dimension(vabs(d));
          shear = get_shear(d, invD, maxDim);
          matDouble = (fbits(c3.w) & 1u) != 0u;
          bnodes = as_global(s.nodes);
          btris = as_global(s.tris);
          blasBase = sp;
          st.put(sp++, fint(c0.w));
          continue;
        }
        if (sp == 0) break;
      }
      const int32_t ref = st.get(--sp);
      if (ref >= 0) {
        const TNode n = (inBlas ? bnodes : tnodes)[ref];
        float d0 = intersect_box(o, f3of(n.a), f3of(n.b), invD, tMax);
        float d1 = intersect_box(o, f3of(n.c), f3of(n.d), invD, tMax);
        if (kCount) cnt->boxes += 2;
        int32_t left = fint(n.a.w), right = fint(n.b.w);
        if (d0 > d1) {
          if (d0 > 0.0f) st.put(sp++, left);
          if (d1 > 0.0f) st.put(sp++, right);
        } else {
          if (d1 > 0.0f) st.put(sp++, right);
          if (d0 > 0.0f) st.put(sp++, left);
        }
      } else if (inBlas) {
        int32_t first = ref & 0x7fffffff;
        int32_t count = fint(btris[first].v1.w);
        for (int32_t k = 0; k < count; k++) {
          int32_t tri = first + k;
          if (bounces > 0 && tri == curTri && slotIdx == curSlot) continue;
          const TriRec r = btris[tri];
          f3 p0 = f3of(r.v0), p1 = f3of(r.v1), p2 = f3of(r.v2);
          if (kCount) cnt->tris++;
          f3 bary;
          float t = intersect_triangle(o, p0, p1, p2, maxDim, shear, tMax, bary);
          if (t > 0.0f) {
            f3 nrm = normalize(cross(p1 - p0, p2 - p0));
            if (!doubleSided && !matDouble && dot(nrm, d) > 0.0f) continue;
            tMax = t;
            h.hit = true;
            h.t = t;
            h.bary = bary;
            h.normal = nrm;
            h.position = o + tMax * d;
            h.slot = slotIdx;
            h.tri = tri;
            if (kAnyHit) return true;
          }
        }
      } else {
        leafFirst = ref & 0x7fffffff;
        leafRemaining = fint(slots[leafFirst].c1.w);
      }
    }
  } else {
    while (true) {
      if (blasBase >= 0 && sp > blasBase) {
        int32_t ref = st.get(--sp);
        if (ref >= 0) {
          const TNode n = bnodes[ref];
          float d0 = intersect_box(o, f3of(n.a), f3of(n.b), invD, tMax);
          float d1 = intersect_box(o, f3of(n.c), f3of(n.d), invD, tMax);
          if (kCount) cnt->boxes += 2;
          int32_t left = fint(n.a.w), right = fint(n.b.w);
          if (d0 > d1) {
            if (d0 > 0.0f) st.put(sp++, left);
            if (d1 > 0.0f) st.put(sp++, right);
          } else {
            if (d1 > 0.0f) st.put(sp++, right);
            if (d0 > 0.0f) st.put(sp++, left);
          }
        } else {
          int32_t first = ref & 0x7fffffff;
          int32_t count = fint(btris[first].v1.w);
          for (int32_t k = 0; k < count; k++) {
            int32_t tri = first + k;
            if (bounces > 0 && tri == curTri && slotIdx == curSlot) continue;
            const TriRec r = btris[tri];
            f3 p0 = f3of(r.v0), p1 = f3of(r.v1), p2 = f3of(r.v2);
            if (kCount) cnt->tris++;
            f3 bary;
            float t = intersect_triangle(o, p0, p1, p2, maxDim, shear, tMax, bary);
            if (t > 0.0f) {
              f3 nrm = normalize(cross(p1 - p0, p2 - p0));
              if (!doubleSided && !matDouble && dot(nrm, d) > 0.0f) continue;
              if (masked && !cov.covers(slotIdx, tri, bary)) continue;
              tMax = t;
              h.hit = true;
              h.t = t;
              h.bary = bary;
              h.normal = nrm;
              h.position = o + tMax * d;
              h.slot = slotIdx;
              h.tri = tri;
              if (kAnyHit) return true;
            }
          }
        }
        continue;
      }
      if (blasBase >= 0) blasBase = -1;
      if (leafRemaining > 0) {
        leafRemaining--;
        slotIdx = leafFirst + leafRemaining;
        const MTX_AS_GLOBAL InstSlot &s = slots[slotIdx];
        if (acc.perInstanceBounds) {
          if (kCount) cnt->boxes++;
          if (intersect_box(ro, f3of(s.bmin), f3of(s.bmax), worldInvD, tMax) <= 0.0f) continue;
        }
        if (kCount) cnt->instances++;
        float4 c0 = s.c0, c1 = s.c1, c2 = s.c2, c3 = s.c3;
        Affine w2l{f3of(c0), f3of(c1), f3of(c2), f3of(c3)};
        o = xform(w2l, ro, 1.0f);
        d = xform(w2l, rd, 0.0f);
        invD = splat(1.0f) / d;
        maxDim = max_dimension(vabs(d));
        shear = get_shear(d, invD, maxDim);
        matDouble = (fbits(c3.w) & 1u) != 0u;
        masked = (fbits(c3.w) & kSlotMasked) != 0u;
        bnodes = as_global(s.nodes);
        btris = as_global(s.tris);
        blasBase = sp;
        st.put(sp++, fint(c0.w));
        continue;
      }
      if (sp == 0) break;
      int32_t ref = st.get(--sp);
      o = ro;
      d = rd;
      invD = worldInvD;
      if (ref >= 0) {
        const TNode n = tnodes[ref];
        float d0 = intersect_box(o, f3of(n.a), f3of(n.b), invD, tMax);
        float d1 = intersect_box(o, f3of(n.c), f3of(n.d), invD, tMax);
        if (kCount) cnt->boxes += 2;
        int32_t left = fint(n.a.w), right = fint(n.b.w);
        if (d0 > d1) {
          if (d0 > 0.0f) st.put(sp++, left);
          if (d1 > 0.0f) st.put(sp++, right);
        } else {
          if (d1 > 0.0f) st.put(sp++, right);
          if (d0 > 0.0f) st.put(sp++, left);
        }
      } else {
        leafFirst = ref & 0x7fffffff;
        leafRemaining = fint(slots[leafFirst].c1.w);
      }
    }
  }
  return h.hit;
}

}  // namespace mtx
