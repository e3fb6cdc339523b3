// Resumable two-level traversal for the persistent wavefront trace kernels.
//
// Walker visits exactly the nodes traverse() (mtx_accel.h) visits, in the
// same order, with the same box and triangle arithmetic and the same
// hit-update rule, so its closest hit (or first opaque hit) is bit-identical.
// Two things differ in how the work is scheduled, not in what is computed:
//   - the node to visit next is kept in a register: after an interior node
//     the nearer hit child becomes `cur` and only the farther one is pushed
//     (the stack-based form pushes both and pops the nearer at once);
//   - the state lives in registers between calls of step(), so a lane whose
//     ray has finished can take a new ray while the rest of its wave keeps
//     traversing (the kernels refill lanes through a wave-aggregated atomic).
#pragma once
#include "mtx_accel.h"

namespace mtx {

// Per-lane stack: kLds entries in LDS (stride = block size), deeper entries
// in the lane's global spill slice.
template <int kLds>
struct WalkStack {
  int32_t *lds;
  int32_t stride;
  int32_t *spill;
  __device__ __forceinline__ void put(int i, int32_t v) {
    if (i < kLds) lds[i * stride] = v; else spill[i - kLds] = v;
  }
  __device__ __forceinline__ int32_t get(int i) const { return i < kLds ? lds[i * stride] : spill[i - kLds]; }
};

template <bool kAnyHit, bool kCount>
struct Walker {
  f3 ro, rd, winv;    // world ray and 1/direction
  f3 o, d, invD;      // ray in the current level's space
  f3 shear;
  int maxDim;
  float tMax;
  int32_t cur;        // ref to process next
  int sp, blasBase, leafFirst, leafRemaining;
  int32_t slotIdx, curTri, curSlot;
  int bounces;
  bool inBlas, done, matDouble, masked, doubleSided;
  const TNode *bnodes;
  const TriRec *btris;
  // result
  bool hit;
  float t;
  f3 bary;
  int32_t hitSlot, hitTri;
  TraceCounters cnt;

  template <class Stack>
  __device__ __forceinline__ void advance(const AccelView &acc, Stack &st) {
    if (inBlas) {
      if (sp > blasBase) { cur = st.get(--sp); return; }
      inBlas = false;
      o = ro;
      d = rd;
      invD = winv;
    }
    while (leafRemaining > 0) {
      leafRemaining--;
      slotIdx = leafFirst + leafRemaining;
      const InstSlot &s = acc.slots[slotIdx];
      if (acc.perInstanceBounds) {
        if (kCount) cnt.boxes++;
        if (intersect_box(ro, f3of(s.bmin), f3of(s.bmax), winv, tMax) <= 0.0f) continue;
      }
      if (kCount) cnt.instances++;
      const float4 c0 = s.c0, c1 = s.c1, c2 = s.c2, c3 = s.c3;
      const Affine w2l{f3of(c0), f3of(c1), f3of(c2), f3of(c3)};
      o = xform(w2l, ro, 1.0f);
      d = xform(w2l, rd, 0.0f);
      invD = splat(1.0f) / d;
      maxDim = max_dimension(vabs(d));
      shear = get_shear(d, invD, maxDim);
      matDouble = (fbits(c3.w) & 1u) != 0u;
      masked = kAnyHit && (fbits(c3.w) & kSlotMasked) != 0u;
      bnodes = s.nodes;
      btris = s.tris;
      inBlas = true;
      blasBase = sp;
      cur = fint(c0.w);
      return;
    }
    if (sp == 0) { done = true; return; }
    cur = st.get(--sp);
  }

  template <class Stack>
  __device__ __forceinline__ void init(const AccelView &acc, f3 rayO, f3 rayD, float tmax, bool dbl, int bnc,
                                       int32_t cTri, int32_t cSlot, Stack &st) {
    ro = rayO; rd = rayD; tMax = tmax; doubleSided = dbl; bounces = bnc; curTri = cTri; curSlot = cSlot;
    hit = false; t = kRayMaxDistance; hitSlot = -1; hitTri = -1; bary = splat(0.0f);
    if (kCount) cnt = TraceCounters{0u, 0u, 0u};
    sp = 0; blasBase = 0; leafFirst = 0; leafRemaining = 0; slotIdx = -1;
    inBlas = false; matDouble = false; masked = false; bnodes = nullptr; btris = nullptr;
    maxDim = 2; shear = splat(0.0f);
    winv = splat(1.0f) / rd;
    o = ro; d = rd; invD = winv;
    done = acc.empty != 0u;
    if (done) return;
    if (kCount) cnt.boxes++;
    if (intersect_box(o, f3{acc.rootMin[0], acc.rootMin[1], acc.rootMin[2]},
                      f3{acc.rootMax[0], acc.rootMax[1], acc.rootMax[2]}, invD, tMax) > 0.0f) {
      cur = acc.tlasRootRef;
    } else {
      done = true;
    }
  }

  // Processes `cur`; returns true once the query has finished. Any-hit
  // queries accept a triangle of a masked instance only where cov.covers()
  // holds (the coverage rule of traverse()).
  template <class Stack, class Cov = NoCoverage>
  __device__ __forceinline__ bool step(const AccelView &acc, Stack &st, const Cov &cov = Cov()) {
    if (cur >= 0) {
      const TNode n = (inBlas ? bnodes : acc.tlasNodes)[cur];
      const float d0 = intersect_box(o, f3of(n.a), f3of(n.b), invD, tMax);
      const float d1 = intersect_box(o, f3of(n.c), f3of(n.d), invD, tMax);
      if (kCount) cnt.boxes += 2;
      const int32_t left = fint(n.a.w), right = fint(n.b.w);
      const bool leftNear = !(d0 > d1);
      const int32_t nearRef = leftNear ? left : right, farRef = leftNear ? right : left;
      const bool nearHit = (leftNear ? d0 : d1) > 0.0f, farHit = (leftNear ? d1 : d0) > 0.0f;
      if (nearHit) {
        if (farHit) st.put(sp++, farRef);
        cur = nearRef;
      } else if (farHit) {
        cur = farRef;
      } else {
        advance(acc, st);
      }
    } else if (inBlas) {
      const int32_t first = cur & 0x7fffffff;
      const int32_t count = fint(btris[first].v1.w);
      for (int32_t k = 0; k < count; k++) {
        const int32_t tri = first + k;
        if (bounces > 0 && tri == curTri && slotIdx == curSlot) continue;
        const TriRec r = btris[tri];
        const f3 p0 = f3of(r.v0), p1 = f3of(r.v1), p2 = f3of(r.v2);
        if (kCount) cnt.tris++;
        f3 b;
        const float tt = intersect_triangle(o, p0, p1, p2, maxDim, shear, tMax, b);
        if (tt > 0.0f) {
          if (!doubleSided && !matDouble) {
            const f3 nrm = normalize(cross(p1 - p0, p2 - p0));
            if (dot(nrm, d) > 0.0f) continue;
          }
          if (kAnyHit && masked && !cov.covers(slotIdx, tri, b)) continue;
          tMax = tt;
          hit = true;
          t = tt;
          bary = b;
          hitSlot = slotIdx;
          hitTri = tri;
          if (kAnyHit) { done = true; return true; }
        }
      }
      advance(acc, st);
    } else {
      leafFirst = cur & 0x7fffffff;
      leafRemaining = fint(acc.slots[leafFirst].c1.w);
      advance(acc, st);
    }
    return done;
  }
};

}  // namespace mtx
