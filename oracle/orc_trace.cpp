// Oracle: triangle/box tests and reference-order two-level BVH2 traversal.
// Test infrastructure only. Restates src/renderer/wgsl/chunks/intersect.wgsl.ts.
#include "orc_trace.h"

namespace orc {

float intersect_box(V3 o, V3 bmin, V3 bmax, V3 invD, float tMax) {
  V3 tBot = (bmin - o) * invD;
  V3 tTop = (bmax - o) * invD;
  V3 tNear = vmin(tBot, tTop);
  V3 tFar = vmax(tBot, tTop);
  float t0 = mxf(tNear.x, mxf(tNear.y, tNear.z));
  float t1 = mnf(tFar.x, mnf(tFar.y, tFar.z));
  if (t0 > t1 || t0 > tMax) return -1.0f;
  return t0 > 0.0f ? t0 : t1;
}

int max_dimension(V3 v) {
  return v.x > v.y ? (v.x > v.z ? 0 : 2) : (v.y > v.z ? 1 : 2);
}

V3 get_shear(V3 d, int maxDim) {
  if (maxDim == 0) return V3(-d.y, -d.z, 1.0f) / d.x;
  if (maxDim == 1) return V3(-d.z, -d.x, 1.0f) / d.y;
  return V3(-d.x, -d.y, 1.0f) / d.z;
}

static inline V3 permute(V3 p, int maxDim) {
  if (maxDim == 0) return V3(p.y, p.z, p.x);
  if (maxDim == 1) return V3(p.z, p.x, p.y);
  return p;
}

TriangleIntersect intersect_triangle(V3 o, V3 d, V3 p0, V3 p1, V3 p2, int maxDim, V3 shear,
                                     float tMax) {
  (void)d;
  TriangleIntersect ti;
  ti.t = 0.0f;
  V3 p0t = permute(p0 - o, maxDim), p1t = permute(p1 - o, maxDim), p2t = permute(p2 - o, maxDim);
  p0t = V3(p0t.x + shear.x * p0t.z, p0t.y + shear.y * p0t.z, p0t.z);
  p1t = V3(p1t.x + shear.x * p1t.z, p1t.y + shear.y * p1t.z, p1t.z);
  p2t = V3(p2t.x + shear.x * p2t.z, p2t.y + shear.y * p2t.z, p2t.z);
  V3 e(p1t.x * p2t.y - p1t.y * p2t.x, p2t.x * p0t.y - p2t.y * p0t.x, p0t.x * p1t.y - p0t.y * p1t.x);
  bool anyNeg = e.x < 0.0f || e.y < 0.0f || e.z < 0.0f;
  bool anyPos = e.x > 0.0f || e.y > 0.0f || e.z > 0.0f;
  if (anyNeg && anyPos) return ti;
  float det = (e.x + e.y) + e.z;
  float z0 = p0t.z * shear.z, z1 = p1t.z * shear.z, z2 = p2t.z * shear.z;
  float tScaled = (e.x * z0 + e.y * z1) + e.z * z2;
  if (fabsf(tScaled) > fabsf(tMax * det)) return ti;
  float invDet = 1.0f / det;
  ti.t = tScaled * invDet;
  ti.bary = e * invDet;
  return ti;
}

static inline V3 ld3(const std::vector<uint32_t> &w, size_t vec4) {
  return V3(u2f(w[vec4 * 4]), u2f(w[vec4 * 4 + 1]), u2f(w[vec4 * 4 + 2]));
}
static inline V3 pos(const TraceScene &s, int32_t i) {
  return V3(s.positions[(size_t)i * 3], s.positions[(size_t)i * 3 + 1], s.positions[(size_t)i * 3 + 2]);
}

// getBVHTraverseCode (intersect.wgsl.ts:213-627), BVH2 layout only.
// anyHit: stop at the first accepted triangle (shadow query).
static ClosestHit traverse(const TraceScene &s, V3 ro, V3 rd, float tMax, bool doubleSided,
                           int bounces, int32_t curTri, int32_t curInst, bool anyHit,
                           TraceStats *st, const Coverage *cov = nullptr) {
  ClosestHit res{};
  res.hit = false; res.instance = -1; res.record = -1; res.t = RAY_MAX_DISTANCE;
  if (s.tlas.empty()) return res;
  V3 o = ro, d = rd;
  const V3 worldInvD = V3(1.0f) / rd;
  V3 invD = worldInvD;
  int maxDim = 2;
  V3 shear;
  int32_t tlasStack[128];
  int32_t blasStack[128];
  int tlasTop = -1;
  if (st) st->boxTests++;
  if (intersect_box(o, ld3(s.tlas, 0), ld3(s.tlas, 1), invD, tMax) > 0.0f) { tlasTop = 0; tlasStack[0] = 0; }
  while (tlasTop >= 0) {
    int32_t tlasNode = tlasStack[tlasTop--];
    o = ro; d = rd; invD = V3(1.0f) / d;
    int32_t ptr = (int32_t)s.tlas[(size_t)tlasNode * 4 + 3];
    if (ptr >= 0) {
      int32_t left = tlasNode + 2, right = ptr;
      float d0 = intersect_box(o, ld3(s.tlas, left), ld3(s.tlas, left + 1), invD, tMax);
      float d1 = intersect_box(o, ld3(s.tlas, right), ld3(s.tlas, right + 1), invD, tMax);
      if (st) st->boxTests += 2;
      if (d0 > d1) {
        if (d0 > 0.0f) tlasStack[++tlasTop] = left;
        if (d1 > 0.0f) tlasStack[++tlasTop] = right;
      } else {
        if (d1 > 0.0f) tlasStack[++tlasTop] = right;
        if (d0 > 0.0f) tlasStack[++tlasTop] = left;
      }
      continue;
    }
    int32_t count = -ptr;
    for (int32_t io = count - 1; io >= 0; io--) {
      size_t rec = (size_t)(tlasNode + 2 + io) * 4;
      int32_t insIdx = (int32_t)s.tlas[rec];
      int32_t blasRoot = (int32_t)s.tlas[rec + 1];
      if (s.perInstanceBounds) {
        int32_t box = (int32_t)s.tlas[rec + 3];
        if (st) st->boxTests++;
        if (intersect_box(ro, ld3(s.tlas, box), ld3(s.tlas, box + 1), worldInvD, tMax) <= 0.0f) continue;
      }
      if (st) st->instanceVisits++;
      const uint32_t *ins = &s.instances[(size_t)insIdx * 40];
      float w2l[16];
      for (int k = 0; k < 16; k++) w2l[k] = u2f(ins[4 + k]);
      o = mat_mul_point(w2l, ro, 1.0f);
      d = mat_mul_point(w2l, rd, 0.0f);
      invD = V3(1.0f) / d;
      maxDim = max_dimension(vabs(d));
      shear = get_shear(d, maxDim);
      int32_t blasOffset = (int32_t)(ins[0] & 0x3fffffffu);
      int32_t vertexOffset = (int32_t)ins[1];
      int32_t material = ((int32_t)ins[2]) >> 8;
      bool matDouble = material >= 0 && (size_t)material < s.materialDoubleSided.size() &&
                       s.materialDoubleSided[material];
      int blasTop = 0;
      blasStack[0] = blasRoot;
      while (blasTop >= 0) {
        int32_t node = blasStack[blasTop--];
        int32_t primitiveCount = 0, primitiveOffset = 0;
        int32_t bptr = (int32_t)s.blas[(size_t)node * 4 + 3];
        if (bptr >= 0) {
          int32_t left = node + 2, right = bptr + blasOffset;
          float d0 = intersect_box(o, ld3(s.blas, left), ld3(s.blas, left + 1), invD, tMax);
          float d1 = intersect_box(o, ld3(s.blas, right), ld3(s.blas, right + 1), invD, tMax);
          if (st) st->boxTests += 2;
          if (d0 > d1) {
            if (d0 > 0.0f) blasStack[++blasTop] = left;
            if (d1 > 0.0f) blasStack[++blasTop] = right;
          } else {
            if (d1 > 0.0f) blasStack[++blasTop] = right;
            if (d0 > 0.0f) blasStack[++blasTop] = left;
          }
        } else {
          primitiveCount = -bptr;
          primitiveOffset = node + 2;
        }
        for (int32_t k = 0; k < primitiveCount; k++) {
          int32_t tri = primitiveOffset + k;
          if (bounces > 0 && tri == curTri && insIdx == curInst) continue;
          size_t r3 = (size_t)tri * 4;
          V3 p0 = pos(s, (int32_t)s.blas[r3] + vertexOffset);
          V3 p1 = pos(s, (int32_t)s.blas[r3 + 1] + vertexOffset);
          V3 p2 = pos(s, (int32_t)s.blas[r3 + 2] + vertexOffset);
          if (st) st->triTests++;
          TriangleIntersect h = intersect_triangle(o, d, p0, p1, p2, maxDim, shear, tMax);
          if (h.t > 0.0f) {
            V3 n = normalize(cross(p1 - p0, p2 - p0));
            if (!doubleSided && !matDouble && dot(n, d) > 0.0f) continue;
            // intersect.wgsl.ts:643-646: a candidate the coverage test
            // rejects leaves tMax at the shadow ray's maximum.
            if (anyHit && cov && !cov->covers(cov->ctx, insIdx, tri, h.bary)) continue;
            tMax = h.t;
            res.hit = true; res.instance = insIdx; res.record = tri; res.t = tMax;
            res.bary = h.bary; res.normal = n; res.position = o + tMax * d;
            if (anyHit) return res;
          }
        }
      }
    }
  }
  return res;
}

ClosestHit trace_closest(const TraceScene &s, V3 ro, V3 rd, float tMax, bool doubleSided,
                         int bounces, int32_t curTri, int32_t curInst, TraceStats *st) {
  return traverse(s, ro, rd, tMax, doubleSided, bounces, curTri, curInst, false, st);
}

bool trace_any(const TraceScene &s, V3 ro, V3 rd, float maxDistance, int32_t curTri,
               int32_t curInst, TraceStats *st, const Coverage *cov) {
  float tMax = fminf(fmaxf(maxDistance, 0.0f), RAY_MAX_DISTANCE);
  return traverse(s, ro, rd, tMax, true, 1, curTri, curInst, true, st, cov).hit;
}

ClosestHit trace_brute(const TraceScene &s, V3 ro, V3 rd, float tMax, bool doubleSided) {
  ClosestHit res{};
  res.hit = false; res.instance = -1; res.record = -1; res.t = RAY_MAX_DISTANCE;
  uint32_t nInst = (uint32_t)(s.instances.size() / 40);
  for (uint32_t insIdx = 0; insIdx < nInst; insIdx++) {
    const uint32_t *ins = &s.instances[(size_t)insIdx * 40];
    float w2l[16];
    for (int k = 0; k < 16; k++) w2l[k] = u2f(ins[4 + k]);
    V3 o = mat_mul_point(w2l, ro, 1.0f), d = mat_mul_point(w2l, rd, 0.0f);
    int maxDim = max_dimension(vabs(d));
    V3 shear = get_shear(d, maxDim);
    int32_t blasOffset = (int32_t)(ins[0] & 0x3fffffffu), vertexOffset = (int32_t)ins[1];
    int32_t material = ((int32_t)ins[2]) >> 8;
    bool matDouble = material >= 0 && (size_t)material < s.materialDoubleSided.size() &&
                     s.materialDoubleSided[material];
    // Walk this BLAS's leaves in buffer order (every triangle exactly once).
    std::vector<int32_t> stack = {blasOffset};
    while (!stack.empty()) {
      int32_t node = stack.back(); stack.pop_back();
      int32_t bptr = (int32_t)s.blas[(size_t)node * 4 + 3];
      if (bptr >= 0) { stack.push_back(bptr + blasOffset); stack.push_back(node + 2); continue; }
      for (int32_t k = 0; k < -bptr; k++) {
        int32_t tri = node + 2 + k;
        size_t r3 = (size_t)tri * 4;
        V3 p0 = pos(s, (int32_t)s.blas[r3] + vertexOffset);
        V3 p1 = pos(s, (int32_t)s.blas[r3 + 1] + vertexOffset);
        V3 p2 = pos(s, (int32_t)s.blas[r3 + 2] + vertexOffset);
        TriangleIntersect h = intersect_triangle(o, d, p0, p1, p2, maxDim, shear, tMax);
        if (h.t > 0.0f) {
          V3 n = normalize(cross(p1 - p0, p2 - p0));
          if (!doubleSided && !matDouble && dot(n, d) > 0.0f) continue;
          tMax = h.t;
          res.hit = true; res.instance = (int32_t)insIdx; res.record = tri; res.t = tMax;
          res.bary = h.bary; res.normal = n; res.position = o + tMax * d;
        }
      }
    }
  }
  return res;
}

}  // namespace orc
