// Oracle: watertight triangle test, slab test and reference-order two-level
// BVH2 traversal over the reference buffer layout. Test infrastructure only.
#pragma once
#include <cstdint>
#include <vector>
#include "orc_common.h"

namespace orc {

// intersect.wgsl.ts:1421-1424
struct TriangleIntersect { float t; V3 bary; };

float intersect_box(V3 o, V3 bmin, V3 bmax, V3 invD, float tMax);   // :1512-1521
int max_dimension(V3 v);                                              // :1523-1525
V3 get_shear(V3 d, int maxDim);                                       // :1527-1537
TriangleIntersect intersect_triangle(V3 o, V3 d, V3 p0, V3 p1, V3 p2,
                                     int maxDim, V3 shear, float tMax); // :1430-1493

// Scene in the reference layout: TLAS words (with web per-instance boxes),
// concatenated BLAS words (offsets in vec4 units), instance words (stride 40)
// and one global vertex position array (xyz f32).
struct TraceScene {
  std::vector<uint32_t> tlas;
  std::vector<uint32_t> blas;
  std::vector<uint32_t> instances;
  std::vector<float> positions;
  std::vector<uint8_t> materialDoubleSided;  // per material index
  bool perInstanceBounds = true;
  // Parity identities: BLAS-region vec4 record -> source triangle,
  // instance target index -> source instance index.
  std::vector<uint32_t> sourceTriangleOfRecord;
  std::vector<uint32_t> sourceInstanceOfTarget;
};

struct TraceStats { uint64_t boxTests = 0, triTests = 0, instanceVisits = 0; };

// querySceneSoftware (intersect.wgsl.ts:660-684) / intersectScene (:696-740):
// closest hit; `record` is the BLAS-region vec4 index of the leaf entry.
struct ClosestHit {
  bool hit; int32_t instance; int32_t record; float t; V3 bary;
  V3 normal;      // object-space geometric normal of the accepted triangle
  V3 position;    // object-space hit position o + t*d
};
ClosestHit trace_closest(const TraceScene &s, V3 ro, V3 rd, float tMax, bool doubleSided,
                         int bounces, int32_t currentTriangle, int32_t currentInstance,
                         TraceStats *stats);
// Coverage test of an any-hit candidate (intersectionAlphaPasses,
// intersectionAlpha.wgsl.ts:28-88): false lets the shadow ray pass the
// triangle `record` of `instance` at barycentrics `bary`.
struct Coverage {
  bool (*covers)(const void *ctx, int32_t instance, int32_t record, V3 bary);
  const void *ctx;
};
// intersectSceneShadow (intersect.wgsl.ts:829-879): first hit that the
// coverage test accepts (every hit when cov is null).
bool trace_any(const TraceScene &s, V3 ro, V3 rd, float maxDistance,
               int32_t currentTriangle, int32_t currentInstance, TraceStats *stats,
               const Coverage *cov = nullptr);
// All-triangles closest hit (same triangle test and tie rule, no BVH).
ClosestHit trace_brute(const TraceScene &s, V3 ro, V3 rd, float tMax, bool doubleSided);

}  // namespace orc
