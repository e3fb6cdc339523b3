// Oracle: binned-SAH BVH2 build, BLAS/TLAS packing and TLAS refit
// (test infrastructure only).
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>

namespace orc {

// src/renderer/binnedSahBvh.ts:23-34
struct BinnedSahBvh {
  uint32_t nodeCount = 0;
  std::vector<float> bounds;       // 6 per node
  std::vector<uint32_t> leftFirst;
  std::vector<uint32_t> primCount;
  std::vector<uint32_t> primIndices;
};

// src/renderer/binnedSahBvh.ts:243-250
struct Bvh2Layout {
  std::vector<uint32_t> order, offsets;
  uint32_t vec4Count = 0, maxDepth = 0, maxLeafPrimitives = 0;
};

BinnedSahBvh build_binned_sah(const float *aabbs, uint32_t count, uint32_t bins);  // :70-236
Bvh2Layout layout_bvh2(const BinnedSahBvh &bvh);                                   // :252-275

// BVH2 BLAS buffer (bvhAccel.ts:200-286), without displacement.
struct FlatBvh {
  std::vector<uint32_t> words;     // vec4Count * 4 words (f32 or i32 lanes)
  uint32_t maxDepth = 0, maxLeafPrimitives = 0;
  double buildSurfaceArea = 0;
  std::vector<uint32_t> instanceWords;  // TLAS only: stride 40 words
};
FlatBvh build_blas(const float *positions, uint32_t vertexCount,
                   const uint32_t *indices, uint32_t triCount,
                   uint32_t bins, std::vector<uint32_t> *sourceTriangleOfSlot);

// refitBLAS on BVH2 words in place; returns the refitted node surface area.
double refit_blas_words(uint32_t *words, std::size_t vec4Count, const float *positions, uint32_t vertexCount);

// One TLAS input (bvhAccel.ts:61-68 TLASInstance).
struct TlasInstance {
  uint32_t blasIndex;
  float bounds[6];
  double transform[16];
  uint32_t material;
  float color[3];
};
// bvhAccel.ts:385-449 (buildTLAS); perInstanceBounds selects buildWebTLAS.
FlatBvh build_tlas(const TlasInstance *inst, uint32_t count,
                   const uint32_t *blasOffsets, const uint32_t *vertexOffsets,
                   uint32_t bins, bool perInstanceBounds);
// bvhAccel.ts:579-740 (full refit, dirty == nullptr) or dirty-ancestor refit.
void refit_tlas(FlatBvh &tlas, const TlasInstance *inst, uint32_t count,
                const uint32_t *blasOffsets, const uint32_t *vertexOffsets,
                const uint32_t *dirty, uint32_t dirtyCount);

// scenePreparation.ts:19-47: world box of the eight transformed BLAS corners.
void instance_world_bounds(const double *m, const float lo[3], const float hi[3], float out[6]);
// src/math/mat4.ts:27-90 (cofactor inverse in f64).
bool invert4(const double *a, double *out);

}  // namespace orc
