// Resident scene objects shared by the library's translation units: GPU-built
// BLASes, scene accelerations (TLAS + instance slots) and their builders.
#pragma once
#include <memory>
#include <vector>

#include "mtx_internal.h"

namespace mtx {

constexpr int kSmall = 1024;   // largest node finished by one workgroup in LDS
constexpr int kMaxBins = 64;

// Build node with the reference builder's fields (binnedSahBvh.ts:23-34):
// interior: leftFirst = left child (right = leftFirst + 1), primCount = 0;
// leaf: leftFirst = first slot in the primitive order, primCount > 0.
struct BNode {
  float lo[3];
  uint32_t leftFirst;
  float hi[3];
  uint32_t primCount;
  uint32_t depth;   // root = 1, as layoutBvh2 counts it
  int32_t parent;
  uint32_t pad[2];
};
static_assert(sizeof(BNode) == 48, "BNode");

__host__ __device__ __forceinline__ float jminf(float a, float b) {
  return a < b ? a : (b < a ? b : (signbit(a) ? a : b));
}
__host__ __device__ __forceinline__ float jmaxf(float a, float b) {
  return a > b ? a : (b > a ? b : (signbit(a) ? b : a));
}

// f64 cofactor inverse of a column-major 4x4 by 2x2 minors, the form of the
// reference's mat4.invert (src/math/mat4.ts:27-90). Host and device use this
// one definition so the f32 world-to-local columns agree bit-for-bit.
MTX_HD bool invert4(const double *a, double *out) {
  double a00 = a[0], a01 = a[1], a02 = a[2], a03 = a[3];
  double a10 = a[4], a11 = a[5], a12 = a[6], a13 = a[7];
  double a20 = a[8], a21 = a[9], a22 = a[10], a23 = a[11];
  double a30 = a[12], a31 = a[13], a32 = a[14], a33 = a[15];
  double b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10;
  double b02 = a00 * a13 - a03 * a10, b03 = a01 * a12 - a02 * a11;
  double b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
  double b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30;
  double b08 = a20 * a33 - a23 * a30, b09 = a21 * a32 - a22 * a31;
  double b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
  double det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
  if (det == 0 || det != det) return false;
  det = 1.0 / det;
  out[0] = (a11 * b11 - a12 * b10 + a13 * b09) * det;
  out[1] = (a02 * b10 - a01 * b11 - a03 * b09) * det;
  out[2] = (a31 * b05 - a32 * b04 + a33 * b03) * det;
  out[3] = (a22 * b04 - a21 * b05 - a23 * b03) * det;
  out[4] = (a12 * b08 - a10 * b11 - a13 * b07) * det;
  out[5] = (a00 * b11 - a02 * b08 + a03 * b07) * det;
  out[6] = (a32 * b02 - a30 * b05 - a33 * b01) * det;
  out[7] = (a20 * b05 - a22 * b02 + a23 * b01) * det;
  out[8] = (a10 * b10 - a11 * b08 + a13 * b06) * det;
  out[9] = (a01 * b08 - a00 * b10 - a03 * b06) * det;
  out[10] = (a30 * b04 - a31 * b02 + a33 * b00) * det;
  out[11] = (a21 * b02 - a20 * b04 - a23 * b00) * det;
  out[12] = (a11 * b07 - a10 * b09 - a12 * b06) * det;
  out[13] = (a00 * b09 - a01 * b07 + a02 * b06) * det;
  out[14] = (a31 * b01 - a30 * b03 - a32 * b00) * det;
  out[15] = (a20 * b03 - a21 * b01 + a22 * b00) * det;
  return true;
}

// World box of an instance: the eight BLAS root-box corners through the
// transform, each rounded to f32 (scenePreparation.ts:19-47,
// vec3.transformMat4 of src/math/vec3.ts:96-106).
MTX_HD void world_bounds(const double *m, const float lo[3], const float hi[3], float out[6]) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int corner = 0; corner < 8; corner++) {
    double x = (corner & 1) ? hi[0] : lo[0];
    double y = (corner & 2) ? hi[1] : lo[1];
    double z = (corner & 4) ? hi[2] : lo[2];
    double w = m[3] * x + m[7] * y + m[11] * z + m[15];
    if (w == 0 || w != w) w = 1.0;
    float p[3];
    for (int k = 0; k < 3; k++) p[k] = (float)((m[k] * x + m[4 + k] * y + m[8 + k] * z + m[12 + k]) / w);
    if (!isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2])) continue;
    for (int k = 0; k < 3; k++) { mn[k] = jminf(mn[k], p[k]); mx[k] = jmaxf(mx[k], p[k]); }
  }
  for (int k = 0; k < 3; k++) { out[k] = mn[k]; out[k + 3] = mx[k]; }
}

struct Topology {
  DevBuf<BNode> nodes;
  DevBuf<uint32_t> counter;  // nodes allocated so far (device atomic during the build)
  DevBuf<uint32_t> prim;    // leaf-ordered compacted primitive ids
  DevBuf<uint32_t> source;  // compacted primitive -> input primitive (triangle / instance)
  DevBuf<float4> plo, phi, cen;
  uint32_t primCount = 0, nodeCount = 0, maxDepth = 0;
};

struct Layout {
  DevBuf<TNode> tnodes;
  uint32_t interiorCount = 0;
  int32_t rootRef = 0;
  float rootLo[3] = {0, 0, 0}, rootHi[3] = {0, 0, 0};
};

// refitBLAS's ratio of refitted to built surface area above which a rebuild
// is advised (DYNAMIC_BLAS_REBUILD_RATIO, bvhAccel.ts:743).
constexpr double kDynamicBlasRebuildRatio = 2.0;

// Refit bookkeeping of a BLAS, made on its first refit: nodes by depth, the
// depth-first TNode index of every node, the build's node surface area.
struct RefitData {
  bool ready = false;
  DevBuf<uint32_t> byDepth, dfs;
  std::vector<uint32_t> levelStart, levelCount;
  double buildArea = 0.0;
};

struct Blas {
  Topology topo;
  Layout layout;
  DevBuf<TriRec> tris;
  DevBuf<uint32_t> triSource;  // leaf slot -> source triangle
  RefitData refit;
  uint32_t maxLeaf = 0;
  uint32_t triangleInput = 0;
  uint32_t vertexInput = 0;
  size_t resident_bytes() const {
    return layout.tnodes.bytes() + tris.bytes() + triSource.bytes();
  }
};

struct Accel {
  Topology topo;
  Layout layout;
  DevBuf<InstSlot> slots;
  DevBuf<AccelView> view;
  DevBuf<double> transforms;      // source order, 16 per instance (f64)
  DevBuf<float> blasBoxes;        // per source instance: BLAS root box (6 floats)
  DevBuf<uint32_t> slotOfSource;
  DevBuf<uint32_t> levelNodes;    // TLAS TNode indices grouped by depth, deepest first
  std::vector<uint32_t> levelStart;
  std::vector<std::shared_ptr<Blas>> blases;
  std::vector<uint32_t> blasIndex, material, slotSource;
  std::vector<uint8_t> doubleSided;
  std::vector<float> color;
  std::vector<double> transformsHost;
  uint32_t instanceCount = 0, bins = 64, flags = 0, maxDepth = 0, blasMaxDepth = 1;
  AccelView hostView{};
  uint32_t max_stack() const { return maxDepth + blasMaxDepth + 2; }
};

struct MeshInput {
  const float *pos;
  uint32_t vcount;
  const uint32_t *idx;
  uint32_t tcount;
  bool deviceInput;
};
std::shared_ptr<Blas> build_blas(const float *pos, uint32_t vcount, const uint32_t *idx, uint32_t tcount,
                                 uint32_t bins, bool deviceInput, hipStream_t s);
std::vector<std::shared_ptr<Blas>> build_blases(const MeshInput *meshes, size_t count, uint32_t bins, hipStream_t s);
size_t export_blas(const Blas &B, const uint32_t *indices, uint32_t *words, size_t capWords,
                   uint32_t *sourceOfRecord, hipStream_t s);
std::shared_ptr<Accel> build_accel(const std::vector<std::shared_ptr<Blas>> &blases,
                                   const mtx_instance_desc *inst, uint32_t count, uint32_t bins,
                                   uint32_t flags, hipStream_t s);
// refitBLAS: new vertex positions (device pointers, the build's vertex count
// and index buffer); returns true when the refitted surface area exceeds
// kDynamicBlasRebuildRatio times the build's (a rebuild is advised).
bool refit_blas(Blas &B, const float *pos, const uint32_t *idx, hipStream_t s);
// After refit_blas: the instances of BLAS `blas` take its new root box and the TLAS is refitted.
void refit_accel_blas(Accel &A, uint32_t blas, hipStream_t s);
void refit_accel(Accel &A, const mtx_instance_desc *inst, uint32_t count, const uint32_t *dirty,
                 uint32_t dirtyCount, hipStream_t s);
void export_tlas(const Accel &A, std::vector<uint32_t> &tlasWords, std::vector<uint32_t> &instanceWords,
                 hipStream_t s);

}  // namespace mtx
