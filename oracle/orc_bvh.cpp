// Oracle: binned-SAH BVH2 build (f64 arithmetic, f32 storage), BVH2 layout,
// BLAS/TLAS packing and TLAS refit. Test infrastructure only.
#include "orc_bvh.h"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace orc {

static const double FAR = 1e30;  // binnedSahBvh.ts:21

// binnedSahBvh.ts:46-49. JS `x | 0` is ToInt32: truncate, wrap mod 2^32, NaN/inf -> 0.
static inline int32_t js_to_int32(double v) {
  if (!std::isfinite(v)) return 0;
  double t = std::trunc(v);
  double m = std::fmod(t, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return (int32_t)(uint32_t)m;
}
static inline int32_t bin_of(double center, double nodeMin, double scale, int32_t bins) {
  int32_t bin = js_to_int32((center - nodeMin) * scale);
  return bin < 0 ? 0 : bin > bins - 1 ? bins - 1 : bin;
}
static inline double jmin(double a, double b) {  // Math.min for non-NaN operands
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}
static inline double jmax(double a, double b) {
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}

BinnedSahBvh build_binned_sah(const float *aabbs, uint32_t count, uint32_t BINS) {
  BinnedSahBvh out;
  uint32_t capacity = std::max<uint32_t>(1, count * 2 - 1);
  if (count == 0) capacity = 1;
  out.bounds.assign((size_t)capacity * 6, 0.0f);
  out.leftFirst.assign(capacity, 0);
  out.primCount.assign(capacity, 0);
  out.primIndices.assign(count, 0);
  if (count == 0) { out.nodeCount = 0; return out; }
  float *bounds = out.bounds.data();
  uint32_t *leftFirst = out.leftFirst.data(), *primCount = out.primCount.data();
  uint32_t *primIndices = out.primIndices.data();

  std::vector<float> centers((size_t)count * 3);
  double rMinX = FAR, rMinY = FAR, rMinZ = FAR, rMaxX = -FAR, rMaxY = -FAR, rMaxZ = -FAR;
  for (uint32_t i = 0; i < count; i++) {
    const float *b = aabbs + (size_t)i * 6;
    double minX = b[0], minY = b[1], minZ = b[2], maxX = b[3], maxY = b[4], maxZ = b[5];
    centers[i * 3 + 0] = (float)((minX + maxX) * 0.5);
    centers[i * 3 + 1] = (float)((minY + maxY) * 0.5);
    centers[i * 3 + 2] = (float)((minZ + maxZ) * 0.5);
    if (minX < rMinX) rMinX = minX;
    if (minY < rMinY) rMinY = minY;
    if (minZ < rMinZ) rMinZ = minZ;
    if (maxX > rMaxX) rMaxX = maxX;
    if (maxY > rMaxY) rMaxY = maxY;
    if (maxZ > rMaxZ) rMaxZ = maxZ;
    primIndices[i] = i;
  }
  bounds[0] = (float)rMinX; bounds[1] = (float)rMinY; bounds[2] = (float)rMinZ;
  bounds[3] = (float)rMaxX; bounds[4] = (float)rMaxY; bounds[5] = (float)rMaxZ;
  leftFirst[0] = 0;
  primCount[0] = count;
  const double minDim[3] = {(rMaxX - rMinX) * 1e-20, (rMaxY - rMinY) * 1e-20, (rMaxZ - rMinZ) * 1e-20};

  std::vector<double> binMin((size_t)3 * BINS * 3), binMax((size_t)3 * BINS * 3);
  std::vector<uint32_t> binCount((size_t)3 * BINS);
  std::vector<double> leftCost(BINS - 1), rightCost(BINS - 1);
  uint32_t newNode = 1;
  std::vector<uint32_t> tasks;
  uint32_t nodeIndex = 0;
  const int32_t B = (int32_t)BINS;
  while (true) {
    while (true) {
      const float *nb = bounds + (size_t)nodeIndex * 6;
      double nMin[3] = {nb[0], nb[1], nb[2]};
      double extent[3] = {(double)nb[3] - nMin[0], (double)nb[4] - nMin[1], (double)nb[5] - nMin[2]};
      double area = extent[0] * extent[1] + extent[1] * extent[2] + extent[2] * extent[0];
      if (area == 0) break;
      uint32_t first = leftFirst[nodeIndex];
      uint32_t nodePrims = primCount[nodeIndex];
      if (nodePrims == 1) break;
      double scale[3];
      for (int a = 0; a < 3; a++) scale[a] = extent[a] > minDim[a] ? (double)BINS / extent[a] : 0.0;
      std::fill(binMin.begin(), binMin.end(), FAR);
      std::fill(binMax.begin(), binMax.end(), -FAR);
      std::fill(binCount.begin(), binCount.end(), 0u);
      for (uint32_t i = first, end = first + nodePrims; i < end; i++) {
        uint32_t p = primIndices[i];
        const float *b = aabbs + (size_t)p * 6;
        for (int a = 0; a < 3; a++) {
          int32_t bin = a * B + bin_of(centers[p * 3 + a], nMin[a], scale[a], B);
          binCount[bin]++;
          double *mn = &binMin[(size_t)bin * 3], *mx = &binMax[(size_t)bin * 3];
          for (int k = 0; k < 3; k++) {
            mn[k] = jmin(mn[k], (double)b[k]);
            mx[k] = jmax(mx[k], (double)b[k + 3]);
          }
        }
      }
      double splitCost = FAR;
      int bestAxis = 0, bestPos = 0;
      for (int axis = 0; axis < 3; axis++) {
        if (!(extent[axis] > minDim[axis])) continue;
        double lMin[3] = {FAR, FAR, FAR}, lMax[3] = {-FAR, -FAR, -FAR};
        double rMin[3] = {FAR, FAR, FAR}, rMax[3] = {-FAR, -FAR, -FAR};
        uint32_t leftN = 0, rightN = 0;
        for (int i = 0; i < B - 1; i++) {
          size_t l = (size_t)(axis * B + i) * 3, r = (size_t)(axis * B + B - 1 - i) * 3;
          for (int k = 0; k < 3; k++) {
            lMin[k] = jmin(lMin[k], binMin[l + k]);
            lMax[k] = jmax(lMax[k], binMax[l + k]);
          }
          for (int k = 0; k < 3; k++) {
            rMin[k] = jmin(rMin[k], binMin[r + k]);
            rMax[k] = jmax(rMax[k], binMax[r + k]);
          }
          leftN += binCount[axis * B + i];
          rightN += binCount[axis * B + B - 1 - i];
          double lx = lMax[0] - lMin[0], ly = lMax[1] - lMin[1], lz = lMax[2] - lMin[2];
          double rx = rMax[0] - rMin[0], ry = rMax[1] - rMin[1], rz = rMax[2] - rMin[2];
          leftCost[i] = leftN == 0 ? FAR : (lx * ly + ly * lz + lz * lx) * (double)leftN;
          rightCost[B - 2 - i] = rightN == 0 ? FAR : (rx * ry + ry * rz + rz * rx) * (double)rightN;
        }
        for (int i = 0; i < B - 1; i++) {
          double cost = leftCost[i] + rightCost[i];
          if (cost < splitCost) { splitCost = cost; bestAxis = axis; bestPos = i; }
        }
      }
      splitCost = 1.0 + 1.0 * splitCost / area;
      if (!(splitCost < (double)nodePrims * 1.0)) break;

      double sc = scale[bestAxis], nm = nMin[bestAxis];
      uint32_t j = first + nodePrims, src = first;
      for (uint32_t i = 0; i < nodePrims; i++) {
        uint32_t p = primIndices[src];
        if (bin_of(centers[p * 3 + bestAxis], nm, sc, B) <= bestPos) {
          src++;
        } else {
          primIndices[src] = primIndices[--j];
          primIndices[j] = p;
        }
      }
      uint32_t leftPrims = src - first, rightPrims = nodePrims - leftPrims;
      if (leftPrims == 0 || rightPrims == 0) break;
      uint32_t n = newNode;
      newNode += 2;
      float *lb = bounds + (size_t)n * 6, *rb = lb + 6;
      for (int k = 0; k < 3; k++) {
        lb[k] = rb[k] = (float)FAR;
        lb[3 + k] = rb[3 + k] = (float)-FAR;
      }
      for (int i = 0; i < B; i++) {
        size_t o = (size_t)(bestAxis * B + i) * 3;
        float *base = i <= bestPos ? lb : rb;
        for (int k = 0; k < 3; k++) {
          if (binMin[o + k] < (double)base[k]) base[k] = (float)binMin[o + k];
          if (binMax[o + k] > (double)base[3 + k]) base[3 + k] = (float)binMax[o + k];
        }
      }
      leftFirst[n] = first; primCount[n] = leftPrims;
      leftFirst[n + 1] = j; primCount[n + 1] = rightPrims;
      leftFirst[nodeIndex] = n; primCount[nodeIndex] = 0;
      tasks.push_back(n + 1);
      nodeIndex = n;
    }
    if (tasks.empty()) break;
    nodeIndex = tasks.back();
    tasks.pop_back();
  }
  out.nodeCount = newNode;
  return out;
}

Bvh2Layout layout_bvh2(const BinnedSahBvh &bvh) {
  Bvh2Layout L;
  L.order.assign(bvh.nodeCount, 0);
  L.offsets.assign(bvh.nodeCount, 0);
  if (!bvh.nodeCount) return L;
  std::vector<uint32_t> stack = {0, 1};
  uint32_t written = 0;
  while (!stack.empty()) {
    uint32_t depth = stack.back(); stack.pop_back();
    uint32_t node = stack.back(); stack.pop_back();
    L.order[written++] = node;
    L.offsets[node] = L.vec4Count;
    if (depth > L.maxDepth) L.maxDepth = depth;
    uint32_t prims = bvh.primCount[node];
    L.vec4Count += 2 + prims;
    if (prims) {
      if (prims > L.maxLeafPrimitives) L.maxLeafPrimitives = prims;
    } else {
      stack.push_back(bvh.leftFirst[node] + 1); stack.push_back(depth + 1);
      stack.push_back(bvh.leftFirst[node]); stack.push_back(depth + 1);
    }
  }
  return L;
}

static inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

FlatBvh build_blas(const float *positions, uint32_t vertexCount, const uint32_t *indices,
                   uint32_t triCount, uint32_t bins, std::vector<uint32_t> *sourceOfSlot) {
  (void)vertexCount;
  FlatBvh F;
  std::vector<float> aabbs((size_t)triCount * 6);
  std::vector<uint32_t> triangles(triCount);
  uint32_t count = 0;
  for (uint32_t t = 0; t < triCount; t++) {  // bvhAccel.ts:126-168 without displacement
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 3; c++) {
      const float *p = positions + (size_t)indices[t * 3 + c] * 3;
      for (int k = 0; k < 3; k++) {
        double v = p[k];
        mn[k] = std::isnan(v) || std::isnan(mn[k]) ? NAN : jmin(mn[k], v);
        mx[k] = std::isnan(v) || std::isnan(mx[k]) ? NAN : jmax(mx[k], v);
      }
    }
    float *o = &aabbs[(size_t)count * 6];
    bool finite = true;
    for (int k = 0; k < 3; k++) {
      o[k] = (float)mn[k]; o[k + 3] = (float)mx[k];
      finite = finite && std::isfinite(mn[k]) && std::isfinite(mx[k]);
    }
    if (finite) triangles[count++] = t;
  }
  if (!count) return F;
  BinnedSahBvh bvh = build_binned_sah(aabbs.data(), count, bins);
  Bvh2Layout L = layout_bvh2(bvh);
  F.words.assign((size_t)L.vec4Count * 4, 0u);
  if (sourceOfSlot) sourceOfSlot->assign(L.vec4Count, 0xffffffffu);
  for (uint32_t k = 0; k < bvh.nodeCount; k++) {
    uint32_t node = L.order[k];
    size_t o = (size_t)L.offsets[node] * 4, b = (size_t)node * 6;
    const float *bd = &bvh.bounds[b];
    F.words[o] = fbits(bd[0]); F.words[o + 1] = fbits(bd[1]); F.words[o + 2] = fbits(bd[2]);
    F.words[o + 4] = fbits(bd[3]); F.words[o + 5] = fbits(bd[4]); F.words[o + 6] = fbits(bd[5]);
    double x = (double)bd[3] - bd[0], y = (double)bd[4] - bd[1], z = (double)bd[5] - bd[2];
    F.buildSurfaceArea += 2 * (x * z + x * y + z * y);
    uint32_t prims = bvh.primCount[node];
    if (!prims) {
      F.words[o + 3] = L.offsets[bvh.leftFirst[node] + 1];
      continue;
    }
    F.words[o + 3] = (uint32_t)(-(int32_t)prims);
    for (uint32_t i = 0; i < prims; i++) {
      uint32_t t = triangles[bvh.primIndices[bvh.leftFirst[node] + i]];
      size_t e = o + 8 + (size_t)i * 4;
      F.words[e] = indices[t * 3]; F.words[e + 1] = indices[t * 3 + 1]; F.words[e + 2] = indices[t * 3 + 2];
      F.words[e + 3] = 0;
      if (sourceOfSlot) (*sourceOfSlot)[e / 4] = t;
    }
  }
  F.maxDepth = L.maxDepth;
  F.maxLeafPrimitives = L.maxLeafPrimitives;
  return F;
}

bool invert4(const double *a, double *out) {
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
  if (det == 0 || std::isnan(det)) return false;
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

void instance_world_bounds(const double *m, const float lo[3], const float hi[3], float out[6]) {
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int corner = 0; corner < 8; corner++) {
    double x = (corner & 1) ? hi[0] : lo[0];
    double y = (corner & 2) ? hi[1] : lo[1];
    double z = (corner & 4) ? hi[2] : lo[2];
    double w = m[3] * x + m[7] * y + m[11] * z + m[15];
    if (w == 0 || std::isnan(w)) w = 1.0;
    float p[3];
    for (int k = 0; k < 3; k++) p[k] = (float)((m[k] * x + m[4 + k] * y + m[8 + k] * z + m[12 + k]) / w);
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
    for (int k = 0; k < 3; k++) { mn[k] = jmin(mn[k], p[k]); mx[k] = jmax(mx[k], p[k]); }
  }
  for (int k = 0; k < 3; k++) { out[k] = (float)mn[k]; out[k + 3] = (float)mx[k]; }
}

static void write_instance_record(uint32_t *w, uint32_t target, const TlasInstance &inst,
                                  const uint32_t *blasOffsets, const uint32_t *vertexOffsets) {
  uint32_t *r = w + (size_t)target * 40;  // INSTANCE_WORD_STRIDE (bvhAccel.ts:16-17)
  r[0] = blasOffsets[inst.blasIndex] | (0u << 30);
  r[1] = vertexOffsets[inst.blasIndex];
  r[2] = inst.material;
  r[3] = fbits(0.0f);
  double inv[16];
  if (!invert4(inst.transform, inv)) std::memset(inv, 0, sizeof inv);
  for (int i = 0; i < 16; i++) {
    r[4 + i] = fbits((float)inv[i]);
    r[20 + i] = fbits((float)inst.transform[i]);
  }
  r[36] = fbits(inst.color[0]); r[37] = fbits(inst.color[1]); r[38] = fbits(inst.color[2]);
  r[39] = fbits(0.0f);
}

FlatBvh build_tlas(const TlasInstance *inst, uint32_t count, const uint32_t *blasOffsets,
                   const uint32_t *vertexOffsets, uint32_t bins, bool perInstanceBounds) {
  FlatBvh F;
  if (!count) return F;
  std::vector<float> aabbs((size_t)count * 6);
  for (uint32_t i = 0; i < count; i++) std::memcpy(&aabbs[(size_t)i * 6], inst[i].bounds, 24);
  BinnedSahBvh bvh = build_binned_sah(aabbs.data(), count, bins);
  Bvh2Layout L = layout_bvh2(bvh);
  F.words.assign((size_t)(L.vec4Count + (perInstanceBounds ? count * 2 : 0)) * 4, 0u);
  F.instanceWords.assign((size_t)count * 40, 0u);
  uint32_t target = 0;
  for (uint32_t k = 0; k < bvh.nodeCount; k++) {
    uint32_t node = L.order[k];
    size_t o = (size_t)L.offsets[node] * 4;
    const float *bd = &bvh.bounds[(size_t)node * 6];
    F.words[o] = fbits(bd[0]); F.words[o + 1] = fbits(bd[1]); F.words[o + 2] = fbits(bd[2]);
    F.words[o + 4] = fbits(bd[3]); F.words[o + 5] = fbits(bd[4]); F.words[o + 6] = fbits(bd[5]);
    uint32_t prims = bvh.primCount[node];
    if (!prims) { F.words[o + 3] = L.offsets[bvh.leftFirst[node] + 1]; continue; }
    F.words[o + 3] = (uint32_t)(-(int32_t)prims);
    for (uint32_t i = 0; i < prims; i++) {
      uint32_t source = bvh.primIndices[bvh.leftFirst[node] + i];
      size_t e = o + 8 + (size_t)i * 4;
      F.words[e] = target;
      F.words[e + 1] = blasOffsets[inst[source].blasIndex];
      F.words[e + 2] = source;
      if (perInstanceBounds) {
        uint32_t box = L.vec4Count + target * 2;
        F.words[e + 3] = box;
        for (int c = 0; c < 3; c++) {
          F.words[(size_t)box * 4 + c] = fbits(aabbs[(size_t)source * 6 + c]);
          F.words[(size_t)box * 4 + 4 + c] = fbits(aabbs[(size_t)source * 6 + 3 + c]);
        }
      }
      write_instance_record(F.instanceWords.data(), target, inst[source], blasOffsets, vertexOffsets);
      target++;
    }
  }
  F.maxDepth = L.maxDepth;
  F.maxLeafPrimitives = L.maxLeafPrimitives;
  return F;
}

void refit_tlas(FlatBvh &F, const TlasInstance *inst, uint32_t count, const uint32_t *blasOffsets,
                const uint32_t *vertexOffsets, const uint32_t *dirty, uint32_t dirtyCount) {
  if (F.words.empty() || F.instanceWords.empty()) return;
  std::vector<uint8_t> isDirty(count, dirty ? 0 : 1);
  for (uint32_t i = 0; i < dirtyCount; i++) if (dirty[i] < count) isDirty[dirty[i]] = 1;
  float *fl = reinterpret_cast<float *>(F.words.data());
  int32_t *in = reinterpret_cast<int32_t *>(F.words.data());
  // bvhAccel.ts:658-687 refitNode, iterative post-order.
  struct Item { uint32_t index; bool expanded; };
  std::vector<Item> stack = {{0, false}};
  while (!stack.empty()) {
    Item it = stack.back(); stack.pop_back();
    size_t base = (size_t)it.index * 4;
    int32_t child = in[base + 3];
    if (child >= 0) {
      if (!it.expanded) {
        stack.push_back({it.index, true});
        stack.push_back({(uint32_t)child, false});
        stack.push_back({it.index + 2, false});
        continue;
      }
      size_t left = base + 8, right = (size_t)child * 4;
      for (int k = 0; k < 3; k++) {
        fl[base + k] = (float)jmin(fl[left + k], fl[right + k]);
        fl[base + 4 + k] = (float)jmax(fl[left + 4 + k], fl[right + 4 + k]);
      }
      continue;
    }
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = 0; i < -child; i++) {
      size_t record = base + 8 + (size_t)i * 4;
      uint32_t source = (uint32_t)in[record + 2];
      if (source >= count) continue;
      const float *b = inst[source].bounds;
      for (int k = 0; k < 3; k++) { mn[k] = jmin(mn[k], b[k]); mx[k] = jmax(mx[k], b[k + 3]); }
      if (isDirty[source]) {
        int32_t box = in[record + 3];
        if (box) {
          for (int k = 0; k < 3; k++) { fl[(size_t)box * 4 + k] = b[k]; fl[(size_t)box * 4 + 4 + k] = b[k + 3]; }
        }
        write_instance_record(F.instanceWords.data(), (uint32_t)in[record], inst[source], blasOffsets, vertexOffsets);
      }
    }
    for (int k = 0; k < 3; k++) { fl[base + k] = (float)mn[k]; fl[base + 4 + k] = (float)mx[k]; }
  }
}

}  // namespace orc

namespace orc {

// refitBLAS (bvhAccel.ts:800-870, without displacement): the BVH2 words'
// node boxes recomputed bottom-up from new positions, topology kept. Leaf
// entries are [i0, i1, i2, h]; a leaf's box spans the distinct finite
// vertices of its triangles (getBlasRefitData, bvhAccel.ts:748-797, collects
// them once per leaf), an interior node's the union of its two children
// (left child after the node's two rows, right child at its row index).
// Returns the refitted node surface area, sum of 2 (xz + xy + zy) in the
// reference's postorder.
double refit_blas_words(uint32_t *words, size_t vec4Count, const float *positions, uint32_t vertexCount) {
  if (!vec4Count) return 0.0;
  float *f = reinterpret_cast<float *>(words);
  const int32_t *ints = reinterpret_cast<const int32_t *>(words);
  std::vector<uint32_t> preorder;
  std::vector<uint32_t> stack = {0};
  while (!stack.empty()) {
    const uint32_t row = stack.back();
    stack.pop_back();
    preorder.push_back(row);
    const int32_t child = ints[row * 4 + 3];
    if (child >= 0) {
      stack.push_back((uint32_t)child);
      stack.push_back(row + 2);
    }
  }
  double area = 0.0;
  for (size_t k = preorder.size(); k-- > 0;) {
    const uint32_t row = preorder[k];
    const size_t base = (size_t)row * 4;
    const int32_t child = ints[base + 3];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (child >= 0) {
      const size_t left = base + 8, right = (size_t)child * 4;
      for (int a = 0; a < 3; a++) {
        mn[a] = jmin(f[left + a], f[right + a]);
        mx[a] = jmax(f[left + 4 + a], f[right + 4 + a]);
      }
    } else {
      std::vector<uint32_t> seen;
      for (int32_t i = 0; i < -child; i++)
        for (int j = 0; j < 3; j++) {
          const uint32_t v = words[base + 8 + (size_t)i * 4 + j];
          if (std::find(seen.begin(), seen.end(), v) == seen.end()) seen.push_back(v);
        }
      for (uint32_t v : seen) {
        if (v >= vertexCount) continue;
        const double x = positions[(size_t)v * 3], y = positions[(size_t)v * 3 + 1], z = positions[(size_t)v * 3 + 2];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z)) continue;
        mn[0] = jmin(mn[0], x); mn[1] = jmin(mn[1], y); mn[2] = jmin(mn[2], z);
        mx[0] = jmax(mx[0], x); mx[1] = jmax(mx[1], y); mx[2] = jmax(mx[2], z);
      }
    }
    const double x = (double)(float)mx[0] - (double)(float)mn[0], y = (double)(float)mx[1] - (double)(float)mn[1],
                 z = (double)(float)mx[2] - (double)(float)mn[2];
    area += 2 * (x * z + x * y + z * y);
    for (int a = 0; a < 3; a++) {
      f[base + a] = (float)mn[a];
      f[base + 4 + a] = (float)mx[a];
    }
  }
  return area;
}

}  // namespace orc
