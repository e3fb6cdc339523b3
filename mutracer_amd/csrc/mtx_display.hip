// Display composition and the joint non-local-means (JNLM) denoiser over a
// render target's FP64 accumulators.
//
// compose: the resolved HDR colour (WGPURayTraceOutput.ts:32-92) through
// createComposeMain (compose.frag.wgsl.ts:100-142: exposure, gradeColor,
// vignette, tone mapping, sRGB / gamma 2.2 encoding), without bloom, flat
// backgrounds and shadow catchers.
// JNLM (src/renderer/jnlm.ts, wgsl/jnlm.compute.ts): guide pass (luminance
// mean and variance of the mean from the moments), 3x3 variance smoothing,
// the composed image decoded to display-linear light
// (createCompositeDenoiseColorMain), `iterations` filter passes with dilation
// 2^pass, and the sRGB encoding of jnlmEncodeWGSL.
// All arithmetic is f32 in the order the WGSL writes it (-ffp-contract=off),
// with the library's deterministic pow/exp/log2 (include/mtx_fmath.h), so the
// CPU oracle's restatement returns the same bits. One thread per pixel.
#include <algorithm>
#include <cmath>

#include "../../include/mtx_fmath.h"
#include "mtx_display.h"

namespace mtx {

namespace {

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float a) { return V3{a, a, a}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return V3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return V3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ V3 vclamp(V3 a, float lo, float hi) { return vmin(vmax(a, v3(lo)), v3(hi)); }
__device__ __forceinline__ float dot3(V3 a, V3 b) {
  float r = a.x * b.x;
  r = r + a.y * b.y;
  return r + a.z * b.z;
}
// WGSL mix(a, b, t) = a * (1 - t) + b * t.
__device__ __forceinline__ V3 vmix(V3 a, V3 b, float t) {
  const float u = 1.0f - t;
  return V3{a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t};
}
// Column-major 3x3 times vector: c0 * v.x + c1 * v.y + c2 * v.z.
__device__ __forceinline__ V3 mat3(const float (&m)[9], V3 v) {
  V3 r;
  r.x = m[0] * v.x; r.x = r.x + m[3] * v.y; r.x = r.x + m[6] * v.z;
  r.y = m[1] * v.x; r.y = r.y + m[4] * v.y; r.y = r.y + m[7] * v.z;
  r.z = m[2] * v.x; r.z = r.z + m[5] * v.y; r.z = r.z + m[8] * v.z;
  return r;
}
__device__ __forceinline__ V3 vpow(V3 a, float e) {
  return V3{mtxfm::powf(a.x, e), mtxfm::powf(a.y, e), mtxfm::powf(a.z, e)};
}

__device__ V3 linear_to_srgb(V3 c) {
  const V3 lin = vmax(c, v3(0.0f));
  const V3 low = lin * v3(12.92f);
  const V3 high = v3(1.055f) * vpow(lin, 1.0f / 2.4f) - v3(0.055f);
  return V3{lin.x <= 0.0031308f ? low.x : high.x, lin.y <= 0.0031308f ? low.y : high.y,
            lin.z <= 0.0031308f ? low.z : high.z};
}

__device__ V3 srgb_to_linear(V3 e) {  // createCompositeDenoiseColorMain's decode
  V3 r;
  const float *in = &e.x;
  float *out = &r.x;
  for (int k = 0; k < 3; k++)
    out[k] = in[k] <= 0.04045f ? in[k] / 12.92f : mtxfm::powf((in[k] + 0.055f) / 1.055f, 2.4f);
  return r;
}

__device__ V3 rrt_and_odt_fit(V3 v) {
  const V3 num = v * (v + v3(0.0245786f)) - v3(0.000090537f);
  const V3 den = v * (v3(0.983729f) * v + v3(0.4329510f)) + v3(0.238081f);
  return num / den;
}

__device__ V3 tone_map(V3 c, uint32_t mode) {
  if (mode == MTX_TONE_ACES) {
    const float in[9] = {0.59719f, 0.07600f, 0.02840f, 0.35458f, 0.90834f, 0.13383f, 0.04823f, 0.01566f, 0.83777f};
    const float out[9] = {1.60475f, -0.10208f, -0.00327f, -0.53108f, 1.10813f, -0.07276f, -0.07367f, -0.00605f,
                          1.07602f};
    const V3 color = mat3(in, c / v3(0.6f));
    return vclamp(mat3(out, rrt_and_odt_fit(color)), 0.0f, 1.0f);
  }
  if (mode == MTX_TONE_CINEON) {
    const V3 color = vmax(v3(0.0f), c - v3(0.004f));
    return vpow((color * (v3(6.2f) * color + v3(0.5f))) / (color * (v3(6.2f) * color + v3(1.7f)) + v3(0.06f)), 2.2f);
  }
  if (mode == MTX_TONE_REINHARD) return vclamp(c / (v3(1.0f) + c), 0.0f, 1.0f);
  if (mode == MTX_TONE_LINEAR) return vclamp(c, 0.0f, 1.0f);
  // AgX
  const float rec2020ToSrgb[9] = {1.6605f, -0.1246f, -0.0182f, -0.5876f, 1.1329f, -0.1006f, -0.0728f, -0.0083f, 1.1187f};
  const float srgbToRec2020[9] = {0.6274f, 0.0691f, 0.0164f, 0.3293f, 0.9195f, 0.0880f, 0.0433f, 0.0113f, 0.8956f};
  const float inset[9] = {0.856627153315983f, 0.137318972929847f, 0.11189821299995f,
                          0.0951212405381588f, 0.761241990602591f, 0.0767994186031903f,
                          0.0482516061458583f, 0.101439036467562f, 0.811302368396859f};
  const float outset[9] = {1.1271005818144368f, -0.1413297634984383f, -0.14132976349843826f,
                           -0.11060664309660323f, 1.157823702216272f, -0.11060664309660294f,
                           -0.016493938717834573f, -0.016493938717834257f, 1.2519364065950405f};
  V3 color = mat3(inset, mat3(srgbToRec2020, c));
  const V3 lg = vmax(color, v3(1e-10f));
  color = vclamp((V3{mtxfm::log2f(lg.x), mtxfm::log2f(lg.y), mtxfm::log2f(lg.z)} + v3(12.47393f)) / v3(16.5f), 0.0f,
                 1.0f);
  const V3 v2 = color * color, v4 = v2 * v2;
  V3 p = v3(15.5f) * v4 * v2;
  p = p - v3(40.14f) * v4 * color;
  p = p + v3(31.96f) * v4;
  p = p - v3(6.868f) * v2 * color;
  p = p + v3(0.4298f) * v2;
  p = p + v3(0.1191f) * color;
  p = p - v3(0.00232f);
  color = mat3(outset, p);
  color = vpow(vmax(color, v3(0.0f)), 2.2f);
  return vclamp(mat3(rec2020ToSrgb, color), 0.0f, 1.0f);
}

__device__ V3 grade(V3 c, const DisplayParams &P) {
  const V3 balance{1.0f + P.temperature * 0.16f + P.tint * 0.04f, 1.0f - P.tint * 0.12f,
                   1.0f - P.temperature * 0.16f + P.tint * 0.04f};
  V3 color = vmax(c * balance, v3(0.0f));
  color = vmax((color - v3(0.18f)) * v3(P.contrast) + v3(0.18f), v3(0.0f));
  const float lum = dot3(color, V3{0.2126f, 0.7152f, 0.0722f});
  return vmax(vmix(v3(lum), color, P.saturation), v3(0.0f));
}

__device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
  const float t = fminf(fmaxf((x - e0) / (e1 - e0), 0.0f), 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

__device__ __forceinline__ uint8_t unorm8(float v) {
  const float c = fminf(fmaxf(v, 0.0f), 1.0f);
  return (uint8_t)floorf(c * 255.0f + 0.5f);
}

// Resolved HDR colour of accumulation pixel i (hdr_resolve_component in FP64,
// rounded once).
__device__ float4 resolved_color(const double *__restrict__ pl, size_t N, size_t i) {
  const double count = pl[11 * N + i];
  if (!(count > 0.0)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const double total = fmax(count, 1.0), hit = pl[7 * N + i];
  const double div = hit > 0.0 ? hit : total;
  return make_float4((float)(pl[i] / div), (float)(pl[N + i] / div), (float)(pl[2 * N + i] / div),
                     (float)(pl[3 * N + i] / total));
}

// Per texture-row pixel: the displayed RGBA8, the display-linear colour for
// the denoiser and the albedo / normal guides (texture rows, top first).
__global__ void k_compose(const double *__restrict__ pl, uint32_t W, uint32_t H, DisplayParams P,
                          uint8_t *__restrict__ rgba8, float4 *__restrict__ linear, float4 *__restrict__ albedo,
                          float4 *__restrict__ normal) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t N = (size_t)W * H, i = (size_t)(H - 1 - y) * W + x, o = (size_t)y * W + x;
  const float4 hdr = resolved_color(pl, N, i);
  V3 graded = grade(V3{hdr.x, hdr.y, hdr.z} * v3(P.exposure), P);
  const float dx = (float)x / (float)W - 0.5f, dy = (float)y / (float)H - 0.5f;
  const float edge = smoothstep(0.2f, 0.72f, sqrtf(dx * dx + dy * dy));
  graded = graded * v3(1.0f - P.vignette * edge);
  graded = tone_map(graded, P.toneMapping);
  const V3 enc = P.gamma22 ? vpow(vmax(graded, v3(0.0f)), 1.0f / 2.2f) : linear_to_srgb(graded);
  if (rgba8) {
    rgba8[o * 4] = unorm8(enc.x);
    rgba8[o * 4 + 1] = unorm8(enc.y);
    rgba8[o * 4 + 2] = unorm8(enc.z);
    rgba8[o * 4 + 3] = unorm8(hdr.w);
  }
  if (linear) {
    const V3 e = vmax(enc, v3(0.0f));
    const V3 lin = P.gamma22 ? vpow(e, 2.2f) : srgb_to_linear(e);
    linear[o] = make_float4(lin.x, lin.y, lin.z, hdr.w);
  }
  if (albedo || normal) {
    const double count = pl[11 * N + i];
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), n = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (count > 0.0) {
      const double total = fmax(count, 1.0), hit = pl[7 * N + i];
      a = make_float4(fminf(fmaxf((float)(pl[4 * N + i] / total), 0.0f), 1.0f),
                      fminf(fmaxf((float)(pl[5 * N + i] / total), 0.0f), 1.0f),
                      fminf(fmaxf((float)(pl[6 * N + i] / total), 0.0f), 1.0f),
                      (float)fmin(fmax(hit / total, 0.0), 1.0));
      const double nx = pl[8 * N + i], ny = pl[9 * N + i], nz = pl[10 * N + i];
      if (nx * nx + ny * ny + nz * nz > 0.1)
        n = make_float4((float)(nx / total * 0.5 + 0.5), (float)(ny / total * 0.5 + 0.5),
                        (float)(nz / total * 0.5 + 0.5), 1.0f);
    }
    if (albedo) albedo[o] = a;
    if (normal) normal[o] = n;
  }
}

// Pass 1 (jnlmGuideWGSL): luminance mean and variance of the mean from the
// moments (FP64 sums, rounded once); -1 with a single sample.
__global__ void k_jnlm_guide(const double *__restrict__ pl, uint32_t W, uint32_t H, float2 *__restrict__ raw) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t N = (size_t)W * H, i = (size_t)(H - 1 - y) * W + x;
  const double n = pl[11 * N + i];
  const double sum = 0.2126 * pl[i] + 0.7152 * pl[N + i] + 0.0722 * pl[2 * N + i];
  const double mean = sum / fmax(n, 1.0);
  double variance = -1.0;
  if (n >= 2.0) variance = fmax(pl[12 * N + i] - n * mean * mean, 0.0) / (n * (n - 1.0));
  raw[(size_t)y * W + x] = make_float2((float)mean, (float)variance);
}

// Pass 2 (jnlmVarianceWGSL): 3x3 variance smoothing; single-sample pixels
// take their neighbourhood's spatial variance.
__global__ void k_jnlm_variance(const float2 *__restrict__ raw, uint32_t W, uint32_t H, float2 *__restrict__ guide) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= W || y >= H) return;
  const float2 center = raw[(size_t)y * W + x];
  float vSum = 0.0f, vCount = 0.0f, mSum = 0.0f, m2Sum = 0.0f;
  for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
      const int qx = std::min(std::max((int)x + dx, 0), (int)W - 1), qy = std::min(std::max((int)y + dy, 0), (int)H - 1);
      const float2 v = raw[(size_t)qy * W + qx];
      mSum = mSum + v.x;
      m2Sum = m2Sum + v.x * v.x;
      if (v.y >= 0.0f) {
        vSum = vSum + v.y;
        vCount = vCount + 1.0f;
      }
    }
  float variance;
  if (center.y >= 0.0f) {
    variance = vSum / vCount;
  } else {
    const float sm = mSum / 9.0f;
    variance = fmaxf(m2Sum / 9.0f - sm * sm, 0.0f) * 9.0f / 8.0f;
  }
  guide[(size_t)y * W + x] = make_float2(center.x, variance);
}

__device__ __forceinline__ size_t pix(int x, int y, int W, int H) {
  x = std::min(std::max(x, 0), W - 1);
  y = std::min(std::max(y, 0), H - 1);
  return (size_t)y * W + x;
}

// Pass 3 (createJNLMFilterWGSL): one joint NL-means pass at `step` spacing.
template <int kPatch>
__global__ void k_jnlm_filter(JnlmParams J, const float2 *__restrict__ guide, const float4 *__restrict__ color,
                              const float4 *__restrict__ albedo, const float4 *__restrict__ normal,
                              float4 *__restrict__ outColor, float2 *__restrict__ outGuide, int step) {
  constexpr int kPatchSize = (2 * kPatch + 1) * (2 * kPatch + 1);
  const int W = (int)J.width, H = (int)J.height;
  const int px = (int)(blockIdx.x * blockDim.x + threadIdx.x), py = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (px >= W || py >= H) return;
  const size_t index = pix(px, py, W, H);
  const float4 cA = albedo[index], cN = normal[index];
  if (cN.w <= 0.0f) {  // primary miss (no shadow catchers): kept as is
    outColor[index] = color[index];
    outGuide[index] = guide[index];
    return;
  }
  float2 centerPatch[kPatchSize];
  {
    int k = 0;
    for (int dy = -kPatch; dy <= kPatch; dy++)
      for (int dx = -kPatch; dx <= kPatch; dx++) centerPatch[k++] = guide[pix(px + dx * step, py + dy * step, W, H)];
  }
  const float k2 = J.colorSensitivity * J.colorSensitivity;
  const float floorA = 0.02f;  // JNLM_ALBEDO_FLOOR
  const V3 centerAlbedo{cA.x, cA.y, cA.z}, centerNormal{cN.x, cN.y, cN.z};
  const V3 featureScale = vmax(centerAlbedo, v3(floorA));
  float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  V3 featureSum = v3(0.0f), featureSquareSum = v3(0.0f), featureColorSum = v3(0.0f);
  float guideSum = 0.0f, guideVarianceSum = 0.0f, weightSum = 0.0f, maxWeight = 0.0f;
  const int S = (int)J.searchRadius;
  for (int sy = -S; sy <= S; sy++)
    for (int sx = -S; sx <= S; sx++) {
      const int qx = px + sx * step, qy = py + sy * step;
      if ((sx == 0 && sy == 0) || qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
      const size_t qi = (size_t)qy * W + qx;
      const float4 qa = albedo[qi], qn = normal[qi];
      // featureDistance
      const V3 albedoRgb{qa.x, qa.y, qa.z};
      const V3 albedoScale = vmax((albedoRgb + centerAlbedo) * v3(0.5f), v3(floorA));
      const V3 albedoDelta = (albedoRgb - centerAlbedo) / albedoScale;
      const float coverageDelta = qa.w - cA.w;
      const V3 normalDelta = V3{qn.x, qn.y, qn.z} - centerNormal;
      const float fa = dot3(albedoDelta, albedoDelta) * J.invAlbedoBandwidth2 +
                       coverageDelta * coverageDelta * J.invCoverageBandwidth2;
      const float features = fmaxf(fa, dot3(normalDelta, normalDelta) * J.invNormalBandwidth2);
      if (features > 16.0f) continue;
      float colorDistance = 0.0f;
      int k = 0;
      for (int dy = -kPatch; dy <= kPatch; dy++)
        for (int dx = -kPatch; dx <= kPatch; dx++) {
          const float2 a = centerPatch[k];
          const float2 b = guide[pix(qx + dx * step, qy + dy * step, W, H)];
          const float delta = a.x - b.x;
          colorDistance = colorDistance + (delta * delta - 1.0f * (a.y + fminf(a.y, b.y))) / (1e-10f + k2 * (a.y + b.y));
          k++;
        }
      colorDistance = colorDistance / (float)kPatchSize;
      const float weight = mtxfm::expf(-fmaxf(fmaxf(colorDistance, 0.0f), features));
      const float4 c = color[qi];
      const float2 g = guide[qi];
      const V3 a = (albedoRgb - centerAlbedo) / featureScale;
      const float4 value = make_float4(c.x * c.w, c.y * c.w, c.z * c.w, c.w);
      sum = make_float4(sum.x + value.x * weight, sum.y + value.y * weight, sum.z + value.z * weight,
                        sum.w + value.w * weight);
      featureSum = featureSum + a * v3(weight);
      featureSquareSum = featureSquareSum + a * a * v3(weight);
      featureColorSum = featureColorSum + a * V3{value.x, value.y, value.z} * v3(weight);
      guideSum = guideSum + g.x * weight;
      guideVarianceSum = guideVarianceSum + g.y * weight * weight;
      weightSum = weightSum + weight;
      maxWeight = fmaxf(maxWeight, weight);
    }
  const float4 center = color[index];
  const float2 centerGuide = guide[index];
  const float centerWeight = fmaxf(maxWeight, 0.01f);  // CENTER_WEIGHT_MIN
  const float totalWeight = weightSum + centerWeight;
  const float4 cv = make_float4(center.x * center.w, center.y * center.w, center.z * center.w, center.w);
  const float4 mean = make_float4((sum.x + cv.x * centerWeight) / totalWeight, (sum.y + cv.y * centerWeight) / totalWeight,
                                  (sum.z + cv.z * centerWeight) / totalWeight, (sum.w + cv.w * centerWeight) / totalWeight);
  const V3 meanRgb{mean.x, mean.y, mean.z};
  const V3 tw = v3(totalWeight);
  const V3 meanFeature = featureSum / tw;
  const V3 featureVariance = vmax(featureSquareSum / tw - meanFeature * meanFeature, v3(0.0f));
  const V3 covariance = featureColorSum / tw - meanFeature * meanRgb;
  const V3 meanAlbedo = centerAlbedo + meanFeature * featureScale;
  const V3 maximumSlope = meanRgb * featureScale / vmax(meanAlbedo, v3(floorA));
  const V3 slope = vmin(vmax(covariance / (featureVariance + v3(0.0001f)), v3(0.0f)), maximumSlope);
  const V3 corrected = meanRgb - slope * meanFeature;
  const V3 filtered = vmax(corrected, v3(0.0f));
  const float fa = fmaxf(mean.w, 0.0f);
  const V3 rgb = fa > 1e-6f ? filtered / v3(fmaxf(fa, 1e-6f)) : V3{center.x, center.y, center.z};
  outColor[index] = make_float4(rgb.x, rgb.y, rgb.z, fa);
  outGuide[index] = make_float2((guideSum + centerGuide.x * centerWeight) / totalWeight,
                                (guideVarianceSum + centerGuide.y * centerWeight * centerWeight) /
                                    (totalWeight * totalWeight));
}

// jnlmEncodeWGSL (premultiply off) to RGBA8.
__global__ void k_jnlm_encode(const float4 *__restrict__ c, uint32_t n, uint8_t *__restrict__ rgba8) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = c[i];
  const V3 e = linear_to_srgb(V3{v.x, v.y, v.z});
  rgba8[(size_t)i * 4] = unorm8(e.x);
  rgba8[(size_t)i * 4 + 1] = unorm8(e.y);
  rgba8[(size_t)i * 4 + 2] = unorm8(e.z);
  rgba8[(size_t)i * 4 + 3] = unorm8(v.w);
}

}  // namespace

void display_compose(const double *planes, uint32_t W, uint32_t H, const DisplayParams &P, uint8_t *rgba8,
                     float4 *linear, float4 *albedo, float4 *normal, hipStream_t s) {
  const dim3 block(8, 8), grid((W + 7) / 8, (H + 7) / 8);
  hipLaunchKernelGGL(k_compose, grid, block, 0, s, planes, W, H, P, rgba8, linear, albedo, normal);
  MTX_HIP(hipGetLastError());
}

void jnlm_denoise(const double *planes, uint32_t W, uint32_t H, const DisplayParams &P, const JnlmParams &J,
                  JnlmBuffers &B, float4 *result, uint8_t *rgba8, hipStream_t s) {
  const size_t N = (size_t)W * H;
  if (B.pixels < N) {
    B.linear.alloc(N);
    B.albedo.alloc(N);
    B.normal.alloc(N);
    B.color2.alloc(N);
    B.guide[0].alloc(N);
    B.guide[1].alloc(N);
    B.pixels = N;
  }
  const dim3 block(8, 8), grid((W + 7) / 8, (H + 7) / 8);
  display_compose(planes, W, H, P, nullptr, B.linear.ptr, B.albedo.ptr, B.normal.ptr, s);
  hipLaunchKernelGGL(k_jnlm_guide, grid, block, 0, s, planes, W, H, B.guide[1].ptr);
  hipLaunchKernelGGL(k_jnlm_variance, grid, block, 0, s, B.guide[1].ptr, W, H, B.guide[0].ptr);
  float4 *colors[2] = {B.linear.ptr, B.color2.ptr};
  float2 *guides[2] = {B.guide[0].ptr, B.guide[1].ptr};
  for (uint32_t pass = 0; pass < J.iterations; pass++) {
    JnlmParams Jp = J;
    const int step = 1 << pass;  // jnlmPassDilation
    switch (J.patchRadius) {
      case 0: hipLaunchKernelGGL(k_jnlm_filter<0>, grid, block, 0, s, Jp, guides[0], colors[0], B.albedo.ptr, B.normal.ptr, colors[1], guides[1], step); break;
      case 1: hipLaunchKernelGGL(k_jnlm_filter<1>, grid, block, 0, s, Jp, guides[0], colors[0], B.albedo.ptr, B.normal.ptr, colors[1], guides[1], step); break;
      case 2: hipLaunchKernelGGL(k_jnlm_filter<2>, grid, block, 0, s, Jp, guides[0], colors[0], B.albedo.ptr, B.normal.ptr, colors[1], guides[1], step); break;
      default: hipLaunchKernelGGL(k_jnlm_filter<3>, grid, block, 0, s, Jp, guides[0], colors[0], B.albedo.ptr, B.normal.ptr, colors[1], guides[1], step); break;
    }
    std::swap(colors[0], colors[1]);
    std::swap(guides[0], guides[1]);
  }
  MTX_HIP(hipGetLastError());
  if (rgba8) hipLaunchKernelGGL(k_jnlm_encode, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, colors[0], (uint32_t)N, rgba8);
  if (result) MTX_HIP(hipMemcpyAsync(result, colors[0], N * sizeof(float4), hipMemcpyDeviceToDevice, s));
  MTX_HIP(hipGetLastError());
}

}  // namespace mtx
