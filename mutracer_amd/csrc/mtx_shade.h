// Device-side building blocks of the path tracer shared by the megakernel
// and the wavefront kernels: Owen-scrambled Sobol / ZSobol sampler, OpenPBR
// lobes with the energy tables, light sampling (alias table, environment CDF,
// rect lights), surface completion and the MIS weights. Each function follows
// the reference WGSL it cites; all arithmetic obeys the contract of mtx_math.h.
#pragma once
#include "../../include/mtx_fmath.h"
#include "mtx_pt.h"

namespace mtx {

// ---------------------------------------------------------------- constants
// Constant expressions are folded in f64 and rounded once, as WGSL folds
// abstract-float constants.
#define KF(x) ((float)(x))
constexpr float kQuarterPi = KF(0.25 * 3.14159265359);
constexpr float kHalfPi = KF(0.5 * 3.14159265359);
constexpr float kAlphaMin = KF(0.001 * 0.001);
constexpr float kFonA = KF(0.5 - 2.0 / (3.0 * 3.14159265359));
constexpr float kFonB = KF(2.0 / 3.0 - 28.0 / (15.0 * 3.14159265359));
constexpr float kEdgeCos = KF(1.0 / 7.0);
constexpr float kEpsLo = 0.0005f;
constexpr float kEpsScale = KF(1.0 - 2.0 * 0.0005);
constexpr int kBlock = 128;

enum { kLutMetal = 1056, kLutMetalAvg = 1088, kLutDiel = 1089, kLutDielAvg = 2113, kLutRatio = 2145 };

__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
__device__ __forceinline__ f3 vpow(f3 a, f3 b) { return f3{mtxfm::powf(a.x, b.x), mtxfm::powf(a.y, b.y), mtxfm::powf(a.z, b.z)}; }
__device__ __forceinline__ f3 vmaxs(f3 a, float s) { return vmax(a, splat(s)); }
__device__ __forceinline__ f3 clamp3(f3 v, float lo, float hi) {
  return f3{clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)};
}
__device__ __forceinline__ f3 reflect3(f3 e1, f3 e2) { return e1 - (2.0f * dot(e2, e1)) * e2; }
__device__ __forceinline__ f3 refract3(f3 e1, f3 e2, float e3) {
  float c = dot(e2, e1);
  float k = 1.0f - e3 * e3 * (1.0f - c * c);
  if (k < 0.0f) return splat(0.0f);
  return e3 * e1 - (e3 * c + sqrtf(k)) * e2;
}

// ---------------------------------------------------------------- sampler
// Burley 2020 hash-based Owen scrambling of the 2D Sobol sequence
// (sobol.wgsl.ts:23-52, 107-168); the second dimension uses the five-step
// Pascal butterfly of sampling.ts:31-34.
__device__ __forceinline__ uint32_t shash(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  return x ^ (x >> 16);
}
__device__ __forceinline__ uint32_t sseed(uint32_t a, uint32_t b) { return shash(a ^ shash(b + 0x9e3779b9u)); }
__device__ __forceinline__ uint32_t lkperm(uint32_t x, uint32_t s) {
  x += s;
  x ^= x * 0x6c50b47cu; x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u; x ^= x * 0x8d22f6e6u;
  return x;
}
__device__ __forceinline__ uint32_t owen_scramble(uint32_t x, uint32_t s) {
  return __builtin_bitreverse32(lkperm(__builtin_bitreverse32(x), s));
}
__device__ __forceinline__ uint32_t sobol_second(uint32_t v) {
  uint32_t r = __builtin_bitreverse32(v);
  r ^= (r << 1) & 0xaaaaaaaau;
  r ^= (r << 2) & 0xccccccccu;
  r ^= (r << 4) & 0xf0f0f0f0u;
  r ^= (r << 8) & 0xff00ff00u;
  r ^= (r << 16) & 0xffff0000u;
  return r;
}
__device__ __forceinline__ uint32_t morton8(uint32_t x, uint32_t y) {
  x &= 255u; y &= 255u;
  x = (x | (x << 4)) & 0x0f0fu; y = (y | (y << 4)) & 0x0f0fu;
  x = (x | (x << 2)) & 0x3333u; y = (y | (y << 2)) & 0x3333u;
  x = (x | (x << 1)) & 0x5555u; y = (y | (y << 1)) & 0x5555u;
  return x | (y << 1);
}
// The 24 base-4 digit permutations of pbrt-v4's ZSobolSampler, four 8-bit
// codes (digit i in bits 2i..2i+1) per word (sampling.ts:20-33).
__device__ __forceinline__ uint32_t zperm(uint32_t p) {
  const uint32_t w = p >> 2;
  const uint32_t word = w == 0u ? 0x78d8b4e4u : w == 1u ? 0xb1e19c6cu : w == 2u ? 0x8d2d39c9u
                      : w == 3u ? 0x72d236c6u : w == 4u ? 0x87271e4eu : 0x93634b1bu;
  return (word >> ((p & 3u) * 8u)) & 255u;
}
__device__ uint32_t zsobol_index(uint32_t morton, uint32_t dim, uint32_t logSPP) {
  uint32_t odd = logSPP & 1u;
  int digits = 8 + (int)((logSPP + 1u) >> 1);
  uint32_t result = 0;
  for (int digit = digits - 1; digit >= (int)odd; digit--) {
    uint32_t shift = 2u * (uint32_t)digit - odd;
    uint32_t higher = shift < 30u ? morton >> (shift + 2u) : 0u;
    uint32_t perm = (shash(higher ^ (0x55555555u * dim)) >> 24) % 24u;
    uint32_t d = (morton >> shift) & 3u;
    result |= ((zperm(perm) >> (2u * d)) & 3u) << shift;
  }
  if (odd) result |= (morton & 1u) ^ (shash((morton >> 1) ^ (0x55555555u * dim)) & 1u);
  return result;
}

// pcg4d (rand.wgsl.ts:13-19; Jarzynski and Olano, "Hash Functions for GPU
// Rendering", JCGT 9(3), 2020): an LCG step on each lane of a 4 x u32 state,
// two rounds of cross-lane multiply-adds around a 16-bit xorshift. Every step
// is invertible, so the map is a bijection of the 128-bit state.
__device__ __forceinline__ void pcg4d(uint4 &v) {
  v.x = v.x * 1664525u + 1013904223u; v.y = v.y * 1664525u + 1013904223u;
  v.z = v.z * 1664525u + 1013904223u; v.w = v.w * 1664525u + 1013904223u;
  v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
  v.x ^= v.x >> 16; v.y ^= v.y >> 16; v.z ^= v.z >> 16; v.w ^= v.w >> 16;
  v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
}

// Radical inverse of `index` in `base` in FP64 with halton.ts:6-15's
// operation order; the random pattern's camera jitter (WGPURenderingPipeline
// .ts:534-541) is (halton(k, 2) - 0.5) / width, (halton(k, 3) - 0.5) / height
// for pixel sample k, rounded to f32.
__device__ __forceinline__ double halton(uint32_t index, uint32_t base) {
  double result = 0.0, f = 1.0 / (double)base;
  for (uint32_t i = index; i > 0u; i /= base) {
    result = result + f * (double)(i % base);
    f = f / (double)base;
  }
  return result;
}

constexpr uint32_t kPatternRandom = 2u;

// Sample streams of one path. Patterns 0 and 1 address Owen-scrambled
// Sobol / ZSobol points by dimension; pattern 2 (random) draws from the
// path's pcg4d state in call order, so `rng` (the path's 16-byte state in
// device memory or in the kernel's registers) advances on every draw.
struct Smp {
  uint32_t px, py, width, index, seed, logSPP, pattern, bounce;
  uint4 *rng;  // pattern 2 only
  // rng_initialize (rand.wgsl.ts:8-11) as called at rayTrace.compute.ts:1193.
  __device__ void seed_rng() const { *rng = make_uint4(px, py, index ^ shash(seed), px + py); }
  // rand (rand.wgsl.ts:21-23): f32(s0.x) / f32(0xffffffff), whose divisor
  // rounds to 2^32, so the result lies in [0, 1].
  __device__ float rnd() const {
    uint4 v = *rng;
    pcg4d(v);
    *rng = v;
    return (float)v.x * (1.0f / 4294967296.0f);
  }
  // hashedSample (sobol.wgsl.ts:158-168) where its value cannot matter: the
  // alpha-coverage test of completeVertex (rayTrace.compute.ts:533) against
  // an opaque material. Only the random pattern's stream moves.
  __device__ void hashed_draw() const {
    if (pattern == kPatternRandom) (void)rnd();
  }
  __device__ uint32_t bdim(uint32_t slot) const { return 3u + bounce * 64u + slot; }
  __device__ uint32_t ldim(uint32_t cand, uint32_t slot) const {
    if (cand == 0u) return slot == 2u ? bdim(8u) : bdim(slot);
    return bdim(9u + (cand - 1u) * 3u + slot);
  }
  __device__ void raw(uint32_t dim, float &x, float &y, bool needY) const {
    uint32_t pixelID = px + py * width;
    uint32_t s = sseed(sseed(pixelID, dim), seed);
    uint32_t idx = index;
    if (pattern == 1u) {
      uint32_t lg = logSPP;
      if (idx == 0u) lg = 0u; else idx -= 1u;
      uint32_t block = idx >> lg;
      uint32_t mask = (1u << lg) - 1u;
      uint32_t m = (morton8(px, py) << lg) | (idx & mask);
      idx = zsobol_index(m, dim, lg);
      s = sseed(sseed(dim, seed), sseed(px >> 8, py >> 8));
      if (index != 0u) block += 1u;
      s = sseed(s, block);
    }
    uint32_t i = owen_scramble(idx, sseed(s, 0u));
    uint32_t xb = __builtin_bitreverse32(lkperm(i, sseed(s, 1u)));
    x = (float)(xb >> 8) * (1.0f / 16777216.0f);
    if (needY) {
      uint32_t yb = owen_scramble(sobol_second(i), sseed(s, 2u));
      y = (float)(yb >> 8) * (1.0f / 16777216.0f);
    }
  }
  __device__ float s1(uint32_t dim) const {
    if (pattern == kPatternRandom) return kEpsLo + kEpsScale * rnd();
    float x, y;
    raw(dim, x, y, false);
    return kEpsLo + kEpsScale * x;
  }
  __device__ void s2(uint32_t dim, float &a, float &b) const {
    if (pattern == kPatternRandom) {  // two sequential draws, x first (sobol.wgsl.ts:150-156)
      a = s1(dim);
      b = s1(dim);
      return;
    }
    float x, y;
    raw(dim, x, y, true);
    a = kEpsLo + kEpsScale * x;
    b = kEpsLo + kEpsScale * y;
  }
  // lightCandidateRandom (rayTrace.compute.ts:251-261): the random pattern
  // discards one draw to keep its stream, the value is always the hash.
  __device__ float candidate_random(uint32_t salt) const {
    if (pattern == kPatternRandom) (void)rnd();
    uint32_t pixelID = px + py * width;
    uint32_t key = sseed(sseed(pixelID, index), seed);
    uint32_t dimension = sseed(bounce, salt) ^ 0xa511e9b3u;
    return (float)(shash(key ^ dimension) >> 8) * (1.0f / 16777216.0f);
  }
};

// ---------------------------------------------------------------- BSDF math
__device__ f3 ggx_vndf(f3 view, float ax, float ay, float sx, float sy) {
  f3 st{ax, ay, 1.0f};
  f3 sv = normalize(st * view);
  float vz = mnf(sv.z, 1.0f);
  float span = 1.0f + vz;
  float nz = (1.0f - sy) * span, omz = sy * span, opz = nz + (1.0f - vz);
  float radial = sqrtf(mxf(0.0f, opz * omz));
  float az = kTwoPi * sx;
  const mtxfm::SinCos a = mtxfm::sincosf(az);
  f3 sn{sv.x + radial * a.c, sv.y + radial * a.s, nz};
  return normalize(st * sn);
}
__device__ __forceinline__ float ggx_ndf(f3 h, float ax, float ay) {
  f3 e{h.x / ax, h.y / ay, h.z};
  float den = dot(e, e);
  return 1.0f / (kPi * ax * ay * den * den);
}
__device__ __forceinline__ float smith_g1(f3 d, float ax, float ay) {
  float nc = fabsf(d.z);
  if (nc <= 1e-8f) return 0.0f;
  float rx = ax * d.x, ry = ay * d.y;
  float tan2 = (rx * rx + ry * ry) / (nc * nc);
  return 2.0f / (1.0f + sqrtf(1.0f + tan2));
}
__device__ __forceinline__ float schlick5(float c) { return mtxfm::powf(clampf(1.0f - c, 0.0f, 1.0f), 5.0f); }
__device__ float fresnel_dielectric(float incidentCos, float eta) {
  float cosI = clampf(fabsf(incidentCos), 0.0f, 1.0f);
  float sinT2 = eta * eta * mxf(0.0f, 1.0f - cosI * cosI);
  if (sinT2 >= 1.0f) return 1.0f;
  float cosT = sqrtf(mxf(0.0f, 1.0f - sinT2));
  float perp = (cosI - eta * cosT) / mxf(cosI + eta * cosT, 1e-8f);
  float par = (eta * cosI - cosT) / mxf(eta * cosI + cosT, 1e-8f);
  return 0.5f * (perp * perp + par * par);
}
__device__ f3 cosine_hemisphere(float sx, float sy) {
  float x = 2.0f * sx - 1.0f, y = 2.0f * sy - 1.0f, dx = 0.0f, dy = 0.0f;
  if (!(x == 0.0f && y == 0.0f)) {
    float r, a;
    if (fabsf(x) > fabsf(y)) { r = x; a = kQuarterPi * (y / x); }
    else { r = y; a = kHalfPi - kQuarterPi * (x / y); }
    const mtxfm::SinCos sc = mtxfm::sincosf(a);
    dx = r * sc.c;
    dy = r * sc.s;
  }
  return f3{dx, dy, sqrtf(mxf(0.0f, 1.0f - (dx * dx + dy * dy)))};
}
__device__ __forceinline__ float power_heuristic(float a, float b) {
  float scale = mxf(a, b);
  if (scale <= 0.0f) return 0.0f;
  float na = a / scale, nb = b / scale;
  float sa = na * na, sb = nb * nb;
  return sa / (sa + sb);
}

struct Basis { f3 t, b, n; };
__device__ __forceinline__ Basis make_basis(f3 n) {
  float signZ = n.z >= 0.0f ? 1.0f : -1.0f;
  float factor = -1.0f / (signZ + n.z);
  float xy = n.x * n.y * factor;
  Basis B;
  B.t = f3{1.0f + signZ * n.x * n.x * factor, signZ * xy, -signZ * n.x};
  B.b = f3{xy, signZ + n.y * n.y * factor, -n.y};
  B.n = n;
  return B;
}
__device__ __forceinline__ f3 to_world(const Basis &B, f3 d) { return d.x * B.t + d.y * B.b + d.z * B.n; }
__device__ __forceinline__ f3 to_local(const Basis &B, f3 d) { return f3{dot(d, B.t), dot(d, B.b), dot(d, B.n)}; }

// ---------------------------------------------------------------- material textures
// wrapMapCoordinate (materialBuffer.wgsl.ts:285-289): repeat, clamp, mirror.
__device__ __forceinline__ float wrap_coord(float v, uint32_t mode) {
  if (mode == MTX_WRAP_REPEAT) return fractf(v);
  if (mode == MTX_WRAP_CLAMP_TO_EDGE) return clampf(v, 0.0f, 1.0f);
  return 1.0f - fabsf(1.0f - (v - 2.0f * floorf(v * 0.5f)));
}
__device__ __forceinline__ float4 map_texel(const GpuScene &S, const GpuTexture &t, int x, int y) {
  const size_t i = (size_t)t.base + (size_t)(uint32_t)y * t.width + (uint32_t)x;
  if (t.flags & 1u) return S.texels32[i];
  const uint32_t w = S.texels8[i];
  return make_float4((float)(w & 255u) / 255.0f, (float)((w >> 8) & 255u) / 255.0f,
                     (float)((w >> 16) & 255u) / 255.0f, (float)(w >> 24) / 255.0f);
}
__device__ __forceinline__ int map_tap(int c, int size, uint32_t mode) {
  return mode == MTX_WRAP_REPEAT ? ((c % size) + size) % size : min(max(c, 0), size - 1);
}
__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
  return make_float4(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t));
}
// sampleMap + loadMapTexel (materialBuffer.wgsl.ts:314-348) for an ordinary
// image: the coordinate uv * scale + offset is wrapped per axis, and the
// texels around (wrapped * size - 0.5) are blended bilinearly in f32, taps
// past the edge repeating or clamping by the axis' mode. (The reference
// leaves interior footprints to the hardware sampler, whose weights carry
// fewer fraction bits; its explicit four-tap form is the one used here for
// every footprint.) A NaN coordinate (an infinite uv) samples at 0.
__device__ float4 sample_map(const GpuScene &S, uint32_t index, float2 uv) {
  const GpuTexture t = S.textures[index];
  const uint32_t ms = (t.flags >> 1) & 3u, mt = (t.flags >> 3) & 3u;
  float wu = wrap_coord(uv.x * t.sx + t.ox, ms), wv = wrap_coord(uv.y * t.sy + t.oy, mt);
  wu = wu == wu ? wu : 0.0f;  // the wrapped value lies in [0, 1] unless it is NaN
  wv = wv == wv ? wv : 0.0f;
  const int W = (int)t.width, H = (int)t.height;
  const float tx = wu * (float)W - 0.5f, ty = wv * (float)H - 0.5f;
  const float fx = floorf(tx), fy = floorf(ty);
  const int bx = (int)fx, by = (int)fy;
  const float ax = tx - fx, ay = ty - fy;
  const int x0 = map_tap(bx, W, ms), x1 = map_tap(bx + 1, W, ms);
  const int y0 = map_tap(by, H, mt), y1 = map_tap(by + 1, H, mt);
  return mix4(mix4(map_texel(S, t, x0, y0), map_texel(S, t, x1, y0), ax),
              mix4(map_texel(S, t, x0, y1), map_texel(S, t, x1, y1), ax), ay);
}
// sRGBToLinear (materialBuffer.wgsl.ts:258-260); decodeMapColor (:262-265)
// leaves RGBA32F (linear) textures as they are.
__device__ __forceinline__ float srgb_to_linear(float c) {
  return c < 0.04045f ? c * 0.0773993808f : mtxfm::powf(c * 0.9478672986f + 0.0521327014f, 2.4f);
}
__device__ __forceinline__ f3 map_color(const GpuScene &S, uint32_t index, float4 c) {
  if (S.textures[index].flags & 1u) return f3{c.x, c.y, c.z};
  return f3{srgb_to_linear(c.x), srgb_to_linear(c.y), srgb_to_linear(c.z)};
}
// Texture of slot k of material m (index + 1 stored; 0 = none), sampled at
// its UV set (getMapUvSet, materialBuffer.wgsl.ts:303-305).
__device__ __forceinline__ uint32_t map_slot(const float *__restrict__ m, int k) {
  return (uint32_t)m[MTX_MAT_MAPS + k];
}
__device__ __forceinline__ float4 map_at(const GpuScene &S, uint32_t slot, float2 uv, float2 uv1) {
  return sample_map(S, slot - 1u, (S.textures[slot - 1u].flags & 32u) ? uv1 : uv);
}
// getMaterialEmission (materialBuffer.wgsl.ts:351-372): the packed emission
// times the emissive map; a coat weight map moves the coat tint, which the
// packed emission then leaves out, to the sample.
__device__ f3 material_emission(const GpuScene &S, const float *__restrict__ m, float2 uv, float2 uv1) {
  f3 e{m[MTX_MAT_EMISSION], m[MTX_MAT_EMISSION + 1], m[MTX_MAT_EMISSION + 2]};
  if (const uint32_t t = map_slot(m, MTX_MAP_EMISSIVE)) e = e * map_color(S, t - 1u, map_at(S, t, uv, uv1));
  if (const uint32_t t = map_slot(m, MTX_MAP_COAT_WEIGHT)) {
    const float w = clampf(m[MTX_MAT_COAT_WEIGHT], 0.0f, 1.0f) * map_at(S, t, uv, uv1).x;
    const f3 c{sqrtf(mxf(m[MTX_MAT_COAT_COLOR], 0.0f)), sqrtf(mxf(m[MTX_MAT_COAT_COLOR + 1], 0.0f)),
               sqrtf(mxf(m[MTX_MAT_COAT_COLOR + 2], 0.0f))};
    e = e * f3{mixf(1.0f, c.x, w), mixf(1.0f, c.y, w), mixf(1.0f, c.z, w)};
  }
  return e;
}

// ---------------------------------------------------------------- materials
struct Mat {
  f3 baseColor, specularColor, transmissionColor, coatColor, emission, fuzzColor, subsurfaceColor;
  float baseWeight, diffuseRoughness, baseMetalness, specularWeight, specularRoughness, specularIor;
  float transmissionWeight, coatWeight, coatRoughness, coatIor, coatDarkening, fuzzWeight, fuzzRoughness;
  float subsurfaceWeight;
  float relativeIor, alphaX, alphaY;
  float alpha;  // coverage after the alpha mode: 1 opaque; mask 0 or 1
  bool inside, thinWalled;
};

__device__ __forceinline__ void rough_alpha(float r, float aniso, float &ax, float &ay) {
  float a = r * r;
  if (aniso == 0.0f) { ax = ay = mxf(kAlphaMin, a); return; }
  float c = 1.0f - aniso;
  float at = a * sqrtf(2.0f / (1.0f + c * c));
  ax = mxf(kAlphaMin, at);
  ay = mxf(kAlphaMin, c * at);
}

// Texture samples of getMaterial (materialBuffer.wgsl.ts:78-160), in its
// order, on the clamped record values and before the coat broadening: the
// base map (color; map alpha as transmission or as coverage), the alpha map
// (coverage *= g), transmission, metallic-roughness, roughness, metalness,
// coat weight and roughness, fuzz color and roughness, specular weight and
// color. The coverage samples multiply ma.alpha before its alpha-mode
// resolution in load_material.
__device__ void apply_maps(const GpuScene &S, const float *__restrict__ m, float2 uv, float2 uv1, Mat &ma) {
  ma.emission = material_emission(S, m, uv, uv1);
  if (const uint32_t t = map_slot(m, MTX_MAP_BASE)) {
    const float4 c = map_at(S, t, uv, uv1);
    ma.baseColor = ma.baseColor * map_color(S, t - 1u, c);
    if (m[MTX_MAT_MAP_ALPHA_USAGE] == 3.0f) ma.transmissionWeight = 1.0f - c.w;
    else if (m[MTX_MAT_MAP_ALPHA_USAGE] == 2.0f) ma.alpha *= c.w;
  }
  if (const uint32_t t = map_slot(m, MTX_MAP_ALPHA)) ma.alpha *= map_at(S, t, uv, uv1).y;
  if (const uint32_t t = map_slot(m, MTX_MAP_TRANSMISSION)) ma.transmissionWeight *= map_at(S, t, uv, uv1).x;
  if (const uint32_t t = map_slot(m, MTX_MAP_METALLIC_ROUGHNESS)) {
    const float4 c = map_at(S, t, uv, uv1);
    ma.baseMetalness *= c.z;
    ma.specularRoughness *= c.y;
  }
  if (const uint32_t t = map_slot(m, MTX_MAP_ROUGHNESS)) ma.specularRoughness *= map_at(S, t, uv, uv1).y;
  if (const uint32_t t = map_slot(m, MTX_MAP_METALNESS)) ma.baseMetalness *= map_at(S, t, uv, uv1).z;
  if (const uint32_t t = map_slot(m, MTX_MAP_COAT_WEIGHT)) ma.coatWeight *= map_at(S, t, uv, uv1).x;
  if (const uint32_t t = map_slot(m, MTX_MAP_COAT_ROUGHNESS)) ma.coatRoughness *= map_at(S, t, uv, uv1).y;
  if (const uint32_t t = map_slot(m, MTX_MAP_FUZZ_COLOR))
    ma.fuzzColor = ma.fuzzColor * map_color(S, t - 1u, map_at(S, t, uv, uv1));
  if (const uint32_t t = map_slot(m, MTX_MAP_FUZZ_ROUGHNESS)) ma.fuzzRoughness *= map_at(S, t, uv, uv1).w;
  if (const uint32_t t = map_slot(m, MTX_MAP_SPECULAR_WEIGHT)) ma.specularWeight *= map_at(S, t, uv, uv1).w;
  if (const uint32_t t = map_slot(m, MTX_MAP_SPECULAR_COLOR))
    ma.specularColor = ma.specularColor * map_color(S, t - 1u, map_at(S, t, uv, uv1));
}

// kFull = false: the scene has no fuzz, subsurface, thin-walled or textured
// material (its kernels are compiled without those lobes and the maps).
// S: the scene whose maps apply at (uv, uv1), or null.
template <bool kFull>
__device__ Mat load_material(const float *__restrict__ m, const float col[3], const GpuScene *S = nullptr,
                             float2 uv = float2{0.0f, 0.0f}, float2 uv1 = float2{0.0f, 0.0f}) {
  Mat ma;
  ma.baseColor = f3{m[1], m[2], m[3]};
  ma.specularRoughness = m[10];
  ma.baseMetalness = m[5];
  float aniso = clampf(m[27], 0.0f, 1.0f);
  ma.inside = false;
  ma.specularIor = m[11];
  ma.transmissionWeight = m[12];
  ma.emission = f3{m[23], m[24], m[25]};
  ma.baseWeight = clampf(m[0], 0.0f, 1.0f);
  ma.diffuseRoughness = clampf(m[4], 0.0f, 1.0f);
  ma.specularWeight = mxf(m[6], 0.0f);
  ma.specularColor = vmaxs(f3{m[7], m[8], m[9]}, 0.0f);
  ma.transmissionColor = vmaxs(f3{m[13], m[14], m[15]}, 0.0f);
  ma.coatWeight = clampf(m[16], 0.0f, 1.0f);
  ma.coatColor = vmaxs(f3{m[17], m[18], m[19]}, 0.0f);
  ma.coatRoughness = clampf(m[20], kRoughnessMin, 1.0f);
  ma.coatIor = mxf(m[21], 1.0f);
  ma.coatDarkening = clampf(m[22], 0.0f, 1.0f);
  ma.fuzzWeight = kFull ? clampf(m[MTX_MAT_FUZZ_WEIGHT], 0.0f, 1.0f) : 0.0f;
  ma.fuzzColor = vmaxs(f3{m[MTX_MAT_FUZZ_COLOR], m[MTX_MAT_FUZZ_COLOR + 1], m[MTX_MAT_FUZZ_COLOR + 2]}, 0.0f);
  ma.fuzzRoughness = clampf(m[MTX_MAT_FUZZ_ROUGHNESS], 0.0f, 1.0f);
  ma.subsurfaceWeight = kFull ? clampf(m[MTX_MAT_SUBSURFACE_WEIGHT], 0.0f, 1.0f) : 0.0f;
  ma.subsurfaceColor = vmaxs(f3{m[MTX_MAT_SUBSURFACE_COLOR], m[MTX_MAT_SUBSURFACE_COLOR + 1],
                                m[MTX_MAT_SUBSURFACE_COLOR + 2]}, 0.0f);
  ma.thinWalled = kFull && m[MTX_MAT_THIN_WALLED] > 0.5f;
  ma.alpha = m[MTX_MAT_OPACITY];
  if (kFull && S != nullptr && S->textures != nullptr) apply_maps(*S, m, uv, uv1, ma);
  // alpha mode (materialBuffer.wgsl.ts:69-77): opaque covers fully; mask
  // covers where the coverage reaches the cutoff. Blend materials are
  // refused when the scene is created.
  ma.alpha = m[MTX_MAT_ALPHA_MODE] == 1.0f ? (ma.alpha >= m[MTX_MAT_ALPHA_CUTOFF] ? 1.0f : 0.0f) : 1.0f;
  if (ma.coatWeight > 0.0f && ma.coatRoughness > 0.0f) {
    float b2 = ma.specularRoughness * ma.specularRoughness, c2 = ma.coatRoughness * ma.coatRoughness;
    ma.specularRoughness = mixf(ma.specularRoughness, mtxfm::powf(mnf(1.0f, b2 * b2 + 2.0f * c2 * c2), 0.25f), ma.coatWeight);
  }
  ma.baseColor = ma.baseColor * f3{col[0], col[1], col[2]};
  ma.specularRoughness = clampf(ma.specularRoughness, kRoughnessMin, 1.0f);
  ma.baseMetalness = clampf(ma.baseMetalness, 0.0f, 1.0f);
  rough_alpha(ma.specularRoughness, aniso, ma.alphaX, ma.alphaY);
  ma.relativeIor = 1.0f;
  return ma;
}

// ---------------------------------------------------------------- energy tables
struct Lut {
  const float *__restrict__ d;
  __device__ __forceinline__ float at(int col, int row) const { return d[row * 32 + col]; }
  __device__ float tri(int base, float eI, float eA, float eC) const {
    float iI = clampf(eI, 0.0f, 31.0f), aI = clampf(eA, 0.0f, 31.0f), cI = clampf(eC, 0.0f, 31.0f);
    int i0 = (int)floorf(iI), a0 = (int)floorf(aI), c0 = (int)floorf(cI);
    int i1 = min(i0 + 1, 31), a1 = min(a0 + 1, 31), c1 = min(c0 + 1, 31);
    float im = iI - (float)i0, am = aI - (float)a0, cm = cI - (float)c0;
    float v000 = at(c0, base + i0 * 32 + a0), v001 = at(c1, base + i0 * 32 + a0);
    float v010 = at(c0, base + i0 * 32 + a1), v011 = at(c1, base + i0 * 32 + a1);
    float v100 = at(c0, base + i1 * 32 + a0), v101 = at(c1, base + i1 * 32 + a0);
    float v110 = at(c0, base + i1 * 32 + a1), v111 = at(c1, base + i1 * 32 + a1);
    float v00 = mixf(v000, v100, im), v01 = mixf(v001, v101, im);
    float v10 = mixf(v010, v110, im), v11 = mixf(v011, v111, im);
    return mixf(mixf(v00, v10, am), mixf(v01, v11, am), cm);
  }
  __device__ float bi(int base, float eI, float eA) const {
    float iI = clampf(eI, 0.0f, 31.0f), aI = clampf(eA, 0.0f, 31.0f);
    int i0 = (int)floorf(iI), a0 = (int)floorf(aI);
    int i1 = min(i0 + 1, 31), a1 = min(a0 + 1, 31);
    float im = iI - (float)i0, am = aI - (float)a0;
    float v00 = at(a0, base + i0), v01 = at(a1, base + i0), v10 = at(a0, base + i1), v11 = at(a1, base + i1);
    return mixf(mixf(v00, v10, im), mixf(v01, v11, im), am);
  }
};
__device__ __forceinline__ float f0_ior(float e) { float v = (e - 1.0f) / (e + 1.0f); return v * v; }
__device__ __forceinline__ float ior_index(float ior) {
  return ior < 1.0f ? 15.0f - ((1.0f / ior - 1.0f) / 1.5f) * 15.0f : 16.0f + ((ior - 1.0f) / 1.5f) * 15.0f;
}
__device__ __forceinline__ float extrapolate(float v, float ior) {
  if (ior > 2.5f || ior < 0.4f) {
    float maxF0 = f0_ior(2.5f);
    float progress = (f0_ior(ior) - maxF0) / (1.0f - maxF0);
    return (1.0f - progress) * v;
  }
  return v;
}
__device__ __forceinline__ float opaque_ec(const Lut &L, float ior, float a, float c) {
  return extrapolate(L.tri(0, ior_index(ior), sqrtf(a) * 31.0f, c * 31.0f), ior);
}
__device__ __forceinline__ float opaque_avg_ec(const Lut &L, float ior, float a) {
  return extrapolate(L.bi(1024, ior_index(ior), sqrtf(a) * 31.0f), ior);
}
__device__ __forceinline__ float ideal_loss(const Lut &L, float ior, float a, float c) {
  return L.tri(kLutDiel, ior_index(ior), sqrtf(a) * 31.0f, c * 31.0f);
}
__device__ __forceinline__ float ideal_avg(const Lut &L, float ior, float a) { return L.bi(kLutDielAvg, ior_index(ior), sqrtf(a) * 31.0f); }
__device__ __forceinline__ float ideal_ratio(const Lut &L, float ior, float a) { return L.bi(kLutRatio, ior_index(ior), sqrtf(a) * 31.0f); }
__device__ float metal_ec(const Lut &L, float a, float c) {
  float rI = clampf(sqrtf(a) * 31.0f, 0.0f, 31.0f), cI = clampf(c * 31.0f, 0.0f, 31.0f);
  int r0 = (int)floorf(rI), c0 = (int)floorf(cI);
  int r1 = min(r0 + 1, 31), c1 = min(c0 + 1, 31);
  float fc = fractf(cI), fr = fractf(rI);
  return mixf(mixf(L.at(c0, kLutMetal + r0), L.at(c1, kLutMetal + r0), fc),
              mixf(L.at(c0, kLutMetal + r1), L.at(c1, kLutMetal + r1), fc), fr);
}
__device__ float metal_avg_ec(const Lut &L, float a) {
  float idx = clampf(sqrtf(a) * 31.0f, 0.0f, 31.0f);
  int i0 = (int)floorf(idx);
  return mixf(L.at(i0, kLutMetalAvg), L.at(min(i0 + 1, 31), kLutMetalAvg), fractf(idx));
}

__device__ __forceinline__ float lum_pbr(f3 c) { return dot(c, f3{0.212671f, 0.715160f, 0.072169f}); }
__device__ __forceinline__ float spec_weighted_ior(float eta, float w) {
  float f0 = mnf(f0_ior(eta) * w, 0.9999f);
  float s = sqrtf(f0);
  float ext = (1.0f + s) / (1.0f - s);
  return eta < 1.0f ? 1.0f / ext : ext;
}
__device__ __forceinline__ float pbr_fresnel(float c, float rel) {
  return fabsf(rel - 1.0f) < 1e-6f ? 0.0f : fresnel_dielectric(c, rel);
}
__device__ f3 metal_fresnel(float c, f3 f0, f3 tint) {
  f3 schlick = mix3(f0, splat(1.0f), schlick5(c));
  if (tint.x == 1.0f && tint.y == 1.0f && tint.z == 1.0f) return schlick;
  float edgeComp = 1.0f - kEdgeCos;
  f3 edgeF = mix3(f0, splat(1.0f), schlick5(kEdgeCos));
  float comp = 1.0f - c;
  float correction = c * schlick5(c) * comp / (kEdgeCos * mtxfm::powf(edgeComp, 6.0f));
  return clamp3(schlick - correction * (edgeF - tint * edgeF), 0.0f, 1.0f);
}
__device__ f3 eval_diffuse(const Mat &m, f3 color, f3 v, f3 l, float comp, float &pdf) {
  pdf = 0.0f;
  if (l.z <= 0.0f) return splat(0.0f);
  float r = m.diffuseRoughness, muI = l.z, muO = v.z;
  float dir = dot(l, v) - muI * muO;
  float dot_ = dir > 0.0f ? dir / mxf(mxf(muI, muO), 1e-7f) : dir;
  float ssA = 1.0f / (1.0f + kFonA * r);
  f3 single = color * kInvPi * ssA * (1.0f + r * dot_);
  float ic = 1.0f - muI, oc = 1.0f - muO;
  float iG = ic * (0.0571085289f + ic * (0.491881867f + ic * (-0.332181442f + ic * 0.0714429953f)));
  float oG = oc * (0.0571085289f + oc * (0.491881867f + oc * (-0.332181442f + oc * 0.0714429953f)));
  float iAlb = (1.0f + r * iG) / (1.0f + kFonA * r);
  float oAlb = (1.0f + r * oG) / (1.0f + kFonA * r);
  float avg = ssA * (1.0f + kFonB * r);
  f3 msAlb = color * color * avg / vmax(splat(1e-7f), splat(1.0f) - color * (1.0f - avg));
  f3 multiple = msAlb * kInvPi * mxf(1e-7f, 1.0f - oAlb) * mxf(1e-7f, 1.0f - iAlb) / mxf(1e-7f, 1.0f - avg);
  pdf = l.z * kInvPi;
  return (single + multiple) * comp;
}
__device__ f3 eval_refl(float ax, float ay, f3 v, f3 l, f3 h, f3 F, float &pdf) {
  pdf = 0.0f;
  if (l.z <= 0.0f) return splat(0.0f);
  float D = ggx_ndf(h, ax, ay), G1 = smith_g1(v, ax, ay), G2 = G1 * smith_g1(l, ax, ay);
  pdf = G1 * D / (4.0f * mxf(v.z, 1e-8f));
  return F * D * G2 / (4.0f * mxf(l.z * v.z, 1e-8f));
}
__device__ f3 eval_refr(float ax, float ay, float rel, f3 tint, f3 v, f3 l, f3 h, f3 F, float &pdf) {
  pdf = 0.0f;
  if (l.z >= 0.0f) return splat(0.0f);
  float lh = dot(l, h), vh = dot(v, h);
  if (vh <= 0.0f || lh >= 0.0f) return splat(0.0f);
  float D = ggx_ndf(h, ax, ay), G1 = smith_g1(v, ax, ay), G2 = G1 * smith_g1(l, ax, ay);
  float den = lh + vh * rel;
  den *= den;
  float jac = fabsf(lh) / mxf(den, 1e-8f);
  float rel2 = rel * rel;
  pdf = G1 * mxf(0.0f, vh) * D * jac / mxf(v.z, 1e-8f);
  return tint * (splat(1.0f) - F) * D * G2 * fabsf(vh) * jac * rel2 / mxf(fabsf(l.z * v.z), 1e-8f);
}
__device__ __forceinline__ float avg_fresnel(float eta) {
  if (fabsf(eta - 1.0f) < 1e-6f) return 0.0f;
  if (eta > 1.0f) return (eta - 1.0f) / (4.08567f + 1.00071f * eta);
  return 0.997118f + 0.1014f * eta - 0.965241f * eta * eta - 0.130607f * eta * eta * eta;
}
__device__ f3 metal_avg_fresnel(f3 f0, f3 tint) {
  f3 edgeF = mix3(f0, splat(1.0f), schlick5(kEdgeCos));
  f3 b = edgeF * (splat(1.0f) - tint) / (kEdgeCos * mtxfm::powf(1.0f - kEdgeCos, 6.0f));
  return clamp3(f0 + (splat(1.0f) - f0) / 21.0f - b / 126.0f, 0.0f, 1.0f);
}
__device__ f3 coat_passage(const Mat &m, float c) {
  if (c <= 0.0f) return splat(1.0f);
  float rc = sqrtf(mxf(0.0f, 1.0f - (1.0f - c * c) / (m.coatIor * m.coatIor)));
  f3 passage = vpow(vmaxs(m.coatColor, 0.0f), splat(0.5f / mxf(rc, 1e-5f)));
  return mix3(splat(1.0f), passage, m.coatWeight);
}

enum { LF = 0, LC = 1, LM = 2, LS = 3, LT = 4, LD = 5, LSS = 6, LN = 7 };

// sqrt(2 * PI) with PI = 3.14159265359 as WGSL's abstract-float constant
// folding computes it (in f64), rounded to f32.
constexpr float kSqrtTwoPi = KF(2.506628274631083);

// openpbr_fuzz_directional_albedo (openPbrBsdf.wgsl.ts:595-607): the fitted
// albedo of the fuzz lobe, a Gaussian in the view cosine whose scale, mean
// and offset are rational functions of the roughness. pow(z, 2) is written
// z * z, its correctly rounded value.
__device__ float fuzz_albedo(float viewCosine, float roughness) {
  const float x = clampf(viewCosine, 0.0f, 1.0f);
  const float y = clampf(roughness, 0.01f, 1.0f);
  const float scale = y * (0.0206607f + 1.58491f * y) / (0.0379424f + y * (1.32227f + y));
  const float mean = y * (-0.193854f + y * (-1.14885f + y * (1.7932f - 0.95943f * y * y))) / (0.046391f + y);
  const float offset = y * (0.000654023f + (-0.0207818f + 0.119681f * y) * y) / (1.26264f + y * (-1.92021f + y));
  const float z = (x - mean) / scale;
  const float gaussian = mtxfm::expf(-0.5f * (z * z)) / (scale * kSqrtTwoPi);
  return mxf(gaussian + offset, 0.0f);
}

struct Prep {
  Mat m;
  Basis basis;
  f3 v, coatBaseScale, metalMsScale, msRefl, msTrans;
  // Fuzz layer, fixed by the view direction: the factor on every lobe
  // under it and its own response weight * color * albedo / pi.
  f3 fuzzResp;
  float fuzzScale;
  float specEta, refrEta, transReflIor, specAlpha, dsComp;
  float metalMsFactor, metalMsProb, msViewLoss, msRatio, msProb;
  float w[LN], p[LN];
};

__device__ f3 coat_darkening(const Prep &P) {
  const Mat &m = P.m;
  float diel = 1.0f - m.baseMetalness;
  float opaque = diel * (1.0f - m.transmissionWeight) * (1.0f - m.subsurfaceWeight);
  float sss = diel * (1.0f - m.transmissionWeight) * m.subsurfaceWeight;
  float trans = diel * m.transmissionWeight;
  float oSpec = avg_fresnel(P.specEta), tSpec = avg_fresnel(1.0f / P.transReflIor);
  float specularity = clampf(opaque * oSpec + (sss + trans) * tSpec + m.baseMetalness, 0.0f, 1.0f);
  float effR = mixf(1.0f, m.specularRoughness, specularity);
  float ks = avg_fresnel(m.coatIor);
  float kr = 1.0f - (1.0f - ks) / (m.coatIor * m.coatIor);
  float k = mixf(ks, kr, effR);
  f3 eb = m.baseMetalness * m.specularWeight * metal_avg_fresnel(m.baseWeight * m.baseColor, m.specularColor) +
          opaque * mix3(m.baseWeight * m.baseColor, splat(1.0f), oSpec) + sss * m.subsurfaceColor + splat(trans);
  f3 delta = splat(1.0f - k) / vmax(splat(1.0f) - eb * k, splat(1e-6f));
  return mix3(splat(1.0f), delta, m.coatWeight * m.coatDarkening);
}
__device__ __forceinline__ f3 coat_base_scale(const Prep &P, float lc) {
  if (P.m.coatWeight <= 0.0f) return splat(1.0f);
  if (P.m.inside) return coat_passage(P.m, -lc);
  return P.coatBaseScale * coat_passage(P.m, lc);
}

// Thin-walled dielectric (openPbrBsdf.wgsl.ts:697-731): a slab's two
// interfaces reflect 2F / (1 + F); its transmission tint is raised to the
// power 1 / cos of the refracted angle.
__device__ __forceinline__ float thin_wall_fresnel(float f) { return 2.0f * f / (1.0f + f); }
__device__ __forceinline__ float thin_wall_reflection(float f, float transmissionWeight) {
  return mixf(f, 2.0f * f / (1.0f + f), clampf(transmissionWeight, 0.0f, 1.0f));
}
__device__ f3 final_transmission_tint(const Mat &m, float incidentCos, float eta) {
  const f3 tint = m.transmissionColor;
  if (!m.thinWalled) return tint;
  const float sinI2 = mxf(0.0f, 1.0f - incidentCos * incidentCos);
  const float sinT2 = sinI2 / mxf(eta * eta, 1e-9f);
  if (sinT2 >= 1.0f) return splat(0.0f);
  const float cosT = sqrtf(1.0f - sinT2);
  return vpow(vmaxs(tint, 0.0f), splat(1.0f / mxf(cosT, 1e-5f)));
}
// openpbr_eval_subsurface_boundary (openPbrBsdf.wgsl.ts:733-750): the
// diffuse transmission into a non-thin subsurface medium.
__device__ f3 eval_subsurface_boundary(const Mat &m, f3 v, f3 l, float &pdf) {
  pdf = 0.0f;
  if ((v.z <= 0.0f) | (l.z >= 0.0f)) return splat(0.0f);
  const float entry = fresnel_dielectric(v.z, m.relativeIor);
  pdf = -l.z * kInvPi;
  return splat(1.0f - entry) * kInvPi;
}

__device__ void prepare(Prep &P, const Mat &m, f3 view, f3 normal, const Lut &L) {
  P.m = m;
  P.basis = make_basis(normal);
  P.v = to_local(P.basis, view);
  float diel = 1.0f - m.baseMetalness, opaqueDiel = diel * (1.0f - m.transmissionWeight);
  P.w[LF] = m.fuzzWeight;
  P.w[LC] = m.inside ? 0.0f : m.coatWeight;
  P.w[LM] = m.baseMetalness;
  P.w[LS] = diel;
  P.w[LT] = diel * m.transmissionWeight;
  P.w[LD] = opaqueDiel * (1.0f - m.subsurfaceWeight) * m.baseWeight;
  P.w[LSS] = opaqueDiel * m.subsurfaceWeight;
  P.specEta = spec_weighted_ior(1.0f / m.relativeIor, m.specularWeight);
  P.m.relativeIor = 1.0f / P.specEta;
  P.refrEta = P.specEta;
  P.transReflIor = P.m.relativeIor;
  if (m.coatWeight > 0.0f) {
    float wCoat = m.inside ? 1.0f : mixf(1.0f, m.coatIor, m.coatWeight);
    float wBase = wCoat * spec_weighted_ior(m.specularIor / wCoat, m.specularWeight);
    float reflEta = wBase / wCoat;
    if ((wCoat > wBase) && (wBase >= 1.0f)) reflEta = 1.0f / reflEta;
    P.specEta = reflEta;
    P.transReflIor = 1.0f / reflEta;
    P.refrEta = wBase;
    if (m.inside) { P.refrEta = 1.0f / wBase; P.transReflIor = wBase; }
    P.m.relativeIor = 1.0f / P.refrEta;
  }
  P.specAlpha = m.specularRoughness * m.specularRoughness;
  float vc = fabsf(P.v.z);
  const float sif = pbr_fresnel(vc, 1.0f / P.specEta);
  f3 dielF = m.specularColor * (m.thinWalled ? mixf(sif, thin_wall_fresnel(sif), m.transmissionWeight) : sif);
  f3 metalF = m.specularWeight * metal_fresnel(vc, m.baseWeight * m.baseColor, m.specularColor);
  float transF = pbr_fresnel(vc, P.transReflIor);
  if (m.thinWalled) transF = thin_wall_fresnel(transF);
  f3 tint = final_transmission_tint(m, vc, P.refrEta);
  P.dsComp = opaque_ec(L, P.specEta, P.specAlpha, vc) / mxf(opaque_avg_ec(L, P.specEta, P.specAlpha), 1e-12f);
  P.p[LF] = 0.0f;
  P.fuzzScale = 1.0f;
  P.fuzzResp = splat(0.0f);
  if (P.w[LF] > 0.0f) {
    const float fa = fuzz_albedo(vc, m.fuzzRoughness);
    P.p[LF] = P.w[LF] * lum_pbr(m.fuzzColor) * mxf(fa, 0.02f);
    P.fuzzScale = mxf(0.0f, 1.0f - P.w[LF] * fa);
    P.fuzzResp = P.w[LF] * (m.fuzzColor * fa * kInvPi);
  }
  P.p[LC] = 0.0f;
  P.coatBaseScale = splat(1.0f);
  if ((m.coatWeight > 0.0f) && !m.inside) {
    float ce = m.coatRoughness <= kRoughnessMin
                   ? pbr_fresnel(vc, 1.0f / m.coatIor)
                   : clampf(1.0f - opaque_ec(L, m.coatIor, m.coatRoughness * m.coatRoughness, vc), 0.0f, 1.0f);
    P.p[LC] = P.w[LC] * ce;
    P.coatBaseScale = coat_passage(m, vc) * (1.0f - P.w[LC] * ce) * coat_darkening(P);
  }
  P.p[LM] = P.w[LM] * lum_pbr(metalF);
  P.metalMsScale = splat(0.0f);
  P.metalMsFactor = 0.0f;
  P.metalMsProb = 0.0f;
  if ((m.specularRoughness >= 0.04f) && (P.w[LM] > 0.0f)) {
    f3 avgF = metal_avg_fresnel(m.baseWeight * m.baseColor, m.specularColor);
    P.metalMsScale = m.specularWeight * avgF * avgF;
    float viewLoss = metal_ec(L, P.specAlpha, vc);
    P.metalMsFactor = viewLoss / mxf(metal_avg_ec(L, P.specAlpha), 1e-12f);
    float est = P.w[LM] * lum_pbr(P.metalMsScale) * viewLoss;
    P.p[LM] += est;
    P.metalMsProb = est / mxf(P.p[LM], 1e-12f);
  }
  P.p[LS] = P.w[LS] * lum_pbr(dielF);
  P.p[LT] = P.w[LT] * (1.0f - transF) * lum_pbr(tint);
  P.msRefl = splat(0.0f);
  P.msTrans = splat(0.0f);
  P.msViewLoss = 0.0f;
  P.msRatio = 0.0f;
  P.msProb = 0.0f;
  if (!m.thinWalled & (m.transmissionWeight > 0.0f) & (P.specAlpha >= 0.0016f) & (P.w[LT] > 0.0f) &
      (fabsf(P.refrEta - 1.0f) > 1e-6f)) {
    float nonOpaque = P.w[LT] + P.w[LSS];
    P.msRefl = (m.inside ? splat(1.0f) : m.specularColor) * nonOpaque;
    P.msTrans = tint * P.w[LT] + splat(P.w[LSS]);
    P.msRatio = ideal_ratio(L, P.refrEta, P.specAlpha);
    P.msViewLoss = ideal_loss(L, P.refrEta, P.specAlpha, vc);
    f3 avgScale = mix3(P.msTrans, P.msRefl, P.msRatio);
    float est = lum_pbr(avgScale) * P.msViewLoss;
    P.p[LT] += est;
    P.msProb = est / mxf(P.p[LT], 1e-12f);
  }
  P.p[LD] = P.w[LD] * lum_pbr(m.baseColor) * P.dsComp;
  P.p[LSS] = P.w[LSS] * (1.0f - transF) * lum_pbr(m.subsurfaceColor) * P.dsComp;
  if ((m.coatWeight > 0.0f) && !m.inside) {
    float bs = maxc(P.coatBaseScale);
    P.p[LM] *= bs; P.p[LS] *= bs; P.p[LT] *= bs; P.p[LD] *= bs; P.p[LSS] *= bs;
  }
  float total = P.p[0];
#pragma unroll
  for (int l = 1; l < LN; l++) total += P.p[l];
  if (total <= 1e-8f) {
#pragma unroll
    for (int l = 0; l < LN; l++) P.p[l] = 0.0f;
  } else {
    float inv = 1.0f / total;
#pragma unroll
    for (int l = 0; l < LN; l++) P.p[l] *= inv;
  }
}

__device__ __forceinline__ float substrate_fresnel(const Prep &P, float c, bool reflection) {
  if (P.m.coatWeight <= 0.0f) return pbr_fresnel(c, P.m.relativeIor);
  float tF = pbr_fresnel(c, P.transReflIor);
  if (!reflection) return tF;
  float oF = pbr_fresnel(c, 1.0f / P.specEta);
  float nonOpaque = P.m.transmissionWeight + (1.0f - P.m.transmissionWeight) * P.m.subsurfaceWeight;
  return mixf(oF, tF, nonOpaque);
}

__device__ f3 eval_bsdf(const Prep &P, f3 lightDir, float &pdf, const Lut &L) {
  const Mat &m = P.m;
  f3 v = P.v, l = to_local(P.basis, lightDir);
  bool reflection = l.z * v.z > 0.0f;
  f3 h = reflection ? normalize(l + v) : normalize(l + v * m.relativeIor);
  if (h.z < 0.0f) h = -h;
  f3 resp = splat(0.0f);
  pdf = 0.0f;
  float lp = 0.0f;
  float vh = fabsf(dot(v, h));
  // Thin-wall transmission mirrors the light to the view's side: its half
  // vector is that of a reflection.
  f3 thinH = h;
  float dielVh = vh;
  if (!reflection & m.thinWalled) {
    thinH = normalize(f3{l.x, l.y, -l.z} + v);
    if (thinH.z < 0.0f) thinH = -thinH;
    dielVh = fabsf(dot(v, thinH));
  }
  bool evalDiel = (reflection & (P.w[LS] > 0.0f)) | (!reflection & (P.w[LT] > 0.0f));
  bool evalMetal = reflection & (P.w[LM] > 0.0f);
  f3 fr0 = splat(0.0f), fr1 = splat(0.0f);
  if (evalDiel) fr0 = splat(substrate_fresnel(P, dielVh, reflection));
  if (evalMetal) fr1 = metal_fresnel(vh, m.baseWeight * m.baseColor, m.specularColor);
  float comp = P.dsComp * opaque_ec(L, P.specEta, P.specAlpha, fabsf(l.z));
  if (reflection) {
    if (P.w[LD] > 0.0f) { resp = resp + P.w[LD] * eval_diffuse(m, m.baseColor, v, l, comp, lp); pdf += P.p[LD] * lp; }
    if ((P.w[LSS] > 0.0f) & m.thinWalled) {
      resp = resp + 0.5f * P.w[LSS] * eval_diffuse(m, m.subsurfaceColor, v, l, comp, lp);
      pdf += P.p[LSS] * 0.5f * lp;
    }
    if (P.w[LS] > 0.0f) {
      f3 dF = fr0;
      if (m.thinWalled) dF = splat(thin_wall_reflection(fr0.x, m.transmissionWeight));
      resp = resp + P.w[LS] * eval_refl(m.alphaX, m.alphaY, v, l, h, m.specularColor * dF, lp);
      pdf += P.p[LS] * lp;
    }
    if (P.w[LM] > 0.0f) {
      resp = resp + P.w[LM] * eval_refl(m.alphaX, m.alphaY, v, l, h, m.specularWeight * fr1, lp);
      if (P.metalMsFactor > 0.0f) {
        float factor = mnf(P.metalMsFactor * metal_ec(L, P.specAlpha, l.z), 1.0f / mxf(l.z, 1e-8f));
        resp = resp + P.w[LM] * P.metalMsScale * factor * kInvPi;
        lp = mixf(lp, 0.5f * kInvPi, P.metalMsProb);
      }
      pdf += P.p[LM] * lp;
    }
  } else {
    if (P.w[LT] > 0.0f) {
      if (m.thinWalled) {
        const f3 window = splat(thin_wall_reflection(fr0.x, 1.0f));
        const f3 tint = final_transmission_tint(m, fabsf(v.z), P.refrEta);
        resp = resp + P.w[LT] * eval_refl(m.alphaX, m.alphaY, v, f3{l.x, l.y, -l.z}, thinH,
                                          tint * (splat(1.0f) - window), lp);
      } else {
        resp = resp + P.w[LT] * eval_refr(m.alphaX, m.alphaY, m.relativeIor, m.transmissionColor, v, l, h, fr0, lp);
      }
      if (P.msProb > 0.0f) lp *= 1.0f - P.msProb;
      pdf += P.p[LT] * lp;
    }
    if (P.w[LSS] > 0.0f) {
      if (m.thinWalled) {
        resp = resp + 0.5f * P.w[LSS] * eval_diffuse(m, m.subsurfaceColor, v, f3{l.x, l.y, -l.z}, comp, lp);
        lp *= 0.5f;
      } else {
        resp = resp + P.w[LSS] * eval_subsurface_boundary(m, v, l, lp);
      }
      pdf += P.p[LSS] * lp;
    }
  }
  if (P.msProb > 0.0f) {
    float c = fabsf(l.z);
    if (c > 0.0f) {
      float ior = reflection ? P.refrEta : 1.0f / P.refrEta;
      float ratio = reflection ? P.msRatio : 1.0f - P.msRatio;
      float factors = P.msViewLoss * ideal_loss(L, ior, P.specAlpha, c) / mxf(ideal_avg(L, ior, P.specAlpha), 1e-12f);
      float factor = mnf(factors * ratio, 1.0f / c);
      resp = resp + (reflection ? P.msRefl : P.msTrans) * factor * kInvPi;
      pdf += P.p[LT] * P.msProb * ratio * 0.5f * kInvPi;
    }
  }
  if (m.coatWeight > 0.0f) resp = resp * coat_base_scale(P, l.z);
  if (reflection & (P.w[LC] > 0.0f)) {
    f3 coatF = splat(pbr_fresnel(vh, 1.0f / m.coatIor));
    float cax, cay;
    rough_alpha(m.coatRoughness, 0.0f, cax, cay);
    resp = resp + P.w[LC] * eval_refl(cax, cay, v, l, h, coatF, lp);
    pdf += P.p[LC] * lp;
  }
  if (reflection & (P.w[LF] > 0.0f)) {  // fuzz on top (openPbrBsdf.wgsl.ts:1400-1418)
    const bool lit = (v.z > 0.0f) & (l.z > 0.0f);
    resp = resp * P.fuzzScale;
    resp = resp + (lit ? P.fuzzResp : splat(0.0f));
    pdf += P.p[LF] * (lit ? l.z * kInvPi : 0.0f);
  }
  return resp * fabsf(l.z);
}

__device__ bool sample_bsdf(const Prep &P, float rx, float ry, float lobe, f3 &dir, bool &transmission) {
  const Mat &m = P.m;
  f3 v = P.v;
  transmission = false;
  dir = P.basis.n;
  float cdf[LN];
  cdf[0] = P.p[0];
#pragma unroll
  for (int i = 1; i < LN; i++) cdf[i] = cdf[i - 1] + P.p[i];
  if (cdf[LSS] <= 0.0f) return false;
  f3 l;
  if ((lobe < cdf[LF]) | ((lobe >= cdf[LT]) & (lobe < cdf[LD]))) {
    l = cosine_hemisphere(rx, ry);
  } else if (lobe < cdf[LC]) {
    float cax, cay;
    rough_alpha(m.coatRoughness, 0.0f, cax, cay);
    f3 h = ggx_vndf(v, cax, cay, rx, ry);
    if (h.z < 0.0f) h = -h;
    l = normalize(reflect3(-v, h));
  } else if (lobe < cdf[LS]) {
    bool isMetal = lobe < cdf[LM];
    float metalDraw = (lobe - cdf[LC]) / mxf(P.p[LM], 1e-12f);
    if (isMetal & (metalDraw < P.metalMsProb)) {
      float z = rx, radial = sqrtf(mxf(0.0f, 1.0f - z * z)), az = kTwoPi * ry;
      const mtxfm::SinCos a = mtxfm::sincosf(az);
      l = f3{radial * a.c, radial * a.s, z};
    } else {
      f3 h = ggx_vndf(v, m.alphaX, m.alphaY, rx, ry);
      if (h.z < 0.0f) h = -h;
      l = normalize(reflect3(-v, h));
    }
  } else if (lobe < cdf[LT]) {
    transmission = true;
    float msDraw = (lobe - cdf[LS]) / mxf(P.p[LT], 1e-12f);
    if ((P.msProb > 0.0f) & (msDraw < P.msProb)) {
      float ratio = P.msRatio;
      bool reflectMs = rx < ratio;
      float z = reflectMs ? mnf(rx / ratio, 0.99999994f) : -mnf((rx - ratio) / (1.0f - ratio), 0.99999994f);
      float radial = sqrtf(mxf(0.0f, 1.0f - z * z)), az = kTwoPi * ry;
      const mtxfm::SinCos a = mtxfm::sincosf(az);
      l = f3{radial * a.c, radial * a.s, z};
      transmission = !reflectMs;
    } else {
      f3 h = ggx_vndf(v, m.alphaX, m.alphaY, rx, ry);
      if (h.z < 0.0f) h = -h;
      if (m.thinWalled) {
        const f3 r = normalize(reflect3(-v, h));
        l = f3{r.x, r.y, -r.z};
      } else {
        l = refract3(-v, h, m.relativeIor);
        if (dot(l, l) <= 1e-8f) { transmission = false; return false; }
        l = normalize(l);
      }
    }
  } else {
    transmission = true;
    if (m.thinWalled) {  // half diffuse reflection, half diffuse transmission
      const bool transmit = rx >= 0.5f;
      const float sx = transmit ? (rx - 0.5f) * 2.0f : rx * 2.0f;
      l = cosine_hemisphere(sx, ry);
      if (transmit) l.z = -l.z; else transmission = false;
    } else {
      l = cosine_hemisphere(rx, ry);
      l.z = -l.z;
    }
  }
  dir = to_world(P.basis, l);
  if ((l.z == 0.0f) | ((l.z < 0.0f) != transmission)) return false;
  return true;
}

// ---------------------------------------------------------------- lights
struct LSample {
  f3 emission, direction, normal;
  float dist, pdf, selPmf;
  uint32_t sourceType;
  bool isDelta;
};

__device__ __forceinline__ float4 env_texel(const GpuScene &S, int x, int y) { return S.env[(size_t)y * S.envW + x]; }
__device__ f3 texture_linear(const GpuScene &S, float u, float v) {
  float fx = u * (float)S.envW - 0.5f, fy = v * (float)S.envH - 0.5f;
  float flx = floorf(fx), fly = floorf(fy);
  float tx = fx - flx, ty = fy - fly;
  int ix = (int)flx, iy = (int)fly, mx = (int)S.envW - 1, my = (int)S.envH - 1;
  int x0 = min(max(ix, 0), mx), x1 = min(max(ix + 1, 0), mx), y0 = min(max(iy, 0), my), y1 = min(max(iy + 1, 0), my);
  float4 a = env_texel(S, x0, y0), b = env_texel(S, x1, y0), c = env_texel(S, x0, y1), d = env_texel(S, x1, y1);
  f3 s1 = f3of(a), s2 = f3of(b), s3 = f3of(c), s4 = f3of(d);
  return mix3(mix3(s1, s2, tx), mix3(s3, s4, tx), ty);
}
__device__ __forceinline__ void env_uv(f3 d, float &u, float &v) {
  float phi = mtxfm::atan2f(-d.z, -d.x);
  if (phi < 0.0f) phi += kTwoPi;
  float theta = mtxfm::acosf(clampf(d.y, -1.0f, 1.0f));
  u = phi * 0.5f * kInvPi;
  v = theta * kInvPi;
}
__device__ __forceinline__ float2 cdf_at(const GpuScene &S, int x, int y) {
  x = min(max(x, 0), (int)S.envW + 1);
  y = min(max(y, 0), (int)S.envH);
  return S.envCdf[(size_t)y * (S.envW + 2) + x];
}
__device__ __forceinline__ float env_solid_pdf(float px, float py, float v) {
  if (!(v > 0.0f && v < 1.0f)) return 0.0f;
  float sinT = mtxfm::sinf(mnf(v, 1.0f - v) * kPi);
  return px * py * kInvPi2 / (2.0f * sinT);
}
__device__ float env_pdf(const GpuScene &S, float u, float v) {
  int x = min(max((int)(u * (float)S.envW), 0), (int)S.envW - 1);
  int y = min(max((int)(v * (float)S.envH), 0), (int)S.envH - 1);
  return env_solid_pdf(cdf_at(S, 1 + x, y).y, cdf_at(S, 0, y).y, v);
}
// Search bracket [lo, hi] of the first CDF entry above r: the guide table's
// entries for the bins around r (the search result is monotone in r, so it
// lies between the results for the bin's two ends), or the whole range.
__device__ __forceinline__ void env_bracket(const uint16_t *guide, uint32_t bins, float r, int full, int &lo, int &hi) {
  lo = 0;
  hi = full;
  if (guide && r >= 0.0f && r < 1.0f) {
    const uint32_t k = min((uint32_t)(r * (float)bins), bins - 1u);
    lo = guide[k];
    hi = guide[k + 1];
  }
}
__device__ f3 sample_env(const GpuScene &S, float r0, float r1, float &u, float &v, float &pdf) {
  int H = (int)S.envH, W = (int)S.envW;
  int left, right;
  env_bracket(S.envGuide, kEnvGuideRows, r0, H + 1, left, right);
  while (left < right) {
    int mid = (left + right) >> 1;
    if (cdf_at(S, 0, mid).x <= r0) left = mid + 1; else right = mid;
  }
  int vo = left - 1;
  float2 a0 = cdf_at(S, 0, vo), a1 = cdf_at(S, 0, vo + 1);
  v = ((float)vo + (r0 - a0.x) / (a1.x - a0.x)) / (float)H;
  const bool rowGuide = (vo >= 0) & (vo < H);
  env_bracket(rowGuide ? S.envGuide + (kEnvGuideRows + 1) + (size_t)vo * (kEnvGuideCols + 1) : nullptr,
              kEnvGuideCols, r1, W + 1, left, right);
  while (left < right) {
    int mid = (left + right) >> 1;
    if (cdf_at(S, 1 + mid, vo).x <= r1) left = mid + 1; else right = mid;
  }
  int uo = left - 1;
  float2 b0 = cdf_at(S, 1 + uo, vo), b1 = cdf_at(S, 1 + uo + 1, vo);
  u = ((float)uo + (r1 - b0.x) / (b1.x - b0.x)) / (float)W;
  float phi = u * kTwoPi, theta = v * kPi;
  const mtxfm::SinCos t = mtxfm::sincosf(theta), f = mtxfm::sincosf(phi);
  const float ct = t.c, st = t.s, cp = f.c, sp = f.s;
  pdf = env_solid_pdf(b0.y, a0.y, v);
  return f3{-st * cp, ct, -st * sp};
}
// Analytic light types (lightConstants.ts:10-15); type 0 is read as a rect.
enum : int { kLightDirectional = 1, kLightSphere = 2, kLightRect = 3, kLightSpot = 4, kLightDisk = 5, kLightCylinder = 6 };
__device__ __forceinline__ int light_type(const float *__restrict__ L) {
  const int t = (int)L[0];
  return t == 0 ? kLightRect : t;
}
constexpr float kInv4Pi = KF(0.07957747154594766);

__device__ void sample_rect(const float *__restrict__ L, f3 p, float sx, float sy, LSample &r) {
  f3 U{L[4], L[5], L[6]}, V{L[7], L[8], L[9]}, pos{L[1], L[2], L[3]};
  f3 n = normalize(cross(U, V));
  f3 lowerLeft = pos - 0.5f * (U + V);
  f3 surf = lowerLeft + U * sx + V * sy;
  f3 off = surf - p;
  float d2 = dot(off, off);
  r.dist = sqrtf(d2);
  r.direction = off / r.dist;
  float ec = dot(-r.direction, n);
  float spread = clampf(L[13], 0.0f, 1.0f);
  float sc = mtxfm::cosf(0.5f * kPi * spread);
  if (ec <= 0.0f || ec < sc) { r.emission = splat(0.0f); r.pdf = 0.0f; return; }
  r.emission = f3{L[10], L[11], L[12]};
  r.pdf = d2 / mxf(L[14] * ec, 1e-8f);
}
// light_sample_sphere (sampleLight.wgsl.ts:10-89): a zero radius is a point
// emitter (delta, emission = intensity / d^2); inside the sphere uniform
// directions, outside cone sampling with the cancellation-free solid angle.
__device__ void sample_sphere(const float *__restrict__ L, f3 p, float rx, float ry, LSample &r) {
  const f3 pos{L[1], L[2], L[3]}, E{L[10], L[11], L[12]};
  const f3 co = pos - p;
  const float cd = len(co);
  if (L[13] <= 0.0f) {
    if (cd <= 1e-6f) {
      r.direction = f3{0.0f, 0.0f, 1.0f};
      r.dist = 0.0f;
      r.pdf = 1.0f;
      r.emission = splat(0.0f);
      r.isDelta = true;
      return;
    }
    r.direction = co / cd;
    r.dist = cd;
    r.pdf = 1.0f;
    r.emission = E / mxf(cd * cd, 1e-8f);
    r.isDelta = true;
    return;
  }
  const float radius = mxf(L[13], 1e-6f);
  r.isDelta = false;
  f3 sd;
  if (cd <= radius) {
    const float z = 1.0f - 2.0f * rx;
    const float radial = sqrtf(mxf(0.0f, 1.0f - z * z));
    const float az = kTwoPi * ry;
    const mtxfm::SinCos a = mtxfm::sincosf(az);
    sd = f3{radial * a.c, radial * a.s, z};
    r.pdf = kInv4Pi;
  } else {
    const f3 axis = co / cd;
    const float s2 = mnf(1.0f, radius * radius / (cd * cd));
    const float cmax = sqrtf(mxf(0.0f, 1.0f - s2));
    const float omc = s2 / (1.0f + cmax);
    const float omct = rx * omc;
    const float ct = 1.0f - omct;
    const float st = sqrtf(mxf(0.0f, omct * (2.0f - omct)));
    const float az = kTwoPi * ry;
    const Basis B = make_basis(axis);
    const mtxfm::SinCos a = mtxfm::sincosf(az);
    sd = axis * ct + B.t * (st * a.c) + B.b * (st * a.s);
    r.pdf = 1.0f / mxf(kTwoPi * omc, 1e-8f);
  }
  const float cp = dot(sd, co), c2 = dot(co, co);
  const float disc = cp * cp - c2 + radius * radius;
  const float nearD = cp - sqrtf(mxf(0.0f, disc)), farD = cp + sqrtf(mxf(0.0f, disc));
  r.dist = nearD > 1e-6f ? nearD : farD;
  r.direction = sd;
  r.emission = E;
}
// evaluateSpotEmissionMultiplier (rayTraceCore.wgsl.ts:365-378), without IES
// profiles: hard cone at cos(outer) = v.x, smoothstep to cos(inner) = v.y.
__device__ float spot_multiplier(const float *__restrict__ L, f3 lightToSurface) {
  const f3 axis = normalize(f3{L[4], L[5], L[6]});
  const float cosT = dot(axis, lightToSurface);
  float cone = cosT >= L[7] ? 1.0f : 0.0f;
  if (L[8] > L[7] + 1e-6f) {
    const float t = clampf((cosT - L[7]) / (L[8] - L[7]), 0.0f, 1.0f);
    cone = t * t * (3.0f - 2.0f * t);
  }
  return cone;
}
// light_sample_disk / light_sample_cylinder (sampleLight.wgsl.ts:163-195):
// area sampling, pdf = d^2 / (area cos).
__device__ void sample_area_point(const float *__restrict__ L, f3 p, f3 surf, f3 normal, LSample &r) {
  const f3 off = surf - p;
  const float d2 = dot(off, off);
  r.dist = sqrtf(d2);
  r.direction = off / mxf(r.dist, 1e-8f);
  const float ec = dot(-r.direction, normal);
  if (ec <= 0.0f) { r.emission = splat(0.0f); r.pdf = 0.0f; return; }
  r.emission = f3{L[10], L[11], L[12]};
  r.pdf = d2 / mxf(L[14] * ec, 1e-8f);
}
__device__ void sample_disk(const float *__restrict__ L, f3 p, float sx, float sy, LSample &r) {
  const f3 U{L[4], L[5], L[6]}, V{L[7], L[8], L[9]}, pos{L[1], L[2], L[3]};
  const float rr = sqrtf(sx), phi = kTwoPi * sy;
  const mtxfm::SinCos a = mtxfm::sincosf(phi);
  const f3 surf = pos + U * (rr * a.c) + V * (rr * a.s);
  sample_area_point(L, p, surf, normalize(cross(U, V)), r);
}
__device__ void sample_cylinder(const float *__restrict__ L, f3 p, float sx, float sy, LSample &r) {
  const f3 U{L[4], L[5], L[6]}, V{L[7], L[8], L[9]}, pos{L[1], L[2], L[3]};
  const f3 axis = normalize(U), r0 = normalize(V), r1 = normalize(cross(axis, r0));
  const float phi = kTwoPi * sy;
  const mtxfm::SinCos a = mtxfm::sincosf(phi);
  const f3 normal = r0 * a.c + r1 * a.s;
  const f3 surf = pos + U * (sx - 0.5f) + normal * L[13];
  sample_area_point(L, p, surf, normal, r);
}
// sampleAnalyticLight (sampleLight.wgsl.ts:223-253).
__device__ void sample_analytic(const float *__restrict__ L, f3 p, float sx, float sy, LSample &r) {
  switch (light_type(L)) {
    case kLightRect: sample_rect(L, p, sx, sy, r); break;
    case kLightSphere: sample_sphere(L, p, sx, sy, r); break;
    case kLightSpot: {
      sample_sphere(L, p, sx, sy, r);
      const f3 surf = p + r.direction * r.dist;
      r.emission = r.emission * spot_multiplier(L, normalize(p - surf));
      break;
    }
    case kLightDirectional:
      r.direction = normalize(-f3{L[4], L[5], L[6]});
      r.emission = f3{L[10], L[11], L[12]};
      r.dist = kInf;
      r.pdf = 1.0f;
      r.isDelta = true;
      break;
    case kLightDisk: sample_disk(L, p, sx, sy, r); break;
    case kLightCylinder: sample_cylinder(L, p, sx, sy, r); break;
    default: break;
  }
}
// sampleEmissiveTriangleLight (sampleLight.wgsl.ts:316-363): uniform point on
// the world-space triangle, one-sided, pdf = selection pmf * d^2 / (area cos);
// textured scenes evaluate getMaterialEmission at the point's interpolated
// (unscaled) UVs.
template <bool kFull>
__device__ void sample_emissive(const GpuScene &S, uint32_t index, f3 p, float rx, float ry, LSample &r) {
  const float4 a = S.emissive[3 * index], b = S.emissive[3 * index + 1], c = S.emissive[3 * index + 2];
  const f3 p0 = f3of(a), p1 = f3of(b), p2 = f3of(c);
  const float sqrtU = sqrtf(rx);
  const f3 bc{1.0f - sqrtU, sqrtU * (1.0f - ry), sqrtU * ry};
  const f3 lp = bc.x * p0 + bc.y * p1 + bc.z * p2;
  const f3 normal = normalize(cross(p1 - p0, p2 - p0));
  f3 dir = lp - p;
  r.dist = len(dir);
  if (r.dist <= kEps) { r.emission = splat(0.0f); r.pdf = 0.0f; return; }
  r.direction = dir / r.dist;
  r.normal = normal;
  const float cosLight = dot(normal, -r.direction);
  if (cosLight <= 0.0f) { r.emission = splat(0.0f); r.pdf = 0.0f; return; }
  const float *m = S.materials + (size_t)(uint32_t)c.w * MTX_MATERIAL_FLOATS;
  if (kFull && S.emissiveUv != nullptr) {
    const float4 u0 = S.emissiveUv[3 * index], u1 = S.emissiveUv[3 * index + 1], u2 = S.emissiveUv[3 * index + 2];
    const float2 uv{bc.x * u0.x + bc.y * u0.z + bc.z * u1.x, bc.x * u0.y + bc.y * u0.w + bc.z * u1.y};
    const float2 uv1{bc.x * u1.z + bc.y * u2.x + bc.z * u2.z, bc.x * u1.w + bc.y * u2.y + bc.z * u2.w};
    r.emission = material_emission(S, m, uv, uv1);
  } else {
    r.emission = f3{m[23], m[24], m[25]};
  }
  r.selPmf = a.w;
  r.pdf = a.w * (r.dist * r.dist) / mxf(b.w * cosLight, 1e-8f);
}
template <bool kFull>
__device__ void sample_light(const GpuScene &S, f3 p, float rx, float ry, float rz, uint32_t cand, uint32_t count,
                             LSample &ls) {
  uint32_t cnt = max(count, 1u);
  rz = (rz + (float)min(cand, cnt - 1u)) / (float)cnt;
  ls.emission = splat(0.0f); ls.pdf = 0.0f; ls.selPmf = 0.0f; ls.dist = kInf; ls.isDelta = false; ls.sourceType = 0;
  int n = S.choiceCount;
  if (n <= 0) return;
  float scaled = rz * (float)n;
  int column = min((int)scaled, n - 1);
  LightChoice c = S.choices[column];
  int sel = fractf(scaled) < c.prob ? column : c.alias;
  LightChoice ch = S.choices[sel];
  if (ch.type == 0 && S.env) {
    float u, v, pdf;
    ls.direction = sample_env(S, rx, ry, u, v, pdf);
    ls.pdf = pdf;
    ls.emission = texture_linear(S, u, v);
    ls.dist = kInf;
    ls.selPmf = S.envPmf;
    ls.pdf *= S.envPmf;
  } else if (ch.type == 1) {
    ls.sourceType = 2;
    sample_analytic(S.lights + (size_t)ch.index * 16, p, rx, ry, ls);
    ls.selPmf = S.lightPmf[ch.index];
    ls.pdf *= ls.selPmf;
  } else if (ch.type == 2) {
    ls.sourceType = 1;
    sample_emissive<kFull>(S, (uint32_t)ch.index, p, rx, ry, ls);
  }
}
// intersectSphere (intersect.wgsl.ts:924-945).
__device__ float intersect_sphere(float rad, f3 pos, f3 ro, f3 rd) {
  const f3 op = pos - ro;
  const float b = dot(op, rd);
  float det = b * b - dot(op, op) + rad * rad;
  if (det < 0.0f) return kInf;
  det = sqrtf(det);
  const float t1 = b - det;
  if (t1 > 0.001f) return t1;
  const float t2 = b + det;
  if (t2 > 0.001f) return t2;
  return kInf;
}
// sphereSolidAnglePdf (intersect.wgsl.ts:966-974).
__device__ float sphere_solid_angle_pdf(f3 origin, f3 center, float radius) {
  const float cd = len(origin - center);
  if (cd <= radius) return kInv4Pi;
  const float s2 = mnf(1.0f, radius * radius / (cd * cd));
  const float cmax = sqrtf(mxf(0.0f, 1.0f - s2));
  return 1.0f / mxf(kTwoPi * (s2 / (1.0f + cmax)), 1e-8f);
}
// intersectLight (intersect.wgsl.ts:976-1122): distance to the analytic light
// along a continuation ray (kInf when missed or not hittable) with the
// light's emission and the conditional density NEE would have used.
__device__ float intersect_light(const float *__restrict__ L, f3 ro, f3 rd, LSample &ls) {
  ls.pdf = 0.0f;
  ls.emission = splat(0.0f);
  ls.isDelta = false;
  const f3 U{L[4], L[5], L[6]}, V{L[7], L[8], L[9]}, P{L[1], L[2], L[3]}, E{L[10], L[11], L[12]};
  const int type = light_type(L);
  if (type == kLightRect) {
    f3 pos = P - U * 0.5f - V * 0.5f;
    f3 n = normalize(cross(U, V));
    if (dot(n, rd) > 0.0f) return kInf;
    float t = (dot(n, pos) - dot(n, ro)) / dot(rd, n);
    float d = kInf;
    if (t > kEps) {
      f3 vi = (ro + rd * t) - pos;
      float a1 = dot(U / dot(U, U), vi);
      if (a1 >= 0.0f && a1 <= 1.0f) {
        float a2 = dot(V / dot(V, V), vi);
        if (a2 >= 0.0f && a2 <= 1.0f) d = t;
      }
    }
    if (d == kInf) return kInf;
    float cosT = dot(-rd, n);
    if (cosT >= mtxfm::cosf(L[13] * kPi / 2.0f)) {
      ls.pdf = (d * d) / (L[14] * cosT);
      ls.emission = E;
      return d;
    }
    return kInf;
  }
  if (type == kLightDisk) {
    const f3 n = normalize(cross(U, V));
    if (dot(n, rd) >= 0.0f) return kInf;
    const float d = dot(P - ro, n) / dot(n, rd);
    const f3 hit = ro + rd * d - P;
    const float radius = mxf(len(U), 1e-8f);
    if (d <= kEps || dot(hit, hit) > radius * radius) return kInf;
    const float cosine = dot(-rd, n);
    ls.pdf = d * d / mxf(L[14] * cosine, 1e-8f);
    ls.emission = E;
    return d;
  }
  if (type == kLightCylinder) {
    const f3 axis = normalize(U);
    const float halfLength = 0.5f * len(U);
    const f3 oc = ro - P;
    const f3 rdp = rd - axis * dot(rd, axis), od = oc - axis * dot(oc, axis);
    const float a = dot(rdp, rdp), b = dot(od, rdp), c = dot(od, od) - L[13] * L[13];
    const float disc = b * b - a * c;
    if (a <= 1e-12f || disc < 0.0f) return kInf;
    const float root = sqrtf(disc);
    const float nearD = (-b - root) / a, farD = (-b + root) / a;
    const float d = nearD > kEps ? nearD : farD;
    const float axial = dot(oc + rd * d, axis);
    if (d <= kEps || fabsf(axial) > halfLength) return kInf;
    const f3 hit = ro + rd * d;
    const f3 n = normalize(hit - P - axis * axial);
    const float cosine = dot(-rd, n);
    if (cosine <= 0.0f) return kInf;
    ls.pdf = d * d / mxf(L[14] * cosine, 1e-8f);
    ls.emission = E;
    return d;
  }
  if (type == kLightSphere || type == kLightSpot) {
    float d;
    if (L[13] <= 0.0f) {
      const f3 toPoint = P - ro;
      const float proj = dot(toPoint, rd);
      const f3 perp = toPoint - proj * rd;
      if (proj <= kEps || dot(perp, perp) > 1e-8f) return kInf;
      d = proj;
      ls.pdf = 1.0f;
      ls.isDelta = true;
    } else {
      d = intersect_sphere(L[13], P, ro, rd);
      if (type == kLightSphere && d > 0.0f && d != kInf) ls.pdf = sphere_solid_angle_pdf(ro, P, L[13]);
    }
    if (type == kLightSphere) {
      if (d < 0.0f) return kInf;
      ls.emission = L[13] <= 0.0f ? E / mxf(d * d, 1e-8f) : E;
      return d;
    }
    if (d != kInf) {
      const f3 hitPt = ro + d * rd;
      const float m = spot_multiplier(L, normalize(ro - hitPt));
      if (m <= 0.0f) return kInf;
      if (L[13] > 0.0f) ls.pdf = sphere_solid_angle_pdf(ro, P, L[13]);
      ls.emission = E * m * (L[13] <= 0.0f ? 1.0f / mxf(d * d, 1e-8f) : 1.0f);
    }
    return d;
  }
  return kInf;  // directional lights are not hit by rays
}
__device__ __forceinline__ f3 offset_surface_ray(f3 p, f3 gn, f3 dir) {
  f3 n = dot(gn, dir) >= 0.0f ? gn : -gn;
  const float lim = KF(1.0 / 32.0);
  float pc[3] = {p.x, p.y, p.z}, nc[3] = {n.x, n.y, n.z}, o[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int32_t io = (int32_t)(256.0f * nc[k]);
    float shifted = intf(fint(pc[k]) + (pc[k] < 0.0f ? -io : io));
    o[k] = fabsf(pc[k]) < lim ? pc[k] + nc[k] / 65536.0f : shifted;
  }
  return f3{o[0], o[1], o[2]};
}

// ---------------------------------------------------------------- integrator
struct Surf {
  f3 position, normal, faceNormal, shadowOffset;
  int32_t tri, slot;
  uint32_t material;
  float color[3];
  float2 uv, uv1;  // texture coordinates of textured scenes (times the material's UV scale)
  bool hit, albedoSampled;
};

// getMatNormal (materialBuffer.wgsl.ts:377-415): the normal map's tangent-
// space normal in the cotangent frame of UV0 (dp1 = p0 - p2, dp2 = p1 - p2,
// duv the raw vertex UV differences), object space.
__device__ f3 mapped_normal(const GpuScene &S, uint32_t slot, const float *__restrict__ m, float2 uv, f3 normal,
                            f3 dp1, f3 dp2, float2 duv1, float2 duv2) {
  const f3 dp2perp = cross(dp2, normal), dp1perp = cross(normal, dp1);
  f3 dpdu = dp2perp * duv1.x + dp1perp * duv2.x;
  f3 dpdv = dp2perp * duv1.y + dp1perp * duv2.y;
  const float g2 = mxf(dot(dpdu, dpdu), dot(dpdv, dpdv));
  if (g2 > 1e-12f) {
    const float inv = 1.0f / sqrtf(g2);
    dpdu = dpdu * inv;
    dpdv = dpdv * inv;
  } else {
    const f3 ref = fabsf(normal.z) < 0.999f ? f3{0.0f, 0.0f, 1.0f} : f3{0.0f, 1.0f, 0.0f};
    dpdu = normalize(cross(ref, normal));
    dpdv = cross(normal, dpdu);
  }
  const float4 c = sample_map(S, slot - 1u, uv);
  f3 n{2.0f * c.x - 1.0f, 2.0f * c.y - 1.0f, 2.0f * c.z - 1.0f};
  n = normalize(f3{n.x * m[MTX_MAT_NORMAL_SCALE], n.y * m[MTX_MAT_NORMAL_SCALE + 1], n.z});
  return normalize(dpdu * n.x + dpdv * n.y + normal * n.z);
}

// completeSurfaceInteraction + fetchAttributes (intersect.wgsl.ts:1233-1419,
// 1539-1601). kFull: textured scenes also interpolate the UV sets, scale
// them by the material's UV scale and apply its normal map.
template <bool kFull = true>
__device__ void complete_surface(const GpuScene &S, const TraceHit &h, Surf &si) {
  const ShadeSlot &sh = S.shade[h.slot];
  const InstSlot &is = S.accel->slots[h.slot];
  const TriRec r = is.tris[h.tri];
  uint32_t src = (uint32_t)fint(r.v0.w);
  uint32_t i0 = sh.indices[src * 3], i1 = sh.indices[src * 3 + 1], i2 = sh.indices[src * 3 + 2];
  f3 p0 = f3of(r.v0), p1 = f3of(r.v1), p2 = f3of(r.v2);
  f3 n0{sh.normals[i0 * 3], sh.normals[i0 * 3 + 1], sh.normals[i0 * 3 + 2]};
  f3 n1{sh.normals[i1 * 3], sh.normals[i1 * 3 + 1], sh.normals[i1 * 3 + 2]};
  f3 n2{sh.normals[i2 * 3], sh.normals[i2 * 3 + 1], sh.normals[i2 * 3 + 2]};
  f3 b = h.bary, position = h.position;
  f3 smooth = b.x * n0 + b.y * n1 + b.z * n2;
  f3 normal = normalize(smooth);
  f3 geometric = normalize(cross(p1 - p0, p2 - p0));
  float A = dot(n2 - n0, p0 - p2), B = dot(n2 - n1, p1 - p2), C = dot(n1 - n0, p1 - p0);
  float hh = A * b.x * (b.x - 1.0f) + (A + B + C) * b.x * b.y + B * b.y * (b.y - 1.0f);
  if (dot(smooth, geometric) > 0.0f) {
    float h0 = mxf(mxf(dot(p1 - p0, n0), dot(p2 - p0, n0)), 0.0f);
    float h1 = mxf(mxf(dot(p0 - p1, n1), dot(p2 - p1, n1)), 0.0f);
    float h2 = mxf(mxf(dot(p0 - p2, n2), dot(p1 - p2, n2)), 0.0f);
    h0 = mxf(dot(p0 - position, n0) + h0, 0.0f);
    h1 = mxf(dot(p1 - position, n1) + h1, 0.0f);
    h2 = mxf(dot(p2 - position, n2) + h2, 0.0f);
    hh = mxf(mnf(mnf(h0, h1), h2), hh * 0.5f);
  } else {
    float h0 = mxf(mxf(dot(p0 - p1, n0), dot(p0 - p2, n0)), 0.0f);
    float h1 = mxf(mxf(dot(p1 - p0, n1), dot(p1 - p2, n1)), 0.0f);
    float h2 = mxf(mxf(dot(p2 - p0, n2), dot(p2 - p1, n2)), 0.0f);
    h0 = mxf(dot(position - p0, n0) + h0, 0.0f);
    h1 = mxf(dot(position - p1, n1) + h1, 0.0f);
    h2 = mxf(dot(position - p2, n2) + h2, 0.0f);
    hh = mnf(-mnf(mnf(h0, h1), h2), hh * 0.5f);
  }
  Affine w2l{f3of(is.c0), f3of(is.c1), f3of(is.c2), f3of(is.c3)};
  Affine l2w{f3of(sh.l2w0), f3of(sh.l2w1), f3of(sh.l2w2), f3of(sh.l2w3)};
  si.hit = true;
  si.tri = h.tri;
  si.slot = h.slot;
  si.material = sh.material;
  si.color[0] = sh.color[0]; si.color[1] = sh.color[1]; si.color[2] = sh.color[2];
  si.uv = si.uv1 = float2{0.0f, 0.0f};
  f3 shading = normal;
  if (kFull && S.textures != nullptr) {
    const float *__restrict__ m = S.materials + (size_t)sh.material * MTX_MATERIAL_FLOATS;
    float2 t0{0.0f, 0.0f}, t1{0.0f, 0.0f}, t2{0.0f, 0.0f};
    if (sh.uvs != nullptr) {
      t0 = float2{sh.uvs[i0 * 2], sh.uvs[i0 * 2 + 1]};
      t1 = float2{sh.uvs[i1 * 2], sh.uvs[i1 * 2 + 1]};
      t2 = float2{sh.uvs[i2 * 2], sh.uvs[i2 * 2 + 1]};
      si.uv = float2{b.x * t0.x + b.y * t1.x + b.z * t2.x, b.x * t0.y + b.y * t1.y + b.z * t2.y};
    }
    if (sh.uvs1 != nullptr) {
      const float2 s0{sh.uvs1[i0 * 2], sh.uvs1[i0 * 2 + 1]}, s1{sh.uvs1[i1 * 2], sh.uvs1[i1 * 2 + 1]};
      const float2 s2{sh.uvs1[i2 * 2], sh.uvs1[i2 * 2 + 1]};
      si.uv1 = float2{b.x * s0.x + b.y * s1.x + b.z * s2.x, b.x * s0.y + b.y * s1.y + b.z * s2.y};
    }
    const float su = m[MTX_MAT_UV_SCALE], sv = m[MTX_MAT_UV_SCALE + 1];
    si.uv = float2{si.uv.x * su, si.uv.y * sv};
    si.uv1 = float2{si.uv1.x * su, si.uv1.y * sv};
    if (const uint32_t t = map_slot(m, MTX_MAP_NORMAL))
      shading = mapped_normal(S, t, m, si.uv, normal, p0 - p2, p1 - p2, float2{t0.x - t2.x, t0.y - t2.y},
                              float2{t1.x - t2.x, t1.y - t2.y});
  }
  si.faceNormal = normalize(xform_normal_transposed(w2l, h.normal));
  si.normal = normalize(xform_normal_transposed(w2l, shading));
  si.shadowOffset = xform(l2w, smooth * hh, 0.0f);
  si.position = xform(l2w, position, 1.0f);
}

// intersectionAlphaPasses (intersectionAlpha.wgsl.ts:28-88) for the
// triangle `tri` of slot `slot` hit at barycentrics b, whose material is
// alpha-masked: coverage = opacity, times the base map's alpha when map alpha
// is coverage, times the alpha map's g; both maps are sampled at UV0 * the
// material's UV scale (the any-hit test reads no UV1 set), and the triangle
// occludes when the coverage reaches the cutoff.
__device__ bool alpha_covers(const GpuScene &S, int32_t slot, int32_t tri, f3 b) {
  const ShadeSlot &sh = S.shade[slot];
  const float *__restrict__ m = S.materials + (size_t)sh.material * MTX_MATERIAL_FLOATS;
  float alpha = m[MTX_MAT_OPACITY];
  if (S.textures != nullptr) {
    const uint32_t base = m[MTX_MAT_MAP_ALPHA_USAGE] == 2.0f ? map_slot(m, MTX_MAP_BASE) : 0u;
    const uint32_t amap = map_slot(m, MTX_MAP_ALPHA);
    if (base | amap) {
      float2 uv{0.0f, 0.0f};
      if (sh.uvs != nullptr) {
        const TriRec r = S.accel->slots[slot].tris[tri];
        const uint32_t src = (uint32_t)fint(r.v0.w);
        const uint32_t i0 = sh.indices[src * 3], i1 = sh.indices[src * 3 + 1], i2 = sh.indices[src * 3 + 2];
        const float2 t0{sh.uvs[i0 * 2], sh.uvs[i0 * 2 + 1]}, t1{sh.uvs[i1 * 2], sh.uvs[i1 * 2 + 1]};
        const float2 t2{sh.uvs[i2 * 2], sh.uvs[i2 * 2 + 1]};
        uv = float2{b.x * t0.x + b.y * t1.x + b.z * t2.x, b.x * t0.y + b.y * t1.y + b.z * t2.y};
      }
      uv = float2{uv.x * m[MTX_MAT_UV_SCALE], uv.y * m[MTX_MAT_UV_SCALE + 1]};
      if (base) alpha *= sample_map(S, base - 1u, uv).w;
      if (amap) alpha *= sample_map(S, amap - 1u, uv).y;
    }
  }
  return alpha >= m[MTX_MAT_ALPHA_CUTOFF];
}
// Coverage test of the shadow queries of a render scene.
struct SceneCoverage {
  const GpuScene *S;
  __device__ __forceinline__ bool covers(int32_t slot, int32_t tri, f3 b) const { return alpha_covers(*S, slot, tri, b); }
};
// The coverage test of kernels compiled for scenes with (kAlpha) or without
// masked materials; without, no instance carries the masked flag.
template <bool kAlpha>
__device__ __forceinline__ auto coverage(const GpuScene *S) {
  if constexpr (kAlpha) return SceneCoverage{S};
  else return NoCoverage();
}

__device__ __forceinline__ float nee_mis(uint32_t count, float lp, float sp) {
  if (count <= 1u) return power_heuristic(lp, sp);
  float w = (float)count * lp;
  return w / mxf(w + sp, 1e-20f);
}
__device__ __forceinline__ float bsdf_mis(uint32_t count, float sp, float lp) {
  if (count <= 1u) return power_heuristic(sp, lp);
  float w = (float)count * lp;
  return sp / mxf(sp + w, 1e-20f);
}

struct Shadow {
  f3 ro, rd, contribution;
  float maxDistance, risTarget, selPmf;
  bool valid;
};

// ---------------------------------------------------------------- path
// State of one path between integrator events (rayTrace.compute.ts Path +
// RayTraceAuxiliary, rayTraceCore.wgsl.ts:192-210, :323-331).
// Bits 0-6 of `flags` are the path's booleans (kIgnored: the last vertex was
// an alpha-masked pass-through, sampleIgnored of rayTrace.compute.ts:730);
// bits 8-15 hold the path depth, the integrator's `depth`, which counts the
// completed vertices that were not passed through.
enum : uint32_t {
  kFromEye = 1u, kIsDelta = 2u, kNextRefr = 4u, kIsBlank = 8u, kAlbedoSampled = 16u, kNotTransparent = 32u,
  kIgnored = 64u, kDepthShift = 8u
};
struct PathRegs {
  f3 ro, rd, li, beta, auxAlbedo, auxNormal;
  float alpha, scatPdf;
  int32_t curTri, curSlot;  // last surface hit (self-intersection skip)
  uint32_t flags;
};
// Shadow ray of one vertex and the radiance it adds when unoccluded.
struct ShadowReq {
  f3 ro, rd, payload;
  float tMax;
  bool valid;
};

// `rng` holds the path's pcg4d state when P.pattern is the random pattern;
// it may be null otherwise.
__device__ __forceinline__ Smp path_sampler(const RenderParams &P, uint32_t px, uint32_t py, uint32_t index,
                                            uint4 *rng) {
  return Smp{px, py, P.width, index, P.seed, P.logSPP, P.pattern, 0u, rng};
}

// sampleCircle (rayTrace.compute.ts:1148-1154): concentric square-to-disk
// map of a 2D sample, with the reference's operation order.
__device__ __forceinline__ void sample_circle(float ux, float uy, float &cx, float &cy) {
  const float px = 2.0f * ux - 1.0f, py = 2.0f * uy - 1.0f;
  const bool greater = fabsf(px) > fabsf(py);
  const float r = greater ? px : py;
  const float theta = greater ? kQuarterPi * py / px : kPi * (0.5f - 0.25f * px / py);
  const mtxfm::SinCos a = mtxfm::sincosf(theta);
  cx = r * a.c;
  cy = r * a.s;
}

// main (rayTrace.compute.ts:1188-1205): pixel jitter (Sobol CAMERA_PIXEL
// sample, or the Halton offset of the random pattern), rng_initialize,
// initCameraRay (:591-610, browser projection-inverse form) and, with a
// positive aperture, the thin lens: focus plane perpendicular to the view
// direction at the focus distance, lens point = aperture *
// sampleCircle(CAMERA_LENS sample).
__device__ void init_path(const RenderParams &P, const Smp &smp, PathRegs &pr) {
  const float Wf = (float)P.width, Hf = (float)P.height;
  float uvx = ((float)smp.px + 0.5f) / Wf, uvy = ((float)smp.py + 0.5f) / Hf;
  float jx, jy;
  if (smp.pattern == kPatternRandom) {
    jx = (float)((halton(smp.index, 2u) - 0.5) / (double)P.width);
    jy = (float)((halton(smp.index, 3u) - 0.5) / (double)P.height);
    smp.seed_rng();
  } else {
    float sx, sy;
    smp.s2(0u, sx, sy);
    jx = (sx - 0.5f) / Wf;
    jy = (sy - 0.5f) / Hf;
  }
  float cx = uvx + jx, cy = uvy + jy;
  float nx = cx * 2.0f - 1.0f, ny = cy * 2.0f - 1.0f;
  float o4[4], c4[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    float a = P.projInv[r] * nx; a = a + P.projInv[4 + r] * ny; a = a + P.projInv[8 + r] * -1.0f; a = a + P.projInv[12 + r] * 1.0f;
    float b = P.projInv[r] * nx; b = b + P.projInv[4 + r] * ny; b = b + P.projInv[8 + r] * 1.0f; b = b + P.projInv[12 + r] * 1.0f;
    o4[r] = a; c4[r] = b;
  }
  f3 dir{c4[0] * o4[3] - o4[0] * c4[3], c4[1] * o4[3] - o4[1] * c4[3], c4[2] * o4[3] - o4[2] * c4[3]};
  f3 op{o4[0] / o4[3], o4[1] / o4[3], o4[2] / o4[3]};
  Affine vi{f3{P.viewInv[0], P.viewInv[1], P.viewInv[2]}, f3{P.viewInv[4], P.viewInv[5], P.viewInv[6]},
            f3{P.viewInv[8], P.viewInv[9], P.viewInv[10]}, f3{P.viewInv[12], P.viewInv[13], P.viewInv[14]}};
  pr.ro = xform(vi, op, 1.0f);
  pr.rd = normalize(xform(vi, dir, 0.0f));
  if (P.aperture > 0.0f) {
    const f3 eye{P.viewInv[12], P.viewInv[13], P.viewInv[14]};
    const f3 forward = -normalize(f3{P.viewInv[8], P.viewInv[9], P.viewInv[10]});
    const f3 focusPoint = eye + pr.rd * (P.focus / dot(pr.rd, forward));
    float ux, uy, lx, ly;
    smp.s2(1u, ux, uy);
    sample_circle(ux, uy, lx, ly);
    const f3 origin = xform(vi, f3{P.aperture * lx, P.aperture * ly, 0.0f}, 1.0f);
    pr.ro = origin;
    pr.rd = normalize(focusPoint - origin);
  }
  pr.li = splat(0.0f);
  pr.beta = splat(1.0f);
  pr.auxAlbedo = splat(0.0f);
  pr.auxNormal = splat(0.0f);
  pr.alpha = 1.0f;
  pr.scatPdf = 0.0f;
  pr.curTri = 0;
  pr.curSlot = 0;
  pr.flags = kFromEye;
}

// The path depth: the `bounces` argument of the next closest-hit query
// (rayTrace.compute.ts:745, before the event's completeVertex increment).
__device__ __forceinline__ int path_depth(uint32_t flags) { return (int)(flags >> kDepthShift); }

// One integrator event after its closest-hit query `h` (rayTrace.compute.ts
// :891-946 = completeVertex :514-589 + shadeVertex :343-512 with
// SampleMaterial, sampleMaterial.wgsl.ts:316-508). Emission reached by the
// ray is added to pr.li; the vertex's NEE shadow ray, if any, is returned in
// `sr` (its payload is added to pr.li by the caller when unoccluded, before
// the next event). Returns true when the path continues with pr.ro / pr.rd.
// kFull = false compiles the event for scenes without fuzz, subsurface and
// thin-walled materials and without the random pattern; the launcher picks
// it only then, so it computes the same values as the full form.
template <bool kFull = true>
__device__ bool shade_event(const GpuScene &S, const RenderParams &P, const Lut &L, PathRegs &pr, const TraceHit &h,
                            int event, Smp &smp, ShadowReq &sr) {
  if (!kFull && smp.pattern == kPatternRandom) __builtin_unreachable();
  sr.valid = false;
  Surf si;
  si.hit = h.hit;
  si.tri = pr.curTri;
  si.slot = pr.curSlot;
  si.albedoSampled = (pr.flags & kAlbedoSampled) != 0u;
  Mat ma;
  if (h.hit) {
    complete_surface<kFull>(S, h, si);
    ma = load_material<kFull>(S.materials + (size_t)si.material * MTX_MATERIAL_FLOATS, si.color, &S, si.uv, si.uv1);
  }
  pr.curTri = si.tri;
  pr.curSlot = si.slot;
  const f3 camPos{P.viewInv[12], P.viewInv[13], P.viewInv[14]};
  const int bounceLimit = (int)P.bounces;
  const int eventLimit = min(bounceLimit * 3, 50);
  int depth = path_depth(pr.flags);
  bool abort = false;
  if (event == 0) {
    if (si.hit) {
      pr.auxAlbedo = ma.transmissionWeight > 0.5f ? splat(1.0f) : ma.baseColor;
      pr.auxNormal = dot(si.normal, si.position - camPos) > 0.0f ? -si.normal : si.normal;
      if (ma.transmissionWeight > 0.5f) pr.flags |= kAlbedoSampled; else pr.flags &= ~kAlbedoSampled;
      si.albedoSampled = ma.transmissionWeight > 0.5f;
      if ((ma.transmissionWeight < 0.001f) & (ma.subsurfaceWeight < 0.001f) & (ma.alpha > 0.99f))
        pr.flags |= kNotTransparent;
    }
  } else if (!(pr.flags & kIgnored)) {
    // completeVertex: the coverage draw (hashedSample(3u) < alpha) lies in
    // (0, 1), so against a coverage of 0 or 1 it is `alpha > 0`; a
    // passed-through surface does not occlude the light or the environment.
    smp.hashed_draw();
    const bool alphaOccluded = ma.alpha > 0.0f;
    for (uint32_t i = 0; i < S.lightCount; i++) {
      LSample ls;
      float d = intersect_light(S.lights + (size_t)i * 16, pr.ro, pr.rd, ls);
      ls.pdf *= S.lightPmf[i];
      bool occluded = si.hit && d > len(si.position - pr.ro);
      if (d != kInf && !(occluded && alphaOccluded)) {
        float mis = bsdf_mis(P.candidates, pr.scatPdf, ls.pdf);
        pr.li = pr.li + mis * pr.beta * ls.emission;
        pr.alpha = 1.0f;
        abort = true;
      }
    }
    if (S.env && !(si.hit && alphaOccluded)) {
      float u, v;
      env_uv(pr.rd, u, v);
      f3 em = texture_linear(S, u, v);
      if (pr.flags & kNextRefr) {
        float mis = (pr.flags & kIsDelta) ? 1.0f : bsdf_mis(P.candidates, pr.scatPdf, env_pdf(S, u, v) * S.envPmf);
        pr.li = pr.li + mis * pr.beta * em * 1.0f;
      } else {
        float mis = bsdf_mis(P.candidates, pr.scatPdf, env_pdf(S, u, v) * S.envPmf);
        pr.li = pr.li + pr.beta * em * mis;
      }
      pr.alpha = 1.0f;
      abort = true;
    }
    depth++;
    pr.flags = (pr.flags & ((1u << kDepthShift) - 1u)) | ((uint32_t)depth << kDepthShift);
  }
  if ((pr.flags & kNotTransparent) && event > bounceLimit) abort = true;
  if (abort || event >= eventLimit) return false;
  pr.flags &= ~kIgnored;
  smp.bounce = (uint32_t)depth;
  if (!si.hit) {
    if (pr.flags & kFromEye) {
      f3 bg = splat(0.0f);
      float ba = 0.0f;
      if (S.env) {
        float u, v;
        env_uv(pr.rd, u, v);
        bg = texture_linear(S, u, v);
        ba = 1.0f;
      }
      pr.li = pr.beta * bg * ba;
      pr.alpha = ba;
      if (ba < 0.001f) pr.flags |= kIsBlank;
    }
    return false;
  }
  if (!(ma.alpha > 0.0f)) {
    // A masked-out surface (shadeVertex :432-458): no material sample, no
    // guides, no roulette; the ray continues behind it from position +
    // d * EPS without consuming a depth.
    pr.flags |= kIgnored;
    pr.ro = si.position + kEps * pr.rd;
    return true;
  }
  // SampleMaterial (sampleMaterial.wgsl.ts:316-508; surface vertex, no media).
  {
    const f3 rd = pr.rd;
    float ndv1 = -dot(si.faceNormal, rd), ndv2 = -dot(si.normal, rd);
    if (((ma.transmissionWeight > 0.0f) | (ma.subsurfaceWeight > 0.0f)) & ma.thinWalled) {
      // a thin wall is always entered from the viewer's side (:328-338)
      if (ndv1 < 0.0f) { si.faceNormal = -si.faceNormal; ndv1 = -ndv1; }
      if (ndv2 < 0.0f) { si.normal = -si.normal; ndv2 = -ndv2; }
    }
    bool useFaceN = (ndv1 >= 0.0f) & (ndv2 < 0.0f);
    float ndv = useFaceN ? ndv1 : ndv2;
    f3 N = useFaceN ? si.faceNormal : si.normal;
    ma.inside = (ndv <= 0.0f) & !ma.thinWalled;
    ma.relativeIor = ndv > 0.0f ? 1.0f / ma.specularIor : ma.specularIor;
    f3 ff = ndv >= 0.0f ? N : -N;
    pr.flags &= ~kIsDelta;
    // Surface emission (sampleMaterial.wgsl.ts:412-422), weighted by
    // emissiveHitMisWeight (:279-303): one-sided unless double-sided, weight 1
    // for camera rays, else MIS against the emissive-triangle light density
    // (selection pmf / area = luminance / total light weight).
    {
      const float cosLight = dot(si.faceNormal, -rd);
      float w = 1.0f;
      if (cosLight <= 0.0f && S.materials[(size_t)si.material * MTX_MATERIAL_FLOATS + MTX_MAT_DOUBLE_SIDED] < 0.5f) {
        w = 0.0f;
      } else if (!((pr.flags & kFromEye) | (S.lightWeightTotal <= 0.0f))) {
        const float *em = S.materials + (size_t)si.material * MTX_MATERIAL_FLOATS + MTX_MAT_EMISSION;
        const float lum = dot(f3{em[0], em[1], em[2]}, f3{0.2126f, 0.7152f, 0.0722f});
        if (!((lum <= 0.0f) | (cosLight <= 0.0f))) {
          const float hitDistance = len(si.position - pr.ro);
          const float lightPdf = lum / S.lightWeightTotal * hitDistance * hitDistance / mxf(cosLight, 1e-8f);
          w = bsdf_mis(P.candidates, pr.scatPdf, lightPdf);
        }
      }
      pr.li = pr.li + w * ma.emission * pr.beta;
    }
    Prep PR;
    prepare(PR, ma, -rd, ff, L);
    const uint32_t cc = max(P.candidates, 1u);
    Shadow sel;
    sel.valid = false;
    float selTarget = 0.0f, wsum = 0.0f;
    f3 sdir = splat(0.0f), sresp = splat(0.0f);
    float spdf = 0.0f;
    bool strans = false, discrete = false;
    for (uint32_t ev = 0; ev <= cc; ev++) {
      const bool cont = ev == cc;
      LSample ls;
      ls.pdf = 0.0f; ls.selPmf = 0.0f; ls.isDelta = false; ls.sourceType = 0; ls.dist = kInf;
      f3 ldir;
      float lpdf = 0.0f;
      f3 resp = splat(0.0f);
      bool disc = false;
      if (cont) {
        float ucw = 1.0f;
        bool valid = sel.valid;
        if (cc > 1u) {
          valid = valid & (selTarget > 0.0f);
          if (valid) ucw = sel.selPmf * wsum / ((float)cc * selTarget);
        }
        if (valid) {
          sr.valid = true;
          sr.ro = sel.ro;
          sr.rd = sel.rd;
          sr.tMax = fminf(fmaxf(sel.maxDistance, 0.0f), kRayMaxDistance);
          sr.payload = sel.contribution * ucw;
        }
        float d0, d1;
        smp.s2(smp.bdim(3u), d0, d1);
        float lobe = smp.s1(smp.bdim(2u));
        disc = !PR.m.thinWalled & (fabsf(PR.m.relativeIor - 1.0f) <= 1e-6f) & (PR.p[LT] >= 1.0f - 1e-6f);
        if (disc) {
          ldir = to_world(PR.basis, -PR.v);
          resp = PR.m.transmissionColor * coat_base_scale(PR, -PR.v.z);
          lpdf = 1.0f;
          strans = true;
        } else {
          bool tr;
          if (!sample_bsdf(PR, d0, d1, lobe, ldir, tr)) { spdf = 0.0f; break; }
          strans = tr;
        }
        discrete = disc;
      } else {
        float a, b;
        smp.s2(smp.ldim(ev, 1u), a, b);
        float c = smp.s1(smp.ldim(ev, 0u));
        sample_light<kFull>(S, si.position, a, b, c, ev, cc, ls);
        ldir = ls.direction;
        if ((ls.pdf <= 0.0f) | (ls.selPmf <= 0.0f) | (maxc(ls.emission) <= 0.0f)) continue;
        if ((dot(ff, ldir) <= 0.0f) & (ma.transmissionWeight <= 0.0f) & (ma.subsurfaceWeight <= 0.0f)) continue;
      }
      if (!disc) resp = eval_bsdf(PR, ldir, lpdf, L);
      if (cont) {
        sdir = ldir;
        sresp = resp;
        spdf = lpdf;
        break;
      }
      // finishDirectLight (sampleMaterial.wgsl.ts:32-151)
      f3 Li = ls.emission;
      if ((ls.pdf <= 0.0f) | (ls.selPmf <= 0.0f) | (maxc(Li) <= 0.0f)) continue;
      if ((lpdf <= 0.0f) | (maxc(resp) <= 0.0f)) continue;
      f3 origin = offset_surface_ray(si.position, si.faceNormal, ldir);
      float nl = dot(si.normal, ldir);
      f3 gn = nl < 0.0f ? -si.faceNormal : si.faceNormal;
      nl = fabsf(nl);
      float gl = dot(gn, ldir);
      float amount = nl < 0.1f ? clampf(2.0f - (gl + nl) / 0.1f, 0.0f, 1.0f) : clampf(1.0f - gl / 0.1f, 0.0f, 1.0f);
      origin = origin + si.shadowOffset * amount;
      if ((dot(ff, ldir) < 0.0f) & (ls.sourceType == 0u)) {
        if (S.env) {
          float u, v;
          env_uv(ldir, u, v);
          Li = texture_linear(S, u, v);
        } else {
          Li = splat(0.0f);
        }
      }
      float mis = ls.isDelta ? 1.0f : nee_mis(P.candidates, ls.pdf, lpdf);
      if (!(mis > 0.0f)) continue;
      f3 un = Li * resp / ls.pdf * pr.beta;
      Shadow cand;
      cand.risTarget = mxf(dot(vmax(un * ls.selPmf, splat(0.0f)), f3{0.2126f, 0.7152f, 0.0722f}), 0.0f);
      cand.ro = origin;
      cand.rd = ldir;
      cand.maxDistance = ls.dist - kEps;
      if (ls.sourceType == 1u) {
        // Emissive triangles: the visibility ray ends EPS off the sampled
        // point along the light normal (sampleMaterial.wgsl.ts:98-110).
        const f3 lightEndpoint = si.position + ls.direction * ls.dist;
        const f3 toOffsetLight = (lightEndpoint + kEps * ls.normal) - origin;
        cand.maxDistance = len(toOffsetLight);
        cand.rd = normalize(toOffsetLight);
      }
      cand.contribution = mis * un;
      cand.selPmf = ls.selPmf;
      cand.valid = true;
      if (cc == 1u) { sel = cand; continue; }
      float sw = cand.risTarget / mxf(cand.selPmf, 1e-20f);
      if (!(sw > 0.0f) | (sw != sw)) continue;
      wsum += sw;
      if (!sel.valid || smp.candidate_random(ev) * wsum <= sw) {
        sel = cand;
        selTarget = cand.risTarget;
      }
    }
    bool finite = isfinite(sresp.x) && isfinite(sresp.y) && isfinite(sresp.z) && fabsf(sresp.x) < 3.402823e38f &&
                  fabsf(sresp.y) < 3.402823e38f && fabsf(sresp.z) < 3.402823e38f;
    if (!(spdf > 0.0f) | !finite | (dot(sdir, sdir) < 0.5f)) {
      pr.scatPdf = 0.0f;
      abort = true;
    } else {
      if (discrete) pr.flags |= kIsDelta;
      pr.beta = pr.beta * (sresp / spdf);
      if (pr.beta.x == 0.0f && pr.beta.y == 0.0f && pr.beta.z == 0.0f) {
        abort = true;
      } else {
        if (strans) pr.flags |= kNextRefr; else pr.flags &= ~kNextRefr;
        pr.ro = si.position + kEps * sdir;
        pr.scatPdf = spdf;
        pr.rd = sdir;
        pr.flags &= ~kFromEye;
      }
    }
  }
  if (!si.albedoSampled) {
    if (!(ma.transmissionWeight > 0.5f || (ma.specularRoughness <= 0.02f && ma.baseMetalness > 0.98f))) {
      pr.auxAlbedo = ma.baseColor;
      pr.auxNormal = dot(si.normal, si.position - camPos) > 0.0f ? -si.normal : si.normal;
      si.albedoSampled = true;
    }
    if (depth > 4) si.albedoSampled = true;
  }
  if (si.albedoSampled) pr.flags |= kAlbedoSampled;
  if (depth >= 2) {
    float q = mxf(0.0f, 1.0f - maxc(pr.beta));
    if (smp.s1(smp.bdim(7u)) < q) {
      pr.alpha = 1.0f;
      abort = true;
    } else {
      pr.beta = pr.beta / (1.0f - q);
    }
  }
  return !abort;
}

// Per-sample record added to the FP64 accumulators after the clamp and the
// NaN/negative guard of rayTrace.compute.ts:1240-1261.
struct SampleRec {
  f3 li, albedo, normal;
  float alpha;
  bool blank;
};
__device__ __forceinline__ SampleRec finish_path(const RenderParams &P, const PathRegs &pr) {
  SampleRec r;
  r.li = pr.li;
  r.alpha = pr.alpha;
  if (P.clamping > 0.0f) r.li = clamp3(r.li, -P.clamping, P.clamping);
  if (!(r.li.x < kInf && r.li.x > -kEps)) { r.li = splat(0.0f); r.alpha = 1.0f; }
  r.albedo = pr.auxAlbedo;
  r.normal = pr.auxNormal;
  r.blank = (pr.flags & kIsBlank) != 0u;
  return r;
}
// Adds one sample to the 13 running sums (color rgba, albedo rgb + hit,
// normal rgb + 1, luminance squared).
// The 13 values one sample adds.
__device__ __forceinline__ void sample_values(const SampleRec &r, double v[13]) {
  v[0] = (double)r.li.x; v[1] = (double)r.li.y; v[2] = (double)r.li.z; v[3] = (double)r.alpha;
  v[4] = (double)r.albedo.x; v[5] = (double)r.albedo.y; v[6] = (double)r.albedo.z;
  v[7] = r.blank ? 0.0 : 1.0;
  v[8] = (double)r.normal.x; v[9] = (double)r.normal.y; v[10] = (double)r.normal.z;
  v[11] = 1.0;
  const double lum = (double)dot(r.li, f3{0.2126f, 0.7152f, 0.0722f});
  v[12] = lum * lum;
}
__device__ __forceinline__ void add_sample(double s[13], const SampleRec &r) {
  double v[13];
  sample_values(r, v);
#pragma unroll
  for (int c = 0; c < 13; c++) s[c] += v[c];
}

}  // namespace mtx
