// Oracle: RGB megakernel integrator, OpenPBR lobes, lights (test infrastructure only).
// See orc_render.h for the reference files restated here.
#include "orc_render.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "orc_sobol.h"
#include "../include/mtx_fmath.h"

namespace orc {

// ------------------------------------------------------------------ constants
// WGSL folds constant expressions in abstract (f64) precision and rounds once.
static const float K_PI = (float)3.14159265359;
static const float K_QUARTER_PI = (float)(0.25 * 3.14159265359);
static const float K_HALF_PI = (float)(0.5 * 3.14159265359);
static const float K_ALPHA_MIN = (float)(0.001 * 0.001);
static const float K_FON_A = (float)(0.5 - 2.0 / (3.0 * 3.14159265359));
static const float K_FON_B = (float)(2.0 / 3.0 - 28.0 / (15.0 * 3.14159265359));
static const float K_EDGE_COS = (float)(1.0 / 7.0);
static const V3 K_LUM(0.2126f, 0.7152f, 0.0722f);  // rayTraceCore.wgsl.ts:45

static const float *gLut = nullptr;
void set_energy_lut(const float *lut) { gLut = lut; }
static inline float lutf(int col, int row) { return gLut[row * 32 + col]; }
enum { LUT_METAL_ROW = 1056, LUT_METAL_AVG_ROW = 1088, LUT_DIEL_ROW = 1089, LUT_DIEL_AVG_ROW = 2113,
       LUT_RATIO_ROW = 2145 };

static inline float fractf(float x) { return x - floorf(x); }
static inline V3 vpow(V3 a, V3 b) { return V3(mtxfm::powf(a.x, b.x), mtxfm::powf(a.y, b.y), mtxfm::powf(a.z, b.z)); }
static inline V3 vmaxs(V3 a, float s) { return vmax(a, V3(s)); }
static inline V3 clamp3(V3 v, float lo, float hi) { return V3(clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)); }
static inline V3 reflect(V3 e1, V3 e2) { return e1 - (2.0f * dot(e2, e1)) * e2; }
static inline V3 refract(V3 e1, V3 e2, float e3) {
  float c = dot(e2, e1);
  float k = 1.0f - e3 * e3 * (1.0f - c * c);
  if (k < 0.0f) return V3(0.0f);
  return e3 * e1 - (e3 * c + std::sqrt(k)) * e2;
}

// ------------------------------------------------------------------ sampling.wgsl.ts
static V3 ggx_visible_normal(V3 view, float ax, float ay, float sx, float sy) {
  V3 stretch(ax, ay, 1.0f);
  V3 sv = normalize(stretch * view);
  float vz = mnf(sv.z, 1.0f);
  float capSpan = 1.0f + vz;
  float nz = (1.0f - sy) * capSpan;
  float omz = sy * capSpan;
  float opz = nz + (1.0f - vz);
  float radial = std::sqrt(mxf(0.0f, opz * omz));
  float az = TWOPI * sx;
  V3 sn(sv.x + radial * mtxfm::cosf(az), sv.y + radial * mtxfm::sinf(az), nz);
  return normalize(stretch * sn);
}
static float ggx_ndf(V3 h, float ax, float ay) {
  V3 e(h.x / ax, h.y / ay, h.z);
  float den = dot(e, e);
  return 1.0f / (PI * ax * ay * den * den);
}
static float smith_g1(V3 d, float ax, float ay) {
  float nc = fabsf(d.z);
  if (nc <= 1e-8f) return 0.0f;
  float rx = ax * d.x, ry = ay * d.y;
  float tan2 = (rx * rx + ry * ry) / (nc * nc);
  return 2.0f / (1.0f + std::sqrt(1.0f + tan2));
}
static float schlick5(float c) { return mtxfm::powf(clampf(1.0f - c, 0.0f, 1.0f), 5.0f); }
static float dielectric_fresnel_s(float incidentCos, float eta) {
  float cosI = clampf(fabsf(incidentCos), 0.0f, 1.0f);
  float sinT2 = eta * eta * mxf(0.0f, 1.0f - cosI * cosI);
  if (sinT2 >= 1.0f) return 1.0f;
  float cosT = std::sqrt(mxf(0.0f, 1.0f - sinT2));
  float perp = (cosI - eta * cosT) / mxf(cosI + eta * cosT, 1e-8f);
  float par = (eta * cosI - cosT) / mxf(eta * cosI + cosT, 1e-8f);
  return 0.5f * (perp * perp + par * par);
}
static void concentric_disk(float sx, float sy, float &ox, float &oy) {
  float x = 2.0f * sx - 1.0f, y = 2.0f * sy - 1.0f;
  if (x == 0.0f && y == 0.0f) { ox = oy = 0.0f; return; }
  float r, a;
  if (fabsf(x) > fabsf(y)) { r = x; a = K_QUARTER_PI * (y / x); }
  else { r = y; a = K_HALF_PI - K_QUARTER_PI * (x / y); }
  ox = r * mtxfm::cosf(a);
  oy = r * mtxfm::sinf(a);
}
static V3 cosine_hemisphere(float sx, float sy) {
  float dx, dy;
  concentric_disk(sx, sy, dx, dy);
  return V3(dx, dy, std::sqrt(mxf(0.0f, 1.0f - (dx * dx + dy * dy))));
}
static float power_heuristic(float a, float b) {
  float scale = mxf(a, b);
  if (scale <= 0.0f) return 0.0f;
  float na = a / scale, nb = b / scale;
  float sa = na * na, sb = nb * nb;
  return sa / (sa + sb);
}
static void orthonormal_basis(V3 n, V3 &t, V3 &b) {
  float signZ = n.z >= 0.0f ? 1.0f : -1.0f;
  float factor = -1.0f / (signZ + n.z);
  float xy = n.x * n.y * factor;
  t = V3(1.0f + signZ * n.x * n.x * factor, signZ * xy, -signZ * n.x);
  b = V3(xy, signZ + n.y * n.y * factor, -n.y);
}

// ------------------------------------------------------------------ materials
struct Mat {
  V3 baseColor;
  float baseWeight, diffuseRoughness, baseMetalness;
  float specularWeight;
  V3 specularColor;
  float specularRoughness, specularIor, specularAnisotropy;
  float transmissionWeight;
  V3 transmissionColor;
  float coatWeight;
  V3 coatColor;
  float coatRoughness, coatIor, coatDarkening;
  float fuzzWeight;
  V3 fuzzColor;
  float fuzzRoughness;
  float subsurfaceWeight;
  V3 subsurfaceColor;
  bool thinWalled;
  V3 emission;
  float alpha;
  bool inside;
  float relativeIor, alphaX, alphaY;
};

static void roughness_to_alpha(float r, float aniso, float &ax, float &ay) {
  float a = r * r;
  if (aniso == 0.0f) { ax = ay = mxf(K_ALPHA_MIN, a); return; }
  float c = 1.0f - aniso;
  float at = a * std::sqrt(2.0f / (1.0f + c * c));
  ax = mxf(K_ALPHA_MIN, at);
  ay = mxf(K_ALPHA_MIN, c * at);
}

// ------------------------------------------------------------------ textures
// UV pair (WGSL vec2f).
struct UV { float x, y; };

// wrapMapCoordinate (materialBuffer.wgsl.ts:285-289): 0 repeat (fract),
// 1 clamp to edge, 2 mirrored repeat.
static float wrap_map_coordinate(float value, uint32_t mode) {
  if (mode == 0u) return value - std::floor(value);
  if (mode == 1u) return clampf(value, 0.0f, 1.0f);
  return 1.0f - fabsf(1.0f - (value - 2.0f * std::floor(value * 0.5f)));
}
// One texel (textureLoad of the atlas): RGBA8 words are unorm bytes, R in
// the low byte, each byte / 255; RGBA32F texels are read as stored.
static void load_texel(const Texture &t, int x, int y, float out[4]) {
  const size_t i = (size_t)(uint32_t)y * t.width + (uint32_t)x;
  if (t.storage == 1u) {
    for (int k = 0; k < 4; k++) out[k] = t.rgba32[i * 4 + k];
    return;
  }
  const uint32_t w = t.rgba8[i];
  for (int k = 0; k < 4; k++) out[k] = (float)((w >> (8 * k)) & 255u) / 255.0f;
}
// loadMapTexel (materialBuffer.wgsl.ts:338-348): repeat wraps the integer
// coordinate modulo the size, the other modes clamp it to the image.
static int map_texel_coordinate(int c, int size, uint32_t mode) {
  if (mode == 0u) return ((c % size) + size) % size;
  return c < 0 ? 0 : (c > size - 1 ? size - 1 : c);
}
// sampleMap (materialBuffer.wgsl.ts:314-336) for an ordinary image: the
// coordinate uv * scale + offset wrapped per axis (wrappedMapUv, :291-297),
// texel = wrapped * size - 0.5, the four texels around it blended with
// mix(mix(t00, t10, fx), mix(t01, t11, fx), fy). The reference hands interior
// footprints to the hardware sampler; this is its explicit four-tap form,
// used for every footprint. A NaN wrapped coordinate reads as 0.
static void sample_map(const TextureSet &T, uint32_t index, UV uv, float out[4]) {
  const Texture &t = T.tex[index];
  float wu = wrap_map_coordinate(uv.x * t.sx + t.ox, t.wrapS);
  float wv = wrap_map_coordinate(uv.y * t.sy + t.oy, t.wrapT);
  if (wu != wu) wu = 0.0f;
  if (wv != wv) wv = 0.0f;
  const int W = (int)t.width, H = (int)t.height;
  const float tx = wu * (float)W - 0.5f, ty = wv * (float)H - 0.5f;
  const float bx = std::floor(tx), by = std::floor(ty);
  const float fx = tx - bx, fy = ty - by;
  const int x0 = map_texel_coordinate((int)bx, W, t.wrapS), x1 = map_texel_coordinate((int)bx + 1, W, t.wrapS);
  const int y0 = map_texel_coordinate((int)by, H, t.wrapT), y1 = map_texel_coordinate((int)by + 1, H, t.wrapT);
  float a[4], b[4], c[4], d[4];
  load_texel(t, x0, y0, a);
  load_texel(t, x1, y0, b);
  load_texel(t, x0, y1, c);
  load_texel(t, x1, y1, d);
  for (int k = 0; k < 4; k++) out[k] = mixf(mixf(a[k], b[k], fx), mixf(c[k], d[k], fx), fy);
}
// sRGBToLinear (materialBuffer.wgsl.ts:258-260) per channel.
static float srgb_to_linear(float c) {
  return c < 0.04045f ? c * 0.0773993808f : mtxfm::powf(c * 0.9478672986f + 0.0521327014f, 2.4f);
}
// decodeMapColor (:262-265): sRGB decode unless the texture stores floats.
static V3 decode_map_color(const TextureSet &T, uint32_t index, const float c[4]) {
  if (T.tex[index].storage == 1u) return V3(c[0], c[1], c[2]);
  return V3(srgb_to_linear(c[0]), srgb_to_linear(c[1]), srgb_to_linear(c[2]));
}
// The texture of slot k (index + 1 stored, 0 = none) sampled at its UV set
// (getMapUvSet, :303-305); returns false when the slot is empty.
static bool material_map(const TextureSet &T, const float *m, int k, UV uv, UV uv1, float out[4],
                         uint32_t *index = nullptr) {
  const uint32_t slot = (uint32_t)m[M_MAPS + k];
  if (slot == 0u) return false;
  sample_map(T, slot - 1u, T.tex[slot - 1u].uvChannel ? uv1 : uv, out);
  if (index) *index = slot - 1u;
  return true;
}
void probe_sample_map(const TextureSet &T, uint32_t index, float u, float v, float rgba[4], float decoded[3]) {
  sample_map(T, index, UV{u, v}, rgba);
  V3 d = decode_map_color(T, index, rgba);
  decoded[0] = d.x; decoded[1] = d.y; decoded[2] = d.z;
}

// getMaterialEmission (materialBuffer.wgsl.ts:351-372).
static V3 material_emission(const TextureSet *T, const float *m, UV uv, UV uv1) {
  V3 e(m[M_EMISSION], m[M_EMISSION + 1], m[M_EMISSION + 2]);
  if (!T) return e;
  float c[4];
  uint32_t index;
  if (material_map(*T, m, MAP_EMISSIVE, uv, uv1, c, &index)) e = e * decode_map_color(*T, index, c);
  if (material_map(*T, m, MAP_COAT_WEIGHT, uv, uv1, c)) {
    float coatWeight = clampf(m[M_COAT_WEIGHT], 0.0f, 1.0f) * c[0];
    V3 tint(std::sqrt(mxf(m[M_COAT_COLOR], 0.0f)), std::sqrt(mxf(m[M_COAT_COLOR + 1], 0.0f)),
            std::sqrt(mxf(m[M_COAT_COLOR + 2], 0.0f)));
    e = e * mix3(V3(1.0f), tint, coatWeight);
  }
  return e;
}

// getMaterial (materialBuffer.wgsl.ts:170-245), with the texture samples of
// :78-160 when the scene has textures and the alpha-mode resolution of
// :66-77 (opaque -> 1, mask -> alpha >= cutoff ? 1 : 0).
static Mat get_material(const float *m, V3 instanceColor, const TextureSet *T = nullptr, UV uv = UV{0.0f, 0.0f},
                        UV uv1 = UV{0.0f, 0.0f}) {
  Mat ma{};
  ma.baseColor = V3(m[M_BASE_COLOR], m[M_BASE_COLOR + 1], m[M_BASE_COLOR + 2]);
  ma.specularRoughness = m[M_SPEC_ROUGHNESS];
  ma.baseMetalness = m[M_METALNESS];
  ma.specularAnisotropy = clampf(m[M_SPEC_ANISOTROPY], 0.0f, 1.0f);
  ma.emission = V3(m[M_EMISSION], m[M_EMISSION + 1], m[M_EMISSION + 2]);
  ma.inside = false;
  ma.specularIor = m[M_SPEC_IOR];
  ma.transmissionWeight = m[M_TRANS_WEIGHT];
  ma.alpha = m[M_OPACITY];
  ma.baseWeight = clampf(m[M_BASE_WEIGHT], 0.0f, 1.0f);
  ma.diffuseRoughness = clampf(m[M_DIFFUSE_ROUGHNESS], 0.0f, 1.0f);
  ma.specularWeight = mxf(m[M_SPEC_WEIGHT], 0.0f);
  ma.specularColor = vmaxs(V3(m[M_SPEC_COLOR], m[M_SPEC_COLOR + 1], m[M_SPEC_COLOR + 2]), 0.0f);
  ma.transmissionColor = vmaxs(V3(m[M_TRANS_COLOR], m[M_TRANS_COLOR + 1], m[M_TRANS_COLOR + 2]), 0.0f);
  ma.coatWeight = clampf(m[M_COAT_WEIGHT], 0.0f, 1.0f);
  ma.coatColor = vmaxs(V3(m[M_COAT_COLOR], m[M_COAT_COLOR + 1], m[M_COAT_COLOR + 2]), 0.0f);
  ma.coatRoughness = clampf(m[M_COAT_ROUGHNESS], ROUGHNESS_MIN, 1.0f);
  ma.coatIor = mxf(m[M_COAT_IOR], 1.0f);
  ma.coatDarkening = clampf(m[M_COAT_DARKENING], 0.0f, 1.0f);
  ma.fuzzWeight = clampf(m[M_FUZZ_WEIGHT], 0.0f, 1.0f);
  ma.fuzzColor = vmaxs(V3(m[M_FUZZ_COLOR], m[M_FUZZ_COLOR + 1], m[M_FUZZ_COLOR + 2]), 0.0f);
  ma.fuzzRoughness = clampf(m[M_FUZZ_ROUGHNESS], 0.0f, 1.0f);
  ma.subsurfaceWeight = clampf(m[M_SSS_WEIGHT], 0.0f, 1.0f);
  ma.subsurfaceColor = vmaxs(V3(m[M_SSS_COLOR], m[M_SSS_COLOR + 1], m[M_SSS_COLOR + 2]), 0.0f);
  ma.thinWalled = m[M_THIN_WALLED] > 0.5f;
  if (T) {
    ma.emission = material_emission(T, m, uv, uv1);
    float c[4];
    uint32_t index;
    if (material_map(*T, m, MAP_BASE, uv, uv1, c, &index)) {
      ma.baseColor = ma.baseColor * decode_map_color(*T, index, c);
      if (m[M_MAP_ALPHA_USAGE] == 3.0f) ma.transmissionWeight = 1.0f - c[3];
      else if (m[M_MAP_ALPHA_USAGE] == 2.0f) ma.alpha = ma.alpha * c[3];
    }
    if (material_map(*T, m, MAP_ALPHA, uv, uv1, c)) ma.alpha = ma.alpha * c[1];
    if (material_map(*T, m, MAP_TRANSMISSION, uv, uv1, c)) ma.transmissionWeight = ma.transmissionWeight * c[0];
    if (material_map(*T, m, MAP_METALLIC_ROUGHNESS, uv, uv1, c)) {
      ma.baseMetalness = ma.baseMetalness * c[2];
      ma.specularRoughness = ma.specularRoughness * c[1];
    }
    if (material_map(*T, m, MAP_ROUGHNESS, uv, uv1, c)) ma.specularRoughness = ma.specularRoughness * c[1];
    if (material_map(*T, m, MAP_METALNESS, uv, uv1, c)) ma.baseMetalness = ma.baseMetalness * c[2];
    if (material_map(*T, m, MAP_COAT_WEIGHT, uv, uv1, c)) ma.coatWeight = ma.coatWeight * c[0];
    if (material_map(*T, m, MAP_COAT_ROUGHNESS, uv, uv1, c)) ma.coatRoughness = ma.coatRoughness * c[1];
    if (material_map(*T, m, MAP_FUZZ_COLOR, uv, uv1, c, &index)) ma.fuzzColor = ma.fuzzColor * decode_map_color(*T, index, c);
    if (material_map(*T, m, MAP_FUZZ_ROUGHNESS, uv, uv1, c)) ma.fuzzRoughness = ma.fuzzRoughness * c[3];
    if (material_map(*T, m, MAP_SPECULAR_WEIGHT, uv, uv1, c)) ma.specularWeight = ma.specularWeight * c[3];
    if (material_map(*T, m, MAP_SPECULAR_COLOR, uv, uv1, c, &index))
      ma.specularColor = ma.specularColor * decode_map_color(*T, index, c);
  }
  if (m[M_ALPHA_MODE] == 0.0f) ma.alpha = 1.0f;
  else if (m[M_ALPHA_MODE] == 1.0f) ma.alpha = ma.alpha >= m[M_ALPHA_CUTOFF] ? 1.0f : 0.0f;
  if (ma.coatWeight > 0.0f && ma.coatRoughness > 0.0f) {
    float b2 = ma.specularRoughness * ma.specularRoughness;
    float c2 = ma.coatRoughness * ma.coatRoughness;
    ma.specularRoughness = mixf(ma.specularRoughness, mtxfm::powf(mnf(1.0f, b2 * b2 + 2.0f * c2 * c2), 0.25f), ma.coatWeight);
  }
  ma.baseColor = ma.baseColor * instanceColor;
  ma.specularRoughness = clampf(ma.specularRoughness, ROUGHNESS_MIN, 1.0f);
  ma.baseMetalness = clampf(ma.baseMetalness, 0.0f, 1.0f);
  roughness_to_alpha(ma.specularRoughness, ma.specularAnisotropy, ma.alphaX, ma.alphaY);
  return ma;
}

// ------------------------------------------------------------------ OpenPBR
enum { L_FUZZ = 0, L_COAT = 1, L_METAL = 2, L_SPEC = 3, L_TRANS = 4, L_DIFF = 5, L_SSS = 6, L_COUNT = 7 };

struct Basis { V3 t, b, n; };
struct Prepared {
  Mat material;
  Basis basis;
  V3 viewLocal, diffuseColor;
  float specularEta, refractionEta, transmissionReflectionIor, specularAlpha;
  float diffuseSpecularEnergyCompensation;
  V3 coatBaseScale, metalMsScale;
  float metalMsFactor, metalMsProbability;
  V3 dielMsReflScale, dielMsTransScale;
  float dielMsViewLoss, dielMsRatio, dielMsProbability;
  float w[L_COUNT], p[L_COUNT];
};

static float lum_pbr(V3 c) { return dot(c, V3(0.212671f, 0.715160f, 0.072169f)); }
static float f0_from_ior(float e) { float v = (e - 1.0f) / (e + 1.0f); return v * v; }
static float apply_specular_weight_to_ior(float eta, float w) {
  float f0 = mnf(f0_from_ior(eta) * w, 0.9999f);
  float s = std::sqrt(f0);
  float ext = (1.0f + s) / (1.0f - s);
  return eta < 1.0f ? 1.0f / ext : ext;
}
static float pbr_dielectric_fresnel(float c, float rel) {
  if (fabsf(rel - 1.0f) < 1e-6f) return 0.0f;
  return dielectric_fresnel_s(c, rel);
}
static float energy_ior_index(float ior) {
  if (ior < 1.0f) return 15.0f - ((1.0f / ior - 1.0f) / 1.5f) * 15.0f;
  return 16.0f + ((ior - 1.0f) / 1.5f) * 15.0f;
}
static float trilinear(int rowBase, float eI, float eA, float eC) {
  float iI = clampf(eI, 0.0f, 31.0f), aI = clampf(eA, 0.0f, 31.0f), cI = clampf(eC, 0.0f, 31.0f);
  int i0 = (int)floorf(iI), a0 = (int)floorf(aI), c0 = (int)floorf(cI);
  int i1 = std::min(i0 + 1, 31), a1 = std::min(a0 + 1, 31), c1 = std::min(c0 + 1, 31);
  float im = iI - (float)i0, am = aI - (float)a0, cm = cI - (float)c0;
  float v000 = lutf(c0, rowBase + i0 * 32 + a0), v001 = lutf(c1, rowBase + i0 * 32 + a0);
  float v010 = lutf(c0, rowBase + i0 * 32 + a1), v011 = lutf(c1, rowBase + i0 * 32 + a1);
  float v100 = lutf(c0, rowBase + i1 * 32 + a0), v101 = lutf(c1, rowBase + i1 * 32 + a0);
  float v110 = lutf(c0, rowBase + i1 * 32 + a1), v111 = lutf(c1, rowBase + i1 * 32 + a1);
  float v00 = mixf(v000, v100, im), v01 = mixf(v001, v101, im);
  float v10 = mixf(v010, v110, im), v11 = mixf(v011, v111, im);
  return mixf(mixf(v00, v10, am), mixf(v01, v11, am), cm);
}
static float bilinear(int rowBase, float eI, float eA) {
  float iI = clampf(eI, 0.0f, 31.0f), aI = clampf(eA, 0.0f, 31.0f);
  int i0 = (int)floorf(iI), a0 = (int)floorf(aI);
  int i1 = std::min(i0 + 1, 31), a1 = std::min(a0 + 1, 31);
  float im = iI - (float)i0, am = aI - (float)a0;
  float v00 = lutf(a0, rowBase + i0), v01 = lutf(a1, rowBase + i0);
  float v10 = lutf(a0, rowBase + i1), v11 = lutf(a1, rowBase + i1);
  return mixf(mixf(v00, v10, im), mixf(v01, v11, im), am);
}
static float ideal_loss(float ior, float a, float c) { return trilinear(LUT_DIEL_ROW, energy_ior_index(ior), std::sqrt(a) * 31.0f, c * 31.0f); }
static float ideal_avg_loss(float ior, float a) { return bilinear(LUT_DIEL_AVG_ROW, energy_ior_index(ior), std::sqrt(a) * 31.0f); }
static float ideal_ratio(float ior, float a) { return bilinear(LUT_RATIO_ROW, energy_ior_index(ior), std::sqrt(a) * 31.0f); }
static float metal_energy_complement(float a, float c) {
  float rI = clampf(std::sqrt(a) * 31.0f, 0.0f, 31.0f), cI = clampf(c * 31.0f, 0.0f, 31.0f);
  int r0 = (int)floorf(rI), c0 = (int)floorf(cI);
  int r1 = std::min(r0 + 1, 31), c1 = std::min(c0 + 1, 31);
  float fc = fractf(cI), fr = fractf(rI);
  return mixf(mixf(lutf(c0, LUT_METAL_ROW + r0), lutf(c1, LUT_METAL_ROW + r0), fc),
              mixf(lutf(c0, LUT_METAL_ROW + r1), lutf(c1, LUT_METAL_ROW + r1), fc), fr);
}
static float metal_average_energy_complement(float a) {
  float idx = clampf(std::sqrt(a) * 31.0f, 0.0f, 31.0f);
  int i0 = (int)floorf(idx);
  return mixf(lutf(i0, LUT_METAL_AVG_ROW), lutf(std::min(i0 + 1, 31), LUT_METAL_AVG_ROW), fractf(idx));
}
static float extrapolate(float v, float ior) {
  if (ior > 2.5f || ior < 0.4f) {
    float maxF0 = f0_from_ior(2.5f);
    float progress = (f0_from_ior(ior) - maxF0) / (1.0f - maxF0);
    return (1.0f - progress) * v;
  }
  return v;
}
static float opaque_energy_complement(float ior, float a, float c) {
  return extrapolate(trilinear(0, energy_ior_index(ior), std::sqrt(a) * 31.0f, c * 31.0f), ior);
}
static float opaque_average_energy_complement(float ior, float a) {
  return extrapolate(bilinear(1024, energy_ior_index(ior), std::sqrt(a) * 31.0f), ior);
}
static V3 to_world(const Basis &B, V3 d) { return d.x * B.t + d.y * B.b + d.z * B.n; }
static V3 to_local(const Basis &B, V3 d) { return V3(dot(d, B.t), dot(d, B.b), dot(d, B.n)); }

static V3 metal_fresnel(float c, V3 f0, V3 tint) {
  V3 schlick = mix3(f0, V3(1.0f), schlick5(c));
  if (tint.x == 1.0f && tint.y == 1.0f && tint.z == 1.0f) return schlick;
  float edgeComp = 1.0f - K_EDGE_COS;
  V3 edgeF = mix3(f0, V3(1.0f), schlick5(K_EDGE_COS));
  float comp = 1.0f - c;
  float correction = c * schlick5(c) * comp / (K_EDGE_COS * mtxfm::powf(edgeComp, 6.0f));
  return clamp3(schlick - correction * (edgeF - tint * edgeF), 0.0f, 1.0f);
}
static V3 eval_diffuse(const Mat &m, V3 color, V3 v, V3 l, float comp, float &pdf) {
  pdf = 0.0f;
  if (l.z <= 0.0f) return V3(0.0f);
  float r = m.diffuseRoughness;
  float muI = l.z, muO = v.z;
  float directional = dot(l, v) - muI * muO;
  float dot_ = directional > 0.0f ? directional / mxf(mxf(muI, muO), 1e-7f) : directional;
  float ssA = 1.0f / (1.0f + K_FON_A * r);
  V3 single = color * INVPI * ssA * (1.0f + r * dot_);
  float ic = 1.0f - muI, oc = 1.0f - muO;
  float iG = ic * (0.0571085289f + ic * (0.491881867f + ic * (-0.332181442f + ic * 0.0714429953f)));
  float oG = oc * (0.0571085289f + oc * (0.491881867f + oc * (-0.332181442f + oc * 0.0714429953f)));
  float iAlb = (1.0f + r * iG) / (1.0f + K_FON_A * r);
  float oAlb = (1.0f + r * oG) / (1.0f + K_FON_A * r);
  float avg = ssA * (1.0f + K_FON_B * r);
  V3 msAlb = color * color * avg / vmax(V3(1e-7f), V3(1.0f) - color * (1.0f - avg));
  V3 multiple = msAlb * INVPI * mxf(1e-7f, 1.0f - oAlb) * mxf(1e-7f, 1.0f - iAlb) / mxf(1e-7f, 1.0f - avg);
  pdf = l.z * INVPI;
  return (single + multiple) * comp;
}
static V3 eval_refl(float ax, float ay, V3 v, V3 l, V3 h, V3 F, float &pdf) {
  pdf = 0.0f;
  if (l.z <= 0.0f) return V3(0.0f);
  float D = ggx_ndf(h, ax, ay);
  float G1 = smith_g1(v, ax, ay);
  float G2 = G1 * smith_g1(l, ax, ay);
  pdf = G1 * D / (4.0f * mxf(v.z, 1e-8f));
  return F * D * G2 / (4.0f * mxf(l.z * v.z, 1e-8f));
}
static V3 eval_refr(float ax, float ay, float rel, V3 tint, V3 v, V3 l, V3 h, V3 F, float &pdf) {
  pdf = 0.0f;
  if (l.z >= 0.0f) return V3(0.0f);
  float lh = dot(l, h), vh = dot(v, h);
  if (vh <= 0.0f || lh >= 0.0f) return V3(0.0f);
  float D = ggx_ndf(h, ax, ay);
  float G1 = smith_g1(v, ax, ay);
  float G2 = G1 * smith_g1(l, ax, ay);
  float den = lh + vh * rel;
  den *= den;
  float jac = fabsf(lh) / mxf(den, 1e-8f);
  float rel2 = rel * rel;
  pdf = G1 * mxf(0.0f, vh) * D * jac / mxf(v.z, 1e-8f);
  return tint * (V3(1.0f) - F) * D * G2 * fabsf(vh) * jac * rel2 / mxf(fabsf(l.z * v.z), 1e-8f);
}
static float average_fresnel(float eta) {
  if (fabsf(eta - 1.0f) < 1e-6f) return 0.0f;
  if (eta > 1.0f) return (eta - 1.0f) / (4.08567f + 1.00071f * eta);
  return 0.997118f + 0.1014f * eta - 0.965241f * eta * eta - 0.130607f * eta * eta * eta;
}
static V3 metal_average_fresnel(V3 f0, V3 tint) {
  V3 edgeF = mix3(f0, V3(1.0f), schlick5(K_EDGE_COS));
  V3 b = edgeF * (V3(1.0f) - tint) / (K_EDGE_COS * mtxfm::powf(1.0f - K_EDGE_COS, 6.0f));
  return clamp3(f0 + (V3(1.0f) - f0) / 21.0f - b / 126.0f, 0.0f, 1.0f);
}
static V3 coat_passage(const Mat &m, float c) {
  if (c <= 0.0f) return V3(1.0f);
  float rc = std::sqrt(mxf(0.0f, 1.0f - (1.0f - c * c) / (m.coatIor * m.coatIor)));
  V3 passage = vpow(vmaxs(m.coatColor, 0.0f), V3(0.5f / mxf(rc, 1e-5f)));
  return mix3(V3(1.0f), passage, m.coatWeight);
}
static float coat_energy(const Mat &m, float c) {
  if (m.coatRoughness <= ROUGHNESS_MIN) return pbr_dielectric_fresnel(c, 1.0f / m.coatIor);
  return clampf(1.0f - opaque_energy_complement(m.coatIor, m.coatRoughness * m.coatRoughness, c), 0.0f, 1.0f);
}
static V3 coat_darkening(const Prepared &P) {
  const Mat &m = P.material;
  float diel = 1.0f - m.baseMetalness;
  float opaque = diel * (1.0f - m.transmissionWeight) * (1.0f - m.subsurfaceWeight);
  float sss = diel * (1.0f - m.transmissionWeight) * m.subsurfaceWeight;
  float trans = diel * m.transmissionWeight;
  float oSpec = average_fresnel(P.specularEta);
  float tSpec = average_fresnel(1.0f / P.transmissionReflectionIor);
  float specularity = clampf(opaque * oSpec + (sss + trans) * tSpec + m.baseMetalness, 0.0f, 1.0f);
  float effR = mixf(1.0f, m.specularRoughness, specularity);
  float ks = average_fresnel(m.coatIor);
  float kr = 1.0f - (1.0f - ks) / (m.coatIor * m.coatIor);
  float k = mixf(ks, kr, effR);
  V3 eb = m.baseMetalness * m.specularWeight * metal_average_fresnel(m.baseWeight * m.baseColor, m.specularColor) +
          opaque * mix3(m.baseWeight * m.baseColor, V3(1.0f), oSpec) + sss * m.subsurfaceColor + V3(trans);
  V3 delta = V3(1.0f - k) / vmax(V3(1.0f) - eb * k, V3(1e-6f));
  return mix3(V3(1.0f), delta, m.coatWeight * m.coatDarkening);
}
static V3 coat_base_scale(const Prepared &P, float lightCos) {
  if (P.material.coatWeight <= 0.0f) return V3(1.0f);
  if (P.material.inside) return coat_passage(P.material, -lightCos);
  return P.coatBaseScale * coat_passage(P.material, lightCos);
}

// openpbr_prepare (openPbrBsdf.wgsl.ts:853-1078) without fuzz, subsurface,
// thin walls, thin film and anisotropy frames.
// openpbr_fuzz_directional_albedo (openPbrBsdf.wgsl.ts:595-607).
static float fuzz_directional_albedo(float viewCosine, float roughness) {
  static const float sqrtTwoPi = (float)std::sqrt(2.0 * 3.14159265359);
  float x = clampf(viewCosine, 0.0f, 1.0f);
  float y = clampf(roughness, 0.01f, 1.0f);
  float scale = y * (0.0206607f + 1.58491f * y) / (0.0379424f + y * (1.32227f + y));
  float mean = y * (-0.193854f + y * (-1.14885f + y * (1.7932f - 0.95943f * y * y))) / (0.046391f + y);
  float offset = y * (0.000654023f + (-0.0207818f + 0.119681f * y) * y) / (1.26264f + y * (-1.92021f + y));
  float z = (x - mean) / scale;
  float gaussian = mtxfm::expf(-0.5f * (z * z)) / (scale * sqrtTwoPi);  // pow(z, 2) correctly rounded
  return mxf(gaussian + offset, 0.0f);
}

float fuzz_albedo(float viewCosine, float roughness) { return fuzz_directional_albedo(viewCosine, roughness); }

// openpbr_eval_fuzz (openPbrBsdf.wgsl.ts:609-625).
static V3 eval_fuzz(const Mat &m, V3 v, V3 l, float &pdf) {
  pdf = 0.0f;
  if ((v.z <= 0.0f) | (l.z <= 0.0f)) return V3(0.0f);
  float albedo = fuzz_directional_albedo(v.z, m.fuzzRoughness);
  pdf = l.z * INVPI;
  return m.fuzzColor * albedo * INVPI;
}

// openpbr_thin_wall_fresnel, openpbr_thin_wall_reflection_coefficient,
// openpbr_final_transmission_tint (openPbrBsdf.wgsl.ts:697-731).
static float thin_wall_fresnel(float f) { return 2.0f * f / (1.0f + f); }
static V3 thin_wall_reflection_coefficient(V3 f, float transmissionWeight) {
  V3 window = 2.0f * f / (V3(1.0f) + f);
  return mix3(f, window, clampf(transmissionWeight, 0.0f, 1.0f));
}
static V3 final_transmission_tint(const Mat &m, float incidentCosine, float etaTOverEtaI) {
  V3 tint = m.transmissionColor;
  if (!m.thinWalled) return tint;
  float incidentSineSquared = mxf(0.0f, 1.0f - incidentCosine * incidentCosine);
  float transmittedSineSquared = incidentSineSquared / mxf(etaTOverEtaI * etaTOverEtaI, 1e-9f);
  if (transmittedSineSquared >= 1.0f) return V3(0.0f);
  float transmittedCosine = std::sqrt(1.0f - transmittedSineSquared);
  float e = 1.0f / mxf(transmittedCosine, 1e-5f);
  V3 b = vmaxs(tint, 0.0f);
  return V3(mtxfm::powf(b.x, e), mtxfm::powf(b.y, e), mtxfm::powf(b.z, e));
}
// openpbr_eval_subsurface_boundary (openPbrBsdf.wgsl.ts:733-750).
static V3 eval_subsurface_boundary(const Mat &m, V3 v, V3 l, float &pdf) {
  pdf = 0.0f;
  if ((v.z <= 0.0f) | (l.z >= 0.0f)) return V3(0.0f);
  float transmittedCosine = -l.z;
  float entryFresnel = dielectric_fresnel_s(v.z, m.relativeIor);
  pdf = transmittedCosine * INVPI;
  return V3(1.0f - entryFresnel) * INVPI;
}

static Prepared prepare(const Mat &m, V3 view, V3 normal) {
  Prepared P{};
  P.material = m;
  orthonormal_basis(normal, P.basis.t, P.basis.b);
  P.basis.n = normal;
  P.viewLocal = to_local(P.basis, view);
  P.diffuseColor = m.baseColor;
  float dielBase = 1.0f - m.baseMetalness;
  float opaqueDiel = dielBase * (1.0f - m.transmissionWeight);
  P.w[L_FUZZ] = m.fuzzWeight;
  P.w[L_COAT] = m.inside ? 0.0f : m.coatWeight;
  P.w[L_METAL] = m.baseMetalness;
  P.w[L_SPEC] = dielBase;
  P.w[L_TRANS] = dielBase * m.transmissionWeight;
  P.w[L_DIFF] = opaqueDiel * (1.0f - m.subsurfaceWeight) * m.baseWeight;
  P.w[L_SSS] = opaqueDiel * m.subsurfaceWeight;
  P.specularEta = apply_specular_weight_to_ior(1.0f / m.relativeIor, m.specularWeight);
  P.material.relativeIor = 1.0f / P.specularEta;
  P.refractionEta = P.specularEta;
  P.transmissionReflectionIor = P.material.relativeIor;
  if (m.coatWeight > 0.0f) {
    float wCoat = 1.0f;
    if (!m.inside) wCoat = mixf(1.0f, m.coatIor, m.coatWeight);
    float wBase = wCoat * apply_specular_weight_to_ior(m.specularIor / wCoat, m.specularWeight);
    float reflEta = wBase / wCoat;
    if ((wCoat > wBase) && (wBase >= 1.0f)) reflEta = 1.0f / reflEta;
    P.specularEta = reflEta;
    P.transmissionReflectionIor = 1.0f / reflEta;
    P.refractionEta = wBase;
    if (m.inside) {
      P.refractionEta = 1.0f / wBase;
      P.transmissionReflectionIor = wBase;
    }
    P.material.relativeIor = 1.0f / P.refractionEta;
  }
  P.specularAlpha = m.specularRoughness * m.specularRoughness;
  float vc = fabsf(P.viewLocal.z);
  float sif = pbr_dielectric_fresnel(vc, 1.0f / P.specularEta);
  V3 dielF = m.specularColor * (m.thinWalled ? mixf(sif, thin_wall_fresnel(sif), m.transmissionWeight) : sif);
  V3 metalF = m.specularWeight * metal_fresnel(vc, m.baseWeight * m.baseColor, m.specularColor);
  float transF = m.thinWalled ? thin_wall_fresnel(pbr_dielectric_fresnel(vc, P.transmissionReflectionIor))
                              : pbr_dielectric_fresnel(vc, P.transmissionReflectionIor);
  V3 transTint = final_transmission_tint(m, vc, P.refractionEta);
  float viewEC = opaque_energy_complement(P.specularEta, P.specularAlpha, vc);
  float avgEC = opaque_average_energy_complement(P.specularEta, P.specularAlpha);
  P.diffuseSpecularEnergyCompensation = viewEC / mxf(avgEC, 1e-12f);
  P.p[L_FUZZ] = P.w[L_FUZZ] * lum_pbr(m.fuzzColor) * mxf(fuzz_directional_albedo(vc, m.fuzzRoughness), 0.02f);
  P.p[L_COAT] = 0.0f;
  P.coatBaseScale = V3(1.0f);
  if ((m.coatWeight > 0.0f) && !m.inside) {
    float ce = coat_energy(m, vc);
    P.p[L_COAT] = P.w[L_COAT] * ce;
    P.coatBaseScale = coat_passage(m, vc) * (1.0f - P.w[L_COAT] * ce) * coat_darkening(P);
  }
  P.p[L_METAL] = P.w[L_METAL] * lum_pbr(metalF);
  P.metalMsScale = V3(0.0f);
  P.metalMsFactor = 0.0f;
  P.metalMsProbability = 0.0f;
  if ((m.specularRoughness >= 0.04f) && (P.w[L_METAL] > 0.0f)) {
    V3 avgF = metal_average_fresnel(m.baseWeight * m.baseColor, m.specularColor);
    P.metalMsScale = m.specularWeight * avgF * avgF;
    float viewLoss = metal_energy_complement(P.specularAlpha, vc);
    P.metalMsFactor = viewLoss / mxf(metal_average_energy_complement(P.specularAlpha), 1e-12f);
    float est = P.w[L_METAL] * lum_pbr(P.metalMsScale) * viewLoss;
    P.p[L_METAL] += est;
    P.metalMsProbability = est / mxf(P.p[L_METAL], 1e-12f);
  }
  P.p[L_SPEC] = P.w[L_SPEC] * lum_pbr(dielF);
  P.p[L_TRANS] = P.w[L_TRANS] * (1.0f - transF) * lum_pbr(transTint);
  P.dielMsReflScale = V3(0.0f);
  P.dielMsTransScale = V3(0.0f);
  P.dielMsViewLoss = 0.0f;
  P.dielMsRatio = 0.0f;
  P.dielMsProbability = 0.0f;
  if (!m.thinWalled & (m.transmissionWeight > 0.0f) & (P.specularAlpha >= 0.0016f) & (P.w[L_TRANS] > 0.0f) &
      (fabsf(P.refractionEta - 1.0f) > 1e-6f)) {
    float nonOpaque = P.w[L_TRANS] + P.w[L_SSS];
    P.dielMsReflScale = (m.inside ? V3(1.0f) : m.specularColor) * nonOpaque;
    P.dielMsTransScale = transTint * P.w[L_TRANS] + V3(P.w[L_SSS]);
    P.dielMsRatio = ideal_ratio(P.refractionEta, P.specularAlpha);
    P.dielMsViewLoss = ideal_loss(P.refractionEta, P.specularAlpha, vc);
    V3 avgScale = mix3(P.dielMsTransScale, P.dielMsReflScale, P.dielMsRatio);
    float est = lum_pbr(avgScale) * P.dielMsViewLoss;
    P.p[L_TRANS] += est;
    P.dielMsProbability = est / mxf(P.p[L_TRANS], 1e-12f);
  }
  P.p[L_DIFF] = P.w[L_DIFF] * lum_pbr(P.diffuseColor) * P.diffuseSpecularEnergyCompensation;
  P.p[L_SSS] = P.w[L_SSS] * (1.0f - transF) * lum_pbr(m.subsurfaceColor) * P.diffuseSpecularEnergyCompensation;
  if ((m.coatWeight > 0.0f) && !m.inside) {
    float bs = maxc(P.coatBaseScale);
    P.p[L_METAL] *= bs; P.p[L_SPEC] *= bs; P.p[L_TRANS] *= bs; P.p[L_DIFF] *= bs; P.p[L_SSS] *= bs;
  }
  float total = P.p[0];
  for (int l = 1; l < L_COUNT; l++) total += P.p[l];
  if (total <= 1e-8f) {
    for (int l = 0; l < L_COUNT; l++) P.p[l] = 0.0f;
  } else {
    float inv = 1.0f / total;
    for (int l = 0; l < L_COUNT; l++) P.p[l] *= inv;
  }
  return P;
}

static float substrate_fresnel(const Prepared &P, float c, bool reflection) {
  if (P.material.coatWeight <= 0.0f) return pbr_dielectric_fresnel(c, P.material.relativeIor);
  float tF = pbr_dielectric_fresnel(c, P.transmissionReflectionIor);
  if (!reflection) return tF;
  float oF = pbr_dielectric_fresnel(c, 1.0f / P.specularEta);
  float nonOpaque = P.material.transmissionWeight + (1.0f - P.material.transmissionWeight) * P.material.subsurfaceWeight;
  return mixf(oF, tF, nonOpaque);
}

// openpbr_eval_prepared_lobe (openPbrBsdf.wgsl.ts:1092-1421), allowedLobe = all.
static V3 eval_prepared(const Prepared &P, V3 lightDir, float &pdf) {
  const Mat &m = P.material;
  V3 v = P.viewLocal, l = to_local(P.basis, lightDir);
  bool reflection = l.z * v.z > 0.0f;
  V3 h = reflection ? normalize(l + v) : normalize(l + v * m.relativeIor);
  if (h.z < 0.0f) h = -h;
  V3 response(0.0f);
  pdf = 0.0f;
  float lobePdf = 0.0f;
  float viewHalf = fabsf(dot(v, h));
  V3 thinWallHalf = h;
  float dielectricViewHalf = viewHalf;
  if (!reflection & m.thinWalled) {
    thinWallHalf = normalize(V3(l.x, l.y, -l.z) + v);
    if (thinWallHalf.z < 0.0f) thinWallHalf = -thinWallHalf;
    dielectricViewHalf = fabsf(dot(v, thinWallHalf));
  }
  bool evalDiel = (reflection & (P.w[L_SPEC] > 0.0f)) | (!reflection & (P.w[L_TRANS] > 0.0f));
  bool evalMetal = reflection & (P.w[L_METAL] > 0.0f);
  V3 fr0(0.0f), fr1(0.0f);
  if (evalDiel) fr0 = V3(substrate_fresnel(P, dielectricViewHalf, reflection));
  if (evalMetal) fr1 = metal_fresnel(viewHalf, m.baseWeight * m.baseColor, m.specularColor);
  float comp = P.diffuseSpecularEnergyCompensation * opaque_energy_complement(P.specularEta, P.specularAlpha, fabsf(l.z));
  if (reflection) {
    if (P.w[L_DIFF] > 0.0f) {
      response = response + P.w[L_DIFF] * eval_diffuse(m, P.diffuseColor, v, l, comp, lobePdf);
      pdf += P.p[L_DIFF] * lobePdf;
    }
    if ((P.w[L_SSS] > 0.0f) & m.thinWalled) {
      response = response + 0.5f * P.w[L_SSS] * eval_diffuse(m, m.subsurfaceColor, v, l, comp, lobePdf);
      pdf += P.p[L_SSS] * 0.5f * lobePdf;
    }
    if (P.w[L_SPEC] > 0.0f) {
      V3 dielectricFresnel = fr0;
      if (m.thinWalled) dielectricFresnel = thin_wall_reflection_coefficient(dielectricFresnel, m.transmissionWeight);
      V3 F = m.specularColor * dielectricFresnel;
      response = response + P.w[L_SPEC] * eval_refl(m.alphaX, m.alphaY, v, l, h, F, lobePdf);
      pdf += P.p[L_SPEC] * lobePdf;
    }
    if (P.w[L_METAL] > 0.0f) {
      V3 F = m.specularWeight * fr1;
      response = response + P.w[L_METAL] * eval_refl(m.alphaX, m.alphaY, v, l, h, F, lobePdf);
      if (P.metalMsFactor > 0.0f) {
        float factor = mnf(P.metalMsFactor * metal_energy_complement(P.specularAlpha, l.z), 1.0f / mxf(l.z, 1e-8f));
        response = response + P.w[L_METAL] * P.metalMsScale * factor * INVPI;
        lobePdf = mixf(lobePdf, 0.5f * INVPI, P.metalMsProbability);
      }
      pdf += P.p[L_METAL] * lobePdf;
    }
  } else {
    if (P.w[L_TRANS] > 0.0f) {
      if (m.thinWalled) {
        V3 mirroredLight(l.x, l.y, -l.z);
        V3 windowFresnel = thin_wall_reflection_coefficient(fr0, 1.0f);
        V3 tint = final_transmission_tint(m, fabsf(v.z), P.refractionEta);
        response = response + P.w[L_TRANS] * eval_refl(m.alphaX, m.alphaY, v, mirroredLight, thinWallHalf,
                                                       tint * (V3(1.0f) - windowFresnel), lobePdf);
      } else {
        response = response + P.w[L_TRANS] * eval_refr(m.alphaX, m.alphaY, m.relativeIor, m.transmissionColor, v, l,
                                                       h, fr0, lobePdf);
      }
      if (P.dielMsProbability > 0.0f) lobePdf *= 1.0f - P.dielMsProbability;
      pdf += P.p[L_TRANS] * lobePdf;
    }
    if (P.w[L_SSS] > 0.0f) {
      if (m.thinWalled) {
        V3 flippedLight(l.x, l.y, -l.z);
        response = response + 0.5f * P.w[L_SSS] * eval_diffuse(m, m.subsurfaceColor, v, flippedLight, comp, lobePdf);
        lobePdf *= 0.5f;
      } else {
        response = response + P.w[L_SSS] * eval_subsurface_boundary(m, v, l, lobePdf);
      }
      pdf += P.p[L_SSS] * lobePdf;
    }
  }
  if (P.dielMsProbability > 0.0f) {
    float c = fabsf(l.z);
    if (c > 0.0f) {
      float ior = reflection ? P.refractionEta : 1.0f / P.refractionEta;
      float ratio = reflection ? P.dielMsRatio : 1.0f - P.dielMsRatio;
      float factors = P.dielMsViewLoss * ideal_loss(ior, P.specularAlpha, c) / mxf(ideal_avg_loss(ior, P.specularAlpha), 1e-12f);
      float factor = mnf(factors * ratio, 1.0f / c);
      V3 scale = reflection ? P.dielMsReflScale : P.dielMsTransScale;
      response = response + scale * factor * INVPI;
      pdf += P.p[L_TRANS] * P.dielMsProbability * ratio * 0.5f * INVPI;
    }
  }
  if (m.coatWeight > 0.0f) response = response * coat_base_scale(P, l.z);
  if (reflection & (P.w[L_COAT] > 0.0f)) {
    V3 coatF(pbr_dielectric_fresnel(viewHalf, 1.0f / m.coatIor));
    float cax, cay;
    roughness_to_alpha(m.coatRoughness, 0.0f, cax, cay);
    response = response + P.w[L_COAT] * eval_refl(cax, cay, v, l, h, coatF, lobePdf);
    pdf += P.p[L_COAT] * lobePdf;
  }
  if (reflection & (P.w[L_FUZZ] > 0.0f)) {
    float fuzzAlbedo = fuzz_directional_albedo(fabsf(v.z), m.fuzzRoughness);
    response = response * vmaxs(V3(1.0f - P.w[L_FUZZ] * fuzzAlbedo), 0.0f);
    response = response + P.w[L_FUZZ] * eval_fuzz(m, v, l, lobePdf);
    pdf += P.p[L_FUZZ] * lobePdf;
  }
  return response * fabsf(l.z);
}

// openpbr_sample_direction (openPbrBsdf.wgsl.ts:1456-1656).
static bool sample_direction(const Prepared &P, float rx, float ry, float lobe, V3 &lightDir, bool &transmission) {
  const Mat &m = P.material;
  V3 v = P.viewLocal;
  transmission = false;
  lightDir = P.basis.n;
  float cdf[L_COUNT];
  cdf[0] = P.p[0];
  for (int i = 1; i < L_COUNT; i++) cdf[i] = cdf[i - 1] + P.p[i];
  if (cdf[L_SSS] <= 0.0f) return false;
  V3 l;
  if ((lobe < cdf[L_FUZZ]) | ((lobe >= cdf[L_TRANS]) & (lobe < cdf[L_DIFF]))) {
    l = cosine_hemisphere(rx, ry);
  } else if (lobe < cdf[L_COAT]) {
    float cax, cay;
    roughness_to_alpha(m.coatRoughness, 0.0f, cax, cay);
    V3 h = ggx_visible_normal(v, cax, cay, rx, ry);
    if (h.z < 0.0f) h = -h;
    l = normalize(reflect(-v, h));
  } else if (lobe < cdf[L_SPEC]) {
    bool isMetal = lobe < cdf[L_METAL];
    float metalDraw = (lobe - cdf[L_COAT]) / mxf(P.p[L_METAL], 1e-12f);
    if (isMetal & (metalDraw < P.metalMsProbability)) {
      float z = rx;
      float radial = std::sqrt(mxf(0.0f, 1.0f - z * z));
      float az = TWOPI * ry;
      l = V3(radial * mtxfm::cosf(az), radial * mtxfm::sinf(az), z);
    } else {
      V3 h = ggx_visible_normal(v, m.alphaX, m.alphaY, rx, ry);
      if (h.z < 0.0f) h = -h;
      l = normalize(reflect(-v, h));
    }
  } else if (lobe < cdf[L_TRANS]) {
    transmission = true;
    float msDraw = (lobe - cdf[L_SPEC]) / mxf(P.p[L_TRANS], 1e-12f);
    if ((P.dielMsProbability > 0.0f) & (msDraw < P.dielMsProbability)) {
      float ratio = P.dielMsRatio;
      bool reflectMs = rx < ratio;
      float z = reflectMs ? mnf(rx / ratio, 0.99999994f) : -mnf((rx - ratio) / (1.0f - ratio), 0.99999994f);
      float radial = std::sqrt(mxf(0.0f, 1.0f - z * z));
      float az = TWOPI * ry;
      l = V3(radial * mtxfm::cosf(az), radial * mtxfm::sinf(az), z);
      transmission = !reflectMs;
    } else {
      V3 h = ggx_visible_normal(v, m.alphaX, m.alphaY, rx, ry);
      if (h.z < 0.0f) h = -h;
      if (m.thinWalled) {
        V3 reflected = normalize(reflect(-v, h));
        l = V3(reflected.x, reflected.y, -reflected.z);
      } else {
        l = refract(-v, h, m.relativeIor);
        if (dot(l, l) <= 1e-8f) { transmission = false; return false; }
        l = normalize(l);
      }
    }
  } else {
    transmission = true;
    if (m.thinWalled) {
      bool transmit = rx >= 0.5f;
      if (transmit) rx = (rx - 0.5f) * 2.0f;
      else rx *= 2.0f;
      l = cosine_hemisphere(rx, ry);
      if (transmit) l.z = -l.z;
      else transmission = false;
    } else {
      l = cosine_hemisphere(rx, ry);
      l.z = -l.z;
    }
  }
  lightDir = to_world(P.basis, l);
  if ((l.z == 0.0f) | ((l.z < 0.0f) != transmission)) return false;
  return true;
}

// ------------------------------------------------------------------ lights
struct LightSample {
  V3 normal, emission, direction;
  float dist, pdf, selectionPmf;
  bool isDelta;
  uint32_t sourceType;
};

// lightSampling.ts:172-212 (Walker alias table) over {env, analytic lights}.
void build_tables(const RenderScene &s, HostTables &T) {
  std::vector<int32_t> types, idxs;
  std::vector<double> weights;
  auto push = [&](int32_t ty, int32_t ix, double w) {
    if (!(w > 0) || !std::isfinite(w)) return;
    types.push_back(ty); idxs.push_back(ix); weights.push_back(w);
  };
  double envPower = 0.0;
  if (s.env) {  // lightSampling.ts:415-431
    double acc = 0.0;
    for (uint32_t y = 0; y < s.envH; y++) {
      double sinT = std::sin(3.141592653589793 * (y + 0.5) / s.envH);
      for (uint32_t x = 0; x < s.envW; x++) {
        const float *p = s.env + ((size_t)y * s.envW + x) * 4;
        double l = (double)p[0] * 0.2126 + (double)p[1] * 0.7152 + (double)p[2] * 0.0722;
        acc += (l > 0 ? l : 0) * sinT;
      }
    }
    envPower = acc * (2 * 3.141592653589793 * 3.141592653589793) / ((double)s.envW * s.envH);
  }
  push(0, 0, envPower);
  for (uint32_t i = 0; i < s.lightCount; i++) {  // estimateAnalyticLightPower, lightSampling.ts:72-107
    const float *L = s.lights + (size_t)i * 16;
    double lum = (double)L[L_EMISSION] * 0.2126 + (double)L[L_EMISSION + 1] * 0.7152 + (double)L[L_EMISSION + 2] * 0.0722;
    lum = lum > 0 ? lum : 0;
    const double pi = 3.141592653589793, area = (double)L[L_AREA], radius = (double)L[L_RADIUS];
    int type = (int)L[L_TYPE];
    if (type == 0) type = LT_RECT;
    double w = 0.0;
    if (lum > 0) {
      if (type == LT_RECT || type == LT_DISK || type == LT_CYLINDER) {
        w = lum * pi * std::max(area, 0.0);
      } else if (type == LT_SPHERE) {
        w = radius == 0.0 ? lum * 4 * pi : lum * pi * std::max(area, 0.0);
      } else if (type == LT_SPOT) {
        // getSpotLightShapingFluxFactor (lightsData.ts:150-154): v = (cos outer, cos inner)
        double flux = 0.5 * (1 - 0.5 * ((double)L[L_V + 1] + (double)L[L_V]));
        w = radius == 0.0 ? lum * 4 * pi * flux : lum * pi * area * flux;
      } else {
        w = lum * 4 * pi;  // directional
      }
    }
    push(1, (int32_t)i, w);
  }
  for (uint32_t i = 0; i < s.emissiveCount; i++) push(2, (int32_t)i, s.emissiveWeight[i]);  // sceneAssembly.ts:47
  size_t n = weights.size();
  T.lightPmf.assign(s.lightCount, 0.0f);
  T.emissivePmf.assign(s.emissiveCount, 0.0f);
  double total = 0.0;
  for (double w : weights) total += w;
  T.totalWeight = n && total > 0 ? total : 0.0;
  T.prob.assign(std::max<size_t>(n, 1), 0.0f);
  T.alias.assign(std::max<size_t>(n, 1), 0);
  T.type.assign(std::max<size_t>(n, 1), 0);
  T.index.assign(std::max<size_t>(n, 1), 0);
  if (!n || !(total > 0)) { T.prob.clear(); return; }
  std::vector<float> pmf(n);
  std::vector<double> scaled(n);
  std::vector<uint32_t> small, large;
  for (size_t i = 0; i < n; i++) {
    pmf[i] = (float)(weights[i] / total);
    scaled[i] = (double)pmf[i] * (double)n;
    if (scaled[i] < 1) small.push_back((uint32_t)i); else large.push_back((uint32_t)i);
  }
  while (!small.empty() && !large.empty()) {
    uint32_t low = small.back(); small.pop_back();
    uint32_t high = large.back(); large.pop_back();
    T.prob[low] = (float)scaled[low];
    T.alias[low] = (int32_t)high;
    scaled[high] = scaled[high] + scaled[low] - 1;
    if (scaled[high] < 1) small.push_back(high); else large.push_back(high);
  }
  for (uint32_t i : small) { T.prob[i] = 1.0f; T.alias[i] = (int32_t)i; }
  for (uint32_t i : large) { T.prob[i] = 1.0f; T.alias[i] = (int32_t)i; }
  for (size_t i = 0; i < n; i++) {
    T.type[i] = types[i];
    T.index[i] = idxs[i];
    if (types[i] == 0) T.envPmf = pmf[i];
    else if (types[i] == 1) T.lightPmf[idxs[i]] = pmf[i];
    else T.emissivePmf[idxs[i]] = pmf[i];
  }
  // Small/large lists pop from the back in the reference (small[--smallCount]).
}

// envMapDistribution (envMapDistribution.ts:57-134).
void build_env_cdf(const RenderScene &s, HostTables &T) {
  if (!s.env) return;
  const uint32_t W = s.envW, H = s.envH, CW = W + 2, CH = H + 1;
  T.cdfW = CW; T.cdfH = CH;
  T.cdf.assign((size_t)CW * CH * 2, 0.0f);
  auto at = [&](uint32_t x, uint32_t y, int c) -> float & { return T.cdf[((size_t)y * CW + x) * 2 + c]; };
  std::vector<float> lum((size_t)W * H), horiz((size_t)W * H);
  for (size_t i = 0; i < (size_t)W * H; i++) {
    const float *p = s.env + i * 4;
    lum[i] = (float)(1.0 * (0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2]));
  }
  for (uint32_t y = 0; y < H; y++)
    for (uint32_t x = 0; x < W; x++) {
      size_t row = (size_t)y * W;
      double a = lum[row + (x > 0 ? x - 1 : 0)], b = lum[row + x], c = lum[row + std::min(W - 1, x + 1)];
      horiz[row + x] = (float)std::max(std::max(0.0, a), std::max(b, c));
    }
  for (uint32_t y = 0; y < H; y++) {
    double sinT = std::sin(3.141592653589793 * (y + 0.5) / H);
    size_t above = (size_t)(y > 0 ? y - 1 : 0) * W, row = (size_t)y * W, below = (size_t)std::min(H - 1, y + 1) * W;
    for (uint32_t x = 0; x < W; x++) {
      double w = std::max(std::max((double)horiz[above + x], (double)horiz[row + x]),
                          std::max((double)horiz[below + x], 1e-12)) * sinT;
      at(x + 2, y, 0) = (float)((double)at(x + 1, y, 0) + w / W);
      at(x + 1, y, 1) = (float)w;
    }
    double rowIntegral = at(CW - 1, y, 0);
    for (uint32_t x = 1; x < CW; x++) {
      at(x, y, 0) = (float)((double)at(x, y, 0) / rowIntegral);
      at(x, y, 1) = (float)((double)at(x, y, 1) / rowIntegral);
    }
    at(0, y + 1, 0) = (float)((double)at(0, y, 0) + rowIntegral / H);
    at(0, y, 1) = (float)rowIntegral;
  }
  double integral = at(0, CH - 1, 0);
  for (uint32_t y = 0; y < CH; y++) {
    at(0, y, 0) = (float)((double)at(0, y, 0) / integral);
    at(0, y, 1) = (float)((double)at(0, y, 1) / integral);
  }
}

struct Ctx {
  const RenderScene &s;
  const HostTables &T;
  uint64_t rays = 0;
};

static V3 env_texel(const RenderScene &s, int x, int y) {
  const float *p = s.env + ((size_t)y * s.envW + x) * 4;
  return V3(p[0], p[1], p[2]);
}
// textureLinear (envMap.wgsl.ts:10-30), rgb only.
static V3 texture_linear(const RenderScene &s, float u, float v) {
  float sx = (float)s.envW, sy = (float)s.envH;
  float fx = u * sx - 0.5f, fy = v * sy - 0.5f;
  float flx = floorf(fx), fly = floorf(fy);
  float tx = fx - flx, ty = fy - fly;
  int ix = (int)flx, iy = (int)fly;
  int mx = (int)s.envW - 1, my = (int)s.envH - 1;
  auto cl = [](int a, int hi) { return a < 0 ? 0 : (a > hi ? hi : a); };
  V3 s1 = env_texel(s, cl(ix, mx), cl(iy, my)), s2 = env_texel(s, cl(ix + 1, mx), cl(iy, my));
  V3 s3 = env_texel(s, cl(ix, mx), cl(iy + 1, my)), s4 = env_texel(s, cl(ix + 1, mx), cl(iy + 1, my));
  return mix3(mix3(s1, s2, tx), mix3(s3, s4, tx), ty);
}
static void environment_uv(V3 d, float &u, float &v) {  // cartesianToEquirect, identity rotation
  float phi = mtxfm::atan2f(-d.z, -d.x);
  if (phi < 0.0f) phi += TWOPI;
  float theta = mtxfm::acosf(clampf(d.y, -1.0f, 1.0f));
  u = phi * 0.5f * INVPI;
  v = theta * INVPI;
}
static float cdfx(const HostTables &T, int x, int y, int c) {
  x = std::max(0, std::min(x, (int)T.cdfW - 1));
  y = std::max(0, std::min(y, (int)T.cdfH - 1));
  return T.cdf[((size_t)y * T.cdfW + x) * 2 + c];
}
static float env_solid_angle_pdf(float px, float py, float v) {
  if (!(v > 0.0f && v < 1.0f)) return 0.0f;
  float sinT = mtxfm::sinf(mnf(v, 1.0f - v) * PI);
  return px * py * INVPI2 / (2.0f * sinT);
}
static float env_map_pdf(const Ctx &c, float u, float v) {
  float sx = (float)c.s.envW, sy = (float)c.s.envH;
  int x = (int)(u * sx), y = (int)(v * sy);
  x = std::max(0, std::min(x, (int)c.s.envW - 1));
  y = std::max(0, std::min(y, (int)c.s.envH - 1));
  return env_solid_angle_pdf(cdfx(c.T, 1 + x, y, 1), cdfx(c.T, 0, y, 1), v);
}
static V3 sample_envmap(const Ctx &c, float r0, float r1, float &u, float &v, float &pdf) {
  int H = (int)c.s.envH, W = (int)c.s.envW;
  int left = 0, right = H + 1;
  while (left < right) {
    int mid = (left + right) >> 1;
    if (cdfx(c.T, 0, mid, 0) <= r0) left = mid + 1; else right = mid;
  }
  int vOff = left - 1;
  float a0 = cdfx(c.T, 0, vOff, 0), a1 = cdfx(c.T, 0, vOff + 1, 0);
  float py = cdfx(c.T, 0, vOff, 1);
  v = ((float)vOff + (r0 - a0) / (a1 - a0)) / (float)H;
  left = 0; right = W + 1;
  while (left < right) {
    int mid = (left + right) >> 1;
    if (cdfx(c.T, 1 + mid, vOff, 0) <= r1) left = mid + 1; else right = mid;
  }
  int uOff = left - 1;
  float b0 = cdfx(c.T, 1 + uOff, vOff, 0), b1 = cdfx(c.T, 1 + uOff + 1, vOff, 0);
  float px = cdfx(c.T, 1 + uOff, vOff, 1);
  u = ((float)uOff + (r1 - b0) / (b1 - b0)) / (float)W;
  float phi = u * TWOPI, theta = v * PI;
  float ct = mtxfm::cosf(theta), st = mtxfm::sinf(theta), cp = mtxfm::cosf(phi), sp = mtxfm::sinf(phi);
  pdf = env_solid_angle_pdf(px, py, v);
  return V3(-st * cp, ct, -st * sp);
}

static void light_sample_rect(const float *L, V3 p, float sx, float sy, LightSample &r) {
  V3 U(L[L_U], L[L_U + 1], L[L_U + 2]), Vv(L[L_V], L[L_V + 1], L[L_V + 2]);
  V3 pos(L[L_POS], L[L_POS + 1], L[L_POS + 2]);
  V3 n = normalize(cross(U, Vv));
  V3 lowerLeft = pos - 0.5f * (U + Vv);
  V3 surf = lowerLeft + U * sx + Vv * sy;
  V3 off = surf - p;
  float d2 = dot(off, off);
  r.dist = std::sqrt(d2);
  r.direction = off / r.dist;
  r.normal = n;
  float ec = dot(-r.direction, n);
  float spread = clampf(L[L_SPREAD], 0.0f, 1.0f);
  float sc = mtxfm::cosf(0.5f * PI * spread);
  if (ec <= 0.0f || ec < sc) { r.emission = V3(0.0f); r.pdf = 0.0f; return; }
  r.emission = V3(L[L_EMISSION], L[L_EMISSION + 1], L[L_EMISSION + 2]);
  r.pdf = d2 / mxf(L[L_AREA] * ec, 1e-8f);
}

static int light_type_of(const float *L) {
  int t = (int)L[L_TYPE];
  return t == 0 ? LT_RECT : t;
}
static const float INV4PI_F = (float)0.07957747154594766;

// light_sample_sphere (sampleLight.wgsl.ts:10-89).
static void light_sample_sphere(const float *L, V3 p, float rx, float ry, LightSample &r) {
  V3 pos(L[L_POS], L[L_POS + 1], L[L_POS + 2]), E(L[L_EMISSION], L[L_EMISSION + 1], L[L_EMISSION + 2]);
  V3 co = pos - p;
  float cd = length(co);
  if (L[L_RADIUS] <= 0.0f) {
    if (cd <= 1e-6f) {
      r.direction = V3(0.0f, 0.0f, 1.0f); r.normal = V3(0.0f, 0.0f, -1.0f); r.dist = 0.0f; r.pdf = 1.0f;
      r.emission = V3(0.0f); r.isDelta = true;
      return;
    }
    r.direction = co / cd; r.normal = -r.direction; r.dist = cd; r.pdf = 1.0f;
    r.emission = E / mxf(cd * cd, 1e-8f);
    r.isDelta = true;
    return;
  }
  float radius = mxf(L[L_RADIUS], 1e-6f);
  r.isDelta = false;
  V3 sd;
  if (cd <= radius) {
    float z = 1.0f - 2.0f * rx;
    float radial = std::sqrt(mxf(0.0f, 1.0f - z * z));
    float az = TWOPI * ry;
    sd = V3(radial * mtxfm::cosf(az), radial * mtxfm::sinf(az), z);
    r.pdf = INV4PI_F;
  } else {
    V3 axis = co / cd;
    float s2 = mnf(1.0f, radius * radius / (cd * cd));
    float cmax = std::sqrt(mxf(0.0f, 1.0f - s2));
    float omc = s2 / (1.0f + cmax);
    float omct = rx * omc;
    float ct = 1.0f - omct;
    float st = std::sqrt(mxf(0.0f, omct * (2.0f - omct)));
    float az = TWOPI * ry;
    V3 t, b;
    orthonormal_basis(axis, t, b);
    sd = axis * ct + t * (st * mtxfm::cosf(az)) + b * (st * mtxfm::sinf(az));
    r.pdf = 1.0f / mxf(TWOPI * omc, 1e-8f);
  }
  float cproj = dot(sd, co), c2 = dot(co, co);
  float disc = cproj * cproj - c2 + radius * radius;
  float nearD = cproj - std::sqrt(mxf(0.0f, disc)), farD = cproj + std::sqrt(mxf(0.0f, disc));
  r.dist = nearD > 1e-6f ? nearD : farD;
  r.direction = sd;
  r.normal = normalize((p + sd * r.dist) - pos);
  r.emission = E;
}

// evaluateSpotEmissionMultiplier (rayTraceCore.wgsl.ts:365-378), no IES profile.
static float spot_emission_multiplier(const float *L, V3 lightToSurface) {
  V3 axis = normalize(V3(L[L_U], L[L_U + 1], L[L_U + 2]));
  float cosTheta = dot(axis, lightToSurface);
  float lo = L[L_V], hi = L[L_V + 1];
  float cone = cosTheta >= lo ? 1.0f : 0.0f;
  if (hi > lo + 1e-6f) {  // WGSL smoothstep
    float t = clampf((cosTheta - lo) / (hi - lo), 0.0f, 1.0f);
    cone = t * t * (3.0f - 2.0f * t);
  }
  return cone;
}

// Area sample of a disk or cylinder point (sampleLight.wgsl.ts:163-195).
static void area_light_point(const float *L, V3 p, V3 surfacePoint, V3 normal, LightSample &r) {
  V3 offset = surfacePoint - p;
  float d2 = dot(offset, offset);
  r.dist = std::sqrt(d2);
  r.direction = offset / mxf(r.dist, 1e-8f);
  r.normal = normal;
  float emitterCosine = dot(-r.direction, normal);
  if (emitterCosine <= 0.0f) { r.emission = V3(0.0f); r.pdf = 0.0f; return; }
  r.emission = V3(L[L_EMISSION], L[L_EMISSION + 1], L[L_EMISSION + 2]);
  r.pdf = d2 / mxf(L[L_AREA] * emitterCosine, 1e-8f);
}

// sampleAnalyticLight (sampleLight.wgsl.ts:223-253).
static void light_sample_analytic(const float *L, V3 p, float sx, float sy, LightSample &r) {
  V3 U(L[L_U], L[L_U + 1], L[L_U + 2]), Vv(L[L_V], L[L_V + 1], L[L_V + 2]);
  V3 pos(L[L_POS], L[L_POS + 1], L[L_POS + 2]);
  switch (light_type_of(L)) {
    case LT_RECT: light_sample_rect(L, p, sx, sy, r); break;
    case LT_SPHERE: light_sample_sphere(L, p, sx, sy, r); break;
    case LT_SPOT: {
      light_sample_sphere(L, p, sx, sy, r);
      V3 surfacePoint = p + r.direction * r.dist;
      r.emission = r.emission * spot_emission_multiplier(L, normalize(p - surfacePoint));
      break;
    }
    case LT_DIRECTIONAL:
      r.direction = normalize(-U);
      r.normal = -r.direction;
      r.emission = V3(L[L_EMISSION], L[L_EMISSION + 1], L[L_EMISSION + 2]);
      r.dist = INF;
      r.pdf = 1.0f;
      r.isDelta = true;
      break;
    case LT_DISK: {
      float rr = std::sqrt(sx), phi = TWOPI * sy;
      V3 surfacePoint = pos + U * (rr * mtxfm::cosf(phi)) + Vv * (rr * mtxfm::sinf(phi));
      area_light_point(L, p, surfacePoint, normalize(cross(U, Vv)), r);
      break;
    }
    case LT_CYLINDER: {
      V3 axis = normalize(U), radial0 = normalize(Vv), radial1 = normalize(cross(axis, radial0));
      float phi = TWOPI * sy;
      V3 normal = radial0 * mtxfm::cosf(phi) + radial1 * mtxfm::sinf(phi);
      V3 surfacePoint = pos + U * (sx - 0.5f) + normal * L[L_RADIUS];
      area_light_point(L, p, surfacePoint, normal, r);
      break;
    }
    default: break;
  }
}

// collectEmissiveTriangles (lightSampling.ts:237-320): vec3.transformMat4 in
// f64 stored to f32, f32 edges and cross product, area = |cross| / 2 in f64
// (Math.hypot), degenerate triangles dropped; weight = area * luminance of
// the material's emission (prepareSceneLightSampling, sceneAssembly.ts:47).
uint32_t emissive_triangles(const float *positions, const uint32_t *vertexOffsets, const uint32_t *indices,
                            const uint32_t *indexOffsets, const uint32_t *triangleCounts, const uint32_t *instanceMesh,
                            const double *transforms, const uint32_t *instanceMaterial, uint32_t instanceCount,
                            const float *materials, float *out, double *weight, uint32_t capacity,
                            uint32_t *corners) {
  uint32_t n = 0;
  for (uint32_t i = 0; i < instanceCount; i++) {
    const float *em = materials + (size_t)instanceMaterial[i] * M_FLOATS + M_EMISSION;
    double lum = (double)em[0] * 0.2126 + (double)em[1] * 0.7152 + (double)em[2] * 0.0722;
    if (lum < 0) lum = 0;
    if (!(lum > 1e-8)) continue;
    const uint32_t mesh = instanceMesh[i];
    const double *m = transforms + (size_t)i * 16;
    for (uint32_t t = 0; t < triangleCounts[mesh]; t++) {
      float world[3][3];
      for (int corner = 0; corner < 3; corner++) {
        uint32_t v = vertexOffsets[mesh] + indices[indexOffsets[mesh] + t * 3 + corner];
        double x = positions[v * 3], y = positions[v * 3 + 1], z = positions[v * 3 + 2];
        double w = m[3] * x + m[7] * y + m[11] * z + m[15];
        if (w == 0.0) w = 1.0;  // `w || 1.0`
        world[corner][0] = (float)((m[0] * x + m[4] * y + m[8] * z + m[12]) / w);
        world[corner][1] = (float)((m[1] * x + m[5] * y + m[9] * z + m[13]) / w);
        world[corner][2] = (float)((m[2] * x + m[6] * y + m[10] * z + m[14]) / w);
      }
      float a[3], b[3], cr[3];
      for (int k = 0; k < 3; k++) {
        a[k] = (float)((double)world[1][k] - (double)world[0][k]);
        b[k] = (float)((double)world[2][k] - (double)world[0][k]);
      }
      cr[0] = (float)((double)a[1] * b[2] - (double)a[2] * b[1]);
      cr[1] = (float)((double)a[2] * b[0] - (double)a[0] * b[2]);
      cr[2] = (float)((double)a[0] * b[1] - (double)a[1] * b[0]);
      double area = 0.5 * std::hypot((double)cr[0], (double)cr[1], (double)cr[2]);
      if (!(area > 1e-10) || !std::isfinite(area)) continue;
      if (n < capacity) {
        float *o = out + (size_t)n * 12;
        for (int corner = 0; corner < 3; corner++)
          for (int k = 0; k < 3; k++) o[corner * 3 + k] = world[corner][k];
        o[9] = (float)area;
        o[10] = (float)instanceMaterial[i];
        o[11] = 0.0f;
        weight[n] = area * lum;
        if (corners)
          for (int corner = 0; corner < 3; corner++)
            corners[(size_t)n * 3 + corner] = vertexOffsets[mesh] + indices[indexOffsets[mesh] + t * 3 + corner];
      }
      n++;
    }
  }
  return n;
}

// sampleEmissiveTriangleLight (sampleLight.wgsl.ts:316-363).
static void sample_emissive_triangle(const Ctx &c, uint32_t index, V3 p, float rx, float ry, LightSample &r) {
  const float *t = c.s.emissive + (size_t)index * 12;
  V3 p0(t[0], t[1], t[2]), p1(t[3], t[4], t[5]), p2(t[6], t[7], t[8]);
  float area = t[9];
  uint32_t material = (uint32_t)t[10];
  float sqrtU = std::sqrt(rx);
  V3 b(1.0f - sqrtU, sqrtU * (1.0f - ry), sqrtU * ry);
  V3 lightPosition = b.x * p0 + b.y * p1 + b.z * p2;
  V3 normal = normalize(cross(p1 - p0, p2 - p0));
  r.direction = lightPosition - p;
  r.dist = length(r.direction);
  if (r.dist <= EPS) { r.emission = V3(0.0f); r.pdf = 0.0f; return; }
  r.direction = r.direction / r.dist;
  r.normal = normal;
  float cosLight = dot(normal, -r.direction);
  if (cosLight <= 0.0f) { r.emission = V3(0.0f); r.pdf = 0.0f; return; }
  const float *m = c.s.materials + (size_t)material * M_FLOATS;
  if (c.s.textures) {
    const float *u = c.s.textures->emissiveUv.data() + (size_t)index * 12;
    UV uv{b.x * u[0] + b.y * u[2] + b.z * u[4], b.x * u[1] + b.y * u[3] + b.z * u[5]};
    UV uv1{b.x * u[6] + b.y * u[8] + b.z * u[10], b.x * u[7] + b.y * u[9] + b.z * u[11]};
    r.emission = material_emission(c.s.textures, m, uv, uv1);
  } else {
    r.emission = V3(m[M_EMISSION], m[M_EMISSION + 1], m[M_EMISSION + 2]);
  }
  r.selectionPmf = c.T.emissivePmf[index];
  r.pdf = r.selectionPmf * (r.dist * r.dist) / mxf(area * cosLight, 1e-8f);
}

// sampleLightChoiceTable (sampleLight.wgsl.ts:365-420) with the candidate's
// stratified selection number (:298-308).
static void sample_light_candidate(const Ctx &c, V3 p, float rx, float ry, float rz, uint32_t cand,
                                   uint32_t count, LightSample &ls) {
  uint32_t cnt = std::max(count, 1u);
  rz = (rz + (float)std::min(cand, cnt - 1u)) / (float)cnt;
  ls.emission = V3(0.0f); ls.pdf = 0.0f; ls.selectionPmf = 0.0f; ls.dist = INF; ls.isDelta = false; ls.sourceType = 0;
  int choiceCount = (int)c.T.prob.size();
  if (choiceCount <= 0) return;
  float scaled = rz * (float)choiceCount;
  int column = std::min((int)scaled, choiceCount - 1);
  int sel = fractf(scaled) < c.T.prob[column] ? column : c.T.alias[column];
  if (c.T.type[sel] == 0 && c.s.env) {
    ls.sourceType = 0;
    float u, v, pdf;
    ls.direction = sample_envmap(c, rx, ry, u, v, pdf);
    ls.pdf = pdf;
    ls.emission = texture_linear(c.s, u, v);
    ls.dist = INF;
    ls.selectionPmf = c.T.envPmf;
    ls.pdf *= c.T.envPmf;
  } else if (c.T.type[sel] == 1) {
    ls.sourceType = 2;
    light_sample_analytic(c.s.lights + (size_t)c.T.index[sel] * 16, p, rx, ry, ls);
    ls.selectionPmf = c.T.lightPmf[c.T.index[sel]];
    ls.pdf *= ls.selectionPmf;
  } else if (c.T.type[sel] == 2) {
    ls.sourceType = 1;
    sample_emissive_triangle(c, (uint32_t)c.T.index[sel], p, rx, ry, ls);
  }
}

// intersectLight for a rect (intersect.wgsl.ts:976-1021).
static float intersect_rect_light(const float *L, V3 ro, V3 rd, LightSample &ls) {
  ls.direction = rd; ls.normal = -rd; ls.dist = INF; ls.pdf = 0.0f; ls.emission = V3(0.0f); ls.isDelta = false;
  V3 U(L[L_U], L[L_U + 1], L[L_U + 2]), Vv(L[L_V], L[L_V + 1], L[L_V + 2]);
  V3 pos = V3(L[L_POS], L[L_POS + 1], L[L_POS + 2]) - U * 0.5f - Vv * 0.5f;
  V3 n = normalize(cross(U, Vv));
  if (dot(n, rd) > 0.0f) return INF;
  float dt = dot(rd, n);
  float t = (dot(n, pos) - dot(n, ro)) / dt;
  float d = INF;
  if (t > EPS) {
    V3 p = ro + rd * t, vi = p - pos;
    float a1 = dot(U / dot(U, U), vi);
    if (a1 >= 0.0f && a1 <= 1.0f) {
      float a2 = dot(Vv / dot(Vv, Vv), vi);
      if (a2 >= 0.0f && a2 <= 1.0f) d = t;
    }
  }
  if (d < 0.0f) return INF;
  if (d == INF) return INF;
  float cosT = dot(-rd, n);
  if (cosT >= mtxfm::cosf(L[L_SPREAD] * PI / 2.0f)) {
    ls.direction = rd; ls.normal = n; ls.dist = d;
    ls.pdf = (d * d) / (L[L_AREA] * cosT);
    ls.emission = V3(L[L_EMISSION], L[L_EMISSION + 1], L[L_EMISSION + 2]);
    return d;
  }
  return INF;
}

// intersectSphere (intersect.wgsl.ts:924-945).
static float intersect_sphere(float rad, V3 pos, V3 ro, V3 rd) {
  V3 op = pos - ro;
  const float eps = 0.001f;
  float b = dot(op, rd);
  float det = b * b - dot(op, op) + rad * rad;
  if (det < 0.0f) return INF;
  det = std::sqrt(det);
  float t1 = b - det;
  if (t1 > eps) return t1;
  float t2 = b + det;
  if (t2 > eps) return t2;
  return INF;
}

// sphereSolidAnglePdf (intersect.wgsl.ts:966-974).
static float sphere_solid_angle_pdf(V3 origin, V3 center, float radius) {
  float centerDistance = length(origin - center);
  if (centerDistance <= radius) return INV4PI_F;
  float s2 = mnf(1.0f, radius * radius / (centerDistance * centerDistance));
  float cmax = std::sqrt(mxf(0.0f, 1.0f - s2));
  float omc = s2 / (1.0f + cmax);
  return 1.0f / mxf(TWOPI * omc, 1e-8f);
}

// intersectLight (intersect.wgsl.ts:976-1122) for every analytic type.
static float intersect_light(const float *L, V3 ro, V3 rd, LightSample &ls) {
  const int type = light_type_of(L);
  if (type == LT_RECT) return intersect_rect_light(L, ro, rd, ls);
  ls.direction = rd; ls.normal = -rd; ls.dist = INF; ls.pdf = 0.0f; ls.emission = V3(0.0f); ls.isDelta = false;
  V3 U(L[L_U], L[L_U + 1], L[L_U + 2]), Vv(L[L_V], L[L_V + 1], L[L_V + 2]);
  V3 pos(L[L_POS], L[L_POS + 1], L[L_POS + 2]), E(L[L_EMISSION], L[L_EMISSION + 1], L[L_EMISSION + 2]);
  const float radiusOrSpread = L[L_RADIUS];
  if (type == LT_DISK) {
    V3 normal = normalize(cross(U, Vv));
    if (dot(normal, rd) >= 0.0f) return INF;
    float d = dot(pos - ro, normal) / dot(normal, rd);
    V3 hit = ro + rd * d - pos;
    float radius = mxf(length(U), 1e-8f);
    if (d <= EPS || dot(hit, hit) > radius * radius) return INF;
    float cosine = dot(-rd, normal);
    ls.direction = rd; ls.normal = normal; ls.dist = d;
    ls.pdf = d * d / mxf(L[L_AREA] * cosine, 1e-8f);
    ls.emission = E;
    return d;
  }
  if (type == LT_CYLINDER) {
    V3 axis = normalize(U);
    float halfLength = 0.5f * length(U);
    V3 oc = ro - pos;
    V3 rdp = rd - axis * dot(rd, axis), od = oc - axis * dot(oc, axis);
    float a = dot(rdp, rdp), b = dot(od, rdp), cc = dot(od, od) - radiusOrSpread * radiusOrSpread;
    float discriminant = b * b - a * cc;
    if (a <= 1e-12f || discriminant < 0.0f) return INF;
    float root = std::sqrt(discriminant);
    float nearD = (-b - root) / a, farD = (-b + root) / a;
    float d = nearD > EPS ? nearD : farD;
    float axial = dot(oc + rd * d, axis);
    if (d <= EPS || fabsf(axial) > halfLength) return INF;
    V3 hit = ro + rd * d;
    V3 normal = normalize(hit - pos - axis * axial);
    float cosine = dot(-rd, normal);
    if (cosine <= 0.0f) return INF;
    ls.direction = rd; ls.normal = normal; ls.dist = d;
    ls.pdf = d * d / mxf(L[L_AREA] * cosine, 1e-8f);
    ls.emission = E;
    return d;
  }
  if (type == LT_SPHERE || type == LT_SPOT) {
    float d;
    if (radiusOrSpread <= 0.0f) {
      V3 toPoint = pos - ro;
      float projection = dot(toPoint, rd);
      V3 perpendicular = toPoint - projection * rd;
      if (projection <= EPS || dot(perpendicular, perpendicular) > 1e-8f) return INF;
      d = projection;
      ls.direction = rd; ls.normal = -rd; ls.dist = d; ls.pdf = 1.0f; ls.isDelta = true;
    } else {
      d = intersect_sphere(radiusOrSpread, pos, ro, rd);
      ls.isDelta = false;
      if (type == LT_SPHERE && d > 0.0f && d != INF) {
        ls.direction = rd; ls.dist = d;
        ls.normal = normalize(ro + rd * d - pos);
        ls.pdf = sphere_solid_angle_pdf(ro, pos, radiusOrSpread);
      }
    }
    if (type == LT_SPHERE) {
      if (d < 0.0f) return INF;
      ls.direction = rd; ls.dist = d;
      ls.emission = radiusOrSpread <= 0.0f ? E / mxf(d * d, 1e-8f) : E;
      return d;
    }
    if (d != INF) {
      V3 hitPt = ro + d * rd;
      float multiplier = spot_emission_multiplier(L, normalize(ro - hitPt));
      if (multiplier <= 0.0f) return INF;
      if (radiusOrSpread > 0.0f) {
        ls.direction = rd; ls.dist = d;
        ls.normal = normalize(ro + rd * d - pos);
        ls.pdf = sphere_solid_angle_pdf(ro, pos, radiusOrSpread);
      }
      ls.emission = E * multiplier * (radiusOrSpread <= 0.0f ? 1.0f / mxf(d * d, 1e-8f) : 1.0f);
    }
    return d;
  }
  return INF;  // directional lights are not intersected
}

// offsetSurfaceRay (common.wgsl.ts:14-19).
static V3 offset_surface_ray(V3 p, V3 gn, V3 dir) {
  V3 n = dot(gn, dir) >= 0.0f ? gn : -gn;
  V3 out;
  for (int k = 0; k < 3; k++) {
    int32_t io = (int32_t)(256.0f * n[k]);
    int32_t bits = f2i(p[k]);
    float shifted = i2f(bits + (p[k] < 0.0f ? -io : io));
    out[k] = fabsf(p[k]) < (float)(1.0 / 32.0) ? p[k] + n[k] / 65536.0f : shifted;
  }
  return out;
}

// ------------------------------------------------------------------ integrator
struct Surface {
  bool hit;
  V3 position, normal, faceNormal, shadowOffset;
  int32_t triangleIndex, instanceIndex, materialIndex;
  bool albedoSampled;
  UV uv, uv1;  // times the material's UV scale (textured scenes)
};

// getMatNormal (materialBuffer.wgsl.ts:377-415), object space: the normal
// map's tangent-space normal in the cotangent frame of UV0
// (thetenthplanet.de/archives/1180), a fixed frame where UV0 has no gradient.
static V3 mat_normal(const TextureSet &T, uint32_t index, const float *m, UV uv, V3 normal, V3 dp1, V3 dp2, UV duv1,
                     UV duv2) {
  V3 dp2perp = cross(dp2, normal), dp1perp = cross(normal, dp1);
  V3 dpdu = dp2perp * duv1.x + dp1perp * duv2.x;
  V3 dpdv = dp2perp * duv1.y + dp1perp * duv2.y;
  float gradientLengthSquared = mxf(dot(dpdu, dpdu), dot(dpdv, dpdv));
  if (gradientLengthSquared > 1e-12f) {
    float invmax = 1.0f / std::sqrt(gradientLengthSquared);
    dpdu = dpdu * invmax;
    dpdv = dpdv * invmax;
  } else {
    V3 referenceAxis = fabsf(normal.z) < 0.999f ? V3(0.0f, 0.0f, 1.0f) : V3(0.0f, 1.0f, 0.0f);
    dpdu = normalize(cross(referenceAxis, normal));
    dpdv = cross(normal, dpdu);
  }
  float c[4];
  sample_map(T, index, uv, c);
  V3 n(2.0f * c[0] - 1.0f, 2.0f * c[1] - 1.0f, 2.0f * c[2] - 1.0f);
  n = normalize(V3(n.x * m[M_NORMAL_SCALE], n.y * m[M_NORMAL_SCALE + 1], n.z));
  return normalize(dpdu * n.x + dpdv * n.y + normal * n.z);
}
struct PathState {
  V3 ro, rd, li, beta;
  float alpha, scatteringPdf;
  bool fromEye, abort, scatteringIsDelta, isNextRefraction;
};
struct Aux { V3 albedo, normal; bool isBlank; };

static inline V3 instance_color(const TraceScene &t, int32_t ins) {
  const uint32_t *r = &t.instances[(size_t)ins * 40];
  return V3(u2f(r[36]), u2f(r[37]), u2f(r[38]));
}

// intersectScene + completeSurfaceInteraction/fetchAttributes (intersect.wgsl.ts:1233-1601).
static void intersect_scene(Ctx &c, const PathState &path, Surface &si, int bounces, bool doubleSided) {
  c.rays++;
  ClosestHit h = trace_closest(*c.s.trace, path.ro, path.rd, RAY_MAX_DISTANCE, doubleSided, bounces, si.triangleIndex,
                               si.instanceIndex, nullptr);
  si.hit = false;
  if (!h.hit) return;
  const TraceScene &t = *c.s.trace;
  const uint32_t *ins = &t.instances[(size_t)h.instance * 40];
  int32_t vo = (int32_t)ins[1];
  size_t r3 = (size_t)h.record * 4;
  int32_t i0 = (int32_t)t.blas[r3] + vo, i1 = (int32_t)t.blas[r3 + 1] + vo, i2 = (int32_t)t.blas[r3 + 2] + vo;
  auto P = [&](int32_t i) { return V3(t.positions[(size_t)i * 3], t.positions[(size_t)i * 3 + 1], t.positions[(size_t)i * 3 + 2]); };
  auto N = [&](int32_t i) { return V3(c.s.normals[(size_t)i * 3], c.s.normals[(size_t)i * 3 + 1], c.s.normals[(size_t)i * 3 + 2]); };
  V3 p0 = P(i0), p1 = P(i1), p2 = P(i2), n0 = N(i0), n1 = N(i1), n2 = N(i2);
  V3 b = h.bary;
  si.hit = true;
  si.triangleIndex = h.record;
  si.instanceIndex = h.instance;
  si.materialIndex = ((int32_t)ins[2]) >> 8;
  V3 position = h.position;
  V3 normal = normalize(b.x * n0 + b.y * n1 + b.z * n2);
  // Shadow-terminator offset (intersect.wgsl.ts:1376-1407), object space.
  V3 geometric = normalize(cross(p1 - p0, p2 - p0));
  V3 smooth = b.x * n0 + b.y * n1 + b.z * n2;
  float A = dot(n2 - n0, p0 - p2), Bv = dot(n2 - n1, p1 - p2), Cv = dot(n1 - n0, p1 - p0);
  float hh = A * b.x * (b.x - 1.0f) + (A + Bv + Cv) * b.x * b.y + Bv * b.y * (b.y - 1.0f);
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
  V3 shadowOffset = smooth * hh;
  si.uv = UV{0.0f, 0.0f};
  si.uv1 = UV{0.0f, 0.0f};
  V3 shading = normal;
  if (const TextureSet *T = c.s.textures) {
    auto at = [&](const std::vector<float> &a, int32_t i) { return UV{a[(size_t)i * 2], a[(size_t)i * 2 + 1]}; };
    UV t0 = at(T->uvs, i0), t1 = at(T->uvs, i1), t2 = at(T->uvs, i2);
    UV s0 = at(T->uvs1, i0), s1 = at(T->uvs1, i1), s2 = at(T->uvs1, i2);
    UV uv{b.x * t0.x + b.y * t1.x + b.z * t2.x, b.x * t0.y + b.y * t1.y + b.z * t2.y};
    UV uv1{b.x * s0.x + b.y * s1.x + b.z * s2.x, b.x * s0.y + b.y * s1.y + b.z * s2.y};
    const float *m = c.s.materials + (size_t)si.materialIndex * M_FLOATS;
    si.uv = UV{uv.x * m[M_UV_SCALE], uv.y * m[M_UV_SCALE + 1]};
    si.uv1 = UV{uv1.x * m[M_UV_SCALE], uv1.y * m[M_UV_SCALE + 1]};
    const uint32_t slot = (uint32_t)m[M_MAPS + MAP_NORMAL];
    if (slot)
      shading = mat_normal(*T, slot - 1u, m, si.uv, normal, p0 - p2, p1 - p2, UV{t0.x - t2.x, t0.y - t2.y},
                           UV{t1.x - t2.x, t1.y - t2.y});
  }
  float l2w[16], w2l[16];
  for (int k = 0; k < 16; k++) { w2l[k] = u2f(ins[4 + k]); l2w[k] = u2f(ins[20 + k]); }
  auto nt = [&](V3 n) {  // transpose(mat3(w2l)) * n
    return V3(dot(V3(w2l[0], w2l[1], w2l[2]), n), dot(V3(w2l[4], w2l[5], w2l[6]), n), dot(V3(w2l[8], w2l[9], w2l[10]), n));
  };
  si.faceNormal = normalize(nt(h.normal));
  si.normal = normalize(nt(shading));
  si.shadowOffset = mat_mul_point(l2w, shadowOffset, 0.0f);
  si.position = mat_mul_point(l2w, position, 1.0f);
}

// intersectionAlphaPasses (intersectionAlpha.wgsl.ts:28-88) for the shadow
// query: alpha = opacity, times the base map's a when map alpha is coverage
// (:43-61) and the alpha map's g (:63-78), each sampled with sampleMap at the
// UV0 barycentric blend times the material's UV scale; then the alpha mode
// (:80-87) and `alpha >= 1`.
static bool alpha_covers(const void *ctx, int32_t instance, int32_t record, V3 b) {
  const Ctx &c = *(const Ctx *)ctx;
  const TraceScene &t = *c.s.trace;
  const uint32_t *ins = &t.instances[(size_t)instance * 40];
  const float *m = c.s.materials + (size_t)(((int32_t)ins[2]) >> 8) * M_FLOATS;
  float alpha = m[M_OPACITY];
  if (const TextureSet *T = c.s.textures) {
    const int32_t vo = (int32_t)ins[1];
    const size_t r3 = (size_t)record * 4;
    const int32_t i0 = (int32_t)t.blas[r3] + vo, i1 = (int32_t)t.blas[r3 + 1] + vo, i2 = (int32_t)t.blas[r3 + 2] + vo;
    auto uv_at = [&]() {
      auto at = [&](int32_t i) { return UV{T->uvs[(size_t)i * 2], T->uvs[(size_t)i * 2 + 1]}; };
      const UV t0 = at(i0), t1 = at(i1), t2 = at(i2);
      UV uv{b.x * t0.x + b.y * t1.x + b.z * t2.x, b.x * t0.y + b.y * t1.y + b.z * t2.y};
      return UV{uv.x * m[M_UV_SCALE], uv.y * m[M_UV_SCALE + 1]};
    };
    float s[4];
    const uint32_t base = (uint32_t)m[M_MAPS + MAP_BASE], amap = (uint32_t)m[M_MAPS + MAP_ALPHA];
    if (m[M_MAP_ALPHA_USAGE] == 2.0f && base) {
      sample_map(*T, base - 1u, uv_at(), s);
      alpha = alpha * s[3];
    }
    if (amap) {
      sample_map(*T, amap - 1u, uv_at(), s);
      alpha = alpha * s[1];
    }
  }
  if (m[M_ALPHA_MODE] == 0.0f) alpha = 1.0f;
  else if (m[M_ALPHA_MODE] == 1.0f) alpha = alpha >= m[M_ALPHA_CUTOFF] ? 1.0f : 0.0f;
  return alpha >= 1.0f;
}

static float nee_mis(uint32_t count, float lightPdf, float scatPdf) {
  if (count <= 1) return power_heuristic(lightPdf, scatPdf);
  float w = (float)count * lightPdf;
  return w / mxf(w + scatPdf, 1e-20f);
}
static float bsdf_mis(uint32_t count, float scatPdf, float lightPdf) {
  if (count <= 1) return power_heuristic(scatPdf, lightPdf);
  float w = (float)count * lightPdf;
  return scatPdf / mxf(scatPdf + w, 1e-20f);
}

struct ShadowWork {
  V3 ro, rd, contribution;
  float maxDistance, risTarget, lightSelectionPmf;
  int32_t triangleIndex;
  bool valid;
};

// finishDirectLight (sampleMaterial.wgsl.ts:32-151), surface vertex.
static bool finish_direct_light(const Ctx &c, const Surface &si, const PathState &path, V3 ff, const LightSample &ls,
                                V3 f, float scatPdf, ShadowWork &w) {
  w.valid = false; w.lightSelectionPmf = 0.0f; w.risTarget = 0.0f;
  V3 Li = ls.emission;
  V3 lightDir = ls.direction;
  float lightPdf = ls.pdf, selPmf = ls.selectionPmf;
  float cosThetaL = dot(ff, lightDir);
  if ((lightPdf <= 0.0f) | (selPmf <= 0.0f) | (maxc(Li) <= 0.0f)) return false;
  if ((scatPdf <= 0.0f) | (maxc(f) <= 0.0f)) return false;
  V3 origin = offset_surface_ray(si.position, si.faceNormal, lightDir);
  const float cutoff = 0.1f;
  float nl = dot(si.normal, lightDir);
  V3 gn = nl < 0.0f ? -si.faceNormal : si.faceNormal;
  nl = fabsf(nl);
  float gl = dot(gn, lightDir);
  float amount = nl < cutoff ? clampf(2.0f - (gl + nl) / cutoff, 0.0f, 1.0f) : clampf(1.0f - gl / cutoff, 0.0f, 1.0f);
  origin = origin + si.shadowOffset * amount;
  float visDist = ls.dist - EPS;
  V3 visDir = lightDir;
  if (ls.sourceType == 1u) {  // sampleMaterial.wgsl.ts:98-110: end EPS off the emitter along its normal
    V3 lightEndpoint = si.position + ls.direction * ls.dist;
    V3 offsetLightEndpoint = lightEndpoint + EPS * ls.normal;
    V3 toOffsetLight = offsetLightEndpoint - origin;
    visDist = length(toOffsetLight);
    visDir = normalize(toOffsetLight);
  }
  if ((cosThetaL < 0.0f) & (ls.sourceType == 0u)) {
    float u, v;
    environment_uv(lightDir, u, v);
    Li = c.s.env ? texture_linear(c.s, u, v) : V3(0.0f);
  }
  float mis = ls.isDelta ? 1.0f : nee_mis(c.s.candidates, lightPdf, scatPdf);
  if (mis > 0.0f) {
    V3 un = Li * f / lightPdf * path.beta;
    w.risTarget = mxf(dot(vmax(un * selPmf, V3(0.0f)), K_LUM), 0.0f);
    w.ro = origin; w.rd = visDir; w.maxDistance = visDist; w.triangleIndex = si.triangleIndex;
    w.contribution = mis * un;
    w.lightSelectionPmf = selPmf;
    w.valid = true;
    return true;
  }
  return false;
}

struct Scatter { V3 direction, response; float pdf; bool transmission; };

// sampleScattering (sampleMaterial.wgsl.ts:160-277), surface vertex.
static Scatter sample_scattering(Ctx &c, Sampler &smp, const Surface &si, const Mat &ma, const Prepared &P,
                                 PathState &path, V3 ff, bool &discreteOut) {
  uint32_t count = std::max(c.s.candidates, 1u);
  ShadowWork sel{};
  sel.valid = false;
  float selTarget = 0.0f, wsum = 0.0f;
  Scatter res{};
  res.pdf = 0.0f;
  discreteOut = false;
  for (uint32_t ev = 0; ev <= count; ev++) {
    bool cont = ev == count;
    LightSample ls{};
    V3 lightDir;
    float scatPdf = 0.0f;
    V3 resp(0.0f);
    bool discrete = false;
    if (cont) {
      float ucw = 1.0f;
      bool valid = sel.valid;
      if (count > 1u) {
        valid = valid & (selTarget > 0.0f);
        if (valid) ucw = sel.lightSelectionPmf * wsum / ((float)count * selTarget);
      }
      if (valid) {
        c.rays++;
        const Coverage cov{alpha_covers, &c};
        if (!trace_any(*c.s.trace, sel.ro, sel.rd, sel.maxDistance, sel.triangleIndex, si.instanceIndex, nullptr, &cov))
          path.li = path.li + sel.contribution * ucw;
      }
      float d2[2];
      smp.sample2d(smp.bounce_dim(BSDF_DIRECTION), d2);
      float lobe = smp.sample1d(smp.bounce_dim(BSDF_LOBE));
      discrete = !P.material.thinWalled & (fabsf(P.material.relativeIor - 1.0f) <= 1e-6f) &
                 (P.p[L_TRANS] >= 1.0f - 1e-6f);
      if (discrete) {
        lightDir = to_world(P.basis, -P.viewLocal);
        resp = P.material.transmissionColor * coat_base_scale(P, -P.viewLocal.z);
        scatPdf = 1.0f;
        res.transmission = true;
      } else {
        bool tr;
        if (!sample_direction(P, d2[0], d2[1], lobe, lightDir, tr)) return res;
        res.transmission = tr;
      }
      discreteOut = discrete;
    } else {
      float l2[2];
      smp.sample2d(smp.light_dim(ev, 1u), l2);
      float l1 = smp.sample1d(smp.light_dim(ev, 0u));
      sample_light_candidate(c, si.position, l2[0], l2[1], l1, ev, count, ls);
      lightDir = ls.direction;
      if ((ls.pdf <= 0.0f) | (ls.selectionPmf <= 0.0f) | (maxc(ls.emission) <= 0.0f)) continue;
      float cosThetaL = dot(ff, lightDir);
      if ((cosThetaL <= 0.0f) & (ma.transmissionWeight <= 0.0f) & (ma.subsurfaceWeight <= 0.0f)) continue;
    }
    if (!discrete) resp = eval_prepared(P, lightDir, scatPdf);
    if (cont) {
      res.direction = lightDir;
      res.response = resp;
      res.pdf = scatPdf;
      return res;
    }
    ShadowWork cand{};
    if (!finish_direct_light(c, si, path, ff, ls, resp, scatPdf, cand)) continue;
    if (count == 1u) { sel = cand; continue; }
    float sw = cand.risTarget / mxf(cand.lightSelectionPmf, 1e-20f);
    if (!(sw > 0.0f) | (sw != sw)) continue;
    wsum += sw;
    if (!sel.valid || smp.light_candidate_random(ev) * wsum <= sw) {
      sel = cand;
      selTarget = cand.risTarget;
    }
  }
  return res;
}

// SampleMaterial (sampleMaterial.wgsl.ts:316-508) without media.
static void sample_material(Ctx &c, Sampler &smp, Surface &si, Mat &ma, PathState &path) {
  float ndv1 = -dot(si.faceNormal, path.rd);
  float ndv2 = -dot(si.normal, path.rd);
  if (((ma.transmissionWeight > 0.0f) | (ma.subsurfaceWeight > 0.0f)) & ma.thinWalled) {
    // Always treat thin material as entering medium.
    if (ndv1 < 0.0f) {
      si.faceNormal = -si.faceNormal;
      ndv1 = -ndv1;
    }
    if (ndv2 < 0.0f) {
      si.normal = -si.normal;
      ndv2 = -ndv2;
    }
  }
  bool useFaceN = (ndv1 >= 0.0f) & (ndv2 < 0.0f);
  float ndv = useFaceN ? ndv1 : ndv2;
  V3 N = useFaceN ? si.faceNormal : si.normal;
  ma.inside = (ndv <= 0.0f) & !ma.thinWalled;
  ma.relativeIor = ndv > 0.0f ? 1.0f / ma.specularIor : ma.specularIor;
  V3 ff = ndv >= 0.0f ? N : -N;
  path.scatteringIsDelta = false;
  // Surface emission (sampleMaterial.wgsl.ts:412-422) with emissiveHitMisWeight (:279-303).
  {
    float cosLight = dot(si.faceNormal, -path.rd);
    float weight = 1.0f;
    const float *m = c.s.materials + (size_t)si.materialIndex * M_FLOATS;
    if (cosLight <= 0.0f && m[M_DOUBLE_SIDED] < 0.5f) {
      weight = 0.0f;
    } else if (!(path.fromEye | ((float)c.T.totalWeight <= 0.0f))) {  // rtUniform.lightWeightTotal is f32
      float samplingLuminance = dot(V3(m[M_EMISSION], m[M_EMISSION + 1], m[M_EMISSION + 2]), K_LUM);
      if (!((samplingLuminance <= 0.0f) | (cosLight <= 0.0f))) {
        float hitDistance = length(si.position - path.ro);
        float lightPdf = samplingLuminance / (float)c.T.totalWeight * hitDistance * hitDistance / mxf(cosLight, 1e-8f);
        weight = bsdf_mis(c.s.candidates, path.scatteringPdf, lightPdf);
      }
    }
    path.li = path.li + weight * ma.emission * path.beta;
  }
  Prepared P = prepare(ma, -path.rd, ff);
  bool discrete = false;
  Scatter sc = sample_scattering(c, smp, si, ma, P, path, ff, discrete);
  V3 f = sc.response;
  bool finite = std::isfinite(f.x) && std::isfinite(f.y) && std::isfinite(f.z) && fabsf(f.x) < 3.402823e38f &&
                fabsf(f.y) < 3.402823e38f && fabsf(f.z) < 3.402823e38f;
  if (!(sc.pdf > 0.0f) | !finite | (dot(sc.direction, sc.direction) < 0.5f)) {
    path.scatteringPdf = 0.0f;
    path.abort = true;
    return;
  }
  path.scatteringIsDelta = discrete;
  path.beta = path.beta * (f / sc.pdf);
  if (path.beta.x == 0.0f && path.beta.y == 0.0f && path.beta.z == 0.0f) { path.abort = true; return; }
  path.isNextRefraction = sc.transmission;
  path.ro = si.position + EPS * sc.direction;
  path.scatteringPdf = sc.pdf;
  path.rd = sc.direction;
  path.fromEye = false;
}

// camera (rayTrace.compute.ts:591-611, browser projectionInv variant).
static void camera_ray(const float *cam, float cx, float cy, V3 &o, V3 &d) {
  const float *projInv = cam, *viewInv = cam + 16;
  float ndcX = cx * 2.0f - 1.0f, ndcY = cy * 2.0f - 1.0f;
  float a[4] = {ndcX, ndcY, -1.0f, 1.0f}, b[4] = {ndcX, ndcY, 1.0f, 1.0f}, og[4], ce[4];
  mat_mul4(projInv, a, og);
  mat_mul4(projInv, b, ce);
  V3 dir(ce[0] * og[3] - og[0] * ce[3], ce[1] * og[3] - og[1] * ce[3], ce[2] * og[3] - og[2] * ce[3]);
  V3 op(og[0] / og[3], og[1] / og[3], og[2] / og[3]);
  o = mat_mul_point(viewInv, op, 1.0f);
  d = normalize(mat_mul_point(viewInv, dir, 0.0f));
}

// sampleCircle (rayTrace.compute.ts:1148-1154).
static void sample_circle(float ux, float uy, float &cx, float &cy) {
  float px = 2.0f * ux - 1.0f, py = 2.0f * uy - 1.0f;
  bool greater = fabsf(px) > fabsf(py);
  float r = greater ? px : py;
  float theta = greater ? K_QUARTER_PI * py / px : K_PI * (0.5f - 0.25f * px / py);
  cx = r * mtxfm::cosf(theta);
  cy = r * mtxfm::sinf(theta);
}

// Thin lens (rayTrace.compute.ts:1196-1205).
static void thin_lens(const float *cam, Sampler &smp, V3 &o, V3 &d) {
  const float *viewInv = cam + 16;
  float aperture = cam[32], focus = cam[33];
  V3 eye(viewInv[12], viewInv[13], viewInv[14]);
  V3 forward = -normalize(V3(viewInv[8], viewInv[9], viewInv[10]));
  V3 focusPoint = eye + d * (focus / dot(d, forward));
  float u[2], lx, ly;
  smp.sample2d(CAMERA_LENS, u);
  sample_circle(u[0], u[1], lx, ly);
  V3 origin = mat_mul_point(viewInv, V3(aperture * lx, aperture * ly, 0.0f), 1.0f);
  o = origin;
  d = normalize(focusPoint - origin);
}

// main (rayTrace.compute.ts:1161-1205): pixel-centre uv, jitter, the
// random pattern's rng_initialize, camera ray, thin lens. The jitter is the
// CAMERA_PIXEL sample for Sobol/ZSobol; for the random pattern it is the
// rtUniform.jitter the pipeline writes before each sample
// (WGPURenderingPipeline.ts:534-541): Halton bases 2 and 3 of the sample
// index, minus 0.5, over the image size, in FP64, stored as f32.
static void primary_ray(const float *camera, uint32_t width, uint32_t height, uint32_t px, uint32_t py,
                        Sampler &smp, V3 &o, V3 &d) {
  float W = (float)width, H = (float)height;
  float uvx = ((float)px + 0.5f) / W, uvy = ((float)py + 0.5f) / H;
  float jx, jy;
  if (smp.pattern == 2u) {
    jx = (float)((halton(smp.pixelSampleIndex, 2) - 0.5) / width);
    jy = (float)((halton(smp.pixelSampleIndex, 3) - 0.5) / height);
  } else {
    float j[2];
    smp.sample2d(CAMERA_PIXEL, j);
    jx = (j[0] - 0.5f) / W;
    jy = (j[1] - 0.5f) / H;
  }
  smp.rng_initialize();
  float cx = uvx + jx, cy = uvy + jy;
  camera_ray(camera, cx, cy, o, d);
  if (camera[32] > 0.0f) thin_lens(camera, smp, o, d);
}

static Sampler make_sampler(uint32_t px, uint32_t py, uint32_t width, uint32_t index, uint32_t pattern,
                            uint32_t seed, uint32_t logSPP) {
  Sampler smp{};
  smp.px = px; smp.py = py; smp.width = width;
  smp.pixelSampleIndex = index;
  smp.sequenceSeed = seed; smp.zSobolLogSPP = logSPP; smp.pattern = pattern; smp.bevelSamples = 0;
  smp.samplingBounce = 0; smp.hashedCounter = 0;
  return smp;
}

void camera_sample(const float *camera, uint32_t width, uint32_t height, uint32_t px, uint32_t py, uint32_t index,
                   uint32_t pattern, uint32_t seed, uint32_t logSPP, float out[6]) {
  Sampler smp = make_sampler(px, py, width, index, pattern, seed, logSPP);
  V3 o, d;
  primary_ray(camera, width, height, px, py, smp, o, d);
  out[0] = o.x; out[1] = o.y; out[2] = o.z; out[3] = d.x; out[4] = d.y; out[5] = d.z;
}

static void integrate_pixel(Ctx &c, uint32_t px, uint32_t py, uint32_t sampleIndex, Accum &acc,
                            uint32_t *drawLog = nullptr, uint32_t drawCap = 0, uint32_t *drawCount = nullptr) {
  const RenderScene &s = c.s;
  Sampler smp = make_sampler(px, py, s.width, sampleIndex + s.sampleOffset, s.pattern, s.seed, s.logSPP);
  smp.log = drawLog;
  smp.logCap = drawCap;
  PathState path{};
  primary_ray(s.camera, s.width, s.height, px, py, smp, path.ro, path.rd);
  Aux aux{V3(0.0f), V3(0.0f), false};
  // integrator (rayTrace.compute.ts:891-946)
  Surface si{};
  si.triangleIndex = 0; si.instanceIndex = 0; si.materialIndex = 0; si.albedoSampled = false;
  Mat ma{};
  path.alpha = 1.0f;
  path.beta = V3(1.0f);
  path.fromEye = true;
  int bounceLimit = (int)s.bounces;
  int eventLimit = std::min(bounceLimit * 3, 50);
  bool isNotTransparent = false;
  int depth = 0;
  bool sampleIgnored = false;
  for (int event = 0; event <= eventLimit; event++) {
    intersect_scene(c, path, si, depth, event != 0);
    if (si.hit) {
      int32_t mi = si.materialIndex;
      ma = get_material(s.materials + (size_t)mi * M_FLOATS, instance_color(*s.trace, si.instanceIndex), s.textures,
                        si.uv, si.uv1);
    }
    if (event == 0) {
      if (si.hit) {
        aux.albedo = ma.transmissionWeight > 0.5f ? V3(1.0f) : ma.baseColor;
        V3 camPos(s.camera[16 + 12], s.camera[16 + 13], s.camera[16 + 14]);
        aux.normal = dot(si.normal, si.position - camPos) > 0.0f ? -si.normal : si.normal;
        si.albedoSampled = ma.transmissionWeight > 0.5f;
      }
      isNotTransparent = ma.transmissionWeight < 0.001f && ma.subsurfaceWeight < 0.001f && ma.alpha > 0.99f;
    } else if (!sampleIgnored) {
      // completeVertex (rayTrace.compute.ts:514-589); it returns at once
      // after a passed-through vertex (:527-529).
      bool isAlphaOccluded = smp.hashed_sample(3u) < ma.alpha;
      for (uint32_t i = 0; i < s.lightCount; i++) {
        LightSample ls{};
        const float *L = s.lights + (size_t)i * 16;
        float d = intersect_light(L, path.ro, path.rd, ls);
        ls.pdf *= c.T.lightPmf[i];
        V3 dd = si.position - path.ro;
        bool occluded = si.hit && d > length(dd);
        if (d != INF && !(occluded && isAlphaOccluded)) {
          float mis = bsdf_mis(s.candidates, path.scatteringPdf, ls.pdf);
          path.li = path.li + mis * path.beta * ls.emission;
          path.alpha = 1.0f;
          path.abort = true;
        }
      }
      if (s.env && !(si.hit && isAlphaOccluded)) {
        float u, v;
        environment_uv(path.rd, u, v);
        V3 em = texture_linear(s, u, v);
        if (path.isNextRefraction) {  // displayed background = environment here
          float mis = path.scatteringIsDelta ? 1.0f
                                             : bsdf_mis(s.candidates, path.scatteringPdf, env_map_pdf(c, u, v) * c.T.envPmf);
          path.li = path.li + mis * path.beta * em * 1.0f;
        } else {
          float mis = bsdf_mis(s.candidates, path.scatteringPdf, env_map_pdf(c, u, v) * c.T.envPmf);
          path.li = path.li + path.beta * em * mis;
        }
        path.alpha = 1.0f;
        path.abort = true;
      }
      depth++;
    }
    if (isNotTransparent && event > bounceLimit) path.abort = true;
    if (path.abort || event >= eventLimit) break;
    // shadeVertex (rayTrace.compute.ts:343-512)
    sampleIgnored = false;
    smp.samplingBounce = (uint32_t)std::max(depth, 0);
    if (!si.hit) {
      if (path.fromEye) {
        V3 bg(0.0f);
        float ba = 0.0f;
        if (s.env) {
          float u, v;
          environment_uv(path.rd, u, v);
          bg = texture_linear(s, u, v);
          ba = 1.0f;
        }
        path.li = path.beta * bg * ba;
        path.alpha = ba;
        aux.isBlank = ba < 0.001f;
      }
      path.abort = true;
      break;
    }
    // Alpha coverage (:432-458): a surface of coverage <= 0 (or one the
    // coverage draw misses) is passed through: no material sample, guides or
    // roulette, and the ray moves on from position + d * EPS.
    const float coverage = clampf(ma.alpha, 0.0f, 1.0f);
    if (coverage < 1.0f && (coverage <= 0.0f || smp.hashed_sample(3u) > coverage)) {
      sampleIgnored = true;
      path.ro = si.position + path.rd * EPS;
      continue;
    }
    sample_material(c, smp, si, ma, path);
    if (!si.albedoSampled) {
      if (!(ma.transmissionWeight > 0.5f || (ma.specularRoughness <= 0.02f && ma.baseMetalness > 0.98f))) {
        aux.albedo = ma.baseColor;
        V3 camPos(s.camera[16 + 12], s.camera[16 + 13], s.camera[16 + 14]);
        aux.normal = dot(si.normal, si.position - camPos) > 0.0f ? -si.normal : si.normal;
        si.albedoSampled = true;
      }
      if (depth > 4) si.albedoSampled = true;
    }
    if (depth >= 2) {
      float q = mxf(0.0f, 1.0f - maxc(path.beta));
      if (smp.sample1d(smp.bounce_dim(RUSSIAN_ROULETTE)) < q) {
        path.alpha = 1.0f;
        path.abort = true;
      } else {
        path.beta = path.beta / (1.0f - q);
      }
    }
    if (path.abort) break;
  }
  V3 li = path.li;
  float alpha = path.alpha;
  if (s.clamping > 0.0f) li = clamp3(li, -s.clamping, s.clamping);
  if (!(li.x < INF && li.x > -EPS)) { li = V3(0.0f); alpha = 1.0f; }
  acc.color[0] += li.x; acc.color[1] += li.y; acc.color[2] += li.z; acc.color[3] += alpha;
  acc.albedo[0] += aux.albedo.x; acc.albedo[1] += aux.albedo.y; acc.albedo[2] += aux.albedo.z;
  acc.albedo[3] += aux.isBlank ? 0.0 : 1.0;
  acc.normal[0] += aux.normal.x; acc.normal[1] += aux.normal.y; acc.normal[2] += aux.normal.z; acc.normal[3] += 1.0;
  double lum = (double)dot(li, K_LUM);
  acc.lum2 += lum * lum;
  if (drawCount) *drawCount = smp.logLen;
}

uint32_t path_draws(const RenderScene &s, uint32_t px, uint32_t py, uint32_t sampleIndex, uint32_t *log,
                    uint32_t cap) {
  HostTables T;
  build_tables(s, T);
  build_env_cdf(s, T);
  Ctx c{s, T, 0};
  Accum acc{};
  uint32_t n = 0;
  integrate_pixel(c, px, py, sampleIndex, acc, log, cap, &n);
  return n;
}

uint64_t render(const RenderScene &s, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t sampleStart,
                uint32_t count, Accum *out) {
  HostTables T;
  build_tables(s, T);
  build_env_cdf(s, T);
  uint64_t rays = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : rays)
  for (int64_t i = 0; i < (int64_t)w * h; i++) {
    uint32_t px = x0 + (uint32_t)(i % w), py = y0 + (uint32_t)(i / w);
    Ctx c{s, T, 0};
    for (uint32_t k = 0; k < count; k++) integrate_pixel(c, px, py, sampleStart + k, out[i]);
    rays += c.rays;
  }
  return rays;
}

// ------------------------------------------------------------ BSDF probes
// The material as sample_material prepares it on the outside of a surface
// whose shading and face normals are +z, seen along `view` (unit vector from
// the surface toward the viewer, view.z > 0). Used by the oracle's own tests
// to check the lobes against closed forms and quadrature.
static Prepared probe_prepare(const float *m, const float *view) {
  Mat ma = get_material(m, V3(1.0f));
  ma.inside = false;
  ma.relativeIor = 1.0f / ma.specularIor;
  return prepare(ma, V3(view[0], view[1], view[2]), V3(0.0f, 0.0f, 1.0f));
}

void probe_bsdf_eval(const float *m, const float *view, uint32_t n, const float *dirs, float *f, float *pdf) {
  Prepared P = probe_prepare(m, view);
  for (uint32_t i = 0; i < n; i++) {
    float p = 0.0f;
    V3 r = eval_prepared(P, V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), p);
    f[3 * i] = r.x;
    f[3 * i + 1] = r.y;
    f[3 * i + 2] = r.z;
    pdf[i] = p;
  }
}

void probe_bsdf_sample(const float *m, const float *view, uint32_t n, const float *u, float *dirs, uint8_t *ok) {
  Prepared P = probe_prepare(m, view);
  for (uint32_t i = 0; i < n; i++) {
    V3 d;
    bool transmission = false;
    ok[i] = sample_direction(P, u[3 * i], u[3 * i + 1], u[3 * i + 2], d, transmission) ? 1 : 0;
    dirs[3 * i] = d.x;
    dirs[3 * i + 1] = d.y;
    dirs[3 * i + 2] = d.z;
  }
}

void probe_lobes(const float *m, const float *view, float *probabilities, float *coatBaseScale) {
  Prepared P = probe_prepare(m, view);
  for (int l = 0; l < L_COUNT; l++) probabilities[l] = P.p[l];
  coatBaseScale[0] = P.coatBaseScale.x;
  coatBaseScale[1] = P.coatBaseScale.y;
  coatBaseScale[2] = P.coatBaseScale.z;
}

void probe_metal_fresnel(float c, const float *f0, const float *tint, float *out, float *average) {
  V3 a(f0[0], f0[1], f0[2]), t(tint[0], tint[1], tint[2]);
  V3 r = metal_fresnel(c, a, t), avg = metal_average_fresnel(a, t);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
  average[0] = avg.x; average[1] = avg.y; average[2] = avg.z;
}

float probe_coat_energy(const float *m, float c) { return coat_energy(get_material(m, V3(1.0f)), c); }

}  // namespace orc
