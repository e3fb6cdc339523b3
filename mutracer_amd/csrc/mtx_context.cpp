// Persistent contexts over the raw scene descriptor: descriptor validation and
// decoding (native/docs/scene-format.md, record layouts of
// native/include/mutracer-scene.h and mutracer-layout.h) and the context / job
// contract of native/include/mutracer-context.h:12-187, on top of the scene
// and render-target entry points of include/mtx.h. Host code only; the GPU work
// is the library's render target on the context's own HIP stream.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mtx.h"

namespace {

// ------------------------------------------------------------------ errors
struct Fail {
  int32_t code;
  std::string msg;
};
[[noreturn]] void fail(int32_t code, const std::string &msg) { throw Fail{code, msg}; }
void need(bool ok, const std::string &msg) {
  if (!ok) fail(MTX_INVALID_ARGUMENT, msg);
}
[[noreturn]] void unsupported(const std::string &what) {
  fail(MTX_UNAVAILABLE, what + " is not supported by this renderer");
}

void write_error(char *err, size_t cap, const std::string &msg) {
  if (!err || !cap) return;
  size_t n = std::min(msg.size(), cap - 1);
  while (n > 0 && n < msg.size() && ((unsigned char)msg[n] & 0xC0) == 0x80) n--;  // whole UTF-8 characters
  std::memcpy(err, msg.data(), n);
  err[n] = 0;
}

// Calls an mtx_* entry point; a nonzero status becomes a Fail with its message.
template <typename F>
void call(F &&f) {
  char msg[512] = {0};
  const int32_t rc = f(msg, sizeof msg);
  if (rc != MTX_OK) fail(rc, msg[0] ? msg : "library call failed");
}

// ---------------------------------------------------------- record layout
// Word offsets and record sizes of the scene descriptor (mutracer-layout.h).
enum : uint32_t {
  kHeaderWords = 32, kGeometryWords = 12, kMeshWords = 21, kMaterialWords = 24, kExtensionWords = 4,
  kLightWords = 32, kLightWordsV5 = 28, kCameraWords = 32, kBatchWords = 4,
  kTextureWords = 13, kTextureWordsV6 = 12, kNone = 0xffffffffu,
};
namespace H {
enum : uint32_t {
  Version = 0, Length, Width, Height, GeometryCount, MeshCount, MaterialCount, LightCount, TextureCount,
  GeometryOffset, MeshOffset, MaterialOffset, LightOffset, TextureOffset, CameraOffset, Reserved0,
  Background = 16, EnvironmentTexture = 20, BackgroundTexture, QualityFlags, BatchOffset, MeshIds, MaterialIds,
  EnvironmentSettings, GroundedBackground, ReservedEnd,
};
}
namespace G {
enum : uint32_t { VertexCount = 0, IndexCount, Position, Indices, Normal, Uv, Uv1, Color, Tangent, MaterialIndex, Reserved };
}
namespace M {
enum : uint32_t { Geometry = 0, Material, MaterialTable, MaterialCount, Transform, Flags = 20 };
}
namespace MAT {
enum : uint32_t {
  BaseColor = 0, Roughness = 3, Metalness, Transmission, Ior, EmissionColor, EmissionLuminance = 10, CoatWeight,
  CoatRoughness, Opacity, ThinWalled, Map, NormalMap, MetallicRoughnessMap, EmissiveMap, ExtensionCount,
  ExtensionOffset, Graph, Reserved,
};
}
namespace L {
enum : uint32_t {
  Type = 0, Color, Intensity = 4, Size, HeightOrInner, SpreadOrOuter, Texture, IesOffset, IesBytes, Reserved,
  Transform, Radius = 28, Length, Normalize, TreatAsLine,
};
}
namespace CAM {
enum : uint32_t {
  Projection = 0, Fov, Near, Far, Left, Right, Bottom, Top, Zoom, FStop, FocusDistance, SensorHeight,
  MetersPerUnit, Reserved, Transform = 16,
};
}
namespace T {
enum : uint32_t { Width = 0, Height, Storage, Pixels, FlipY, UvChannel, UvScale, UvOffset = 8, WrapS = 10, WrapT, Udim };
}

// Read-only view of the descriptor words with the validator's checks.
struct Words {
  const uint32_t *w;
  size_t n;
  uint32_t u(size_t i) const { return w[i]; }
  float f(size_t i) const {
    float v;
    std::memcpy(&v, &w[i], 4);
    return v;
  }
  void span(size_t off, size_t count, const char *what) const {
    need(off >= kHeaderWords && off < n && count <= n - off,
         std::string(what) + " lies outside the descriptor");
  }
  void finite(size_t off, size_t count, const char *what) const {
    for (size_t i = 0; i < count; i++) need(std::isfinite(f(off + i)), std::string(what) + " holds a non-finite float");
  }
  void zero(size_t off, size_t count, const char *what) const {
    for (size_t i = 0; i < count; i++) need(w[off + i] == 0, std::string(what) + ": reserved words must be 0");
  }
  // Column-major affine, invertible 4x4 (scene-format.md "Layout").
  void affine(size_t off, const char *what) const {
    finite(off, 16, what);
    need(f(off + 3) == 0.0f && f(off + 7) == 0.0f && f(off + 11) == 0.0f && f(off + 15) == 1.0f,
         std::string(what) + " is not affine");
    const double a = f(off), b = f(off + 4), c = f(off + 8), d = f(off + 1), e = f(off + 5), g = f(off + 9),
                 h = f(off + 2), i = f(off + 6), k = f(off + 10);
    const double det = a * (e * k - g * i) - b * (d * k - g * h) + c * (d * i - e * h);
    need(std::isfinite(det) && std::fabs(det) > 1e-30, std::string(what) + " is not invertible");
  }
};

// ------------------------------------------------------------ decoded scene
constexpr int kMatFloats = MTX_MATERIAL_FLOATS;

struct MeshRef {  // one instance: which mesh record and batch entry it came from
  uint32_t mesh, batchEntry;  // batchEntry = kNone for an ordinary mesh
};

struct Decoded {
  uint32_t version = 0, width = 0, height = 0;
  std::vector<std::vector<float>> positions, normals, uvs, uvs1;  // per BLAS mesh (uvs: empty when absent)
  std::vector<std::vector<uint32_t>> indices;
  std::vector<mtx_texture_desc> textures;        // material textures (pixels set by scene_desc)
  std::vector<std::vector<uint32_t>> texPixels;  // their pixel words (RGBA8 words or RGBA32F float bits)
  std::vector<mtx_instance_desc> instances;
  std::vector<MeshRef> instanceSource;
  std::vector<float> materials, lights;
  std::vector<float> env;
  uint32_t envW = 0, envH = 0;
  mtx_camera_desc camera{};
  size_t cameraOffset = 0;
};

// Material record + extensions -> the packed OpenPBR material
// (decodeMaterial, native/src/scene-records.ts:40-56, and
// applyNativeMaterialProperties, native/src/material-properties.ts:5-88;
// defaults of src/scene/openPbrSchema.ts:32-69). Texture slots hold the
// descriptor's texture index + 1 (0 = none); the caller remaps them to the
// scene's material texture table.
void decode_material(const Words &W, size_t o, uint32_t version, uint32_t textureCount, float *m) {
  W.finite(o + MAT::BaseColor, 14, "material");
  const uint32_t thin = W.u(o + MAT::ThinWalled);
  need(thin <= 1, "material thin_walled must be 0 or 1");
  uint32_t maps[MTX_MAP_COUNT];
  std::fill(maps, maps + MTX_MAP_COUNT, kNone);
  const int recordSlot[4] = {MTX_MAP_BASE, MTX_MAP_NORMAL, MTX_MAP_METALLIC_ROUGHNESS, MTX_MAP_EMISSIVE};
  for (uint32_t t = MAT::Map; t <= MAT::EmissiveMap; t++) {
    const uint32_t tex = W.u(o + t);
    need(tex == kNone || tex < textureCount, "material texture index out of range");
    maps[recordSlot[t - MAT::Map]] = tex;
  }
  if (W.u(o + MAT::Graph)) {
    need(version >= 8, "material graph_offset must be 0 before version 8");
    unsupported("material graphs");
  }
  W.zero(o + MAT::Reserved, 2, "material");
  struct P {
    float baseWeight = 1, baseColor[3], diffuseRoughness = 0, metalness, specularWeight = 1,
          specularColor[3] = {1, 1, 1}, roughness, ior, anisotropy = 0, transmission, transmissionColor[3] = {1, 1, 1},
          subsurfaceWeight = 0, subsurfaceColor[3] = {0.8f, 0.8f, 0.8f}, fuzzWeight = 0, fuzzColor[3] = {1, 1, 1},
          fuzzRoughness = 0.5f, coatWeight, coatColor[3] = {1, 1, 1}, coatRoughness, coatIor = 1.6f,
          coatDarkening = 1, emissionColor[3], emissionLuminance, opacity, alphaCutoff = 0,
          normalScale[2] = {1, 1}, uvScale[2] = {1, 1};
    uint32_t alphaMode = 0, mapAlphaUsage = 0;  // native enums: auto/opaque/mask/blend, opaque/alpha/transmission
    bool thinWalled, doubleSided = false;
  } p;
  for (int k = 0; k < 3; k++) {
    p.baseColor[k] = W.f(o + MAT::BaseColor + k);
    p.emissionColor[k] = W.f(o + MAT::EmissionColor + k);
  }
  p.roughness = W.f(o + MAT::Roughness);
  p.metalness = W.f(o + MAT::Metalness);
  p.transmission = W.f(o + MAT::Transmission);
  p.ior = W.f(o + MAT::Ior);
  p.emissionLuminance = W.f(o + MAT::EmissionLuminance);
  p.coatWeight = W.f(o + MAT::CoatWeight);
  p.coatRoughness = W.f(o + MAT::CoatRoughness);
  p.opacity = W.f(o + MAT::Opacity);
  p.thinWalled = thin != 0;
  const uint32_t ec = W.u(o + MAT::ExtensionCount), eo = W.u(o + MAT::ExtensionOffset);
  if (ec) W.span(eo, (size_t)ec * kExtensionWords, "material extensions");
  bool seen[128] = {false};
  for (uint32_t i = 0; i < ec; i++) {
    const size_t e = eo + (size_t)i * kExtensionWords;
    const uint32_t id = W.u(e);
    need(id >= 1 && id <= 76, "unknown material property id " + std::to_string(id));
    need(!seen[id], "material property " + std::to_string(id) + " appears twice");
    seen[id] = true;
    auto x = [&] { return W.f(e + 1); };
    auto color = [&](float *c) {
      W.finite(e + 1, 3, "material property");
      for (int k = 0; k < 3; k++) c[k] = W.f(e + 1 + k);
    };
    auto scalar = [&](float &v) {
      W.finite(e + 1, 1, "material property");
      W.zero(e + 2, 2, "material property");
      v = x();
    };
    auto flag = [&](bool &v) {
      need(W.u(e + 1) <= 1, "boolean material property must be 0 or 1");
      W.zero(e + 2, 2, "material property");
      v = W.u(e + 1) != 0;
    };
    // Properties whose default leaves the render unchanged may be given at
    // their default; any other value selects a feature this build lacks.
    auto only_default = [&](float def, const char *what) {
      float v;
      scalar(v);
      if (v != def) unsupported(what);
    };
    auto only_default3 = [&](const float *def, const char *what) {
      float c[3];
      color(c);
      if (c[0] != def[0] || c[1] != def[1] || c[2] != def[2]) unsupported(what);
    };
    auto texture = [&]() {
      const uint32_t t = W.u(e + 1);
      W.zero(e + 2, 2, "material property");
      need(t == kNone || t < textureCount, "material texture index out of range");
      return t;
    };
    auto no_texture = [&](const char *what) {
      if (texture() != kNone) unsupported(what);
    };
    auto vec2 = [&](float *v) {
      W.finite(e + 1, 2, "material property");
      W.zero(e + 3, 1, "material property");
      v[0] = W.f(e + 1);
      v[1] = W.f(e + 2);
    };
    static const float kZero3[3] = {0, 0, 0};
    switch (id) {
      case 1: scalar(p.baseWeight); break;
      case 2: color(p.baseColor); break;
      case 3: scalar(p.diffuseRoughness); break;
      case 4: scalar(p.metalness); break;
      case 5: scalar(p.specularWeight); break;
      case 6: color(p.specularColor); break;
      case 7: scalar(p.roughness); break;
      case 8: scalar(p.ior); break;
      case 9: scalar(p.anisotropy); break;
      case 10: scalar(p.transmission); break;
      case 11: color(p.transmissionColor); break;
      case 12: only_default(0, "transmission depth (participating media)"); break;
      case 13: only_default3(kZero3, "transmission scatter (participating media)"); break;
      case 14: only_default(0, "transmission scatter anisotropy"); break;
      case 15: only_default(0, "dispersion"); break;
      case 16: { float v; scalar(v); } break;  // Abbe number: only used with dispersion
      case 17: scalar(p.subsurfaceWeight); break;
      case 18: color(p.subsurfaceColor); break;
      case 19: { float v; scalar(v); } break;  // subsurface radius: the medium behind the boundary is not built
      case 20: { float c[3]; color(c); } break;  // radius scale of the same medium
      case 21: only_default(0, "subsurface anisotropy"); break;
      case 22: scalar(p.fuzzWeight); break;
      case 23: color(p.fuzzColor); break;
      case 24: scalar(p.fuzzRoughness); break;
      case 25: scalar(p.coatWeight); break;
      case 26: color(p.coatColor); break;
      case 27: scalar(p.coatRoughness); break;
      case 28: only_default(0, "coat anisotropy"); break;
      case 29: scalar(p.coatIor); break;
      case 30: scalar(p.coatDarkening); break;
      case 31: only_default(0, "thin film"); break;
      case 32: case 33: case 62: { float v; scalar(v); } break;  // thin-film shape: inert at weight 0
      case 34: color(p.emissionColor); break;
      case 35: scalar(p.emissionLuminance); break;
      case 36: scalar(p.opacity); break;
      case 37: flag(p.thinWalled); break;
      case 38: flag(p.doubleSided); break;
      case 41: {
        need(W.u(e + 1) <= 3, "alpha mode must be 0..3");
        W.zero(e + 2, 2, "material property");
        p.alphaMode = W.u(e + 1);
      } break;
      case 42: scalar(p.alphaCutoff); break;
      case 45: vec2(p.normalScale); break;
      case 56: vec2(p.uvScale); break;
      case 75: { float v[2]; vec2(v); } break;  // coat normal scale: the coat normal map is not built
      case 51: { bool v; flag(v); if (v) unsupported("shadow catchers"); } break;
      case 52: { bool v; flag(v); if (v) unsupported("unlit materials"); } break;
      case 53: only_default(0, "bevel"); break;
      case 54: {
        need(W.u(e + 1) <= 2, "map alpha usage must be 0..2");
        W.zero(e + 2, 2, "material property");
        p.mapAlphaUsage = W.u(e + 1);
      } break;
      case 55: case 57: {
        need(W.u(e + 1) <= 3, "UV projection must be 0..3");
        W.zero(e + 2, 2, "material property");
        if (W.u(e + 1) != 0) unsupported("UV projections");
      } break;
      case 58: case 60: case 61: case 71: case 72: case 76: { float v; scalar(v); } break;
      case 63: case 73: {
        need(W.u(e + 1) <= 3, "enum material property out of range");
        W.zero(e + 2, 2, "material property");
      } break;
      case 39: maps[MTX_MAP_BASE] = texture(); break;
      case 40: maps[MTX_MAP_ALPHA] = texture(); break;
      case 43: maps[MTX_MAP_EMISSIVE] = texture(); break;
      case 44: maps[MTX_MAP_NORMAL] = texture(); break;
      case 46: maps[MTX_MAP_METALLIC_ROUGHNESS] = texture(); break;
      case 47: maps[MTX_MAP_ROUGHNESS] = texture(); break;
      case 48: maps[MTX_MAP_METALNESS] = texture(); break;
      case 49: maps[MTX_MAP_TRANSMISSION] = texture(); break;
      case 64: maps[MTX_MAP_COAT_WEIGHT] = texture(); break;
      case 65: maps[MTX_MAP_COAT_ROUGHNESS] = texture(); break;
      case 66: maps[MTX_MAP_FUZZ_COLOR] = texture(); break;
      case 67: maps[MTX_MAP_FUZZ_ROUGHNESS] = texture(); break;
      case 68: maps[MTX_MAP_SPECULAR_WEIGHT] = texture(); break;
      case 69: maps[MTX_MAP_SPECULAR_COLOR] = texture(); break;
      case 50: no_texture("thin-film maps"); break;
      case 59: no_texture("displacement maps"); break;
      case 70: no_texture("anisotropy maps"); break;
      case 74: no_texture("coat normal maps"); break;
      default: break;
    }
  }
  // resolveMaterialAlphaMode (src/scene/MaterialState.ts:251-264), to the
  // packed modes 0 opaque, 1 mask, 2 blend; blend (stochastic coverage) is
  // not built.
  uint32_t resolvedAlpha = p.alphaMode == 1 ? 0u : p.alphaMode == 2 ? 1u : 2u;
  if (p.alphaMode == 0)
    resolvedAlpha = p.alphaCutoff > 0.0f ? 1u
                    : (p.opacity < 1.0f || p.mapAlphaUsage == 1 || maps[MTX_MAP_ALPHA] != kNone) ? 2u : 0u;
  if (resolvedAlpha == 2u) unsupported("blend alpha coverage (stochastic transparency)");
  if (p.anisotropy != 0.0f) unsupported("specular anisotropy");
  // materialEmission (src/renderer/materialEmission.ts:10-22), in f64 as JS;
  // a coat weight map leaves the coat tint to the sample.
  double em[3];
  for (int k = 0; k < 3; k++) em[k] = (double)p.emissionColor[k] * (double)p.emissionLuminance;
  if (p.coatWeight > 0.0f && maps[MTX_MAP_COAT_WEIGHT] == kNone) {
    const double w = std::min((double)p.coatWeight, 1.0);
    for (int k = 0; k < 3; k++) em[k] = em[k] * (1.0 + w * (std::sqrt(std::max((double)p.coatColor[k], 0.0)) - 1.0));
  }
  std::fill(m, m + kMatFloats, 0.0f);
  auto put3 = [&](int off, const float *c) { for (int k = 0; k < 3; k++) m[off + k] = c[k]; };
  m[MTX_MAT_BASE_WEIGHT] = p.baseWeight;
  put3(MTX_MAT_BASE_COLOR, p.baseColor);
  m[MTX_MAT_BASE_DIFFUSE_ROUGHNESS] = p.diffuseRoughness;
  m[MTX_MAT_BASE_METALNESS] = p.metalness;
  m[MTX_MAT_SPECULAR_WEIGHT] = p.specularWeight;
  put3(MTX_MAT_SPECULAR_COLOR, p.specularColor);
  m[MTX_MAT_SPECULAR_ROUGHNESS] = p.roughness;
  m[MTX_MAT_SPECULAR_IOR] = p.ior;
  m[MTX_MAT_TRANSMISSION_WEIGHT] = p.transmission;
  put3(MTX_MAT_TRANSMISSION_COLOR, p.transmissionColor);
  m[MTX_MAT_COAT_WEIGHT] = p.coatWeight;
  put3(MTX_MAT_COAT_COLOR, p.coatColor);
  m[MTX_MAT_COAT_ROUGHNESS] = p.coatRoughness;
  m[MTX_MAT_COAT_IOR] = p.coatIor;
  m[MTX_MAT_COAT_DARKENING] = p.coatDarkening;
  for (int k = 0; k < 3; k++) m[MTX_MAT_EMISSION + k] = (float)em[k];
  m[MTX_MAT_DOUBLE_SIDED] = p.doubleSided ? 1.0f : 0.0f;
  m[MTX_MAT_SPECULAR_ANISOTROPY] = p.anisotropy;
  m[MTX_MAT_FUZZ_WEIGHT] = p.fuzzWeight;
  put3(MTX_MAT_FUZZ_COLOR, p.fuzzColor);
  m[MTX_MAT_FUZZ_ROUGHNESS] = p.fuzzRoughness;
  m[MTX_MAT_SUBSURFACE_WEIGHT] = p.subsurfaceWeight;
  put3(MTX_MAT_SUBSURFACE_COLOR, p.subsurfaceColor);
  m[MTX_MAT_THIN_WALLED] = p.thinWalled ? 1.0f : 0.0f;
  m[MTX_MAT_OPACITY] = p.opacity;
  m[MTX_MAT_ALPHA_MODE] = (float)resolvedAlpha;
  m[MTX_MAT_ALPHA_CUTOFF] = std::max(p.alphaCutoff, 0.0f);
  m[MTX_MAT_MAP_ALPHA_USAGE] = p.mapAlphaUsage == 0 ? 0.0f : p.mapAlphaUsage == 1 ? 2.0f : 3.0f;
  m[MTX_MAT_NORMAL_SCALE] = p.normalScale[0];
  m[MTX_MAT_NORMAL_SCALE + 1] = p.normalScale[1];
  m[MTX_MAT_UV_SCALE] = p.uvScale[0];
  m[MTX_MAT_UV_SCALE + 1] = p.uvScale[1];
  for (int k = 0; k < MTX_MAP_COUNT; k++) m[MTX_MAT_MAPS + k] = maps[k] == kNone ? 0.0f : (float)(maps[k] + 1u);
}

// getRigidLightFrame (src/renderer/lightsData.ts:112-124) in f64.
struct Frame {
  double x[3], y[3], dir[3];
};
Frame rigid_frame(const Words &W, size_t t) {
  auto norm = [](const double *v, const double *fb, double *out) {
    const double l = std::hypot(v[0], v[1], v[2]);
    for (int k = 0; k < 3; k++) out[k] = l > 1e-12 ? v[k] / l : fb[k];
  };
  const double m0[3] = {W.f(t), W.f(t + 1), W.f(t + 2)}, m2[3] = {W.f(t + 8), W.f(t + 9), W.f(t + 10)};
  const double zfb[3] = {0, 0, 1}, xfb[3] = {1, 0, 0};
  double z[3];
  norm(m2, zfb, z);
  const double proj = m0[0] * z[0] + m0[1] * z[1] + m0[2] * z[2];
  double x[3] = {m0[0] - proj * z[0], m0[1] - proj * z[1], m0[2] - proj * z[2]};
  if (std::hypot(x[0], x[1], x[2]) <= 1e-12) {
    if (std::fabs(z[1]) < 0.99) { x[0] = z[2]; x[1] = 0; x[2] = -z[0]; }
    else { x[0] = 1; x[1] = 0; x[2] = 0; }
  }
  Frame F;
  norm(x, xfb, F.x);
  F.y[0] = z[1] * F.x[2] - z[2] * F.x[1];
  F.y[1] = z[2] * F.x[0] - z[0] * F.x[2];
  F.y[2] = z[0] * F.x[1] - z[1] * F.x[0];
  for (int k = 0; k < 3; k++) F.dir[k] = -z[k] + 0.0;  // `-z || 0` turns -0 into +0
  return F;
}

// Light record -> packed analytic light (lightFromRecord,
// native/src/scene-records.ts:100-115, and packLightsData,
// src/renderer/lightsData.ts:18-109).
void decode_light(const Words &W, size_t o, uint32_t version, float *out) {
  const uint32_t words = version >= 6 ? kLightWords : kLightWordsV5;
  const uint32_t type = W.u(o + L::Type);
  need(type <= 5, "unknown light type");
  W.finite(o + L::Color, 7, "light");
  need(W.f(o + L::Intensity) >= 0.0f, "light intensity must be nonnegative");
  if (W.u(o + L::Texture)) unsupported("textured rect lights");
  if (W.u(o + L::IesOffset) || W.u(o + L::IesBytes)) {
    need(type == 3, "only spot lights accept an IES profile");
    unsupported("IES profiles");
  }
  W.zero(o + L::Reserved, 1, "light");
  W.affine(o + L::Transform, "light matrix");
  float radius = 0, length = 0;
  bool normalize = false;
  if (words == kLightWords) {
    W.finite(o + L::Radius, 2, "light");
    need(W.u(o + L::Normalize) <= 1 && W.u(o + L::TreatAsLine) <= 1, "light flags must be 0 or 1");
    radius = W.f(o + L::Radius);
    length = W.f(o + L::Length);
    normalize = W.u(o + L::Normalize) != 0;
  } else if (type == 4 || type == 5) {
    radius = W.f(o + L::Size);
    length = type == 5 ? W.f(o + L::HeightOrInner) : 0.0f;
  }
  const size_t t = o + L::Transform;
  const Frame F = rigid_frame(W, t);
  const double size = W.f(o + L::Size), hOrIn = W.f(o + L::HeightOrInner), sOrOut = W.f(o + L::SpreadOrOuter);
  const double inten = W.f(o + L::Intensity);
  const double kPi = 3.141592653589793;
  std::fill(out, out + MTX_LIGHT_FLOATS, 0.0f);
  for (int k = 0; k < 3; k++) out[MTX_LIGHT_POSITION + k] = W.f(t + 12 + k);
  double scale = inten;
  auto setU = [&](const double *v, double s) { for (int k = 0; k < 3; k++) out[MTX_LIGHT_U + k] = (float)(v[k] * s); };
  auto setV = [&](const double *v, double s) { for (int k = 0; k < 3; k++) out[MTX_LIGHT_V + k] = (float)(v[k] * s); };
  switch (type) {
    case 0: {  // rect
      need(size > 0 && hOrIn > 0, "rect lights need positive width and height");
      out[MTX_LIGHT_TYPE] = MTX_LIGHT_RECT;
      setU(F.x, size);
      setV(F.y, -hOrIn);
      out[MTX_LIGHT_SPREAD] = (float)sOrOut;
      out[MTX_LIGHT_AREA] = (float)(size * hOrIn);
    } break;
    case 1: {  // sphere
      need(size >= 0, "sphere radius must be nonnegative");
      out[MTX_LIGHT_TYPE] = MTX_LIGHT_SPHERE;
      out[MTX_LIGHT_RADIUS] = (float)size;
      out[MTX_LIGHT_AREA] = (float)(4 * kPi * size * size);
    } break;
    case 2: {  // directional
      out[MTX_LIGHT_TYPE] = MTX_LIGHT_DIRECTIONAL;
      setU(F.dir, 1.0);
    } break;
    case 3: {  // spot
      need(hOrIn >= 0 && hOrIn <= sOrOut && sOrOut > 0 && sOrOut < 90,
           "spot angles need 0 <= inner <= outer < 90 and outer > 0");
      need(size >= 0, "spot radius must be nonnegative");
      out[MTX_LIGHT_TYPE] = MTX_LIGHT_SPOT;
      setU(F.dir, 1.0);
      out[MTX_LIGHT_V] = (float)std::cos(sOrOut * kPi / 180.0);
      out[MTX_LIGHT_V + 1] = (float)std::cos(std::min(hOrIn, sOrOut) * kPi / 180.0);
      out[MTX_LIGHT_RADIUS] = (float)size;
      out[MTX_LIGHT_AREA] = (float)(4 * kPi * size * size);
      if (size != 0.0) scale = inten / (kPi * size * size);
    } break;
    case 4: {  // disk
      need(radius > 0, "disk lights need a positive radius");
      const double area = kPi * (double)radius * radius;
      out[MTX_LIGHT_TYPE] = MTX_LIGHT_DISK;
      setU(F.x, radius);
      setV(F.y, -(double)radius);
      out[MTX_LIGHT_RADIUS] = radius;
      out[MTX_LIGHT_AREA] = (float)area;
      if (normalize) scale = inten / std::max(area, 1e-8);
    } break;
    default: {  // cylinder
      need(radius > 0 && length > 0, "cylinder lights need a positive radius and length");
      const double area = 2 * kPi * (double)radius * length;
      out[MTX_LIGHT_TYPE] = MTX_LIGHT_CYLINDER;
      setU(F.x, length);
      setV(F.y, radius);
      out[MTX_LIGHT_RADIUS] = radius;
      out[MTX_LIGHT_AREA] = (float)area;
      if (normalize) scale = inten / std::max(area, 1e-8);
    } break;
  }
  for (int k = 0; k < 3; k++) out[MTX_LIGHT_EMISSION + k] = (float)((double)W.f(o + L::Color + k) * scale);
}

// Camera record -> the browser camera block: projection inverse and
// camera-to-world (Camera.getProjectionMatrix, src/scene/Camera.ts:127-155;
// cameraApertureRadius, src/renderer/cameraLens.ts). The inverse is written in
// closed form in f64.
void decode_camera(const Words &W, size_t o, uint32_t width, uint32_t height, mtx_camera_desc &c) {
  const uint32_t proj = W.u(o + CAM::Projection);
  need(proj <= 1, "unknown camera projection");
  W.finite(o + CAM::Fov, 12, "camera");
  W.zero(o + CAM::Reserved, 3, "camera");
  W.affine(o + CAM::Transform, "camera matrix");
  const double fov = W.f(o + CAM::Fov), nearP = W.f(o + CAM::Near), farP = W.f(o + CAM::Far);
  const double zoom = std::max((double)W.f(o + CAM::Zoom), 2.220446049250313e-16);
  const double kPi = 3.141592653589793;
  std::memset(&c, 0, sizeof c);
  c.struct_size = sizeof(mtx_camera_desc);
  double inv[16] = {0};
  if (proj == 0) {
    need(fov > 0 && fov < 180 && nearP > 0 && farP > nearP, "perspective camera needs 0 < fov < 180, 0 < near < far");
    const double aspect = (double)width / (double)height;
    const double f = zoom / std::tan(fov * kPi / 180.0 / 2.0);
    const double nf = 1.0 / (nearP - farP);
    const double A = (farP + nearP) * nf, B = 2.0 * farP * nearP * nf;
    inv[0] = aspect / f;
    inv[5] = 1.0 / f;
    inv[11] = 1.0 / B;
    inv[14] = -1.0;
    inv[15] = A / B;
  } else {
    const double l0 = W.f(o + CAM::Left), r0 = W.f(o + CAM::Right), b0 = W.f(o + CAM::Bottom), t0 = W.f(o + CAM::Top);
    need(r0 != l0 && t0 != b0 && farP != nearP, "orthographic camera has an empty volume");
    const double l = (l0 + r0) / 2 - (r0 - l0) / (2 * zoom), r = (l0 + r0) / 2 + (r0 - l0) / (2 * zoom);
    const double b = (b0 + t0) / 2 - (t0 - b0) / (2 * zoom), t = (b0 + t0) / 2 + (t0 - b0) / (2 * zoom);
    const double lr = 1 / (l - r), bt = 1 / (b - t), nf = 1 / (nearP - farP);
    const double p0 = -2 * lr, p5 = -2 * bt, p10 = 2 * nf, p12 = (l + r) * lr, p13 = (t + b) * bt, p14 = (farP + nearP) * nf;
    inv[0] = 1 / p0;
    inv[5] = 1 / p5;
    inv[10] = 1 / p10;
    inv[12] = -p12 / p0;
    inv[13] = -p13 / p5;
    inv[14] = -p14 / p10;
    inv[15] = 1;
  }
  for (int k = 0; k < 16; k++) {
    c.projection_inverse[k] = (float)inv[k];
    c.view_inverse[k] = W.f(o + CAM::Transform + k);
  }
  double aperture = 0.0;
  const double fStop = W.f(o + CAM::FStop), mpu = W.f(o + CAM::MetersPerUnit);
  if (proj == 0 && fStop > 0 && mpu > 0) {
    const double focal = 0.5 * W.f(o + CAM::SensorHeight) * zoom / std::tan(fov * kPi / 360.0);
    const double r = focal / (2 * fStop) / 1000 / mpu;
    aperture = std::isfinite(r) && r > 0 ? r : 0.0;
  }
  c.aperture_radius = (float)aperture;
  c.focus_distance = W.f(o + CAM::FocusDistance);
}

// accumulateVertexNormals (src/util/vertexNormals.ts:9-50): area-weighted
// face normals summed per vertex in f32 storage, products and the cross
// product evaluated in f64 as JS does, then normalized.
std::vector<float> vertex_normals(const float *pos, const uint32_t *idx, size_t indexCount, uint32_t vertexCount) {
  std::vector<float> n((size_t)vertexCount * 3, 0.0f);
  for (size_t f = 0; f + 2 < indexCount; f += 3) {
    const uint32_t i[3] = {idx[f], idx[f + 1], idx[f + 2]};
    float v21[3], v32[3];
    for (int k = 0; k < 3; k++) {
      v21[k] = (float)((double)pos[i[0] * 3 + k] - (double)pos[i[1] * 3 + k]);
      v32[k] = (float)((double)pos[i[1] * 3 + k] - (double)pos[i[2] * 3 + k]);
    }
    const float c[3] = {(float)((double)v21[1] * v32[2] - (double)v21[2] * v32[1]),
                        (float)((double)v21[2] * v32[0] - (double)v21[0] * v32[2]),
                        (float)((double)v21[0] * v32[1] - (double)v21[1] * v32[0])};
    for (int v = 0; v < 3; v++)
      for (int k = 0; k < 3; k++) n[(size_t)i[v] * 3 + k] = (float)((double)n[(size_t)i[v] * 3 + k] + c[k]);
  }
  for (size_t v = 0; v < vertexCount; v++) {
    double x = n[v * 3], y = n[v * 3 + 1], z = n[v * 3 + 2];
    double len = x * x + y * y + z * z;
    if (len > 0) len = 1 / std::sqrt(len);
    n[v * 3] = (float)(x * len);
    n[v * 3 + 1] = (float)(y * len);
    n[v * 3 + 2] = (float)(z * len);
  }
  return n;
}

void mat4_mul(const double *a, const double *b, double *out) {  // column-major a*b
  double r[16];
  for (int c = 0; c < 4; c++)
    for (int rr = 0; rr < 4; rr++) {
      double s = 0;
      for (int k = 0; k < 4; k++) s += a[k * 4 + rr] * b[c * 4 + k];
      r[c * 4 + rr] = s;
    }
  std::memcpy(out, r, sizeof r);
}

// Whole-descriptor validation and decode (scene-format.md "Layout" .. "Limits").
Decoded decode(const uint8_t *bytes, size_t length) {
  need(bytes != nullptr, "descriptor is NULL");
  need(length % 4 == 0 && length >= kHeaderWords * 4, "descriptor length must be a multiple of 4 and hold the header");
  need(length / 4 <= 0xffffffffull, "descriptor exceeds 32-bit word offsets");
  std::vector<uint32_t> copy(length / 4);
  std::memcpy(copy.data(), bytes, length);
  Words W{copy.data(), copy.size()};
  Decoded D;
  D.version = W.u(H::Version);
  need(D.version >= 5 && D.version <= 8, "scene descriptor version must be 5..8");
  need(W.u(H::Length) == W.n, "total_words does not match the descriptor length");
  D.width = W.u(H::Width);
  D.height = W.u(H::Height);
  need(D.width >= 1 && D.width <= 8192 && D.height >= 1 && D.height <= 8192 &&
           (uint64_t)D.width * D.height <= 16777216ull,
       "output size must be 1..8192 per edge and at most 16,777,216 pixels");
  const uint32_t nGeo = W.u(H::GeometryCount), nMesh = W.u(H::MeshCount), nMat = W.u(H::MaterialCount),
                 nLight = W.u(H::LightCount), nTex = W.u(H::TextureCount);
  need(nGeo >= 1 && nMesh >= 1 && nMat >= 1, "a scene needs at least one geometry, mesh and material");
  need(nGeo <= 65536 && nMat <= 65536 && nTex <= 65536 && nLight <= 1024 && nMesh <= 16777216,
       "record count above its limit");
  const uint32_t texWords = D.version >= 7 ? kTextureWords : kTextureWordsV6;
  const uint32_t lightWords = D.version >= 6 ? kLightWords : kLightWordsV5;
  W.span(W.u(H::GeometryOffset), (size_t)nGeo * kGeometryWords, "geometry table");
  W.span(W.u(H::MeshOffset), (size_t)nMesh * kMeshWords, "mesh table");
  W.span(W.u(H::MaterialOffset), (size_t)nMat * kMaterialWords, "material table");
  W.span(W.u(H::LightOffset), (size_t)nLight * lightWords, "light table");
  W.span(W.u(H::TextureOffset), (size_t)nTex * texWords, "texture table");
  W.span(W.u(H::CameraOffset), kCameraWords, "camera");
  W.zero(H::Reserved0, 1, "header");
  W.zero(H::ReservedEnd, 4, "header");
  W.finite(H::Background, 4, "background color");
  const uint32_t quality = W.u(H::QualityFlags);
  need((quality & ~7u) == 0, "unknown quality flags");
  if (quality & 1u) unsupported("bevel");
  if (quality & 2u) unsupported("dispersion");
  if (quality & 4u) unsupported("a hidden background");
  if (W.u(H::GroundedBackground)) unsupported("grounded background projection");

  // Textures: the equirectangular environment, and the material textures
  // (copied below, once the materials name them).
  const size_t texOff = W.u(H::TextureOffset);
  for (uint32_t i = 0; i < nTex; i++) {
    const size_t o = texOff + (size_t)i * texWords;
    const uint32_t w = W.u(o + T::Width), h = W.u(o + T::Height), storage = W.u(o + T::Storage);
    need(w >= 1 && h >= 1 && w <= 16384 && h <= 16384, "texture size must be 1..16384 per edge");
    need(storage <= 1, "unknown texture storage");
    W.span(W.u(o + T::Pixels), (size_t)w * h * (storage ? 4 : 1), "texture pixels");
    if (storage) W.finite(W.u(o + T::Pixels), (size_t)w * h * 4, "texture pixels");
    need(W.u(o + T::FlipY) <= 1 && W.u(o + T::UvChannel) <= 1 && W.u(o + T::WrapS) <= 2 && W.u(o + T::WrapT) <= 2,
         "texture flag out of range");
    W.finite(o + T::UvScale, 4, "texture transform");
    if (texWords == kTextureWords) need(W.u(o + T::Udim) <= 1, "texture udim must be 0 or 1");
  }
  const uint32_t envT = W.u(H::EnvironmentTexture), bgT = W.u(H::BackgroundTexture);
  need(envT <= nTex && bgT <= nTex, "environment/background texture index out of range");
  if (bgT != envT) unsupported("a visible background independent of the environment lighting");
  // Without a map, camera rays that escape return transparent black; a
  // background color other than that is not rendered.
  if (!envT && (W.f(H::Background) != 0.0f || W.f(H::Background + 1) != 0.0f || W.f(H::Background + 2) != 0.0f ||
                W.f(H::Background + 3) != 0.0f))
    unsupported("a background color");
  double envScale = 1.0;
  if (const uint32_t es = W.u(H::EnvironmentSettings)) {
    W.span(es, 4, "environment settings");
    W.finite(es, 4, "environment settings");
    need(W.f(es) >= 0 && W.f(es + 2) >= 0, "environment intensities must be nonnegative");
    if (W.f(es + 1) != 0.0f || W.f(es + 3) != 0.0f) unsupported("environment yaw");
    if (W.f(es) != W.f(es + 2)) unsupported("separate background intensity");
    envScale = W.f(es);
  }
  if (envT) {
    const size_t o = texOff + (size_t)(envT - 1) * texWords;
    need(W.u(o + T::Storage) == 1, "environment maps must be RGBA32F");
    if (texWords == kTextureWords) need(W.u(o + T::Udim) == 0, "environment maps cannot be UDIM sets");
    D.envW = W.u(o + T::Width);
    D.envH = W.u(o + T::Height);
    const size_t px = W.u(o + T::Pixels), n = (size_t)D.envW * D.envH * 4;
    D.env.resize(n);
    for (size_t i = 0; i < n; i++) D.env[i] = envScale == 1.0 ? W.f(px + i) : (float)(W.f(px + i) * envScale);
  }
  if (const uint32_t ids = W.u(H::MeshIds)) {
    W.span(ids, nMesh, "mesh AOV ids");
    std::vector<uint32_t> v(copy.begin() + ids, copy.begin() + ids + nMesh);
    std::sort(v.begin(), v.end());
    need(std::adjacent_find(v.begin(), v.end()) == v.end() && v.back() != kNone, "mesh AOV ids must be unique and not UINT32_MAX");
  }
  if (const uint32_t ids = W.u(H::MaterialIds)) {
    W.span(ids, nMat, "material AOV ids");
    std::vector<uint32_t> v(copy.begin() + ids, copy.begin() + ids + nMat);
    std::sort(v.begin(), v.end());
    need(std::adjacent_find(v.begin(), v.end()) == v.end() && v.back() != kNone,
         "material AOV ids must be unique and not UINT32_MAX");
  }

  // Materials.
  D.materials.resize((size_t)nMat * kMatFloats);
  const size_t matOff = W.u(H::MaterialOffset);
  for (uint32_t i = 0; i < nMat; i++)
    decode_material(W, matOff + (size_t)i * kMaterialWords, D.version, nTex, &D.materials[(size_t)i * kMatFloats]);
  // Material textures: the descriptor textures that some slot references,
  // in first-reference order, copied with their sampling parameters.
  std::map<uint32_t, uint32_t> texOf;
  for (size_t i = 0; i < (size_t)nMat * MTX_MAP_COUNT; i++) {
    float &slot = D.materials[(i / MTX_MAP_COUNT) * kMatFloats + MTX_MAT_MAPS + i % MTX_MAP_COUNT];
    if (slot == 0.0f) continue;
    const uint32_t t = (uint32_t)slot - 1u;
    auto it = texOf.find(t);
    if (it == texOf.end()) {
      const size_t o = texOff + (size_t)t * texWords;
      if (texWords == kTextureWords && W.u(o + T::Udim)) unsupported("UDIM material textures");
      mtx_texture_desc d{};
      d.struct_size = sizeof d;
      d.width = W.u(o + T::Width);
      d.height = W.u(o + T::Height);
      d.storage = W.u(o + T::Storage);
      d.flip_y = W.u(o + T::FlipY);
      d.uv_channel = W.u(o + T::UvChannel);
      d.wrap_s = W.u(o + T::WrapS);
      d.wrap_t = W.u(o + T::WrapT);
      for (int k = 0; k < 2; k++) {
        d.uv_scale[k] = W.f(o + T::UvScale + k);
        d.uv_offset[k] = W.f(o + T::UvOffset + k);
      }
      const size_t px = W.u(o + T::Pixels), n = (size_t)d.width * d.height * (d.storage ? 4 : 1);
      D.texPixels.emplace_back(copy.begin() + px, copy.begin() + px + n);
      D.textures.push_back(d);
      it = texOf.emplace(t, (uint32_t)D.textures.size() - 1u).first;
    }
    slot = (float)(it->second + 1u);
  }

  // Geometry: validate, generate missing normals, split multi-material
  // geometry into one BLAS per (geometry, mesh-local material slot).
  struct Geo {
    uint32_t vc, ic;
    size_t pos, idx, uv = 0, uv1 = 0;
    std::vector<float> normals;
    std::vector<uint32_t> triSlot;  // per triangle: mesh-local material slot (empty: one slot)
    uint32_t slots = 1;
  };
  std::vector<Geo> geos(nGeo);
  const size_t geoOff = W.u(H::GeometryOffset);
  for (uint32_t g = 0; g < nGeo; g++) {
    const size_t o = geoOff + (size_t)g * kGeometryWords;
    Geo &Gg = geos[g];
    Gg.vc = W.u(o + G::VertexCount);
    Gg.ic = W.u(o + G::IndexCount);
    need(Gg.vc >= 1 && Gg.vc <= 4000000 && Gg.ic <= 12000000 && Gg.ic % 3 == 0,
         "geometry needs 1..4,000,000 vertices and a multiple of 3 indices up to 12,000,000");
    need(W.u(o + G::Position) && W.u(o + G::Indices), "geometry positions and indices are required");
    Gg.pos = W.u(o + G::Position);
    Gg.idx = W.u(o + G::Indices);
    W.span(Gg.pos, (size_t)Gg.vc * 3, "geometry positions");
    W.finite(Gg.pos, (size_t)Gg.vc * 3, "geometry positions");
    W.span(Gg.idx, Gg.ic, "geometry indices");
    for (uint32_t i = 0; i < Gg.ic; i++) need(W.u(Gg.idx + i) < Gg.vc, "geometry index out of range");
    W.zero(o + G::Reserved, 2, "geometry");
    const float *pos = reinterpret_cast<const float *>(copy.data() + Gg.pos);
    if (const uint32_t no = W.u(o + G::Normal)) {
      W.span(no, (size_t)Gg.vc * 3, "geometry normals");
      W.finite(no, (size_t)Gg.vc * 3, "geometry normals");
      Gg.normals.assign(reinterpret_cast<const float *>(copy.data() + no),
                        reinterpret_cast<const float *>(copy.data() + no) + (size_t)Gg.vc * 3);
    } else {
      Gg.normals = vertex_normals(pos, copy.data() + Gg.idx, Gg.ic, Gg.vc);
    }
    if (const uint32_t uo = W.u(o + G::Uv)) {
      W.span(uo, (size_t)Gg.vc * 2, "geometry uv");
      W.finite(uo, (size_t)Gg.vc * 2, "geometry uv");
      Gg.uv = uo;
    }
    if (const uint32_t uo = W.u(o + G::Uv1)) {
      W.span(uo, (size_t)Gg.vc * 2, "geometry uv1");
      W.finite(uo, (size_t)Gg.vc * 2, "geometry uv1");
      Gg.uv1 = uo;
    }
    if (const uint32_t to = W.u(o + G::Tangent)) { W.span(to, (size_t)Gg.vc * 4, "geometry tangents"); W.finite(to, (size_t)Gg.vc * 4, "geometry tangents"); }
    if (W.u(o + G::Color)) unsupported("vertex colors");
    if (const uint32_t mo = W.u(o + G::MaterialIndex)) {
      W.span(mo, Gg.vc, "geometry material indices");
      Gg.triSlot.resize(Gg.ic / 3);
      uint32_t mx = 0;
      for (uint32_t t = 0; t < Gg.ic / 3; t++) {
        const uint32_t a = W.u(mo + W.u(Gg.idx + 3 * t)), b = W.u(mo + W.u(Gg.idx + 3 * t + 1)),
                       c = W.u(mo + W.u(Gg.idx + 3 * t + 2));
        need(a == b && b == c, "all three vertices of a triangle must select the same material");
        Gg.triSlot[t] = a;
        mx = std::max(mx, a);
      }
      Gg.slots = mx + 1;
    }
  }

  // BLAS meshes: (geometry, slot) -> index, built lazily for used pairs.
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> blasOf;
  auto blas_for = [&](uint32_t g, uint32_t slot) -> uint32_t {
    auto key = std::make_pair(g, slot);
    auto it = blasOf.find(key);
    if (it != blasOf.end()) return it->second;
    const Geo &Gg = geos[g];
    std::vector<uint32_t> idx;
    std::vector<uint32_t> remap(Gg.vc, kNone);
    std::vector<float> pos, nrm, uv, uv1;
    const float *P = reinterpret_cast<const float *>(copy.data() + Gg.pos);
    const float *U = Gg.uv ? reinterpret_cast<const float *>(copy.data() + Gg.uv) : nullptr;
    const float *U1 = Gg.uv1 ? reinterpret_cast<const float *>(copy.data() + Gg.uv1) : nullptr;
    const bool whole = Gg.triSlot.empty();
    for (uint32_t t = 0; t < Gg.ic / 3; t++) {
      if (!whole && Gg.triSlot[t] != slot) continue;
      for (int k = 0; k < 3; k++) {
        const uint32_t v = W.u(Gg.idx + 3 * t + k);
        if (whole) { idx.push_back(v); continue; }
        if (remap[v] == kNone) {
          remap[v] = (uint32_t)(pos.size() / 3);
          pos.insert(pos.end(), P + (size_t)v * 3, P + (size_t)v * 3 + 3);
          nrm.insert(nrm.end(), Gg.normals.begin() + (size_t)v * 3, Gg.normals.begin() + (size_t)v * 3 + 3);
          if (U) uv.insert(uv.end(), U + (size_t)v * 2, U + (size_t)v * 2 + 2);
          if (U1) uv1.insert(uv1.end(), U1 + (size_t)v * 2, U1 + (size_t)v * 2 + 2);
        }
        idx.push_back(remap[v]);
      }
    }
    if (whole) {
      pos.assign(P, P + (size_t)Gg.vc * 3);
      nrm = Gg.normals;
      if (U) uv.assign(U, U + (size_t)Gg.vc * 2);
      if (U1) uv1.assign(U1, U1 + (size_t)Gg.vc * 2);
    }
    const uint32_t b = (uint32_t)D.positions.size();
    D.positions.push_back(std::move(pos));
    D.normals.push_back(std::move(nrm));
    D.uvs.push_back(std::move(uv));
    D.uvs1.push_back(std::move(uv1));
    D.indices.push_back(std::move(idx));
    blasOf[key] = b;
    return b;
  };

  // Meshes and batches -> instances.
  const size_t meshOff = W.u(H::MeshOffset), batchOff = W.u(H::BatchOffset);
  if (batchOff) W.span(batchOff, (size_t)nMesh * kBatchWords, "batch table");
  for (uint32_t mi = 0; mi < nMesh; mi++) {
    const size_t o = meshOff + (size_t)mi * kMeshWords;
    const uint32_t g = W.u(o + M::Geometry);
    need(g < nGeo, "mesh geometry index out of range");
    W.affine(o + M::Transform, "mesh matrix");
    const uint32_t flags = W.u(o + M::Flags);
    need((flags & ~1u) == 0, "mesh flags: only bit 0 is defined");
    const uint32_t tc = W.u(o + M::MaterialCount), to = W.u(o + M::MaterialTable);
    std::vector<uint32_t> table;
    if (tc) {
      W.span(to, tc, "mesh material table");
      for (uint32_t k = 0; k < tc; k++) {
        need(W.u(to + k) < nMat, "mesh material table entry out of range");
        table.push_back(W.u(to + k));
      }
    } else {
      need(W.u(o + M::Material) < nMat, "mesh material index out of range");
      table.push_back(W.u(o + M::Material));
    }
    const Geo &Gg = geos[g];
    if (!Gg.triSlot.empty()) need(Gg.slots <= table.size() || tc == 0, "geometry material index beyond the mesh table");
    // Instances of this mesh: an ordinary mesh, or its batch.
    std::vector<std::pair<uint32_t, std::vector<double>>> locals;  // (batch entry, world matrix)
    double world[16];
    for (int k = 0; k < 16; k++) world[k] = W.f(o + M::Transform + k);
    const float *colors = nullptr;
    if (batchOff) {
      const size_t bo = batchOff + (size_t)mi * kBatchWords;
      const uint32_t count = W.u(bo);
      W.zero(bo + 3, 1, "batch");
      if (count == kNone || count == 0) {
        need(W.u(bo + 1) == 0 && W.u(bo + 2) == 0, "ordinary or empty batch offsets must be 0");
        if (count == kNone) locals.push_back({kNone, std::vector<double>(world, world + 16)});
      } else {
        const uint32_t tfo = W.u(bo + 1), cfo = W.u(bo + 2);
        W.span(tfo, (size_t)count * 16, "batch transforms");
        if (cfo) {
          W.span(cfo, (size_t)count * 3, "batch colors");
          W.finite(cfo, (size_t)count * 3, "batch colors");
          colors = reinterpret_cast<const float *>(copy.data() + cfo);
        }
        for (uint32_t k = 0; k < count; k++) {
          W.affine(tfo + (size_t)k * 16, "batch matrix");
          double local[16], m[16];
          for (int q = 0; q < 16; q++) local[q] = W.f(tfo + (size_t)k * 16 + q);
          mat4_mul(world, local, m);
          locals.push_back({k, std::vector<double>(m, m + 16)});
        }
      }
    } else {
      locals.push_back({kNone, std::vector<double>(world, world + 16)});
    }
    if (flags & 1u) continue;  // hidden mesh or batch
    const uint32_t slots = Gg.triSlot.empty() ? 1u : Gg.slots;
    for (const auto &lm : locals) {
      for (uint32_t s = 0; s < slots; s++) {
        if (!Gg.triSlot.empty() && std::find(Gg.triSlot.begin(), Gg.triSlot.end(), s) == Gg.triSlot.end()) continue;
        const uint32_t mat = table[std::min<size_t>(s, table.size() - 1)];
        mtx_instance_desc I{};
        I.struct_size = sizeof I;
        I.blas_index = blas_for(g, Gg.triSlot.empty() ? 0u : s);
        I.material_word = mat << 8;
        I.double_sided = D.materials[(size_t)mat * kMatFloats + MTX_MAT_DOUBLE_SIDED] > 0.5f ? 1u : 0u;
        std::memcpy(I.transform, lm.second.data(), sizeof I.transform);
        for (int k = 0; k < 3; k++) I.color[k] = colors && lm.first != kNone ? colors[(size_t)lm.first * 3 + k] : 1.0f;
        D.instances.push_back(I);
        D.instanceSource.push_back({mi, lm.first});
      }
    }
  }
  need(D.instances.size() <= 16777216, "more than 16,777,216 instances");

  // Lights.
  D.lights.resize((size_t)nLight * MTX_LIGHT_FLOATS);
  for (uint32_t i = 0; i < nLight; i++)
    decode_light(W, W.u(H::LightOffset) + (size_t)i * lightWords, D.version, &D.lights[(size_t)i * MTX_LIGHT_FLOATS]);

  D.cameraOffset = W.u(H::CameraOffset);
  decode_camera(W, D.cameraOffset, D.width, D.height, D.camera);
  return D;
}

// ------------------------------------------------------------- contexts
constexpr uint32_t kMaxTotalSamples = 16384;

struct Job {
  uint64_t id = 0;
  uint32_t state = MTX_JOB_QUEUED;
  uint32_t requested = 0, submitted = 0, completed = 0, base = 0;
  bool cancelled = false;
  uint32_t flight = 0;  // samples of the batch in flight (0: none)
  bool active() const {
    return state == MTX_JOB_QUEUED || state == MTX_JOB_RUNNING || state == MTX_JOB_CANCELLING;
  }
};

struct Context {
  std::mutex busy;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  mtx_handle scene = 0, target = 0;
  mtx_render_settings settings{};
  std::vector<uint32_t> words;  // last committed descriptor
  Decoded decoded;
  uint32_t samples = 0, lastCommit = MTX_COMMIT_NONE;
  bool failed = false, injectFailure = false;
  Job job;
  bool hasJob = false;
  ~Context() {
    if (target) mtx_target_destroy(target);
    if (scene) mtx_scene_destroy(scene);
    if (done) (void)hipEventDestroy(done);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

std::mutex gMutex;
std::unordered_map<uint64_t, std::shared_ptr<Context>> gContexts;
std::unordered_map<uint64_t, uint64_t> gJobs;  // job -> context
uint64_t gNextHandle = (1ull << 48);           // context/job handles: monotonic, disjoint from other objects

uint64_t next_handle() {
  std::lock_guard<std::mutex> g(gMutex);
  return gNextHandle++;
}

std::shared_ptr<Context> find_context(uint64_t h) {
  std::lock_guard<std::mutex> g(gMutex);
  auto it = gContexts.find(h);
  if (it == gContexts.end()) fail(MTX_INVALID_HANDLE, "unknown context handle");
  return it->second;
}

std::shared_ptr<Context> find_job_context(uint64_t job) {
  std::lock_guard<std::mutex> g(gMutex);
  auto it = gJobs.find(job);
  if (it == gJobs.end()) fail(MTX_INVALID_HANDLE, "unknown or expired job handle");
  auto c = gContexts.find(it->second);
  if (c == gContexts.end()) fail(MTX_INVALID_HANDLE, "unknown or expired job handle");
  return c->second;
}

void expire_job(Context &C) {
  if (!C.hasJob) return;
  std::lock_guard<std::mutex> g(gMutex);
  gJobs.erase(C.job.id);
  C.hasJob = false;
}

// A GPU-side operation: a failure (or an injected one) marks the context failed.
template <typename F>
void gpu_op(Context &C, F &&f) {
  if (C.injectFailure) {
    C.injectFailure = false;
    C.failed = true;
    fail(MTX_CONTEXT_FAILED, "injected GPU failure");
  }
  try {
    f();
  } catch (const Fail &e) {
    if (e.code == MTX_CONTEXT_FAILED) C.failed = true;
    throw;
  }
}

void ready(const Context &C) {
  if (C.failed) fail(MTX_INVALID_STATE, "Scene failed; commit a new scene before rendering or reading");
  if (!C.scene) fail(MTX_INVALID_STATE, "No committed scene");
}
void idle(const Context &C) {
  if (!C.failed && C.hasJob && C.job.active())
    fail(MTX_INVALID_STATE, "Context has an active render job; cancel and poll it first");
}

mtx_target_settings target_settings(const Decoded &D, const mtx_render_settings &s) {
  mtx_target_settings t{};
  t.struct_size = sizeof t;
  t.width = D.width;
  t.height = D.height;
  t.bounces = std::min(std::max(s.bounces, 1u), 20u);
  t.direct_light_candidates = std::min(std::max(s.direct_light_candidates, 1u), 16u);
  t.sampling_pattern = s.sampling_pattern;
  t.sequence_seed = s.seed;
  t.max_samples = s.max_samples;
  t.sample_sequence_offset = s.sample_sequence_offset;
  t.clamping = (float)s.clamping;
  t.kernel = MTX_KERNEL_WAVEFRONT;
  t.max_paths_in_flight = 0;
  return t;
}

void check_settings(const mtx_render_settings &s) {
  need(s.struct_size == sizeof(mtx_render_settings), "mtx_render_settings.struct_size must be sizeof(mtx_render_settings)");
  need(s.sampling_pattern <= 2, "sampling_pattern must be 0 (Sobol), 1 (ZSobol) or 2 (random)");
  need(std::isfinite(s.clamping) && s.clamping >= 0.0, "clamping must be finite and nonnegative");
  need(std::isfinite(s.transmissive_output_alpha_min), "transmissive_output_alpha_min must be finite");
  need(std::isfinite(s.photon_map_radius) && s.photon_map_radius >= 0.0f, "photon_map_radius must be finite and nonnegative");
  if (s.caustic_flags) unsupported("caustic estimators");
  if (s.light_photons || s.photon_map_photons) unsupported("photon counts");
  if (s.transmissive_output_alpha_min != 1.0) unsupported("transmissive output alpha below 1");
}

mtx_scene_desc scene_desc(const Decoded &D, std::vector<mtx_mesh_desc> &meshes,
                          std::vector<mtx_texture_desc> &textures) {
  meshes.resize(D.positions.size());
  for (size_t i = 0; i < meshes.size(); i++) {
    mtx_mesh_desc &m = meshes[i];
    m = mtx_mesh_desc{};
    m.struct_size = sizeof m;
    m.vertex_count = (uint32_t)(D.positions[i].size() / 3);
    m.triangle_count = (uint32_t)(D.indices[i].size() / 3);
    m.positions = D.positions[i].data();
    m.normals = D.normals[i].data();
    m.indices = D.indices[i].data();
    m.uvs = D.uvs[i].empty() ? nullptr : D.uvs[i].data();
    m.uvs1 = D.uvs1[i].empty() ? nullptr : D.uvs1[i].data();
  }
  textures = D.textures;
  for (size_t i = 0; i < textures.size(); i++) textures[i].pixels = D.texPixels[i].data();
  mtx_scene_desc s{};
  s.struct_size = sizeof s;
  s.mesh_count = (uint32_t)meshes.size();
  s.meshes = meshes.data();
  s.instance_count = (uint32_t)D.instances.size();
  s.instances = D.instances.data();
  s.material_count = (uint32_t)(D.materials.size() / kMatFloats);
  s.materials = D.materials.data();
  s.light_count = (uint32_t)(D.lights.size() / MTX_LIGHT_FLOATS);
  s.lights = D.lights.data();
  s.env_width = D.envW;
  s.env_height = D.envH;
  s.env_rgba = D.env.empty() ? nullptr : D.env.data();
  s.blas_bins = 8;   // native TLAS/BLAS defaults (bvhAccel.ts:385-454)
  s.tlas_bins = 64;
  s.flags = MTX_ACCEL_PER_INSTANCE_BOUNDS;
  s.texture_count = (uint32_t)textures.size();
  s.textures = textures.empty() ? nullptr : textures.data();
  return s;
}

// Waits for the in-flight batch (if any) and credits it.
void drain(Context &C) {
  if (C.hasJob && C.job.flight) {
    gpu_op(C, [&] {
      const hipError_t e = hipEventSynchronize(C.done);
      if (e != hipSuccess) fail(MTX_CONTEXT_FAILED, std::string("batch failed: ") + hipGetErrorString(e));
    });
    C.samples = C.job.base + C.job.submitted;
    C.job.completed = C.job.submitted;
    C.job.flight = 0;
  }
}

// Differences between two descriptors that an incremental commit can apply:
// only camera words, or only mesh matrices (and the camera). Returns 0 if
// none apply (full commit), 1 for camera-only, 2 for transforms.
int incremental_kind(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b, const Decoded &old) {
  if (a.size() != b.size() || a.empty()) return 0;
  const size_t cam = old.cameraOffset;
  const size_t meshOff = a[H::MeshOffset], nMesh = a[H::MeshCount];
  bool other = false, transforms = false;
  for (size_t i = 0; i < a.size() && !other; i++) {
    if (a[i] == b[i]) continue;
    if (i >= cam && i < cam + kCameraWords) continue;
    if (i >= meshOff && i < meshOff + nMesh * kMeshWords) {
      const size_t w = (i - meshOff) % kMeshWords;
      if (w >= M::Transform && w < M::Transform + 16) { transforms = true; continue; }
    }
    other = true;
  }
  if (other) return 0;
  return transforms ? 2 : 1;
}

void commit(Context &C, const uint8_t *bytes, size_t length, const mtx_render_settings *settings) {
  mtx_render_settings s{};
  if (settings) {
    need(settings->struct_size == sizeof(mtx_render_settings),
         "mtx_render_settings.struct_size must be sizeof(mtx_render_settings)");
    s = *settings;
    check_settings(s);
  } else {
    mtx_render_settings_default(&s);
  }
  Decoded D = decode(bytes, length);
  std::vector<uint32_t> words(length / 4);
  std::memcpy(words.data(), bytes, length);
  // A successful publication drains the old job's in-flight batch first.
  drain(C);
  expire_job(C);
  const bool settingsSame = std::memcmp(&s, &C.settings, sizeof s) == 0;
  const int kind = (C.scene && !C.failed) ? incremental_kind(C.words, words, C.decoded) : 0;
  if (kind) {
    // Incremental: keep geometry and acceleration structures.
    if (kind == 2) {
      std::vector<double> xf(D.instances.size() * 16);
      for (size_t i = 0; i < D.instances.size(); i++) std::memcpy(&xf[i * 16], D.instances[i].transform, 128);
      call([&](char *e, size_t n) {
        return mtx_scene_set_transforms(C.scene, xf.data(), (uint32_t)D.instances.size(), nullptr, 0, e, n);
      });
    }
    if (settingsSame) {
      call([&](char *, size_t) { return mtx_target_set_camera(C.target, &D.camera); });
      call([&](char *, size_t) { return mtx_target_reset(C.target); });
    } else {
      mtx_target_settings ts = target_settings(D, s);
      mtx_handle t = 0;
      call([&](char *e, size_t n) { return mtx_target_create(C.scene, &ts, &D.camera, &t, e, n); });
      mtx_target_destroy(C.target);
      C.target = t;
    }
    C.lastCommit = MTX_COMMIT_INCREMENTAL;
  } else {
    std::vector<mtx_mesh_desc> meshes;
    std::vector<mtx_texture_desc> textures;
    mtx_scene_desc sd = scene_desc(D, meshes, textures);
    mtx_handle scene = 0, target = 0;
    call([&](char *e, size_t n) { return mtx_scene_create(&sd, &scene, e, n); });
    mtx_target_settings ts = target_settings(D, s);
    int32_t rc;
    char msg[512] = {0};
    rc = mtx_target_create(scene, &ts, &D.camera, &target, msg, sizeof msg);
    if (rc != MTX_OK) {
      mtx_scene_destroy(scene);
      fail(rc, msg);
    }
    if (C.target) mtx_target_destroy(C.target);
    if (C.scene) mtx_scene_destroy(C.scene);
    C.scene = scene;
    C.target = target;
    C.lastCommit = MTX_COMMIT_FULL;
  }
  C.settings = s;
  C.words.swap(words);
  C.decoded = std::move(D);
  C.samples = 0;
  C.failed = false;
}

// Launches `n` samples on the context's stream and records the batch event.
void launch(Context &C, uint32_t n) {
  gpu_op(C, [&] {
    call([&](char *, size_t) { return mtx_target_render(C.target, n, C.stream); });
    const hipError_t e = hipEventRecord(C.done, C.stream);
    if (e != hipSuccess) fail(MTX_CONTEXT_FAILED, std::string("hipEventRecord: ") + hipGetErrorString(e));
  });
}

void poll(Context &C, uint32_t budget) {
  Job &J = C.job;
  if (!J.active()) return;
  if (J.flight) {
    hipError_t q = hipSuccess;
    gpu_op(C, [&] {
      q = hipEventQuery(C.done);
      if (q != hipSuccess && q != hipErrorNotReady) {
        (void)hipGetLastError();
        fail(MTX_CONTEXT_FAILED, std::string("batch failed: ") + hipGetErrorString(q));
      }
    });
    if (q == hipErrorNotReady) {
      (void)hipGetLastError();
      return;
    }
    J.completed = J.submitted;
    C.samples = J.base + J.completed;
    J.flight = 0;
  }
  if (J.completed == J.requested) { J.state = MTX_JOB_COMPLETE; return; }
  if (J.cancelled) { J.state = MTX_JOB_CANCELLED; return; }
  if (budget == 0) return;
  const uint32_t count = std::min(budget, J.requested - J.submitted);
  launch(C, count);
  J.submitted += count;
  J.flight = count;
  J.state = MTX_JOB_RUNNING;
}

template <typename F>
int32_t guarded(char *err, size_t cap, F &&f) {
  try {
    f();
    return MTX_OK;
  } catch (const Fail &e) {
    write_error(err, cap, e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    write_error(err, cap, "out of host memory");
    return MTX_ERROR;
  } catch (const std::exception &e) {
    write_error(err, cap, e.what());
    return MTX_ERROR;
  }
}

// Runs f with the context's operation lock; an overlapping call gets BUSY.
template <typename F>
int32_t with_context(uint64_t h, char *err, size_t cap, F &&f) {
  return guarded(err, cap, [&] {
    auto C = find_context(h);
    std::unique_lock<std::mutex> lk(C->busy, std::try_to_lock);
    if (!lk.owns_lock()) fail(MTX_BUSY, "another operation on this context is running");
    f(*C);
  });
}

}  // namespace

extern "C" {

int32_t mtx_render_settings_default(mtx_render_settings *s) {
  if (!s) return MTX_INVALID_ARGUMENT;
  std::memset(s, 0, sizeof *s);
  s->struct_size = sizeof *s;
  s->sampling_pattern = 0;
  s->seed = 12345;
  s->max_samples = 64;
  s->bounces = 6;
  s->direct_light_candidates = 2;
  s->bevel_samples = 8;
  s->sample_sequence_offset = 0;
  s->clamping = 0.0;
  s->transmissive_output_alpha_min = 1.0;
  return MTX_OK;
}

int32_t mtx_context_create(mtx_handle *out, char *err, size_t cap) {
  return guarded(err, cap, [&] {
    need(out != nullptr, "Null context output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
      (void)hipGetLastError();
      fail(MTX_UNAVAILABLE, "no HIP device");
    }
    auto C = std::make_shared<Context>();
    if (hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&C->done, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      fail(MTX_UNAVAILABLE, "could not create the context's HIP stream");
    }
    mtx_render_settings_default(&C->settings);
    const uint64_t h = next_handle();
    std::lock_guard<std::mutex> g(gMutex);
    gContexts[h] = C;
    *out = h;
  });
}

int32_t mtx_context_destroy(mtx_handle h, char *err, size_t cap) {
  return guarded(err, cap, [&] {
    auto C = find_context(h);
    std::unique_lock<std::mutex> lk(C->busy, std::try_to_lock);
    if (!lk.owns_lock()) fail(MTX_BUSY, "another operation on this context is running");
    idle(*C);
    if (C->hasJob && C->job.flight) (void)hipEventSynchronize(C->done);
    expire_job(*C);
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    std::lock_guard<std::mutex> g(gMutex);
    gContexts.erase(h);
  });
}

int32_t mtx_context_commit(mtx_handle h, const uint8_t *input, size_t length, char *err, size_t cap) {
  return with_context(h, err, cap, [&](Context &C) { commit(C, input, length, nullptr); });
}

int32_t mtx_context_commit_with_settings(mtx_handle h, const uint8_t *input, size_t length,
                                         const mtx_render_settings *settings, char *err, size_t cap) {
  return with_context(h, err, cap, [&](Context &C) {
    need(settings != nullptr, "Null render settings");
    commit(C, input, length, settings);
  });
}

int32_t mtx_context_render(mtx_handle h, uint32_t additional, char *err, size_t cap) {
  return with_context(h, err, cap, [&](Context &C) {
    ready(C);
    idle(C);
    need(additional >= 1 && additional <= kMaxTotalSamples - C.samples, "Total samples must be 1..16384");
    expire_job(C);
    gpu_op(C, [&] {
      call([&](char *, size_t) { return mtx_target_render(C.target, additional, C.stream); });
      const hipError_t e = hipStreamSynchronize(C.stream);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(MTX_CONTEXT_FAILED, std::string("render failed: ") + hipGetErrorString(e));
      }
    });
    C.samples += additional;
  });
}

int32_t mtx_context_reset(mtx_handle h, char *err, size_t cap) {
  return with_context(h, err, cap, [&](Context &C) {
    ready(C);
    idle(C);
    gpu_op(C, [&] {
      call([&](char *, size_t) { return mtx_target_reset(C.target); });
      call([&](char *, size_t) { return mtx_target_synchronize(C.target); });
    });
    expire_job(C);
    C.samples = 0;
  });
}

int32_t mtx_context_wait(mtx_handle h, uint32_t timeout_ms, char *err, size_t cap) {
  return with_context(h, err, cap, [&](Context &C) {
    ready(C);
    idle(C);
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
      const hipError_t q = hipStreamQuery(C.stream);
      if (q == hipSuccess) return;
      if (q == hipErrorNotReady) (void)hipGetLastError();
      if (q != hipErrorNotReady) {
        (void)hipGetLastError();
        C.failed = true;
        fail(MTX_CONTEXT_FAILED, std::string("GPU work failed: ") + hipGetErrorString(q));
      }
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (timeout_ms != MTX_WAIT_INDEFINITELY && ms >= (double)timeout_ms) fail(MTX_TIMEOUT, "wait timed out");
      std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
  });
}

int32_t mtx_context_info(mtx_handle h, mtx_context_snapshot *info, char *err, size_t cap) {
  return with_context(h, err, cap, [&](Context &C) {
    need(info != nullptr && info->struct_size == sizeof(mtx_context_snapshot), "Context info size mismatch");
    info->width = C.scene ? C.decoded.width : 0;
    info->height = C.scene ? C.decoded.height : 0;
    info->samples = C.samples;
    info->state = C.failed ? MTX_CONTEXT_FAILED_STATE
                  : !C.scene ? MTX_CONTEXT_EMPTY
                  : (C.hasJob && C.job.active()) ? MTX_CONTEXT_JOB_ACTIVE
                                                 : MTX_CONTEXT_COMMITTED;
    info->last_commit = C.lastCommit;
  });
}

int32_t mtx_context_read_pixels(mtx_handle h, uint32_t kind, void *pixels, size_t size, size_t *written, char *err,
                                size_t cap) {
  return with_context(h, err, cap, [&](Context &C) {
    need(written != nullptr, "Null pixel length output");
    *written = 0;
    ready(C);
    idle(C);
    need(kind <= MTX_IMAGE_ALBEDO, "Unknown image kind");
    if (C.samples == 0) fail(MTX_INVALID_STATE, "No completed samples");
    const size_t N = (size_t)C.decoded.width * C.decoded.height, required = N * 16;
    *written = required;
    if (!pixels && size == 0) return;
    need(pixels != nullptr, "Null pixel destination");
    if (size < required) fail(MTX_BUFFER_TOO_SMALL, "Pixel destination is too small");
    std::vector<double> sums(N * 13);
    gpu_op(C, [&] {
      call([&](char *, size_t) { return mtx_target_read_accumulators(C.target, sums.data(), sums.size(), nullptr, 0); });
    });
    // Accumulator planes: color rgba 0..3, albedo rgb + hit 4..7, normal rgb + 1 8..11.
    const int base = kind == MTX_IMAGE_HDR ? 0 : kind == MTX_IMAGE_ALBEDO ? 4 : 8;
    std::vector<float> out(N * 4);
    const double s = (double)C.samples;
    for (size_t i = 0; i < N; i++)
      for (int k = 0; k < 4; k++) {
        const float v = (float)(sums[i * 13 + base + k] / s);
        if (!std::isfinite(v)) {
          C.failed = true;
          fail(MTX_CONTEXT_FAILED, "Non-finite native output");
        }
        out[i * 4 + k] = v;
      }
    std::memcpy(pixels, out.data(), required);
  });
}

int32_t mtx_context_submit(mtx_handle h, uint32_t additional, mtx_handle *job, char *err, size_t cap) {
  return with_context(h, err, cap, [&](Context &C) {
    need(job != nullptr, "Null render job output");
    ready(C);
    idle(C);
    need(additional >= 1 && additional <= kMaxTotalSamples - C.samples, "Total samples must be 1..16384");
    expire_job(C);
    C.job = Job{};
    C.job.id = next_handle();
    C.job.base = C.samples;
    C.job.requested = additional;
    C.hasJob = true;
    {
      std::lock_guard<std::mutex> g(gMutex);
      gJobs[C.job.id] = h;
    }
    *job = C.job.id;
  });
}

int32_t mtx_job_poll(mtx_handle job, uint32_t budget, char *err, size_t cap) {
  return guarded(err, cap, [&] {
    need(budget <= 64, "Poll dispatch budget must be 0..64 samples");
    auto C = find_job_context(job);
    std::unique_lock<std::mutex> lk(C->busy, std::try_to_lock);
    if (!lk.owns_lock()) fail(MTX_BUSY, "another operation on this context is running");
    if (!C->hasJob || C->job.id != job) fail(MTX_INVALID_HANDLE, "unknown or expired job handle");
    if (C->failed) fail(MTX_INVALID_STATE, "Render scene failed; commit a new scene");
    try {
      poll(*C, budget);
    } catch (const Fail &) {
      C->job.state = MTX_JOB_FAILED;
      C->job.flight = 0;
      throw;
    }
  });
}

int32_t mtx_job_info(mtx_handle job, mtx_job_snapshot *info, char *err, size_t cap) {
  return guarded(err, cap, [&] {
    auto C = find_job_context(job);
    std::unique_lock<std::mutex> lk(C->busy, std::try_to_lock);
    if (!lk.owns_lock()) fail(MTX_BUSY, "another operation on this context is running");
    if (!C->hasJob || C->job.id != job) fail(MTX_INVALID_HANDLE, "unknown or expired job handle");
    need(info != nullptr && info->struct_size == sizeof(mtx_job_snapshot), "Job info size mismatch");
    info->state = C->failed ? MTX_JOB_FAILED : C->job.state;
    info->requested_samples = C->job.requested;
    info->submitted_samples = C->job.submitted;
    info->completed_samples = C->job.completed;
    info->base_samples = C->job.base;
  });
}

int32_t mtx_job_cancel(mtx_handle job, char *err, size_t cap) {
  return guarded(err, cap, [&] {
    auto C = find_job_context(job);
    std::unique_lock<std::mutex> lk(C->busy, std::try_to_lock);
    if (!lk.owns_lock()) fail(MTX_BUSY, "another operation on this context is running");
    if (!C->hasJob || C->job.id != job) fail(MTX_INVALID_HANDLE, "unknown or expired job handle");
    Job &J = C->job;
    if (!J.active()) return;
    J.cancelled = true;
    J.state = J.flight ? MTX_JOB_CANCELLING : MTX_JOB_CANCELLED;
  });
}

int32_t mtx_descriptor_validate(const uint8_t *input, size_t length, uint32_t info[8], char *err, size_t cap) {
  return guarded(err, cap, [&] {
    const Decoded D = decode(input, length);
    if (info) {
      info[0] = D.width;
      info[1] = D.height;
      info[2] = (uint32_t)D.positions.size();
      info[3] = (uint32_t)D.instances.size();
      info[4] = (uint32_t)(D.materials.size() / kMatFloats);
      info[5] = (uint32_t)(D.lights.size() / MTX_LIGHT_FLOATS);
      info[6] = D.envW;
      info[7] = D.envH;
    }
  });
}

int32_t mtx_context_inject_failure(mtx_handle h) {
  return with_context(h, nullptr, 0, [&](Context &C) { C.injectFailure = true; });
}

}  // extern "C"
