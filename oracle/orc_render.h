// Oracle: RGB megakernel integrator (test infrastructure only).
// Restates src/renderer/wgsl/rayTrace.compute.ts (integrator :891-946,
// shadeVertex :343-512, completeVertex :514-589, main :1157-1263),
// sampleMaterial.wgsl.ts, sampleLight.wgsl.ts, envMap.wgsl.ts,
// openPbrBsdf.wgsl.ts (lobes used by the configurations: coat, metal with
// F82 tint and multiple scattering, dielectric specular, solid transmission
// with dielectric multiple scattering, EON diffuse) and the host tables of
// lightSampling.ts / envMapDistribution.ts.
#pragma once
#include <cstdint>
#include <vector>

#include "orc_trace.h"

namespace orc {

// Packed OpenPBR material parameters (M_FLOATS floats), shared input format.
enum MatField {
  M_BASE_WEIGHT = 0, M_BASE_COLOR = 1, M_DIFFUSE_ROUGHNESS = 4, M_METALNESS = 5,
  M_SPEC_WEIGHT = 6, M_SPEC_COLOR = 7, M_SPEC_ROUGHNESS = 10, M_SPEC_IOR = 11,
  M_TRANS_WEIGHT = 12, M_TRANS_COLOR = 13, M_COAT_WEIGHT = 16, M_COAT_COLOR = 17,
  M_COAT_ROUGHNESS = 20, M_COAT_IOR = 21, M_COAT_DARKENING = 22, M_EMISSION = 23,
  M_DOUBLE_SIDED = 26, M_SPEC_ANISOTROPY = 27, M_FUZZ_WEIGHT = 28, M_FUZZ_COLOR = 29, M_FUZZ_ROUGHNESS = 32,
  M_SSS_WEIGHT = 33, M_SSS_COLOR = 34, M_THIN_WALLED = 37, M_OPACITY = 38, M_ALPHA_MODE = 39, M_ALPHA_CUTOFF = 40,
  M_MAP_ALPHA_USAGE = 41, M_NORMAL_SCALE = 42, M_UV_SCALE = 44, M_MAPS = 46, M_FLOATS = 64
};
// Texture slots (texture index + 1 at M_MAPS + slot; 0 = none).
enum MapSlot {
  MAP_BASE = 0, MAP_ALPHA, MAP_TRANSMISSION, MAP_METALLIC_ROUGHNESS, MAP_ROUGHNESS, MAP_METALNESS, MAP_COAT_WEIGHT,
  MAP_COAT_ROUGHNESS, MAP_FUZZ_COLOR, MAP_FUZZ_ROUGHNESS, MAP_SPECULAR_WEIGHT, MAP_SPECULAR_COLOR, MAP_EMISSIVE,
  MAP_NORMAL, MAP_COUNT
};
// Material textures and vertex UVs of a scene. Texel rows are in sampling
// order (flip_y already applied); uvs / uvs1 hold xy per global vertex (zeros
// where a mesh has none); emissiveUv holds 12 floats per emissive triangle:
// uv0, uv1, uv2 and the corners' UV1 (UV0 where the mesh has no UV1).
struct Texture {
  uint32_t width, height, storage, wrapS, wrapT, uvChannel;
  float sx, sy, ox, oy;
  std::vector<uint32_t> rgba8;
  std::vector<float> rgba32;
};
struct TextureSet {
  std::vector<Texture> tex;
  std::vector<float> uvs, uvs1, emissiveUv;
};
// Packed light (16 floats): type, position, u, v, emission, spread or radius,
// area (the reference's Light record, rayTraceCore.wgsl.ts:228-248). Types are
// lightConstants.ts:10-15; 0 is read as a rect.
enum LightField { L_TYPE = 0, L_POS = 1, L_U = 4, L_V = 7, L_EMISSION = 10, L_SPREAD = 13, L_RADIUS = 13, L_AREA = 14 };
enum LightType { LT_DIRECTIONAL = 1, LT_SPHERE = 2, LT_RECT = 3, LT_SPOT = 4, LT_DISK = 5, LT_CYLINDER = 6 };

struct RenderScene {
  const TraceScene *trace;
  const float *normals;       // per global vertex xyz
  const float *materials;     // 32 per material
  uint32_t materialCount;
  const float *lights;        // 16 per light
  uint32_t lightCount;
  const float *env;           // RGBA f32 env map (nullptr: none)
  uint32_t envW, envH;
  const float *camera;        // viewProjInv[16], viewInv[16], aperture, focus
  const float *emissive;      // 12 per emissive triangle: p0, p1, p2 (world), area, material, unused
  const double *emissiveWeight;  // alias weight (area * emitted luminance) per emissive triangle
  uint32_t emissiveCount;
  uint32_t width, height;
  uint32_t bounces, candidates, pattern, seed, logSPP, sampleOffset;
  float clamping;
  const TextureSet *textures = nullptr;  // null: no material samples a texture
};

struct Accum {
  double color[4], albedo[4], normal[4], lum2;
};

// Host light tables: Walker alias table over {environment, rect lights}
// (lightSampling.ts:172-212, power estimates :72-107, :415-431) and the
// environment CDF (envMapDistribution.ts:57-134).
struct HostTables {
  // alias table entries: probability, alias, type (0 env, 1 analytic), index
  std::vector<float> prob;
  std::vector<int32_t> alias, type, index;
  std::vector<float> lightPmf;  // per analytic light
  std::vector<float> emissivePmf;  // per emissive triangle
  float envPmf = 0.0f;
  double totalWeight = 0.0;
  std::vector<float> cdf;       // (W+2) x (H+1) x 2
  uint32_t cdfW = 0, cdfH = 0;
};
void build_tables(const RenderScene &s, HostTables &T);
// collectEmissiveTriangles (lightSampling.ts:237-320) over instances in order:
// writes up to `capacity` records of 12 floats and their alias weights (and,
// when `corners` is given, the global vertex index of each corner), returns
// the number of emissive triangles.
uint32_t emissive_triangles(const float *positions, const uint32_t *vertexOffsets, const uint32_t *indices,
                            const uint32_t *indexOffsets, const uint32_t *triangleCounts, const uint32_t *instanceMesh,
                            const double *transforms, const uint32_t *instanceMaterial, uint32_t instanceCount,
                            const float *materials, float *out, double *weight, uint32_t capacity,
                            uint32_t *corners = nullptr);
// Primary ray of (pixel, sample index): origin xyz, direction xyz.
void camera_sample(const float *camera, uint32_t width, uint32_t height, uint32_t px, uint32_t py, uint32_t index,
                   uint32_t pattern, uint32_t seed, uint32_t logSPP, float out[6]);
void build_env_cdf(const RenderScene &s, HostTables &T);

void set_energy_lut(const float *lut);  // 32 x 2177 floats
// sampleMap of texture `index` at (u, v): filtered RGBA and its decoded color.
void probe_sample_map(const TextureSet &T, uint32_t index, float u, float v, float rgba[4], float decoded[3]);
// Fitted directional albedo of the fuzz lobe (openPbrBsdf.wgsl.ts:595-607).
float fuzz_albedo(float viewCosine, float roughness);
// Renders samples [sampleStart, sampleStart+count) for the pixel rectangle
// and adds them to out (one Accum per pixel of the rectangle, row-major).
// Returns the number of rays traced (camera + continuation + shadow).
uint64_t render(const RenderScene &s, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                uint32_t sampleStart, uint32_t count, Accum *out);
// Integrates one path and records the tag of each random-pattern draw in
// log (up to cap); returns the number of draws.
uint32_t path_draws(const RenderScene &s, uint32_t px, uint32_t py, uint32_t sampleIndex, uint32_t *log,
                    uint32_t cap);

// BSDF probes: material m (M_FLOATS floats) on the outside of a surface with
// normal +z, seen along the unit vector `view` (view.z > 0).
// probe_bsdf_eval: f = BSDF * |cos| (RGB) and pdf of n world directions.
void probe_bsdf_eval(const float *m, const float *view, uint32_t n, const float *dirs, float *f, float *pdf);
// probe_bsdf_sample: directions for n triples (rx, ry, lobe) of [0,1) numbers;
// ok[i] = 0 where the sampler rejects the draw.
void probe_bsdf_sample(const float *m, const float *view, uint32_t n, const float *u, float *dirs, uint8_t *ok);
// Lobe selection probabilities (fuzz, coat, metal, specular, transmission,
// diffuse, subsurface) and the coat's view-side base scale.
void probe_lobes(const float *m, const float *view, float *probabilities, float *coatBaseScale);
// F82-tint metal Fresnel at cosine c, and its hemispherical average.
void probe_metal_fresnel(float c, const float *f0, const float *tint, float *out, float *average);
// Directional energy the coat reflects at cosine c.
float probe_coat_energy(const float *m, float c);

}  // namespace orc
