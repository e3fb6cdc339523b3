// Oracle C entry points for the Python test harness (ctypes). Test
// infrastructure only: nothing in the product links this library.
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/mtx_fmath.h"
#include "orc_bvh.h"
#include "orc_render.h"
#include "orc_sobol.h"
#include "orc_trace.h"

using namespace orc;

extern "C" {

// ---------------------------------------------------------------- sampler
uint32_t orc_sample_hash(uint32_t x) { return sample_hash(x); }
uint32_t orc_sample_seed(uint32_t a, uint32_t b) { return sample_seed(a, b); }
uint32_t orc_owen(uint32_t x, uint32_t s) { return owen(x, s); }
uint32_t orc_morton2(uint32_t x, uint32_t y) { return morton2(x, y); }
uint32_t orc_zsobol_index(uint32_t m, uint32_t dim, uint32_t logSPP) { return zsobol_index(m, dim, logSPP); }
uint32_t orc_sobol_d1(uint32_t i) { return sobol_d1(i); }
void orc_pcg4d(uint32_t *v) { pcg4d(v); }
float orc_fuzz_albedo(float viewCosine, float roughness) { return fuzz_albedo(viewCosine, roughness); }
double orc_halton(uint32_t index, uint32_t base) { return halton(index, base); }

void orc_sobol2d_batch(const uint32_t *index, uint32_t n, uint32_t seed, uint32_t *out) {
  for (uint32_t i = 0; i < n; i++) sobol2d_bits(index[i], seed, out + 2 * i);
}

// Structured sample for (pixel, sample index, dimension) with the EPS remap,
// as sobolReference.ts:52-68 referenceSample (hashed = structured dims off).
void orc_structured_sample(uint32_t pattern, uint32_t px, uint32_t py, uint32_t width,
                           uint32_t index, uint32_t dim, uint32_t logSPP, uint32_t seed,
                           float *out) {
  Sampler s{};
  s.px = px; s.py = py; s.width = width; s.pixelSampleIndex = index;
  s.sequenceSeed = seed; s.zSobolLogSPP = logSPP; s.pattern = pattern;
  s.sample2d(dim, out);
}

// ---------------------------------------------------------------- BVH
struct OrcBvh { BinnedSahBvh bvh; Bvh2Layout layout; };

void *orc_bvh_build(const float *aabbs, uint32_t count, uint32_t bins) {
  OrcBvh *h = new OrcBvh;
  h->bvh = build_binned_sah(aabbs, count, bins);
  h->layout = layout_bvh2(h->bvh);
  return h;
}
void orc_bvh_info(void *hp, uint32_t *info) {
  OrcBvh *h = (OrcBvh *)hp;
  info[0] = h->bvh.nodeCount; info[1] = h->layout.vec4Count;
  info[2] = h->layout.maxDepth; info[3] = h->layout.maxLeafPrimitives;
}
void orc_bvh_get(void *hp, float *bounds, uint32_t *leftFirst, uint32_t *primCount,
                 uint32_t *primIndices, uint32_t *order, uint32_t *offsets) {
  OrcBvh *h = (OrcBvh *)hp;
  uint32_t n = h->bvh.nodeCount;
  if (bounds) std::memcpy(bounds, h->bvh.bounds.data(), (size_t)n * 24);
  if (leftFirst) std::memcpy(leftFirst, h->bvh.leftFirst.data(), (size_t)n * 4);
  if (primCount) std::memcpy(primCount, h->bvh.primCount.data(), (size_t)n * 4);
  if (primIndices) std::memcpy(primIndices, h->bvh.primIndices.data(), h->bvh.primIndices.size() * 4);
  if (order) std::memcpy(order, h->layout.order.data(), (size_t)n * 4);
  if (offsets) std::memcpy(offsets, h->layout.offsets.data(), (size_t)n * 4);
}
void orc_bvh_free(void *hp) { delete (OrcBvh *)hp; }

struct OrcFlat { FlatBvh f; std::vector<uint32_t> source; };

void *orc_blas_build(const float *positions, uint32_t vertexCount, const uint32_t *indices,
                     uint32_t triCount, uint32_t bins) {
  OrcFlat *h = new OrcFlat;
  h->f = build_blas(positions, vertexCount, indices, triCount, bins, &h->source);
  return h;
}

double orc_blas_refit_words(uint32_t *words, size_t vec4Count, const float *positions, uint32_t vertexCount) {
  return orc::refit_blas_words(words, vec4Count, positions, vertexCount);
}

void *orc_tlas_build(const uint32_t *blasIndex, const float *bounds, const double *transforms,
                     const uint32_t *materials, const float *colors, uint32_t count,
                     const uint32_t *blasOffsets, const uint32_t *vertexOffsets, uint32_t bins,
                     uint32_t perInstanceBounds) {
  std::vector<TlasInstance> inst(count);
  for (uint32_t i = 0; i < count; i++) {
    inst[i].blasIndex = blasIndex[i];
    std::memcpy(inst[i].bounds, bounds + (size_t)i * 6, 24);
    std::memcpy(inst[i].transform, transforms + (size_t)i * 16, 128);
    inst[i].material = materials[i];
    for (int k = 0; k < 3; k++) inst[i].color[k] = colors ? colors[i * 3 + k] : 1.0f;
  }
  OrcFlat *h = new OrcFlat;
  h->f = build_tlas(inst.data(), count, blasOffsets, vertexOffsets, bins, perInstanceBounds != 0);
  return h;
}

void orc_tlas_refit(void *hp, const uint32_t *blasIndex, const float *bounds,
                    const double *transforms, const uint32_t *materials, const float *colors,
                    uint32_t count, const uint32_t *blasOffsets, const uint32_t *vertexOffsets,
                    const uint32_t *dirty, uint32_t dirtyCount) {
  std::vector<TlasInstance> inst(count);
  for (uint32_t i = 0; i < count; i++) {
    inst[i].blasIndex = blasIndex[i];
    std::memcpy(inst[i].bounds, bounds + (size_t)i * 6, 24);
    std::memcpy(inst[i].transform, transforms + (size_t)i * 16, 128);
    inst[i].material = materials[i];
    for (int k = 0; k < 3; k++) inst[i].color[k] = colors ? colors[i * 3 + k] : 1.0f;
  }
  OrcFlat *h = (OrcFlat *)hp;
  refit_tlas(h->f, inst.data(), count, blasOffsets, vertexOffsets, dirty, dirtyCount);
}

void orc_flat_info(void *hp, uint64_t *info) {
  OrcFlat *h = (OrcFlat *)hp;
  info[0] = h->f.words.size(); info[1] = h->f.instanceWords.size();
  info[2] = h->f.maxDepth; info[3] = h->f.maxLeafPrimitives;
  double a = h->f.buildSurfaceArea; std::memcpy(&info[4], &a, 8);
}
void orc_flat_get(void *hp, uint32_t *words, uint32_t *instanceWords, uint32_t *source) {
  OrcFlat *h = (OrcFlat *)hp;
  if (words) std::memcpy(words, h->f.words.data(), h->f.words.size() * 4);
  if (instanceWords) std::memcpy(instanceWords, h->f.instanceWords.data(), h->f.instanceWords.size() * 4);
  if (source && !h->source.empty()) std::memcpy(source, h->source.data(), h->source.size() * 4);
}
void orc_flat_free(void *hp) { delete (OrcFlat *)hp; }

void orc_instance_world_bounds(const double *m, const float *lo, const float *hi, float *out) {
  instance_world_bounds(m, lo, hi, out);
}

// ---------------------------------------------------------------- traversal
void *orc_scene_create(const uint32_t *tlas, uint64_t tlasWords, const uint32_t *blas,
                       uint64_t blasWords, const uint32_t *instances, uint64_t instanceWords,
                       const float *positions, uint64_t vertexCount, const uint8_t *matDouble,
                       uint32_t materialCount, uint32_t perInstanceBounds) {
  TraceScene *s = new TraceScene;
  s->tlas.assign(tlas, tlas + tlasWords);
  s->blas.assign(blas, blas + blasWords);
  s->instances.assign(instances, instances + instanceWords);
  s->positions.assign(positions, positions + vertexCount * 3);
  if (matDouble) s->materialDoubleSided.assign(matDouble, matDouble + materialCount);
  s->perInstanceBounds = perInstanceBounds != 0;
  return s;
}
void orc_scene_free(void *s) { delete (TraceScene *)s; }

// rays: n x 8 floats (o.xyz, tmax, d.xyz, unused). hits: n x 8 words
// (hit, instance target, record, t bits, bary xyz bits, 0). stats: n x 3 (optional).
void orc_trace_closest_batch(void *sp, const float *rays, uint32_t n, uint32_t doubleSided,
                             uint32_t *hits, uint32_t *stats) {
  TraceScene *s = (TraceScene *)sp;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const float *r = rays + i * 8;
    TraceStats st;
    ClosestHit h = trace_closest(*s, V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), r[3],
                                 doubleSided != 0, 0, -1, -1, stats ? &st : nullptr);
    uint32_t *o = hits + i * 8;
    o[0] = h.hit; o[1] = (uint32_t)h.instance; o[2] = (uint32_t)h.record; o[3] = f2u(h.t);
    o[4] = f2u(h.bary.x); o[5] = f2u(h.bary.y); o[6] = f2u(h.bary.z); o[7] = 0;
    if (stats) { stats[i * 3] = (uint32_t)st.boxTests; stats[i * 3 + 1] = (uint32_t)st.triTests; stats[i * 3 + 2] = (uint32_t)st.instanceVisits; }
  }
}

void orc_trace_any_batch(void *sp, const float *rays, uint32_t n, uint32_t *occluded) {
  TraceScene *s = (TraceScene *)sp;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const float *r = rays + i * 8;
    occluded[i] = trace_any(*s, V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), r[3], -1, -1, nullptr);
  }
}

void orc_trace_brute_batch(void *sp, const float *rays, uint32_t n, uint32_t doubleSided,
                           uint32_t *hits) {
  TraceScene *s = (TraceScene *)sp;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const float *r = rays + i * 8;
    ClosestHit h = trace_brute(*s, V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), r[3], doubleSided != 0);
    uint32_t *o = hits + i * 8;
    o[0] = h.hit; o[1] = (uint32_t)h.instance; o[2] = (uint32_t)h.record; o[3] = f2u(h.t);
    o[4] = f2u(h.bary.x); o[5] = f2u(h.bary.y); o[6] = f2u(h.bary.z); o[7] = 0;
  }
}

// ---------------------------------------------------------------- integrator
void orc_set_energy_lut(const float *lut) { set_energy_lut(lut); }

void orc_bsdf_eval(const float *m, const float *view, uint32_t n, const float *dirs, float *f, float *pdf) {
  probe_bsdf_eval(m, view, n, dirs, f, pdf);
}
void orc_bsdf_sample(const float *m, const float *view, uint32_t n, const float *u, float *dirs, uint8_t *ok) {
  probe_bsdf_sample(m, view, n, u, dirs, ok);
}
void orc_bsdf_lobes(const float *m, const float *view, float *probabilities, float *coatBaseScale) {
  probe_lobes(m, view, probabilities, coatBaseScale);
}
void orc_metal_fresnel(float c, const float *f0, const float *tint, float *out, float *average) {
  probe_metal_fresnel(c, f0, tint, out, average);
}
float orc_coat_energy(const float *m, float c) { return probe_coat_energy(m, c); }

// settings: width, height, bounces, candidates, pattern, seed, logSPP, sampleOffset.
// out: w*h*13 doubles per pixel (color rgba, albedo rgba, normal rgba, lum^2).
uint32_t orc_emissive_triangles(const float *positions, const uint32_t *vertexOffsets, const uint32_t *indices,
                                const uint32_t *indexOffsets, const uint32_t *triangleCounts,
                                const uint32_t *instanceMesh, const double *transforms,
                                const uint32_t *instanceMaterial, uint32_t instanceCount, const float *materials,
                                float *out, double *weight, uint32_t capacity, uint32_t *corners) {
  return emissive_triangles(positions, vertexOffsets, indices, indexOffsets, triangleCounts, instanceMesh,
                            transforms, instanceMaterial, instanceCount, materials, out, weight, capacity, corners);
}

// Material textures and UVs for orc_render: n textures of 11 words each
// (width, height, storage (0 RGBA8, 1 RGBA32F), flip_y, uv_channel, wrap_s,
// wrap_t, then the f32 bits of uv_scale[2] and uv_offset[2]) with their
// row-major pixels, UV0 / UV1 per global vertex (2 floats each) and 12 floats
// per emissive triangle (uv0, uv1, uv2, then the corners' UV1).
void *orc_texset_create(uint32_t n, const uint32_t *desc, const void *const *pixels, uint32_t vertexCount,
                        const float *uvs, const float *uvs1, uint32_t emissiveCount, const float *emissiveUv) {
  auto *T = new TextureSet();
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t *d = desc + (size_t)i * 11;
    Texture t{};
    t.width = d[0]; t.height = d[1]; t.storage = d[2];
    const bool flip = d[3] != 0u;
    t.uvChannel = d[4]; t.wrapS = d[5]; t.wrapT = d[6];
    std::memcpy(&t.sx, d + 7, 4); std::memcpy(&t.sy, d + 8, 4);
    std::memcpy(&t.ox, d + 9, 4); std::memcpy(&t.oy, d + 10, 4);
    const size_t row = (size_t)t.width * (t.storage == 1u ? 4 : 1);
    for (uint32_t y = 0; y < t.height; y++) {
      const size_t src = (size_t)(flip ? t.height - 1 - y : y) * row;
      if (t.storage == 1u) {
        const float *f = static_cast<const float *>(pixels[i]) + src;
        t.rgba32.insert(t.rgba32.end(), f, f + row);
      } else {
        const uint32_t *w = static_cast<const uint32_t *>(pixels[i]) + src;
        t.rgba8.insert(t.rgba8.end(), w, w + row);
      }
    }
    T->tex.push_back(std::move(t));
  }
  T->uvs.assign(uvs, uvs + (size_t)vertexCount * 2);
  T->uvs1.assign(uvs1, uvs1 + (size_t)vertexCount * 2);
  T->emissiveUv.assign(emissiveUv, emissiveUv + (size_t)emissiveCount * 12);
  return T;
}
void orc_texset_destroy(void *T) { delete static_cast<TextureSet *>(T); }

// sampleMap of texture `index` at n UVs (RGBA, 4 floats each), and the
// decoded color (decodeMapColor, 3 floats each).
void orc_sample_map(void *T, uint32_t index, uint32_t n, const float *uv, float *rgba, float *decoded) {
  const TextureSet &S = *static_cast<TextureSet *>(T);
  for (uint32_t i = 0; i < n; i++) probe_sample_map(S, index, uv[2 * i], uv[2 * i + 1], rgba + 4 * i, decoded + 3 * i);
}

static RenderScene render_scene(void *trace, const float *normals, const float *materials, uint32_t materialCount,
                                const float *lights, uint32_t lightCount, const float *env, uint32_t envW,
                                uint32_t envH, const float *camera, const uint32_t *settings, float clamping,
                                const float *emissive, const double *emissiveWeight, uint32_t emissiveCount,
                                void *textures) {
  RenderScene s{};
  s.textures = static_cast<const TextureSet *>(textures);
  s.emissive = emissive;
  s.emissiveWeight = emissiveWeight;
  s.emissiveCount = emissiveCount;
  s.trace = (const TraceScene *)trace;
  s.normals = normals;
  s.materials = materials; s.materialCount = materialCount;
  s.lights = lights; s.lightCount = lightCount;
  s.env = env; s.envW = envW; s.envH = envH;
  s.camera = camera;
  s.width = settings[0]; s.height = settings[1]; s.bounces = settings[2]; s.candidates = settings[3];
  s.pattern = settings[4]; s.seed = settings[5]; s.logSPP = settings[6]; s.sampleOffset = settings[7];
  s.clamping = clamping;
  return s;
}

uint64_t orc_render(void *trace, const float *normals, const float *materials, uint32_t materialCount,
                    const float *lights, uint32_t lightCount, const float *env, uint32_t envW, uint32_t envH,
                    const float *camera, const uint32_t *settings, float clamping, uint32_t x0, uint32_t y0,
                    uint32_t w, uint32_t h, uint32_t sampleStart, uint32_t count, double *out,
                    const float *emissive, const double *emissiveWeight, uint32_t emissiveCount, void *textures) {
  RenderScene s = render_scene(trace, normals, materials, materialCount, lights, lightCount, env, envW, envH, camera,
                               settings, clamping, emissive, emissiveWeight, emissiveCount, textures);
  static_assert(sizeof(Accum) == 13 * sizeof(double), "Accum layout");
  return render(s, x0, y0, w, h, sampleStart, count, reinterpret_cast<Accum *>(out));
}

// Draw tags of one path of the random pattern (same scene arguments as
// orc_render); returns the number of draws.
uint32_t orc_path_draws(void *trace, const float *normals, const float *materials, uint32_t materialCount,
                        const float *lights, uint32_t lightCount, const float *env, uint32_t envW, uint32_t envH,
                        const float *camera, const uint32_t *settings, float clamping, uint32_t px, uint32_t py,
                        uint32_t sampleIndex, uint32_t *log, uint32_t cap, const float *emissive,
                        const double *emissiveWeight, uint32_t emissiveCount, void *textures) {
  RenderScene s = render_scene(trace, normals, materials, materialCount, lights, lightCount, env, envW, envH, camera,
                               settings, clamping, emissive, emissiveWeight, emissiveCount, textures);
  return path_draws(s, px, py, sampleIndex, log, cap);
}

}  // extern "C"

// OpenMP threads the batch entry points use (for the CPU-baseline report).
extern "C" int32_t orc_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Elementary functions of include/mtx_fmath.h on the host: kind 0 sin, 1 cos,
// 2 pow(x, y), 3 atan2(x, y) (x is the ordinate), 4 acos.
extern "C" void orc_math_eval(uint32_t kind, const float *x, const float *y, float *out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    switch (kind) {
      case 0: out[i] = mtxfm::sinf(x[i]); break;
      case 1: out[i] = mtxfm::cosf(x[i]); break;
      case 2: out[i] = mtxfm::powf(x[i], y[i]); break;
      case 3: out[i] = mtxfm::atan2f(x[i], y[i]); break;
      case 4: out[i] = mtxfm::acosf(x[i]); break;
      case 5: out[i] = mtxfm::expf(x[i]); break;
      case 6: out[i] = mtxfm::sincosf(x[i]).s; break;
      default: out[i] = mtxfm::sincosf(x[i]).c; break;
    }
  }
}

// Host light tables of a scene's lights and environment. Returns the alias
// table length n; prob/alias/type/index get n entries (capacity cap),
// lightPmf one per light, scalars[0] = env pmf, totalOut = total weight.
extern "C" uint32_t orc_light_tables(const float *lights, uint32_t lightCount, const float *env, uint32_t envW,
                                     uint32_t envH, float *prob, int32_t *alias, int32_t *type, int32_t *index,
                                     uint32_t cap, float *lightPmf, float *envPmf, double *totalOut) {
  RenderScene s{};
  s.lights = lights; s.lightCount = lightCount;
  s.env = env; s.envW = envW; s.envH = envH;
  HostTables T;
  build_tables(s, T);
  uint32_t n = (uint32_t)T.prob.size();
  for (uint32_t i = 0; i < n && i < cap; i++) {
    prob[i] = T.prob[i]; alias[i] = T.alias[i]; type[i] = T.type[i]; index[i] = T.index[i];
  }
  for (uint32_t i = 0; i < lightCount; i++) lightPmf[i] = T.lightPmf[i];
  *envPmf = T.envPmf;
  *totalOut = T.totalWeight;
  return n;
}

// Environment CDF: (W+2) x (H+1) x 2 floats into out.
extern "C" void orc_env_cdf(const float *env, uint32_t envW, uint32_t envH, float *out) {
  RenderScene s{};
  s.env = env; s.envW = envW; s.envH = envH;
  HostTables T;
  build_env_cdf(s, T);
  for (size_t i = 0; i < T.cdf.size(); i++) out[i] = T.cdf[i];
}

extern "C" void orc_camera_rays(const float *camera, uint32_t width, uint32_t height, const uint32_t *pix,
                                uint64_t n, uint32_t pattern, uint32_t seed, uint32_t logSPP, float *out) {
  for (uint64_t i = 0; i < n; i++)
    orc::camera_sample(camera, width, height, pix[3 * i], pix[3 * i + 1], pix[3 * i + 2], pattern, seed, logSPP,
                       out + 6 * i);
}
