// C ABI for resident render scenes and render targets (progressive FP64
// accumulation), plus the host-side light tables: the Walker alias table over
// {environment, rect lights} (lightSampling.ts:72-107,172-212,415-431) and the
// environment CDF with its 3x3 max-dilated luminance (envMapDistribution.ts:57-134).
#include "mtx_display.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <deque>
#include <mutex>

#include "mtx_pt.h"
#include "mtx_registry.h"

namespace mtx {

extern const uint32_t kEnergyLutBits[32 * 2177];

struct RenderSceneObj {
  std::vector<std::shared_ptr<Blas>> blases;
  std::shared_ptr<Accel> accel;
  std::vector<DevBuf<float>> normals, uvs, uvs1;
  std::vector<DevBuf<uint32_t>> indices;
  DevBuf<ShadeSlot> shade;
  DevBuf<float> materials, lights, lightPmf, lut;
  DevBuf<LightChoice> choices;
  DevBuf<float4> env, emissive, emissiveUv, texels32;
  DevBuf<float2> envCdf;
  DevBuf<uint16_t> envGuide;
  DevBuf<GpuTexture> textures;
  DevBuf<uint32_t> texels8;
  std::vector<mtx_instance_desc> inst;
  GpuScene view{};
  uint64_t triangles = 0;
  size_t bytes() const {
    size_t b = shade.bytes() + materials.bytes() + lights.bytes() + lut.bytes() + env.bytes() + envCdf.bytes() +
               envGuide.bytes() + emissive.bytes() + emissiveUv.bytes() + textures.bytes() + texels8.bytes() +
               texels32.bytes();
    for (auto &x : normals) b += x.bytes();
    for (auto &x : uvs) b += x.bytes();
    for (auto &x : uvs1) b += x.bytes();
    for (auto &x : indices) b += x.bytes();
    for (auto &x : blases) b += x->resident_bytes();
    return b;
  }
};

struct TargetObj {
  std::shared_ptr<RenderSceneObj> scene;
  RenderParams params{};
  DevBuf<double> planes;
  DevBuf<uint32_t> count;
  DevBuf<uint64_t> rays;
  uint32_t samplesDone = 0;
  uint32_t kernel = MTX_KERNEL_WAVEFRONT;
  WfBuffers::Owner wf;
  WfConfig wfConfig;
  WfGraph graph;
  DevBuf<float> targets;  // per-pixel sample limits (adaptive sampling)
  JnlmBuffers jnlm;       // work images of the JNLM denoiser
  bool useTargets = false;
  hipStream_t stream = nullptr;
  std::mutex busy;
};

static std::mutex gRenderMutex;
static std::unordered_map<uint64_t, std::shared_ptr<RenderSceneObj>> gScenes;
static std::unordered_map<uint64_t, std::shared_ptr<TargetObj>> gTargets;

std::shared_ptr<RenderSceneObj> get_scene(uint64_t h) {
  std::lock_guard<std::mutex> g(gRenderMutex);
  auto it = gScenes.find(h);
  if (it == gScenes.end()) throw Error(MTX_INVALID_HANDLE, "unknown scene handle");
  return it->second;
}
std::shared_ptr<TargetObj> get_target(uint64_t h) {
  std::lock_guard<std::mutex> g(gRenderMutex);
  auto it = gTargets.find(h);
  if (it == gTargets.end()) throw Error(MTX_INVALID_HANDLE, "unknown render target handle");
  return it->second;
}

// ---------------------------------------------------------------- light tables
struct LightTables {
  std::vector<LightChoice> choices;
  std::vector<float> lightPmf, emissivePmf;
  float envPmf = 0.0f;
  double total = 0.0;
};

// collectEmissiveTriangles (lightSampling.ts:237-320): every triangle of an
// instance whose material emits (luminance of the packed emission above
// 1e-8), in instance then triangle order, moved to world space with the
// instance transform (f64, f32 storage as vec3.transformMat4 does), with its
// area (half the f64 length of the f32 edge cross product; degenerate
// triangles are skipped) and alias weight area * luminance
// (prepareSceneLightSampling, sceneAssembly.ts:46).
// With textures each triangle also records its vertex UVs (3 float4:
// (uv0, uv1), (uv2, uv1_0), (uv1_1, uv1_2)); UV1 falls back to UV0, and a
// mesh without UVs records zeros (lightSampling.ts:262-299).
struct EmissiveList {
  std::vector<float4> tris;  // 3 per triangle: (p0, pmf), (p1, area), (p2, material)
  std::vector<float4> uvs;
  std::vector<double> weight;
};
// UV0 / UV1 arrays of a mesh descriptor (absent in descriptors that end before them).
static const float *mesh_uvs(const mtx_mesh_desc &m) {
  return m.struct_size >= offsetof(mtx_mesh_desc, uvs) + sizeof(m.uvs) ? m.uvs : nullptr;
}
static const float *mesh_uvs1(const mtx_mesh_desc &m) {
  return m.struct_size >= offsetof(mtx_mesh_desc, uvs1) + sizeof(m.uvs1) ? m.uvs1 : nullptr;
}
static EmissiveList collect_emissive(const mtx_scene_desc &D, bool withUvs) {
  EmissiveList E;
  for (uint32_t i = 0; i < D.instance_count; i++) {
    const mtx_instance_desc &d = D.instances[i];
    const uint32_t mat = d.material_word >> 8;
    const float *em = D.materials + (size_t)mat * MTX_MATERIAL_FLOATS + MTX_MAT_EMISSION;
    const double lum = std::max((double)em[0] * 0.2126 + (double)em[1] * 0.7152 + (double)em[2] * 0.0722, 0.0);
    if (!(lum > 1e-8)) continue;
    const mtx_mesh_desc &m = D.meshes[d.blas_index];
    const double *M = d.transform;
    for (uint32_t t = 0; t < m.triangle_count; t++) {
      float w[3][3];
      for (int c = 0; c < 3; c++) {
        const float *v = m.positions + (size_t)m.indices[(size_t)t * 3 + c] * 3;
        const double x = v[0], y = v[1], z = v[2];
        double h = M[3] * x + M[7] * y + M[11] * z + M[15];
        if (h == 0.0) h = 1.0;
        w[c][0] = (float)((M[0] * x + M[4] * y + M[8] * z + M[12]) / h);
        w[c][1] = (float)((M[1] * x + M[5] * y + M[9] * z + M[13]) / h);
        w[c][2] = (float)((M[2] * x + M[6] * y + M[10] * z + M[14]) / h);
      }
      float e0[3], e1[3];
      for (int k = 0; k < 3; k++) {
        e0[k] = (float)((double)w[1][k] - (double)w[0][k]);
        e1[k] = (float)((double)w[2][k] - (double)w[0][k]);
      }
      const float cx = (float)((double)e0[1] * e1[2] - (double)e0[2] * e1[1]);
      const float cy = (float)((double)e0[2] * e1[0] - (double)e0[0] * e1[2]);
      const float cz = (float)((double)e0[0] * e1[1] - (double)e0[1] * e1[0]);
      const double area = 0.5 * std::hypot((double)cx, (double)cy, (double)cz);
      if (!(area > 1e-10) || !std::isfinite(area)) continue;
      E.tris.push_back(make_float4(w[0][0], w[0][1], w[0][2], 0.0f));
      E.tris.push_back(make_float4(w[1][0], w[1][1], w[1][2], (float)area));
      E.tris.push_back(make_float4(w[2][0], w[2][1], w[2][2], (float)mat));
      E.weight.push_back(area * lum);
      if (withUvs) {
        const float *u0 = mesh_uvs(m), *u1 = mesh_uvs1(m);
        if (u1 == nullptr) u1 = u0;
        float q[12] = {};  // uv0, uv1, uv2, then UV1 of the three corners
        for (int c = 0; c < 3; c++) {
          const size_t v = m.indices[(size_t)t * 3 + c];
          if (u0 != nullptr) { q[2 * c] = u0[v * 2]; q[2 * c + 1] = u0[v * 2 + 1]; }
          if (u1 != nullptr) { q[6 + 2 * c] = u1[v * 2]; q[6 + 2 * c + 1] = u1[v * 2 + 1]; }
        }
        for (int k = 0; k < 3; k++) E.uvs.push_back(make_float4(q[4 * k], q[4 * k + 1], q[4 * k + 2], q[4 * k + 3]));
      }
    }
  }
  return E;
}

// estimateAnalyticLightPower (lightSampling.ts:72-107) from the packed light:
// the emission holds color * intensity, or the radiance of disk, cylinder and
// finite spot lights, so emitted luminance * pi * area covers the area lights;
// point, spot (flux factor of its cone) and directional lights use 4 pi.
static double analytic_power(const float *L) {
  const double kPiD = 3.141592653589793;
  const double lum = std::max((double)L[MTX_LIGHT_EMISSION] * 0.2126 + (double)L[MTX_LIGHT_EMISSION + 1] * 0.7152 +
                                  (double)L[MTX_LIGHT_EMISSION + 2] * 0.0722, 0.0);
  if (!(lum > 0)) return 0.0;
  const int type = (int)L[MTX_LIGHT_TYPE];
  const double area = (double)L[MTX_LIGHT_AREA], radius = (double)L[MTX_LIGHT_RADIUS];
  switch (type) {
    case 0:
    case MTX_LIGHT_RECT:
    case MTX_LIGHT_DISK:
    case MTX_LIGHT_CYLINDER: return lum * kPiD * std::max(area, 0.0);
    case MTX_LIGHT_SPHERE: return radius == 0.0 ? lum * 4 * kPiD : lum * kPiD * std::max(area, 0.0);
    case MTX_LIGHT_SPOT: {
      const double flux = 0.5 * (1 - 0.5 * ((double)L[MTX_LIGHT_V + 1] + (double)L[MTX_LIGHT_V]));
      return radius == 0.0 ? lum * 4 * kPiD * flux : lum * kPiD * area * flux;
    }
    default: return lum * 4 * kPiD;  // directional
  }
}

static LightTables light_tables(const float *lights, uint32_t nLights, const float *env, uint32_t W, uint32_t H,
                                const std::vector<double> &emissiveWeight = {}) {
  struct In { int32_t type, index; double w; };
  std::vector<In> in;
  auto push = [&](int32_t t, int32_t i, double w) {
    if (w > 0 && std::isfinite(w)) in.push_back({t, i, w});
  };
  double envPower = 0.0;
  if (env && W && H) {
    const double kPiD = 3.141592653589793;
    for (uint32_t y = 0; y < H; y++) {
      double sinT = std::sin(kPiD * (y + 0.5) / H);
      for (uint32_t x = 0; x < W; x++) {
        const float *p = env + ((size_t)y * W + x) * 4;
        double l = (double)p[0] * 0.2126 + (double)p[1] * 0.7152 + (double)p[2] * 0.0722;
        envPower += (l > 0 ? l : 0) * sinT;
      }
    }
    envPower = envPower * (2 * kPiD * kPiD) / ((double)W * H);
  }
  push(0, 0, envPower);
  for (uint32_t i = 0; i < nLights; i++) push(1, (int32_t)i, analytic_power(lights + (size_t)i * MTX_LIGHT_FLOATS));
  for (size_t i = 0; i < emissiveWeight.size(); i++) push(2, (int32_t)i, emissiveWeight[i]);
  LightTables T;
  T.lightPmf.assign(nLights, 0.0f);
  T.emissivePmf.assign(emissiveWeight.size(), 0.0f);
  const size_t n = in.size();
  double total = 0.0;
  for (auto &e : in) total += e.w;
  if (!n || !(total > 0)) return T;
  T.total = total;
  T.choices.resize(n);
  std::vector<float> pmf(n);
  std::vector<double> scaled(n);
  std::vector<uint32_t> small, large;
  for (size_t i = 0; i < n; i++) {
    pmf[i] = (float)(in[i].w / total);
    scaled[i] = (double)pmf[i] * (double)n;
    (scaled[i] < 1 ? small : large).push_back((uint32_t)i);
  }
  while (!small.empty() && !large.empty()) {
    uint32_t lo = small.back(), hi = large.back();
    small.pop_back();
    large.pop_back();
    T.choices[lo].prob = (float)scaled[lo];
    T.choices[lo].alias = (int32_t)hi;
    scaled[hi] = scaled[hi] + scaled[lo] - 1;
    (scaled[hi] < 1 ? small : large).push_back(hi);
  }
  for (uint32_t i : small) { T.choices[i].prob = 1.0f; T.choices[i].alias = (int32_t)i; }
  for (uint32_t i : large) { T.choices[i].prob = 1.0f; T.choices[i].alias = (int32_t)i; }
  for (size_t i = 0; i < n; i++) {
    T.choices[i].type = in[i].type;
    T.choices[i].index = in[i].index;
    if (in[i].type == 0) T.envPmf = pmf[i];
    else if (in[i].type == 1) T.lightPmf[in[i].index] = pmf[i];
    else T.emissivePmf[in[i].index] = pmf[i];
  }
  return T;
}

static std::vector<float2> env_cdf(const float *env, uint32_t W, uint32_t H) {
  const uint32_t CW = W + 2, CH = H + 1;
  std::vector<float2> cdf((size_t)CW * CH, make_float2(0.0f, 0.0f));
  auto X = [&](uint32_t x, uint32_t y) -> float & { return cdf[(size_t)y * CW + x].x; };
  auto Y = [&](uint32_t x, uint32_t y) -> float & { return cdf[(size_t)y * CW + x].y; };
  std::vector<float> lum((size_t)W * H), horiz((size_t)W * H);
  for (size_t i = 0; i < (size_t)W * H; i++) {
    const float *p = env + i * 4;
    lum[i] = (float)(0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2]);
  }
  for (uint32_t y = 0; y < H; y++)
    for (uint32_t x = 0; x < W; x++) {
      size_t r = (size_t)y * W;
      double a = lum[r + (x ? x - 1 : 0)], b = lum[r + x], c = lum[r + std::min(W - 1, x + 1)];
      horiz[r + x] = (float)std::max(std::max(0.0, a), std::max(b, c));
    }
  for (uint32_t y = 0; y < H; y++) {
    double sinT = std::sin(3.141592653589793 * (y + 0.5) / H);
    size_t above = (size_t)(y ? y - 1 : 0) * W, row = (size_t)y * W, below = (size_t)std::min(H - 1, y + 1) * W;
    for (uint32_t x = 0; x < W; x++) {
      double w = std::max(std::max((double)horiz[above + x], (double)horiz[row + x]),
                          std::max((double)horiz[below + x], 1e-12)) * sinT;
      X(x + 2, y) = (float)((double)X(x + 1, y) + w / W);
      Y(x + 1, y) = (float)w;
    }
    double rowIntegral = X(CW - 1, y);
    for (uint32_t x = 1; x < CW; x++) {
      X(x, y) = (float)((double)X(x, y) / rowIntegral);
      Y(x, y) = (float)((double)Y(x, y) / rowIntegral);
    }
    X(0, y + 1) = (float)((double)X(0, y) + rowIntegral / H);
    Y(0, y) = (float)rowIntegral;
  }
  double integral = X(0, CH - 1);
  for (uint32_t y = 0; y < CH; y++) {
    X(0, y) = (float)((double)X(0, y) / integral);
    Y(0, y) = (float)((double)Y(0, y) / integral);
  }
  return cdf;
}

// Guide tables of the environment CDF (see kEnvGuideRows): for each bin
// edge r = k / bins, the result of sample_env's search (first entry with
// CDF > r, or the end of the range) over the marginal CDF (column 0) and
// over each row's conditional CDF. Empty when an index does not fit 16 bits.
static std::vector<uint16_t> env_guides(const std::vector<float2> &cdf, uint32_t W, uint32_t H) {
  std::vector<uint16_t> g;
  if (W + 1 > 65535u || H + 1 > 65535u) return g;
  const uint32_t CW = W + 2;
  auto search = [](auto at, uint32_t n, float r) {  // first i in [0, n) with at(i) > r, else n
    uint32_t i = 0;
    while (i < n && !(at(i) > r)) i++;
    return (uint16_t)i;
  };
  g.reserve((kEnvGuideRows + 1) + (size_t)H * (kEnvGuideCols + 1));
  auto marginal = [&](uint32_t i) { return cdf[(size_t)i * CW].x; };
  for (uint32_t k = 0; k <= kEnvGuideRows; k++)
    g.push_back(search(marginal, H + 1, (float)k / (float)kEnvGuideRows));
  for (uint32_t y = 0; y < H; y++) {
    auto row = [&](uint32_t i) { return cdf[(size_t)y * CW + 1 + i].x; };
    // one pass per row: bin edges are increasing
    uint32_t i = 0;
    for (uint32_t k = 0; k <= kEnvGuideCols; k++) {
      const float r = (float)k / (float)kEnvGuideCols;
      while (i < W + 1 && !(row(i) > r)) i++;
      g.push_back((uint16_t)i);
    }
  }
  return g;
}

static void fill_shade(RenderSceneObj &R) {
  const Accel &A = *R.accel;
  std::vector<ShadeSlot> sh(std::max<uint32_t>(1, A.instanceCount));
  for (uint32_t q = 0; q < A.instanceCount; q++) {
    uint32_t src = A.slotSource[q];
    const mtx_instance_desc &d = R.inst[src];
    const double *m = d.transform;
    ShadeSlot s{};
    s.l2w0 = make_float4((float)m[0], (float)m[1], (float)m[2], 0.0f);
    s.l2w1 = make_float4((float)m[4], (float)m[5], (float)m[6], 0.0f);
    s.l2w2 = make_float4((float)m[8], (float)m[9], (float)m[10], 0.0f);
    s.l2w3 = make_float4((float)m[12], (float)m[13], (float)m[14], 0.0f);
    s.normals = R.normals[d.blas_index].ptr;
    s.indices = R.indices[d.blas_index].ptr;
    s.uvs = R.uvs[d.blas_index].count ? R.uvs[d.blas_index].ptr : nullptr;
    s.uvs1 = R.uvs1[d.blas_index].count ? R.uvs1[d.blas_index].ptr : nullptr;
    s.material = d.material_word >> 8;
    for (int k = 0; k < 3; k++) s.color[k] = d.color[k];
    sh[q] = s;
  }
  R.shade.upload_min1(sh.data(), A.instanceCount, 0);
}

// Material textures (the descriptor's texture table, mutracer-scene.h:87-95):
// validated, rows put in sampling order (flip_y reverses them, as
// textureAtlasPixels.ts:35-37 does) and packed into one pool per storage.
// Returns the texture count.
static uint32_t upload_textures(RenderSceneObj &R, const mtx_scene_desc &D) {
  const uint32_t n =
      D.struct_size >= offsetof(mtx_scene_desc, textures) + sizeof(D.textures) ? D.texture_count : 0u;
  if (n == 0) return 0;
  MTX_REQUIRE(D.textures, MTX_INVALID_ARGUMENT, "textures is NULL");
  std::vector<GpuTexture> T(n);
  size_t n8 = 0, n32 = 0;
  for (uint32_t i = 0; i < n; i++) {
    const mtx_texture_desc &t = D.textures[i];
    MTX_REQUIRE(t.struct_size >= sizeof(mtx_texture_desc), MTX_INVALID_ARGUMENT,
                "mtx_texture_desc.struct_size too small");
    MTX_REQUIRE(t.width >= 1 && t.height >= 1 && t.width <= 16384 && t.height <= 16384, MTX_INVALID_ARGUMENT,
                "texture size must be 1..16384 per edge");
    MTX_REQUIRE(t.storage <= MTX_TEXTURE_RGBA32F, MTX_INVALID_ARGUMENT, "unknown texture storage");
    MTX_REQUIRE(t.flip_y <= 1 && t.uv_channel <= 1 && t.wrap_s <= 2 && t.wrap_t <= 2, MTX_INVALID_ARGUMENT,
                "texture flag out of range");
    MTX_REQUIRE(std::isfinite(t.uv_scale[0]) && std::isfinite(t.uv_scale[1]) && std::isfinite(t.uv_offset[0]) &&
                    std::isfinite(t.uv_offset[1]),
                MTX_INVALID_ARGUMENT, "texture transform must be finite");
    MTX_REQUIRE(t.pixels, MTX_INVALID_ARGUMENT, "texture pixels are NULL");
    const size_t px = (size_t)t.width * t.height;
    size_t &pool = t.storage == MTX_TEXTURE_RGBA32F ? n32 : n8;
    MTX_REQUIRE(pool + px <= 0xffffffffull, MTX_INVALID_ARGUMENT, "texture pool exceeds 2^32 texels per storage");
    T[i] = GpuTexture{t.uv_scale[0], t.uv_scale[1], t.uv_offset[0], t.uv_offset[1], t.width, t.height,
                      t.storage | (t.wrap_s << 1) | (t.wrap_t << 3) | (t.uv_channel << 5), (uint32_t)pool};
    pool += px;
  }
  std::vector<uint32_t> p8(n8);
  std::vector<float4> p32(n32);
  for (uint32_t i = 0; i < n; i++) {
    const mtx_texture_desc &t = D.textures[i];
    for (uint32_t y = 0; y < t.height; y++) {
      const size_t src = (size_t)(t.flip_y ? t.height - 1 - y : y) * t.width;
      const size_t dst = (size_t)T[i].base + (size_t)y * t.width;
      if (t.storage == MTX_TEXTURE_RGBA32F) {
        const float *f = static_cast<const float *>(t.pixels) + src * 4;
        for (size_t k = 0; k < (size_t)t.width * 4; k++)
          MTX_REQUIRE(std::isfinite(f[k]), MTX_INVALID_ARGUMENT, "texture pixels must be finite");
        std::memcpy(&p32[dst], f, (size_t)t.width * sizeof(float4));
      } else {
        std::memcpy(&p8[dst], static_cast<const uint32_t *>(t.pixels) + src, (size_t)t.width * sizeof(uint32_t));
      }
    }
  }
  R.textures.upload(T.data(), T.size(), 0);
  R.texels8.upload_min1(p8.data(), p8.size(), 0);
  R.texels32.upload_min1(p32.data(), p32.size(), 0);
  return n;
}

// Texture-related material fields: resolved alpha mode opaque or mask (blend
// coverage is not built), map alpha usage, slot indices within the texture
// table, finite UV and normal scales. Returns whether some material samples a
// texture.
static bool check_material_maps(const float *mat, uint32_t textureCount) {
  MTX_REQUIRE(mat[MTX_MAT_ALPHA_MODE] != 2.0f, MTX_UNAVAILABLE,
              "blend alpha coverage (stochastic transparency) is not supported by this renderer");
  MTX_REQUIRE(mat[MTX_MAT_ALPHA_MODE] == 0.0f || mat[MTX_MAT_ALPHA_MODE] == 1.0f, MTX_INVALID_ARGUMENT,
              "material alpha mode must be 0 (opaque), 1 (mask) or 2 (blend)");
  const float usage = mat[MTX_MAT_MAP_ALPHA_USAGE];
  MTX_REQUIRE(usage == 0.0f || usage == 2.0f || usage == 3.0f, MTX_INVALID_ARGUMENT,
              "map alpha usage must be 0, 2 or 3");
  for (int k = 0; k < 2; k++)
    MTX_REQUIRE(std::isfinite(mat[MTX_MAT_UV_SCALE + k]) && std::isfinite(mat[MTX_MAT_NORMAL_SCALE + k]),
                MTX_INVALID_ARGUMENT, "material UV and normal scales must be finite");
  bool textured = false;
  for (int k = 0; k < MTX_MAP_COUNT; k++) {
    const float v = mat[MTX_MAT_MAPS + k];
    MTX_REQUIRE(v >= 0.0f && v == std::floor(v) && v <= (float)textureCount, MTX_INVALID_ARGUMENT,
                "material texture slot out of range");
    textured |= v > 0.0f;
  }
  return textured;
}

static std::shared_ptr<RenderSceneObj> create_scene(const mtx_scene_desc &D) {
  MTX_REQUIRE(D.struct_size >= offsetof(mtx_scene_desc, texture_count), MTX_INVALID_ARGUMENT,
              "mtx_scene_desc.struct_size too small");
  MTX_REQUIRE(D.mesh_count == 0 || D.meshes, MTX_INVALID_ARGUMENT, "meshes is NULL");
  MTX_REQUIRE(D.instance_count == 0 || D.instances, MTX_INVALID_ARGUMENT, "instances is NULL");
  MTX_REQUIRE(D.material_count > 0 && D.materials, MTX_INVALID_ARGUMENT, "at least one material is required");
  MTX_REQUIRE(D.light_count == 0 || D.lights, MTX_INVALID_ARGUMENT, "lights is NULL");
  auto R = std::make_shared<RenderSceneObj>();
  const uint32_t bins = D.blas_bins ? D.blas_bins : 8u, tbins = D.tlas_bins ? D.tlas_bins : 64u;
  std::vector<MeshInput> meshes;
  for (uint32_t i = 0; i < D.mesh_count; i++) {
    const mtx_mesh_desc &m = D.meshes[i];
    MTX_REQUIRE(m.struct_size >= offsetof(mtx_mesh_desc, uvs), MTX_INVALID_ARGUMENT,
                "mtx_mesh_desc.struct_size too small");
    MTX_REQUIRE(m.positions && m.normals && m.indices, MTX_INVALID_ARGUMENT, "mesh arrays are NULL");
    for (const float *u : {mesh_uvs(m), mesh_uvs1(m)})
      if (u != nullptr)
        for (size_t k = 0; k < (size_t)m.vertex_count * 2; k++)
          MTX_REQUIRE(std::isfinite(u[k]), MTX_INVALID_ARGUMENT, "mesh UVs must be finite");
    for (size_t k = 0; k < (size_t)m.triangle_count * 3; k++)
      MTX_REQUIRE(m.indices[k] < m.vertex_count, MTX_INVALID_ARGUMENT, "mesh index out of range");
    meshes.push_back(MeshInput{m.positions, m.vertex_count, m.indices, m.triangle_count, false});
  }
  R->blases = build_blases(meshes.data(), meshes.size(), bins, 0);
  for (uint32_t i = 0; i < D.mesh_count; i++) {
    const mtx_mesh_desc &m = D.meshes[i];
    R->normals.emplace_back();
    R->normals.back().upload_min1(m.normals, (size_t)m.vertex_count * 3, 0);
    R->indices.emplace_back();
    R->indices.back().upload_min1(m.indices, (size_t)m.triangle_count * 3, 0);
    R->uvs.emplace_back();
    if (mesh_uvs(m)) R->uvs.back().upload_min1(mesh_uvs(m), (size_t)m.vertex_count * 2, 0);
    R->uvs1.emplace_back();
    if (mesh_uvs1(m)) R->uvs1.back().upload_min1(mesh_uvs1(m), (size_t)m.vertex_count * 2, 0);
  }
  const uint32_t textureCount = upload_textures(*R, D);
  bool textured = false;
  for (uint32_t m = 0; m < D.material_count; m++)
    textured |= check_material_maps(D.materials + (size_t)m * MTX_MATERIAL_FLOATS, textureCount);
  R->inst.assign(D.instances, D.instances + D.instance_count);
  for (auto &d : R->inst) {
    MTX_REQUIRE(d.blas_index < D.mesh_count, MTX_INVALID_ARGUMENT, "instance mesh index out of range");
    uint32_t mat = d.material_word >> 8;
    MTX_REQUIRE(mat < D.material_count, MTX_INVALID_ARGUMENT, "instance material out of range");
    if (D.materials[(size_t)mat * MTX_MATERIAL_FLOATS + MTX_MAT_DOUBLE_SIDED] > 0.5f) d.double_sided = 1;
    // Shadow queries run the coverage test on instances of masked materials.
    d.material_word = (d.material_word & ~kSlotMasked) |
                      (D.materials[(size_t)mat * MTX_MATERIAL_FLOATS + MTX_MAT_ALPHA_MODE] == 1.0f ? kSlotMasked : 0u);
    R->triangles += D.meshes[d.blas_index].triangle_count;
  }
  R->accel = build_accel(R->blases, R->inst.data(), D.instance_count, tbins, D.flags, 0);
  MTX_REQUIRE(R->accel->max_stack() <= (uint32_t)kMaxStack, MTX_INVALID_ARGUMENT, "scene deeper than the traversal stack");
  fill_shade(*R);
  R->materials.upload(D.materials, (size_t)D.material_count * MTX_MATERIAL_FLOATS, 0);
  R->lights.upload_min1(D.lights, (size_t)D.light_count * MTX_LIGHT_FLOATS, 0);
  const bool hasEnv = D.env_rgba && D.env_width && D.env_height;
  EmissiveList E = collect_emissive(D, textured);
  LightTables T = light_tables(D.lights, D.light_count, hasEnv ? D.env_rgba : nullptr, D.env_width, D.env_height,
                               E.weight);
  for (size_t i = 0; i < E.weight.size(); i++) E.tris[3 * i].w = T.emissivePmf[i];
  R->emissive.upload_min1(E.tris.data(), E.tris.size(), 0);
  if (textured) R->emissiveUv.upload_min1(E.uvs.data(), E.uvs.size(), 0);
  R->lightPmf.upload_min1(T.lightPmf.data(), T.lightPmf.size(), 0);
  R->choices.upload_min1(T.choices.data(), T.choices.size(), 0);
  if (hasEnv) {
    R->env.upload(reinterpret_cast<const float4 *>(D.env_rgba), (size_t)D.env_width * D.env_height, 0);
    std::vector<float2> cdf = env_cdf(D.env_rgba, D.env_width, D.env_height);
    R->envCdf.upload(cdf.data(), cdf.size(), 0);
    std::vector<uint16_t> guide = env_guides(cdf, D.env_width, D.env_height);
    if (!guide.empty()) R->envGuide.upload(guide.data(), guide.size(), 0);
  }
  R->lut.upload(reinterpret_cast<const float *>(kEnergyLutBits), 32 * 2177, 0);
  GpuScene &v = R->view;
  v.accel = R->accel->view.ptr;
  v.shade = R->shade.ptr;
  v.materials = R->materials.ptr;
  v.lights = R->lights.ptr;
  v.lightPmf = R->lightPmf.ptr;
  v.choices = R->choices.ptr;
  v.choiceCount = (int32_t)T.choices.size();
  v.lightCount = D.light_count;
  v.envPmf = T.envPmf;
  v.lightWeightTotal = (float)T.total;
  v.env = hasEnv ? R->env.ptr : nullptr;
  v.envCdf = hasEnv ? R->envCdf.ptr : nullptr;
  v.envGuide = hasEnv ? R->envGuide.ptr : nullptr;
  v.emissive = E.weight.empty() ? nullptr : R->emissive.ptr;
  v.emissiveCount = (uint32_t)E.weight.size();
  v.features = 0;
  for (uint32_t m = 0; m < D.material_count; m++) {
    const float *mat = D.materials + (size_t)m * MTX_MATERIAL_FLOATS;
    if (mat[MTX_MAT_FUZZ_WEIGHT] > 0.0f || mat[MTX_MAT_SUBSURFACE_WEIGHT] > 0.0f || mat[MTX_MAT_THIN_WALLED] > 0.5f)
      v.features |= kSceneFull;
    if (mat[MTX_MAT_ALPHA_MODE] == 1.0f) v.features |= kSceneMasked;
  }
  if (textured) {
    v.features |= kSceneFull | kSceneTextured;
    v.textures = R->textures.ptr;
    v.texels8 = R->texels8.ptr;
    v.texels32 = R->texels32.ptr;
    v.emissiveUv = E.weight.empty() ? nullptr : R->emissiveUv.ptr;
  }
  v.envW = hasEnv ? D.env_width : 0;
  v.envH = hasEnv ? D.env_height : 0;
  v.lut = R->lut.ptr;
  MTX_HIP(hipStreamSynchronize(0));
  return R;
}

static uint32_t log2spp(uint32_t pattern, uint32_t maxSamples) {
  double m = pattern == 1 ? (double)maxSamples - 1 : (double)maxSamples;
  if (m < 1) m = 1;
  double l = std::ceil(std::log2(m));
  return (uint32_t)std::min(16.0, std::max(0.0, l));
}

static void set_camera(RenderParams &p, const mtx_camera_desc &c) {
  MTX_REQUIRE(c.struct_size >= sizeof(mtx_camera_desc), MTX_INVALID_ARGUMENT, "mtx_camera_desc.struct_size too small");
  std::memcpy(p.projInv, c.projection_inverse, 64);
  std::memcpy(p.viewInv, c.view_inverse, 64);
  p.aperture = c.aperture_radius;
  p.focus = c.focus_distance;
}

void target_reset(TargetObj &T) {
  MTX_HIP(hipMemsetAsync(T.planes.ptr, 0, T.planes.bytes(), T.stream));
  MTX_HIP(hipMemsetAsync(T.count.ptr, 0, T.count.bytes(), T.stream));
  MTX_HIP(hipMemsetAsync(T.rays.ptr, 0, T.rays.bytes(), T.stream));
  T.samplesDone = 0;
}

// Validated render parameters of a settings + camera pair (whole image).
static RenderParams params_from(const mtx_target_settings &s, const mtx_camera_desc &cam) {
  MTX_REQUIRE(s.struct_size >= sizeof(mtx_target_settings), MTX_INVALID_ARGUMENT,
              "mtx_target_settings.struct_size too small");
  MTX_REQUIRE(s.width > 0 && s.height > 0 && (uint64_t)s.width * s.height <= (1ull << 28), MTX_INVALID_ARGUMENT,
              "invalid target size");
  MTX_REQUIRE(s.sampling_pattern <= 2, MTX_INVALID_ARGUMENT,
              "sampling_pattern must be 0 (sobol), 1 (zsobol) or 2 (random)");
  MTX_REQUIRE(s.bounces <= 50, MTX_INVALID_ARGUMENT, "bounces above 50");
  MTX_REQUIRE(s.kernel <= MTX_KERNEL_MEGAKERNEL, MTX_INVALID_ARGUMENT, "unknown kernel mode");
  MTX_REQUIRE(std::isfinite(s.clamping) && s.clamping >= 0.0f, MTX_INVALID_ARGUMENT, "clamping must be finite");
  RenderParams p{};
  set_camera(p, cam);
  p.width = s.width;
  p.height = s.height;
  p.bounces = s.bounces;
  p.candidates = std::max(1u, s.direct_light_candidates);
  p.pattern = s.sampling_pattern;
  p.seed = s.sequence_seed;
  p.logSPP = log2spp(s.sampling_pattern, s.max_samples);
  p.sampleOffset = s.sample_sequence_offset;
  p.clamping = s.clamping;
  p.pixelBase = 0;
  p.pixelCount = s.width * s.height;
  return p;
}

// Resolved RGBA f32, top row first, from 13 planes + counts (bottom row
// first): the normalize pass of WGPURayTraceOutput.ts:32-92 with
// hdr_resolve_component (compose.frag.wgsl.ts:264-272), evaluated on the FP64
// sums and rounded once. total = max(sample count, 1); hit = albedo plane A.
//   color:  rgb / hit when hit > 0, else / total; a / total
//   albedo: rgb / total, a = clamp(hit / total, 0, 1)
//   normal: rgb / total * 0.5 + 0.5 and a = 1 when |sum|^2 > 0.1, else 0
// Pixels without samples resolve to 0 in every kind.
static void resolve_pixels(const double *pl, const uint32_t *cnt, uint32_t W, uint32_t H, uint32_t kind,
                           float *pixels) {
  const size_t N = (size_t)W * H;
  for (uint32_t y = 0; y < H; y++)
    for (uint32_t x = 0; x < W; x++) {
      size_t i = (size_t)y * W + x, o = ((size_t)(H - 1 - y) * W + x) * 4;
      double r[4] = {0.0, 0.0, 0.0, 0.0};
      const double count = pl[11 * N + i];
      if (kind == MTX_PIXELS_VARIANCE) {
        double n = cnt[i] ? (double)cnt[i] : 1.0;
        double mean = (0.2126 * pl[i] + 0.7152 * pl[N + i] + 0.0722 * pl[2 * N + i]) / n;
        r[0] = std::max(0.0, pl[12 * N + i] / n - mean * mean) / n;
        r[3] = 1.0;
      } else if (count > 0.0) {
        const double total = std::max(count, 1.0), hit = pl[7 * N + i];
        if (kind == MTX_PIXELS_COLOR) {
          const double div = hit > 0.0 ? hit : total;
          for (int k = 0; k < 3; k++) r[k] = pl[(size_t)k * N + i] / div;
          r[3] = pl[3 * N + i] / total;
        } else if (kind == MTX_PIXELS_ALBEDO) {
          for (int k = 0; k < 3; k++) r[k] = pl[(size_t)(4 + k) * N + i] / total;
          r[3] = std::min(std::max(hit / total, 0.0), 1.0);
        } else {
          const double nx = pl[8 * N + i], ny = pl[9 * N + i], nz = pl[10 * N + i];
          if (nx * nx + ny * ny + nz * nz > 0.1) {
            r[0] = nx / total * 0.5 + 0.5;
            r[1] = ny / total * 0.5 + 0.5;
            r[2] = nz / total * 0.5 + 0.5;
            r[3] = 1.0;
          }
        }
      }
      for (int k = 0; k < 4; k++) pixels[o + k] = (float)r[k];
    }
}

// Pixel-major copy of the 13 planes (13 doubles per pixel).
static void interleave_planes(const std::vector<double> &planes, size_t N, double *sums) {
  for (size_t i = 0; i < N; i++)
    for (int k = 0; k < 13; k++) sums[i * 13 + k] = planes[(size_t)k * N + i];
}

std::shared_ptr<TargetObj> create_target(std::shared_ptr<RenderSceneObj> scene, const mtx_target_settings &s,
                                         const mtx_camera_desc &cam) {
  auto T = std::make_shared<TargetObj>();
  T->scene = scene;
  T->params = params_from(s, cam);
  T->kernel = s.kernel;
  if (s.max_paths_in_flight) T->wfConfig.maxPaths = s.max_paths_in_flight;
  const size_t N = (size_t)s.width * s.height;
  T->planes.alloc(N * 13);
  T->count.alloc(N);
  T->rays.alloc(1);
  target_reset(*T);
  MTX_HIP(hipStreamSynchronize(T->stream));
  return T;
}

void target_render(TargetObj &T, uint32_t samples, hipStream_t s, uint64_t *work = nullptr,
                   LaunchTimer *timer = nullptr) {
  RenderParams p = T.params;
  p.samples = samples;
  Accumulators acc{T.planes.ptr, T.count.ptr, T.rays.ptr, p.width * p.height, work};
  acc.targets = T.useTargets ? T.targets.ptr : nullptr;
  if (T.kernel == MTX_KERNEL_MEGAKERNEL)
    launch_render(T.scene->view, p, acc, s, timer);
  else
    launch_render_wavefront(T.scene->view, p, acc, T.wf, T.wfConfig, s, timer, work ? nullptr : &T.graph);
  if (!s && T.graph.own && T.kernel != MTX_KERNEL_MEGAKERNEL) s = T.graph.own;
  T.stream = s;
  T.samplesDone += samples;
}

// One thread per tile (offlineAdaptive.compute.ts:29-106), in FP64.
__global__ void k_plan_adaptive(const double *__restrict__ planes, const uint32_t *__restrict__ count, uint32_t W,
                                uint32_t H, uint32_t N, uint32_t tile, uint32_t cols, uint32_t rows, uint32_t minS,
                                uint32_t maxS, double absErr, double relErr, double conf, float *__restrict__ targets,
                                double *__restrict__ tileVar, uint32_t *__restrict__ tileTarget) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cols * rows) return;
  const uint32_t tx = t % cols, ty = t / cols;
  const uint32_t x0 = tx * tile, y0 = ty * tile, x1 = min(x0 + tile, W), y1 = min(y0 + tile, H);
  double meanSum = 0.0, maxVar = 0.0;
  uint32_t n = 0;
  for (uint32_t y = y0; y < y1; y++)
    for (uint32_t x = x0; x < x1; x++) {
      const uint32_t i = y * W + x;
      const double s = (double)count[i];
      if (s <= 0.0) continue;
      const double mean = (0.2126 * planes[i] + 0.7152 * planes[N + i] + 0.0722 * planes[2 * N + i]) / s;
      const double second = planes[12 * (size_t)N + i] / s;
      double var = fmax(second - mean * mean, 0.0);
      if (s > 1.0) var = var * s / (s - 1.0);
      meanSum += mean;
      maxVar = fmax(maxVar, var);
      n++;
    }
  const double mean = meanSum / (double)max(n, 1u);
  const double tol = absErr + relErr * fabs(mean);
  const double required = ceil(conf * conf * maxVar / fmax(tol * tol, 1e-16));
  const double tgt = fmin((double)maxS, fmax((double)minS, required));
  tileVar[t] = maxVar;
  tileTarget[t] = (uint32_t)tgt;
  for (uint32_t y = y0; y < y1; y++)
    for (uint32_t x = x0; x < x1; x++) targets[y * W + x] = (float)tgt;
}

}  // namespace mtx

using namespace mtx;

#define API_TRY try {
#define API_CATCH(err, cap)                       \
  }                                               \
  catch (const mtx::Error &e) {                   \
    set_error(err, cap, e.what());                \
    return e.code;                                \
  }                                               \
  catch (const std::exception &e) {               \
    set_error(err, cap, e.what());                \
    return MTX_ERROR;                             \
  }

extern "C" {

int32_t mtx_scene_create(const mtx_scene_desc *desc, mtx_handle *out, char *err, size_t cap) {
  API_TRY
  MTX_REQUIRE(desc && out, MTX_INVALID_ARGUMENT, "desc/out is NULL");
  auto R = create_scene(*desc);
  std::lock_guard<std::mutex> g(gRenderMutex);
  uint64_t h = registry().next();
  gScenes[h] = R;
  *out = h;
  return MTX_OK;
  API_CATCH(err, cap)
}

int32_t mtx_scene_set_transforms(mtx_handle scene, const double *transforms, uint32_t count, const uint32_t *dirty,
                                 uint32_t dirtyCount, char *err, size_t cap) {
  API_TRY
  auto R = get_scene(scene);
  MTX_REQUIRE(transforms, MTX_INVALID_ARGUMENT, "transforms is NULL");
  MTX_REQUIRE(count == R->inst.size(), MTX_INVALID_ARGUMENT, "instance count mismatch");
  for (uint32_t i = 0; i < count; i++) std::memcpy(R->inst[i].transform, transforms + (size_t)i * 16, 128);
  refit_accel(*R->accel, R->inst.data(), count, dirty, dirtyCount, 0);
  fill_shade(*R);
  MTX_HIP(hipStreamSynchronize(0));
  return MTX_OK;
  API_CATCH(err, cap)
}

int32_t mtx_scene_update_mesh(mtx_handle scene, uint32_t mesh, const float *positions, const float *normals,
                              uint32_t vertex_count, uint32_t *rebuild_advised, char *err, size_t cap) {
  API_TRY
  auto R = get_scene(scene);
  MTX_REQUIRE(positions && normals, MTX_INVALID_ARGUMENT, "positions/normals are NULL");
  MTX_REQUIRE(mesh < R->blases.size(), MTX_INVALID_ARGUMENT, "mesh index out of range");
  Blas &B = *R->blases[mesh];
  MTX_REQUIRE(vertex_count == B.vertexInput, MTX_INVALID_ARGUMENT, "vertex count differs from the commit's");
  DevBuf<float> pos;
  pos.upload(positions, (size_t)vertex_count * 3, 0);
  R->normals[mesh].upload(normals, (size_t)vertex_count * 3, 0);
  const bool advise = refit_blas(B, pos.ptr, R->indices[mesh].ptr, 0);
  refit_accel_blas(*R->accel, mesh, 0);
  if (rebuild_advised) *rebuild_advised = advise ? 1u : 0u;
  return MTX_OK;
  API_CATCH(err, cap)
}

int32_t mtx_scene_info(mtx_handle scene, uint64_t info[4]) {
  API_TRY
  auto R = get_scene(scene);
  MTX_REQUIRE(info, MTX_INVALID_ARGUMENT, "info is NULL");
  info[0] = R->triangles;
  info[1] = R->inst.size();
  info[2] = R->blases.size();
  info[3] = (R->bytes() + (1u << 20) - 1) >> 20;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_scene_destroy(mtx_handle scene) {
  std::lock_guard<std::mutex> g(gRenderMutex);
  return gScenes.erase(scene) ? MTX_OK : MTX_INVALID_HANDLE;
}

int32_t mtx_target_create(mtx_handle scene, const mtx_target_settings *settings, const mtx_camera_desc *camera,
                          mtx_handle *out, char *err, size_t cap) {
  API_TRY
  MTX_REQUIRE(settings && camera && out, MTX_INVALID_ARGUMENT, "settings/camera/out is NULL");
  auto T = create_target(get_scene(scene), *settings, *camera);
  std::lock_guard<std::mutex> g(gRenderMutex);
  uint64_t h = registry().next();
  gTargets[h] = T;
  *out = h;
  return MTX_OK;
  API_CATCH(err, cap)
}

int32_t mtx_target_render(mtx_handle target, uint32_t samples, void *stream) {
  API_TRY
  auto T = get_target(target);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  target_render(*T, samples, (hipStream_t)stream);
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_render_counted(mtx_handle target, uint32_t samples, uint64_t work[MTX_WORK_COUNTERS],
                                  void *stream) {
  API_TRY
  MTX_REQUIRE(work, MTX_INVALID_ARGUMENT, "work is NULL");
  auto T = get_target(target);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  hipStream_t s = (hipStream_t)stream;
  static_assert(kWorkSlots == MTX_WORK_COUNTERS && kWorkGroup == MTX_WORK_GROUP, "work counter layout");
  DevBuf<uint64_t> w(kWorkSlots);
  MTX_HIP(hipMemsetAsync(w.ptr, 0, w.bytes(), s));
  target_render(*T, samples, s, w.ptr);
  w.download(work, kWorkSlots, s);
  MTX_HIP(hipStreamSynchronize(s));
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_render_timed(mtx_handle target, uint32_t samples, double ms[MTX_TIMER_CLASSES], void *stream) {
  API_TRY
  MTX_REQUIRE(ms, MTX_INVALID_ARGUMENT, "ms is NULL");
  auto T = get_target(target);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  LaunchTimer timer;
  timer.stream = (hipStream_t)stream;
  target_render(*T, samples, (hipStream_t)stream, nullptr, &timer);
  timer.collect(ms);
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_synchronize(mtx_handle target) {
  API_TRY
  auto T = get_target(target);
  MTX_HIP(hipStreamSynchronize(T->stream));
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_read_accumulators(mtx_handle target, double *sums, size_t capSums, uint32_t *counts,
                                     size_t capCounts) {
  API_TRY
  auto T = get_target(target);
  const size_t N = (size_t)T->params.width * T->params.height;
  MTX_HIP(hipStreamSynchronize(T->stream));
  if (sums) {
    if (capSums < N * 13) return MTX_BUFFER_TOO_SMALL;
    interleave_planes(T->planes.to_host(N * 13, T->stream), N, sums);
  }
  if (counts) {
    if (capCounts < N) return MTX_BUFFER_TOO_SMALL;
    T->count.download(counts, N, T->stream);
    MTX_HIP(hipStreamSynchronize(T->stream));
  }
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_read_pixels(mtx_handle target, uint32_t kind, float *pixels, size_t cap, size_t *written) {
  API_TRY
  auto T = get_target(target);
  const uint32_t W = T->params.width, H = T->params.height;
  const size_t N = (size_t)W * H;
  if (written) *written = N * 4;
  if (!pixels) return MTX_OK;
  if (cap < N * 4) return MTX_BUFFER_TOO_SMALL;
  MTX_REQUIRE(kind <= MTX_PIXELS_VARIANCE, MTX_INVALID_ARGUMENT, "unknown pixel kind");
  MTX_HIP(hipStreamSynchronize(T->stream));
  std::vector<double> pl = T->planes.to_host(N * 13, T->stream);
  std::vector<uint32_t> cnt = T->count.to_host(N, T->stream);
  resolve_pixels(pl.data(), cnt.data(), W, H, kind, pixels);
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_stats(mtx_handle target, uint64_t stats[2]) {
  API_TRY
  auto T = get_target(target);
  MTX_REQUIRE(stats, MTX_INVALID_ARGUMENT, "stats is NULL");
  MTX_HIP(hipStreamSynchronize(T->stream));
  T->rays.download(&stats[0], 1, T->stream);
  MTX_HIP(hipStreamSynchronize(T->stream));
  stats[1] = T->samplesDone;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_set_camera(mtx_handle target, const mtx_camera_desc *camera) {
  API_TRY
  auto T = get_target(target);
  MTX_REQUIRE(camera, MTX_INVALID_ARGUMENT, "camera is NULL");
  set_camera(T->params, *camera);
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_set_sample_offset(mtx_handle target, uint32_t offset) {
  API_TRY
  auto T = get_target(target);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  T->params.sampleOffset = offset;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_reset(mtx_handle target) {
  API_TRY
  auto T = get_target(target);
  target_reset(*T);
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_set_tuning(mtx_handle target, const mtx_wavefront_tuning *t) {
  API_TRY
  MTX_REQUIRE(t && t->struct_size >= sizeof(mtx_wavefront_tuning), MTX_INVALID_ARGUMENT, "bad tuning struct");
  MTX_REQUIRE(t->trace_kernel <= 1, MTX_INVALID_ARGUMENT, "trace_kernel must be 0 or 1");
  MTX_REQUIRE(t->lds_stack == 16 || t->lds_stack == 24 || t->lds_stack == 32, MTX_INVALID_ARGUMENT,
              "lds_stack must be 16, 24 or 32");
  MTX_REQUIRE(t->lds_stack_closest == 0 || t->lds_stack_closest == 16 || t->lds_stack_closest == 24 ||
                  t->lds_stack_closest == 32,
              MTX_INVALID_ARGUMENT, "lds_stack_closest must be 0, 16, 24 or 32");
  MTX_REQUIRE(t->trace_oversubscription >= 1 && t->trace_oversubscription <= 16 && t->shade_oversubscription >= 1 &&
                  t->shade_oversubscription <= 16,
              MTX_INVALID_ARGUMENT, "oversubscription must be in 1..16");
  MTX_REQUIRE(t->shade_waves_per_simd >= 1 && t->shade_waves_per_simd <= 4, MTX_INVALID_ARGUMENT,
              "shade_waves_per_simd must be in 1..4");
  auto T = get_target(target);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  if (t->max_paths_in_flight) T->wfConfig.maxPaths = t->max_paths_in_flight;
  T->wfConfig.walker = t->trace_kernel == 1;
  T->wfConfig.ldsStack = (int)t->lds_stack;
  T->wfConfig.traceOversub = (int)t->trace_oversubscription;
  T->wfConfig.shadeOversub = (int)t->shade_oversubscription;
  T->wfConfig.shadeWaves = (int)t->shade_waves_per_simd;
  T->wfConfig.graphs = t->use_graphs != 0;
  T->wfConfig.tailPaths = t->tail_paths;
  T->wfConfig.queueSort = t->queue_sort != 0;
  T->wfConfig.ldsStackClosest = (int)t->lds_stack_closest;
  T->graph.reset();
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_get_tuning(mtx_handle target, mtx_wavefront_tuning *t) {
  API_TRY
  MTX_REQUIRE(t, MTX_INVALID_ARGUMENT, "tuning is NULL");
  auto T = get_target(target);
  t->struct_size = sizeof(mtx_wavefront_tuning);
  t->max_paths_in_flight = T->wfConfig.maxPaths;
  t->trace_kernel = T->wfConfig.walker ? 1u : 0u;
  t->lds_stack = (uint32_t)T->wfConfig.ldsStack;
  t->trace_oversubscription = (uint32_t)T->wfConfig.traceOversub;
  t->shade_oversubscription = (uint32_t)T->wfConfig.shadeOversub;
  t->shade_waves_per_simd = (uint32_t)T->wfConfig.shadeWaves;
  t->use_graphs = T->wfConfig.graphs ? 1u : 0u;
  t->graph_replays = (uint32_t)T->graph.replays;
  t->graph_captures = (uint32_t)T->graph.captures;
  t->tail_paths = T->wfConfig.tailPaths;
  t->queue_sort = T->wfConfig.queueSort ? 1u : 0u;
  t->lds_stack_closest = (uint32_t)T->wfConfig.ldsStackClosest;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_set_sample_targets(mtx_handle target, const float *targets, size_t count) {
  API_TRY
  auto T = get_target(target);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  const size_t N = (size_t)T->params.width * T->params.height;
  if (!targets) {
    T->useTargets = false;
    return MTX_OK;
  }
  MTX_REQUIRE(count == N, MTX_INVALID_ARGUMENT, "targets must hold width*height values");
  T->targets.upload(targets, N, T->stream);
  MTX_HIP(hipStreamSynchronize(T->stream));
  T->useTargets = true;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_plan_adaptive(mtx_handle target, const mtx_adaptive_options *o, double plan[2],
                                 double *tileVariance, size_t tileCap) {
  API_TRY
  MTX_REQUIRE(o && o->struct_size >= sizeof(mtx_adaptive_options) && plan, MTX_INVALID_ARGUMENT, "bad arguments");
  auto T = get_target(target);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  // normalizeOfflineAdaptiveSamplingOptions (OfflineAdaptiveSampling.ts:75-108)
  const uint32_t maxS = std::max(1u, o->max_samples);
  const uint32_t minS = std::min(maxS, o->min_samples ? o->min_samples : std::min(maxS, 16u));
  const uint32_t tile = o->tile_size ? o->tile_size : 16u;
  const double absErr = std::max(o->absolute_error > 0 ? o->absolute_error : 0.002, 1e-8);
  const double relErr = std::max(o->relative_error > 0 ? o->relative_error : 0.03, 0.0);
  const double conf = std::max(o->confidence > 0 ? o->confidence : 1.96, 0.01);
  const uint32_t W = T->params.width, H = T->params.height, N = W * H;
  const uint32_t cols = (W + tile - 1) / tile, rows = (H + tile - 1) / tile;
  if (T->targets.count < N) T->targets.alloc(N);
  DevBuf<double> tv(cols * rows);
  DevBuf<uint32_t> tt(cols * rows);
  hipStream_t s = T->stream;
  hipLaunchKernelGGL(k_plan_adaptive, dim3((cols * rows + 127) / 128), dim3(128), 0, s, T->planes.ptr, T->count.ptr, W,
                     H, N, tile, cols, rows, minS, maxS, absErr, relErr, conf, T->targets.ptr, tv.ptr, tt.ptr);
  MTX_HIP(hipGetLastError());
  std::vector<uint32_t> tiles = tt.to_host(cols * rows, s);
  std::vector<double> vars = tv.to_host(cols * rows, s);
  double mx = 0, sum = 0;
  for (uint32_t t = 0; t < cols * rows; t++) {
    const uint32_t tx = t % cols, ty = t / cols;
    const uint32_t w = std::min(tile, W - tx * tile), h = std::min(tile, H - ty * tile);
    mx = std::max(mx, (double)tiles[t]);
    sum += (double)tiles[t] * w * h;
  }
  plan[0] = mx;
  plan[1] = sum / N;
  if (tileVariance) {
    if (tileCap < (size_t)cols * rows) return MTX_BUFFER_TOO_SMALL;
    std::memcpy(tileVariance, vars.data(), vars.size() * sizeof(double));
  }
  T->useTargets = true;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_display_settings_default(mtx_display_settings *s) {
  if (!s) return MTX_INVALID_ARGUMENT;
  std::memset(s, 0, sizeof *s);
  s->struct_size = sizeof *s;
  s->tone_mapping = MTX_TONE_AGX;
  s->exposure = 1.0f;
  s->contrast = 1.0f;
  s->saturation = 1.0f;
  return MTX_OK;
}

int32_t mtx_jnlm_options_default(mtx_jnlm_options *o) {
  if (!o) return MTX_INVALID_ARGUMENT;
  std::memset(o, 0, sizeof *o);
  o->struct_size = sizeof *o;
  o->strength = 1.0f;
  o->search_radius = 5;
  o->patch_radius = 1;
  o->albedo_bandwidth = 0.4f;
  o->normal_bandwidth = 0.05f;
  o->iterations = 2;
  return MTX_OK;
}

}  // extern "C"

namespace mtx {

static DisplayParams display_params(const mtx_display_settings *d) {
  mtx_display_settings def;
  mtx_display_settings_default(&def);
  if (!d) d = &def;
  MTX_REQUIRE(d->struct_size == sizeof(mtx_display_settings), MTX_INVALID_ARGUMENT,
              "mtx_display_settings.struct_size must be sizeof(mtx_display_settings)");
  MTX_REQUIRE(d->tone_mapping <= MTX_TONE_LINEAR, MTX_INVALID_ARGUMENT, "unknown tone mapping");
  for (float v : {d->exposure, d->contrast, d->saturation, d->temperature, d->tint, d->vignette})
    MTX_REQUIRE(std::isfinite(v), MTX_INVALID_ARGUMENT, "display parameters must be finite");
  return DisplayParams{d->tone_mapping, d->exposure, d->contrast, d->saturation, d->temperature, d->tint, d->vignette,
                       d->gamma22 ? 1u : 0u};
}

// A WGSL constant written as (value).toFixed(6), parsed as f32.
static float fixed6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.6f", v);
  return std::strtof(buf, nullptr);
}

static JnlmParams jnlm_params(const mtx_jnlm_options *o, uint32_t W, uint32_t H) {
  mtx_jnlm_options def;
  mtx_jnlm_options_default(&def);
  if (!o) o = &def;
  MTX_REQUIRE(o->struct_size == sizeof(mtx_jnlm_options), MTX_INVALID_ARGUMENT,
              "mtx_jnlm_options.struct_size must be sizeof(mtx_jnlm_options)");
  MTX_REQUIRE(std::isfinite(o->strength) && o->strength > 0.0f, MTX_INVALID_ARGUMENT, "strength must be finite and positive");
  MTX_REQUIRE(o->search_radius >= 1 && o->search_radius <= 16, MTX_INVALID_ARGUMENT, "searchRadius must be 1..16");
  MTX_REQUIRE(o->patch_radius <= 3, MTX_INVALID_ARGUMENT, "patchRadius must be 0..3");
  MTX_REQUIRE(o->iterations >= 1 && o->iterations <= 4, MTX_INVALID_ARGUMENT, "iterations must be 1..4");
  for (float b : {o->albedo_bandwidth, o->normal_bandwidth})
    MTX_REQUIRE(std::isfinite(b) && b >= 0.0001f && b <= 10.0f, MTX_INVALID_ARGUMENT, "bandwidths must be 0.0001..10");
  JnlmParams J{};
  J.width = W;
  J.height = H;
  J.colorSensitivity = (float)(0.6 * std::max((double)o->strength, 1e-3));
  J.searchRadius = o->search_radius;
  J.patchRadius = o->patch_radius;
  J.iterations = o->iterations;
  const double ab = o->albedo_bandwidth, nb = o->normal_bandwidth;
  J.invAlbedoBandwidth2 = fixed6(1.0 / (ab * ab));
  J.invCoverageBandwidth2 = fixed6(1.0 / (0.1 * 0.1));  // JNLM_COVERAGE_BANDWIDTH
  J.invNormalBandwidth2 = fixed6(1.0 / (nb * nb));
  return J;
}

}  // namespace mtx

extern "C" {

int32_t mtx_target_compose(mtx_handle target, const mtx_display_settings *display, uint8_t *rgba8, size_t capacity,
                           size_t *written) {
  API_TRY
  auto T = get_target(target);
  const uint32_t W = T->params.width, H = T->params.height;
  const size_t bytes = (size_t)W * H * 4;
  if (written) *written = bytes;
  if (!rgba8) return MTX_OK;
  if (capacity < bytes) return MTX_BUFFER_TOO_SMALL;
  const DisplayParams P = display_params(display);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  DevBuf<uint8_t> out(bytes);
  display_compose(T->planes.ptr, W, H, P, out.ptr, nullptr, nullptr, nullptr, T->stream);
  out.download(rgba8, bytes, T->stream);
  MTX_HIP(hipStreamSynchronize(T->stream));
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_target_denoise_jnlm(mtx_handle target, const mtx_display_settings *display, const mtx_jnlm_options *options,
                                float *linear_rgba, uint8_t *rgba8, size_t pixels, char *err, size_t cap) {
  API_TRY
  auto T = get_target(target);
  const uint32_t W = T->params.width, H = T->params.height;
  const size_t N = (size_t)W * H;
  MTX_REQUIRE(linear_rgba || rgba8, MTX_INVALID_ARGUMENT, "no output buffer");
  if (pixels < N) return MTX_BUFFER_TOO_SMALL;
  const DisplayParams P = display_params(display);
  const JnlmParams J = jnlm_params(options, W, H);
  std::unique_lock<std::mutex> lk(T->busy, std::try_to_lock);
  if (!lk.owns_lock()) return MTX_BUSY;
  DevBuf<float4> lin;
  DevBuf<uint8_t> enc;
  if (linear_rgba) lin.alloc(N);
  if (rgba8) enc.alloc(N * 4);
  jnlm_denoise(T->planes.ptr, W, H, P, J, T->jnlm, lin.ptr, enc.ptr, T->stream);
  if (linear_rgba) lin.download(reinterpret_cast<float4 *>(linear_rgba), N, T->stream);
  if (rgba8) enc.download(rgba8, N * 4, T->stream);
  MTX_HIP(hipStreamSynchronize(T->stream));
  return MTX_OK;
  API_CATCH(err, cap)
}

int32_t mtx_target_destroy(mtx_handle target) {
  std::lock_guard<std::mutex> g(gRenderMutex);
  return gTargets.erase(target) ? MTX_OK : MTX_INVALID_HANDLE;
}

}  // extern "C"

// ====================================================================== server
namespace mtx {

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ServerJob {
  std::shared_ptr<RenderSceneObj> scene;
  RenderParams params{};
  DevBuf<double> planes;
  DevBuf<uint32_t> count;
  uint32_t requested = 0, submitted = 0, completed = 0;
  uint32_t state = MTX_JOB_QUEUED;
  bool cancel = false, admitted = false;
  double submitMs = 0, startMs = 0, completeMs = 0;
  size_t bytes() const { return (size_t)params.pixelCount * (13 * sizeof(double) + sizeof(uint32_t)); }
};

struct ServerObj {
  std::mutex m;
  uint32_t maxPaths = 1u << 24, maxJobs = kMaxJobsPerRound, maxSamplesPerRound = 64;
  uint64_t budget = 200ull << 30, used = 0, rounds = 0;
  std::deque<uint64_t> queued;   // submission order
  std::vector<uint64_t> running; // admission order
  std::unordered_map<uint64_t, std::shared_ptr<ServerJob>> jobs;
  WfBuffers::Owner wf;
  WfConfig cfg;
  WfGraph graph;  // the round's launch sequence, replayed while the round's shape repeats
  DevBuf<JobSlot> slots;
  DevBuf<uint64_t> rays;
};

static std::unordered_map<uint64_t, std::shared_ptr<ServerObj>> gServers;

static std::shared_ptr<ServerObj> get_server(uint64_t h) {
  std::lock_guard<std::mutex> g(gRenderMutex);
  auto it = gServers.find(h);
  if (it == gServers.end()) throw Error(MTX_INVALID_HANDLE, "unknown server handle");
  return it->second;
}
static std::shared_ptr<ServerJob> get_job(ServerObj &S, uint64_t h) {
  auto it = S.jobs.find(h);
  if (it == S.jobs.end()) throw Error(MTX_INVALID_HANDLE, "unknown job handle");
  return it->second;
}

// One round: admission, slicing, one multi-job generation.
static uint32_t server_round(ServerObj &S, hipStream_t s, bool &idle) {
  // Admission in arrival order while the accumulators fit the budget.
  while (!S.queued.empty()) {
    auto J = S.jobs.at(S.queued.front());
    if (J->cancel) { J->state = MTX_JOB_CANCELLED; J->completeMs = now_ms(); S.queued.pop_front(); continue; }
    if (S.used + J->bytes() > S.budget) break;  // waits for running or unreleased jobs to free memory
    J->planes.alloc((size_t)J->params.pixelCount * 13);
    J->count.alloc(J->params.pixelCount);
    MTX_HIP(hipMemsetAsync(J->planes.ptr, 0, J->planes.bytes(), s));
    MTX_HIP(hipMemsetAsync(J->count.ptr, 0, J->count.bytes(), s));
    J->admitted = true;
    S.used += J->bytes();
    S.running.push_back(S.queued.front());
    S.queued.pop_front();
  }
  // Slices of the round's path budget, oldest job first.
  std::vector<JobSlot> slots;
  std::vector<std::pair<std::shared_ptr<ServerJob>, uint32_t>> parts;
  uint32_t paths = 0, pixels = 0;
  int eventLimit = 0;
  for (uint64_t h : S.running) {
    if (slots.size() >= S.maxJobs) break;
    auto J = S.jobs.at(h);
    if (J->state != MTX_JOB_QUEUED && J->state != MTX_JOB_RUNNING) continue;
    if (J->cancel) continue;
    const uint32_t left = J->requested - J->completed;
    const uint32_t fit = (S.maxPaths - paths) / J->params.pixelCount;
    const uint32_t g = std::min(std::min(left, fit), S.maxSamplesPerRound);
    if (g == 0) {
      if (fit == 0) break;  // round full: later jobs wait (arrival order is kept)
      continue;
    }
    JobSlot js{};
    js.scene = J->scene->view;
    js.params = J->params;
    js.params.samples = g;
    js.acc = Accumulators{J->planes.ptr, J->count.ptr, S.rays.ptr, J->params.pixelCount, nullptr};
    js.pathBase = paths;
    js.pixelPrefix = pixels;
    slots.push_back(js);
    parts.emplace_back(J, g);
    paths += g * J->params.pixelCount;
    pixels += J->params.pixelCount;
    eventLimit = std::max(eventLimit, std::min((int)J->params.bounces * 3, 50));
    if (J->state == MTX_JOB_QUEUED) { J->state = MTX_JOB_RUNNING; J->startMs = now_ms(); }
  }
  idle = slots.empty();
  uint32_t finished = 0;
  if (!idle) {
    bool full = false;
    bool masked = false;
    for (const JobSlot &js : slots) {
      full = full || needs_full_shading(js.scene, js.params);
      masked = masked || (js.scene.features & kSceneMasked) != 0u;
    }
    // A fixed-size slot table keeps its address, so the captured round
    // replays whenever the next round has the same shape.
    if (S.slots.count < (size_t)kMaxJobsPerRound) S.slots.alloc(kMaxJobsPerRound);
    // The null stream cannot be captured: rounds then run on the server's
    // own blocking stream (ordered with the null stream's admission work).
    hipStream_t rs = s;
    if (!rs && S.cfg.graphs) {
      if (!S.graph.own) MTX_HIP(hipStreamCreate(&S.graph.own));
      rs = S.graph.own;
    }
    S.slots.upload(slots.data(), slots.size(), rs);
    launch_generation_multi(S.slots.ptr, (uint32_t)slots.size(), paths, pixels, eventLimit, S.rays.ptr, S.wf, S.cfg,
                            rs, full, masked, nullptr, &S.graph);
    MTX_HIP(hipStreamSynchronize(rs));
    const double t = now_ms();
    for (auto &pj : parts) {
      pj.first->submitted += pj.second;
      pj.first->completed += pj.second;
      if (pj.first->completed >= pj.first->requested) {
        pj.first->state = MTX_JOB_COMPLETE;
        pj.first->completeMs = t;
        finished++;
      }
    }
    S.rounds++;
  }
  // Cancelled jobs end between rounds: running ones, and admitted ones that
  // never got a slot (max_jobs or the round's path budget ran out).
  for (uint64_t h : S.running) {
    auto J = S.jobs.at(h);
    if (J->cancel && (J->state == MTX_JOB_RUNNING || J->state == MTX_JOB_QUEUED)) {
      J->state = MTX_JOB_CANCELLED;
      J->completeMs = now_ms();
    }
  }
  std::vector<uint64_t> keep;
  for (uint64_t h : S.running) {
    auto st = S.jobs.at(h)->state;
    if (st == MTX_JOB_QUEUED || st == MTX_JOB_RUNNING) keep.push_back(h);
  }
  S.running.swap(keep);
  return finished;
}

}  // namespace mtx

extern "C" {

int32_t mtx_server_create(const mtx_server_config *cfg, mtx_handle *out, char *err, size_t cap) {
  API_TRY
  MTX_REQUIRE(cfg && out, MTX_INVALID_ARGUMENT, "config/out is NULL");
  MTX_REQUIRE(cfg->struct_size >= sizeof(mtx_server_config), MTX_INVALID_ARGUMENT, "mtx_server_config.struct_size too small");
  auto S = std::make_shared<ServerObj>();
  if (cfg->max_paths_per_round) S->maxPaths = cfg->max_paths_per_round;
  if (cfg->max_jobs_per_round) S->maxJobs = std::min<uint32_t>(cfg->max_jobs_per_round, kMaxJobsPerRound);
  if (cfg->max_samples_per_job_round) S->maxSamplesPerRound = cfg->max_samples_per_job_round;
  if (cfg->hbm_budget_bytes) S->budget = cfg->hbm_budget_bytes;
  S->rays.alloc(1);
  MTX_HIP(hipMemset(S->rays.ptr, 0, 8));
  std::lock_guard<std::mutex> g(gRenderMutex);
  uint64_t h = registry().next();
  gServers[h] = S;
  *out = h;
  return MTX_OK;
  API_CATCH(err, cap)
}

int32_t mtx_server_submit(mtx_handle server, mtx_handle scene, const mtx_target_settings *settings,
                          const mtx_camera_desc *camera, uint32_t samples, mtx_handle *job, char *err, size_t cap) {
  API_TRY
  MTX_REQUIRE(settings && camera && job, MTX_INVALID_ARGUMENT, "settings/camera/job is NULL");
  MTX_REQUIRE(samples >= 1 && samples <= 16384, MTX_INVALID_ARGUMENT, "samples must be in 1..16384");
  auto S = get_server(server);
  auto J = std::make_shared<ServerJob>();
  J->scene = get_scene(scene);
  J->params = params_from(*settings, *camera);
  std::lock_guard<std::mutex> g(S->m);
  MTX_REQUIRE(J->params.pixelCount <= S->maxPaths, MTX_INVALID_ARGUMENT, "image larger than the round's path budget");
  MTX_REQUIRE(J->bytes() <= S->budget, MTX_INVALID_ARGUMENT, "job accumulators exceed the HBM budget");
  J->requested = samples;
  J->submitMs = now_ms();
  uint64_t h = registry().next();
  S->jobs[h] = J;
  S->queued.push_back(h);
  *job = h;
  return MTX_OK;
  API_CATCH(err, cap)
}

int32_t mtx_server_step(mtx_handle server, void *stream, uint32_t *completed, uint32_t *idle) {
  API_TRY
  auto S = get_server(server);
  std::lock_guard<std::mutex> g(S->m);
  bool isIdle = true;
  uint32_t n = server_round(*S, (hipStream_t)stream, isIdle);
  if (completed) *completed = n;
  if (idle) *idle = isIdle ? 1u : 0u;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_server_job_info(mtx_handle server, mtx_handle job, mtx_served_job_info *info) {
  API_TRY
  MTX_REQUIRE(info, MTX_INVALID_ARGUMENT, "info is NULL");
  auto S = get_server(server);
  std::lock_guard<std::mutex> g(S->m);
  auto J = get_job(*S, job);
  info->struct_size = sizeof(mtx_served_job_info);
  info->state = J->state;
  info->requested_samples = J->requested;
  info->submitted_samples = J->submitted;
  info->completed_samples = J->completed;
  info->base_samples = 0;
  info->submit_ms = J->submitMs;
  info->start_ms = J->startMs;
  info->complete_ms = J->completeMs;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_server_job_read_pixels(mtx_handle server, mtx_handle job, uint32_t kind, float *pixels, size_t cap,
                                   size_t *written) {
  API_TRY
  auto S = get_server(server);
  std::lock_guard<std::mutex> g(S->m);
  auto J = get_job(*S, job);
  const uint32_t W = J->params.width, H = J->params.height;
  const size_t N = (size_t)W * H;
  if (written) *written = N * 4;
  if (!pixels) return MTX_OK;
  if (cap < N * 4) return MTX_BUFFER_TOO_SMALL;
  MTX_REQUIRE(kind <= MTX_PIXELS_VARIANCE, MTX_INVALID_ARGUMENT, "unknown pixel kind");
  if (!J->admitted || !J->planes.ptr) return MTX_INVALID_STATE;
  std::vector<double> pl = J->planes.to_host(N * 13, 0);
  std::vector<uint32_t> cnt = J->count.to_host(N, 0);
  resolve_pixels(pl.data(), cnt.data(), W, H, kind, pixels);
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_server_job_read_accumulators(mtx_handle server, mtx_handle job, double *sums, size_t capSums,
                                         uint32_t *counts, size_t capCounts) {
  API_TRY
  auto S = get_server(server);
  std::lock_guard<std::mutex> g(S->m);
  auto J = get_job(*S, job);
  const size_t N = J->params.pixelCount;
  if (!J->admitted || !J->planes.ptr) return MTX_INVALID_STATE;
  if (sums) {
    if (capSums < N * 13) return MTX_BUFFER_TOO_SMALL;
    interleave_planes(J->planes.to_host(N * 13, 0), N, sums);
  }
  if (counts) {
    if (capCounts < N) return MTX_BUFFER_TOO_SMALL;
    J->count.download(counts, N, 0);
    MTX_HIP(hipStreamSynchronize(0));
  }
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_server_job_cancel(mtx_handle server, mtx_handle job) {
  API_TRY
  auto S = get_server(server);
  std::lock_guard<std::mutex> g(S->m);
  auto J = get_job(*S, job);
  if (J->state == MTX_JOB_QUEUED || J->state == MTX_JOB_RUNNING) J->cancel = true;
  // Queued work stops at once (mutracer-context.h:155-157), admitted or not:
  // no sample of it is on the device between rounds. Running jobs end when
  // the next round returns.
  if (J->state == MTX_JOB_QUEUED) {
    J->state = MTX_JOB_CANCELLED;
    J->completeMs = now_ms();
    S->queued.erase(std::remove(S->queued.begin(), S->queued.end(), job), S->queued.end());
    S->running.erase(std::remove(S->running.begin(), S->running.end(), job), S->running.end());
  }
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_server_job_release(mtx_handle server, mtx_handle job) {
  API_TRY
  auto S = get_server(server);
  std::lock_guard<std::mutex> g(S->m);
  auto J = get_job(*S, job);
  if (J->state == MTX_JOB_QUEUED || J->state == MTX_JOB_RUNNING) return MTX_INVALID_STATE;
  if (J->admitted) S->used -= J->bytes();
  S->jobs.erase(job);
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_server_stats(mtx_handle server, uint64_t stats[5]) {
  API_TRY
  MTX_REQUIRE(stats, MTX_INVALID_ARGUMENT, "stats is NULL");
  auto S = get_server(server);
  std::lock_guard<std::mutex> g(S->m);
  stats[0] = S->rounds;
  S->rays.download(&stats[1], 1, 0);
  MTX_HIP(hipStreamSynchronize(0));
  stats[2] = S->running.size();
  stats[3] = S->queued.size();
  stats[4] = S->used;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_server_graph_stats(mtx_handle server, uint64_t stats[2]) {
  API_TRY
  MTX_REQUIRE(stats, MTX_INVALID_ARGUMENT, "stats is NULL");
  auto S = get_server(server);
  std::lock_guard<std::mutex> g(S->m);
  stats[0] = S->graph.replays;
  stats[1] = S->graph.captures;
  return MTX_OK;
  API_CATCH(nullptr, 0)
}

int32_t mtx_server_destroy(mtx_handle server) {
  std::lock_guard<std::mutex> g(gRenderMutex);
  return gServers.erase(server) ? MTX_OK : MTX_INVALID_HANDLE;
}

}  // extern "C"
