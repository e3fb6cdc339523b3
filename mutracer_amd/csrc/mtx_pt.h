// Path tracer: resident render scene, render targets and kernel entry points.
#pragma once
#include "mtx_internal.h"
#include "mtx_scene.h"

namespace mtx {

// Shading data per instance slot (leaf order, parallel to InstSlot).
struct ShadeSlot {
  float4 l2w0, l2w1, l2w2, l2w3;  // local-to-world columns (xyz)
  const float *normals;           // source mesh vertex normals (xyz)
  const uint32_t *indices;        // source mesh triangle indices
  uint32_t material;
  float color[3];
  const float *uvs, *uvs1;        // source mesh UV0 / UV1 (xy), or null
};
static_assert(sizeof(ShadeSlot) == 112, "ShadeSlot");

// A material texture: the sample transform, size, flags (bit 0 RGBA32F,
// bits 1-2 wrap s, bits 3-4 wrap t, bit 5 UV1) and its first texel in the
// pool of its storage (rows already in sampling order).
struct GpuTexture {
  float sx, sy, ox, oy;
  uint32_t width, height, flags, base;
};
static_assert(sizeof(GpuTexture) == 32, "GpuTexture");

// Alias table entry (lightSampling.ts:172-212): probability, alias, type, index.
struct LightChoice {
  float prob;
  int32_t alias, type, index;
};

// Packed OpenPBR material (32 floats) and rect light (16 floats); the field
// offsets are documented with mtx_material_desc / mtx_light_desc in mtx.h.
struct GpuScene {
  const AccelView *accel;
  const ShadeSlot *shade;
  const float *materials;
  const float *lights;
  const float *lightPmf;       // per light selection pmf
  const LightChoice *choices;
  int32_t choiceCount;
  uint32_t lightCount;
  float envPmf;
  float lightWeightTotal;
  const float4 *env;           // RGBA f32, envH rows of envW
  const float2 *envCdf;        // (envW + 2) x (envH + 1)
  uint32_t envW, envH;
  const float *lut;            // 32 x 2177 energy tables
  const float4 *emissive;      // 3 per emissive triangle: (p0, pmf), (p1, area), (p2, material), world space
  uint32_t emissiveCount;
  uint32_t features;           // kSceneFull when a material has fuzz, subsurface, a thin wall or a texture
  const uint16_t *envGuide;    // search brackets of envCdf (kEnvGuideRows + 1, then envH x (kEnvGuideCols + 1)), or null
  const GpuTexture *textures;  // material textures, or null (no material has a map)
  const uint32_t *texels8;     // RGBA8 texel pool
  const float4 *texels32;      // RGBA32F texel pool
  const float4 *emissiveUv;    // 3 per emissive triangle: (uv0, uv1), (uv2, uv1_0), (uv1_1, uv1_2), with textures
};
// Bins of the environment CDF guide tables: bin k of the marginal (of row
// y's conditional) CDF holds the index its search returns for the sample
// k / bins, so a sample in bin k searches only between bins k and k + 1.
constexpr uint32_t kSceneFull = 1u;
constexpr uint32_t kSceneTextured = 2u;  // set with kSceneFull: the full kernels sample the maps
constexpr uint32_t kSceneMasked = 4u;    // a material is alpha-masked: shadow rays run the coverage test
constexpr uint32_t kEnvGuideRows = 1024;
constexpr uint32_t kEnvGuideCols = 128;

struct RenderParams {
  float projInv[16], viewInv[16];
  float aperture, focus;
  uint32_t width, height;
  uint32_t bounces, candidates, pattern, seed, logSPP, sampleOffset;
  float clamping;
  uint32_t samples;            // samples per pixel this launch
  uint32_t pixelBase;          // first pixel of this launch (row-major)
  uint32_t pixelCount;
};

// FP64 accumulators, one plane per channel (13 planes) + u32 sample counts.
struct Accumulators {
  double *planes;  // 13 * pixels
  uint32_t *count;
  uint64_t *rays;  // traced rays (camera + continuation + shadow)
  uint32_t pixels;
  uint64_t *work;  // optional: box tests, triangle tests, instance visits, hits, shadow rays
  const float *targets = nullptr;  // optional per-pixel sample targets (adaptive sampling)
};

struct RenderSceneObj;
struct TargetObj;

// Wavefront path state (structure of arrays, one entry per path in flight):
//   o    ray origin, w = scattering pdf        d    ray direction, w = alpha
//   li   radiance, w = path flags (bits)       beta throughput, w = last hit triangle
//   alb  albedo AOV, w = last hit slot         nrm  normal AOV
//   hitA closest hit t, barycentrics           hitB (instance slot or -1, triangle)
//   qA/qB path-id queues of even/odd events
//   s0..s3 shadow queue: (origin, tMax), (direction, path id), (payload, tri), slot
//   counts[e] = length of event e's queue, counts[kWfShadowBase + e] = its shadow queue,
//   counts[kWfFetchBase + 2e (+1)] = next entry the persistent trace (shadow) kernel takes
//   counts[kWfXcdBase + (2e (+1)) * 8 + x] = next 64-entry chunk of XCD x's segment of
//                                           event e's (shadow) queue
//   counts[kWfTailFlag] = 1 + the event at which the tail kernel took over the
//                         generation's remaining paths (0 while it has not)
constexpr int kWfMaxEvents = 51;
constexpr int kWfShadowBase = kWfMaxEvents + 1;
constexpr int kWfFetchBase = kWfShadowBase + kWfMaxEvents + 1;
constexpr int kWfXcdBase = kWfFetchBase + 2 * (kWfMaxEvents + 1);
constexpr int kWfXcds = 8;  // XCDs of an MI355X (counter segments per queue)
constexpr int kWfTailFlag = kWfXcdBase + 2 * (kWfMaxEvents + 1) * kWfXcds;
// Queue sorting: the shade kernel appends each continuing path to the bin of
// its ray direction's octant, counts[kWfSortBase + kWfSortBins * e + o]
// entries for queue e, and a gather kernel concatenates the bins into the
// queue, so a wave's rays share a direction octant (near/far child order and
// nodes in common) while keeping their pixel order within the octant.
constexpr int kWfSortBins = 8;
constexpr int kWfSortBase = kWfTailFlag + 1;
constexpr int kWfCounters = kWfSortBase + kWfSortBins * (kWfMaxEvents + 2);
struct WfBuffers {
  float4 *o, *d, *li, *beta, *alb, *nrm, *hitA;
  int2 *hitB;
  uint32_t *qA, *qB;
  float4 *s0, *s1, *s2;
  int32_t *s3;
  uint16_t *item;   // job slot of each path (rounds with several jobs)
  uint4 *rng;       // pcg4d state of each path (random sampling pattern)
  uint32_t *qSort;  // kWfSortBins bins of nPaths entries (queue sorting on), else null
  uint32_t *counts;
  uint32_t nPaths;  // paths in flight this generation
  struct Owner {
    DevBuf<float4> o, d, li, beta, alb, nrm, hitA, s0, s1, s2;
    DevBuf<int2> hitB;
    DevBuf<uint32_t> qA, qB, counts;
    DevBuf<int32_t> s3;
    DevBuf<uint16_t> item;
    DevBuf<uint4> rng;
    DevBuf<uint32_t> qSort;
    uint32_t capacity = 0, sortCapacity = 0;
    void ensure(uint32_t paths, bool sort);
    size_t bytes() const;
    WfBuffers view(uint32_t nPaths, bool sort) const;
  };
};

// One job's share of a wavefront generation: its scene, parameters (samples =
// samples per pixel in this generation), accumulators, and where its paths
// and pixels sit in the generation's concatenated path / pixel ranges. Paths
// of a job are ordered pixel-major: path = pathBase + local * samples + s.
struct JobSlot {
  GpuScene scene;
  RenderParams params;
  Accumulators acc;
  uint32_t pathBase, pixelPrefix;
};
constexpr uint32_t kMaxJobsPerRound = 1024;

// What every wavefront kernel of one generation receives: a single job by
// value, or a device table of jobs (item[] maps paths to it).
struct WfLaunch {
  JobSlot one;
  const JobSlot *jobs;
  uint32_t nJobs;
  uint32_t totalPixels;
  uint64_t *rays;  // traced-ray counter
  uint64_t *work;  // optional traversal work totals (8 counters)
  uint32_t masked; // a job's scene has masked materials: the host launches the coverage-testing
                   // any-hit kernels (the kernels do not read this field)
};

// Launch shape of the wavefront kernels (the autotuner's search space).
struct WfConfig {
  uint32_t maxPaths = 1u << 25;  // paths in flight per generation
  int traceOversub = 1;          // trace/shadow grid = CUs * resident blocks * this
  int shadeOversub = 1;
  bool walker = false;           // persistent lane-refill traversal (else one ray per thread)
  int ldsStack = 16;             // LDS stack entries per lane of the any-hit (and walker) kernels (16, 24 or 32)
  int ldsStackClosest = 0;       // the same for the closest-hit trace kernel (0 = ldsStack)
  int shadeWaves = 2;            // occupancy target of the shade kernel (1..4 waves per SIMD)
  bool graphs = true;            // capture each generation's launch sequence into a hipGraph and replay it
  uint32_t tailPaths = 65536;    // the tail kernel finishes the paths once a queue is this short (0 = never)
  bool queueSort = false;        // continuation queues grouped by ray direction octant (measured slower at 1080p)
};

// Traversal work totals of an instrumented launch (acc.work): one group of
// kWorkGroup counters for the wavefront (or megakernel) kernels, one for the
// wavefront tail kernel at kWorkTailBase.
enum : int {
  kWorkClosestRays = 0, kWorkBoxes, kWorkTris, kWorkInstances, kWorkHits, kWorkShadowRays, kWorkShadowBoxes,
  kWorkShadowTris, kWorkShadowInstances, kWorkGroup, kWorkTailBase = kWorkGroup, kWorkSlots = 2 * kWorkGroup
};

// Optional per-kernel-class timing with HIP events on the launch stream:
// class 0 generate + accumulate, 1 closest-hit trace, 2 shade, 3 shadow
// trace, 4 megakernel, 5 wavefront tail kernel.
constexpr int kTimerClasses = 6;
struct LaunchTimer {
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> events;  // begin/end pairs
  std::vector<int> classes;
  void begin(int cls);
  void end();
  void collect(double ms[kTimerClasses]);  // synchronizes, sums, releases the events
  ~LaunchTimer();
};

// A captured generation (hipGraph) and the launch arguments it was captured
// with; replayed while they are unchanged.
struct WfGraph {
  hipGraphExec_t exec = nullptr;
  uint64_t key = 0;
  hipStream_t own = nullptr;  // capture/replay stream when the caller passes the null stream
  uint64_t replays = 0, captures = 0;
  void reset();
  ~WfGraph();
};

void launch_render(const GpuScene &scene, const RenderParams &p, const Accumulators &acc, hipStream_t s,
                   LaunchTimer *timer = nullptr);
void launch_render_wavefront(const GpuScene &scene, const RenderParams &p, const Accumulators &acc,
                             WfBuffers::Owner &buf, const WfConfig &cfg, hipStream_t s, LaunchTimer *timer = nullptr,
                             WfGraph *graph = nullptr);
// One generation of several jobs (slots on the device, totals given): paths
// = sum of pathCount, pixels = sum of pixelCount, events = the largest job's
// event limit; `full`: some job's scene has kSceneFull or its sampling
// pattern is the random one.
// With `graph` (and cfg.graphs), the launch sequence is captured once and
// replayed while its shape (job count, totals, buffers, stream) is unchanged:
// the kernels read the slots from device memory.
void launch_generation_multi(const JobSlot *devSlots, uint32_t nJobs, uint32_t nPaths, uint32_t totalPixels,
                             int eventLimit, uint64_t *rays, WfBuffers::Owner &buf, const WfConfig &cfg,
                             hipStream_t s, bool full, bool masked, LaunchTimer *timer = nullptr,
                             WfGraph *graph = nullptr);
// Whether one job needs the full shading kernels.
inline bool needs_full_shading(const GpuScene &scene, const RenderParams &p) {
  return (scene.features & kSceneFull) != 0u || p.pattern == 2u;
}

}  // namespace mtx
