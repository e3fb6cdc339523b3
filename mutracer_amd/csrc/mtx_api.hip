// C ABI of libmtx: handle registry, BLAS / scene acceleration objects and the
// ray-query kernels.
#include <atomic>
#include <map>
#include <mutex>

#include "mtx_fetch.h"
#include "mtx_registry.h"

namespace mtx {

// ------------------------------------------------------------------ registry
Registry &registry() {
  static Registry r;
  return r;
}

uint64_t Registry::next() { return nextHandle.fetch_add(1); }

// ------------------------------------------------------------- trace kernels
constexpr int kTraceBlock = 128;

__global__ void k_zero_counters(uint32_t *__restrict__ c) {
  if (threadIdx.x < (uint32_t)kXcds) c[threadIdx.x] = 0u;
}

// Ray queries: the same traversal, occupancy targets and XCD-segmented work
// distribution as the wavefront trace (closest hit) and shadow (any hit)
// kernels; `ctr` holds the kXcds chunk counters, zeroed before the launch.
template <int kMode>  // 0 closest, 1 any, 2 counters
__global__ void __launch_bounds__(kTraceBlock, kMode == 1 ? kShadowWaves : kTraceWaves)
k_trace(const AccelView *__restrict__ viewp, const mtx_ray *__restrict__ rays, uint32_t n,
        mtx_hit *__restrict__ hits, uint32_t *__restrict__ out, uint32_t flags, int32_t *__restrict__ spill,
        uint32_t *__restrict__ ctr) {
  __shared__ int32_t lds[kMinLdsStack * kTraceBlock];
  const AccelView acc = *viewp;
  LaneStackT<kMinLdsStack> st{lds + threadIdx.x, kTraceBlock,
                              spill + ((size_t)blockIdx.x * kTraceBlock + threadIdx.x) * (kMaxStack - kMinLdsStack)};
  const bool doubleSided = (flags & MTX_TRACE_DOUBLE_SIDED) != 0;
  auto ray = [&](uint32_t i) {
    const float4 *r4 = reinterpret_cast<const float4 *>(rays + i);
    float4 a = r4[0], b = r4[1];
    f3 o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
    TraceHit h;
    TraceCounters c{0, 0, 0};
    if (kMode == 1) {
      float tMax = fminf(fmaxf(a.w, 0.0f), kRayMaxDistance);
      out[i] = traverse<true, false>(acc, o, d, tMax, true, 1, -1, -1, st, h, nullptr) ? 1u : 0u;
      return;
    }
    const bool counting = kMode == 2 || (flags & MTX_TRACE_COUNT);
    if (counting) traverse<false, true>(acc, o, d, kRayMaxDistance, doubleSided, 0, -1, -1, st, h, &c);
    else traverse<false, false>(acc, o, d, kRayMaxDistance, doubleSided, 0, -1, -1, st, h, nullptr);
    if (kMode == 2) {
      out[(size_t)i * 3] = c.boxes;
      out[(size_t)i * 3 + 1] = c.tris;
      out[(size_t)i * 3 + 2] = c.instances;
      return;
    }
    mtx_hit o4;
    o4.hit = h.hit ? 1u : 0u;
    if (h.hit) {
      const InstSlot &s = acc.slots[h.slot];
      o4.instance = fint(s.c2.w);
      o4.primitive = fint(s.tris[h.tri].v0.w);
      o4.t = h.t;
      o4.barycentric[0] = h.bary.x;
      o4.barycentric[1] = h.bary.y;
      o4.barycentric[2] = h.bary.z;
    } else {
      o4.instance = -1;
      o4.primitive = -1;
      o4.t = kRayMaxDistance;
      o4.barycentric[0] = o4.barycentric[1] = o4.barycentric[2] = 0.0f;
    }
    o4.visits = counting ? c.boxes + c.tris : 0u;
    hits[i] = o4;
  };
  for_each_ray<kMode == 1 ? kShadowFetch : kTraceFetch>(n, ctr, ray);
}

// Spill slices for the traversal stacks and chunk counters of the ray
// queries, one set per stream.
struct SpillPool {
  std::mutex m;
  std::map<std::pair<int, hipStream_t>, std::shared_ptr<DevBuf<int32_t>>> bufs;
  std::map<std::pair<int, hipStream_t>, std::shared_ptr<DevBuf<uint32_t>>> counters;
};
static SpillPool gSpill;

static uint32_t *counters_for(hipStream_t s) {
  int dev = 0;
  MTX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(gSpill.m);
  auto &b = gSpill.counters[{dev, s}];
  if (!b) b = std::make_shared<DevBuf<uint32_t>>(kXcds);
  return b->ptr;
}

uint32_t trace_grid() {
  static uint32_t grid = 0;
  if (!grid) {
    int dev = 0, cus = 0;
    MTX_HIP(hipGetDevice(&dev));
    MTX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid = (uint32_t)std::max(8, cus) * 8u;
    grid = (grid + 7u) & ~7u;
  }
  return grid;
}

int32_t *spill_for(hipStream_t s, uint32_t blocksNeeded, uint32_t blockThreads) {
  int dev = 0;
  MTX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(gSpill.m);
  auto &b = gSpill.bufs[{dev, s}];
  // Each lane owns kMaxStack - kLds entries at lane * (kMaxStack - kLds):
  // sized for the smallest LDS depth any kernel uses.
  size_t need = (size_t)blocksNeeded * blockThreads * (kMaxStack - kMinLdsStack);
  if (!b || b->count < need) b = std::make_shared<DevBuf<int32_t>>(need);
  return b->ptr;
}

static void launch_trace(int mode, const Accel &A, const mtx_ray *rays, uint32_t n, mtx_hit *hits, uint32_t *out,
                         uint32_t flags, hipStream_t s) {
  MTX_REQUIRE(A.max_stack() <= (uint32_t)kMaxStack, MTX_INVALID_STATE, "scene deeper than the traversal stack");
  if (!n) return;
  // Persistent grid: the resident blocks of the kernel on every CU.
  static uint32_t bpc[3] = {0, 0, 0};
  const void *k = mode == 0 ? (const void *)k_trace<0> : mode == 1 ? (const void *)k_trace<1> : (const void *)k_trace<2>;
  if (!bpc[mode]) {
    int b = 0;
    MTX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, kTraceBlock, 0));
    bpc[mode] = (uint32_t)std::max(1, b);
  }
  uint32_t grid = trace_grid() / 8u * bpc[mode];
  uint32_t rayBlocks = (n + kTraceBlock - 1) / kTraceBlock;
  grid = std::min<uint32_t>(grid, ((rayBlocks + 7u) / 8u) * 8u);
  int32_t *spill = spill_for(s, grid, kTraceBlock);
  uint32_t *ctr = counters_for(s);
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, ctr);
  if (mode == 0)
    hipLaunchKernelGGL(k_trace<0>, dim3(grid), dim3(kTraceBlock), 0, s, A.view.ptr, rays, n, hits, out, flags, spill,
                       ctr);
  else if (mode == 1)
    hipLaunchKernelGGL(k_trace<1>, dim3(grid), dim3(kTraceBlock), 0, s, A.view.ptr, rays, n, hits, out, flags, spill,
                       ctr);
  else
    hipLaunchKernelGGL(k_trace<2>, dim3(grid), dim3(kTraceBlock), 0, s, A.view.ptr, rays, n, hits, out, flags, spill,
                       ctr);
  MTX_HIP(hipGetLastError());
}

// ----------------------------------------------- allocation-free ray query
// mtx_trace_closest_noalloc runs k_trace<0> on a fixed persistent grid of
// kNoAllocBlocks blocks whose traversal-stack spill slices and chunk counters
// live in `scratch`, a file-scope device array: each device holds its own copy
// from module load until the process ends, and nothing is allocated per call.
constexpr uint32_t kNoAllocBlocks = 1024;
constexpr size_t kNoAllocSpillInts = (size_t)kNoAllocBlocks * kTraceBlock * (kMaxStack - kMinLdsStack);
constexpr size_t kNoAllocScratchBytes = kNoAllocSpillInts * sizeof(int32_t) + 256;  // + kXcds chunk counters
constexpr uint32_t kNoAllocMaxRays = 1u << 30;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Device address of `scratch` per device, looked up once with that device
// current (the host-side symbol is only a stand-in, not a device address).
static std::mutex gScratchMutex;
static std::map<int, uint8_t *> gScratchPtr;

static uint8_t *scratch_for_current_device() {
  int dev = 0;
  MTX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(gScratchMutex);
  auto it = gScratchPtr.find(dev);
  if (it != gScratchPtr.end()) return it->second;
  void *p = nullptr;
  MTX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  MTX_REQUIRE(p != nullptr, MTX_CONTEXT_FAILED, "scratch symbol has no device address");
  gScratchPtr[dev] = (uint8_t *)p;
  return (uint8_t *)p;
}

static void launch_trace_noalloc(const Accel &A, const mtx_ray *rays, uint32_t n, mtx_hit *hits, uint32_t flags,
                                 hipStream_t s) {
  MTX_REQUIRE(A.max_stack() <= (uint32_t)kMaxStack, MTX_INVALID_ARGUMENT, "scene deeper than the traversal stack");
  MTX_REQUIRE(n <= kNoAllocMaxRays, MTX_INVALID_ARGUMENT, "more than 2^30 rays in one allocation-free query");
  uint8_t *base = scratch_for_current_device();
  if (!n) return;
  int32_t *spill = reinterpret_cast<int32_t *>(base);
  uint32_t *ctr = reinterpret_cast<uint32_t *>(base + kNoAllocSpillInts * sizeof(int32_t));
  static uint32_t bpc = 0;
  if (!bpc) {
    int b = 0;
    MTX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void *)k_trace<0>, kTraceBlock, 0));
    bpc = (uint32_t)std::max(1, b);
  }
  uint32_t grid = std::min<uint32_t>(trace_grid() / 8u * bpc, kNoAllocBlocks);
  const uint32_t rayBlocks = (n + kTraceBlock - 1) / kTraceBlock;
  grid = std::min<uint32_t>(grid, ((rayBlocks + 7u) / 8u) * 8u);
  grid = std::max<uint32_t>(8u, grid & ~7u);
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, ctr);
  hipLaunchKernelGGL(k_trace<0>, dim3(grid), dim3(kTraceBlock), 0, s, A.view.ptr, rays, n, hits, nullptr, flags, spill,
                     ctr);
  MTX_HIP(hipGetLastError());
}

}  // namespace mtx

using namespace mtx;

#define MTX_API_BEGIN try {
#define MTX_API_END(err, cap)                                     \
  }                                                               \
  catch (const mtx::Error &e) {                                   \
    set_error(err, cap, e.what());                                \
    return e.code;                                                \
  }                                                               \
  catch (const std::bad_alloc &) {                                \
    set_error(err, cap, "out of host memory");                    \
    return MTX_ERROR;                                             \
  }                                                               \
  catch (const std::exception &e) {                               \
    set_error(err, cap, e.what());                                \
    return MTX_ERROR;                                             \
  }

extern "C" {

uint32_t mtx_abi_version(void) { return MTX_ABI_VERSION; }

int32_t mtx_device_info(char *name, size_t cap, int32_t *cus) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MTX_UNAVAILABLE;
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return MTX_UNAVAILABLE;
  set_error(name, cap, std::string(p.name) + " " + p.gcnArchName);
  if (cus) *cus = p.multiProcessorCount;
  return MTX_OK;
}

int32_t mtx_blas_build(const float *positions, uint32_t vertex_count, const uint32_t *indices,
                       uint32_t triangle_count, uint32_t bins, uint32_t flags, mtx_handle *out, char *err,
                       size_t cap) {
  MTX_API_BEGIN
  MTX_REQUIRE(out, MTX_INVALID_ARGUMENT, "out_blas is NULL");
  MTX_REQUIRE(triangle_count == 0 || (positions && indices), MTX_INVALID_ARGUMENT, "positions/indices are NULL");
  MTX_REQUIRE(bins >= 2 && bins <= 64, MTX_INVALID_ARGUMENT, "bins must be in [2, 64]");
  auto B = build_blas(positions, vertex_count, indices, triangle_count, bins, (flags & MTX_INPUT_DEVICE) != 0, 0);
  auto &R = registry();
  std::lock_guard<std::mutex> g(R.m);
  uint64_t h = R.next();
  R.blas[h] = B;
  *out = h;
  return MTX_OK;
  MTX_API_END(err, cap)
}

int32_t mtx_blas_info(mtx_handle blas, uint32_t info[8]) {
  MTX_API_BEGIN
  auto B = registry().get_blas(blas);
  MTX_REQUIRE(info, MTX_INVALID_ARGUMENT, "info is NULL");
  info[0] = B->topo.nodeCount;
  info[1] = B->topo.nodeCount * 2 + B->topo.primCount;
  info[2] = B->topo.maxDepth;
  info[3] = B->maxLeaf;
  info[4] = B->topo.primCount;
  uint64_t bytes = B->resident_bytes();
  info[5] = (uint32_t)bytes;
  info[6] = (uint32_t)(bytes >> 32);
  info[7] = B->layout.interiorCount;
  return MTX_OK;
  MTX_API_END(nullptr, 0)
}

int32_t mtx_blas_refit(mtx_handle blas, const float *positions, uint32_t vertex_count, const uint32_t *indices,
                       uint32_t triangle_count, uint32_t flags, uint32_t *rebuild_advised, char *err, size_t cap) {
  MTX_API_BEGIN
  auto B = registry().get_blas(blas);
  MTX_REQUIRE(positions && indices, MTX_INVALID_ARGUMENT, "positions/indices are NULL");
  MTX_REQUIRE(vertex_count == B->vertexInput && triangle_count == B->triangleInput, MTX_INVALID_ARGUMENT,
              "refit needs the build's vertex and triangle counts");
  const bool dev = (flags & MTX_INPUT_DEVICE) != 0;
  DevBuf<float> hpos;
  DevBuf<uint32_t> hidx;
  const float *dpos = positions;
  const uint32_t *didx = indices;
  if (!dev) {
    hpos.upload(positions, (size_t)vertex_count * 3, 0);
    hidx.upload(indices, (size_t)triangle_count * 3, 0);
    dpos = hpos.ptr;
    didx = hidx.ptr;
  }
  const bool advise = refit_blas(*B, dpos, didx, 0);
  if (rebuild_advised) *rebuild_advised = advise ? 1u : 0u;
  return MTX_OK;
  MTX_API_END(err, cap)
}

int32_t mtx_blas_export(mtx_handle blas, const uint32_t *indices, uint32_t *words, size_t capacity,
                        size_t *written, uint32_t *source_of_record) {
  MTX_API_BEGIN
  auto B = registry().get_blas(blas);
  size_t need = export_blas(*B, indices, nullptr, 0, nullptr, 0);
  if (written) *written = need;
  if (!words) return MTX_OK;
  if (capacity < need) return MTX_BUFFER_TOO_SMALL;
  export_blas(*B, indices, words, capacity, source_of_record, 0);
  return MTX_OK;
  MTX_API_END(nullptr, 0)
}

int32_t mtx_blas_destroy(mtx_handle blas) {
  auto &R = registry();
  std::lock_guard<std::mutex> g(R.m);
  return R.blas.erase(blas) ? MTX_OK : MTX_INVALID_HANDLE;
}

int32_t mtx_accel_create(const mtx_handle *blases, uint32_t blas_count, const mtx_instance_desc *instances,
                         uint32_t instance_count, uint32_t tlas_bins, uint32_t flags, mtx_handle *out, char *err,
                         size_t cap) {
  MTX_API_BEGIN
  MTX_REQUIRE(out, MTX_INVALID_ARGUMENT, "out_accel is NULL");
  MTX_REQUIRE(instance_count == 0 || instances, MTX_INVALID_ARGUMENT, "instances is NULL");
  MTX_REQUIRE(tlas_bins >= 2 && tlas_bins <= 64, MTX_INVALID_ARGUMENT, "tlas_bins must be in [2, 64]");
  for (uint32_t i = 0; i < instance_count; i++)
    MTX_REQUIRE(instances[i].struct_size >= sizeof(mtx_instance_desc), MTX_INVALID_ARGUMENT,
                "mtx_instance_desc.struct_size too small");
  std::vector<std::shared_ptr<Blas>> bl;
  for (uint32_t i = 0; i < blas_count; i++) bl.push_back(registry().get_blas(blases[i]));
  auto A = build_accel(bl, instances, instance_count, tlas_bins, flags, 0);
  auto &R = registry();
  std::lock_guard<std::mutex> g(R.m);
  uint64_t h = R.next();
  R.accel[h] = A;
  *out = h;
  return MTX_OK;
  MTX_API_END(err, cap)
}

int32_t mtx_accel_refit(mtx_handle accel, const mtx_instance_desc *instances, uint32_t count, const uint32_t *dirty,
                        uint32_t dirty_count, char *err, size_t cap) {
  MTX_API_BEGIN
  auto A = registry().get_accel(accel);
  MTX_REQUIRE(count == 0 || instances, MTX_INVALID_ARGUMENT, "instances is NULL");
  refit_accel(*A, instances, count, dirty, dirty_count, 0);
  return MTX_OK;
  MTX_API_END(err, cap)
}

int32_t mtx_accel_export_tlas(mtx_handle accel, uint32_t *tlas_words, size_t tlas_capacity, size_t *tlas_written,
                              uint32_t *instance_words, size_t instance_capacity, size_t *instance_written) {
  MTX_API_BEGIN
  auto A = registry().get_accel(accel);
  std::vector<uint32_t> tw, iw;
  export_tlas(*A, tw, iw, 0);
  if (tlas_written) *tlas_written = tw.size();
  if (instance_written) *instance_written = iw.size();
  if (tlas_words) {
    if (tlas_capacity < tw.size()) return MTX_BUFFER_TOO_SMALL;
    std::memcpy(tlas_words, tw.data(), tw.size() * 4);
  }
  if (instance_words) {
    if (instance_capacity < iw.size()) return MTX_BUFFER_TOO_SMALL;
    std::memcpy(instance_words, iw.data(), iw.size() * 4);
  }
  return MTX_OK;
  MTX_API_END(nullptr, 0)
}

int32_t mtx_accel_destroy(mtx_handle accel) {
  auto &R = registry();
  std::lock_guard<std::mutex> g(R.m);
  return R.accel.erase(accel) ? MTX_OK : MTX_INVALID_HANDLE;
}

int32_t mtx_trace_closest(mtx_handle accel, const mtx_ray *rays, uint32_t count, mtx_hit *hits, uint32_t flags,
                          void *stream) {
  MTX_API_BEGIN
  auto A = registry().get_accel(accel);
  MTX_REQUIRE(count == 0 || (rays && hits), MTX_INVALID_ARGUMENT, "rays/hits are NULL");
  launch_trace(0, *A, rays, count, hits, nullptr, flags, (hipStream_t)stream);
  return MTX_OK;
  MTX_API_END(nullptr, 0)
}

int32_t mtx_trace_closest_noalloc(mtx_handle accel, const mtx_ray *rays, uint32_t count, mtx_hit *hits,
                                  uint32_t flags, void *stream) {
  MTX_API_BEGIN
  auto A = registry().get_accel(accel);
  MTX_REQUIRE(count == 0 || (rays && hits), MTX_INVALID_ARGUMENT, "rays/hits are NULL");
  launch_trace_noalloc(*A, rays, count, hits, flags, (hipStream_t)stream);
  return MTX_OK;
  MTX_API_END(nullptr, 0)
}

int32_t mtx_trace_any(mtx_handle accel, const mtx_ray *rays, uint32_t count, uint32_t *occluded, void *stream) {
  MTX_API_BEGIN
  auto A = registry().get_accel(accel);
  MTX_REQUIRE(count == 0 || (rays && occluded), MTX_INVALID_ARGUMENT, "rays/occluded are NULL");
  launch_trace(1, *A, rays, count, nullptr, occluded, 0, (hipStream_t)stream);
  return MTX_OK;
  MTX_API_END(nullptr, 0)
}

int32_t mtx_trace_count(mtx_handle accel, const mtx_ray *rays, uint32_t count, uint32_t *counters, uint32_t flags,
                        void *stream) {
  MTX_API_BEGIN
  auto A = registry().get_accel(accel);
  MTX_REQUIRE(count == 0 || (rays && counters), MTX_INVALID_ARGUMENT, "rays/counters are NULL");
  launch_trace(2, *A, rays, count, nullptr, counters, flags, (hipStream_t)stream);
  return MTX_OK;
  MTX_API_END(nullptr, 0)
}

}  // extern "C"
