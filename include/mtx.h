/*
 * mtx: MI355X-native render hot path for MuTracer -- C ABI.
 *
 * One shared library (libmtx.so, HIP for gfx950) serves the TypeScript host
 * (through a thin N-API addon) and native hosts. Conventions follow the
 * reference's native C ABI:
 *   - every call returns an int32_t status (mutracer-status.h:14-45);
 *   - optional caller-owned `char *error, size_t error_capacity` receive a
 *     NUL-terminated, UTF-8-truncated message;
 *   - every struct starts with `uint32_t struct_size`
 *     (mutracer-status.h:60-65);
 *   - handles are uint64_t, monotonic and never reused;
 *   - inputs are copied, no caller pointer is retained after a call returns;
 *   - no exception crosses the ABI; a GPU error marks the object failed.
 * Pointers documented as "device" are HIP device addresses (HBM); all others
 * are host addresses.
 */
#ifndef MTX_H
#define MTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTX_ABI_VERSION 1u

/* Status codes: values of mutracer-status.h:14-45. */
enum {
  MTX_OK = 0,
  MTX_ERROR = 1,
  MTX_INVALID_ARGUMENT = 2,
  MTX_INVALID_HANDLE = 3,
  MTX_BUSY = 4,
  MTX_INVALID_STATE = 5,
  MTX_CANCELLED = 7,
  MTX_BUFFER_TOO_SMALL = 8,
  MTX_CONTEXT_FAILED = 9,
  MTX_UNAVAILABLE = 10,
  MTX_TIMEOUT = 11
};

typedef uint64_t mtx_handle;

/* Library ABI version (MTX_ABI_VERSION). No reference counterpart. */
uint32_t mtx_abi_version(void);

/* Device name and compute-unit count of the current HIP device.
 * Returns MTX_UNAVAILABLE when no HIP device is present. */
int32_t mtx_device_info(char *name, size_t name_capacity, int32_t *compute_units);

/* ------------------------------------------------------------------ BLAS */
/* Input pointers are device addresses when flags has MTX_INPUT_DEVICE. */
#define MTX_INPUT_DEVICE 1u

/* Binned-SAH BVH2 build of one triangle mesh on the GPU. Replaces
 * buildBLAS (src/renderer/bvhAccel.ts:200-286) with buildBinnedSahBvh
 * (src/renderer/binnedSahBvh.ts:70-236); the tree (topology, bounds, leaf
 * order) is identical to the reference builder's for the same input and
 * `bins` (8 is the reference default). Triangles with non-finite bounds are
 * skipped as in bvhAccel.ts:162-168. The BVH stays resident on the device. */
int32_t mtx_blas_build(const float *positions, uint32_t vertex_count, const uint32_t *indices,
                       uint32_t triangle_count, uint32_t bins, uint32_t flags, mtx_handle *out_blas,
                       char *error, size_t error_capacity);

/* info[0] node count, [1] reference BVH2 vec4 count (2*nodes + prims),
 * [2] max depth (root = 1), [3] max leaf primitives, [4] primitive count,
 * [5] resident bytes (low 32 bits), [6] resident bytes (high 32 bits). */
int32_t mtx_blas_info(mtx_handle blas, uint32_t info[8]);

/* Writes the BVH in the reference's BVH2 vec4 layout (bvhAccel.ts:248-275:
 * node = [min.xyz, right or -count][max.xyz, 0], leaf entries [i0,i1,i2,0]
 * after their leaf, depth-first with the left child adjacent). words == NULL
 * queries the size through *written_words. source_of_record (optional,
 * vec4_count entries) receives the source triangle of each leaf entry, or
 * 0xffffffff for node rows. Parity/interop only; not used on the hot path. */
int32_t mtx_blas_export(mtx_handle blas, const uint32_t *indices, uint32_t *words,
                        size_t capacity_words, size_t *written_words, uint32_t *source_of_record);

int32_t mtx_blas_destroy(mtx_handle blas);

/* ----------------------------------------------------- scene acceleration */
typedef struct mtx_instance_desc {
  uint32_t struct_size;
  uint32_t blas_index;      /* index into the blas array of mtx_accel_create */
  uint32_t material_word;   /* packed material word (material << 8) */
  uint32_t double_sided;    /* material's doubleSided flag (no back-face culling) */
  double transform[16];     /* local-to-world, column-major (f64, as the reference) */
  float color[3];           /* instance color (instancesBuffer row 9) */
  float reserved;
} mtx_instance_desc;

#define MTX_ACCEL_PER_INSTANCE_BOUNDS 1u /* buildWebTLAS leaf boxes (bvhAccel.ts:452) */

/* TLAS over instances (buildTLAS / buildWebTLAS, bvhAccel.ts:385-454), built
 * with the same GPU binned-SAH builder (tlas_bins: 64 web, 8 native). */
int32_t mtx_accel_create(const mtx_handle *blases, uint32_t blas_count,
                         const mtx_instance_desc *instances, uint32_t instance_count,
                         uint32_t tlas_bins, uint32_t flags, mtx_handle *out_accel,
                         char *error, size_t error_capacity);

/* Transform-only update (refitTLAS, bvhAccel.ts:579-740): new instance
 * transforms for the listed source instances (dirty == NULL: all), world boxes
 * recomputed on the GPU and TLAS node boxes refitted bottom-up; topology kept. */
int32_t mtx_accel_refit(mtx_handle accel, const mtx_instance_desc *instances, uint32_t instance_count,
                        const uint32_t *dirty, uint32_t dirty_count, char *error, size_t error_capacity);

/* Reference-layout TLAS words (incl. web leaf boxes) and instance records
 * (stride 40 words, bvhAccel.ts:323-353). NULL outputs query sizes. */
int32_t mtx_accel_export_tlas(mtx_handle accel, uint32_t *tlas_words, size_t tlas_capacity,
                              size_t *tlas_written, uint32_t *instance_words,
                              size_t instance_capacity, size_t *instance_written);

int32_t mtx_accel_destroy(mtx_handle accel);

/* ------------------------------------------------------------ ray queries */
typedef struct mtx_ray {
  float origin[3];
  float tmax;
  float direction[3];
  uint32_t flags;
} mtx_ray; /* 32 bytes */

/* Closest-hit record: the 32-byte record of native/host/src/hit_diagnostics.rs:135
 * with the BVH-layout-specific `record` replaced by layout-free identities. */
typedef struct mtx_hit {
  uint32_t hit;
  int32_t instance;   /* source instance index, -1 on miss */
  int32_t primitive;  /* source triangle index in the instance's mesh, -1 on miss */
  float t;
  float barycentric[3];
  uint32_t visits;    /* boxes + triangles tested (when MTX_TRACE_COUNT) */
} mtx_hit;

#define MTX_TRACE_DOUBLE_SIDED 1u /* disable back-face culling */
#define MTX_TRACE_COUNT 2u        /* fill mtx_hit.visits */

/* querySceneSoftware (intersect.wgsl.ts:660-684): closest hit for `count`
 * device-resident rays into device-resident hits on `stream` (hipStream_t,
 * NULL = default). Asynchronous; synchronize the stream before reading. */
int32_t mtx_trace_closest(mtx_handle accel, const mtx_ray *rays, uint32_t count, mtx_hit *hits,
                          uint32_t flags, void *stream);

/* intersectSceneShadow (intersect.wgsl.ts:829-879): occluded[i] = 1 when an
 * opaque hit lies within [0, min(max(tmax,0), 9999)]. Device pointers. */
int32_t mtx_trace_any(mtx_handle accel, const mtx_ray *rays, uint32_t count, uint32_t *occluded,
                      void *stream);

/* Per-ray traversal counters (box tests, triangle tests, instance visits;
 * 3 uint32 per ray, device pointer), for the roofline's algorithmic bytes. */
int32_t mtx_trace_count(mtx_handle accel, const mtx_ray *rays, uint32_t count, uint32_t *counters,
                        uint32_t flags, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MTX_H */
