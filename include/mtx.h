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

/* refitBLAS (src/renderer/bvhAccel.ts:800-870): new vertex positions of the
 * mesh the BLAS was built from (same vertex count, same index buffer, host or
 * device per MTX_INPUT_DEVICE); node boxes are recomputed bottom-up on the GPU
 * (a leaf spans the finite vertices of its triangles, an interior node its
 * children), the triangle records re-gathered, topology and leaf order kept.
 * *rebuild_advised (optional) = 1 when the refitted node surface area exceeds
 * twice the build's (DYNAMIC_BLAS_REBUILD_RATIO, bvhAccel.ts:743): a dynamic
 * mesh should then be rebuilt. Synchronous. */
int32_t mtx_blas_refit(mtx_handle blas, const float *positions, uint32_t vertex_count, const uint32_t *indices,
                       uint32_t triangle_count, uint32_t flags, uint32_t *rebuild_advised, char *error,
                       size_t error_capacity);

int32_t mtx_blas_destroy(mtx_handle blas);

/* ----------------------------------------------------- scene acceleration */
typedef struct mtx_instance_desc {
  uint32_t struct_size;
  uint32_t blas_index;      /* index into the blas array of mtx_accel_create */
  uint32_t material_word;   /* packed material word (material << 8); bit 1 set: the material is
                               alpha-masked (render scenes set it from MTX_MAT_ALPHA_MODE; the
                               raw query of mtx_trace_any does not test coverage) */
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

/* mtx_trace_closest without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * traversal stacks' spill slices and the work counters live in a static
 * device array of 72 MiB + 256 B (1024 persistent blocks x 128 lanes x 144
 * stack entries) that every device holds from library load until the process
 * ends; its device address is looked up once per device on the first call.
 * Same results as mtx_trace_closest, bit for bit. At most 2^30 rays per call
 * (MTX_INVALID_ARGUMENT beyond). Every call on one device shares that array:
 * calls that overlap on one GPU (different streams, other threads) race. */
int32_t mtx_trace_closest_noalloc(mtx_handle accel, const mtx_ray *rays, uint32_t count, mtx_hit *hits,
                                  uint32_t flags, void *stream);

/* intersectSceneShadow (intersect.wgsl.ts:829-879): occluded[i] = 1 when an
 * opaque hit lies within [0, min(max(tmax,0), 9999)]. Device pointers. */
int32_t mtx_trace_any(mtx_handle accel, const mtx_ray *rays, uint32_t count, uint32_t *occluded,
                      void *stream);

/* Per-ray traversal counters (box tests, triangle tests, instance visits;
 * 3 uint32 per ray, device pointer), for the roofline's algorithmic bytes. */
int32_t mtx_trace_count(mtx_handle accel, const mtx_ray *rays, uint32_t count, uint32_t *counters,
                        uint32_t flags, void *stream);


/* ------------------------------------------------------------- rendering */
/* Packed OpenPBR material: MTX_MATERIAL_FLOATS floats at these offsets (reference parameter
 * names and defaults: src/scene/openPbrSchema.ts:32-69). */
#define MTX_MAT_BASE_WEIGHT 0
#define MTX_MAT_BASE_COLOR 1          /* 3 floats */
#define MTX_MAT_BASE_DIFFUSE_ROUGHNESS 4
#define MTX_MAT_BASE_METALNESS 5
#define MTX_MAT_SPECULAR_WEIGHT 6
#define MTX_MAT_SPECULAR_COLOR 7      /* 3 floats (F82 tint for metals) */
#define MTX_MAT_SPECULAR_ROUGHNESS 10
#define MTX_MAT_SPECULAR_IOR 11
#define MTX_MAT_TRANSMISSION_WEIGHT 12
#define MTX_MAT_TRANSMISSION_COLOR 13 /* 3 floats */
#define MTX_MAT_COAT_WEIGHT 16
#define MTX_MAT_COAT_COLOR 17         /* 3 floats */
#define MTX_MAT_COAT_ROUGHNESS 20
#define MTX_MAT_COAT_IOR 21
#define MTX_MAT_COAT_DARKENING 22
#define MTX_MAT_EMISSION 23           /* 3 floats: materialEmission() (color * luminance, coat-tinted);
                                         instances with emitting materials are triangle lights */
#define MTX_MAT_DOUBLE_SIDED 26
#define MTX_MAT_SPECULAR_ANISOTROPY 27
#define MTX_MAT_FUZZ_WEIGHT 28
#define MTX_MAT_FUZZ_COLOR 29         /* 3 floats */
#define MTX_MAT_FUZZ_ROUGHNESS 32
#define MTX_MAT_SUBSURFACE_WEIGHT 33
#define MTX_MAT_SUBSURFACE_COLOR 34   /* 3 floats */
#define MTX_MAT_THIN_WALLED 37        /* > 0.5: thin-walled surface */
#define MTX_MAT_OPACITY 38            /* alpha (materialParams.ts:214) */
#define MTX_MAT_ALPHA_MODE 39         /* resolved mode: 0 opaque, 1 mask, 2 blend (materialConstants.ts:12-16,
                                         resolveMaterialAlphaMode, MaterialState.ts:251-264). Mask coverage =
                                         opacity x base-map a (map alpha usage 2) x alpha-map g, covered where
                                         >= the cutoff: camera and bounce rays pass uncovered points without a
                                         depth, shadow rays ignore them. A scene with a blend material is
                                         rejected with MTX_UNAVAILABLE */
#define MTX_MAT_ALPHA_CUTOFF 40       /* max(0, alphaCutoff) (materialParams.ts:218) */
#define MTX_MAT_MAP_ALPHA_USAGE 41    /* base map alpha: 0 none, 2 alpha, 3 transmission = 1 - a
                                         (mapAlphaUsageConstants, materialConstants.ts:2) */
#define MTX_MAT_NORMAL_SCALE 42       /* 2 floats: normal map xy scale */
#define MTX_MAT_UV_SCALE 44           /* 2 floats: UV0/UV1 multiplier of surface hits (uvProjectionScale,
                                         materialParams.ts:202; default 1, 1) */
#define MTX_MAT_MAPS 46               /* MTX_MAP_COUNT floats: scene texture index + 1 per slot, 0 = none */
#define MTX_MATERIAL_FLOATS 64        /* floats 60..63 are reserved (zero) */

/* Texture slots of a material, in the order their samples apply
 * (materialBuffer.wgsl.ts:78-160, slot table materialTextureSlots.ts:8-140).
 * Color slots (base, emissive, fuzz color, specular color) decode RGBA8
 * texels from sRGB; RGBA32F texels and the other slots are linear. */
#define MTX_MAP_BASE 0                /* baseColor *= rgb; a per MTX_MAT_MAP_ALPHA_USAGE */
#define MTX_MAP_ALPHA 1               /* alpha *= g */
#define MTX_MAP_TRANSMISSION 2        /* transmission weight *= r */
#define MTX_MAP_METALLIC_ROUGHNESS 3  /* metalness *= b, specular roughness *= g */
#define MTX_MAP_ROUGHNESS 4           /* specular roughness *= g */
#define MTX_MAP_METALNESS 5           /* metalness *= b */
#define MTX_MAP_COAT_WEIGHT 6         /* coat weight *= r (emission tint follows it) */
#define MTX_MAP_COAT_ROUGHNESS 7      /* coat roughness *= g */
#define MTX_MAP_FUZZ_COLOR 8          /* fuzz color *= rgb */
#define MTX_MAP_FUZZ_ROUGHNESS 9      /* fuzz roughness *= a */
#define MTX_MAP_SPECULAR_WEIGHT 10    /* specular weight *= a */
#define MTX_MAP_SPECULAR_COLOR 11     /* specular color *= rgb */
#define MTX_MAP_EMISSIVE 12           /* emission *= rgb */
#define MTX_MAP_NORMAL 13             /* tangent-space normal (cotangent frame of UV0) */
#define MTX_MAP_COUNT 14

/* Texture storage and wrap modes (mutracer-scene.h:26-27). */
#define MTX_TEXTURE_RGBA8_SRGB 0u     /* one word per pixel, R in the low byte */
#define MTX_TEXTURE_RGBA32F 1u        /* four floats per pixel */
#define MTX_WRAP_REPEAT 0u
#define MTX_WRAP_CLAMP_TO_EDGE 1u
#define MTX_WRAP_MIRRORED_REPEAT 2u

/* A material texture (mt_texture_v4, mutracer-scene.h:87-95, without UDIM):
 * row-major pixels, row 0 first (flip_y reverses the rows at upload). A
 * sample at uv reads the wrapped coordinate (uv * uv_scale + uv_offset) with
 * bilinear filtering between texel centers (sampleMap,
 * materialBuffer.wgsl.ts:314-336). */
typedef struct mtx_texture_desc {
  uint32_t struct_size;
  uint32_t width, height;       /* 1..16384 */
  uint32_t storage;             /* MTX_TEXTURE_* */
  uint32_t flip_y;              /* 0 or 1 */
  uint32_t uv_channel;          /* 0: UV0, 1: UV1 */
  uint32_t wrap_s, wrap_t;      /* MTX_WRAP_* */
  float uv_scale[2], uv_offset[2];
  const void *pixels;
} mtx_texture_desc;

/* Packed rect light (RectLight.ts:5-22): 16 floats. */
/* Packed analytic light, the fields of the reference's Light record
 * (rayTraceCore.wgsl.ts:228-248, packed as lightsData.ts:17-100). Types use
 * the reference's constants (lightConstants.ts:10-15); 0 is read as a rect.
 *   rect:        u, v edge vectors (normal = normalize(u x v)), spread in [0,1], area |u x v|
 *   sphere:      radius (0 = point light, emission = intensity), area 4 pi r^2
 *   spot:        u = unit direction, v = (cos outer, cos inner, 0), radius, area 4 pi r^2;
 *                emission = color * intensity / (pi r^2) (r > 0) or color * intensity
 *   directional: u = unit direction the light travels
 *   disk:        u = x axis * radius, v = y axis * -radius (normal = normalize(u x v)), area pi r^2
 *   cylinder:    u = axis * length, v = radial axis * radius, radius, area 2 pi r length
 * Disk and cylinder emission is their radiance (intensity / area when normalized). */
#define MTX_LIGHT_DIRECTIONAL 1
#define MTX_LIGHT_SPHERE 2
#define MTX_LIGHT_RECT 3
#define MTX_LIGHT_SPOT 4
#define MTX_LIGHT_DISK 5
#define MTX_LIGHT_CYLINDER 6
#define MTX_LIGHT_TYPE 0              /* one of MTX_LIGHT_* */
#define MTX_LIGHT_POSITION 1          /* center, 3 floats */
#define MTX_LIGHT_U 4                 /* 3 floats, per type */
#define MTX_LIGHT_V 7                 /* 3 floats, per type */
#define MTX_LIGHT_EMISSION 10         /* 3 floats */
#define MTX_LIGHT_SPREAD 13           /* rect spread */
#define MTX_LIGHT_RADIUS 13           /* sphere / spot / cylinder radius */
#define MTX_LIGHT_AREA 14
#define MTX_LIGHT_FLOATS 16

typedef struct mtx_mesh_desc {
  uint32_t struct_size;
  uint32_t vertex_count;
  uint32_t triangle_count;
  uint32_t reserved;
  const float *positions;   /* vertex_count * 3 */
  const float *normals;     /* vertex_count * 3 */
  const uint32_t *indices;  /* triangle_count * 3 */
  const float *uvs;         /* vertex_count * 2, or NULL (UV0 = 0) */
  const float *uvs1;        /* vertex_count * 2, or NULL (UV1 = 0; emissive triangles then use UV0) */
} mtx_mesh_desc;
/* A struct_size that ends before `uvs` is accepted: the mesh has no UVs. */

typedef struct mtx_scene_desc {
  uint32_t struct_size;
  uint32_t mesh_count;
  const mtx_mesh_desc *meshes;
  uint32_t instance_count;  /* instances[i].blas_index is a mesh index */
  const mtx_instance_desc *instances;
  uint32_t material_count;
  const float *materials;   /* MTX_MATERIAL_FLOATS each */
  uint32_t light_count;
  const float *lights;      /* MTX_LIGHT_FLOATS each */
  uint32_t env_width, env_height;
  const float *env_rgba;    /* env_height rows of env_width RGBA f32, or NULL */
  uint32_t blas_bins;       /* 8 = reference default */
  uint32_t tlas_bins;       /* 64 = web TLAS */
  uint32_t flags;           /* MTX_ACCEL_PER_INSTANCE_BOUNDS */
  uint32_t texture_count;   /* material textures (MTX_MAT_MAPS index them) */
  const mtx_texture_desc *textures;
} mtx_scene_desc;
/* A struct_size that ends before `texture_count` is accepted: no textures. */

/* Scene commit (the GPU side of Engine.update / WGPURayTracePass.setMeshes,
 * WGPURayTracePass.ts:2246, setLights :3180): builds every BLAS on the GPU,
 * the TLAS, the light alias table (lightSampling.ts:172-212) and the
 * environment CDF (envMapDistribution.ts:57-134), and keeps them resident. */
int32_t mtx_scene_create(const mtx_scene_desc *desc, mtx_handle *out_scene, char *error, size_t error_capacity);

/* Transform-only edit of the listed source instances (refitTLAS). */
int32_t mtx_scene_set_transforms(mtx_handle scene, const double *transforms, uint32_t instance_count,
                                 const uint32_t *dirty, uint32_t dirty_count, char *error, size_t error_capacity);

/* Vertex deformation of mesh `mesh` of a committed scene (refitBLAS plus
 * refitTLAS, as WGPURayTracePass.ts:2735 takes the refit path when the
 * topology is unchanged): new object-space positions and normals of the
 * commit's vertex count; the mesh's BLAS is refitted, every instance of it
 * takes the new root box and the TLAS is refitted. Emissive triangles keep
 * the world-space list of the commit. *rebuild_advised as mtx_blas_refit. */
int32_t mtx_scene_update_mesh(mtx_handle scene, uint32_t mesh, const float *positions, const float *normals,
                              uint32_t vertex_count, uint32_t *rebuild_advised, char *error, size_t error_capacity);

/* info[0] triangles, [1] instances, [2] BLAS count, [3] resident MiB. */
int32_t mtx_scene_info(mtx_handle scene, uint64_t info[4]);
int32_t mtx_scene_destroy(mtx_handle scene);

typedef struct mtx_target_settings {
  uint32_t struct_size;
  uint32_t width, height;
  uint32_t bounces;                 /* RayTraceUniform.bounces */
  uint32_t direct_light_candidates; /* NEE RIS candidates (render.ts:21) */
  uint32_t sampling_pattern;        /* 0 sobol, 1 zsobol, 2 random: pcg4d per path with a Halton camera
                                       jitter per sample index (sampling.ts:3-8, rand.wgsl.ts:8-23) */
  uint32_t sequence_seed;
  uint32_t max_samples;             /* ZSobol sample budget */
  uint32_t sample_sequence_offset;  /* Render.ts:151 */
  float clamping;                   /* 0 = off */
  uint32_t kernel;                  /* MTX_KERNEL_WAVEFRONT (0) or MTX_KERNEL_MEGAKERNEL (1) */
  uint32_t max_paths_in_flight;     /* wavefront generation size; 0 = 8M paths */
} mtx_target_settings;

#define MTX_KERNEL_WAVEFRONT 0u
#define MTX_KERNEL_MEGAKERNEL 1u

/* Camera block (cameraUniform.ts:2-17, browser projection-inverse form). */
typedef struct mtx_camera_desc {
  uint32_t struct_size;
  float projection_inverse[16];
  float view_inverse[16];
  float aperture_radius;
  float focus_distance;
} mtx_camera_desc;

/* Render target: resident FP64 per-pixel running sums (color RGBA, albedo,
 * normal, luminance^2) and the per-pixel sample counter that gives each
 * pixel's next Sobol index (rayTrace.compute.ts:1188). Replaces
 * RenderTarget.trace (RenderingPipeline.ts:218) / WGPURayTracePass.createPass. */
int32_t mtx_target_create(mtx_handle scene, const mtx_target_settings *settings, const mtx_camera_desc *camera,
                          mtx_handle *out_target, char *error, size_t error_capacity);

/* Adds `samples` path samples to every pixel on `stream` (NULL = default);
 * asynchronous. The target's kernel setting chooses the work: the wavefront
 * form (per bounce: trace, shade, shadow and compaction kernels, replayed
 * from a HIP graph) or one megakernel launch per call. */
int32_t mtx_target_render(mtx_handle target, uint32_t samples, void *stream);

/* mtx_target_render plus per-launch totals of the traversal work, in two
 * groups of MTX_WORK_GROUP counters: work[0..8] for the wavefront trace and
 * shadow kernels (or the megakernel), work[9..17] for the wavefront tail
 * kernel. Within a group: [0] closest-hit rays, [1] closest-hit box tests,
 * [2] closest-hit triangle tests, [3] closest-hit instance visits, [4]
 * closest hits, [5] shadow rays, [6] shadow box tests, [7] shadow triangle
 * tests, [8] shadow instance visits. Synchronous. No reference counterpart:
 * it feeds the algorithmic-bytes figure of the roofline. */
#define MTX_WORK_GROUP 9
#define MTX_WORK_COUNTERS 18
int32_t mtx_target_render_counted(mtx_handle target, uint32_t samples, uint64_t work[MTX_WORK_COUNTERS],
                                  void *stream);

/* mtx_target_render with HIP events around every kernel launch on `stream`;
 * returns the summed GPU time per kernel class: ms[0] generate + accumulate,
 * [1] closest-hit trace, [2] shade, [3] shadow trace, [4] megakernel, [5]
 * wavefront tail kernel. Synchronous. No reference counterpart (measurement). */
#define MTX_TIMER_CLASSES 6
int32_t mtx_target_render_timed(mtx_handle target, uint32_t samples, double ms[MTX_TIMER_CLASSES], void *stream);

/* Blocks until the target's submitted work is done. */
int32_t mtx_target_synchronize(mtx_handle target);

#define MTX_PIXELS_COLOR 0   /* readPixels('hdr'): RGB / hit count (or / samples without hits), A / samples */
#define MTX_PIXELS_ALBEDO 1  /* readPixels('albedo'): RGB / samples, A = clamp(hits / samples, 0, 1) */
#define MTX_PIXELS_NORMAL 2  /* readPixels('normal'): RGB / samples * 0.5 + 0.5, A = 1 (0 where |sum| is ~0) */
#define MTX_PIXELS_VARIANCE 3 /* luminance variance of the mean in R (no reference counterpart) */

/* Resolved RGBA f32, top row first (Render.readPixels, Render.ts:782-801),
 * normalized as the output pass does (WGPURayTraceOutput.ts:32-92,
 * hdr_resolve_component in compose.frag.wgsl.ts:264-272); pixels without
 * samples are 0. pixels == NULL queries the float count through *written. */
int32_t mtx_target_read_pixels(mtx_handle target, uint32_t kind, float *pixels, size_t capacity_floats,
                               size_t *written);

/* Raw FP64 sums, 13 per pixel in bottom-row-first order (color rgba, albedo
 * rgba, normal rgba, luminance^2), plus samples per pixel. */
int32_t mtx_target_read_accumulators(mtx_handle target, double *sums, size_t capacity_doubles, uint32_t *counts,
                                     size_t capacity_counts);

/* Launch shape of the wavefront kernels (what the autotuner searches). No
 * reference counterpart: the reference's equivalent knobs are its tile size
 * and time budget (TileRender.ts:302-490). */
typedef struct mtx_wavefront_tuning {
  uint32_t struct_size;
  uint32_t max_paths_in_flight;    /* paths per generation (0 = keep) */
  uint32_t trace_kernel;           /* 0 one ray per thread, 1 persistent lane-refill walker */
  uint32_t lds_stack;              /* traversal LDS stack entries per lane: 16, 24 or 32 */
  uint32_t trace_oversubscription; /* trace grid = CUs x resident blocks x this (>= 1) */
  uint32_t shade_oversubscription; /* shade grid likewise (>= 1) */
  uint32_t shade_waves_per_simd;   /* register budget of the shade kernel: 1..4 resident waves per SIMD */
  uint32_t use_graphs;             /* 1: capture each generation's launches into a hipGraph, replay while unchanged */
  uint32_t graph_replays;          /* get only: generations launched from a graph */
  uint32_t graph_captures;         /* get only: graphs captured */
  uint32_t tail_paths;             /* once an event's queue holds at most this many paths, one kernel
                                      finishes them all (0 = never) */
  uint32_t queue_sort;             /* 1: group each event's continuation queue by ray direction octant */
  uint32_t lds_stack_closest;      /* LDS stack entries of the closest-hit trace kernel: 0 (= lds_stack),
                                      16, 24 or 32; lds_stack then sets the any-hit kernel's */
} mtx_wavefront_tuning;
int32_t mtx_target_set_tuning(mtx_handle target, const mtx_wavefront_tuning *tuning);
int32_t mtx_target_get_tuning(mtx_handle target, mtx_wavefront_tuning *tuning);

/* Adaptive sampling (two-stage converge, Render.ts:886-895): after a pilot
 * render, plan per-pixel sample targets from the FP64 luminance moments
 * (offlineAdaptive.compute.ts:29-106, options of OfflineAdaptiveSampling.ts
 * :75-108): per tile of tile_size^2 pixels, the largest unbiased per-pixel
 * variance v and the mean luminance m give
 *   target = clamp(ceil(confidence^2 * v / (absolute_error + relative_error*|m|)^2),
 *                  min_samples, max_samples).
 * The plan is installed as the target's per-pixel limits: later renders skip
 * pixels whose sample count has reached their target (rayTrace.compute.ts
 * :1171-1180). Zero fields take the reference defaults (min 16 capped at max,
 * tile 16, absolute 0.002, relative 0.03, confidence 1.96). */
typedef struct mtx_adaptive_options {
  uint32_t struct_size;
  uint32_t max_samples, min_samples, tile_size;
  double absolute_error, relative_error, confidence;
} mtx_adaptive_options;
/* plan[0] = largest target, plan[1] = average target per pixel. tile_variance
 * (columns x rows doubles, optional) receives each tile's variance. */
int32_t mtx_target_plan_adaptive(mtx_handle target, const mtx_adaptive_options *options, double plan[2],
                                 double *tile_variance, size_t tile_capacity);
/* Installs caller targets (width*height floats, bottom row first) or, with
 * NULL, removes the limits. */
int32_t mtx_target_set_sample_targets(mtx_handle target, const float *targets, size_t count);

/* stats[0] rays traced, [1] samples per pixel completed. */
int32_t mtx_target_stats(mtx_handle target, uint64_t stats[2]);
int32_t mtx_target_set_camera(mtx_handle target, const mtx_camera_desc *camera);
/* setSampleSequenceOffset (WGPURayTracePass.ts:2085): first Sobol index of later samples. */
int32_t mtx_target_set_sample_offset(mtx_handle target, uint32_t sample_sequence_offset);
int32_t mtx_target_reset(mtx_handle target);
int32_t mtx_target_destroy(mtx_handle target);

/* ------------------------------------------------------- display and JNLM */
/* Display composition of a render target's resolved HDR image (compose.frag
 * .wgsl.ts:100-142 without bloom, flat backgrounds and shadow catchers):
 * exposure, gradeColor (contrast, saturation, temperature, tint), vignette,
 * tone mapping (TONE_MAPPING_MODES: AgX 0, ACES 1, Cineon 2, Reinhard 3,
 * linear 4), then sRGB (or gamma 2.2) encoding. Defaults of
 * src/settings/display.ts:12-18. */
#define MTX_TONE_AGX 0u
#define MTX_TONE_ACES 1u
#define MTX_TONE_CINEON 2u
#define MTX_TONE_REINHARD 3u
#define MTX_TONE_LINEAR 4u
typedef struct mtx_display_settings {
  uint32_t struct_size;
  uint32_t tone_mapping;  /* MTX_TONE_* (default AgX) */
  float exposure;         /* default 1 */
  float contrast;         /* default 1 */
  float saturation;       /* default 1 */
  float temperature;      /* default 0 */
  float tint;             /* default 0 */
  float vignette;         /* default 0 */
  uint32_t gamma22;       /* 0: sRGB encoding (default), 1: gamma 2.2 */
} mtx_display_settings;
int32_t mtx_display_settings_default(mtx_display_settings *output);

/* The displayed image, RGBA8 (RGB encoded, A = coverage), top row first:
 * width*height*4 bytes; rgba8 == NULL queries the size through *written. */
int32_t mtx_target_compose(mtx_handle target, const mtx_display_settings *display, uint8_t *rgba8, size_t capacity,
                           size_t *written);

/* Joint non-local-means denoiser (src/renderer/jnlm.ts, wgsl/jnlm.compute.ts):
 * the per-pixel luminance mean and variance of the mean from the FP64
 * moments, a 3x3 variance smoothing pass, then `iterations` filter passes of
 * the composed display-linear image with doubling dilation, guided by the
 * albedo (relative contrast, coverage) and encoded normals, with the bounded
 * albedo regression; the result is sRGB-encoded (jnlmEncodeWGSL). Defaults
 * of jnlm.ts:23-46 (strength 1, search radius 5, patch radius 1, albedo
 * bandwidth 0.4, normal bandwidth 0.05, 2 iterations). */
typedef struct mtx_jnlm_options {
  uint32_t struct_size;
  float strength;            /* > 0 */
  uint32_t search_radius;    /* 1..16 */
  uint32_t patch_radius;     /* 0..3 */
  float albedo_bandwidth;    /* 0.0001..10 */
  float normal_bandwidth;    /* 0.0001..10 */
  uint32_t iterations;       /* 1..4 */
} mtx_jnlm_options;
int32_t mtx_jnlm_options_default(mtx_jnlm_options *output);
/* Filters the target's composed image. linear_rgba (optional, width*height*4
 * f32, top row first) receives the filtered display-linear RGBA; rgba8
 * (optional, width*height*4 bytes) the encoded result; `pixels` is the pixel
 * capacity of each given buffer. */
int32_t mtx_target_denoise_jnlm(mtx_handle target, const mtx_display_settings *display,
                                const mtx_jnlm_options *options, float *linear_rgba, uint8_t *rgba8,
                                size_t pixels, char *error, size_t error_capacity);

/* ------------------------------------------------------------ context API */
/* Persistent contexts over the raw scene descriptor (versions 5..8,
 * native/docs/scene-format.md), with the semantics of mutracer-context.h:12-187:
 * commit copies and validates a descriptor (INVALID_ARGUMENT names the
 * problem; a feature this build does not render returns UNAVAILABLE), a failed
 * commit keeps the previous scene, every successful commit restarts
 * accumulation; blocking renders and one asynchronous job per context
 * (submit / poll / info / cancel); read_pixels returns the accumulated sums
 * divided by the completed sample count, RGBA f32, bottom row first. A GPU
 * failure during render, poll, read or reset marks the context failed and
 * returns MTX_CONTEXT_FAILED; later calls return MTX_INVALID_STATE until a
 * successful commit. Calls may come from any thread; an operation that
 * overlaps another on the same context returns MTX_BUSY. Every context owns
 * one HIP stream (default priority) for its GPU work. */

/* mt_render_settings (mutracer-context.h:110-127), 64 bytes, same field
 * order and meaning. bevel_samples, caustic_flags, the photon fields and
 * transmissive_output_alpha_min below 1 select estimators this build does
 * not have: nonzero caustic flags or photon counts and an alpha minimum other
 * than 1 return MTX_UNAVAILABLE. */
typedef struct mtx_render_settings {
  uint32_t struct_size, sampling_pattern, seed, max_samples;
  uint32_t bounces, direct_light_candidates, bevel_samples, sample_sequence_offset;
  double clamping, transmissive_output_alpha_min;
  uint32_t caustic_flags;
  uint32_t light_photons, photon_map_photons;
  float photon_map_radius;
} mtx_render_settings;
#if defined(__cplusplus) && __cplusplus >= 201103L
static_assert(sizeof(mtx_render_settings) == 64, "mtx_render_settings layout");
#elif defined(__STDC_VERSION__) && __STDC_VERSION__ >= 201112L
_Static_assert(sizeof(mtx_render_settings) == 64, "mtx_render_settings layout");
#endif

/* mt_render_settings_default (mutracer-context.h:128-129): Sobol, seed 12345,
 * max_samples 64, bounces 6, candidates 2, bevel samples 8, offset 0,
 * clamping 0, transmissive alpha minimum 1. No device needed. */
int32_t mtx_render_settings_default(mtx_render_settings *output);

/* mt_context_snapshot (mutracer-context.h:57-62), 24 bytes. */
typedef struct mtx_context_snapshot {
  uint32_t struct_size;
  uint32_t width, height, samples;
  uint32_t state;       /* MTX_CONTEXT_* */
  uint32_t last_commit; /* MTX_COMMIT_* */
} mtx_context_snapshot;
#define MTX_CONTEXT_EMPTY 0u
#define MTX_CONTEXT_COMMITTED 1u
#define MTX_CONTEXT_FAILED_STATE 2u
#define MTX_CONTEXT_JOB_ACTIVE 3u
#define MTX_COMMIT_NONE 0u
#define MTX_COMMIT_FULL 1u
#define MTX_COMMIT_INCREMENTAL 2u /* camera-only or mesh-transform-only edit (TLAS refit) */

/* mt_job_snapshot (mutracer-context.h:168-172), 24 bytes. */
typedef struct mtx_job_snapshot {
  uint32_t struct_size;
  uint32_t state; /* MTX_JOB_* */
  uint32_t requested_samples, submitted_samples, completed_samples, base_samples;
} mtx_job_snapshot;

/* mt_image_kind values accepted by mtx_context_read_pixels. */
#define MTX_IMAGE_HDR 0u
#define MTX_IMAGE_NORMAL 1u
#define MTX_IMAGE_ALBEDO 2u

/* mt_context_create / mt_context_destroy (mutracer-context.h:78,81).
 * Destroy rejects an active job (MTX_INVALID_STATE) unless the context failed. */
int32_t mtx_context_create(mtx_handle *output, char *error, size_t error_capacity);
int32_t mtx_context_destroy(mtx_handle context, char *error, size_t error_capacity);
/* mt_context_commit / mt_context_commit_with_settings (:92,130). A commit
 * while a job is active waits for its in-flight batch and expires the job. */
int32_t mtx_context_commit(mtx_handle context, const uint8_t *input, size_t length, char *error,
                           size_t error_capacity);
int32_t mtx_context_commit_with_settings(mtx_handle context, const uint8_t *input, size_t length,
                                         const mtx_render_settings *settings, char *error, size_t error_capacity);
/* mt_context_render (:132): adds samples and waits; total 1..16384. */
int32_t mtx_context_render(mtx_handle context, uint32_t additional_samples, char *error, size_t error_capacity);
/* mt_context_reset (:133). */
int32_t mtx_context_reset(mtx_handle context, char *error, size_t error_capacity);
/* mt_context_wait (:140-141): waits up to timeout_ms for submitted GPU work
 * (MTX_WAIT_INDEFINITELY: no deadline, 0: only checks); MTX_TIMEOUT on expiry. */
#define MTX_WAIT_INDEFINITELY UINT32_MAX
int32_t mtx_context_wait(mtx_handle context, uint32_t timeout_ms, char *error, size_t error_capacity);
/* mt_context_info (:142). info->struct_size must be sizeof(mtx_context_snapshot). */
int32_t mtx_context_info(mtx_handle context, mtx_context_snapshot *info, char *error, size_t error_capacity);
/* mt_context_read_pixels (:143-144): kind MTX_IMAGE_*, `size` in bytes,
 * *written = width*height*16 (required). pixels == NULL with size 0 queries
 * the size; a short destination returns MTX_BUFFER_TOO_SMALL. */
int32_t mtx_context_read_pixels(mtx_handle context, uint32_t kind, void *pixels, size_t size, size_t *written,
                                char *error, size_t error_capacity);
/* mt_context_submit / mt_job_poll / mt_job_info / mt_job_cancel (:173-179).
 * submit records the request; poll (max_new_samples 0..64) credits a finished
 * batch without waiting, then dispatches the next batch asynchronously on the
 * context's stream when none is in flight; info reports the last polled
 * state; cancel stops queued work at once (an in-flight batch finishes:
 * MTX_JOB_CANCELLING until the next poll). A job handle expires on the next
 * submit, blocking render, reset, commit or context destruction. */
int32_t mtx_context_submit(mtx_handle context, uint32_t additional_samples, mtx_handle *job, char *error,
                           size_t error_capacity);
int32_t mtx_job_poll(mtx_handle job, uint32_t max_new_samples, char *error, size_t error_capacity);
int32_t mtx_job_info(mtx_handle job, mtx_job_snapshot *info, char *error, size_t error_capacity);
int32_t mtx_job_cancel(mtx_handle job, char *error, size_t error_capacity);
/* The validation and decoding mt_context_commit performs, without a device
 * or a context (native/host/src/scene.rs validate, scene-format.md rules).
 * info (optional): [0] width, [1] height, [2] BLAS meshes after splitting
 * multi-material geometry, [3] instances, [4] materials, [5] analytic lights,
 * [6] environment width, [7] environment height. */
int32_t mtx_descriptor_validate(const uint8_t *input, size_t length, uint32_t info[8], char *error,
                                size_t error_capacity);
/* Test hook, no reference counterpart beyond its test-only injected flight
 * failures (context.rs inject_flight_failure): the context's next GPU
 * operation (render, poll dispatch, read or reset) fails as a GPU error would,
 * before anything is launched, so failure handling can be tested without
 * faulting a device. */
int32_t mtx_context_inject_failure(mtx_handle context);

/* -------------------------------------------------------------- serving */
/* Batched render server: many render jobs (mixed scenes, resolutions, sample
 * counts) share every wavefront generation. The reference's native job API
 * runs one job per context and one sample batch per dispatch
 * (mt_context_submit / mt_job_poll / mt_job_info / mt_job_cancel,
 * mutracer-context.h:146-187); here a round gives every running job a slice
 * of the round's path budget, so new jobs join while others are in flight
 * (continuous batching). A job's result is identical to rendering it alone:
 * its samples keep their Sobol indices and FP64 accumulation order. */
typedef struct mtx_server_config {
  uint32_t struct_size;
  uint32_t max_paths_per_round;    /* path budget of one round (0 = 2^24) */
  uint32_t max_jobs_per_round;     /* 1 = one job at a time (the plain form); 0 = 1024 */
  uint32_t max_samples_per_job_round; /* per-job samples per round (0 = 64, as mt_job_poll) */
  uint64_t hbm_budget_bytes;       /* accumulators of admitted jobs (0 = 200 GB) */
} mtx_server_config;

/* Job states: values of mt_job_state (mutracer-status.h:123-126). */
#define MTX_JOB_QUEUED 0u
#define MTX_JOB_RUNNING 1u
#define MTX_JOB_COMPLETE 2u
#define MTX_JOB_CANCELLED 3u
#define MTX_JOB_FAILED 4u
#define MTX_JOB_CANCELLING 5u /* context jobs: cancelled with a batch still in flight */

typedef struct mtx_served_job_info {
  uint32_t struct_size;
  uint32_t state;
  uint32_t requested_samples, submitted_samples, completed_samples, base_samples;
  double submit_ms, start_ms, complete_ms; /* host steady clock, milliseconds */
} mtx_served_job_info;

int32_t mtx_server_create(const mtx_server_config *config, mtx_handle *out_server, char *error, size_t error_capacity);
/* Queues a job of `samples` samples per pixel on a committed scene (inputs copied). */
int32_t mtx_server_submit(mtx_handle server, mtx_handle scene, const mtx_target_settings *settings,
                          const mtx_camera_desc *camera, uint32_t samples, mtx_handle *out_job, char *error,
                          size_t error_capacity);
/* Runs one round on `stream` and waits for it: admits queued jobs in arrival
 * order while their accumulators fit the HBM budget, gives each running job
 * up to max_samples_per_job_round samples within the path budget, launches
 * one wavefront generation over all of them. *completed = jobs finished by
 * this round; *idle = 1 when there was nothing to run. */
int32_t mtx_server_step(mtx_handle server, void *stream, uint32_t *completed, uint32_t *idle);
int32_t mtx_server_job_info(mtx_handle server, mtx_handle job, mtx_served_job_info *info);
int32_t mtx_server_job_read_pixels(mtx_handle server, mtx_handle job, uint32_t kind, float *pixels,
                                   size_t capacity_floats, size_t *written);
int32_t mtx_server_job_read_accumulators(mtx_handle server, mtx_handle job, double *sums, size_t capacity_doubles,
                                         uint32_t *counts, size_t capacity_counts);
/* Queued jobs stop at once; running jobs stop after the current round. */
int32_t mtx_server_job_cancel(mtx_handle server, mtx_handle job);
/* Frees a finished job's accumulators (returns them to the HBM budget). */
int32_t mtx_server_job_release(mtx_handle server, mtx_handle job);
/* stats[0] rounds, [1] rays traced, [2] running jobs, [3] queued jobs, [4] admitted bytes. */
int32_t mtx_server_stats(mtx_handle server, uint64_t stats[5]);
/* stats[0] rounds launched from a captured HIP graph (a round whose job
 * count, path and pixel totals, event limit and lobe set match the
 * previous round's replays its capture), [1] captures. No reference
 * counterpart (measurement). */
int32_t mtx_server_graph_stats(mtx_handle server, uint64_t stats[2]);
int32_t mtx_server_destroy(mtx_handle server);

/* ------------------------------------------------------------- denoiser */
/* OIDN RT U-Nets (the TZA weights the reference ships) on MFMA with FP32
 * accumulation: the "UNet" shape of the standard and small models
 * (rt_hdr_alb_nrm, rt_hdr_alb_nrm_small) and the "UNetLarge" shape of the
 * clean-auxiliary model and its guide prefilters (rt_hdr_calb_cnrm_large,
 * rt_alb_large, rt_nrm_large), told apart by their layer names, channel
 * counts read from the weights. Replaces mt_denoiser_create/run/destroy
 * (mutracer-denoise.h:41-53) and the browser path's tileExecute calls
 * (WGPURenderingPipeline.ts:1352-1476). */
#define MTX_DENOISE_FP16 1u
#define MTX_DENOISE_BF16 2u
/* weights: the bytes of a version-2 tensor archive (.tza). */
/* mt_denoiser_create (mutracer-denoise.h:36-37), same signature: weights_path
 * is the .tza network to load (the reference's `library` argument names the
 * OIDN runtime; here the runtime is this library and the argument selects the
 * weights); NULL uses the MTX_DENOISER_WEIGHTS environment variable, then the
 * standard network shipped beside libmtx.so (data/rt_hdr_alb_nrm.tza).
 * backend: MTX_DENOISE_BACKEND_AUTO or _HIP run the FP16 MFMA network; _CPU,
 * _METAL and _CUDA return MTX_UNAVAILABLE (explicit backends never fall back). */
#define MTX_DENOISE_BACKEND_AUTO 0u
#define MTX_DENOISE_BACKEND_CPU 1u
#define MTX_DENOISE_BACKEND_METAL 2u
#define MTX_DENOISE_BACKEND_CUDA 3u
#define MTX_DENOISE_BACKEND_HIP 4u
int32_t mtx_denoiser_create(const char *weights_path, uint32_t backend, mtx_handle *output, char *error,
                            size_t error_capacity);
/* mt_denoiser_backend (mutracer-denoise.h:38): always MTX_DENOISE_BACKEND_HIP. */
int32_t mtx_denoiser_backend(mtx_handle denoiser, uint32_t *output, char *error, size_t error_capacity);
/* The network from TZA bytes at a chosen precision (MTX_DENOISE_FP16 or
 * MTX_DENOISE_BF16). No reference counterpart: it exposes the precision the
 * reference leaves to its runtime ('auto'). */
int32_t mtx_denoiser_create_from_weights(const void *weights, size_t size, uint32_t precision,
                                         mtx_handle *out_denoiser, char *error, size_t error_capacity);
/* mt_denoiser_run: color, normal, albedo, output are RGBA f32 (host), count =
 * width*height*4; OIDN's RT input processing: HDR color times the input
 * scale (autoexposure unless set) through the 'pu' transfer, albedo clamped
 * to [0, 1], normals in [-1, 1] mapped to [0, 1]; output alpha = color
 * alpha; output may alias an input; nothing is written on failure. */
int32_t mtx_denoiser_run(mtx_handle denoiser, uint32_t width, uint32_t height, const float *color,
                         const float *normal, const float *albedo, float *output, size_t count, char *error,
                         size_t error_capacity);
/* The network alone: input H x W x C f32 (host, H and W multiples of 16, C
 * the network's input channels, 9 or 3) -> output H x W x 3 f32;
 * layer_outputs (one pointer per convolution, or NULL) receive each
 * convolution's output (H_l x W_l x C_l f32, NULL entries skipped). No
 * reference counterpart: it exposes the per-layer precision. */
int32_t mtx_denoiser_forward(mtx_handle denoiser, uint32_t width, uint32_t height, const float *input,
                             float *output, float *const *layer_outputs, char *error, size_t error_capacity);
/* Average GPU time of `iterations` network passes at width x height (HIP events). */
int32_t mtx_denoiser_benchmark(mtx_handle denoiser, uint32_t width, uint32_t height, uint32_t iterations,
                               double *ms_per_forward);
/* Launch shape of layer `layer`'s convolution (network order): each
 * wave computes pixel_tiles x 16 pixels of a row (1, 2, 4 or 8) for
 * channel_tiles of the layer's 16-channel output tiles (1..7); pixel_tiles
 * or'ed with MTX_DENOISER_STAGED selects the kernel whose blocks of four
 * waves (four rows) stage their input rows and weights in LDS; 0, 0 restores
 * the automatic choice. Every shape gives the same bits. No reference
 * counterpart: the denoiser autotuner's knob. */
#define MTX_DENOISER_STAGED 16u
int32_t mtx_denoiser_set_tiling(mtx_handle denoiser, uint32_t layer, uint32_t pixel_tiles, uint32_t channel_tiles);
/* The launch shape `layer` uses at width x height. */
int32_t mtx_denoiser_get_tiling(mtx_handle denoiser, uint32_t layer, uint32_t width, uint32_t height,
                                uint32_t *pixel_tiles, uint32_t *channel_tiles);
/* Prefilters of the guides (the clean-auxiliary pipeline,
 * WGPURenderingPipeline.ts:1352-1440 and oidn_webgpu.rs:84-150): two other
 * denoisers made from 3-channel networks (rt_alb_*, rt_nrm_*) of this one's
 * precision. mtx_denoiser_run then passes the albedo (sRGB-encoded, clamped
 * to [0, 1]) and the normal (mapped to [0, 1]) through them
 * (oidnAuxiliary.wgsl.ts:2-16), decodes the results (sRGB to linear; [0, 1]
 * back to [-1, 1]) and uses them as this network's guides. Both 0 removes
 * them. No C-ABI counterpart: the reference's loader takes them as
 * auxiliaryModelUrls (oidnNetwork.ts:94-141). */
int32_t mtx_denoiser_set_auxiliary(mtx_handle denoiser, mtx_handle albedo_filter, mtx_handle normal_filter,
                                   char *error, size_t error_capacity);
/* Image tiling of mtx_denoiser_run (oidn-webgpu's maxTileSize / tileOverlap,
 * WGPURenderingPipeline.ts:1330-1342): tiles of max_tile_size pixels (0 = the
 * whole image in one pass, the default), each run on its region grown by
 * `overlap` pixels per side (clipped to the image) and kept inside the tile;
 * multiples of 16. A tile whose region covers the image gives the untiled
 * result; smaller regions differ near tile edges (the network sees less
 * context there). */
int32_t mtx_denoiser_set_tiles(mtx_handle denoiser, uint32_t max_tile_size, uint32_t overlap);
/* Scale of the HDR color before the transfer (OIDN's inputScale; the output
 * is divided by it): NaN (the default) computes OIDN's autoexposure of each
 * image, 0.18 over the geometric mean of its 16 x 16-bin mean luminances. */
int32_t mtx_denoiser_set_input_scale(mtx_handle denoiser, float scale);
/* info[0] convolution layers (16 or 19), [1] input channels (9 or 3), [2] 1
 * for the UNetLarge shape, [3] precision (MTX_DENOISE_FP16 / _BF16). */
int32_t mtx_denoiser_info(mtx_handle denoiser, uint32_t info[4]);
/* mt_denoiser_destroy (mutracer-denoise.h:53); MTX_BUSY while a run is in progress. */
int32_t mtx_denoiser_destroy(mtx_handle denoiser, char *error, size_t error_capacity);

/* ------------------------------------------------------------ diagnostics */
/* Evaluates the path tracer's elementary functions (include/mtx_fmath.h) on
 * the device: kind 0 sin(x), 1 cos(x), 2 pow(x, y), 3 atan2(x, y) with x the
 * ordinate, 4 acos(x), 5 exp(x), 6 and 7 the sine and cosine parts of sincos(x). Device pointers, n elements, asynchronous on `stream`.
 * No reference counterpart: it lets hosts check that device and host builds
 * of these functions agree bit for bit. */
int32_t mtx_math_eval(uint32_t kind, const float *x, const float *y, float *out, uint64_t n, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MTX_H */
