/*
 * rayon_hip.h — C ABI of the MI355X path-tracing library (librayon_hip.so).
 *
 * The reference drives its GPU render mode through the extern "C" functions of
 * src/rayon/gpu_renderers/renderer_cuda_device.cuh:8-68 plus the C++ statics
 * Scene::CudaSceneBuilder::buildGPUScene/freeGPUScene (scenes/scene_builder.hpp:234-239).
 * This header is the re-designed, reentrant boundary: every call takes a context,
 * scenes and jobs are opaque handles owned by the library, every function returns a
 * rayon_status and leaves a thread-local message for rayon_last_error().
 * The reference's own entry-point names are exported as well; they are declared in
 * rayon_legacy.h and implemented on top of this API.
 *
 * Plain pointers and sizes only: no C++ or framework types cross this boundary.
 */
#ifndef RAYON_HIP_H
#define RAYON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RAYON_ABI_VERSION 1

typedef enum rayon_status {
   RAYON_OK = 0,
   RAYON_ERR_INVALID = 1,   /* bad argument, size or handle                              */
   RAYON_ERR_HIP = 2,       /* a HIP runtime call failed (message names it)               */
   RAYON_ERR_OOM = 3,       /* device or host allocation failed                          */
   RAYON_ERR_BUDGET = 4,    /* admission refused: the HBM budget would be exceeded       */
   RAYON_ERR_IO = 5,        /* file could not be read                                    */
   RAYON_ERR_PARSE = 6,     /* scene file could not be parsed                            */
   RAYON_ERR_UNSUPPORTED = 7
} rayon_status;

/* ---------------------------------------------------------------------------------
 * Scene description (host memory, read during rayon_scene_create only).
 * ------------------------------------------------------------------------------- */

/* Material types. Numbering follows CudaScene::MaterialType (cuda_scene.cuh:26-41);
 * the last two are additions of this library (no counterpart in the reference). */
enum {
   RAYON_MAT_LAMBERTIAN = 0,
   RAYON_MAT_METAL = 1,
   RAYON_MAT_MIRROR = 2,
   RAYON_MAT_ROUGH_MIRROR = 3,
   RAYON_MAT_GLASS = 4,
   RAYON_MAT_DIELECTRIC = 5,
   RAYON_MAT_LIGHT = 6,
   RAYON_MAT_CONSTANT = 7,
   RAYON_MAT_SHOW_NORMALS = 8,
   RAYON_MAT_SDF_MATERIAL = 9,
   RAYON_MAT_ANISOTROPIC_METAL = 10,
   RAYON_MAT_THIN_FILM = 11,
   RAYON_MAT_CLEAR_COAT = 12,
   RAYON_MAT_GGX_CONDUCTOR = 13,  /* isotropic GGX conductor, VNDF sampled             */
   RAYON_MAT_GGX_DIELECTRIC = 14  /* rough GGX dielectric (reflect/refract, VNDF)      */
};

/* Procedural patterns, numbering of CudaScene::ProceduralPattern (cuda_scene.cuh:46-52). */
enum { RAYON_PATTERN_NONE = 0, RAYON_PATTERN_FIBONACCI_DOTS = 1 };

typedef struct rayon_material {
   int32_t type;
   float albedo[3];
   float emission[3];
   float roughness;
   float metallic;
   float refractive_index;
   float transmission;
   float anisotropy;
   float eta[3];          /* conductor complex IOR, real part                            */
   float k[3];            /* conductor complex IOR, imaginary part                       */
   float film_thickness;  /* nm                                                          */
   float film_ior;
   int32_t pattern;
   float pattern_color[3];
   float pattern_param1;
   float pattern_param2;
} rayon_material;

/* Geometry types, numbering of CudaScene::GeometryType (cuda_scene.cuh:95-104). */
enum {
   RAYON_GEOM_SPHERE = 0,
   RAYON_GEOM_RECTANGLE = 1,
   RAYON_GEOM_DISPLACED_SPHERE = 3,
   RAYON_GEOM_TRIANGLE = 4
};

typedef struct rayon_geometry {
   int32_t type;
   int32_t material_id;
   int32_t visible;      /* 0: camera rays pass through (cuda_raytracer.cuh:879-883)    */
   int32_t has_normals;  /* triangles: interpolate n0..n2                               */
   float p[9];           /* sphere: cx,cy,cz,r | rectangle: corner,u,v | triangle: v0,v1,v2 */
   float n[9];           /* triangle vertex normals n0,n1,n2                            */
} rayon_geometry;

/* Indexed triangle mesh; every triangle becomes one primitive. */
typedef struct rayon_mesh {
   const float* positions;   /* 3 floats per vertex                                  */
   const float* normals;     /* 3 floats per vertex, or NULL (flat shading)          */
   const int32_t* indices;   /* 3 per triangle                                       */
   int64_t num_vertices;
   int64_t num_triangles;
   int32_t material_id;
   int32_t visible;
} rayon_mesh;

/* Analytic SDF primitives, combined by polynomial smooth union into one field. */
enum { RAYON_SDF_SPHERE = 0, RAYON_SDF_ROUNDED_BOX = 1, RAYON_SDF_TORUS = 2 };

typedef struct rayon_sdf_prim {
   int32_t type;
   float center[3];
   float params[4];  /* sphere: r | rounded box: hx,hy,hz,round | torus: R,r (XZ plane) */
} rayon_sdf_prim;

typedef struct rayon_scene_desc {
   const rayon_material* materials;
   int32_t num_materials;
   const rayon_geometry* geometries;
   int64_t num_geometries;
   const rayon_mesh* meshes;
   int32_t num_meshes;
   const rayon_sdf_prim* sdf_prims;
   int32_t num_sdf_prims;
   float sdf_smooth_k;       /* smooth-union k                                          */
   int32_t sdf_material_id;
   int32_t sdf_max_steps;    /* sphere-tracing step cap                                 */
   float sdf_eps;            /* hit threshold                                           */
} rayon_scene_desc;

/* BVH build options. Default: binary tree stored as FP32 pair nodes (64 B per pair of
 * children), built on the host by binned SAH. RAYON_BVH_QUANTIZED also collapses it into
 * an 8-wide tree of 80 B nodes with 8-bit child boxes rounded outward, and traversal uses
 * that tree instead (2.4x fewer node bytes per ray; more box-test arithmetic per node).
 * RAYON_BVH_DEVICE builds the binary tree on the GPU instead (Morton-ordered linear BVH:
 * milliseconds for a million triangles, at some traversal cost against SAH); images are
 * identical to the host tree's. It falls back to the host build if the tree would be
 * deeper than the traversal stack; rayon_scene_info.builder says which one ran. With both
 * flags the device-built tree is read back, collapsed into the 8-wide layout by the same
 * surface-area rule, and its records are permuted into the wide order on the device. */
enum {
   RAYON_BVH_DEFAULT = 0,
   RAYON_BVH_QUANTIZED = 1 << 0,
   RAYON_BVH_DEVICE = 1 << 1
};
enum { RAYON_BVH_BUILDER_HOST = 0, RAYON_BVH_BUILDER_DEVICE = 1 };

typedef struct rayon_scene_info {
   int64_t num_prims;
   int64_t num_nodes;         /* binary pair nodes                                       */
   int64_t num_qnodes;        /* quantized wide nodes                                    */
   int32_t max_depth;         /* binary tree depth (root = 0)                            */
   int32_t num_lights;
   int64_t device_bytes;      /* device memory held by the scene                         */
   double build_ms;           /* BVH build time (host, or device including its sort)      */
   double upload_ms;          /* host-to-device copies of the scene                      */
   int32_t builder;           /* RAYON_BVH_BUILDER_HOST or RAYON_BVH_BUILDER_DEVICE        */
   int32_t wide_depth;        /* quantized wide tree depth (root = 0; 0 without one)     */
} rayon_scene_info;

/* ---------------------------------------------------------------------------------
 * Camera (reference CameraBase::initialize, camera/camera_base.hpp:76-99, FP64).
 * ------------------------------------------------------------------------------- */
typedef struct rayon_camera {
   double center[3];
   double pixel00[3];
   double delta_u[3];
   double delta_v[3];
   double u[3], v[3], w[3];
} rayon_camera;

rayon_status rayon_camera_frame(const double look_from[3], const double look_at[3], const double vup[3],
                                double vfov_deg, int32_t width, int32_t height, rayon_camera* out);

/* ---------------------------------------------------------------------------------
 * Jobs: one image (camera, resolution, sample budget) over one scene.
 * ------------------------------------------------------------------------------- */
enum {
   RAYON_FLAG_REF_QUIRKS = 1 << 0,    /* metal roughness 0 -> 0.3, fuzz x metal_fuzziness,
                                          glass uses glass_ior (cuda_raytracer.cuh:432-437,
                                          :647-660)                                          */
   RAYON_FLAG_REF_RNG = 1 << 1,       /* sequential per-pixel stream of cuda_utils.cuh:24-36
                                          seeded as cuda_utils.cu:26; default: counter-based */
   RAYON_FLAG_NEE = 1 << 2,           /* next-event estimation on Lambertian surfaces       */
   RAYON_FLAG_ADAPTIVE_REF = 1 << 3,  /* batch-to-batch luminance change test
                                          (render_acc_kernel.cu:240-271)                     */
   RAYON_FLAG_ADAPTIVE_STDERR = 1 << 4,/* per-pixel relative standard error test            */
   RAYON_FLAG_DOF = 1 << 5,           /* thin-lens depth of field (render_acc_kernel.cu:209-221) */
   RAYON_FLAG_COUNT_TRAVERSAL = 1 << 6 /* also count node/primitive/material fetches
                                          (rayon_job_stats.node_fetches...); slower      */
};

typedef struct rayon_job_params {
   int32_t width, height;
   float camera_center[3];
   float pixel00[3];
   float delta_u[3];
   float delta_v[3];
   float cam_u[3];
   float cam_v[3];
   int32_t max_depth;
   int32_t spp;                  /* sample budget per pixel                              */
   uint64_t seed;
   uint32_t flags;
   float light_intensity;        /* reference default 1.0                                */
   float background_intensity;   /* reference default 1.0                                */
   float metal_fuzziness;        /* used with RAYON_FLAG_REF_QUIRKS; reference 0.8        */
   float glass_ior;              /* used with RAYON_FLAG_REF_QUIRKS; reference 1.5        */
   float dof_aperture;
   float dof_focus_distance;
   int32_t min_adaptive_samples; /* adaptive tests start at this many samples            */
   float adaptive_threshold;     /* REF: relative luminance change; STDERR: rel. std-error */
} rayon_job_params;

/* Fills params with the reference's defaults (renderer_cuda_device.cu:18-26, :24 of the header). */
void rayon_job_params_default(rayon_job_params* p);
/* Copies an FP64 camera frame into the FP32 job fields (the cast of renderer_cuda_device.cu:279-284). */
void rayon_job_params_set_camera(rayon_job_params* p, const rayon_camera* cam);

typedef struct rayon_job_stats {
   int64_t samples_done;         /* uniform samples per pixel completed                  */
   int64_t primary_samples;      /* camera samples traced (adaptive pixels may skip)    */
   int64_t segments;             /* extension + shadow ray segments traced              */
   int64_t shadow_segments;
   int64_t converged_pixels;
   int32_t done;                 /* sample budget exhausted or every pixel converged    */
   int32_t pad_;
   int64_t node_fetches;         /* 64 B pair nodes read (RAYON_FLAG_COUNT_TRAVERSAL)    */
   int64_t prim_fetches;         /* 48 B primitive records read                         */
   int64_t material_fetches;     /* 96 B material records read                          */
   int64_t light_fetches;        /* emitter primitive records read for NEE              */
} rayon_job_stats;

typedef struct rayon_ctx rayon_ctx;
typedef struct rayon_scene rayon_scene;
typedef struct rayon_job rayon_job;

const char* rayon_last_error(void);
int32_t rayon_abi_version(void);

rayon_status rayon_ctx_create(int32_t device, rayon_ctx** out);
rayon_status rayon_ctx_destroy(rayon_ctx* ctx);
/* The HIP stream every kernel of this context is launched on (hipStream_t). */
void* rayon_ctx_stream(rayon_ctx* ctx);
rayon_status rayon_ctx_sync(rayon_ctx* ctx);

/* Builds the BVH (binned SAH, 16 bins, leaves <= 4 primitives) on the host and uploads
 * primitives, materials, nodes and lights. Replaces buildGPUScene (scene_builder_cuda.cu:148-303). */
rayon_status rayon_scene_create(rayon_ctx* ctx, const rayon_scene_desc* desc, uint32_t bvh_flags,
                                rayon_scene** out);
rayon_status rayon_scene_info_get(const rayon_scene* scene, rayon_scene_info* out);
/* Host only, no device needed: builds the BVH of a description and verifies its
 * invariants (every primitive in exactly one leaf, leaves of 1..4, every child box
 * containing what lies below it, depth within the traversal stack). Fills counts. */
rayon_status rayon_bvh_check(const rayon_scene_desc* desc, rayon_scene_info* out);
/* The same invariants for a created scene's binary tree (either builder), read back from
 * the device and checked against the description it was created from, and for its 8-wide
 * quantized tree when it has one (quantized boxes containing what lies below them, every
 * primitive in exactly one wide leaf, the reported wide depth); the depth of the binary
 * tree's deepest leaf through `depth`. */
rayon_status rayon_scene_check(const rayon_scene* scene, const rayon_scene_desc* desc, int32_t* depth);
/* Index (into rayon_scene_desc order: geometries first, then mesh triangles in order)
 * of the primitive stored at BVH slot `slot`; for host-side checks. */
rayon_status rayon_scene_prim_order(const rayon_scene* scene, int64_t* out_orig_ids, int64_t n);
rayon_status rayon_scene_destroy(rayon_scene* scene);

rayon_status rayon_job_create(rayon_ctx* ctx, rayon_scene* scene, const rayon_job_params* params,
                              rayon_job** out);
rayon_status rayon_job_destroy(rayon_job* job);
/* Clears accumulation, counts and RNG state (camera change; resetDeviceAccumBuffer). */
rayon_status rayon_job_reset(rayon_job* job);

/* Plain path: one launch of the per-pixel megakernel adds `samples` samples to every
 * unconverged pixel of the job (renderAccKernel, render_acc_kernel.cu:128-282).
 * Asynchronous on the context stream. */
rayon_status rayon_job_render_plain(rayon_ctx* ctx, rayon_job* job, int32_t samples);

/* Batched path: one progressive pass over `n` jobs, each advanced by `samples` samples
 * (clamped to its budget). Jobs of wavefront scenes go through generate / extend / shade /
 * shadow / accumulate stages over shared device queues, jobs of megakernel scenes share
 * one launch (rayon_scene_set_batch_path). Asynchronous on the context's stream. The jobs
 * may belong to any context of the same device (the caller orders their use across
 * streams); such passes are not captured in graphs. A job must not be in two passes at
 * once. */
rayon_status rayon_render_batched(rayon_ctx* ctx, rayon_job* const* jobs, int32_t n, int32_t samples);

/* Context options for the batched path. */
enum {
   RAYON_OPT_WAVE_CAPACITY = 0,   /* paths in flight per round (default 128M, 164 B each,
                                     allocated as passes need them)                         */
   RAYON_OPT_WAVE_WG_PER_CU = 1,  /* persistent workgroups per CU of the queue stages
                                     (0 = as many as each stage's occupancy allows)         */
   RAYON_OPT_GRAPH = 2,           /* capture each pass into a hipGraph and replay it (1)      */
   RAYON_OPT_STAGE_TIMING = 3,    /* time every stage launch with HIP events (0; disables graphs) */
   RAYON_OPT_COUNT_TRAVERSAL = 4, /* count node/primitive fetches of the queue stages (0)     */
   RAYON_OPT_WAVE_CHUNK = 5,      /* 64-ray batches a trace wave claims per queue atomic
                                     (0: 8 on binary trees, 32 on quantized ones)            */
   RAYON_OPT_SHADE_CHUNK = 6,     /* 64-hit batches a shade wave claims per queue atomic (16) */
   RAYON_OPT_TRACE_REFILL = 7,    /* idle lanes (of 64) at which a trace wave takes new rays
                                     from the queue; 64 traces whole batches (0: 32 on binary
                                     trees, 16 on quantized ones)                              */
   RAYON_OPT_TRACE_LEAF_EXIT = 8  /* a trace wave tests the leaves its lanes hold once at most
                                     min(this, lanes holding rays / 4) lanes (/ 3 on quantized
                                     trees) are still descending to their first leaf; 0 waits
                                     for every lane (-1: 12 on binary trees, 16 on quantized) */
};
rayon_status rayon_ctx_set_option(rayon_ctx* ctx, int32_t option, int64_t value);
rayon_status rayon_ctx_get_option(rayon_ctx* ctx, int32_t option, int64_t* value);

/* Accumulated statistics of the batched path since the last reset. Stages: 0 generate,
 * 1 trace (extension closest hits + shadow any hits), 2 shade, 3 megakernel (jobs of
 * megakernel scenes), 4 accumulate, 5 bookkeeping, 6 SDF marching, 7 (unused). */
typedef struct rayon_wave_stats {
   double stage_ms[8];            /* summed HIP-event time per stage (RAYON_OPT_STAGE_TIMING) */
   int64_t launches[8];
   int64_t rounds;
   int64_t extend_node_fetches;   /* RAYON_OPT_COUNT_TRAVERSAL                               */
   int64_t extend_prim_fetches;
   int64_t shadow_node_fetches;
   int64_t shadow_prim_fetches;
   int64_t material_fetches;
   int64_t light_fetches;
   int64_t sdf_marches;           /* rays that started an SDF march (RAYON_OPT_COUNT_TRAVERSAL) */
   int64_t sdf_steps;             /* field evaluations (sphere-tracing steps)                */
   int64_t sdf_prim_evals;        /* primitive distance evaluations within those steps       */
   int64_t sdf_wave_slots;        /* lane-steps the waves spent (64 x the wave's longest
                                     lane per batch): sdf_steps / sdf_wave_slots = lane use */
} rayon_wave_stats;
rayon_status rayon_ctx_wave_stats(rayon_ctx* ctx, rayon_wave_stats* out, int32_t reset);

/* Copies the per-pixel mean linear radiance (3 floats per pixel) to host memory. */
rayon_status rayon_job_resolve_linear(rayon_ctx* ctx, rayon_job* job, float* rgb_out);
/* Gamma-corrected 8-bit image (gammaCorrectKernel, render_acc_kernel.cu:11-44). */
rayon_status rayon_job_resolve_u8(rayon_ctx* ctx, rayon_job* job, uint8_t* out, int32_t channels, float gamma);
/* Allocation-free variant of rayon_job_resolve_u8, with the same result. It has no
 * counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * image is staged in a fixed 64 MiB device array of the library (one per GPU, held until
 * the process ends) instead of a buffer allocated per call, so width * height * channels
 * must not exceed 64 MiB (RAYON_ERR_INVALID otherwise). Every call on one GPU uses the same
 * array: calls that overlap on one GPU race, so make them from one thread per device. */
rayon_status rayon_job_resolve_u8_noalloc(rayon_ctx* ctx, rayon_job* job, uint8_t* out, int32_t channels,
                                          float gamma);
/* Per-pixel sample counts (negative = converged, as the reference's count buffer). */
rayon_status rayon_job_sample_counts(rayon_ctx* ctx, rayon_job* job, int32_t* out);
rayon_status rayon_job_stats_get(rayon_ctx* ctx, rayon_job* job, rayon_job_stats* out);

/* Closest-hit query for explicit rays (origins/dirs: 3 floats each, host memory).
 * Writes the original primitive index (-1 on miss) and hit distance. For parity tests
 * of traversal and intersection; runs the same device code as the renderers. */
rayon_status rayon_trace_closest(rayon_ctx* ctx, rayon_scene* scene, const float* origins, const float* dirs,
                                 int64_t n, float t_min, float t_max, int64_t* out_prim, float* out_t);

/* Material probe: scatters n rays with the renderers' own device code for material m
 * (scatter_material, cuda_raytracer.cuh:613-850, after apply_material's mapping :418-485),
 * each at a surface point with normal normals[i] (facing the ray), tangent tangents[i] and
 * front-face flag front[i], drawing its random numbers in order from rnd[i*nrnd ...]
 * instead of the path's stream. flags / fuzziness / glass IOR / light intensity act as
 * the job's. All arrays are host memory, 3 floats per vector. For parity tests against
 * a CPU restatement; out_used receives how many numbers each sample drew. */
rayon_status rayon_scatter_probe(rayon_ctx* ctx, const rayon_material* m, uint32_t flags, float metal_fuzziness,
                                 float glass_ior, float light_intensity, const float* dirs, const float* normals,
                                 const float* tangents, const int32_t* front, const float* rnd, int32_t nrnd,
                                 int64_t n, float* out_dir, float* out_atten, float* out_emitted,
                                 int32_t* out_scattered, int32_t* out_used);

/* Accumulation probe: folds n sample values (RGB, host memory) in order into one pixel's
 * compensated FP32 sums with the renderers' own device code (the acc += ray_color of
 * render_acc_kernel.cu:224, compensated), returning the sums (rgb, luminance^2) and
 * their compensation terms. For tests of the summation against exact sums. */
rayon_status rayon_accumulate_probe(rayon_ctx* ctx, const float* rgb, int64_t n, float out_sum[4],
                                    float out_comp[4]);

/* Primary camera rays of sample `sample` of a job, exactly as the renderers generate
 * them (origins/dirs 3 floats per pixel). For parity tests. */
rayon_status rayon_job_primary_rays(rayon_ctx* ctx, rayon_job* job, int32_t sample, float* origins,
                                    float* dirs);

/* Path a scene's jobs take in batched passes. WAVEFRONT: the staged queue pipeline
 * (generate / trace / shade / accumulate); MEGAKERNEL: every path of a pixel in one lane,
 * all such jobs of a pass in one launch (the plain kernel's arithmetic). AUTO: megakernel
 * for scenes of at most 8 MiB on the device, wavefront otherwise; rayon_autotune measures
 * both on AUTO scenes and records the faster. A job keeps the path of its first pass
 * until it is reset. Both compute the same estimator; their FP32 rounding differs. */
enum { RAYON_PATH_AUTO = 0, RAYON_PATH_WAVEFRONT = 1, RAYON_PATH_MEGAKERNEL = 2 };
rayon_status rayon_scene_set_batch_path(rayon_scene* scene, int32_t path);
/* The path the scene's new jobs take (AUTO resolved). */
int32_t rayon_scene_batch_path(const rayon_scene* scene);

/* ---------------------------------------------------------------------------------
 * Autotuner of batched passes (no counterpart in the reference, whose launch shape is
 * fixed at 32x8 threads per block, renderer_cuda_device.cu:95-96). On clones of the given
 * jobs (same scenes and parameters, fresh buffers, so the jobs themselves are untouched)
 * it (1) times each RAYON_PATH_AUTO scene's jobs on both batched paths and records the
 * faster on the scene, then (2) times wavefront launch shapes (path-slot capacity, queue
 * chunk, shade chunk, workgroups per CU) by coordinate descent, sets the fastest on the context and
 * caches it under a signature of the workload (job count, total work, depth, scenes,
 * flags). The shape changes only the speed of a pass, never its results.
 * ------------------------------------------------------------------------------- */
typedef struct rayon_tune_result {
   int64_t capacity;     /* RAYON_OPT_WAVE_CAPACITY chosen                                */
   int32_t chunk;        /* RAYON_OPT_WAVE_CHUNK chosen                                   */
   int32_t wg_per_cu;    /* RAYON_OPT_WAVE_WG_PER_CU chosen (0 = occupancy)               */
   double best_ms;       /* one pass of `samples` with the chosen shape                  */
   double start_ms;      /* the same pass with the shape the context had before          */
   int32_t candidates;   /* passes timed (0 when taken from the cache)                    */
   int32_t cached;
   int32_t mega_scenes;  /* scenes of the jobs on the megakernel path after tuning        */
   int32_t wave_scenes;  /* scenes of the jobs on the wavefront path after tuning         */
   int32_t shade_chunk;  /* RAYON_OPT_SHADE_CHUNK chosen                                  */
   int32_t refill;       /* RAYON_OPT_TRACE_REFILL chosen                                 */
} rayon_tune_result;

rayon_status rayon_autotune(rayon_ctx* ctx, rayon_job* const* jobs, int32_t n, int32_t samples,
                            rayon_tune_result* out);

/* ---------------------------------------------------------------------------------
 * Headless progressive renderer of one job: the render loop of the reference's
 * interactive mode (RendererCUDAProgressive, renderer_cuda_progressive_host.hpp:440-920)
 * without a window. Each step renders one batch sized by its frame-time controller
 * (:681-686), optionally at its progressive depth schedule (:853-876), and resolves the
 * display image on a second stream into pinned host memory while the caller goes on.
 * ------------------------------------------------------------------------------- */
typedef struct rayon_progressive rayon_progressive;

typedef struct rayon_progressive_config {
   double target_fps;       /* frame-time target of the batch controller (60)               */
   int32_t max_batch;       /* cap of samples per batch (the reference's slider, 64)        */
   int32_t first_batch;     /* samples of the first batch (4)                               */
   int32_t adaptive_depth;  /* 1: shallow paths first (depth 4 .. the job's own)           */
   int32_t channels;        /* display channels, 3 or 4 (4)                                  */
   float gamma;             /* display gamma (2)                                             */
   int32_t pad_;
} rayon_progressive_config;

typedef struct rayon_progressive_frame {
   int64_t samples;         /* samples per pixel after this step                           */
   int64_t segments;        /* ray segments traced so far                                  */
   double frame_ms;         /* wall time of the batch                                      */
   double convergence;      /* percent of pixels converged (adaptive jobs), from device counters */
   int32_t batch;           /* samples rendered by this step                               */
   int32_t next_batch;      /* the controller's next batch                                 */
   int32_t depth;           /* max depth used by this step                                 */
   int32_t done;            /* sample budget spent or every pixel converged                */
} rayon_progressive_frame;

void rayon_progressive_config_default(rayon_progressive_config* cfg);
rayon_status rayon_progressive_create(rayon_ctx* ctx, rayon_scene* scene, const rayon_job_params* params,
                                      const rayon_progressive_config* cfg, rayon_progressive** out);
rayon_status rayon_progressive_step(rayon_progressive* p, rayon_progressive_frame* frame);
/* Latest display image (width*height*channels bytes) and its samples per pixel. */
rayon_status rayon_progressive_display(rayon_progressive* p, uint8_t* out, int64_t* samples);
/* Camera change: new camera (or NULL to keep it), sums and counts cleared, first batch again. */
rayon_status rayon_progressive_reset(rayon_progressive* p, const rayon_camera* camera);
/* The job being rendered (for resolves and statistics); owned by the renderer. */
rayon_job* rayon_progressive_job(rayon_progressive* p);
rayon_status rayon_progressive_destroy(rayon_progressive* p);

/* ---------------------------------------------------------------------------------
 * Render server: continuous batching of many jobs on one context (no counterpart in the
 * reference). A server thread admits queued jobs by HBM budget (seven eighths of it; the
 * rest holds the first frames of jobs still queued), renders each job's first frame as
 * soon as it is admitted or its frame fits that reserve (one kernel on a first-frame
 * stream, overlapping the pass in flight, resolved on the device), advances the oldest
 * active jobs together in batched passes whose work (pixels x samples) a controller sizes
 * toward target_pass_ms, and retires jobs whose sample budget is spent or whose pixels all
 * converged. The first-frame streams belong to the context (created by its first server,
 * destroyed with the context).
 * ------------------------------------------------------------------------------- */
typedef struct rayon_server rayon_server;

enum {
   RAYON_JOB_QUEUED = 0,
   RAYON_JOB_RUNNING = 1,
   RAYON_JOB_DONE = 2,
   RAYON_JOB_FAILED = 3,
   RAYON_JOB_CANCELLED = 4
};

typedef struct rayon_server_config {
   int64_t hbm_budget_bytes;    /* accumulation + frame buffers + referenced scenes      */
   int32_t max_active_jobs;
   int32_t first_pass_samples;  /* samples of a job's first frame (the plain kernel's
                                   arithmetic for its first samples, outside its sums)  */
   int32_t max_pass_samples;    /* cap of the controlled pass size                       */
   int32_t pad_;
   double target_pass_ms;       /* controller target for one batched pass               */
   int64_t max_pass_pixels;     /* cap of the pixels a refinement pass advances (the
                                   oldest active jobs; at least one job)                */
} rayon_server_config;

typedef struct rayon_server_job_info {
   int32_t state;
   int32_t first_kept;          /* 1: its first frame's samples are its first samples      */
   int64_t samples_done;
   int64_t segments;
   double queue_ms;             /* arrival -> admission                                  */
   double ttff_ms;              /* arrival -> first frame resolved on the device (-1)    */
   double ttc_ms;               /* arrival -> finished (-1)                              */
} rayon_server_job_info;

typedef struct rayon_server_stats {
   int64_t passes, first_passes, completed, failed, queued, active, segments, used_bytes;
   double busy_s, wall_s;
   int32_t pass_samples;        /* samples of the last refinement pass                    */
   int32_t pass_jobs;           /* jobs of the last refinement pass                       */
   double pass_ms;              /* summed wall time of refinement passes (start to done)  */
   double first_ms;             /* summed wall time of first passes                       */
   double check_ms;             /* summed time reading job statistics after passes        */
   double pass_paths;           /* summed pixel-samples of refinement passes              */
   int64_t first_kept;          /* first frames kept as their jobs' first samples          */
   int64_t first_kept_paths;    /* pixel-samples of those first frames                     */
} rayon_server_stats;

void rayon_server_config_default(rayon_server_config* cfg);
rayon_status rayon_server_create(rayon_ctx* ctx, const rayon_server_config* cfg, rayon_server** out);
/* Thread-safe. Fails with RAYON_ERR_BUDGET if the job alone exceeds the budget. */
rayon_status rayon_server_submit(rayon_server* s, rayon_scene* scene, const rayon_job_params* params,
                                 uint64_t* id);
rayon_status rayon_server_job_info_get(rayon_server* s, uint64_t id, rayon_server_job_info* out);
/* Mean linear radiance (3 floats per pixel) of a finished, unreleased job. */
rayon_status rayon_server_result(rayon_server* s, uint64_t id, float* rgb);
/* The first frame of a job (gamma-2 RGBA8, width*height*4 bytes) once ttff_ms >= 0. */
rayon_status rayon_server_first_frame(rayon_server* s, uint64_t id, uint8_t* rgba);
rayon_status rayon_server_cancel(rayon_server* s, uint64_t id);
/* Frees a finished job's device buffers (deferred until it finishes otherwise). */
rayon_status rayon_server_release(rayon_server* s, uint64_t id);
/* Waits until no job is queued or active. */
rayon_status rayon_server_drain(rayon_server* s, double timeout_s);
rayon_status rayon_server_stats_get(rayon_server* s, rayon_server_stats* out);
rayon_status rayon_server_destroy(rayon_server* s);

/* ---------------------------------------------------------------------------------
 * Synthetic scene generators (host) and the YAML scene loader.
 * ------------------------------------------------------------------------------- */
typedef struct rayon_generated rayon_generated;

enum {
   RAYON_GEN_BOOK_SPHERES = 0,   /* ground + 3 large + jittered r=0.2 grid             */
   RAYON_GEN_CORNELL_ICOSPHERE = 1,
   RAYON_GEN_SDF_TERRAIN = 2
};

typedef struct rayon_gen_params {
   int32_t kind;
   uint64_t seed;
   int32_t detail;     /* book: unused | cornell: icosphere subdivision level | sdf: terrain grid */
   int32_t count;      /* sdf: number of SDF primitives                                  */
} rayon_gen_params;

rayon_status rayon_generate(const rayon_gen_params* params, rayon_generated** out);
/* Scene file loader (Scene::loadSceneFromYAML, yaml_scene_loader.cc:505-576, with its OBJ
 * loader obj_loader.hpp:34-171): the same keys, defaults, limits and fall-backs, and the same
 * geometry list (one triangle per OBJ face triangle, in file order). Errors: RAYON_ERR_IO
 * (file), RAYON_ERR_PARSE (a number that does not parse, no materials, no geometry). */
rayon_status rayon_load_yaml(const char* path, rayon_generated** out);
enum { RAYON_YAML_LIFT_CAPS = 1 << 0 };  /* read past 100 materials / 1000 geometry entries */
rayon_status rayon_load_yaml_ex(const char* path, uint32_t flags, rayon_generated** out);
/* Writes a scene file that rayon_load_yaml reads back to the same description, camera and
 * settings (the role of saveSceneToYAML, yaml_scene_loader.cc:578-744, without its loss:
 * every number is written with 9 significant digits and triangles are kept). */
rayon_status rayon_save_yaml(const rayon_generated* g, const char* path);
/* Settings of a loaded scene file (defaults for generated scenes). */
rayon_status rayon_generated_settings(const rayon_generated* g, float background_color[3],
                                      float* background_intensity, float* ambient_light, int32_t* use_bvh,
                                      int32_t* adaptive_sampling);
/* ---- Output files of a finished render -------------------------------------------------
 * 8-bit image (channels 1 gray, 2 gray+alpha, 3 RGB, 4 RGBA; rows top to bottom, tightly
 * packed) written as a PNG file, creating missing parent directories. Replaces
 * FileUtils::writeImage (utils.hpp:34-48, stb with RGB) and ensureDirectoryExists
 * (utils.hpp:22-31). Errors: RAYON_ERR_INVALID (sizes), RAYON_ERR_IO (file). */
rayon_status rayon_write_png(const char* path, const uint8_t* pixels, int32_t width, int32_t height,
                             int32_t channels);
/* Resolves a job's image (gamma-corrected RGB, as rayon_job_resolve_u8) and writes it as a
 * PNG: the reference's output of a render (main.cc:468-471). */
rayon_status rayon_job_write_png(rayon_ctx* ctx, rayon_job* job, const char* path, float gamma);

/* Statistics of one render, the fields of FileUtils::writeRenderStats (utils.hpp:118-172). */
typedef struct rayon_render_stats {
   int32_t samples_per_pixel;
   int32_t width;
   int32_t height;
   int32_t max_depth;
   int64_t rays_traced;
   int64_t image_size_bytes;  /* -1: the size of the file at image_path (0 if it is absent) */
   int64_t render_time_ns;
} rayon_render_stats;
/* Writes the statistics next to the image: image_path with its extension replaced by .txt
 * and by .json, with the keys, order and number formats of writeRenderStats
 * (utils.hpp:118-172): rays_per_second rounded to the nearest integer (0 for a zero
 * duration), render_time_ms truncated, render_time_pretty as rayon_format_duration. */
rayon_status rayon_write_render_stats(const char* image_path, const rayon_render_stats* stats);
/* Duration text of formatDuration (utils.hpp:90-116): "<m>m <s>s" from one minute, "<s>s"
 * from 10 s, "<s.ss>s" from 1 s, "<ms>ms" below. Writes at most buf_len bytes with the NUL. */
rayon_status rayon_format_duration(int64_t duration_ns, char* buf, int32_t buf_len);

/* The description stays valid until rayon_generated_free. */
const rayon_scene_desc* rayon_generated_desc(const rayon_generated* g);
/* Camera from the generator or the YAML camera section: look_from, look_at, vup, vfov. */
rayon_status rayon_generated_camera(const rayon_generated* g, double look_from[3], double look_at[3],
                                    double vup[3], double* vfov);
float rayon_generated_background_intensity(const rayon_generated* g);
void rayon_generated_free(rayon_generated* g);

#ifdef __cplusplus
}
#endif

#endif /* RAYON_HIP_H */
