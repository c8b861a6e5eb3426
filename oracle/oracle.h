/*
 * oracle.h — FP64 CPU restatement of the reference's GPU path-tracing semantics.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for librayon_hip.so: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark load it. The
 * product never links, calls or falls back to it.
 *
 * Scene input uses the same plain structs as the product's C ABI (include/rayon_hip.h):
 * the oracle reads the description itself and builds its own, independent BVH.
 */
#ifndef RAYON_ORACLE_H
#define RAYON_ORACLE_H

#include "../include/rayon_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

oracle_scene* oracle_scene_create(const rayon_scene_desc* desc);
void oracle_scene_free(oracle_scene* s);
int64_t oracle_scene_num_prims(const oracle_scene* s);
int32_t oracle_scene_num_lights(const oracle_scene* s);

/* Closest hit for explicit rays, in FP64 on the FP32 inputs. brute != 0 tests every
 * primitive in index order; otherwise the oracle's own BVH is used. Exact t ties go to
 * the larger primitive index (the linear-scan rule of cuda_raytracer.cuh:568-585 with
 * the inclusive t bound of :215-220). */
void oracle_trace_closest(const oracle_scene* s, const float* origins, const float* dirs, int64_t n,
                          double t_min, double t_max, int32_t brute, int64_t* out_prim, double* out_t);

/* FP64 hit distance of one primitive (input order index) along a ray, -1 if missed. */
double oracle_intersect_prim(const oracle_scene* s, int64_t prim, const float* origin, const float* dir,
                             double t_min, double t_max);

/* 1 if any primitive other than an emitter that NEE samples (LIGHT-material rectangle,
 * triangle or sphere) is hit in [t_min, t_max]. */
void oracle_trace_any(const oracle_scene* s, const float* origins, const float* dirs, int64_t n, double t_min,
                      double t_max, int32_t* out_hit);

/* Adds samples [sample0, sample0+nsamples) to each listed pixel (pixel index y*W+x).
 * cam may be NULL (then the FP32 camera of params is used). sums: 3 doubles per listed
 * pixel, accumulated in place. With RAYON_FLAG_REF_RNG, rng_state holds one uint32 per
 * listed pixel, initialised by the caller with oracle_ref_rng_init and carried across
 * calls; otherwise it may be NULL. segments (optional) receives the ray-segment count.
 * threads <= 0 uses every hardware thread. */
void oracle_render(const oracle_scene* s, const rayon_job_params* params, const rayon_camera* cam,
                   const int64_t* pixels, int64_t npix, int32_t sample0, int32_t nsamples, int32_t threads,
                   double* sums, uint32_t* rng_state, int64_t* segments);

/* Per-sample radiance instead of sums: out has 3 doubles per (pixel, sample). */
void oracle_render_samples(const oracle_scene* s, const rayon_job_params* params, const rayon_camera* cam,
                           const int64_t* pixels, int64_t npix, int32_t sample0, int32_t nsamples,
                           int32_t threads, double* out);

/* RNG restatements. */
uint32_t oracle_ref_rng_init(int64_t pixel_index);       /* cuda_utils.cu:26            */
float oracle_ref_rng_next(uint32_t* state);               /* cuda_utils.cuh:24-36        */
uint32_t oracle_counter_rng_init(uint64_t seed, uint32_t pixel, uint32_t sample);
float oracle_counter_rng_next(uint32_t* state);

/* Camera frame, camera_base.hpp:76-99. */
void oracle_camera_frame(const double look_from[3], const double look_at[3], const double vup[3], double vfov,
                         int32_t width, int32_t height, rayon_camera* out);

/* Adaptive tests. REF: render_acc_kernel.cu:240-271 — returns the new signed count.
 * STDERR: 1 when sqrt(var/n)/max(mean,1e-3) < threshold over the luminance samples. */
int32_t oracle_adaptive_ref(int32_t count_before, const double acc_before[3], const double acc_after[3],
                            int32_t samples_added, int32_t min_samples, double threshold);
int32_t oracle_adaptive_stderr(double lum_sum, double lum_sq_sum, int64_t n, int32_t min_samples,
                               double threshold);

/* Gamma resolve of one channel value, render_acc_kernel.cu:30-40. */
uint8_t oracle_gamma_u8(double sum, int64_t n, double gamma);

/* Material sampling probe for unit tests: scatters a ray arriving along dir at a surface
 * with normal n (facing the ray), front_face, tangent t, using the given random numbers
 * in order. Returns 1 if scattered; writes outgoing dir, attenuation, emitted and the
 * number of random numbers consumed. */
int32_t oracle_scatter_probe(const rayon_material* m, uint32_t flags, float metal_fuzziness, float glass_ior,
                             float light_intensity, const double dir[3], const double n[3], int32_t front_face,
                             const double tangent[3], const float* rnd, int32_t nrnd, double out_dir[3],
                             double atten[3], double emitted[3], int32_t* used);

/* Pinning hooks: the pieces of the material models on their own, so the tests can hold
 * them against closed forms, analytic densities and numerical integrals. */
void oracle_set_fault(int32_t fault); /* plant a known mistake (tests of the tests); 0 = none */
/* VNDF sample (microfacet_ggx.cuh:45-76) for each (u1, u2) pair of u; out: n unit normals. */
void oracle_sample_vndf(const double wo[3], double ax, double ay, const double* u, int64_t n, double* out);
double oracle_ggx_lambda(const double w[3], double ax, double ay); /* microfacet_ggx.cuh:21-28 */
double oracle_fresnel_dielectric(double cos_i, double eta);        /* eta = n_incident / n_transmitted */
double oracle_fresnel_conductor(double cos_i, double eta, double k); /* FrComplex1, :83-101 */
/* random_in_unit_disk (cuda_raytracer.cuh:114-122) n times from the given uniforms; returns
 * the number consumed, or -1 if they ran out. */
int64_t oracle_unit_disk(const float* rnd, int64_t nrnd, int64_t n, double* out_xy);
/* oracle_scatter_probe for `count` independent samples, nrnd uniforms each. */
void oracle_scatter_batch(const rayon_material* m, uint32_t flags, float metal_fuzziness, float glass_ior,
                          float light_intensity, const double dir[3], const double n[3], int32_t front_face,
                          const double tangent[3], const float* rnd, int32_t nrnd, int64_t count,
                          double* out_dir, double* atten, int32_t* scattered);

#ifdef __cplusplus
}
#endif

#endif
