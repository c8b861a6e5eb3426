// legacy.cpp — the reference's extern "C" GPU-render entry points
// (renderer_cuda_device.cuh:8-68), implemented on the reentrant API of rayon_hip.h.
//
// The caller-owned `void*` handles keep the reference's lazy-allocation protocol
// (renderer_cuda_device.cu:198-234): *d_rand_states_ptr receives a library object that
// owns one job (per-pixel reference random streams, compensated sums, counts), and
// *d_accum_buffer_ptr receives that job's device float4 sum buffer, so
// convertAccumToDisplayCUDA / resetDeviceAccumBuffer can be handed the same pointer as
// in the reference. The global setters of renderer_cuda_device.cu:18-79 become fields
// applied to the job at every render call.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/rayon_legacy.h"
#include "internal.h"
#include "kernels.h"

namespace {

struct LegacyState {
   rayon_job* job = nullptr;
};

struct Knobs {
   float light = 1.0f, background = 1.0f, fuzz = 0.8f, glass = 1.5f;  // renderer_cuda_device.cu:18-21
   bool dof = false;                                                  // :24-26
   float aperture = 0.1f, focus = 10.0f;
};

std::mutex g_mu;
rayon_ctx* g_ctx = nullptr;
Knobs g_knobs;
std::unordered_map<void*, LegacyState*> g_by_acc;

void report(const char* where) { fprintf(stderr, "rayon (%s): %s\n", where, rayon_last_error()); }

rayon_ctx* ctx_locked() {
   if (!g_ctx && rayon_ctx_create(0, &g_ctx) != RAYON_OK) {
      report("context");
      g_ctx = nullptr;
   }
   return g_ctx;
}

LegacyState* state_of_acc(void* acc) {
   auto it = g_by_acc.find(acc);
   return it == g_by_acc.end() ? nullptr : it->second;
}

}  // namespace

extern "C" {

void* rayon_legacy_scene_create(const rayon_scene_desc* desc) {
   std::lock_guard<std::mutex> lk(g_mu);
   rayon_ctx* c = ctx_locked();
   if (!c) return nullptr;
   rayon_scene* s = nullptr;
   if (rayon_scene_create(c, desc, RAYON_BVH_DEFAULT, &s) != RAYON_OK) {
      report("scene upload");
      return nullptr;
   }
   return s;
}

void rayon_legacy_scene_destroy(void* scene) {
   std::lock_guard<std::mutex> lk(g_mu);
   rayon_scene_destroy((rayon_scene*)scene);
}

void initCudaStreams(void) {
   std::lock_guard<std::mutex> lk(g_mu);
   ctx_locked();
}

void cleanupCudaStreams(void) {
   // The library-owned context lives until process exit; nothing is left pending.
   std::lock_guard<std::mutex> lk(g_mu);
   if (g_ctx) (void)rayon_ctx_sync(g_ctx);
}

unsigned long long renderPixelsCUDAAccumulative(
    unsigned char* image, float* accum_buffer, void* scene, int width, int height, double ccx, double ccy,
    double ccz, double p00x, double p00y, double p00z, double dux, double duy, double duz, double dvx, double dvy,
    double dvz, int samples_to_add, int total_samples_so_far, int max_depth, void** d_rand_states_ptr,
    void** d_accum_buffer_ptr, double cux, double cuy, double cuz, double cvx, double cvy, double cvz,
    void* d_pixel_sample_counts, int min_adaptive_samples, float adaptive_threshold) {
   (void)image;  // ignored, as in the reference
   (void)total_samples_so_far;
   if (!scene) return 0ULL;  // renderer_cuda_device.cu:179-180
   if (!d_rand_states_ptr || !d_accum_buffer_ptr || width <= 0 || height <= 0 || samples_to_add < 0) {
      fprintf(stderr, "rayon (renderPixelsCUDAAccumulative): bad arguments\n");
      return 0ULL;
   }
   std::lock_guard<std::mutex> lk(g_mu);
   rayon_ctx* c = ctx_locked();
   if (!c) return 0ULL;
   const size_t npix = (size_t)width * height;
   auto* st = (LegacyState*)*d_rand_states_ptr;
   if (!st) {
      rayon_job_params p;
      rayon_job_params_default(&p);
      p.width = width;
      p.height = height;
      p.spp = 1 << 30;
      p.max_depth = max_depth > 0 ? max_depth : 1;
      p.flags = RAYON_FLAG_REF_RNG | RAYON_FLAG_REF_QUIRKS;
      st = new LegacyState();
      if (rayon_job_create(c, (rayon_scene*)scene, &p, &st->job) != RAYON_OK) {
         report("renderPixelsCUDAAccumulative");
         delete st;
         return 0ULL;
      }
      *d_rand_states_ptr = st;
   }
   rayon_job* j = st->job;
   if (j->params.width != width || j->params.height != height) {
      fprintf(stderr, "rayon (renderPixelsCUDAAccumulative): state was created for %dx%d, called with %dx%d\n",
              j->params.width, j->params.height, width, height);
      return 0ULL;
   }
   const hipStream_t stream = c->stream;
   if (*d_accum_buffer_ptr == nullptr) {
      *d_accum_buffer_ptr = j->view.acc;
      g_by_acc[j->view.acc] = st;
      if (accum_buffer) {  // batch path: upload the host float3 sums once (:213-223)
         std::vector<float4> h(npix);
         for (size_t i = 0; i < npix; ++i)
            h[i] = float4{accum_buffer[3 * i], accum_buffer[3 * i + 1], accum_buffer[3 * i + 2], 0.0f};
         if (hipMemcpy(j->view.acc, h.data(), npix * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess) {
            fprintf(stderr, "rayon (renderPixelsCUDAAccumulative): accumulation upload failed\n");
            return 0ULL;
         }
      }
   }
   rayon_job_params& P = j->params;
   const double cc[3] = {ccx, ccy, ccz}, p0[3] = {p00x, p00y, p00z}, du[3] = {dux, duy, duz},
                dv[3] = {dvx, dvy, dvz}, cu[3] = {cux, cuy, cuz}, cv[3] = {cvx, cvy, cvz};
   for (int k = 0; k < 3; ++k) {  // the FP64 -> FP32 cast at launch, renderer_cuda_device.cu:279-284
      P.camera_center[k] = (float)cc[k];
      P.pixel00[k] = (float)p0[k];
      P.delta_u[k] = (float)du[k];
      P.delta_v[k] = (float)dv[k];
      P.cam_u[k] = (float)cu[k];
      P.cam_v[k] = (float)cv[k];
   }
   P.max_depth = max_depth > 0 ? max_depth : 1;
   P.light_intensity = g_knobs.light;
   P.background_intensity = g_knobs.background;
   P.metal_fuzziness = g_knobs.fuzz;
   P.glass_ior = g_knobs.glass;
   P.dof_aperture = g_knobs.aperture;
   P.dof_focus_distance = g_knobs.focus;
   P.flags &= ~(uint32_t)(RAYON_FLAG_DOF | RAYON_FLAG_ADAPTIVE_REF);
   if (g_knobs.dof) P.flags |= RAYON_FLAG_DOF;
   if (d_pixel_sample_counts) {
      P.flags |= RAYON_FLAG_ADAPTIVE_REF;
      P.min_adaptive_samples = min_adaptive_samples;
      P.adaptive_threshold = adaptive_threshold;
      if (hipMemcpyAsync(j->view.counts, d_pixel_sample_counts, npix * sizeof(int), hipMemcpyDeviceToDevice,
                         stream) != hipSuccess) {
         fprintf(stderr, "rayon (renderPixelsCUDAAccumulative): count buffer copy failed\n");
         return 0ULL;
      }
   }
   j->samples_done = 0;  // the budget is the caller's: every call adds samples_to_add
   if (rayon_job_render_plain(c, j, samples_to_add) != RAYON_OK) {
      report("renderPixelsCUDAAccumulative");
      return 0ULL;
   }
   if (d_pixel_sample_counts &&
       hipMemcpyAsync(d_pixel_sample_counts, j->view.counts, npix * sizeof(int), hipMemcpyDeviceToDevice, stream) !=
           hipSuccess) {
      fprintf(stderr, "rayon (renderPixelsCUDAAccumulative): count buffer copy failed\n");
      return 0ULL;
   }
   if (hipStreamSynchronize(stream) != hipSuccess) {
      fprintf(stderr, "rayon (renderPixelsCUDAAccumulative): kernel execution failed\n");
      return 0ULL;
   }
   if (accum_buffer) {  // download every call, float4 -> host float3 (:300-311)
      std::vector<float4> a(npix), cm(npix);
      if (hipMemcpy(a.data(), j->view.acc, npix * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(cm.data(), j->view.comp, npix * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess) {
         fprintf(stderr, "rayon (renderPixelsCUDAAccumulative): accumulation download failed\n");
         return 0ULL;
      }
      for (size_t i = 0; i < npix; ++i) {
         accum_buffer[3 * i] = a[i].x + cm[i].x;
         accum_buffer[3 * i + 1] = a[i].y + cm[i].y;
         accum_buffer[3 * i + 2] = a[i].z + cm[i].z;
      }
   }
   return (unsigned long long)npix * (unsigned long long)samples_to_add;  // primary samples (:313-322)
}

void freeDeviceRandomStates(void* d_rand_states) {
   if (!d_rand_states) return;
   std::lock_guard<std::mutex> lk(g_mu);
   auto* st = (LegacyState*)d_rand_states;
   for (auto it = g_by_acc.begin(); it != g_by_acc.end();) {
      if (it->second == st)
         it = g_by_acc.erase(it);
      else
         ++it;
   }
   rayon_job_destroy(st->job);
   delete st;
}

void freeDeviceAccumBuffer(void* d_accum_buffer) {
   // The sum buffer belongs to the state object and is released with it.
   (void)d_accum_buffer;
}

void resetDeviceAccumBuffer(void* d_accum_buffer, int num_pixels) {
   if (!d_accum_buffer || num_pixels <= 0) return;
   std::lock_guard<std::mutex> lk(g_mu);
   rayon_ctx* c = ctx_locked();
   if (!c) return;
   LegacyState* st = state_of_acc(d_accum_buffer);
   if (st) {  // zero sums, compensation and counts; the random streams continue (as in the reference)
      rayon_job* j = st->job;
      const long long n = (long long)j->params.width * j->params.height;
      if (ry::launch_fill_u32((uint32_t*)j->view.acc, 0u, 4 * n, c->stream) != hipSuccess ||
          ry::launch_fill_u32((uint32_t*)j->view.comp, 0u, 4 * n, c->stream) != hipSuccess ||
          ry::launch_fill_u32((uint32_t*)j->view.counts, 0u, n, c->stream) != hipSuccess)
         fprintf(stderr, "rayon (resetDeviceAccumBuffer): launch failed\n");
   } else if (ry::launch_fill_u32((uint32_t*)d_accum_buffer, 0u, 4ll * num_pixels, c->stream) != hipSuccess) {
      fprintf(stderr, "rayon (resetDeviceAccumBuffer): launch failed\n");
   }
   (void)hipStreamSynchronize(c->stream);
}

void setLightIntensity(float v) {
   std::lock_guard<std::mutex> lk(g_mu);
   g_knobs.light = v;
}
void setBackgroundIntensity(float v) {
   std::lock_guard<std::mutex> lk(g_mu);
   g_knobs.background = v;
}
void setMetalFuzziness(float v) {
   std::lock_guard<std::mutex> lk(g_mu);
   g_knobs.fuzz = v;
}
void setGlassRefractionIndex(float v) {
   std::lock_guard<std::mutex> lk(g_mu);
   g_knobs.glass = v;
}
void setDOFEnabled(bool v) {
   std::lock_guard<std::mutex> lk(g_mu);
   g_knobs.dof = v;
}
void setDOFAperture(float v) {
   std::lock_guard<std::mutex> lk(g_mu);
   g_knobs.aperture = v;
}
void setDOFFocusDistance(float v) {
   std::lock_guard<std::mutex> lk(g_mu);
   g_knobs.focus = v;
}

void convertAccumToDisplayCUDA(void* d_accum_buffer, unsigned char* display_image, int width, int height,
                               int channels, int num_samples, float gamma, void* d_pixel_sample_counts) {
   if (!d_accum_buffer || !display_image || num_samples <= 0) return;  // renderer_cuda_device.cu:343-344
   if (width <= 0 || height <= 0 || (channels != 3 && channels != 4) || !(gamma > 0.0f)) {
      fprintf(stderr, "rayon (convertAccumToDisplayCUDA): bad arguments\n");
      return;
   }
   std::lock_guard<std::mutex> lk(g_mu);
   rayon_ctx* c = ctx_locked();
   if (!c) return;
   LegacyState* st = state_of_acc(d_accum_buffer);
   const float4* comp = st ? st->job->view.comp : nullptr;
   const int npix = width * height;
   const size_t bytes = (size_t)npix * channels;
   uint8_t* d = nullptr;
   hipError_t e = hipMalloc(&d, bytes);
   if (e == hipSuccess)
      e = ry::launch_resolve_u8_raw((const float4*)d_accum_buffer, comp, (const int*)d_pixel_sample_counts,
                                    num_samples, npix, d, channels, gamma, c->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(display_image, d, bytes, hipMemcpyDeviceToHost, c->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
   if (d) (void)hipFree(d);
   if (e != hipSuccess) fprintf(stderr, "rayon (convertAccumToDisplayCUDA): %s\n", hipGetErrorString(e));
}

void allocateAdaptiveBuffer(void** d_pixel_sample_counts, int num_pixels) {
   if (!d_pixel_sample_counts || *d_pixel_sample_counts || num_pixels <= 0) return;
   std::lock_guard<std::mutex> lk(g_mu);
   rayon_ctx* c = ctx_locked();
   if (!c) return;
   void* p = nullptr;
   if (hipMalloc(&p, (size_t)num_pixels * sizeof(int)) != hipSuccess ||
       ry::launch_fill_u32((uint32_t*)p, 0u, num_pixels, c->stream) != hipSuccess ||
       hipStreamSynchronize(c->stream) != hipSuccess) {
      fprintf(stderr, "rayon (allocateAdaptiveBuffer): allocation failed\n");
      if (p) (void)hipFree(p);
      return;
   }
   *d_pixel_sample_counts = p;
}

void resetAdaptiveBuffer(void* d_pixel_sample_counts, int num_pixels) {
   if (!d_pixel_sample_counts || num_pixels <= 0) return;
   std::lock_guard<std::mutex> lk(g_mu);
   rayon_ctx* c = ctx_locked();
   if (!c) return;
   if (ry::launch_fill_u32((uint32_t*)d_pixel_sample_counts, 0u, num_pixels, c->stream) != hipSuccess)
      fprintf(stderr, "rayon (resetAdaptiveBuffer): launch failed\n");
   (void)hipStreamSynchronize(c->stream);
}

void freeAdaptiveBuffer(void* d_pixel_sample_counts) {
   if (d_pixel_sample_counts) (void)hipFree(d_pixel_sample_counts);
}

int countConvergedPixels(void* d_pixel_sample_counts, int num_pixels) {
   if (!d_pixel_sample_counts || num_pixels <= 0) return 0;
   std::vector<int> h((size_t)num_pixels);
   if (hipMemcpy(h.data(), d_pixel_sample_counts, h.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      fprintf(stderr, "rayon (countConvergedPixels): copy failed\n");
      return 0;
   }
   int n = 0;
   for (int v : h) n += v < 0;
   return n;
}

void renderSampleHeatmapCUDA(void* d_pixel_sample_counts, unsigned char* display_image, int width, int height,
                             int channels, int max_samples_for_scale) {
   if (!d_pixel_sample_counts || !display_image) return;
   if (width <= 0 || height <= 0 || (channels != 3 && channels != 4)) return;
   std::lock_guard<std::mutex> lk(g_mu);
   rayon_ctx* c = ctx_locked();
   if (!c) return;
   const int npix = width * height;
   const size_t bytes = (size_t)npix * channels;
   uint8_t* d = nullptr;
   hipError_t e = hipMalloc(&d, bytes);
   if (e == hipSuccess)
      e = ry::launch_heatmap((const int*)d_pixel_sample_counts, npix, d, channels, max_samples_for_scale, c->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(display_image, d, bytes, hipMemcpyDeviceToHost, c->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
   if (d) (void)hipFree(d);
   if (e != hipSuccess) fprintf(stderr, "rayon (renderSampleHeatmapCUDA): %s\n", hipGetErrorString(e));
}

}  // extern "C"
