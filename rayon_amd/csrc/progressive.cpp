// progressive.cpp — headless progressive renderer of one job: the render loop of the
// reference's interactive mode (RendererCUDAProgressive, renderer_cuda_progressive_host.hpp:
// 440-920) without a window. Each step renders one batch of samples, sized by the
// frame-time controller of :681-686 (scale the batch by target/measured frame time,
// clamped to [0.5, 2], at least 4 samples, at most the caller's cap), optionally at the
// progressive depth of calculateProgressiveMaxDepth (:853-876). The display image is
// resolved on a second stream into pinned host memory while the next batch renders, and
// the convergence percentage comes from the device's converged-pixel counters instead of
// a host scan of the count buffer (:688-693).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "internal.h"
#include "kernels.h"

struct rayon_progressive {
   rayon_ctx* ctx = nullptr;
   rayon_job* job = nullptr;
   rayon_progressive_config cfg{};
   int max_depth = 0;  // the job's own depth (the schedule's last step)
   int batch = 0;
   hipStream_t display_stream = nullptr;
   hipEvent_t rendered = nullptr, shown = nullptr;
   uint8_t* d_display = nullptr;
   uint8_t* h_display = nullptr;  // pinned
   size_t display_bytes = 0;
   bool display_pending = false;
   int64_t displayed_samples = 0;
};

namespace {

double now_ms() {
   return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// calculateProgressiveMaxDepth (renderer_cuda_progressive_host.hpp:853-876): shallow paths
// for the first samples, the full depth (at most 512) after 1024.
int progressive_depth(int64_t samples, int max_depth) {
   if (samples <= 4) return 4;
   if (samples <= 16) return 5;
   if (samples <= 32) return 6;
   if (samples <= 64) return 7;
   if (samples <= 128) return 8;
   if (samples <= 512) return 16;
   if (samples <= 1024) return 24;
   return std::min(512, max_depth);
}

void destroy(rayon_progressive* p) {
   if (!p) return;
   if (p->display_stream) (void)hipStreamSynchronize(p->display_stream);
   if (p->job) rayon_job_destroy(p->job);
   if (p->rendered) (void)hipEventDestroy(p->rendered);
   if (p->shown) (void)hipEventDestroy(p->shown);
   if (p->display_stream) (void)hipStreamDestroy(p->display_stream);
   if (p->d_display) (void)hipFree(p->d_display);
   if (p->h_display) (void)hipHostFree(p->h_display);
   delete p;
}

}  // namespace

extern "C" {

void rayon_progressive_config_default(rayon_progressive_config* c) {
   std::memset(c, 0, sizeof(*c));
   c->target_fps = 60.0;
   c->max_batch = 64;
   c->first_batch = 4;
   c->adaptive_depth = 0;
   c->channels = 4;
   c->gamma = 2.0f;
}

rayon_status rayon_progressive_create(rayon_ctx* ctx, rayon_scene* scene, const rayon_job_params* params,
                                      const rayon_progressive_config* cfg, rayon_progressive** out) {
   if (!ctx || !scene || !params || !out) {
      ry::set_error("rayon_progressive_create: NULL argument");
      return RAYON_ERR_INVALID;
   }
   auto* p = new rayon_progressive();
   p->ctx = ctx;
   if (cfg)
      p->cfg = *cfg;
   else
      rayon_progressive_config_default(&p->cfg);
   const rayon_progressive_config& c = p->cfg;
   if (!(c.target_fps > 0) || c.max_batch < 1 || c.first_batch < 1 || (c.channels != 3 && c.channels != 4) ||
       !(c.gamma > 0)) {
      delete p;
      ry::set_error("rayon_progressive_create: bad configuration");
      return RAYON_ERR_INVALID;
   }
   rayon_status st = rayon_job_create(ctx, scene, params, &p->job);
   if (st != RAYON_OK) {
      delete p;
      return st;
   }
   p->max_depth = params->max_depth;
   p->batch = std::min(c.first_batch, c.max_batch);
   p->display_bytes = (size_t)params->width * params->height * c.channels;
   hipError_t e = hipSetDevice(ctx->device);
   if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->display_stream, hipStreamNonBlocking);
   if (e == hipSuccess) e = hipEventCreateWithFlags(&p->rendered, hipEventDisableTiming);
   if (e == hipSuccess) e = hipEventCreateWithFlags(&p->shown, hipEventDisableTiming);
   if (e == hipSuccess) e = hipMalloc(&p->d_display, p->display_bytes);
   if (e == hipSuccess) e = hipHostMalloc(&p->h_display, p->display_bytes, hipHostMallocDefault);
   if (e != hipSuccess) {
      destroy(p);
      return ry::hip_fail(e, "rayon_progressive_create");
   }
   *out = p;
   return RAYON_OK;
}

rayon_status rayon_progressive_step(rayon_progressive* p, rayon_progressive_frame* f) {
   if (!p || !f) {
      ry::set_error("rayon_progressive_step: NULL argument");
      return RAYON_ERR_INVALID;
   }
   std::memset(f, 0, sizeof(*f));
   rayon_job* j = p->job;
   rayon_ctx* ctx = p->ctx;
   const int64_t left = (int64_t)j->params.spp - j->samples_done;
   rayon_job_stats s0;
   rayon_status st = rayon_job_stats_get(ctx, j, &s0);
   if (st != RAYON_OK) return st;
   const int64_t npix = (int64_t)j->params.width * j->params.height;
   if (left <= 0 || s0.converged_pixels >= npix) {
      f->samples = j->samples_done;
      f->done = 1;
      f->convergence = 100.0 * (double)s0.converged_pixels / (double)npix;
      return RAYON_OK;
   }
   const int batch = (int)std::min<int64_t>(p->batch, left);
   // the depth of this batch (the job's own when the schedule is off)
   j->params.max_depth = p->cfg.adaptive_depth ? progressive_depth(j->samples_done + batch, p->max_depth)
                                               : p->max_depth;
   // the display copy of the previous batch must finish reading the sums before they change
   RY_HIP(hipStreamWaitEvent(ctx->stream, p->shown, 0));
   const double t0 = now_ms();
   if (j->params.flags & RAYON_FLAG_REF_RNG)
      st = rayon_job_render_plain(ctx, j, batch);
   else
      st = rayon_render_batched(ctx, &j, 1, batch);
   if (st != RAYON_OK) return st;
   RY_HIP(hipEventRecord(p->rendered, ctx->stream));
   // display: resolve on the second stream into pinned memory, overlapping the next batch
   RY_HIP(hipStreamWaitEvent(p->display_stream, p->rendered, 0));
   ry::fill_jobview(j);
   RY_HIP(ry::launch_resolve_u8(j->view, p->d_display, p->cfg.channels, p->cfg.gamma, p->display_stream));
   RY_HIP(hipMemcpyAsync(p->h_display, p->d_display, p->display_bytes, hipMemcpyDeviceToHost, p->display_stream));
   RY_HIP(hipEventRecord(p->shown, p->display_stream));
   p->display_pending = true;
   // the frame time includes the batch's own completion, as the reference's does
   RY_HIP(hipEventSynchronize(p->rendered));
   const double ms = now_ms() - t0;
   rayon_job_stats s1;
   st = rayon_job_stats_get(ctx, j, &s1);
   if (st != RAYON_OK) return st;
   const double target = 1000.0 / p->cfg.target_fps;
   const double scale = std::min(2.0, std::max(0.5, target / std::max(ms, 1e-3)));
   p->batch = std::min(p->cfg.max_batch, std::max(4, (int)std::lround(batch * scale)));
   p->displayed_samples = j->samples_done;
   f->samples = j->samples_done;
   f->batch = batch;
   f->depth = j->params.max_depth;
   f->next_batch = p->batch;
   f->frame_ms = ms;
   f->segments = s1.segments;
   f->convergence = 100.0 * (double)s1.converged_pixels / (double)npix;
   f->done = (j->samples_done >= j->params.spp || s1.converged_pixels >= npix) ? 1 : 0;
   j->params.max_depth = p->max_depth;
   return RAYON_OK;
}

rayon_status rayon_progressive_display(rayon_progressive* p, uint8_t* out, int64_t* samples) {
   if (!p || !out) {
      ry::set_error("rayon_progressive_display: NULL argument");
      return RAYON_ERR_INVALID;
   }
   if (!p->display_pending) {
      ry::set_error("rayon_progressive_display: no batch rendered yet");
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipEventSynchronize(p->shown));
   std::memcpy(out, p->h_display, p->display_bytes);
   if (samples) *samples = p->displayed_samples;
   return RAYON_OK;
}

rayon_status rayon_progressive_reset(rayon_progressive* p, const rayon_camera* camera) {
   if (!p) {
      ry::set_error("rayon_progressive_reset: NULL argument");
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipEventSynchronize(p->shown));
   if (camera) rayon_job_params_set_camera(&p->job->params, camera);
   ry::fill_jobview(p->job);
   p->batch = std::min(p->cfg.first_batch, p->cfg.max_batch);
   p->display_pending = false;
   return rayon_job_reset(p->job);
}

rayon_job* rayon_progressive_job(rayon_progressive* p) { return p ? p->job : nullptr; }

rayon_status rayon_progressive_destroy(rayon_progressive* p) {
   destroy(p);
   return RAYON_OK;
}

}  // extern "C"
