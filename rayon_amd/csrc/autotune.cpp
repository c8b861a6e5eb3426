// autotune.cpp — autotuner of batched passes.
//
// It picks (1) the batched path of every scene left on automatic choice, by timing its
// jobs on the wavefront and on the one-launch megakernel, and (2) the wavefront's launch
// shape (path-slot capacity, queue chunk, workgroups per CU) by coordinate descent. The
// shapes change only the speed of a pass: every job accumulates its samples per pixel in
// sample order, whatever the round split, chunk or worker count. All timing runs on clones
// of the caller's jobs; the caller's jobs are never advanced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "internal.h"

namespace {

double now_ms() {
   return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Shape {
   long long capacity;
   int chunk;
   int wg;
   int shade_chunk;
   int refill;
   bool operator==(const Shape& o) const {
      return capacity == o.capacity && chunk == o.chunk && wg == o.wg && shade_chunk == o.shade_chunk &&
             refill == o.refill;
   }
};

rayon_status apply(rayon_ctx* ctx, const Shape& s) {
   rayon_status st = rayon_ctx_set_option(ctx, RAYON_OPT_WAVE_CAPACITY, s.capacity);
   if (st == RAYON_OK) st = rayon_ctx_set_option(ctx, RAYON_OPT_WAVE_CHUNK, s.chunk);
   if (st == RAYON_OK) st = rayon_ctx_set_option(ctx, RAYON_OPT_WAVE_WG_PER_CU, s.wg);
   if (st == RAYON_OK) st = rayon_ctx_set_option(ctx, RAYON_OPT_SHADE_CHUNK, s.shade_chunk);
   if (st == RAYON_OK) st = rayon_ctx_set_option(ctx, RAYON_OPT_TRACE_REFILL, s.refill);
   return st;
}

// Wall time of one pass over the clones (after an untimed pass that captures the graph
// and grows the wavefront buffers for this shape). path >= 0 forces the batched path.
rayon_status time_pass(rayon_ctx* ctx, std::vector<rayon_job*>& clones, int samples, double* ms, int path = -1) {
   for (int rep = 0; rep < 2; ++rep) {
      for (rayon_job* c : clones) {
         rayon_status st = rayon_job_reset(c);
         if (st != RAYON_OK) return st;
         if (path >= 0) c->path = path;
      }
      rayon_status st = rayon_ctx_sync(ctx);
      if (st != RAYON_OK) return st;
      const double t0 = now_ms();
      st = rayon_render_batched(ctx, clones.data(), (int)clones.size(), samples);
      if (st == RAYON_OK) st = rayon_ctx_sync(ctx);
      if (st != RAYON_OK) return st;
      *ms = now_ms() - t0;
   }
   return RAYON_OK;
}

}  // namespace

extern "C" rayon_status rayon_autotune(rayon_ctx* ctx, rayon_job* const* jobs, int32_t n, int32_t samples,
                                       rayon_tune_result* out) {
   if (!ctx || !jobs || n <= 0 || samples <= 0 || !out) {
      ry::set_error("rayon_autotune: bad arguments");
      return RAYON_ERR_INVALID;
   }
   // workload signature: job count and total work (both rounded to powers of two), the
   // largest depth, the set of scenes and the union of flags
   long long work = 0;
   int depth = 0;
   uint32_t flags = 0;
   std::vector<long long> scenes;
   for (int k = 0; k < n; ++k) {
      const rayon_job* j = jobs[k];
      if (!j || j->ctx != ctx) {
         ry::set_error("rayon_autotune: job %d is NULL or belongs to another context", k);
         return RAYON_ERR_INVALID;
      }
      work += (long long)std::min<long long>(samples, j->params.spp) * j->params.width * j->params.height;
      depth = std::max(depth, j->params.max_depth);
      flags |= j->params.flags;
      scenes.push_back((long long)(uintptr_t)j->scene);
   }
   std::sort(scenes.begin(), scenes.end());
   scenes.erase(std::unique(scenes.begin(), scenes.end()), scenes.end());
   std::vector<long long> key = {(long long)std::ceil(std::log2((double)n)),
                                 (long long)std::ceil(std::log2((double)std::max(work, 1ll))), depth,
                                 (long long)flags};
   key.insert(key.end(), scenes.begin(), scenes.end());
   {
      std::lock_guard<std::mutex> lk(ctx->tune_mu);
      auto it = ctx->tune_cache.find(key);
      if (it != ctx->tune_cache.end()) {
         *out = it->second;
         out->candidates = 0;
         out->cached = 1;
         return apply(ctx, {out->capacity, out->chunk, out->wg_per_cu, out->shade_chunk, out->refill});
      }
   }
   Shape start{};
   {
      int64_t v = 0;
      rayon_status st = rayon_ctx_get_option(ctx, RAYON_OPT_WAVE_CAPACITY, &v);
      start.capacity = v;
      if (st == RAYON_OK) st = rayon_ctx_get_option(ctx, RAYON_OPT_WAVE_CHUNK, &v);
      start.chunk = (int)v;
      if (st == RAYON_OK) st = rayon_ctx_get_option(ctx, RAYON_OPT_WAVE_WG_PER_CU, &v);
      start.wg = (int)v;
      if (st == RAYON_OK) st = rayon_ctx_get_option(ctx, RAYON_OPT_SHADE_CHUNK, &v);
      start.shade_chunk = (int)v;
      if (st == RAYON_OK) st = rayon_ctx_get_option(ctx, RAYON_OPT_TRACE_REFILL, &v);
      start.refill = (int)v;
      if (st != RAYON_OK) return st;
   }
   std::vector<rayon_job*> clones;
   auto cleanup = [&]() {
      for (rayon_job* c : clones) rayon_job_destroy(c);
      clones.clear();
   };
   for (int k = 0; k < n; ++k) {
      rayon_job* c = nullptr;
      rayon_status st = rayon_job_create(ctx, jobs[k]->scene, &jobs[k]->params, &c);
      if (st != RAYON_OK) {
         cleanup();
         return st;
      }
      clones.push_back(c);
   }
   int timed = 0;
   // batched path per scene left on AUTO: time both paths on that scene's jobs at two
   // sample counts (about 4 M and 8 M paths) and compare the added time per added sample.
   // In the real pass every job shares the wavefront's fixed cost per level, so the
   // marginal cost, not the cost of a pass of one scene alone, decides.
   {
      std::vector<rayon_scene*> done;
      for (int k = 0; k < n; ++k) {
         rayon_scene* sc = jobs[k]->scene;
         if (sc->batch_path != RAYON_PATH_AUTO || std::find(done.begin(), done.end(), sc) != done.end()) continue;
         done.push_back(sc);
         std::vector<rayon_job*> sub;
         long long px = 0;
         int spp = 1 << 30;
         for (int m = 0; m < n && sub.size() < 8; ++m)
            if (jobs[m]->scene == sc) {
               sub.push_back(clones[m]);
               px += (long long)jobs[m]->params.width * jobs[m]->params.height;
               spp = std::min(spp, jobs[m]->params.spp);
            }
         // the larger probe (2x) must stay within the jobs' sample budgets
         const int probe = (int)std::max<long long>(
             1, std::min<long long>({64, std::max(1, spp / 2), (4ll << 20) / std::max(px, 1ll) + 1}));
         double w1 = 0, w2 = 0, m1 = 0, m2 = 0;
         rayon_status st = time_pass(ctx, sub, probe, &w1, RAYON_PATH_WAVEFRONT);
         if (st == RAYON_OK) st = time_pass(ctx, sub, 2 * probe, &w2, RAYON_PATH_WAVEFRONT);
         if (st == RAYON_OK) st = time_pass(ctx, sub, probe, &m1, RAYON_PATH_MEGAKERNEL);
         if (st == RAYON_OK) st = time_pass(ctx, sub, 2 * probe, &m2, RAYON_PATH_MEGAKERNEL);
         timed += 4;
         if (st != RAYON_OK) {
            cleanup();
            return st;
         }
         sc->batch_path = (m2 - m1) < (w2 - w1) ? RAYON_PATH_MEGAKERNEL : RAYON_PATH_WAVEFRONT;
      }
      for (rayon_job* c : clones) c->path = -1;
   }
   auto measure = [&](const Shape& s, double* ms) -> rayon_status {
      rayon_status st = apply(ctx, s);
      if (st == RAYON_OK) st = time_pass(ctx, clones, samples, ms);
      ++timed;
      return st;
   };
   Shape best = start;
   double best_ms = 0, start_ms = 0;
   std::vector<Shape> tried = {start};
   rayon_status st = measure(start, &start_ms);
   best_ms = start_ms;
   // coordinate descent: trace chunk, shade chunk, workers per CU, capacity, then the
   // trace waves' refill threshold
   const long long round_cap = std::min<long long>(1ll << 28, std::max<long long>(1 << 16, work));
   const std::vector<std::vector<long long>> axes = {
      {2, 4, 8, 16, 32},
      {4, 8, 16, 32, 64},
      {0, 1, 2, 4, 8},
      {1ll << 22, 1ll << 24, 1ll << 25, 1ll << 26, 1ll << 27},
      {8, 16, 32, 64},
   };
   for (int ax = 0; ax < 5 && st == RAYON_OK; ++ax) {
      for (long long v : axes[ax]) {
         Shape s = best;
         if (ax == 0) s.chunk = (int)v;
         if (ax == 1) s.shade_chunk = (int)v;
         if (ax == 2) s.wg = (int)v;
         if (ax == 4) s.refill = (int)v;
         if (ax == 3) {
            // capacities above the work of one pass behave alike: try the exact fit once
            s.capacity = std::min(v, round_cap);
            if (s.capacity < (1 << 16)) s.capacity = 1 << 16;
         }
         if (std::find(tried.begin(), tried.end(), s) != tried.end()) continue;
         tried.push_back(s);
         double ms = 0;
         st = measure(s, &ms);
         if (st != RAYON_OK) break;
         if (ms < best_ms * 0.98) {  // a change must win by more than the noise
            best_ms = ms;
            best = s;
         }
      }
   }
   cleanup();
   if (st != RAYON_OK) {
      (void)apply(ctx, start);
      return st;
   }
   st = apply(ctx, best);
   if (st != RAYON_OK) return st;
   rayon_tune_result r{};
   r.capacity = best.capacity;
   r.chunk = best.chunk;
   r.wg_per_cu = best.wg;
   r.shade_chunk = best.shade_chunk;
   r.refill = best.refill;
   r.best_ms = best_ms;
   r.start_ms = start_ms;
   r.candidates = timed;
   r.cached = 0;
   {
      std::lock_guard<std::mutex> lk(ctx->tune_mu);
      ctx->tune_cache[key] = r;
   }
   {
      std::vector<rayon_scene*> seen;
      for (int k = 0; k < n; ++k) {
         rayon_scene* sc = jobs[k]->scene;
         if (std::find(seen.begin(), seen.end(), sc) != seen.end()) continue;
         seen.push_back(sc);
         (ry::scene_batch_path(sc) == RAYON_PATH_MEGAKERNEL ? r.mega_scenes : r.wave_scenes) += 1;
      }
   }
   *out = r;
   return RAYON_OK;
}
