// server.cpp — continuous-batching render server (no counterpart in the reference, whose
// renderers run one job per process; the closest is the progressive loop of
// renderer_cuda_progressive_host.hpp:440-920 with its FPS-driven batch controller
// :681-686, here generalised to many concurrent jobs).
//
// One server thread owns the device work of a context. Each step it
//   1. admits queued jobs while their device memory (accumulation, counts and a u8 frame,
//      plus each newly referenced scene) fits the HBM budget and the active-job cap;
//   2. renders a first frame for each job admitted in this step: first_pass_samples
//      samples of the job by the plain kernel's arithmetic (all new jobs in one launch),
//      resolved on the device; those samples are the job's first accumulated samples
//      (the same streams and arithmetic as the batched passes, so the result does not
//      change); time-to-first-frame is stamped when that work completes. Queued jobs
//      that have no buffers yet get a first frame too, which they cannot keep;
//   3. advances the oldest other active jobs together by one batched pass whose work
//      (pixels x samples) a controller steers toward target_pass_ms: the jobs are taken
//      oldest first until their pixels reach the work budget (at least one job, at most
//      max_pass_pixels), and the sample count fills the budget (so a step, and with it
//      the wait for a first frame, stays short under any load);
//   4. retires jobs whose sample budget is exhausted or whose pixels have all converged
//      (time-to-converge), and jobs that were cancelled.
// Jobs keep their buffers until released, so results can be read after completion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"
#include "kernels.h"

namespace {

double now_s() {
   return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SJob {
   uint64_t id = 0;
   rayon_scene* scene = nullptr;
   rayon_job_params params{};
   rayon_job* job = nullptr;
   uint8_t* frame = nullptr;  // device u8 RGBA of the first frame
   int32_t state = RAYON_JOB_QUEUED;
   double t_arrive = 0, t_admit = 0, t_first = 0, t_done = 0;
   int64_t bytes = 0;
   int64_t samples = 0;
   int64_t segments = 0;
   int64_t live_px = 0;  // pixels still sampling (all, until an adaptive job reports converged ones)
   bool cancel = false;
   bool released = false;
   bool accounted = false;  // bytes (and a scene reference) are held in the budget
   bool needs_clear = false;  // accumulation buffers not cleared yet (before the first pass)
   bool counters_only = false;  // its first frame holds its first samples: clear the counters only
   bool pv_accum = false;       // its first frame in flight stores the job's first samples
   bool first_kept = false;     // its first frame's samples are its first samples
   bool previewing = false;   // its first frame is being rendered
   bool preview_held = false;  // its frame is held in the first-frame reserve (not admitted)
};

int64_t frame_bytes(const SJob* j) { return (int64_t)j->params.width * j->params.height * 4; }

constexpr int kPreviewStreams = 2;

int64_t job_bytes(const rayon_job_params& p) {
   const int64_t npix = (int64_t)p.width * p.height;
   int64_t b = npix * (16 + 16 + 4 + 4) + 8 * 8 * ry::kCounterSlots;
   if (p.flags & RAYON_FLAG_ADAPTIVE_REF) b += npix * 16;
   return b;
}

}  // namespace

struct rayon_server {
   rayon_ctx* ctx = nullptr;
   rayon_ctx* first_ctx = nullptr;  // job and frame allocations (ctx->serve_first)
   hipStream_t pv_stream[kPreviewStreams] = {};  // first frames: first_ctx's stream and ctx->serve_pv
   int next_pv = 0;
   rayon_server_config cfg{};
   std::thread th;
   std::mutex mu;
   std::condition_variable cv, done_cv;
   bool stop = false;
   uint64_t next_id = 1;
   std::deque<SJob*> queue;
   std::vector<SJob*> active;
   std::unordered_map<uint64_t, SJob*> jobs;
   std::unordered_map<rayon_scene*, int> scene_refs;
   int64_t used_bytes = 0;     // admitted jobs and their scenes
   int64_t preview_bytes = 0;  // first frames of jobs not admitted yet
   uint64_t wake_seq = 0;      // arrivals and cancellations, for the server thread
   int pass_samples = 1;
   double pass_work = 4.0 * (1 << 20);  // pixel-samples of the next refinement pass
   // statistics
   int64_t passes = 0, first_passes = 0, total_segments = 0, completed = 0, failed = 0;
   int64_t first_kept = 0, first_kept_paths = 0;
   double busy_s = 0, t_start = 0;
   double pass_ms_sum = 0, first_ms_sum = 0, check_ms_sum = 0, pass_paths_sum = 0;
   int pass_jobs = 0;
   std::string last_error;

   int64_t scene_bytes(rayon_scene* s) { return s->info.device_bytes; }

   void fail_all(std::vector<SJob*>& js, const char* what) {
      std::lock_guard<std::mutex> lk(mu);
      last_error = std::string(what) + ": " + rayon_last_error();
      for (SJob* j : js) {
         j->state = RAYON_JOB_FAILED;
         j->t_done = now_s();
         ++failed;
      }
   }

   // Releases device memory of a finished job and its share of the budget.
   void free_job_locked(SJob* j) {
      if (j->job) {
         rayon_job_destroy(j->job);
         j->job = nullptr;
      }
      if (j->accounted) {
         j->accounted = false;
         used_bytes -= j->bytes;
         if (--scene_refs[j->scene] == 0) {
            used_bytes -= scene_bytes(j->scene);
            scene_refs.erase(j->scene);
         }
      }
      // a frame still being rendered is freed when its first frame completes
      if (j->frame && !j->previewing) {
         (void)hipFreeAsync(j->frame, first_ctx->stream);
         j->frame = nullptr;
      }
      if (!j->frame && j->preview_held) {  // (also a reserve whose allocation failed)
         preview_bytes -= frame_bytes(j);
         j->preview_held = false;
      }
   }

   // A job whose set-up failed, queued or admitted.
   void fail_locked(SJob* j) {
      last_error = rayon_last_error();
      j->state = RAYON_JOB_FAILED;
      j->t_done = now_s();
      ++failed;
      auto q = std::find(queue.begin(), queue.end(), j);
      if (q != queue.end()) queue.erase(q);
      auto a = std::find(active.begin(), active.end(), j);
      if (a != active.end()) active.erase(a);
      free_job_locked(j);
      done_cv.notify_all();
   }

   // The view a first frame renders: the job's own, or one built from its parameters
   // (a job not admitted yet has no buffers; the preview kernel touches none).
   ry::JobView preview_view(const SJob* j) {
      if (j->job) return j->job->view;
      rayon_job tmp;
      tmp.ctx = first_ctx;
      tmp.scene = j->scene;
      tmp.params = j->params;
      ry::fill_jobview(&tmp);
      return tmp.view;
   }

   // Marks finished jobs DONE/FAILED and takes them out of the active set.
   void finish_locked(const std::vector<SJob*>& finished) {
      const double t = now_s();
      for (SJob* j : finished) {
         if (j->state != RAYON_JOB_FAILED) {
            j->state = RAYON_JOB_DONE;
            ++completed;
         }
         j->t_done = t;
         total_segments += j->segments;
         active.erase(std::find(active.begin(), active.end(), j));
         if (j->released) free_job_locked(j);
      }
      if (!finished.empty()) done_cv.notify_all();
   }

   // Reads the statistics of jobs a pass on `c` touched (the pass has completed) and
   // returns those that are done. Fixed-budget jobs that have samples left need no read.
   std::vector<SJob*> check(rayon_ctx* c, const std::vector<SJob*>& touched) {
      std::vector<SJob*> finished;
      for (SJob* j : touched) {
         if (j->state == RAYON_JOB_FAILED) {
            finished.push_back(j);
            continue;
         }
         const bool adaptive = (j->params.flags & (RAYON_FLAG_ADAPTIVE_REF | RAYON_FLAG_ADAPTIVE_STDERR)) != 0;
         if (!adaptive && j->job->samples_done < j->params.spp) {
            std::lock_guard<std::mutex> lk(mu);
            j->samples = j->job->samples_done;
            continue;
         }
         rayon_job_stats s;
         if (rayon_job_stats_get(c, j->job, &s) != RAYON_OK) continue;
         std::lock_guard<std::mutex> lk(mu);
         j->samples = s.samples_done;
         j->segments = s.segments;
         j->live_px = std::max<int64_t>(1, (int64_t)j->params.width * j->params.height - s.converged_pixels);
         if (s.done) finished.push_back(j);
      }
      return finished;
   }

   // First frames in flight: one launch on a preview stream, and the event after it. Jobs
   // whose first frame accumulates have a slice of `part` (pinned host memory) for the
   // statistics of their tiles.
   struct FirstBatch {
      std::vector<SJob*> jobs;
      std::vector<size_t> part_off;  // per job: first entry in `part`, or kNoPart
      unsigned long long* part = nullptr;
      size_t part_bytes = 0;
      hipEvent_t ready = nullptr, done = nullptr;
      double t0 = 0;
   };
   static constexpr size_t kNoPart = ~(size_t)0;
   // Pinned statistics buffers, kept for the server's lifetime and reused.
   std::vector<std::pair<unsigned long long*, size_t>> pinned_free;
   unsigned long long* take_pinned(size_t bytes, size_t& got) {
      for (size_t i = 0; i < pinned_free.size(); ++i)
         if (pinned_free[i].second >= bytes) {
            auto r = pinned_free[i];
            pinned_free.erase(pinned_free.begin() + (long)i);
            got = r.second;
            return r.first;
         }
      got = std::max<size_t>(bytes, (size_t)1 << 20);
      void* p = nullptr;
      if (hipHostMalloc(&p, got, hipHostMallocDefault) != hipSuccess) {
         got = 0;
         return nullptr;
      }
      return (unsigned long long*)p;
   }
   std::vector<hipEvent_t> event_pool;
   hipEvent_t take_event() {
      hipEvent_t e = nullptr;
      if (!event_pool.empty()) {
         e = event_pool.back();
         event_pool.pop_back();
      } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
         e = nullptr;
      }
      return e;
   }

   // The server thread. First frames run on streams of their own, so they overlap the
   //   refinement pass in flight on the main context instead of waiting for it, and the
   //   thread does not wait for them either:
   //   - admit queued jobs by the HBM budget less a first-frame reserve (1/8), create them;
   //   - launch first frames on the next preview stream (no wait) for admitted jobs and
   //     for queued jobs whose frames fit the reserve, so a first frame does not wait for
   //     admission; a job's time-to-first-frame is stamped when the thread sees that
   //     launch complete;
   //   - when no refinement pass is in flight, start one over the oldest jobs that have
   //     a first frame;
   //   - when the pass in flight has completed, check its jobs and steer the next size.
   void loop() {
      (void)hipSetDevice(ctx->device);
      std::vector<SJob*> in_flight;
      std::vector<FirstBatch> firsts;
      bool flying = false;
      double t_pass = 0;
      hipEvent_t pass_done = nullptr, alloc_done = nullptr;
      (void)hipEventCreateWithFlags(&pass_done, hipEventDisableTiming);
      (void)hipEventCreateWithFlags(&alloc_done, hipEventDisableTiming);
      uint64_t seen_seq = 0;
      for (;;) {
         std::vector<SJob*> fresh, previews;
         {
            std::unique_lock<std::mutex> lk(mu);
            if (flying || !firsts.empty())  // wake up for arrivals, or soon to look at the GPU work
               cv.wait_for(lk, std::chrono::microseconds(200), [&] { return stop || wake_seq != seen_seq; });
            else
               cv.wait(lk, [&] {  // a queued job is admitted whatever its size once nothing is active
                  return stop || wake_seq != seen_seq || !active.empty() || !queue.empty();
               });
            seen_seq = wake_seq;
            if (stop) break;
            // retire cancelled jobs that no pass is using
            for (auto it = active.begin(); it != active.end();) {
               SJob* j = *it;
               if (j->cancel && !j->previewing && std::find(in_flight.begin(), in_flight.end(), j) == in_flight.end()) {
                  rayon_job_stats js;
                  if (j->job && !j->needs_clear && rayon_job_stats_get(ctx, j->job, &js) == RAYON_OK) {
                     j->samples = js.samples_done;
                     j->segments = js.segments;
                  }
                  j->state = RAYON_JOB_CANCELLED;
                  j->t_done = now_s();
                  free_job_locked(j);
                  it = active.erase(it);
                  done_cv.notify_all();
               } else {
                  ++it;
               }
            }
            // admission by HBM budget and active-job cap (FIFO: a job that does not fit
            // blocks the ones behind it, so large jobs are not starved)
            const int64_t reserve = cfg.hbm_budget_bytes / 8;
            while (!queue.empty() && (int)active.size() < cfg.max_active_jobs) {
               SJob* j = queue.front();
               if (j->cancel) {
                  queue.pop_front();
                  j->state = RAYON_JOB_CANCELLED;
                  j->t_done = now_s();
                  free_job_locked(j);
                  done_cv.notify_all();
                  continue;
               }
               const int64_t need = j->bytes + (scene_refs.count(j->scene) ? 0 : scene_bytes(j->scene));
               if (used_bytes + need > cfg.hbm_budget_bytes - reserve && (!active.empty() || !fresh.empty())) break;
               queue.pop_front();
               used_bytes += need;
               scene_refs[j->scene]++;
               j->accounted = true;
               if (j->preview_held) {  // the frame is part of the job's bytes from here
                  preview_bytes -= frame_bytes(j);
                  j->preview_held = false;
               }
               j->t_admit = now_s();
               fresh.push_back(j);
               active.push_back(j);  // counted as active from here, so drain() waits for it
            }
            for (SJob* j : fresh)
               if (!j->frame) previews.push_back(j);
            for (SJob* j : queue) {
               if (j->cancel || j->frame) continue;
               if (preview_bytes + frame_bytes(j) > reserve) break;
               preview_bytes += frame_bytes(j);
               j->preview_held = true;
               previews.push_back(j);
            }
         }
         // (1) admitted jobs: buffers on first_ctx. Its allocations must not queue behind
         // the refinement pass in flight on the main stream; the main stream waits for them
         // by event (the stream-ordered pool may hand out memory whose earlier use on
         // first_ctx's stream is still in flight), and so does a preview stream. The first
         // frame writes the job's sums, compensations and counts whole; the counters (and,
         // for a job whose first frame came before its buffers, everything) are cleared on
         // the main stream just before the job's first refinement pass, so a first frame
         // waits for nothing but its own kernel.
         for (SJob* j : fresh) {
            rayon_status st = ry::job_create(first_ctx, j->scene, &j->params, &j->job, false, false);
            std::lock_guard<std::mutex> lk(mu);
            if (st != RAYON_OK) {
               fail_locked(j);
               continue;
            }
            j->needs_clear = true;
            j->state = RAYON_JOB_RUNNING;
            j->live_px = (int64_t)j->params.width * j->params.height;
         }
         if (!fresh.empty() && (hipEventRecord(alloc_done, first_ctx->stream) != hipSuccess ||
                                hipStreamWaitEvent(ctx->stream, alloc_done, 0) != hipSuccess)) {
            (void)hipStreamSynchronize(first_ctx->stream);  // the same order, by the host
         }
         // (1a) first frames: the job's first samples rendered and resolved to RGBA8 by one
         // launch of the preview kernel for all of them (the plain kernel's per-pixel
         // arithmetic, which the batched path matches bit for bit, tables in the kernel
         // arguments). An admitted job keeps them as its first samples (the refinement
         // passes go on from there), so a served job's result equals the job rendered
         // alone; a queued job's first frame is only a frame: its samples are rendered
         // again once it has buffers.
         if (!previews.empty()) {
            FirstBatch fb;
            fb.t0 = now_s();
            std::vector<ry::JobView> views;
            std::vector<uint8_t*> frames;
            std::vector<unsigned long long*> parts;
            {
               std::lock_guard<std::mutex> lk(mu);
               size_t entries = 0;
               for (SJob* j : previews) {
                  if (j->state == RAYON_JOB_FAILED || j->state == RAYON_JOB_CANCELLED) continue;
                  if (hipMallocAsync((void**)&j->frame, (size_t)frame_bytes(j), first_ctx->stream) != hipSuccess) {
                     j->frame = nullptr;
                     ry::set_error("allocation of the first-frame buffer failed");
                     fail_locked(j);
                     continue;
                  }
                  views.push_back(preview_view(j));
                  frames.push_back(j->frame);
                  j->previewing = true;
                  // an admitted job whose buffers are still uncleared takes the first frame's
                  // samples as its own first samples (same streams, same arithmetic as the
                  // batched passes that follow)
                  j->pv_accum = j->job && j->needs_clear && !j->counters_only &&
                                !(j->params.flags & RAYON_FLAG_REF_RNG) && j->params.spp >= cfg.first_pass_samples;
                  fb.part_off.push_back(j->pv_accum ? entries : kNoPart);
                  if (j->pv_accum)
                     entries += (size_t)ry::plain_tiles(j->params.width, j->params.height) * ry::kPreviewTileCounts;
                  fb.jobs.push_back(j);
               }
               if (entries) fb.part = take_pinned(entries * sizeof(unsigned long long), fb.part_bytes);
               for (size_t k = 0; k < fb.jobs.size(); ++k) {
                  if (!fb.part && fb.part_off[k] != kNoPart) {  // no pinned memory: a first frame only
                     fb.part_off[k] = kNoPart;
                     fb.jobs[k]->pv_accum = false;
                  }
                  parts.push_back(fb.part_off[k] == kNoPart ? nullptr : fb.part + fb.part_off[k]);
               }
            }
            if (!fb.jobs.empty()) {
               fb.ready = take_event();
               fb.done = take_event();
               hipStream_t ps = pv_stream[next_pv++ % kPreviewStreams];
               const bool ok = fb.ready && fb.done && hipEventRecord(fb.ready, first_ctx->stream) == hipSuccess &&
                               hipStreamWaitEvent(ps, fb.ready, 0) == hipSuccess &&
                               ry::launch_preview(views.data(), frames.data(), parts.data(), (int)views.size(),
                                                  cfg.first_pass_samples, 2.0f, ps) == hipSuccess &&
                               hipEventRecord(fb.done, ps) == hipSuccess;
               if (!ok) {
                  (void)hipStreamSynchronize(ps);
                  if (fb.done) event_pool.push_back(fb.done);
                  fb.done = nullptr;  // completes below as failed
               }
               firsts.push_back(std::move(fb));
            }
         }
         // (1b) first frames that have completed
         for (size_t k = 0; k < firsts.size();) {
            FirstBatch& fb = firsts[k];
            const hipError_t q = fb.done ? hipEventQuery(fb.done) : hipErrorInvalidValue;
            if (q == hipErrorNotReady) {
               ++k;
               continue;
            }
            const double t = now_s();
            std::lock_guard<std::mutex> lk(mu);
            if (q != hipSuccess) last_error = std::string("first frame: ") + hipGetErrorString(q);
            ++first_passes;
            first_ms_sum += (t - fb.t0) * 1e3;
            for (size_t m = 0; m < fb.jobs.size(); ++m) {
               SJob* j = fb.jobs[m];
               j->previewing = false;
               const bool accum = j->pv_accum;
               j->pv_accum = false;
               if (j->state == RAYON_JOB_CANCELLED || j->state == RAYON_JOB_FAILED) {
                  if (j->released || j->state == RAYON_JOB_CANCELLED) free_job_locked(j);  // the frame now
               } else if (q == hipSuccess) {
                  j->t_first = t;
                  if (accum) {  // the job's first samples are in its buffers; its counters are not
                     const unsigned long long* pp = fb.part + fb.part_off[m];
                     const size_t tiles = (size_t)ry::plain_tiles(j->params.width, j->params.height);
                     int64_t c[ry::kPreviewTileCounts] = {0, 0, 0, 0};
                     for (size_t i = 0; i < tiles; ++i)
                        for (int k = 0; k < ry::kPreviewTileCounts; ++k) c[k] += (int64_t)pp[i * ry::kPreviewTileCounts + k];
                     for (int k = 0; k < ry::kPreviewTileCounts; ++k) j->job->pre_counts[k] += c[k];
                     j->job->samples_done = cfg.first_pass_samples;
                     j->samples = j->job->samples_done;
                     j->segments = c[0];
                     j->counters_only = true;
                     j->first_kept = true;
                     ++first_kept;
                     first_kept_paths += c[2];
                  }
               } else {
                  fail_locked(j);
               }
            }
            if (fb.part) pinned_free.push_back({fb.part, fb.part_bytes});
            for (hipEvent_t e : {fb.ready, fb.done})
               if (e) event_pool.push_back(e);
            firsts.erase(firsts.begin() + (long)k);
         }
         // (2) the refinement pass in flight, once complete
         if (flying && hipEventQuery(pass_done) == hipSuccess) {
            const double ms = (now_s() - t_pass) * 1e3;
            // controller toward target_pass_ms (the ratio rule of
            // renderer_cuda_progressive_host.hpp:681-686, clamped to [0.5, 2] per step)
            const double ratio = std::min(2.0, std::max(0.5, cfg.target_pass_ms / std::max(ms, 0.05)));
            double done = 0;
            for (SJob* j : in_flight) done += (double)j->live_px;
            pass_work = std::min(std::max(done * pass_samples * ratio, 65536.0), 1e12);
            const double tc = now_s();
            std::vector<SJob*> finished = check(ctx, in_flight);
            std::lock_guard<std::mutex> lk(mu);
            ++passes;
            busy_s += ms * 1e-3;
            pass_ms_sum += ms;
            check_ms_sum += (now_s() - tc) * 1e3;
            pass_paths_sum += done * pass_samples;
            pass_jobs = (int)in_flight.size();
            finish_locked(finished);
            in_flight.clear();
            flying = false;
         }
         // (3) start the next refinement pass over the oldest jobs that have a frame
         if (!flying) {
            std::vector<SJob*> run;
            {
               std::lock_guard<std::mutex> lk(mu);
               int64_t px = 0;
               for (SJob* a : active) {
                  if (a->state != RAYON_JOB_RUNNING || a->t_first <= 0 || a->cancel) continue;
                  const int64_t p = a->live_px;
                  if (!run.empty() && (px + p > cfg.max_pass_pixels || (double)px >= pass_work)) break;
                  run.push_back(a);
                  px += p;
               }
               if (px > 0)
                  pass_samples = (int)std::min<double>(cfg.max_pass_samples, std::max(1.0, pass_work / (double)px));
            }
            if (!run.empty()) {
               std::vector<rayon_job*> h;
               std::vector<uint32_t*> cp;
               std::vector<long long> cn;
               for (SJob* j : run) {
                  h.push_back(j->job);
                  if (j->needs_clear) {
                     uint32_t* p[8];
                     long long n[8];
                     const int m = ry::job_clear_ranges(j->job, p, n);
                     const int first = j->counters_only ? m - 1 : 0;  // the counters are the last range
                     cp.insert(cp.end(), p + first, p + m);
                     cn.insert(cn.end(), n + first, n + m);
                  }
               }
               t_pass = now_s();
               const bool cleared =
                   cp.empty() || ry::launch_fill_ranges(cp.data(), cn.data(), (int)cp.size(), 0u, ctx->stream) == hipSuccess;
               if (cleared)
                  for (SJob* j : run) j->needs_clear = j->counters_only = false;
               if (!cleared || rayon_render_batched(ctx, h.data(), (int)h.size(), pass_samples) != RAYON_OK ||
                   hipEventRecord(pass_done, ctx->stream) != hipSuccess) {
                  fail_all(run, "batched pass");
                  (void)rayon_ctx_sync(ctx);
                  std::lock_guard<std::mutex> lk(mu);
                  finish_locked(run);
               } else {
                  in_flight = run;
                  flying = true;
               }
            }
         }
      }
      (void)rayon_ctx_sync(ctx);
      for (hipStream_t ps : pv_stream) (void)hipStreamSynchronize(ps);
      for (FirstBatch& fb : firsts) {
         for (SJob* j : fb.jobs) j->previewing = false;
         for (hipEvent_t e : {fb.ready, fb.done})
            if (e) event_pool.push_back(e);
      }
      for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
      event_pool.clear();
      for (FirstBatch& fb : firsts)
         if (fb.part) pinned_free.push_back({fb.part, fb.part_bytes});
      for (auto& pb : pinned_free) (void)hipHostFree(pb.first);
      pinned_free.clear();
      (void)hipEventDestroy(pass_done);
      (void)hipEventDestroy(alloc_done);
   }
};

extern "C" {

void rayon_server_config_default(rayon_server_config* c) {
   std::memset(c, 0, sizeof(*c));
   c->hbm_budget_bytes = 200ll << 30;  // 200 GiB of the 288 GB
   c->max_active_jobs = 1 << 16;
   c->first_pass_samples = 1;
   c->max_pass_samples = 64;
   c->target_pass_ms = 80.0;
   c->max_pass_pixels = 16ll << 20;
}

rayon_status rayon_server_create(rayon_ctx* ctx, const rayon_server_config* cfg, rayon_server** out) {
   if (!ctx || !out) {
      ry::set_error("rayon_server_create: NULL argument");
      return RAYON_ERR_INVALID;
   }
   auto* s = new rayon_server();
   s->ctx = ctx;
   if (cfg)
      s->cfg = *cfg;
   else
      rayon_server_config_default(&s->cfg);
   if (s->cfg.max_active_jobs < 1 || s->cfg.first_pass_samples < 1 || s->cfg.max_pass_samples < 1 ||
       !(s->cfg.target_pass_ms > 0) || s->cfg.hbm_budget_bytes <= 0 || s->cfg.max_pass_pixels <= 0) {
      delete s;
      ry::set_error("rayon_server_create: bad configuration");
      return RAYON_ERR_INVALID;
   }
   // The first-frame streams live with the context (created once, in this order, after the
   // context's own stream): a process then holds the same four streams whatever number of
   // servers it runs, and first frames behave the same in each. With three streams created
   // and destroyed per server, first frames took 2.7 times longer on average in a process's
   // first server than in its second (p99 time-to-first-frame 153 vs 40 ms at 200 jobs/s):
   // how a process's streams sit on the device's hardware queues decides how those queues
   // share the compute units.
   if (!ctx->serve_first && rayon_ctx_create(ctx->device, &ctx->serve_first) != RAYON_OK) {
      delete s;
      return RAYON_ERR_HIP;
   }
   if (!ctx->serve_pv && hipStreamCreateWithFlags(&ctx->serve_pv, hipStreamNonBlocking) != hipSuccess) {
      ctx->serve_pv = nullptr;
      ry::set_error("rayon_server_create: hipStreamCreateWithFlags failed");
      delete s;
      return RAYON_ERR_HIP;
   }
   s->first_ctx = ctx->serve_first;
   s->pv_stream[0] = ctx->serve_first->stream;
   s->pv_stream[1] = ctx->serve_pv;
   // passes change their job set every step: replaying captured graphs would not pay
   rayon_ctx_set_option(ctx, RAYON_OPT_GRAPH, 0);
   rayon_ctx_set_option(s->first_ctx, RAYON_OPT_GRAPH, 0);
   s->t_start = now_s();
   s->th = std::thread([s] { s->loop(); });
   *out = s;
   return RAYON_OK;
}

rayon_status rayon_server_submit(rayon_server* s, rayon_scene* scene, const rayon_job_params* p, uint64_t* id) {
   if (!s || !scene || !p) {
      ry::set_error("rayon_server_submit: NULL argument");
      return RAYON_ERR_INVALID;
   }
   if (p->width <= 0 || p->height <= 0 || p->spp <= 0 || (p->flags & RAYON_FLAG_REF_RNG)) {
      ry::set_error("rayon_server_submit: bad resolution or spp, or a reference-stream job (not batchable)");
      return RAYON_ERR_INVALID;
   }
   auto* j = new SJob();
   j->scene = scene;
   j->params = *p;
   j->bytes = job_bytes(*p) + (int64_t)p->width * p->height * 4;
   j->t_arrive = now_s();
   std::lock_guard<std::mutex> lk(s->mu);
   if (j->bytes + s->scene_bytes(scene) > s->cfg.hbm_budget_bytes) {
      delete j;
      ry::set_error("rayon_server_submit: job needs more than the whole HBM budget");
      return RAYON_ERR_BUDGET;
   }
   j->id = s->next_id++;
   s->jobs[j->id] = j;
   s->queue.push_back(j);
   if (id) *id = j->id;
   ++s->wake_seq;
   s->cv.notify_one();
   return RAYON_OK;
}

rayon_status rayon_server_job_info_get(rayon_server* s, uint64_t id, rayon_server_job_info* out) {
   if (!s || !out) {
      ry::set_error("rayon_server_job_info_get: NULL argument");
      return RAYON_ERR_INVALID;
   }
   std::lock_guard<std::mutex> lk(s->mu);
   auto it = s->jobs.find(id);
   if (it == s->jobs.end()) {
      ry::set_error("rayon_server_job_info_get: unknown job %llu", (unsigned long long)id);
      return RAYON_ERR_INVALID;
   }
   const SJob* j = it->second;
   out->state = j->state;
   out->first_kept = j->first_kept ? 1 : 0;
   out->samples_done = j->samples;
   out->segments = j->segments;
   out->ttff_ms = j->t_first > 0 ? (j->t_first - j->t_arrive) * 1e3 : -1.0;
   out->ttc_ms = (j->state == RAYON_JOB_DONE && j->t_done > 0) ? (j->t_done - j->t_arrive) * 1e3 : -1.0;
   out->queue_ms = j->t_admit > 0 ? (j->t_admit - j->t_arrive) * 1e3 : -1.0;
   return RAYON_OK;
}

rayon_status rayon_server_result(rayon_server* s, uint64_t id, float* rgb) {
   if (!s || !rgb) {
      ry::set_error("rayon_server_result: NULL argument");
      return RAYON_ERR_INVALID;
   }
   rayon_job* job = nullptr;
   {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->jobs.find(id);
      if (it == s->jobs.end() || it->second->state != RAYON_JOB_DONE || !it->second->job) {
         ry::set_error("rayon_server_result: job %llu is not finished or was released", (unsigned long long)id);
         return RAYON_ERR_INVALID;
      }
      job = it->second->job;
   }
   return rayon_job_resolve_linear(s->ctx, job, rgb);
}

rayon_status rayon_server_first_frame(rayon_server* s, uint64_t id, uint8_t* rgba) {
   if (!s || !rgba) {
      ry::set_error("rayon_server_first_frame: NULL argument");
      return RAYON_ERR_INVALID;
   }
   std::lock_guard<std::mutex> lk(s->mu);  // keeps the frame alive while it is copied
   auto it = s->jobs.find(id);
   if (it == s->jobs.end() || it->second->t_first <= 0 || !it->second->frame) {
      ry::set_error("rayon_server_first_frame: job %llu has no first frame (yet)", (unsigned long long)id);
      return RAYON_ERR_INVALID;
   }
   const SJob* j = it->second;
   (void)hipSetDevice(s->ctx->device);
   RY_HIP(hipMemcpy(rgba, j->frame, (size_t)j->params.width * j->params.height * 4, hipMemcpyDeviceToHost));
   return RAYON_OK;
}

rayon_status rayon_server_cancel(rayon_server* s, uint64_t id) {
   std::lock_guard<std::mutex> lk(s->mu);
   auto it = s->jobs.find(id);
   if (it == s->jobs.end()) {
      ry::set_error("rayon_server_cancel: unknown job");
      return RAYON_ERR_INVALID;
   }
   it->second->cancel = true;
   ++s->wake_seq;
   s->cv.notify_one();
   return RAYON_OK;
}

rayon_status rayon_server_release(rayon_server* s, uint64_t id) {
   std::lock_guard<std::mutex> lk(s->mu);
   auto it = s->jobs.find(id);
   if (it == s->jobs.end()) {
      ry::set_error("rayon_server_release: unknown job");
      return RAYON_ERR_INVALID;
   }
   SJob* j = it->second;
   j->released = true;
   if (j->state == RAYON_JOB_DONE || j->state == RAYON_JOB_FAILED || j->state == RAYON_JOB_CANCELLED) {
      (void)hipSetDevice(s->ctx->device);
      s->free_job_locked(j);
      ++s->wake_seq;  // budget freed: queued jobs may fit now
      s->cv.notify_one();
   }
   return RAYON_OK;
}

rayon_status rayon_server_drain(rayon_server* s, double timeout_s) {
   std::unique_lock<std::mutex> lk(s->mu);
   const bool ok = s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
      return s->queue.empty() && s->active.empty();
   });
   if (!ok) {
      ry::set_error("rayon_server_drain: timed out with %zu queued and %zu active jobs", s->queue.size(),
                    s->active.size());
      return RAYON_ERR_INVALID;
   }
   return RAYON_OK;
}

rayon_status rayon_server_stats_get(rayon_server* s, rayon_server_stats* out) {
   std::lock_guard<std::mutex> lk(s->mu);
   out->passes = s->passes;
   out->first_passes = s->first_passes;
   out->first_kept = s->first_kept;
   out->first_kept_paths = s->first_kept_paths;
   out->completed = s->completed;
   out->failed = s->failed;
   out->queued = (int64_t)s->queue.size();
   out->active = (int64_t)s->active.size();
   out->segments = s->total_segments;
   out->used_bytes = s->used_bytes;
   out->busy_s = s->busy_s;
   out->wall_s = now_s() - s->t_start;
   out->pass_samples = s->pass_samples;
   out->pass_jobs = s->pass_jobs;
   out->pass_ms = s->pass_ms_sum;
   out->first_ms = s->first_ms_sum;
   out->check_ms = s->check_ms_sum;
   out->pass_paths = s->pass_paths_sum;
   return RAYON_OK;
}

rayon_status rayon_server_destroy(rayon_server* s) {
   if (!s) return RAYON_OK;
   {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
      s->cv.notify_all();
   }
   if (s->th.joinable()) s->th.join();
   (void)hipSetDevice(s->ctx->device);
   (void)rayon_ctx_sync(s->ctx);
   for (auto& kv : s->jobs) {
      s->free_job_locked(kv.second);
      delete kv.second;
   }
   (void)rayon_ctx_sync(s->first_ctx);
   for (hipStream_t ps : s->pv_stream) (void)hipStreamSynchronize(ps);
   delete s;
   return RAYON_OK;
}

}  // extern "C"
