// wavefront.hip — batched wavefront renderer: many jobs advance together through
// generate -> (extend -> shade -> shadow) x depth -> accumulate over shared device queues.
//
// This is the MI355X replacement for running renderAccKernel once per job
// (shaders/render_acc_kernel.cu:128-282): the per-thread megakernel loop is split into
// stages so that each stage keeps few registers (traversal needs the LDS stack, shading
// does not), waves stay full as paths terminate, and jobs of different sizes share every
// launch.
//
// Work of one pass = for every job, `samples` samples of every pixel, enumerated
// job-major, then sample, then pixel. It is processed in rounds of at most `capacity`
// paths; path slot i of a round holds work item w0 + i, and its radiance is summed at
// index i. The live paths of a bounce level are packed densely: the shade stage writes
// each continuing path's state, with its slot, to the next position of the other of two
// state sets, and each shadow ray to the next position of the shadow arrays, so every
// stage reads its rays' state contiguously however few paths remain (indexing by slot
// leaves deep levels reading one live record per cache line). Positions come from
// per-wave reservations (one atomic per reservation); positions a wave reserved and did
// not fill are holes, marked so the next stage skips them. The extend/shade/shadow
// kernels are persistent: a fixed grid of workgroups whose waves pull 64 positions at a
// time until the level (whose length the previous stage left in device memory) is
// drained, so a whole pass needs no host synchronisation and can be captured once into a
// hipGraph and replayed.
//
// Determinism: a path's random numbers depend only on (job seed, pixel, sample); its
// radiance is accumulated into the pixel by the accumulate stage in sample order. A job
// therefore produces the same image alone or batched with any other jobs.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "pathtrace.h"
#include "wavefront.h"

namespace ry {

constexpr int kWfBlock = 256;
// Binary trees up to 23 levels deep take a 23-entry stack (22 LDS entries per lane + the
// register top), which fits 7 workgroups of the trace kernel per CU instead of 5.
constexpr int kShortStack = 23;
// LDS words per lane of the quantized-tree trace kernel for trees at most kShortWide / 2
// levels deep (two words per partly visited node): 20 KiB per workgroup, so LDS allows 8
// workgroups per CU and registers 6 waves per SIMD, against 5 with the full 32-word stack.
constexpr int kShortWide = 20;
constexpr int kSdfSlice = 8;   // steps of the first marching slice; slice k takes kSdfSlice << k
constexpr int kSdfSlices = 8;  // slices per level at most (the last one runs to the end)
constexpr long long kMegaSceneBytes = 8ll << 20;  // automatic path choice: megakernel up to this scene size
constexpr int kCtl = 8;  // control words per level
// Path flags, kept in the w word of the throughput record
constexpr uint32_t kPrevLambert = 2u;  // the last scatter was Lambertian
constexpr uint32_t kHole = 0xFFFFFFFFu;  // job word of an unused position

struct WfView {
   // Path state of the rays of a level, at dense positions; set (level & 1). The job and
   // the flags travel in the otherwise unused w words.
   float4* o[2];    // origin, rng state (bits)
   float4* d[2];    // direction, job (bits; kHole at an unused position)
   float4* thr[2];  // throughput, path flags (bits)
   int* ps[2];      // path slot (work item index in the round), -1 at an unused position
   float4* rad;     // radiance, by path slot
   float4* hit;     // closest hit of the level's ray at each position: t, primitive slot (bits), u, v
   float4* so;      // shadow rays of a level at dense positions: origin, path slot (bits; -1 unused)
   float4* sd;      // direction, job (bits)
   float4* sc;      // contribution if unoccluded
   unsigned* ctl;           // per level (stride kCtl): positions, extend head, shade head,
                            // shadow positions, shadow head
   int* qs0;                // SDF marching queues (positions of the level), alternating by slice
   int* qs1;
   unsigned* sctl;          // per level and slice (stride 2): marching queue count, head
   const JobView* jobs;
   const long long* base;   // work prefix per job (njobs + 1 entries)
   unsigned long long* stats;  // traversal statistics (optional)
   int njobs;
   int levels;
   int chunk;  // 64-entry batches a wave claims per queue-head atomic (trace, marching)
   int shade_chunk;  // the same for the shade stage
   int shadow_refill;  // shadow rays with lane refill (no shadow ray may cross an SDF field)
   int dense;          // no job of the pass is adaptive: every level-0 position holds a path
                       // whose slot is the position, throughput 1 and flags 0 (not stored)
   int refill;         // idle lanes (of 64) at which a trace wave takes new rays
   unsigned res_most;  // largest output reservation of a shade wave (multiple of 64)
   unsigned leaf_exit;  // a trace wave stops descending when at most
   unsigned leaf_div;   // min(leaf_exit, lanes holding rays / leaf_div) lanes have no leaf
};

struct Wave {
   int capacity = 0;
   int levels = 0;
   int max_jobs = 0;
   int round = 0;  // path slots of one round of the current pass (<= capacity)
   std::vector<DevBuf> bufs, bufs_slots, bufs_ctl, bufs_tables;
   WfView v{};
   JobView* jobs_dev = nullptr;
   long long* base_dev = nullptr;
   unsigned long long* stats_dev = nullptr;
   // tables of the jobs that take the one-launch megakernel (k_multi)
   JobView* mjobs_dev = nullptr;
   int* mtile_dev = nullptr;
   int* mtodo_dev = nullptr;
   // options
   int opt_capacity = 1 << 27;
   int opt_wg_per_cu = 0;  // 0: as many as fit (occupancy of each stage kernel)
   int opt_chunk = 0;  // 0: kChunkBin or kChunkWide by the pass's tree layout
   int opt_shade_chunk = 16;
   int opt_refill = 0;      // 0: kRefillBin or kRefillWide by the pass's tree layout
   int opt_leaf_exit = -1;  // -1: kLeafExitBin or kLeafExitWide
   int occ_trace = 0, occ_trace_wide = 0, occ_trace_short = 0, occ_trace_wide_short = 0, occ_shade = 0, occ_shade_lean = 0, occ_sdf = 0, occ_sdf_rest = 0;
   struct Pending {
      int stage;
      hipEvent_t e0, e1;
   };
   std::vector<Pending> pending;  // stage events not yet read back
   int opt_graph = 1;
   int opt_timing = 0;
   int opt_count = 0;
   int cus = 0;
   // timing
   double stage_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
   long long launches[8] = {0, 0, 0, 0, 0, 0, 0, 0};
   long long rounds = 0;
   // graph cache
   struct GraphEntry {
      hipGraphExec_t exec = nullptr;
      hipGraph_t graph = nullptr;
   };
   std::map<std::vector<long long>, GraphEntry> graphs;
   std::vector<JobView> host_jobs;
   std::vector<long long> host_base;
   std::vector<JobView> host_mjobs;
   std::vector<int> host_mtile, host_mtodo;
};

// ---------------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Position of this lane's entry in a queue, or -1; one atomic per wave.
__device__ __forceinline__ int wave_push(unsigned* counter, bool pred) {
   const unsigned long long m = __ballot(pred);
   if (m == 0) return -1;
   const int leader = __ffsll((long long)m) - 1;
   unsigned base = 0;
   if (lane_id() == leader) base = atomicAdd(counter, (unsigned)__popcll(m));
   base = __shfl(base, leader);
   const unsigned long long below = m & ((1ull << lane_id()) - 1ull);
   return pred ? (int)(base + (unsigned)__popcll(below)) : -1;
}

constexpr int kMarchResMax = 512;  // marching-queue positions a wave reserves at a time, at most
// Positions beyond the marches themselves that a marching queue must hold: every wave of a
// trace or marching launch may leave at most kMarchResMax - 1 reserved positions unfilled
// (at most 32 workgroups of 4 waves per CU).
constexpr long long kMarchHeadroomPerCu = 32ll * (kWfBlock / 64) * kMarchResMax;
// Positions beyond a level's paths that its arrays hold: every wave of a shade launch may
// leave its last reservation partly unfilled, so the reservations are capped at this
// headroom divided by the launch's waves (1024 for 16 waves per CU, 128 at 32 workgroups
// of 4 waves per CU).
constexpr long long kPosHeadroomPerCu = 16ll * 1024;

// Reservation size for a wave's marching-queue pushes in a launch over `count` rays: half
// of a wave's share (not every ray marches), in multiples of 64 from 64 to kMarchResMax.
// Atomics on the shared head serialise, so a busy launch takes large reservations; the
// unfilled tails stay a few percent of the queue.
__device__ __forceinline__ unsigned march_res(unsigned count) {
   const unsigned waves = gridDim.x * (blockDim.x >> 6);
   const unsigned per = count / (128u * waves);
   return min(64u * max(per, 1u), (unsigned)kMarchResMax);
}

// Appends to a marching queue through positions the wave reserves `res` at a time (a
// multiple of 64, wave-uniform), so the queue's head counter, shared by every wave of a
// launch, takes one atomic per reservation instead of one per push. close() marks the
// positions the wave reserved and did not fill with -1; the consumer skips them.
struct QueueRes {
   unsigned res = 64;
   unsigned next = 0, end = 0;  // wave-uniform
   __device__ __forceinline__ void push(unsigned* head, int* q, bool pred, int value) {
      const unsigned long long m = __ballot(pred);
      if (m == 0) return;
      const unsigned k = (unsigned)__popcll(m);
      const unsigned rank = (unsigned)__popcll(m & ((1ull << lane_id()) - 1ull));
      const unsigned have = end - next;
      if (pred && rank < have) q[next + rank] = value;
      if (k > have) {
         unsigned base = 0;
         if (lane_id() == 0) base = atomicAdd(head, res);
         base = __shfl(base, 0);
         if (pred && rank >= have) q[base + (rank - have)] = value;
         next = base + (k - have);
         end = base + res;
      } else {
         next += k;
      }
   }
   __device__ __forceinline__ void close(int* q) {
      for (unsigned i = next + (unsigned)lane_id(); i < end; i += 64) q[i] = -1;
   }
};

// Positions for a wave's outputs (continuing paths, shadow rays) from reservations of
// `res` positions (a multiple of 64, wave-uniform) taken with one atomic on the level's
// shared counter; what a wave reserved and did not fill it marks unused when it ends.
struct PosRes {
   unsigned next = 0, end = 0;  // wave-uniform
   // Position of each lane with `pred` set, else -1.
   __device__ __forceinline__ int take(unsigned* head, bool pred, unsigned res) {
      const unsigned long long m = __ballot(pred);
      if (m == 0) return -1;
      const unsigned k = (unsigned)__popcll(m);
      const unsigned rank = (unsigned)__popcll(m & ((1ull << lane_id()) - 1ull));
      const unsigned have = end - next;
      int pos = (pred && rank < have) ? (int)(next + rank) : -1;
      if (k > have) {
         unsigned base = 0;
         if (lane_id() == 0) base = atomicAdd(head, res);
         base = __shfl(base, 0);
         if (pred && rank >= have) pos = (int)(base + (rank - have));
         next = base + (k - have);
         end = base + res;
      } else {
         next += k;
      }
      return pos;
   }
};

// Reservation size for the outputs of a level of `count` positions: what a wave shades, in
// multiples of 64, at most `most`. Atomics on one address serialise (about 13 ns each), so
// small reservations made the counter the bottleneck of busy levels; the holes they leave
// come as runs at the ends of the waves' last reservations, which the next stages skip a
// batch at a time.
__device__ __forceinline__ unsigned pos_res(unsigned count, unsigned most) {
   const unsigned waves = gridDim.x * (blockDim.x >> 6);
   const unsigned per = count / (64u * waves);
   return min(64u * max(per, 1u), most);
}

// Persistent-worker loop: each wave claims `chunk` batches of 64 queue entries with one
// atomic on the queue head and runs body(index) on each; body sees indices >= count on
// the last, partial batch and must ignore them. Returns when the queue is drained.
// All the waves of a launch contend for the one head counter, so a wave does not claim
// when it can tell the queue is already drained: an empty queue, a head already past the
// end (a plain read, which can only lag behind the true head), or a claim of its own that
// reached the end.
// Batches of 64 queue entries a wave claims at once: `chunk`, or fewer when the queue is too
// short to give every wave of the launch that many, so a short queue (a pass's deep levels,
// a small pass) spreads over all waves instead of lining up behind a few of them.
__device__ __forceinline__ int fit_chunk(unsigned count, int chunk) {
   const unsigned waves = gridDim.x * (blockDim.x >> 6);
   const unsigned per = (count + 64u * waves - 1u) / (64u * waves);
   return (int)min((unsigned)chunk, max(per, 1u));
}

template <class F>
__device__ __forceinline__ void drain(unsigned* head, unsigned count, int chunk, F&& body) {
   if (count == 0) return;
   chunk = fit_chunk(count, chunk);
   unsigned seen = 0;
   if (lane_id() == 0) seen = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   if (__shfl(seen, 0) >= count) return;
   const unsigned step = 64u * (unsigned)chunk;
   for (;;) {
      unsigned base = 0;
      if (lane_id() == 0) base = atomicAdd(head, step);
      base = __shfl(base, 0);
      if (base >= count) return;
      for (int k = 0; k < chunk; ++k) {
         const unsigned b = base + 64u * (unsigned)k;
         if (b >= count) break;
         body(b + (unsigned)lane_id());
      }
      if (base + step >= count) return;
   }
}

// (sample, pixel) of work item l of a job; 32-bit division whenever it fits.
__device__ __forceinline__ void split_work(long long l, unsigned npix, int& s, int& p) {
   if (l < (1ll << 32)) {
      const unsigned l32 = (unsigned)l, q = l32 / npix;
      s = (int)q;
      p = (int)(l32 - q * npix);
   } else {
      const long long q = l / (long long)npix;
      s = (int)q;
      p = (int)(l - q * (long long)npix);
   }
}

__device__ __forceinline__ int find_job(const WfView& V, long long w) {
   int lo = 0, hi = V.njobs - 1;
   while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (V.base[mid] <= w)
         lo = mid;
      else
         hi = mid - 1;
   }
   return lo;
}

// ---------------------------------------------------------------------------------
// stages
// ---------------------------------------------------------------------------------
// A path's level-0 position is its slot. A converged pixel (possible only for adaptive jobs,
// and rare: their passes cover the pixels still sampling when the pass starts) leaves a hole
// there, which the trace and shade stages skip; packing the live paths instead took one
// atomic per wave on the level's counter, which serialised (5.7 ms per serving pass).
// dense: no job of the pass can skip a pixel (no adaptive test), so no hole is written.
__global__ __launch_bounds__(kWfBlock) void k_wf_generate(WfView V, long long w0, int n, int dense) {
   const int i = blockIdx.x * kWfBlock + threadIdx.x;
   bool live = false;
   float4 o4, d4;
   if (i < n) {
      const long long w = w0 + i;
      const int j = find_job(V, w);
      const JobView& J = V.jobs[j];
      const unsigned units = J.nlive >= 0 ? (unsigned)J.nlive : (unsigned)J.width * (unsigned)J.height;
      int s, p;
      split_work(w - V.base[j], units, s, p);
      if (J.nlive >= 0) p = J.live[p];
      const int sample = J.sample_base + s;
      if (J.counts[p] >= 0) {  // a converged pixel takes no path (the accumulate stage skips it)
         Rng R{counter_init(J.seed_lo, J.seed_hi, (uint32_t)p, (uint32_t)sample), false};
         V3 o, d;
         camera_ray(J, p % J.width, p / J.width, R, o, d);
         o4 = make_float4(o.x, o.y, o.z, __uint_as_float(R.s));
         d4 = make_float4(d.x, d.y, d.z, __uint_as_float((uint32_t)j));
         V.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
         live = true;
      }
   }
   if (i == 0) V.ctl[0] = (unsigned)n;
   if (live) {
      V.o[0][i] = o4;
      V.d[0][i] = d4;
      if (!dense) {  // a dense pass's level 0 implies them (k_wf_shade)
         V.thr[0][i] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
         V.ps[0][i] = i;
      }
   } else if (!dense && i < n) {
      V.d[0][i] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kHole));
      V.ps[0][i] = -1;
   }
}

// Each marching wave keeps one field of at most kSdfMaskMax primitives in its own LDS slot
// (records, and for the set-up the bounding spheres): every step of every march reads a
// record per primitive near the ray, and from LDS those reads no longer wait on the cache
// hierarchy. A wave's rays nearly always share a field (queues keep work order); when the
// first lane with a field has another one than the slot holds, the wave reloads the slot
// (one record per lane) before any lane marches. Lanes of other fields read the scene.
struct FieldSlot {
   SdfRec* rec;
   float4* sph;             // null: no spheres kept
   const SdfRec* held = nullptr;  // the field in the slot (the same in every lane)
};

// True when this lane's field (S may be null for an idle lane) is the one in the slot.
__device__ __forceinline__ bool hold_field(FieldSlot& f, const SceneView* S) {
   const SdfRec* mine = (S != nullptr && S->num_sdf > 0 && S->num_sdf <= kSdfMaskMax) ? S->sdf : nullptr;
   const unsigned long long has = __ballot(mine != nullptr);
   if (has == 0) return false;
   const int first = __ffsll((long long)has) - 1;
   const SdfRec* want = (const SdfRec*)__shfl((unsigned long long)(uintptr_t)mine, first);
   if (want != f.held) {
      const SceneView* Sf = (const SceneView*)__shfl((unsigned long long)(uintptr_t)S, first);
      const int n = Sf->num_sdf, i = lane_id();
      if (i < n) {
         f.rec[i] = gload(want + i);
         if (f.sph) f.sph[i] = gload(Sf->sdf_sph + i);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      f.held = want;
   }
   return mine != nullptr && mine == f.held;
}

// A march of `steps` more steps, with the records from the wave's slot when `held`.
__device__ __forceinline__ int march_held(const FieldSlot& f, bool held, const SceneView& S, V3 o, V3 d,
                                          SdfMarch& m, int steps, float& t) {
   if (held)
      return sdf_march_run(S, o, d, m, steps, t, SdfLdsRecs{(const __attribute__((address_space(3))) gvec4*)f.rec});
   return sdf_march_run(S, o, d, m, steps, t);
}

// SDF marching stage of extension rays: slice `it` takes up to kSdfSlice << it steps of
// every ray in its queue (the last slice runs to the end); rays still marching go to the next
// slice's queue, so each slice's waves hold only live rays. A hit replaces the ray's hit
// record (the march ends at the closest other hit, and an exact tie goes to the field,
// which has the largest original index). Queues hold the rays' positions in the level; the
// march state lives in the shadow-ray arrays at the same position, free between the trace
// stage (which consumed the previous level's shadow rays) and the shade stage.
template <bool CNT>
__global__ __launch_bounds__(kWfBlock) void k_wf_sdf(WfView V, int level, int it, int last) {
   __shared__ SdfRec lrec[kWfBlock / 64][kSdfMaskMax];
   __shared__ float4 lsph[kWfBlock / 64][kSdfMaskMax];  // bounding spheres, for the set-up
   FieldSlot fs{lrec[threadIdx.x / 64], it == 0 ? lsph[threadIdx.x / 64] : nullptr};
   unsigned long long n_begin = 0, n_steps = 0, n_evals = 0, n_slots = 0;
   unsigned* cin = V.sctl + 2 * ((kSdfSlices + 1) * level + it);
   unsigned* cout = cin + 2;
   const unsigned count = cin[0];
   const int* qin = (it & 1) ? V.qs1 : V.qs0;
   int* qout = (it & 1) ? V.qs0 : V.qs1;
   QueueRes res;  // rays still marching after this slice
   res.res = march_res(count);
   drain(&cin[1], count, V.chunk, [&](unsigned idx) {
      int slot = -1;
      bool more = false;
      unsigned taken = 0;
      if (idx < count) slot = qin[idx];  // -1: a position a trace wave reserved and left unused
      float4 d4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const SceneView* Sp = nullptr;
      if (slot >= 0) {
         d4 = V.d[level & 1][slot];
         Sp = &V.jobs[__float_as_uint(d4.w)].scene;
      }
      const bool held = hold_field(fs, Sp);
      if (slot >= 0) {
         const SceneView& S = *Sp;
         const V3 o = mk(V.o[level & 1][slot]), d = mk(d4);
         SdfMarch m;
         bool live = true;
         if (it == 0) {
            const PrimRec fp = gload(S.prims + S.sdf_prim);
            if (held)
               live = sdf_march_begin(S, fp, o, d, 0.001f, V.hit[slot].x, m,
                                      SdfLdsSph{(const __attribute__((address_space(3))) gvec4*)fs.sph});
            else
               live = sdf_march_begin(S, fp, o, d, 0.001f, V.hit[slot].x, m);
         } else {
            const float4 a = V.so[slot], b = V.sc[slot];
            m.tw = a.x;
            m.tend = a.y;
            m.cap = a.z;
            m.used = __float_as_int(a.w);
            m.mask = (uint64_t)__float_as_uint(b.x) | ((uint64_t)__float_as_uint(b.y) << 32);
         }
         if (live) {
            float t = 0.0f;
            const int used0 = m.used;
            const int r = march_held(fs, held, S, o, d, m, last ? S.sdf_steps : (kSdfSlice << it), t);
            if constexpr (CNT) {
               // steps evaluated: the loop's count, plus the step that returned from inside it
               taken = (unsigned)(m.used - used0 + (r >= 0 && m.used < S.sdf_steps ? 1 : 0));
               n_begin += it == 0;
               n_steps += taken;
               n_evals += (unsigned long long)taken * (S.num_sdf > kSdfMaskMax ? S.num_sdf : __popcll(m.mask));
            }
            if (r == 1) {
               V.hit[slot] = make_float4(t, __int_as_float(S.sdf_prim), 0.0f, 0.0f);
            } else if (r < 0) {
               more = true;
               V.so[slot] = make_float4(m.tw, m.tend, m.cap, __int_as_float(m.used));
               V.sc[slot] = make_float4(__uint_as_float((uint32_t)m.mask), __uint_as_float((uint32_t)(m.mask >> 32)),
                                        0.0f, 0.0f);
            }
         }
      }
      if constexpr (CNT) n_slots += 64ull * wave_max_u32(taken);  // lane-steps the wave's SIMD spent
      res.push(&cout[0], qout, more, slot);
   });
   res.close(qout);
   if constexpr (CNT) {
      const unsigned long long a = wave_sum(n_begin), b = wave_sum(n_steps), c = wave_sum(n_evals);
      if (lane_id() == 0) {
         atomicAdd(&V.stats[8], a);
         atomicAdd(&V.stats[9], b);
         atomicAdd(&V.stats[10], c);
         atomicAdd(&V.stats[11], n_slots);
      }
   }
}

// ---- Binary-tree traversal with lane refill (dynamic ray fetch, after Aila and Laine,
// "Understanding the efficiency of ray traversal on GPUs", HPG 2009). A wave keeps a pool of
// queue entries it has claimed; whenever at least V.refill of its lanes have no ray, those
// lanes take the next entries of the pool (claiming another chunk when it runs out), so a
// wave is not held by the slowest of 64 rays. Each lane visits the nodes closest_hit_bin /
// any_hit_bin visit, with leaves postponed until nearly every lane of the wave holds one
// (leaf_limit), and tests one leaf per iteration with the same tie rule, so every ray's
// result is what the run-to-completion traversal returns.

struct WavePool {
   unsigned next = 0, end = 0;  // wave-uniform: entries claimed and not yet handed out
   bool drained = false;        // wave-uniform: the queue has no unclaimed entries left

   // Hands entries to the idle lanes (in lane order) when at least `at` lanes are idle;
   // returns the lane's new queue index, or -1.
   __device__ __forceinline__ long long refill(unsigned* head, unsigned count, unsigned step, bool idle_lane,
                                               int at) {
      const unsigned long long idle = __ballot(idle_lane);
      const int nidle = __popcll(idle);
      if (nidle < at) return -1;
      if (next == end && !drained) {
         unsigned base = 0;
         if (lane_id() == 0) base = atomicAdd(head, step);
         base = __shfl(base, 0);
         if (base >= count) {
            drained = true;
         } else {
            next = base;
            end = min(base + step, count);
            drained = end == count;  // no entries beyond this chunk
         }
      }
      const unsigned avail = end - next;
      const unsigned rank = (unsigned)__popcll(idle & ((1ull << lane_id()) - 1ull));
      const long long mine = (idle_lane && rank < avail) ? (long long)(next + rank) : -1;
      next += min(avail, (unsigned)nidle);
      return mine;
   }
};

// One wave's state at the start of a drain: an empty queue or one already drained by
// other waves hands out nothing.
__device__ __forceinline__ WavePool start_pool(unsigned* head, unsigned count) {
   WavePool p;
   unsigned seen = count;
   if (count != 0 && lane_id() == 0) seen = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   p.drained = __shfl(seen, 0) >= count;
   return p;
}

// SDF marching after the first slice: one persistent launch per level marches every ray
// that slice 0 left unfinished to its end, from the state slice 0 saved, in bursts of
// kSdfBurst steps; lanes whose march ended take the next rays of the queue (WavePool), so
// a few long grazing marches do not hold whole waves, and no further slices are launched.
constexpr int kSdfBurst = 4;

template <bool CNT>
__global__ __launch_bounds__(kWfBlock) void k_wf_sdf_rest(WfView V, int level) {
   __shared__ SdfRec lrec[kWfBlock / 64][kSdfMaskMax];
   FieldSlot fs{lrec[threadIdx.x / 64], nullptr};
   unsigned long long n_steps = 0, n_evals = 0, n_slots = 0;
   unsigned* cin = V.sctl + 2 * ((kSdfSlices + 1) * level + 1);
   const unsigned count = cin[0];
   const unsigned step = 64u * (unsigned)fit_chunk(count, V.chunk);
   WavePool pool = start_pool(&cin[1], count);
   int slot = -1;
   const SceneView* S = nullptr;
   V3 o, d;
   SdfMarch m;
   for (;;) {
      if (!pool.drained || pool.next != pool.end) {
         const long long idx = pool.refill(&cin[1], count, step, slot < 0, V.refill);
         const int s = idx >= 0 ? V.qs1[idx] : -1;  // -1 also marks unused reserved positions
         if (s >= 0) {
            slot = s;
            const float4 d4 = V.d[level & 1][slot];
            S = &V.jobs[__float_as_uint(d4.w)].scene;
            o = mk(V.o[level & 1][slot]);
            d = mk(d4);
            const float4 a = V.so[slot], b = V.sc[slot];
            m.tw = a.x;
            m.tend = a.y;
            m.cap = a.z;
            m.used = __float_as_int(a.w);
            m.mask = (uint64_t)__float_as_uint(b.x) | ((uint64_t)__float_as_uint(b.y) << 32);
         }
      }
      if (__ballot(slot >= 0) == 0) {
         if (pool.drained && pool.next == pool.end) break;
         continue;
      }
      unsigned taken = 0;
      const bool held = hold_field(fs, slot >= 0 ? S : nullptr);
      if (slot >= 0) {
         float t = 0.0f;
         const int used0 = m.used;
         const int r = march_held(fs, held, *S, o, d, m, kSdfBurst, t);
         if constexpr (CNT) {
            taken = (unsigned)(m.used - used0 + (r >= 0 && m.used < S->sdf_steps ? 1 : 0));
            n_steps += taken;
            n_evals += (unsigned long long)taken * (S->num_sdf > kSdfMaskMax ? S->num_sdf : __popcll(m.mask));
         }
         if (r == 1) V.hit[slot] = make_float4(t, __int_as_float(S->sdf_prim), 0.0f, 0.0f);
         if (r >= 0) slot = -1;
      }
      if constexpr (CNT) n_slots += 64ull * wave_max_u32(taken);
   }
   if constexpr (CNT) {
      const unsigned long long b = wave_sum(n_steps), c = wave_sum(n_evals);
      if (lane_id() == 0) {
         atomicAdd(&V.stats[9], b);
         atomicAdd(&V.stats[10], c);
         atomicAdd(&V.stats[11], n_slots);
      }
   }
}

// When a trace wave stops descending to test the leaves its lanes hold (postponed leaves):
// once no more than min(leaf_exit, lanes holding rays / leaf_div) of its lanes are still
// looking for their first leaf. Waiting for the last lane (a limit of 0) leaves two thirds
// of the lanes idle per node step on the 1.31M-triangle mesh; the lanes still looking go on
// after the leaf tests. Measured defaults per tree layout, with the refill threshold that
// suits each (a wave that leaves the node loop more often refills better in larger groups).
constexpr unsigned kLeafExitBin = 12, kLeafDivBin = 4, kLeafExitWide = 16, kLeafDivWide = 3;
constexpr int kRefillBin = 32, kRefillWide = 16;
// 64-entry batches a trace wave claims per queue atomic: a quantized-tree wave runs through
// its rays faster (2.6 node visits per primary ray against 10 pair nodes), so it claims more
// at once (chunk 32: +3 % on the 1080p mesh frame; binary trees gain nothing beyond 8).
constexpr int kChunkBin = 8, kChunkWide = 32;

__device__ __forceinline__ unsigned leaf_limit(const WfView& V) {
   const unsigned lanes = (unsigned)__popcll(__ballot(1));  // the lanes holding rays
   return min(V.leaf_exit, lanes / V.leaf_div);
}

// Extension rays of one level: closest hit (SDF fields deferred to k_wf_sdf).
template <bool CNT, int STACK>
__device__ void ext_rays_bin(const WfView& V, int level, int* lds, TravCount& tc) {
   unsigned* ctl = V.ctl + kCtl * level;
   const unsigned count = ctl[0];
   const float4* O = V.o[level & 1];
   const float4* D = V.d[level & 1];
   unsigned* mctl = V.sctl + 2 * (kSdfSlices + 1) * level;
   const unsigned step = 64u * (unsigned)fit_chunk(count, V.chunk);
   WavePool pool = start_pool(&ctl[1], count);
   int slot = -1, cur = kSentinel;  // slot: the lane's ray's position in the level
   // slot of a finished ray whose SDF march is still to be queued. The lanes' marches are
   // queued together before each refill, into positions the wave reserved in the marching
   // queue march_res(count) at a time (one atomic per reservation, not per refill): the queue head
   // is shared by every wave of the launch. Positions reserved but not filled by the wave's
   // end hold -1, which k_wf_sdf skips.
   int pend = -1;
   QueueRes res;
   res.res = march_res(count);
   auto queue_marches = [&]() {
      res.push(&mctl[0], V.qs0, pend >= 0, pend);
      pend = -1;
   };
   const Node2* nodes = nullptr;
   const PrimRec* prims = nullptr;
   RayInv r;
   HitMin best;
   BinStack<kWfBlock, STACK> st(lds);
   for (;;) {
      if (!pool.drained || pool.next != pool.end) {
         if (__popcll(__ballot(slot < 0)) >= V.refill) queue_marches();
         const long long idx = pool.refill(&ctl[1], count, step, slot < 0, V.refill);
         // origin and direction in one round trip (a hole's origin is read and ignored)
         float4 d4 = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kHole)), o4 = d4;
         if (idx >= 0) {
            d4 = D[idx];
            o4 = O[idx];
         }
         if (__float_as_uint(d4.w) != kHole) {
            slot = (int)idx;
            const SceneView& S = V.jobs[__float_as_uint(d4.w)].scene;
            nodes = S.nodes;
            prims = S.prims;
            cur = S.root;
            r = make_rayinv(mk(o4), mk(d4));
            best.t = 3.402823466e38f;
            best.slot = -1;
            best.orig = -1;
            best.u = best.v = 0.0f;
            best.sdf = -1;
            st.n = 0;
            st.top = kSentinel;
         }
      }
      if (__ballot(slot >= 0) == 0) {  // the entries handed out may all have been holes
         if (pool.drained && pool.next == pool.end) break;
         continue;
      }
      if (slot >= 0) {
         bool finished;
         // Internal nodes until nearly every lane of the wave holds a leaf or has finished
         // (leaf_limit): the first leaf a lane meets is set aside and the lane goes on with
         // its next node (postponed leaves, Aila and Laine 2009), so lanes that reach a leaf
         // early keep working while the others descend. The closest hit does not depend on
         // the order of the primitive tests (equal distances go to the larger original index).
         int leaf = kSentinel;
         const unsigned lim = leaf_limit(V);
         for (;;) {
            if (cur >= 0) {
               const Node2 n = gload(nodes + cur);
               if constexpr (CNT) ++tc.nodes;
               float tn0, tn1;
               bool h0, h1;
               box2(r, n, 0.001f, best.t, tn0, tn1, h0, h1);
               const int r0 = n.meta.z ? pack_leaf(n.meta.x, n.meta.z) : n.meta.x;
               const int r1 = n.meta.w ? pack_leaf(n.meta.y, n.meta.w) : n.meta.y;
               if (h0 && h1) {
                  const bool swap = tn1 < tn0;
                  st.push(swap ? r0 : r1);
                  cur = swap ? r1 : r0;
               } else if (h0) {
                  cur = r0;
               } else if (h1) {
                  cur = r1;
               } else {
                  cur = st.pop();
               }
            } else if (cur != kSentinel && leaf == kSentinel) {
               leaf = cur;
               cur = st.pop();
            }
            if (__popcll(__ballot(leaf == kSentinel && cur != kSentinel)) <= lim) break;
         }
         if (leaf != kSentinel) {  // one leaf
            const int packed = ~leaf;
            const int first = packed >> 3, cnt = (packed & 7) + 1;
            for (int i = first; i < first + cnt; ++i) {
               const PrimRec p = gload(prims + i);
               if constexpr (CNT) ++tc.prims;
               const uint32_t kind = info_kind(__float_as_uint(p.a.w));
               float t = 0.0f, u = 0.0f, v = 0.0f;
               bool hit = false;
               if (kind == K_TRI)
                  hit = isect_tri(p, r.o, r.d, 0.001f, best.t, t, u, v);
               else if (kind == K_SPHERE || kind == K_DISPLACED)
                  hit = isect_sphere(mk(p.a), p.b.x, r.o, r.d, 0.001f, best.t, t);
               else if (kind == K_RECT)
                  hit = isect_rect(mk(p.a), mk(p.b), mk(p.c), r.o, r.d, 0.001f, best.t, t);
               else
                  best.sdf = i;  // marched later by k_wf_sdf
               if (hit) {
                  const int orig = __float_as_int(p.b.w);
                  if (t < best.t || orig > best.orig) {
                     best.t = t;
                     best.slot = i;
                     best.orig = orig;
                     best.u = u;
                     best.v = v;
                  }
               }
            }
         }
         finished = cur == kSentinel;
         if (finished) {
            V.hit[slot] = make_float4(best.t, __int_as_float(best.slot), best.u, best.v);
            if (best.sdf >= 0) pend = slot;  // an SDF field lies on the way
            slot = -1;
         }
      }
   }
   queue_marches();
   res.close(V.qs0);
}

// Shadow rays of one level: any non-emitter hit; unoccluded contributions are added to the
// path's radiance.
template <bool CNT, int STACK>
__device__ void shadow_rays_bin(const WfView& V, int level, int* lds, TravCount& tc) {
   unsigned* sctl = V.ctl + kCtl * level + 3;
   const unsigned count = sctl[0];
   const unsigned step = 64u * (unsigned)fit_chunk(count, V.chunk);
   WavePool pool = start_pool(&sctl[1], count);
   int slot = -1, cur = kSentinel;
   const SceneView* S = nullptr;  // read for SDF fields only
   const Node2* nodes = nullptr;  // the scene's arrays in registers, as in ext_rays_bin
   const PrimRec* prims = nullptr;
   RayInv r;
   BinStack<kWfBlock, STACK> st(lds);
   constexpr float tmin = 1e-4f, tmax = 1.0f - 1e-4f;
   for (;;) {
      if (!pool.drained || pool.next != pool.end) {
         const long long idx = pool.refill(&sctl[1], count, step, slot < 0, V.refill);
         float4 o4 = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1)), d4 = o4;  // one round trip
         if (idx >= 0) {
            o4 = V.so[idx];
            d4 = V.sd[idx];
         }
         if (__float_as_int(o4.w) >= 0) {
            slot = (int)idx;  // the shadow ray's position; o4.w holds its path's slot
            S = &V.jobs[__float_as_uint(d4.w)].scene;
            nodes = S->nodes;
            prims = S->prims;
            cur = S->root;
            r = make_rayinv(mk(o4), mk(d4));
            st.n = 0;
            st.top = kSentinel;
         }
      }
      if (__ballot(slot >= 0) == 0) {  // the entries handed out may all have been holes
         if (pool.drained && pool.next == pool.end) break;
         continue;
      }
      if (slot >= 0) {
         bool occluded = false;
         int leaf = kSentinel;  // postponed leaves, as for extension rays
         const unsigned lim = leaf_limit(V);
         for (;;) {
            if (cur >= 0) {
               const Node2 n = gload(nodes + cur);
               if constexpr (CNT) ++tc.nodes;
               float tn0, tn1;
               bool h0, h1;
               box2(r, n, tmin, tmax, tn0, tn1, h0, h1);
               const int r0 = n.meta.z ? pack_leaf(n.meta.x, n.meta.z) : n.meta.x;
               const int r1 = n.meta.w ? pack_leaf(n.meta.y, n.meta.w) : n.meta.y;
               if (h0 && h1) {
                  st.push(r1);
                  cur = r0;
               } else if (h0) {
                  cur = r0;
               } else if (h1) {
                  cur = r1;
               } else {
                  cur = st.pop();
               }
            } else if (cur != kSentinel && leaf == kSentinel) {
               leaf = cur;
               cur = st.pop();
            }
            if (__popcll(__ballot(leaf == kSentinel && cur != kSentinel)) <= lim) break;
         }
         if (leaf != kSentinel) {  // one leaf
            const int packed = ~leaf;
            const int first = packed >> 3, cnt = (packed & 7) + 1;
            for (int i = first; i < first + cnt && !occluded; ++i) {
               const PrimRec p = gload(prims + i);
               if constexpr (CNT) ++tc.prims;
               const uint32_t info = __float_as_uint(p.a.w);
               if (info & F_EMITTER) continue;
               const uint32_t kind = info_kind(info);
               float t, u, v;
               if (kind == K_TRI)
                  occluded = isect_tri(p, r.o, r.d, tmin, tmax, t, u, v);
               else if (kind == K_SPHERE || kind == K_DISPLACED)
                  occluded = isect_sphere(mk(p.a), p.b.x, r.o, r.d, tmin, tmax, t);
               else if (kind == K_RECT)
                  occluded = isect_rect(mk(p.a), mk(p.b), mk(p.c), r.o, r.d, tmin, tmax, t);
               else
                  occluded = isect_sdf(*S, p, r.o, r.d, tmin, tmax, t);
            }
            if (occluded) cur = kSentinel;
         }
         if (cur == kSentinel) {
            if (!occluded) {
               const int ps = __float_as_int(V.so[slot].w);
               const float4 c = V.sc[slot];
               const float4 rd = V.rad[ps];
               const V3 L = mk(rd) + mk(c);
               V.rad[ps] = make_float4(L.x, L.y, L.z, 0.0f);
            }
            slot = -1;
         }
      }
   }
}

// The 8-wide quantized tree with lane refill: one traversal step visits internal nodes
// until one with hit leaf children has had its primitives tested (or the traversal ends),
// then the wave refills idle lanes. The visiting order, the stack (two LDS words per
// partly visited node: its first internal child, the remaining hit mask with the node's
// internal-child mask) and the primitive tests are those of closest_hit_wide and
// any_hit_wide, so a ray gets the same hit as there.
constexpr uint32_t kWideDone = 0xFFFFFFFFu;

struct WideRay {
   V3 o, d, inv;
   uint32_t perm;
};

__device__ __forceinline__ WideRay make_wideray(V3 o, V3 d) {
   WideRay r;
   r.o = o;
   r.d = d;
   r.inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
   r.perm = 7u ^ ((d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u));
   return r;
}

// The next node after a node whose internal children `hit_int` (visiting order) were hit:
// the first of them (the rest pushed), else the nearest pending child of an ancestor, else
// kWideDone. STACK: LDS words per lane (trees are chosen no deeper than STACK / 2 levels).
template <int STACK>
__device__ __forceinline__ uint32_t wide_next(uint32_t hit_int, uint32_t base, uint32_t imask, uint32_t perm,
                                              int* stack, int& sp) {
   if (!hit_int) {
      if (sp == 0) return kWideDone;
      base = (uint32_t)stack[(2 * (sp - 1)) * kWfBlock];
      const uint32_t bits = (uint32_t)stack[(2 * (sp - 1) + 1) * kWfBlock];
      hit_int = bits & 0xFFu;
      const uint32_t im = bits >> 8;
      const int i = __ffs(hit_int) - 1;
      hit_int &= hit_int - 1u;
      const uint32_t sl = (uint32_t)i ^ perm;
      if (hit_int)
         stack[(2 * (sp - 1) + 1) * kWfBlock] = (int)(hit_int | (im << 8));
      else
         --sp;
      return base + (uint32_t)__popc(im & ((1u << sl) - 1u));
   }
   const int i = __ffs(hit_int) - 1;
   hit_int &= hit_int - 1u;
   const uint32_t sl = (uint32_t)i ^ perm;
   if (hit_int && sp < STACK / 2) {
      stack[(2 * sp) * kWfBlock] = (int)base;
      stack[(2 * sp + 1) * kWfBlock] = (int)(hit_int | (imask << 8));
      ++sp;
   }
   return base + (uint32_t)__popc(imask & ((1u << sl) - 1u));
}

constexpr uint32_t kNoHeld = ~0u;  // no node set aside (masks take 16 bits)

template <bool CNT, int STACK>
__device__ void ext_rays_wide(const WfView& V, int level, int* lds, TravCount& tc) {
   unsigned* ctl = V.ctl + kCtl * level;
   const unsigned count = ctl[0];
   const float4* O = V.o[level & 1];
   const float4* D = V.d[level & 1];
   unsigned* mctl = V.sctl + 2 * (kSdfSlices + 1) * level;
   const unsigned step = 64u * (unsigned)fit_chunk(count, V.chunk);
   WavePool pool = start_pool(&ctl[1], count);
   int slot = -1, pend = -1, sp = 0;  // slot: the lane's ray's position in the level
   uint32_t node = kWideDone;
   uint32_t held_masks = kNoHeld;  // child test of the node set aside for the lane's leaves
   QueueRes res;
   res.res = march_res(count);
   auto queue_marches = [&]() {  // as in ext_rays_bin
      res.push(&mctl[0], V.qs0, pend >= 0, pend);
      pend = -1;
   };
   const SceneView* S = nullptr;
   const WNode* wn = nullptr;  // the scene's arrays in registers, as in ext_rays_bin
   const PrimRec* prims = nullptr;
   WideRay r;
   HitMin best;
   for (;;) {
      if (!pool.drained || pool.next != pool.end) {
         if (__popcll(__ballot(slot < 0)) >= V.refill) queue_marches();
         const long long idx = pool.refill(&ctl[1], count, step, slot < 0, V.refill);
         float4 d4 = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kHole)), o4 = d4;  // as in ext_rays_bin
         if (idx >= 0) {
            d4 = D[idx];
            o4 = O[idx];
         }
         if (__float_as_uint(d4.w) != kHole) {
            slot = (int)idx;
            S = &V.jobs[__float_as_uint(d4.w)].scene;
            wn = S->wnodes;
            prims = S->prims;
            r = make_wideray(mk(o4), mk(d4));
            best.t = 3.402823466e38f;
            best.slot = -1;
            best.orig = -1;
            best.u = best.v = 0.0f;
            best.sdf = -1;
            sp = 0;
            node = 0;
            held_masks = kNoHeld;
         }
      }
      if (__ballot(slot >= 0) == 0) {  // the entries handed out may all have been holes
         if (pool.drained && pool.next == pool.end) break;
         continue;
      }
      if (slot >= 0) {
         // Postponed leaves, as on binary trees: the hit leaf children of the first node that
         // has any are set aside (mask, meta words, primitive base) and the lane goes on with
         // its next node, until nearly every lane of the wave holds leaves or has finished
         // (leaf_limit; the others go on after the leaf tests). A lane
         // that meets another node with hit leaves meanwhile waits on it, keeping that node's
         // child masks, and takes the node up again after its leaves without testing its
         // children again (masks found with a farther closest hit are a superset of the
         // current ones, so the closest hit is the same).
         uint32_t pmask = 0, pm0 = 0, pm1 = 0, pbase = 0;
         bool held = false;
         const unsigned lim = leaf_limit(V);
         for (;;) {
            if (node != kWideDone && !held) {
               const WNode n = load_wnode(wn, node);
               uint32_t hit_int, hit_leaf;
               if (held_masks != kNoHeld) {
                  hit_int = held_masks & 0xFFu;
                  hit_leaf = held_masks >> 8;
                  held_masks = kNoHeld;
               } else {
                  wide_children(n, r.o, r.inv, r.perm, 0.001f, best.t, hit_int, hit_leaf);
               }
               if (hit_leaf && pmask) {
                  held = true;
                  held_masks = hit_int | (hit_leaf << 8);
               } else {
                  if constexpr (CNT) ++tc.nodes;
                  if (hit_leaf) {
                     pmask = hit_leaf;
                     pm0 = n.meta[0];
                     pm1 = n.meta[1];
                     pbase = n.prim_base;
                  }
                  node = wide_next<STACK>(hit_int, n.child_base, n.e_imask >> 24, r.perm, lds, sp);
               }
            }
            if (__popcll(__ballot(pmask == 0 && node != kWideDone)) <= lim) break;
         }
         while (pmask) {
            const int sl = __ffs(pmask) - 1;
            pmask &= pmask - 1u;
            const uint32_t m = ((sl < 4 ? pm0 : pm1) >> (8 * (sl & 3))) & 0xFFu;
            const int first = (int)pbase + (int)(m & 31u);
            const int cnt = (int)(m >> 5) + 1;
            for (int i = first; i < first + cnt; ++i)
               test_closest<CNT, true>(*S, prims, i, r.o, r.d, 0.001f, best, &tc);
         }
         if (node == kWideDone) {
            V.hit[slot] = make_float4(best.t, __int_as_float(best.slot), best.u, best.v);
            if (best.sdf >= 0) pend = slot;  // an SDF field lies on the way
            slot = -1;
         }
      }
   }
   queue_marches();
   res.close(V.qs0);
}

template <bool CNT, int STACK>
__device__ void shadow_rays_wide(const WfView& V, int level, int* lds, TravCount& tc) {
   unsigned* sctl = V.ctl + kCtl * level + 3;
   const unsigned count = sctl[0];
   const unsigned step = 64u * (unsigned)fit_chunk(count, V.chunk);
   WavePool pool = start_pool(&sctl[1], count);
   int slot = -1, sp = 0;
   uint32_t node = kWideDone;
   uint32_t held_masks = kNoHeld;  // as for extension rays
   const SceneView* S = nullptr;  // read for SDF fields only
   const WNode* wn = nullptr;     // the scene's arrays in registers, as in ext_rays_bin
   const PrimRec* prims = nullptr;
   WideRay r;
   constexpr float tmin = 1e-4f, tmax = 1.0f - 1e-4f;
   for (;;) {
      if (!pool.drained || pool.next != pool.end) {
         const long long idx = pool.refill(&sctl[1], count, step, slot < 0, V.refill);
         float4 o4 = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1)), d4 = o4;  // one round trip
         if (idx >= 0) {
            o4 = V.so[idx];
            d4 = V.sd[idx];
         }
         if (__float_as_int(o4.w) >= 0) {
            slot = (int)idx;  // the shadow ray's position; o4.w holds its path's slot
            S = &V.jobs[__float_as_uint(d4.w)].scene;
            wn = S->wnodes;
            prims = S->prims;
            r = make_wideray(mk(o4), mk(d4));
            sp = 0;
            node = 0;
            held_masks = kNoHeld;
         }
      }
      if (__ballot(slot >= 0) == 0) {  // the entries handed out may all have been holes
         if (pool.drained && pool.next == pool.end) break;
         continue;
      }
      if (slot >= 0) {
         bool occluded = false;
         uint32_t pmask = 0, pm0 = 0, pm1 = 0, pbase = 0;  // postponed leaves, as for extension rays
         bool held = false;
         const unsigned lim = leaf_limit(V);
         for (;;) {
            if (node != kWideDone && !held) {
               const WNode n = load_wnode(wn, node);
               uint32_t hit_int, hit_leaf;
               if (held_masks != kNoHeld) {
                  hit_int = held_masks & 0xFFu;
                  hit_leaf = held_masks >> 8;
                  held_masks = kNoHeld;
               } else {
                  wide_children(n, r.o, r.inv, r.perm, tmin, tmax, hit_int, hit_leaf);
               }
               if (hit_leaf && pmask) {
                  held = true;
                  held_masks = hit_int | (hit_leaf << 8);
               } else {
                  if constexpr (CNT) ++tc.nodes;
                  if (hit_leaf) {
                     pmask = hit_leaf;
                     pm0 = n.meta[0];
                     pm1 = n.meta[1];
                     pbase = n.prim_base;
                  }
                  node = wide_next<STACK>(hit_int, n.child_base, n.e_imask >> 24, r.perm, lds, sp);
               }
            }
            if (__popcll(__ballot(pmask == 0 && node != kWideDone)) <= lim) break;
         }
         while (pmask && !occluded) {
            const int sl = __ffs(pmask) - 1;
            pmask &= pmask - 1u;
            const uint32_t m = ((sl < 4 ? pm0 : pm1) >> (8 * (sl & 3))) & 0xFFu;
            const int first = (int)pbase + (int)(m & 31u);
            const int cnt = (int)(m >> 5) + 1;
            for (int i = first; i < first + cnt && !occluded; ++i)
               occluded = test_any<CNT>(*S, prims, i, r.o, r.d, tmin, tmax, &tc);
         }
         if (occluded) {
            node = kWideDone;
            held_masks = kNoHeld;
         }
         if (node == kWideDone) {
            if (!occluded) {
               const int ps = __float_as_int(V.so[slot].w);
               const float4 c = V.sc[slot];
               const float4 rd = V.rad[ps];
               const V3 L = mk(rd) + mk(c);
               V.rad[ps] = make_float4(L.x, L.y, L.z, 0.0f);
            }
            slot = -1;
         }
      }
   }
}

// Shadow rays of one level, each wave tracing batches of 64 rays to completion.
template <bool CNT, int LAYOUT, int STACK>
__device__ void shadow_rays_drain(const WfView& V, int level, int* lds, TravCount& ts) {
   unsigned* sctl = V.ctl + kCtl * level + 3;
   const unsigned count = sctl[0];
   drain(&sctl[1], count, V.chunk, [&](unsigned idx) {
      const float4 o = idx < count ? V.so[idx] : make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
      const int ps = __float_as_int(o.w);  // the path's slot; -1 at an unused position
      if (ps >= 0) {
         const float4 d = V.sd[idx];
         const SceneView& S = V.jobs[__float_as_uint(d.w)].scene;
         if (!any_hit<kWfBlock, CNT, LAYOUT, STACK>(S, mk(o), mk(d), 1e-4f, 1.0f - 1e-4f, lds, &ts)) {
            const float4 c = V.sc[idx];
            const float4 r = V.rad[ps];
            const V3 L = mk(r) + mk(c);
            V.rad[ps] = make_float4(L.x, L.y, L.z, 0.0f);
         }
      }
   });
}

// Traversal stage: first the extension rays of level `ext_level` (closest hit), then the
// shadow rays of level `shadow_level` (any hit; unoccluded contributions are added to the
// path's radiance). Both queues were filled by the preceding shade launch and touch
// disjoint state, so one persistent launch drains both and waves that finish their
// extension work fill the launch's tail with shadow work. Either level may be -1.
template <bool CNT, int LAYOUT, int STACK = kStack>
__global__ __launch_bounds__(kWfBlock, (LAYOUT == L_BIN && STACK == kShortStack)  ? 7
                                       : (LAYOUT == L_WIDE && STACK == kShortWide) ? 6
                                                                                   : 1) void k_wf_trace(
    WfView V, int ext_level, int shadow_level) {
   // binary traversal: STACK - 1 LDS entries per lane (the top stays in a register); the
   // wide one keeps STACK / 2 pairs (the per-ray traversal of mixed layouts kStack / 2)
   __shared__ int stack[(LAYOUT == L_BIN ? STACK - 1 : (LAYOUT == L_WIDE ? STACK : kStack)) * kWfBlock];
   TravCount te, ts;
   if constexpr (LAYOUT == L_BIN) {
      if (ext_level >= 0) ext_rays_bin<CNT, STACK>(V, ext_level, stack + threadIdx.x, te);
      if (shadow_level >= 0) {
         // a shadow ray across an SDF field marches it inline; lanes refilled at different
         // times would march one by one, so passes with fields keep whole batches together
         if (V.shadow_refill)
            shadow_rays_bin<CNT, STACK>(V, shadow_level, stack + threadIdx.x, ts);
         else
            shadow_rays_drain<CNT, LAYOUT, STACK>(V, shadow_level, stack + threadIdx.x, ts);
      }
   } else if constexpr (LAYOUT == L_WIDE) {
      // shadow rays refill lanes also in passes with SDF fields (marched inline): the
      // per-ray drain's traversal would raise the kernel to 136 VGPRs (3 waves per SIMD,
      // against 5 with the refilling traversals alone)
      if (ext_level >= 0) ext_rays_wide<CNT, STACK>(V, ext_level, stack + threadIdx.x, te);
      if (shadow_level >= 0) shadow_rays_wide<CNT, STACK>(V, shadow_level, stack + threadIdx.x, ts);
   } else {
   if (ext_level >= 0) {
      unsigned* ctl = V.ctl + kCtl * ext_level;
      const unsigned count = ctl[0];
      const float4* O = V.o[ext_level & 1];
      const float4* D = V.d[ext_level & 1];
      unsigned* mctl = V.sctl + 2 * (kSdfSlices + 1) * ext_level;
      drain(&ctl[1], count, V.chunk, [&](unsigned idx) {
         const int slot = (int)idx;  // the ray's position in the level
         bool march = false;
         const float4 d = idx < count ? D[idx] : make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kHole));
         if (__float_as_uint(d.w) != kHole) {
            const float4 o = O[idx];
            const SceneView& S = V.jobs[__float_as_uint(d.w)].scene;
            const HitMin h = closest_hit<kWfBlock, CNT, LAYOUT, true, STACK>(S, mk(o), mk(d), 0.001f,
                                                                            3.402823466e38f, stack + threadIdx.x, &te);
            V.hit[slot] = make_float4(h.t, __int_as_float(h.slot), h.u, h.v);
            march = h.sdf >= 0;  // an SDF field lies on the way: march it in k_wf_sdf
         }
         const int pos = wave_push(&mctl[0], march);
         if (march) V.qs0[pos] = slot;
      });
   }
   if (shadow_level >= 0) shadow_rays_drain<CNT, LAYOUT, STACK>(V, shadow_level, stack + threadIdx.x, ts);
   }
   if constexpr (CNT) {
      const unsigned long long a = wave_sum((unsigned long long)te.nodes), b = wave_sum((unsigned long long)te.prims);
      const unsigned long long c = wave_sum((unsigned long long)ts.nodes), d = wave_sum((unsigned long long)ts.prims);
      if (lane_id() == 0) {
         atomicAdd(&V.stats[0], a);
         atomicAdd(&V.stats[1], b);
         atomicAdd(&V.stats[2], c);
         atomicAdd(&V.stats[3], d);
      }
   }
}

constexpr int kMaxChunk = 64;  // batches of 64 a wave may claim per queue atomic

// Copy of a record at a wave-uniform address, every 32-bit word moved to scalar registers
// (words the caller never reads are dropped with their loads).
template <class T>
__device__ __forceinline__ T uniform_copy(const T* p) {
   static_assert(sizeof(T) % 4 == 0, "records are whole 32-bit words");
   constexpr int n = (int)(sizeof(T) / 4);
   const uint32_t* s = (const uint32_t*)p;
   uint32_t w[n];
#pragma unroll
   for (int k = 0; k < n; ++k) w[k] = (uint32_t)__builtin_amdgcn_readfirstlane((int)s[k]);
   T v;
   __builtin_memcpy(&v, w, sizeof(T));
   return v;
}

// Adds a wave's segment tallies of one job to the job's counters (extension + shadow
// segments, shadow segments), one atomic pair per wave and job.
__device__ __forceinline__ void flush_segments(const WfView& V, int job, unsigned long long& ext,
                                               unsigned long long& sh) {
   if (job >= 0 && ext + sh != 0 && lane_id() == 0) {
      unsigned long long* ctr = V.jobs[job].counters + 8 * (blockIdx.x % kCounterSlots);
      atomicAdd(&ctr[0], ext + sh);
      atomicAdd(&ctr[1], sh);
   }
   ext = sh = 0;
}

// Shade stage: one bounce (shade_step) per path of the level. Each continuing path's state
// goes, with its slot, to the next position of the other state set, and each shadow ray to
// the next position of the shadow arrays; positions come from per-wave reservations.
template <bool CNT, uint32_t FT>
__global__ __launch_bounds__(kWfBlock, 4) void k_wf_shade(WfView V, int level) {
   // SDF normals read the job's field from a per-wave LDS slot (kernels with SDF code only)
   constexpr int kSlots = (FT & SF_SDF) ? kWfBlock / 64 : 1, kRecs = (FT & SF_SDF) ? kSdfMaskMax : 1;
   __shared__ SdfRec frec[kSlots][kRecs];
   __shared__ float4 fsph[kSlots][kRecs];
   FieldSlot fs{frec[(FT & SF_SDF) ? threadIdx.x / 64 : 0], fsph[(FT & SF_SDF) ? threadIdx.x / 64 : 0]};
   const int lane = lane_id();
   const int cur = level & 1, nxt = cur ^ 1;
   unsigned* ctl = V.ctl + kCtl * level;
   const unsigned count = ctl[0];
   unsigned* ctl_next = V.ctl + kCtl * (level + 1);
   unsigned* sctl = ctl + 3;
   const int chunk = fit_chunk(count, V.shade_chunk);
   const unsigned res = pos_res(count, V.res_most);
   PosRes rn, rs;  // next-level and shadow positions
   PathCounts pc;
   int tally_job = -1;  // segments shaded by this wave, not yet added to the job's counters
   unsigned long long tally_ext = 0, tally_sh = 0;
   unsigned seen = count;  // skip the claims of an empty or drained level, as drain() does
   if (count != 0 && lane == 0) seen = __hip_atomic_load(&ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   bool more = __shfl(seen, 0) < count;
   while (more) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(&ctl[2], 64u * (unsigned)chunk);
      base = __shfl(base, 0);
      if (base >= count) break;
      more = base + 64u * (unsigned)chunk < count;
      for (int k = 0; k < chunk; ++k) {
         const unsigned bk = base + 64u * (unsigned)k;
         if (bk >= count) break;
         const unsigned idx = bk + (unsigned)lane;
         int slot = -1;
         float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f), d = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kHole));
         float4 t = o, hf = o;
         if (idx < count) {  // the whole path state at once: the loads overlap
            o = V.o[cur][idx];
            d = V.d[cur][idx];
            hf = V.hit[idx];
            if (level == 0 && V.dense) {  // fresh camera paths, at their slots
               slot = (int)idx;
               t = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
            } else {
               slot = V.ps[cur][idx];
               t = V.thr[cur][idx];
            }
         }
         const uint32_t job = __float_as_uint(d.w);
         bool pending = slot >= 0;
         bool cont = false, shadow = false;
         PathState ps;
         ShadowReq sh;
         Rng R{__float_as_uint(o.w), false};
         uint32_t flags = __float_as_uint(t.w);
         // One job at a time: the lanes of the first pending lane's job shade together with
         // that job's view read once into scalar registers (a wave's positions nearly always
         // belong to one or two jobs: the levels keep work order).
         for (;;) {
            const unsigned long long m = __ballot(pending);
            if (m == 0) break;
            const int j0 = __builtin_amdgcn_readlane((int)job, __ffsll((long long)m) - 1);
            const bool mine = pending && (int)job == j0;
            if (j0 != tally_job) flush_segments(V, tally_job, tally_ext, tally_sh);
            tally_job = j0;
            ShadeField fld;
            if constexpr ((FT & SF_SDF) != 0) {
               fld.held = hold_field(fs, mine ? &V.jobs[j0].scene : nullptr);
               fld.rec = (const __attribute__((address_space(3))) gvec4*)fs.rec;
               fld.sph = (const __attribute__((address_space(3))) gvec4*)fs.sph;
            }
            if (mine) {
               pending = false;
               const JobView J = uniform_copy(V.jobs + j0);
               // the step's radiance starts at zero and is added to the slot's only if it
               // adds any (one term per step: rad + (0 + T*e) is rad + T*e bit for bit)
               ps = PathState{mk(o), mk(d), mk(t), mk(0.0f, 0.0f, 0.0f), (flags & kPrevLambert) != 0};
               HitMin h;
               h.t = hf.x;
               h.slot = __float_as_int(hf.y);
               h.u = hf.z;
               h.v = hf.w;
               cont = shade_step<CNT, FT>(J, h, level, ps, R, sh, pc, fld);
               cont = cont && level + 1 < J.max_depth;
               shadow = sh.valid;
               flags = (flags & ~kPrevLambert) | (ps.prev_lambert ? kPrevLambert : 0u);
               if (ps.L.x != 0.0f || ps.L.y != 0.0f || ps.L.z != 0.0f) {
                  const float4 r = V.rad[slot];
                  const V3 L = mk(r) + ps.L;
                  V.rad[slot] = make_float4(L.x, L.y, L.z, 0.0f);
               }
            }
            tally_ext += (unsigned long long)__popcll(__ballot(mine));
            tally_sh += (unsigned long long)__popcll(__ballot(mine && shadow));
         }
         const int pn = rn.take(&ctl_next[0], cont, res);
         if (cont) {
            V.o[nxt][pn] = make_float4(ps.o.x, ps.o.y, ps.o.z, __uint_as_float(R.s));
            V.d[nxt][pn] = make_float4(ps.d.x, ps.d.y, ps.d.z, d.w);
            V.thr[nxt][pn] = make_float4(ps.T.x, ps.T.y, ps.T.z, __uint_as_float(flags));
            V.ps[nxt][pn] = slot;
         }
         const int psh = rs.take(sctl, shadow, res);
         if (shadow) {
            V.so[psh] = make_float4(sh.o.x, sh.o.y, sh.o.z, __int_as_float(slot));
            V.sd[psh] = make_float4(sh.d.x, sh.d.y, sh.d.z, d.w);
            V.sc[psh] = make_float4(sh.contrib.x, sh.contrib.y, sh.contrib.z, 0.0f);
         }
      }
   }
   flush_segments(V, tally_job, tally_ext, tally_sh);
   // positions reserved and not filled: holes the next stages skip
   for (unsigned i = rn.next + (unsigned)lane; i < rn.end; i += 64) {
      V.d[nxt][i] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kHole));
      V.ps[nxt][i] = -1;
   }
   for (unsigned i = rs.next + (unsigned)lane; i < rs.end; i += 64)
      V.so[i] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
   if constexpr (CNT) {
      unsigned long long a = wave_sum((unsigned long long)pc.mats), b = wave_sum((unsigned long long)pc.light_prims);
      if (lane == 0) {
         atomicAdd(&V.stats[4], a);
         atomicAdd(&V.stats[5], b);
      }
   }
}

// One thread per pixel present in the round: adds that pixel's samples of the round in
// sample order (float-float compensated sums), updates its count and the job's count of
// primary samples. A pixel that had converged before the round took no paths (its count
// does not change during a pass's rounds, so the generate stage saw the same sign).
__global__ __launch_bounds__(kWfBlock) void k_wf_accumulate(WfView V, long long w0, int n) {
   const int i = blockIdx.x * kWfBlock + threadIdx.x;
   bool active = false;
   int j = -1;
   unsigned long long prim = 0;
   if (i < n) {
      const long long w = w0 + i;
      j = find_job(V, w);
      const JobView& J = V.jobs[j];
      const long long units = J.nlive >= 0 ? (long long)J.nlive : (long long)J.width * J.height;
      const long long l = w - V.base[j];
      if (l < units || w - units < w0) {  // first sample of this pixel within the round
         active = true;
         int s, p;
         split_work(l, (unsigned)units, s, p);
         if (J.nlive >= 0) p = J.live[p];
         const int n0 = J.counts[p];
         if (n0 >= 0) {
            const long long end = min(w0 + n, V.base[j + 1]);
            float4 a = J.acc[p], c = J.comp[p];
            int added = 0;
            for (long long wk = w; wk < end; wk += units) {
               const float4 r = V.rad[wk - w0];
               accumulate(a, c, mk(r));
               ++added;
            }
            J.acc[p] = a;
            J.comp[p] = c;
            J.counts[p] = n0 + added;
            prim = (unsigned long long)added;
         }
      }
   }
   // per-job counters: one atomic per wave when the wave's pixels belong to one job
   const unsigned long long act = __ballot(active);
   if (act == 0) return;
   const int leader = __ffsll((long long)act) - 1;
   const int jl = __shfl(j, leader);
   const bool uniform = __ballot(active && j != jl) == 0;
   if (uniform) {
      const unsigned long long s2 = wave_sum(prim);
      if (lane_id() == leader && s2) atomicAdd(&V.jobs[jl].counters[8 * (blockIdx.x % kCounterSlots) + 2], s2);
   } else if (active && prim) {
      atomicAdd(&V.jobs[j].counters[8 * (blockIdx.x % kCounterSlots) + 2], prim);
   }
}

// Pass start for batch jobs with the reference adaptive test: keeps the sums before the
// pass, which that test compares against (render_acc_kernel.cu:185-191).
__global__ void k_wf_pass_begin(JobView J, float4* prev) {
   const int p = blockIdx.x * blockDim.x + threadIdx.x;
   if (p < J.width * J.height) prev[p] = J.acc[p];
}

// Pass end: the adaptive test on the pixels that received this pass's samples.
__global__ void k_wf_pass_end(JobView J, const float4* prev, int todo) {
   const int p = blockIdx.x * blockDim.x + threadIdx.x;
   int conv = 0;
   if (p < J.width * J.height) {
      const int n = J.counts[p];
      if (n > 0 && todo > 0) {
         const float4 a = J.acc[p], c = J.comp[p];
         const float4 a0 = prev ? prev[p] : a;
         const int m = adaptive_count(J, a0, n - todo, a, c, n);
         if (m < 0) {
            J.counts[p] = m;
            conv = 1;
         }
      }
   }
   const unsigned long long s = wave_sum((unsigned long long)conv);
   if ((threadIdx.x & 63) == 0 && s) atomicAdd(&J.counters[8 * (blockIdx.x % kCounterSlots) + 3], s);
}

// The pixels of an adaptive job that still sample (count >= 0), in any order; each pass's
// work then covers these only. The order does not reach the results: every pixel still
// folds its samples in sample order, and streams depend on (pixel, sample) only.
__global__ void k_wf_live(JobView J, int* live, unsigned* n) {
   const int p = blockIdx.x * blockDim.x + threadIdx.x;
   const bool on = p < J.width * J.height && J.counts[p] >= 0;
   const int pos = wave_push(n, on);
   if (on) live[pos] = p;
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
static size_t ctl_words(int levels) { return (size_t)(levels + 1) * (kCtl + 2 * (kSdfSlices + 1)); }

static rayon_status wave_reserve(rayon_ctx* ctx, int capacity, int levels, int njobs) {
   Wave*& W = ctx->wave;
   if (!W) W = new Wave();
   if (!W->cus) {
      int cus = 0;
      RY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
      W->cus = cus > 0 ? cus : 256;
      int n = 0;
      RY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_trace<false, L_BIN>, kWfBlock, 0));
      W->occ_trace = n > 0 ? n : 1;
      RY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_trace<false, L_WIDE>, kWfBlock, 0));
      W->occ_trace_wide = n > 0 ? n : 1;
      RY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_trace<false, L_BIN, kShortStack>, kWfBlock, 0));
      W->occ_trace_short = n > 0 ? n : 1;
      RY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_trace<false, L_WIDE, kShortWide>, kWfBlock, 0));
      W->occ_trace_wide_short = n > 0 ? n : 1;
      RY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_shade<false, SF_ALL>, kWfBlock, 0));
      W->occ_shade = n > 0 ? n : 1;
      RY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_shade<false, SF_LEAN>, kWfBlock, 0));
      W->occ_shade_lean = n > 0 ? n : 1;
      RY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_sdf<false>, kWfBlock, 0));
      W->occ_sdf = n > 0 ? n : 1;
      RY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_sdf_rest<false>, kWfBlock, 0));
      W->occ_sdf_rest = n > 0 ? n : 1;
   }
   if (W->capacity >= capacity && W->levels >= levels && W->max_jobs >= njobs) return RAYON_OK;
   // Three groups grow independently and geometrically (path slots, control words, job
   // tables), so a growing job count or pass size does not reallocate gigabytes of slot
   // state each time. Captured graphs hold the old pointers: drop them.
   RY_HIP(hipStreamSynchronize(ctx->stream));
   for (auto& kv : W->graphs) {
      (void)hipGraphExecDestroy(kv.second.exec);
      (void)hipGraphDestroy(kv.second.graph);
   }
   W->graphs.clear();
   auto release = [](std::vector<DevBuf>& g) {
      for (auto& b : g)
         if (b.p) (void)hipFree(b.p);
      g.clear();
   };
   auto alloc = [&](std::vector<DevBuf>& g, size_t bytes, void** p) -> rayon_status {
      hipError_t e = hipMalloc(p, bytes);
      if (e != hipSuccess) return hip_fail(e, "hipMalloc (wavefront state)");
      g.push_back({*p, bytes});
      return RAYON_OK;
   };
   void* p;
   rayon_status st;
   if (!W->stats_dev) {
      if ((st = alloc(W->bufs, 16 * sizeof(unsigned long long), &p)) != RAYON_OK) return st;
      W->stats_dev = (unsigned long long*)p;
      RY_HIP(launch_fill_u32((uint32_t*)W->stats_dev, 0u, 32, ctx->stream));
   }
   if (capacity > W->capacity) {
      release(W->bufs_slots);
      const int cap = (int)std::max<long long>(capacity, std::min<long long>(2ll * W->capacity, W->opt_capacity));
      // positions of a level: its paths plus the holes of the shade waves' reservations
      const size_t npos = (size_t)cap + (size_t)(kPosHeadroomPerCu * W->cus);
      if ((st = alloc(W->bufs_slots, (size_t)cap * sizeof(float4), &p)) != RAYON_OK) return st;
      W->v.rad = (float4*)p;
      float4** f4[] = {&W->v.o[0], &W->v.o[1], &W->v.d[0], &W->v.d[1], &W->v.thr[0], &W->v.thr[1],
                       &W->v.hit, &W->v.so, &W->v.sd, &W->v.sc};
      for (float4** f : f4) {
         if ((st = alloc(W->bufs_slots, npos * sizeof(float4), &p)) != RAYON_OK) return st;
         *f = (float4*)p;
      }
      int** qs[] = {&W->v.ps[0], &W->v.ps[1], &W->v.qs0, &W->v.qs1};
      for (int** q : qs) {
         // the marching queues also hold the unused tails of the trace waves' reservations
         const bool marching = q == &W->v.qs0 || q == &W->v.qs1;
         const size_t n = npos + (marching ? (size_t)(kMarchHeadroomPerCu * W->cus) : 0);
         if ((st = alloc(W->bufs_slots, n * sizeof(int), &p)) != RAYON_OK) return st;
         *q = (int*)p;
      }
      W->capacity = cap;
   }
   if (levels > W->levels) {
      release(W->bufs_ctl);
      const int lv = std::max(levels, std::max(2 * W->levels, 16));
      // control words: kCtl per level, then 2 per (level, marching slice), cleared together
      if ((st = alloc(W->bufs_ctl, ctl_words(lv) * sizeof(unsigned), &p)) != RAYON_OK) return st;
      W->v.ctl = (unsigned*)p;
      W->v.sctl = W->v.ctl + (size_t)(lv + 1) * kCtl;
      W->levels = lv;
   }
   if (njobs > W->max_jobs) {
      release(W->bufs_tables);
      const int mj = std::max(njobs, std::max(2 * W->max_jobs, 64));
      if ((st = alloc(W->bufs_tables, (size_t)mj * sizeof(JobView), &p)) != RAYON_OK) return st;
      W->jobs_dev = (JobView*)p;
      if ((st = alloc(W->bufs_tables, (size_t)(mj + 1) * sizeof(long long), &p)) != RAYON_OK) return st;
      W->base_dev = (long long*)p;
      if ((st = alloc(W->bufs_tables, (size_t)mj * sizeof(JobView), &p)) != RAYON_OK) return st;
      W->mjobs_dev = (JobView*)p;
      if ((st = alloc(W->bufs_tables, (size_t)(mj + 1) * sizeof(int), &p)) != RAYON_OK) return st;
      W->mtile_dev = (int*)p;
      if ((st = alloc(W->bufs_tables, (size_t)mj * sizeof(int), &p)) != RAYON_OK) return st;
      W->mtodo_dev = (int*)p;
      W->max_jobs = mj;
   }
   W->v.jobs = W->jobs_dev;
   W->v.base = W->base_dev;
   W->v.stats = W->stats_dev;
   return RAYON_OK;
}

void wave_destroy(rayon_ctx* ctx) {
   Wave* W = ctx->wave;
   if (!W) return;
   for (auto& pe : W->pending) {
      (void)hipEventDestroy(pe.e0);
      (void)hipEventDestroy(pe.e1);
   }
   for (auto& kv : W->graphs) {
      (void)hipGraphExecDestroy(kv.second.exec);
      (void)hipGraphDestroy(kv.second.graph);
   }
   for (auto* g : {&W->bufs, &W->bufs_slots, &W->bufs_ctl, &W->bufs_tables})
      for (auto& b : *g)
         if (b.p) (void)hipFree(b.p);
   delete W;
   ctx->wave = nullptr;
}

namespace {

// Launches one stage kernel. With stage timing on, the launch goes through
// hipExtLaunchKernelGGL with start/stop events, which the runtime stamps at the kernel's
// own dispatch and completion; the events are read back later (rayon_ctx_wave_stats),
// so timing adds no host synchronisation.
template <typename K, typename... Args>
void launch_stage(Wave* W, int stage, K kernel, dim3 grid, dim3 block, hipStream_t st, Args... args) {
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if (W->opt_timing && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
      hipExtLaunchKernelGGL(kernel, grid, block, 0, st, e0, e1, 0, args...);
      W->pending.push_back({stage, e0, e1});
   } else {
      if (e0) (void)hipEventDestroy(e0);
      hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
   }
   W->launches[stage] += 1;
}

}  // namespace

// Shadow rays refill lanes unless one of them may cross an SDF field: a field marched inline
// by lanes refilled at different times runs one lane at a time. Only jobs with next-event
// estimation (and lights) cast shadow rays, so fields in other jobs' scenes do not matter
// (the 256-job mix: Cornell jobs' shadow rays beside SDF jobs without NEE).
static bool shadow_refill_ok(const std::vector<rayon_job*>& jobs) {
   for (rayon_job* j : jobs)
      if (j->scene->view.num_sdf > 0 && (j->params.flags & RAYON_FLAG_NEE) && j->scene->view.num_lights > 0)
         return false;
   return true;
}

// Enqueues every launch of one pass (the sequence a graph captures).
static hipError_t enqueue_pass(rayon_ctx* ctx, const std::vector<rayon_job*>& jobs, const std::vector<int>& todo,
                               long long total, int levels, int nmega, int mega_tiles, int mbin, int mbin_tiles) {
   Wave* W = ctx->wave;
   hipStream_t st = ctx->stream;
   if (nmega > 0) {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      const bool timed = W->opt_timing && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
      if (timed) (void)hipEventRecord(e0, st);
      hipError_t e = launch_multi(W->mjobs_dev, W->mtile_dev, W->mtodo_dev, mbin, mbin_tiles, nmega, mega_tiles,
                                  W->opt_count != 0, st);
      if (e != hipSuccess) return e;
      if (timed) {
         (void)hipEventRecord(e1, st);
         W->pending.push_back({3, e0, e1});
      } else {
         if (e0) (void)hipEventDestroy(e0);
         if (e1) (void)hipEventDestroy(e1);
      }
      W->launches[3] += 1;
   }
   if (total <= 0) return hipSuccess;
   WfView V = W->v;
   V.njobs = (int)jobs.size();
   V.levels = levels;
   V.shade_chunk = W->opt_shade_chunk;
   V.shadow_refill = shadow_refill_ok(jobs) ? 1 : 0;
   // persistent grids: the resident workgroups, but no more than one lane per path of a
   // round (small passes then launch and drain quickly)
   const long long per_round = std::min<long long>(std::max(total, 1ll), W->round);
   const int needed = (int)std::max<long long>(1, (per_round + kWfBlock - 1) / kWfBlock);
   auto grid_of = [&](int occ) {
      return std::min(needed, W->cus * (W->opt_wg_per_cu > 0 ? W->opt_wg_per_cu : occ));
   };
   const dim3 g_shade(grid_of(W->occ_shade)), blk(kWfBlock);
   // the shade kernel without the code of features no scene of the pass uses
   uint32_t feat = 0;
   for (rayon_job* j : jobs) feat |= j->scene->features;
   const bool lean = feat == SF_LEAN;
   const dim3 g_shade_lean(grid_of(W->occ_shade_lean));
   // output reservations of the shade waves: their unfilled tails must fit the headroom
   {
      const long long waves = (long long)std::max(g_shade.x, g_shade_lean.x) * (kWfBlock / 64);
      const long long most = (kPosHeadroomPerCu * W->cus / waves) / 64 * 64;
      V.res_most = (unsigned)std::min<long long>(1024, std::max<long long>(64, most));
   }
   // one layout for the whole pass when every scene shares it
   int nwide = 0;
   for (rayon_job* j : jobs) nwide += j->scene->view.wnodes != nullptr;
   const int layout = nwide == 0 ? L_BIN : (nwide == (int)jobs.size() ? L_WIDE : L_ANY);
   const bool wide = layout == L_WIDE;
   V.chunk = W->opt_chunk > 0 ? W->opt_chunk : (wide ? kChunkWide : kChunkBin);
   V.refill = W->opt_refill > 0 ? W->opt_refill : (wide ? kRefillWide : kRefillBin);
   V.leaf_exit = W->opt_leaf_exit >= 0 ? (unsigned)W->opt_leaf_exit : (wide ? kLeafExitWide : kLeafExitBin);
   V.leaf_div = wide ? kLeafDivWide : kLeafDivBin;
   // binary trees no deeper than the short stack take the kernel with the smaller LDS stack
   int depth = 0;
   for (rayon_job* j : jobs) depth = std::max(depth, j->scene->info.max_depth);
   const bool short_stack = layout == L_BIN && depth <= kShortStack;
   // quantized trees no deeper than kShortWide / 2 levels: the wide kernel with the short stack
   int wdepth = 0;
   for (rayon_job* j : jobs) wdepth = std::max(wdepth, j->scene->wide_depth);
   const bool short_wide = layout == L_WIDE && wdepth <= kShortWide / 2;
   const dim3 g_trace(grid_of(short_stack       ? W->occ_trace_short
                              : short_wide      ? W->occ_trace_wide_short
                              : layout == L_BIN ? W->occ_trace
                                                : std::min(W->occ_trace, W->occ_trace_wide)));
   auto trace = [&](int ext, int sh) {
      if (W->opt_count) {
         if (layout == L_BIN) launch_stage(W, 1, k_wf_trace<true, L_BIN>, g_trace, blk, st, V, ext, sh);
         else if (layout == L_WIDE) launch_stage(W, 1, k_wf_trace<true, L_WIDE>, g_trace, blk, st, V, ext, sh);
         else launch_stage(W, 1, k_wf_trace<true, L_ANY>, g_trace, blk, st, V, ext, sh);
      } else {
         if (short_stack) launch_stage(W, 1, k_wf_trace<false, L_BIN, kShortStack>, g_trace, blk, st, V, ext, sh);
         else if (layout == L_BIN) launch_stage(W, 1, k_wf_trace<false, L_BIN>, g_trace, blk, st, V, ext, sh);
         else if (short_wide) launch_stage(W, 1, k_wf_trace<false, L_WIDE, kShortWide>, g_trace, blk, st, V, ext, sh);
         else if (layout == L_WIDE) launch_stage(W, 1, k_wf_trace<false, L_WIDE>, g_trace, blk, st, V, ext, sh);
         else launch_stage(W, 1, k_wf_trace<false, L_ANY>, g_trace, blk, st, V, ext, sh);
      }
   };
   // SDF marching slices after each extension trace, when a scene of the pass has a field
   int sdf_steps = 0;
   for (rayon_job* j : jobs)
      if (j->scene->view.num_sdf > 0) sdf_steps = std::max(sdf_steps, j->scene->view.sdf_steps);
   // slice 0 (set-up and the first kSdfSlice steps of every march, all lanes busy), then one
   // refilling launch that finishes the longer marches
   const int nslices = sdf_steps <= 0 ? 0 : (sdf_steps <= kSdfSlice ? 1 : 2);
   const dim3 g_sdf(grid_of(W->occ_sdf)), g_sdf_rest(grid_of(W->occ_sdf_rest));
   auto march = [&](int level) {
      if (level < 0 || nslices == 0) return;
      const int last = nslices == 1 ? 1 : 0;
      if (W->opt_count) launch_stage(W, 6, k_wf_sdf<true>, g_sdf, blk, st, V, level, 0, last);
      else launch_stage(W, 6, k_wf_sdf<false>, g_sdf, blk, st, V, level, 0, last);
      if (nslices < 2) return;
      if (W->opt_count) launch_stage(W, 6, k_wf_sdf_rest<true>, g_sdf_rest, blk, st, V, level);
      else launch_stage(W, 6, k_wf_sdf_rest<false>, g_sdf_rest, blk, st, V, level);
   };
   hipError_t e = hipSuccess;
   int dense = 1;
   for (rayon_job* j : jobs)
      if (j->params.flags & (RAYON_FLAG_ADAPTIVE_REF | RAYON_FLAG_ADAPTIVE_STDERR)) dense = 0;
   V.dense = dense;
   for (size_t k = 0; k < jobs.size(); ++k) {
      rayon_job* j = jobs[k];
      if (todo[k] > 0 && (j->params.flags & RAYON_FLAG_ADAPTIVE_REF) && j->prev) {
         const int np = j->params.width * j->params.height;
         launch_stage(W, 5, k_wf_pass_begin, dim3((np + 255) / 256), dim3(256), st, j->view, j->prev);
      }
   }
   for (long long w0 = 0; w0 < total; w0 += W->round) {
      const int n = (int)std::min<long long>(W->round, total - w0);
      ++W->rounds;
      if ((e = launch_fill_u32(V.ctl, 0u, (long long)ctl_words(W->levels), st)) != hipSuccess) return e;
      W->launches[5] += 1;
      const dim3 g_round((n + kWfBlock - 1) / kWfBlock);
      launch_stage(W, 0, k_wf_generate, g_round, blk, st, V, w0, n, dense);
      trace(0, -1);
      march(0);
      for (int lv = 0; lv < levels; ++lv) {
         if (W->opt_count)
            launch_stage(W, 2, k_wf_shade<true, SF_ALL>, g_shade, blk, st, V, lv);
         else if (lean)
            launch_stage(W, 2, k_wf_shade<false, SF_LEAN>, g_shade_lean, blk, st, V, lv);
         else
            launch_stage(W, 2, k_wf_shade<false, SF_ALL>, g_shade, blk, st, V, lv);
         trace(lv + 1 < levels ? lv + 1 : -1, lv);
         march(lv + 1 < levels ? lv + 1 : -1);
      }
      launch_stage(W, 4, k_wf_accumulate, g_round, blk, st, V, w0, n);
      if ((e = hipGetLastError()) != hipSuccess) return e;
   }
   for (size_t k = 0; k < jobs.size(); ++k) {
      rayon_job* j = jobs[k];
      if (todo[k] <= 0 || !(j->params.flags & (RAYON_FLAG_ADAPTIVE_REF | RAYON_FLAG_ADAPTIVE_STDERR))) continue;
      const int np = j->params.width * j->params.height;
      launch_stage(W, 5, k_wf_pass_end, dim3((np + 255) / 256), dim3(256), st, j->view,
                   (const float4*)j->prev, todo[k]);
   }
   return hipGetLastError();
}

// Path of a scene's jobs in batched passes: its explicit choice, else the megakernel for
// scenes small enough to stay in the L2 caches (there a whole path per lane wastes no
// bandwidth on queue state) and the wavefront otherwise.
int scene_batch_path(const rayon_scene* s) {
   if (s->batch_path == RAYON_PATH_WAVEFRONT || s->batch_path == RAYON_PATH_MEGAKERNEL) return s->batch_path;
   return s->info.device_bytes <= kMegaSceneBytes ? RAYON_PATH_MEGAKERNEL : RAYON_PATH_WAVEFRONT;
}

rayon_status render_batched(rayon_ctx* ctx, rayon_job* const* jobs_in, int n, int samples) {
   if (!ctx || (n > 0 && !jobs_in) || n < 0 || samples < 0) {
      set_error("rayon_render_batched: bad arguments");
      return RAYON_ERR_INVALID;
   }
   std::vector<rayon_job*> jobs;
   std::vector<int> todo;
   for (int k = 0; k < n; ++k) {
      rayon_job* j = jobs_in[k];
      if (!j || j->ctx->device != ctx->device) {
         set_error("rayon_render_batched: job %d is NULL or lives on another device", k);
         return RAYON_ERR_INVALID;
      }
      if (j->params.flags & RAYON_FLAG_REF_RNG) {
         set_error("rayon_render_batched: job %d uses the sequential reference stream (RAYON_FLAG_REF_RNG), "
                   "whose samples cannot run concurrently; render it with rayon_job_render_plain", k);
         return RAYON_ERR_UNSUPPORTED;
      }
      if (j->params.max_depth > 4000) {
         set_error("rayon_render_batched: job %d max_depth %d exceeds 4000", k, j->params.max_depth);
         return RAYON_ERR_INVALID;
      }
      for (int m = 0; m < k; ++m)
         if (jobs_in[m] == j) {
            set_error("rayon_render_batched: job %d listed twice", k);
            return RAYON_ERR_INVALID;
         }
      const int t = (int)std::max<long long>(0, std::min<long long>(samples, j->params.spp - j->samples_done));
      jobs.push_back(j);
      todo.push_back(t);
   }
   // each job keeps the path it took in its first pass (the two paths round differently)
   for (rayon_job* j : jobs)
      if (j->path < 0) j->path = scene_batch_path(j->scene);
   RY_HIP(hipSetDevice(ctx->device));
   std::lock_guard<std::mutex> lk(ctx->mu);
   // Adaptive jobs on the wavefront whose pixels may have converged: list the pixels still
   // sampling, so the pass's work covers those only (units of work per sample).
   std::vector<long long> units(jobs.size());
   std::vector<size_t> listed;
   for (size_t k = 0; k < jobs.size(); ++k) {
      rayon_job* j = jobs[k];
      units[k] = (long long)j->params.width * j->params.height;
      const bool adaptive = (j->params.flags & (RAYON_FLAG_ADAPTIVE_REF | RAYON_FLAG_ADAPTIVE_STDERR)) != 0;
      if (todo[k] <= 0 || j->path == RAYON_PATH_MEGAKERNEL || !adaptive ||
          j->samples_done < std::max(1, j->params.min_adaptive_samples))
         continue;
      if (!j->live) {
         void* p = nullptr;
         const size_t bytes = (size_t)units[k] * sizeof(int) + 256;
         RY_HIP(hipMalloc(&p, bytes));
         j->bufs.push_back({p, bytes});
         j->live_n = (unsigned*)p;
         j->live = (int*)((char*)p + 256);
      }
      RY_HIP(launch_fill_u32(j->live_n, 0u, 1, ctx->stream));
      hipLaunchKernelGGL(k_wf_live, dim3((unsigned)((units[k] + 255) / 256)), dim3(256), 0, ctx->stream, j->view,
                         j->live, j->live_n);
      RY_HIP(hipGetLastError());
      listed.push_back(k);
   }
   if (!listed.empty()) {
      std::vector<unsigned> counts(listed.size());
      for (size_t i = 0; i < listed.size(); ++i)
         RY_HIP(hipMemcpyAsync(&counts[i], jobs[listed[i]]->live_n, sizeof(unsigned), hipMemcpyDeviceToHost,
                               ctx->stream));
      RY_HIP(hipStreamSynchronize(ctx->stream));
      for (size_t i = 0; i < listed.size(); ++i) {
         units[listed[i]] = counts[i];
         if (counts[i] == 0) todo[listed[i]] = 0;  // every pixel has converged
      }
   }
   long long total = 0;
   int levels = 1, nmega = 0;
   for (size_t k = 0; k < jobs.size(); ++k) {
      if (todo[k] <= 0) continue;
      if (jobs[k]->path == RAYON_PATH_MEGAKERNEL) {
         ++nmega;
         continue;
      }
      total += (long long)todo[k] * units[k];
      levels = std::max(levels, jobs[k]->params.max_depth);
   }
   if (total == 0 && nmega == 0) return RAYON_OK;
   if (!ctx->wave) ctx->wave = new Wave();
   const int cap = std::max((int)std::min<long long>(ctx->wave->opt_capacity, std::max(total, 1ll)), 1 << 16);
   rayon_status st = wave_reserve(ctx, cap, levels, n);
   if (st != RAYON_OK) return st;
   Wave* W = ctx->wave;
   W->round = cap;  // rounds follow the capacity option, whatever has been allocated
   for (size_t k = 0; k < jobs.size(); ++k) {
      rayon_job* j = jobs[k];
      if (todo[k] > 0 && j->path != RAYON_PATH_MEGAKERNEL && (j->params.flags & RAYON_FLAG_ADAPTIVE_REF) && !j->prev) {
         void* p = nullptr;
         const size_t bytes = (size_t)j->params.width * j->params.height * sizeof(float4);
         RY_HIP(hipMalloc(&p, bytes));
         j->bufs.push_back({p, bytes});
         j->prev = (float4*)p;
      }
   }
   // job tables (only jobs with samples to add take part in the pass)
   std::vector<rayon_job*> active, mactive;
   std::vector<int> atodo, mtodo;
   W->host_jobs.clear();
   W->host_base.assign(1, 0);
   W->host_mjobs.clear();
   W->host_mtile.assign(1, 0);
   W->host_mtodo.clear();
   // megakernel jobs: binary-tree jobs first, then quantized-tree ones (a launch per layout)
   int mbin = 0;
   for (int wide = 0; wide < 2; ++wide)
      for (size_t k = 0; k < jobs.size(); ++k) {
         rayon_job* j = jobs[k];
         if (todo[k] <= 0 || j->path != RAYON_PATH_MEGAKERNEL || (j->scene->view.wnodes != nullptr) != (wide != 0))
            continue;
         fill_jobview(j);
         j->view.sample_base = (int)j->samples_done;
         mactive.push_back(j);
         mtodo.push_back(todo[k]);
         mbin += wide == 0;
         W->host_mjobs.push_back(j->view);
         W->host_mtile.push_back(W->host_mtile.back() + plain_tiles(j->params.width, j->params.height));
         W->host_mtodo.push_back(todo[k]);
      }
   for (size_t k = 0; k < jobs.size(); ++k) {
      if (todo[k] <= 0) continue;
      rayon_job* j = jobs[k];
      if (j->path == RAYON_PATH_MEGAKERNEL) continue;
      fill_jobview(j);
      j->view.sample_base = (int)j->samples_done;
      if (std::find(listed.begin(), listed.end(), k) != listed.end()) {
         j->view.live = j->live;
         j->view.nlive = (int)units[k];
      }
      active.push_back(j);
      atodo.push_back(todo[k]);
      W->host_jobs.push_back(j->view);
      W->host_base.push_back(W->host_base.back() + (long long)todo[k] * units[k]);
   }
   if (!active.empty()) {
      RY_HIP(hipMemcpyAsync(W->jobs_dev, W->host_jobs.data(), W->host_jobs.size() * sizeof(JobView),
                            hipMemcpyHostToDevice, ctx->stream));
      RY_HIP(hipMemcpyAsync(W->base_dev, W->host_base.data(), W->host_base.size() * sizeof(long long),
                            hipMemcpyHostToDevice, ctx->stream));
   }
   if (nmega > 0) {
      RY_HIP(hipMemcpyAsync(W->mjobs_dev, W->host_mjobs.data(), W->host_mjobs.size() * sizeof(JobView),
                            hipMemcpyHostToDevice, ctx->stream));
      RY_HIP(hipMemcpyAsync(W->mtile_dev, W->host_mtile.data(), W->host_mtile.size() * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
      RY_HIP(hipMemcpyAsync(W->mtodo_dev, W->host_mtodo.data(), W->host_mtodo.size() * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
   }
   // the table copies read pageable host memory: finish them before the vectors change again
   RY_HIP(hipStreamSynchronize(ctx->stream));
   const int mtiles = W->host_mtile.back(), mbin_tiles = W->host_mtile[mbin];
   // a graph keyed by adaptive jobs must be forgotten with them, which only their own
   // context's job destruction does: passes over other contexts' jobs run eagerly
   bool foreign = false;
   for (rayon_job* j : jobs) foreign = foreign || j->ctx != ctx;
   // passes over pixel lists change size every time: not worth a graph each
   const bool use_graph = W->opt_graph && !W->opt_timing && !foreign && listed.empty();
   if (!use_graph) {
      RY_HIP(enqueue_pass(ctx, active, atodo, total, levels, nmega, mtiles, mbin, mbin_tiles));
   } else {
      // The captured launches read every job through the device tables, so a graph fits
      // any job set with the same work per job; only the adaptive pass-begin/end kernels
      // take a job by value, so adaptive jobs are part of the key.
      std::vector<long long> key;
      for (size_t k = 0; k < active.size(); ++k) {
         const bool adaptive = (active[k]->params.flags & (RAYON_FLAG_ADAPTIVE_REF | RAYON_FLAG_ADAPTIVE_STDERR)) != 0;
         key.push_back(adaptive ? (long long)(uintptr_t)active[k] : 0);
         key.push_back(W->host_base[k + 1] - W->host_base[k]);
      }
      key.push_back(levels);
      key.push_back(W->round);
      key.push_back(W->opt_wg_per_cu);
      key.push_back(W->opt_chunk);
      key.push_back(W->opt_shade_chunk);
      key.push_back(W->opt_refill);
      key.push_back(W->opt_leaf_exit);
      key.push_back(W->opt_count);
      for (rayon_job* j : active) key.push_back(j->scene->view.wnodes != nullptr);
      int sdf_steps = 0;  // the marching slices the pass enqueues
      for (rayon_job* j : active)
         if (j->scene->view.num_sdf > 0) sdf_steps = std::max(sdf_steps, j->scene->view.sdf_steps);
      key.push_back(sdf_steps);
      key.push_back(shadow_refill_ok(active));
      int depth = 0;  // picks the trace kernel's stack size
      for (rayon_job* j : active) depth = std::max(depth, j->scene->info.max_depth);
      key.push_back(depth <= kShortStack);
      int wdepth = 0;  // picks the quantized-tree kernel's stack size
      for (rayon_job* j : active) wdepth = std::max(wdepth, j->scene->wide_depth);
      key.push_back(wdepth <= kShortWide / 2);
      uint32_t feat = 0;  // picks the shade kernel's feature set
      for (rayon_job* j : active) feat |= j->scene->features;
      key.push_back(feat == SF_LEAN);
      key.push_back(-1);
      key.push_back(nmega);
      key.push_back(mtiles);
      key.push_back(mbin);
      key.push_back(mbin_tiles);
      auto it = W->graphs.find(key);
      if (it == W->graphs.end() && W->graphs.size() >= 64) {  // bounded cache: start over
         for (auto& kv : W->graphs) {
            (void)hipGraphExecDestroy(kv.second.exec);
            (void)hipGraphDestroy(kv.second.graph);
         }
         W->graphs.clear();
         it = W->graphs.end();
      }
      if (it == W->graphs.end()) {
         Wave::GraphEntry ge;
         RY_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
         hipError_t e = enqueue_pass(ctx, active, atodo, total, levels, nmega, mtiles, mbin, mbin_tiles);
         hipGraph_t g = nullptr;
         hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
         if (e != hipSuccess) return hip_fail(e, "wavefront pass capture");
         if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
         ge.graph = g;
         RY_HIP(hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0));
         it = W->graphs.emplace(key, ge).first;
      }
      RY_HIP(hipGraphLaunch(it->second.exec, ctx->stream));
   }
   for (size_t k = 0; k < mactive.size(); ++k) mactive[k]->samples_done += mtodo[k];
   for (size_t k = 0; k < active.size(); ++k) active[k]->samples_done += atodo[k];
   return RAYON_OK;
}

void wave_forget_job(rayon_ctx* ctx, rayon_job* job) {
   Wave* W = ctx ? ctx->wave : nullptr;
   if (!W) return;
   for (auto it = W->graphs.begin(); it != W->graphs.end();) {
      bool uses = false;
      for (size_t k = 0; k + 1 < it->first.size(); k += 2)
         if (it->first[k] == (long long)(uintptr_t)job) uses = true;
      if (uses) {
         (void)hipGraphExecDestroy(it->second.exec);
         (void)hipGraphDestroy(it->second.graph);
         it = W->graphs.erase(it);
      } else {
         ++it;
      }
   }
}

rayon_status wave_set_option(rayon_ctx* ctx, int opt, long long value) {
   if (!ctx->wave) ctx->wave = new Wave();
   Wave* W = ctx->wave;
   switch (opt) {
   case RAYON_OPT_WAVE_CAPACITY:
      if (value < (1 << 16) || value > (1ll << 28)) {
         set_error("RAYON_OPT_WAVE_CAPACITY must be in [65536, 2^28]");
         return RAYON_ERR_INVALID;
      }
      W->opt_capacity = (int)value;
      break;
   case RAYON_OPT_WAVE_WG_PER_CU:
      if (value < 0 || value > 32) {
         set_error("RAYON_OPT_WAVE_WG_PER_CU must be in [0, 32] (0: occupancy of each stage)");
         return RAYON_ERR_INVALID;
      }
      W->opt_wg_per_cu = (int)value;
      break;
   case RAYON_OPT_WAVE_CHUNK:
      if (value < 0 || value > kMaxChunk) {
         set_error("RAYON_OPT_WAVE_CHUNK must be in [0, 64] (0: the tree layout's default)");
         return RAYON_ERR_INVALID;
      }
      W->opt_chunk = (int)value;
      break;
   case RAYON_OPT_SHADE_CHUNK:
      if (value < 1 || value > kMaxChunk) {
         set_error("RAYON_OPT_SHADE_CHUNK must be in [1, 64]");
         return RAYON_ERR_INVALID;
      }
      W->opt_shade_chunk = (int)value;
      break;
   case RAYON_OPT_TRACE_REFILL:
      if (value < 0 || value > 64) {
         set_error("RAYON_OPT_TRACE_REFILL must be in [0, 64] (0: the tree layout's default)");
         return RAYON_ERR_INVALID;
      }
      W->opt_refill = (int)value;
      break;
   case RAYON_OPT_TRACE_LEAF_EXIT:
      if (value < -1 || value > 64) {
         set_error("RAYON_OPT_TRACE_LEAF_EXIT must be in [-1, 64] (-1: the tree layout's default)");
         return RAYON_ERR_INVALID;
      }
      W->opt_leaf_exit = (int)value;
      break;
   case RAYON_OPT_GRAPH:
      W->opt_graph = value != 0;
      break;
   case RAYON_OPT_STAGE_TIMING:
      W->opt_timing = value != 0;
      break;
   case RAYON_OPT_COUNT_TRAVERSAL:
      W->opt_count = value != 0;
      break;
   default:
      set_error("unknown option %d", opt);
      return RAYON_ERR_INVALID;
   }
   return RAYON_OK;
}

rayon_status wave_get_option(rayon_ctx* ctx, int opt, long long* value) {
   if (!ctx->wave) ctx->wave = new Wave();
   Wave* W = ctx->wave;
   switch (opt) {
   case RAYON_OPT_WAVE_CAPACITY: *value = W->opt_capacity; break;
   case RAYON_OPT_WAVE_WG_PER_CU: *value = W->opt_wg_per_cu; break;
   case RAYON_OPT_WAVE_CHUNK: *value = W->opt_chunk; break;
   case RAYON_OPT_SHADE_CHUNK: *value = W->opt_shade_chunk; break;
   case RAYON_OPT_TRACE_REFILL: *value = W->opt_refill; break;
   case RAYON_OPT_TRACE_LEAF_EXIT: *value = W->opt_leaf_exit; break;
   case RAYON_OPT_GRAPH: *value = W->opt_graph; break;
   case RAYON_OPT_STAGE_TIMING: *value = W->opt_timing; break;
   case RAYON_OPT_COUNT_TRAVERSAL: *value = W->opt_count; break;
   default:
      set_error("unknown option %d", opt);
      return RAYON_ERR_INVALID;
   }
   return RAYON_OK;
}

rayon_status wave_stats(rayon_ctx* ctx, rayon_wave_stats* out, int reset) {
   std::memset(out, 0, sizeof(*out));
   Wave* W = ctx->wave;
   if (!W) return RAYON_OK;
   RY_HIP(hipStreamSynchronize(ctx->stream));
   for (auto& pe : W->pending) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, pe.e0, pe.e1) == hipSuccess) W->stage_ms[pe.stage] += ms;
      (void)hipEventDestroy(pe.e0);
      (void)hipEventDestroy(pe.e1);
   }
   W->pending.clear();
   for (int k = 0; k < 8; ++k) {
      out->stage_ms[k] = W->stage_ms[k];
      out->launches[k] = W->launches[k];
   }
   out->rounds = W->rounds;
   if (W->stats_dev) {
      unsigned long long s[16];
      RY_HIP(hipMemcpy(s, W->stats_dev, sizeof(s), hipMemcpyDeviceToHost));
      out->extend_node_fetches = (int64_t)s[0];
      out->extend_prim_fetches = (int64_t)s[1];
      out->shadow_node_fetches = (int64_t)s[2];
      out->shadow_prim_fetches = (int64_t)s[3];
      out->material_fetches = (int64_t)s[4];
      out->light_fetches = (int64_t)s[5];
      out->sdf_marches = (int64_t)s[8];
      out->sdf_steps = (int64_t)s[9];
      out->sdf_prim_evals = (int64_t)s[10];
      out->sdf_wave_slots = (int64_t)s[11];
   }
   if (reset) {
      for (int k = 0; k < 8; ++k) {
         W->stage_ms[k] = 0;
         W->launches[k] = 0;
      }
      W->rounds = 0;
      if (W->stats_dev) RY_HIP(launch_fill_u32((uint32_t*)W->stats_dev, 0u, 32, ctx->stream));
   }
   return RAYON_OK;
}

}  // namespace ry

extern "C" {

rayon_status rayon_render_batched(rayon_ctx* ctx, rayon_job* const* jobs, int32_t n, int32_t samples) {
   return ry::render_batched(ctx, jobs, n, samples);
}

rayon_status rayon_ctx_set_option(rayon_ctx* ctx, int32_t opt, int64_t value) {
   if (!ctx) {
      ry::set_error("rayon_ctx_set_option: NULL context");
      return RAYON_ERR_INVALID;
   }
   std::lock_guard<std::mutex> lk(ctx->mu);
   return ry::wave_set_option(ctx, opt, value);
}

rayon_status rayon_ctx_get_option(rayon_ctx* ctx, int32_t opt, int64_t* value) {
   if (!ctx || !value) {
      ry::set_error("rayon_ctx_get_option: NULL argument");
      return RAYON_ERR_INVALID;
   }
   std::lock_guard<std::mutex> lk(ctx->mu);
   long long v = 0;
   rayon_status s = ry::wave_get_option(ctx, opt, &v);
   *value = v;
   return s;
}

rayon_status rayon_ctx_wave_stats(rayon_ctx* ctx, rayon_wave_stats* out, int32_t reset) {
   if (!ctx || !out) {
      ry::set_error("rayon_ctx_wave_stats: NULL argument");
      return RAYON_ERR_INVALID;
   }
   std::lock_guard<std::mutex> lk(ctx->mu);
   return ry::wave_stats(ctx, out, reset);
}

}  // extern "C"
