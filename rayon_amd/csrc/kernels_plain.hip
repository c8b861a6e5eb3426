// kernels_plain.hip — the plain path: one launch of a per-pixel megakernel per job and
// pass (the role of renderAccKernel, shaders/render_acc_kernel.cu:128-282), plus the
// resolve, closest-hit query and primary-ray kernels used by the C ABI.
//
// Launch shape: 256-lane workgroups (4 wave64) covering a 16x16 pixel tile, one lane per
// pixel, samples looped in the lane. Workgroups are dispatched round-robin over the 8
// XCDs; the tile index is remapped so each XCD walks a contiguous band of tiles and its
// L2 sees spatially coherent BVH traffic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kernels.h"
#include "pathtrace.h"

namespace ry {

constexpr int kTile = 16;
constexpr int kBlock = kTile * kTile;

__device__ __forceinline__ int xcd_tile(int b, int per_xcd) { return (b & 7) * per_xcd + (b >> 3); }

// One 16x16 tile of a job: every lane runs `samples` whole paths of its pixel in sample
// order and folds them into the compensated sums once, then the wave's statistics go to
// counter slot `slot`.
template <bool CNT, int LAYOUT>
__device__ __forceinline__ void plain_tile(const JobView& J, int samples, int tiles_x, int tile, int slot,
                                           int* stack) {
   const int tid = threadIdx.x;
   PathCounts pc;
   int primary = 0, converged = 0;
   const int x = (tile % tiles_x) * kTile + (tid & (kTile - 1));
   const int y = (tile / tiles_x) * kTile + (tid / kTile);
   if (x < J.width && y < J.height) {
      const int pix = y * J.width + x;
      const int cnt = J.counts[pix];
      int todo = samples;
      if (J.spp_target > 0) todo = min(todo, J.spp_target - cnt);
      if (cnt >= 0 && todo > 0) {
         const float4 a0 = J.acc[pix];
         float4 a = a0, c = J.comp[pix];
         const bool refrng = (J.flags & RAYON_FLAG_REF_RNG) != 0;
         Rng R{refrng ? J.rng[pix] : 0u, refrng};
         for (int s = 0; s < todo; ++s) {
            if (!refrng) R.s = counter_init(J.seed_lo, J.seed_hi, (uint32_t)pix, (uint32_t)(cnt + s));
            V3 o, d;
            camera_ray(J, x, y, R, o, d);
            const V3 L = trace_path<kBlock, CNT, LAYOUT>(J, o, d, R, stack + tid, pc);
            accumulate(a, c, L);
         }
         primary = todo;
         J.acc[pix] = a;
         J.comp[pix] = c;
         if (refrng) J.rng[pix] = R.s;
         const int n = adaptive_count(J, a0, cnt, a, c, cnt + todo);
         converged = n < 0 ? 1 : 0;
         J.counts[pix] = n;
      }
   }
   unsigned long long v0 = wave_sum((unsigned long long)pc.segments + (unsigned long long)pc.shadow);
   unsigned long long v1 = wave_sum((unsigned long long)pc.shadow);
   unsigned long long v2 = wave_sum((unsigned long long)primary);
   unsigned long long v3 = wave_sum((unsigned long long)converged);
   unsigned long long* ctr = J.counters + 8 * (slot % kCounterSlots);
   if ((tid & 63) == 0) {
      if (v0) atomicAdd(&ctr[0], v0);
      if (v1) atomicAdd(&ctr[1], v1);
      if (v2) atomicAdd(&ctr[2], v2);
      if (v3) atomicAdd(&ctr[3], v3);
   }
   if constexpr (CNT) {
      unsigned long long c4 = wave_sum((unsigned long long)pc.trav.nodes);
      unsigned long long c5 = wave_sum((unsigned long long)pc.trav.prims);
      unsigned long long c6 = wave_sum((unsigned long long)pc.mats);
      unsigned long long c7 = wave_sum((unsigned long long)pc.light_prims);
      if ((tid & 63) == 0) {
         atomicAdd(&ctr[4], c4);
         atomicAdd(&ctr[5], c5);
         atomicAdd(&ctr[6], c6);
         atomicAdd(&ctr[7], c7);
      }
   }
}

template <bool CNT, int LAYOUT>
__global__ __launch_bounds__(kBlock) void k_plain(JobView J, int samples, int tiles_x, int ntiles, int per_xcd) {
   __shared__ int stack[kStack * kBlock];
   const int tile = xcd_tile(blockIdx.x, per_xcd);
   if (tile >= ntiles) return;
   plain_tile<CNT, LAYOUT>(J, samples, tiles_x, tile, blockIdx.x, stack);
}

// Many jobs in one launch: the tiles of all jobs are numbered consecutively (tile_base
// holds each job's first tile, njobs + 1 entries), and each workgroup finds its job by a
// binary search over that prefix. The per-pixel arithmetic is the plain kernel's, so a job
// gets the same bits from this launch as from its own.
// tile0 is the first tile of this launch: jobs whose trees have another layout run in a
// launch of their own over the rest of the same tables (tile_base holds absolute tiles).
template <bool CNT, int LAYOUT>
__global__ __launch_bounds__(kBlock) void k_multi(const JobView* __restrict__ jobs, const int* __restrict__ tile_base,
                                                  const int* __restrict__ todo, int njobs, int per_xcd, int tile0) {
   __shared__ int stack[kStack * kBlock];
   const int tile = tile0 + xcd_tile(blockIdx.x, per_xcd);
   if (tile >= tile_base[njobs]) return;
   int lo = 0, hi = njobs - 1;
   while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tile_base[mid] <= tile) lo = mid;
      else hi = mid - 1;
   }
   const JobView& J = jobs[lo];
   const int tiles_x = (J.width + kTile - 1) / kTile;
   plain_tile<CNT, LAYOUT>(J, todo[lo], tiles_x, tile - tile_base[lo], blockIdx.x, stack);
}

// gammaCorrectKernel semantics (render_acc_kernel.cu:11-44) on the compensated sums of
// n samples.
__device__ __forceinline__ void resolve_px(float4 a, float4 c, int n, float inv_gamma, uint8_t* o, int channels) {
   if (n <= 0) n = 1;
   const float inv = 1.0f / (float)n;
   const float r = fminf(powf(fmaxf((a.x + c.x) * inv, 0.0f), inv_gamma), 0.999f);
   const float g = fminf(powf(fmaxf((a.y + c.y) * inv, 0.0f), inv_gamma), 0.999f);
   const float b = fminf(powf(fmaxf((a.z + c.z) * inv, 0.0f), inv_gamma), 0.999f);
   o[0] = (uint8_t)(256.0f * r);
   o[1] = (uint8_t)(256.0f * g);
   o[2] = (uint8_t)(256.0f * b);
   if (channels == 4) o[3] = 255;
}

__global__ void k_resolve_u8(JobView J, uint8_t* out, int channels, float inv_gamma, int n_total) {
   const int pix = blockIdx.x * blockDim.x + threadIdx.x;
   if (pix >= J.width * J.height) return;
   resolve_px(J.acc[pix], J.comp[pix], J.counts ? abs(J.counts[pix]) : n_total, inv_gamma,
              out + (size_t)pix * channels, channels);
}

// First frames of up to kPreviewJobs jobs in one launch. The job table travels in the
// kernel arguments and the resolve is fused, so a first frame is this one kernel: no table
// upload, no cleared accumulation buffers, no separate resolve, nothing that would wait
// for a free compute unit between steps while another stream's kernels hold the chip.
// Each lane runs `samples` paths of its pixel from sample 0 in sample order, folds them
// into compensated sums that start at zero, and writes the RGBA8 pixel: the bits of the
// plain kernel's `samples`-sample render of the job resolved by k_resolve_u8. For a job
// with a statistics array the lane also stores those sums, compensations and the pixel's
// count as the job's first samples (what the plain kernel leaves after its first launch on
// cleared buffers), and the workgroup writes its tile's statistics to that array.
template <int LAYOUT>
__global__ __launch_bounds__(kBlock, 4) void k_preview(const PreviewBatch B) {
   __shared__ int stack[kStack * kBlock];
   __shared__ unsigned long long tile_counts[kBlock / 64][kPreviewTileCounts];
   const int tile = xcd_tile(blockIdx.x, B.per_xcd);
   if (tile >= B.ntiles) return;
   int k = 0;
   while (k + 1 < B.njobs && B.job[k + 1].first_tile <= tile) ++k;
   const PreviewJob& PJ = B.job[k];
   const JobView& J = PJ.view;
   const int tiles_x = (J.width + kTile - 1) / kTile;
   const int t = tile - PJ.first_tile;
   const int tid = threadIdx.x;
   const int x = (t % tiles_x) * kTile + (tid & (kTile - 1));
   const int y = (t / tiles_x) * kTile + (tid / kTile);
   PathCounts pc;
   int converged = 0;
   if (x < J.width && y < J.height) {
      const int pix = y * J.width + x;
      float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c = a;
      Rng R{0u, false};
      for (int s = 0; s < B.samples; ++s) {
         R.s = counter_init(J.seed_lo, J.seed_hi, (uint32_t)pix, (uint32_t)s);
         V3 o, d;
         camera_ray(J, x, y, R, o, d);
         const V3 L = trace_path<kBlock, false, LAYOUT>(J, o, d, R, stack + tid, pc);
         accumulate(a, c, L);
      }
      resolve_px(a, c, B.samples, B.inv_gamma, PJ.frame + (size_t)pix * 4, 4);
      if (PJ.part) {
         J.acc[pix] = a;
         J.comp[pix] = c;
         const int n = adaptive_count(J, make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0, a, c, B.samples);
         J.counts[pix] = n;
         converged = n < 0 ? 1 : 0;
      }
   }
   if (PJ.part) {  // one job per workgroup: the branch is uniform
      const int inside = x < J.width && y < J.height ? 1 : 0;
      const unsigned long long v0 = wave_sum((unsigned long long)pc.segments + (unsigned long long)pc.shadow);
      const unsigned long long v1 = wave_sum((unsigned long long)pc.shadow);
      const unsigned long long v2 = wave_sum((unsigned long long)(inside * B.samples));
      const unsigned long long v3 = wave_sum((unsigned long long)converged);
      if ((tid & 63) == 0) {
         tile_counts[tid >> 6][0] = v0;
         tile_counts[tid >> 6][1] = v1;
         tile_counts[tid >> 6][2] = v2;
         tile_counts[tid >> 6][3] = v3;
      }
      __syncthreads();
      if (tid < kPreviewTileCounts) {
         unsigned long long v = 0;
         for (int w = 0; w < kBlock / 64; ++w) v += tile_counts[w][tid];
         PJ.part[(size_t)t * kPreviewTileCounts + tid] = v;
      }
   }
}

// One launch clears up to kFillRanges buffers (words of 32 bits): blockIdx.y picks the
// buffer, the x blocks stride over it.
__global__ void k_fill_ranges(const FillRanges R, uint32_t v) {
   uint32_t* p = R.p[blockIdx.y];
   const long long n = R.n[blockIdx.y];
   const long long stride = (long long)gridDim.x * blockDim.x;
   for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_resolve_linear(JobView J, float* out) {
   const int pix = blockIdx.x * blockDim.x + threadIdx.x;
   if (pix >= J.width * J.height) return;
   const float4 a = J.acc[pix], c = J.comp[pix];
   int n = abs(J.counts[pix]);
   const float inv = n > 0 ? 1.0f / (float)n : 0.0f;
   out[3 * (size_t)pix + 0] = (a.x + c.x) * inv;
   out[3 * (size_t)pix + 1] = (a.y + c.y) * inv;
   out[3 * (size_t)pix + 2] = (a.z + c.z) * inv;
}

__global__ __launch_bounds__(256) void k_trace_closest(SceneView S, const float* o, const float* d, long long n,
                                                       float tmin, float tmax, long long* prim, float* t) {
   __shared__ int stack[kStack * 256];
   const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
   if (i >= n) return;
   const V3 oo = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
   const V3 dd = mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
   const HitMin h = S.wnodes ? closest_hit<256, false, L_WIDE>(S, oo, dd, tmin, tmax, stack + threadIdx.x)
                             : closest_hit<256, false, L_BIN>(S, oo, dd, tmin, tmax, stack + threadIdx.x);
   prim[i] = h.slot < 0 ? -1 : (long long)h.orig;
   t[i] = h.slot < 0 ? -1.0f : h.t;
}

__global__ void k_primary(JobView J, int sample, float* o, float* d) {
   const int pix = blockIdx.x * blockDim.x + threadIdx.x;
   if (pix >= J.width * J.height) return;
   const bool refrng = (J.flags & RAYON_FLAG_REF_RNG) != 0;
   Rng R{refrng ? J.rng[pix] : counter_init(J.seed_lo, J.seed_hi, (uint32_t)pix, (uint32_t)sample), refrng};
   V3 oo, dd;
   camera_ray(J, pix % J.width, pix / J.width, R, oo, dd);
   o[3 * (size_t)pix] = oo.x;
   o[3 * (size_t)pix + 1] = oo.y;
   o[3 * (size_t)pix + 2] = oo.z;
   d[3 * (size_t)pix] = dd.x;
   d[3 * (size_t)pix + 1] = dd.y;
   d[3 * (size_t)pix + 2] = dd.z;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, long long n) {
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

// Reference stream seeds: (uint32)(1234 + idx * 747796405), cuda_utils.cu:26.
__global__ void k_ref_rng_init(uint32_t* st, int n) {
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) st[i] = 1234u + (uint32_t)i * 747796405u;
}

// gammaCorrectKernel (render_acc_kernel.cu:11-44) on raw buffers: sums (+ optional
// compensation), per-pixel signed counts or one uniform count.
__global__ void k_resolve_u8_raw(const float4* acc, const float4* comp, const int* counts, int n_uniform, int npix,
                                 uint8_t* out, int channels, float inv_gamma) {
   const int pix = blockIdx.x * blockDim.x + threadIdx.x;
   if (pix >= npix) return;
   float4 a = acc[pix];
   if (comp) {
      const float4 c = comp[pix];
      a.x += c.x;
      a.y += c.y;
      a.z += c.z;
   }
   int n = counts ? abs(counts[pix]) : n_uniform;
   if (n <= 0) n = 1;
   const float inv = 1.0f / (float)n;
   uint8_t* o = out + (size_t)pix * channels;
   o[0] = (uint8_t)(256.0f * fminf(powf(fmaxf(a.x * inv, 0.0f), inv_gamma), 0.999f));
   o[1] = (uint8_t)(256.0f * fminf(powf(fmaxf(a.y * inv, 0.0f), inv_gamma), 0.999f));
   o[2] = (uint8_t)(256.0f * fminf(powf(fmaxf(a.z * inv, 0.0f), inv_gamma), 0.999f));
   if (channels == 4) o[3] = 255;
}

// Sample-count heat map: |count| / max through an 8-key piecewise-linear plasma ramp
// (the key table of render_acc_kernel.cu:50-82), 255-scaled (:110-112).
__global__ void k_heatmap(const int* counts, int npix, uint8_t* out, int channels, int max_samples) {
   const int pix = blockIdx.x * blockDim.x + threadIdx.x;
   if (pix >= npix) return;
   const float key[8][4] = {{0.000f, 0.050f, 0.030f, 0.530f}, {0.143f, 0.230f, 0.015f, 0.660f},
                            {0.286f, 0.450f, 0.005f, 0.660f}, {0.429f, 0.650f, 0.060f, 0.540f},
                            {0.571f, 0.820f, 0.170f, 0.360f}, {0.714f, 0.940f, 0.340f, 0.170f},
                            {0.857f, 0.990f, 0.560f, 0.040f}, {1.000f, 0.940f, 0.975f, 0.130f}};
   const int c = abs(counts[pix]);
   float t = max_samples > 0 ? fminf((float)c / (float)max_samples, 1.0f) : 0.0f;
   t = fminf(fmaxf(t, 0.0f), 1.0f);
   float r = key[7][1], g = key[7][2], b = key[7][3];
   for (int i = 0; i < 7; ++i) {
      if (t <= key[i + 1][0]) {
         const float u = (t - key[i][0]) / (key[i + 1][0] - key[i][0]);
         r = key[i][1] + u * (key[i + 1][1] - key[i][1]);
         g = key[i][2] + u * (key[i + 1][2] - key[i][2]);
         b = key[i][3] + u * (key[i + 1][3] - key[i][3]);
         break;
      }
   }
   uint8_t* o = out + (size_t)pix * channels;
   o[0] = (uint8_t)(255.0f * r);
   o[1] = (uint8_t)(255.0f * g);
   o[2] = (uint8_t)(255.0f * b);
   if (channels == 4) o[3] = 255;
}

// ---- host launchers --------------------------------------------------------------
hipError_t launch_plain(const JobView& J, int samples, hipStream_t st) {
   const int tx = (J.width + kTile - 1) / kTile, ty = (J.height + kTile - 1) / kTile;
   const int ntiles = tx * ty;
   const int per = (ntiles + 7) / 8;
   const bool cnt = (J.flags & RAYON_FLAG_COUNT_TRAVERSAL) != 0, wide = J.scene.wnodes != nullptr;
   const dim3 g(8 * per), b(kBlock);
   if (cnt && wide)
      hipLaunchKernelGGL((k_plain<true, L_WIDE>), g, b, 0, st, J, samples, tx, ntiles, per);
   else if (cnt)
      hipLaunchKernelGGL((k_plain<true, L_BIN>), g, b, 0, st, J, samples, tx, ntiles, per);
   else if (wide)
      hipLaunchKernelGGL((k_plain<false, L_WIDE>), g, b, 0, st, J, samples, tx, ntiles, per);
   else
      hipLaunchKernelGGL((k_plain<false, L_BIN>), g, b, 0, st, J, samples, tx, ntiles, per);
   return hipGetLastError();
}

hipError_t launch_multi(const JobView* jobs, const int* tile_base, const int* todo, int nbin, int bin_tiles,
                        int njobs, int total_tiles, bool count, hipStream_t st) {
   if (njobs <= 0 || total_tiles <= 0) return hipSuccess;
   const dim3 b(kBlock);
   if (count) {
      const int per = (total_tiles + 7) / 8;
      hipLaunchKernelGGL((k_multi<true, L_ANY>), dim3(8 * per), b, 0, st, jobs, tile_base, todo, njobs, per, 0);
      return hipGetLastError();
   }
   // one launch per tree layout, so a job runs the same kernel whatever it is batched with
   if (nbin > 0) {
      const int per = (bin_tiles + 7) / 8;
      hipLaunchKernelGGL((k_multi<false, L_BIN>), dim3(8 * per), b, 0, st, jobs, tile_base, todo, nbin, per, 0);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
   }
   if (njobs > nbin) {
      const int per = (total_tiles - bin_tiles + 7) / 8;
      hipLaunchKernelGGL((k_multi<false, L_WIDE>), dim3(8 * per), b, 0, st, jobs + nbin, tile_base + nbin, todo + nbin,
                         njobs - nbin, per, bin_tiles);
      return hipGetLastError();
   }
   return hipSuccess;
}

int plain_tiles(int width, int height) { return ((width + kTile - 1) / kTile) * ((height + kTile - 1) / kTile); }

hipError_t launch_resolve_u8(const JobView& J, uint8_t* out, int channels, float gamma, hipStream_t st) {
   const int n = J.width * J.height;
   hipLaunchKernelGGL(k_resolve_u8, dim3((n + 255) / 256), dim3(256), 0, st, J, out, channels, 1.0f / gamma, 1);
   return hipGetLastError();
}

hipError_t launch_preview(const JobView* views, uint8_t* const* frames, unsigned long long* const* parts, int n,
                          int samples, float gamma, hipStream_t st) {
   // binary-tree jobs first, then quantized-tree jobs, kPreviewJobs of one layout per launch
   std::vector<int> order;
   for (int wide = 0; wide < 2; ++wide)
      for (int k = 0; k < n; ++k)
         if ((views[k].scene.wnodes != nullptr) == (wide != 0)) order.push_back(k);
   for (size_t base = 0; base < order.size();) {
      const bool wide = views[order[base]].scene.wnodes != nullptr;
      PreviewBatch B{};
      B.samples = samples;
      B.inv_gamma = 1.0f / gamma;
      while (base < order.size() && B.njobs < kPreviewJobs && (views[order[base]].scene.wnodes != nullptr) == wide) {
         const int k = order[base++];
         const JobView& J = views[k];
         unsigned long long* part = parts ? parts[k] : nullptr;
         if ((J.flags & RAYON_FLAG_REF_RNG) || J.width <= 0 || J.height <= 0 || !frames[k]) return hipErrorInvalidValue;
         if (part && (!J.acc || !J.comp || !J.counts || (J.spp_target > 0 && J.spp_target < samples)))
            return hipErrorInvalidValue;
         B.job[B.njobs].view = J;
         B.job[B.njobs].frame = frames[k];
         B.job[B.njobs].part = part;
         B.job[B.njobs].first_tile = B.ntiles;
         B.ntiles += plain_tiles(J.width, J.height);
         ++B.njobs;
      }
      B.per_xcd = (B.ntiles + 7) / 8;
      const dim3 g(8 * B.per_xcd), b(kBlock);
      if (wide)
         hipLaunchKernelGGL((k_preview<L_WIDE>), g, b, 0, st, B);
      else
         hipLaunchKernelGGL((k_preview<L_BIN>), g, b, 0, st, B);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
   }
   return hipSuccess;
}

hipError_t launch_fill_ranges(uint32_t* const* p, const long long* n, int count, uint32_t v, hipStream_t st) {
   for (int base = 0; base < count; base += kFillRanges) {
      FillRanges R{};
      long long most = 0;
      const int m = std::min(kFillRanges, count - base);
      for (int k = 0; k < m; ++k) {
         R.p[k] = p[base + k];
         R.n[k] = n[base + k];
         most = std::max(most, n[base + k]);
      }
      if (most <= 0) continue;
      const dim3 g((unsigned)std::min<long long>((most + 255) / 256, 1024), m), b(256);
      hipLaunchKernelGGL(k_fill_ranges, g, b, 0, st, R, v);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
   }
   return hipSuccess;
}

hipError_t launch_resolve_linear(const JobView& J, float* out, hipStream_t st) {
   const int n = J.width * J.height;
   hipLaunchKernelGGL(k_resolve_linear, dim3((n + 255) / 256), dim3(256), 0, st, J, out);
   return hipGetLastError();
}

// Staging area of the allocation-free resolve (rayon_job_resolve_u8_noalloc): a fixed
// array of kNoAllocBytes per device, held until the process ends. Host code never uses
// this name as an address: it asks the runtime for the device address once per device
// (noalloc_scratch_address, with that device current) and passes that pointer on.
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

hipError_t noalloc_scratch_address(void** ptr) { return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)); }

hipError_t launch_trace_closest(const SceneView& S, const float* o, const float* d, long long n, float tmin,
                                float tmax, long long* prim, float* t, hipStream_t st) {
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_trace_closest, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, S, o, d, n, tmin, tmax,
                      prim, t);
   return hipGetLastError();
}

hipError_t launch_primary(const JobView& J, int sample, float* o, float* d, hipStream_t st) {
   const int n = J.width * J.height;
   hipLaunchKernelGGL(k_primary, dim3((n + 255) / 256), dim3(256), 0, st, J, sample, o, d);
   return hipGetLastError();
}

hipError_t launch_ref_rng_init(uint32_t* p, int n, hipStream_t st) {
   hipLaunchKernelGGL(k_ref_rng_init, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
   return hipGetLastError();
}

hipError_t launch_resolve_u8_raw(const float4* acc, const float4* comp, const int* counts, int n_uniform, int npix,
                                 uint8_t* out, int channels, float gamma, hipStream_t st) {
   if (npix <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_resolve_u8_raw, dim3((npix + 255) / 256), dim3(256), 0, st, acc, comp, counts, n_uniform, npix,
                      out, channels, 1.0f / gamma);
   return hipGetLastError();
}

hipError_t launch_heatmap(const int* counts, int npix, uint8_t* out, int channels, int max_samples, hipStream_t st) {
   if (npix <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_heatmap, dim3((npix + 255) / 256), dim3(256), 0, st, counts, npix, out, channels, max_samples);
   return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long long n, hipStream_t st) {
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
   return hipGetLastError();
}

}  // namespace ry
