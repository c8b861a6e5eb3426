// kernels.h — host-side launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace ry {

// First-frame launches: up to kPreviewJobs jobs per launch, their views in the kernel
// arguments (8 x 336 bytes, within the 4 KiB argument limit).
constexpr int kPreviewJobs = 8;
// Per-tile statistics of an accumulating first frame: segments (extension + shadow),
// shadow segments, primary samples, converged pixels of one 16x16 tile.
constexpr int kPreviewTileCounts = 4;
struct PreviewJob {
   JobView view;
   uint8_t* frame;            // RGBA8, width x height
   unsigned long long* part;  // accumulating: kPreviewTileCounts per tile of the job; else NULL
   int first_tile;            // first 16x16 tile of this job in the launch
   int pad_;
};
struct PreviewBatch {
   PreviewJob job[kPreviewJobs];
   int njobs, ntiles, per_xcd, samples;
   float inv_gamma;
};
constexpr int kFillRanges = 32;
struct FillRanges {
   uint32_t* p[kFillRanges];
   long long n[kFillRanges];
};

hipError_t launch_plain(const JobView& J, int samples, hipStream_t st);
// All tiles of `njobs` jobs (device tables: job views, tile prefix with njobs + 1 entries,
// samples per job); the first nbin jobs (bin_tiles tiles) have binary trees, the rest
// quantized ones, and each group runs in a launch of its own layout.
hipError_t launch_multi(const JobView* jobs, const int* tile_base, const int* todo, int nbin, int bin_tiles,
                        int njobs, int total_tiles, bool count, hipStream_t st);
int plain_tiles(int width, int height);
hipError_t launch_resolve_u8(const JobView& J, uint8_t* out, int channels, float gamma, hipStream_t st);
hipError_t launch_resolve_linear(const JobView& J, float* out, hipStream_t st);
hipError_t launch_trace_closest(const SceneView& S, const float* o, const float* d, long long n, float tmin,
                                float tmax, long long* prim, float* t, hipStream_t st);
hipError_t launch_primary(const JobView& J, int sample, float* o, float* d, hipStream_t st);
hipError_t launch_ref_rng_init(uint32_t* p, int n, hipStream_t st);
hipError_t launch_resolve_u8_raw(const float4* acc, const float4* comp, const int* counts, int n_uniform, int npix,
                                 uint8_t* out, int channels, float gamma, hipStream_t st);
hipError_t launch_heatmap(const int* counts, int npix, uint8_t* out, int channels, int max_samples, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long long n, hipStream_t st);
// dst[s] = src[perm[s]] for n primitive records (and their 3 vertex normals each when nsrc).
hipError_t launch_permute_recs(const PrimRec* src, const float4* nsrc, const int* perm, int n, PrimRec* dst,
                               float4* ndst, hipStream_t st);
// One launch per kPreviewJobs jobs: `samples` paths per pixel from sample 0, resolved with
// `gamma` into each job's RGBA8 frame. A job with a `parts` entry (NULL: none) also stores
// those samples as its first `samples` accumulated samples: its sums, compensations and
// per-pixel counts are written whole (no clear needed beforehand) with the plain kernel's
// arithmetic, and each tile's statistics go to parts[k][tile * kPreviewTileCounts + i] (a
// host-visible array, summed by the caller; the job's device counters are not touched).
hipError_t launch_preview(const JobView* views, uint8_t* const* frames, unsigned long long* const* parts, int n,
                          int samples, float gamma, hipStream_t st);
// Fixed staging array of the allocation-free resolve, one per device (64 MiB: a 4096 x 4096
// RGBA8 image); noalloc_scratch_address returns its address on the current device.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
hipError_t noalloc_scratch_address(void** ptr);
// Device material and accumulation code on caller-given inputs (parity probes).
hipError_t launch_scatter_probe(const MatRec& M, const JobView& J, const float* dirs, const float* nrm,
                                const float* tan, const int* front, const float* rnd, int nrnd, long long n,
                                float* od, float* at, float* em, int* ok, int* used, hipStream_t st);
hipError_t launch_accumulate_probe(const float* rgb, long long n, float4* out, hipStream_t st);
// Sets `count` buffers of n[k] 32-bit words to v, kFillRanges buffers per launch.
hipError_t launch_fill_ranges(uint32_t* const* p, const long long* n, int count, uint32_t v, hipStream_t st);

}  // namespace ry
