// kernels.h — host-side launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace ry {

hipError_t launch_plain(const JobView& J, int samples, hipStream_t st);
// All tiles of `njobs` jobs in one launch (device tables: job views, tile prefix with
// njobs + 1 entries, samples per job); layout: 0 binary, 1 wide, 2 mixed.
hipError_t launch_multi(const JobView* jobs, const int* tile_base, const int* todo, int njobs, int total_tiles,
                        bool count, int layout, hipStream_t st);
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

}  // namespace ry
