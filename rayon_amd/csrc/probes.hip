// probes.hip — the renderers' own device functions run on caller-given inputs, for
// parity tests: material scattering (scatter_material, cuda_raytracer.cuh:613-850) with
// a table of random numbers in place of the path's stream, and the compensated
// accumulation of a pixel (render_acc_kernel.cu:224's acc += ray_color) over a given
// series of sample values.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "pathtrace.h"

namespace ry {

namespace {

// Random numbers from a table, in consumption order; 0.5 once the table runs out (the
// count of numbers drawn is reported, so a caller sees that).
struct TableRng {
   const float* r;
   int n;
   int used;
   __device__ __forceinline__ float next() {
      const float v = used < n ? r[used] : 0.5f;
      ++used;
      return v;
   }
};

__global__ __launch_bounds__(256) void k_scatter_probe(MatRec M, JobView J, const float* dirs, const float* nrm,
                                                       const float* tan, const int* front, const float* rnd,
                                                       int nrnd, long long n, float* od, float* at, float* em,
                                                       int* ok, int* used) {
   const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
   if (i >= n) return;
   Surf s;
   s.p = mk(0.0f, 0.0f, 0.0f);
   s.n = ld3(nrm + 3 * i);
   s.tangent = ld3(tan + 3 * i);
   s.center = mk(0.0f, 0.0f, 0.0f);
   s.front = front[i] != 0;
   s.info = 0;
   TableRng R{rnd + i * nrnd, nrnd, 0};
   Scatter sc;
   scatter(M, mk(M.m0), J, s, ld3(dirs + 3 * i), R, sc);
   od[3 * i] = sc.d.x;
   od[3 * i + 1] = sc.d.y;
   od[3 * i + 2] = sc.d.z;
   at[3 * i] = sc.atten.x;
   at[3 * i + 1] = sc.atten.y;
   at[3 * i + 2] = sc.atten.z;
   em[3 * i] = sc.emitted.x;
   em[3 * i + 1] = sc.emitted.y;
   em[3 * i + 2] = sc.emitted.z;
   ok[i] = sc.scattered ? 1 : 0;
   used[i] = R.used;
}

// One lane folds the whole series, in order, exactly as a pixel's samples are folded.
__global__ void k_accumulate_probe(const float* rgb, long long n, float4* out) {
   if (blockIdx.x != 0 || threadIdx.x != 0) return;
   float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
   for (long long i = 0; i < n; ++i) accumulate(a, c, ld3(rgb + 3 * i));
   out[0] = a;
   out[1] = c;
}

}  // namespace

hipError_t launch_scatter_probe(const MatRec& M, const JobView& J, const float* dirs, const float* nrm,
                                const float* tan, const int* front, const float* rnd, int nrnd, long long n,
                                float* od, float* at, float* em, int* ok, int* used, hipStream_t st) {
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_scatter_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, M, J, dirs, nrm, tan,
                      front, rnd, nrnd, n, od, at, em, ok, used);
   return hipGetLastError();
}

hipError_t launch_accumulate_probe(const float* rgb, long long n, float4* out, hipStream_t st) {
   hipLaunchKernelGGL(k_accumulate_probe, dim3(1), dim3(64), 0, st, rgb, n, out);
   return hipGetLastError();
}

}  // namespace ry
