// bvh_device.hip — BVH build on the device (RAYON_BVH_DEVICE): a linear BVH over 63-bit
// Morton codes, so a new scene's acceleration structure takes milliseconds of GPU time
// instead of the host binned-SAH build (the reference builds a full-sweep SAH on the host,
// scenes/scene_description.hpp:793-925, before any pixel is rendered).
//
//   1. centroid bounds (block reduction, order-preserving integer atomics)
//   2. Morton code of each centroid, 21 bits per axis, primitive index as the value
//   3. radix sort of (code, index) pairs (rocPRIM through hipCUB)
//   4. radix tree over the sorted codes: Karras, "Maximizing Parallelism in the
//      Construction of BVHs, Octrees, and k-d Trees" (HPG 2012), equal codes ordered by
//      their position so every internal node covers a contiguous range
//   5. boxes bottom-up: one thread per leaf climbs while it is the second child to arrive
//   6. subtrees of at most four primitives become leaves (the host builder's leaf rule);
//      the remaining internal nodes are compacted by a prefix sum and written as 64-byte
//      pair nodes, the same layout the host builder emits
//
// Primitive boxes are the host's outward-rounded FP32 boxes, so every child box contains
// its primitives exactly as in the host tree. Closest-hit ties resolve by original index
// independently of traversal order, so renders over this tree equal those over the host
// tree bit for bit; only the traversal cost differs (a Morton tree is not SAH-optimised).
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "internal.h"
#include "kernels.h"

namespace ry {

namespace {

constexpr int kLeafMax = 4;  // primitives per leaf, as the host builder
constexpr int kBlock = 256;

__device__ __forceinline__ int ordered(float f) {
   const int i = __float_as_int(f);
   return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float unordered(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void k_init(int* cb, int* maxd) {
   if (threadIdx.x < 3) {
      cb[threadIdx.x] = 0x7FFFFFFF;       // min of ordered centroid coordinates
      cb[3 + threadIdx.x] = (int)0x80000000;  // max
   }
   if (threadIdx.x == 0) *maxd = 0;
}

__global__ __launch_bounds__(kBlock) void k_centroid_bounds(const float4* lo, const float4* hi, int n, int* cb) {
   __shared__ int s[6][kBlock / 64];
   int mn[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, mx[3] = {(int)0x80000000, (int)0x80000000, (int)0x80000000};
   for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
      const float4 a = lo[i], b = hi[i];
      const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
         mn[k] = min(mn[k], ordered(c[k]));
         mx[k] = max(mx[k], ordered(c[k]));
      }
   }
#pragma unroll
   for (int k = 0; k < 3; ++k) {
      for (int off = 32; off > 0; off >>= 1) {
         mn[k] = min(mn[k], __shfl_xor(mn[k], off));
         mx[k] = max(mx[k], __shfl_xor(mx[k], off));
      }
   }
   const int w = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 0) {
      for (int k = 0; k < 3; ++k) {
         s[k][w] = mn[k];
         s[3 + k][w] = mx[k];
      }
   }
   __syncthreads();
   if (threadIdx.x < 3) {
      int a = s[threadIdx.x][0], b = s[3 + threadIdx.x][0];
      for (int j = 1; j < kBlock / 64; ++j) {
         a = min(a, s[threadIdx.x][j]);
         b = max(b, s[3 + threadIdx.x][j]);
      }
      atomicMin(&cb[threadIdx.x], a);
      atomicMax(&cb[3 + threadIdx.x], b);
   }
}

// 21 bits of x spread to every third bit.
__device__ __forceinline__ uint64_t spread21(uint32_t x) {
   uint64_t v = x & 0x1FFFFFu;
   v = (v | v << 32) & 0x1F00000000FFFFull;
   v = (v | v << 16) & 0x1F0000FF0000FFull;
   v = (v | v << 8) & 0x100F00F00F00F00Full;
   v = (v | v << 4) & 0x10C30C30C30C30C3ull;
   v = (v | v << 2) & 0x1249249249249249ull;
   return v;
}

__global__ __launch_bounds__(kBlock) void k_morton(const float4* lo, const float4* hi, int n, const int* cb,
                                                   uint64_t* keys, int* vals) {
   const int i = blockIdx.x * kBlock + threadIdx.x;
   if (i >= n) return;
   const float4 a = lo[i], b = hi[i];
   const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
   uint64_t key = 0;
#pragma unroll
   for (int k = 0; k < 3; ++k) {
      const float l = unordered(cb[k]), h = unordered(cb[3 + k]);
      const float e = h - l;
      float q = e > 0.0f ? (c[k] - l) / e * 2097152.0f : 0.0f;
      q = fminf(fmaxf(q, 0.0f), 2097151.0f);
      key |= spread21((uint32_t)q) << k;
   }
   keys[i] = key;
   vals[i] = i;
}

// Common prefix length of sorted codes i and j (positions break ties); -1 outside [0, n).
__device__ __forceinline__ int delta(const uint64_t* k, int n, int i, int j) {
   if (j < 0 || j >= n) return -1;
   const uint64_t a = k[i], b = k[j];
   if (a != b) return __clzll((long long)(a ^ b));
   return 64 + __clz(i ^ j);
}

// Children: >= 0 internal node, < 0 leaf ~position. Parents: internal nodes in
// parent[0, n-1), leaves in parent[n-1 + position].
__global__ __launch_bounds__(kBlock) void k_radix_tree(const uint64_t* k, int n, int2* child, int2* range,
                                                       int* parent) {
   const int i = blockIdx.x * kBlock + threadIdx.x;
   if (i >= n - 1) return;
   const int d = delta(k, n, i, i + 1) - delta(k, n, i, i - 1) >= 0 ? 1 : -1;
   const int dmin = delta(k, n, i, i - d);
   int lmax = 2;
   while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
   int l = 0;
   for (int t = lmax >> 1; t >= 1; t >>= 1)
      if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
   const int j = i + l * d;
   const int dnode = delta(k, n, i, j);
   int s = 0, t = l;
   do {
      t = (t + 1) >> 1;
      if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
   } while (t > 1);
   const int g = i + s * d + min(d, 0);
   const int first = min(i, j), last = max(i, j);
   const int left = first == g ? ~g : g, right = last == g + 1 ? ~(g + 1) : g + 1;
   child[i] = make_int2(left, right);
   range[i] = make_int2(first, last);
   parent[left >= 0 ? left : n - 1 + ~left] = i;
   parent[right >= 0 ? right : n - 1 + ~right] = i;
   if (i == 0) parent[0] = -1;
}

__device__ __forceinline__ void child_box(int c, const float4* nlo, const float4* nhi, const float4* lo,
                                          const float4* hi, const int* vals, float4& a, float4& b, bool coherent) {
   if (c < 0) {
      const int p = vals[~c];
      a = lo[p];
      b = hi[p];
   } else if (coherent) {
      // written by another thread of this launch: read past the non-coherent L1
      float* pa = (float*)&nlo[c];
      float* pb = (float*)&nhi[c];
      a = make_float4(__hip_atomic_load(pa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(pa + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(pa + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0.0f);
      b = make_float4(__hip_atomic_load(pb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(pb + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(pb + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0.0f);
   } else {
      a = nlo[c];
      b = nhi[c];
   }
}

// One thread per leaf: climbs to the root, stopping at the first node whose other child has
// not been finished yet (that child's thread will continue from there).
__global__ __launch_bounds__(kBlock) void k_refit(const int2* child, const int* parent, int n, const float4* lo,
                                                  const float4* hi, const int* vals, float4* nlo, float4* nhi,
                                                  unsigned* arrived) {
   const int leaf = blockIdx.x * kBlock + threadIdx.x;
   if (leaf >= n) return;
   int p = parent[n - 1 + leaf];
   while (p >= 0) {
      __threadfence();
      if (atomicAdd(&arrived[p], 1u) == 0u) return;
      __threadfence();
      const int2 c = child[p];
      float4 a0, b0, a1, b1;
      child_box(c.x, nlo, nhi, lo, hi, vals, a0, b0, true);
      child_box(c.y, nlo, nhi, lo, hi, vals, a1, b1, true);
      nlo[p] = make_float4(fminf(a0.x, a1.x), fminf(a0.y, a1.y), fminf(a0.z, a1.z), 0.0f);
      nhi[p] = make_float4(fmaxf(b0.x, b1.x), fmaxf(b0.y, b1.y), fmaxf(b0.z, b1.z), 0.0f);
      p = parent[p];
   }
}

// Internal nodes kept as pair nodes: those over more than kLeafMax primitives. Also the
// depth of the deepest leaf: a kept node with a leaf child, one level below it.
__global__ __launch_bounds__(kBlock) void k_keep(const int2* child, const int2* range, const int* parent, int n,
                                                 int* keep, int* maxd) {
   const int i = blockIdx.x * kBlock + threadIdx.x;
   if (i >= n - 1) return;
   const int2 r = range[i];
   const bool k = r.y - r.x + 1 > kLeafMax;
   keep[i] = k ? 1 : 0;
   if (!k) return;
   const int2 c = child[i];
   auto kept = [&](int x) { return x >= 0 && range[x].y - range[x].x + 1 > kLeafMax; };
   if (kept(c.x) && kept(c.y)) return;
   int depth = 0;
   for (int p = parent[i]; p >= 0; p = parent[p]) ++depth;
   atomicMax(maxd, depth + 1);
}

__global__ __launch_bounds__(kBlock) void k_emit(const int2* child, const int2* range, const int* keep,
                                                 const int* index, int n, const float4* lo, const float4* hi,
                                                 const int* vals, const float4* nlo, const float4* nhi, Node2* out) {
   const int i = blockIdx.x * kBlock + threadIdx.x;
   if (i >= n - 1 || !keep[i]) return;
   const int2 c = child[i];
   Node2 node;
   float* bx = &node.bx.x;
   float* by = &node.by.x;
   float* bz = &node.bz.x;
   int* meta = &node.meta.x;
   const int cs[2] = {c.x, c.y};
#pragma unroll
   for (int w = 0; w < 2; ++w) {
      float4 a, b;
      child_box(cs[w], nlo, nhi, lo, hi, vals, a, b, false);
      bx[2 * w] = a.x;
      bx[2 * w + 1] = b.x;
      by[2 * w] = a.y;
      by[2 * w + 1] = b.y;
      bz[2 * w] = a.z;
      bz[2 * w + 1] = b.z;
      if (cs[w] < 0) {  // single primitive
         meta[w] = ~cs[w];
         meta[2 + w] = 1;
      } else if (keep[cs[w]]) {
         meta[w] = index[cs[w]];
         meta[2 + w] = 0;
      } else {  // small subtree: its contiguous range is one leaf
         const int2 r = range[cs[w]];
         meta[w] = r.x;
         meta[2 + w] = r.y - r.x + 1;
      }
   }
   out[index[i]] = node;
}

__global__ __launch_bounds__(kBlock) void k_permute(const PrimRec* src, const float4* nsrc, const int* vals, int n,
                                                    PrimRec* dst, float4* ndst) {
   const int s = blockIdx.x * kBlock + threadIdx.x;
   if (s >= n) return;
   const int p = vals[s];
   dst[s] = src[p];
   if (nsrc) {
      ndst[3 * s] = nsrc[3 * p];
      ndst[3 * s + 1] = nsrc[3 * p + 1];
      ndst[3 * s + 2] = nsrc[3 * p + 2];
   }
}

template <class T>
hipError_t salloc(T** p, size_t count, hipStream_t st) {
   return hipMallocAsync((void**)p, std::max<size_t>(count, 1) * sizeof(T), st);
}

}  // namespace

hipError_t launch_permute_recs(const PrimRec* src, const float4* nsrc, const int* perm, int n, PrimRec* dst,
                               float4* ndst, hipStream_t st) {
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_permute, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, src, nsrc, perm, n,
                      dst, ndst);
   return hipGetLastError();
}

rayon_status build_bvh_device(rayon_ctx* ctx, const float4* d_lo, const float4* d_hi, const PrimRec* d_recs,
                              const float4* d_normals, int n, DeviceTree& out) {
   hipStream_t st = ctx->stream;
   out = DeviceTree{};
   if (n <= 0) {
      out.root = kSentinel;
      return RAYON_OK;
   }
   if (n > (1 << 30)) {  // positions and range searches are 32-bit
      set_error("device BVH build: %d primitives exceed 2^30", n);
      return RAYON_ERR_INVALID;
   }
   uint64_t *keys = nullptr, *keys2 = nullptr;
   int *vals = nullptr, *vals2 = nullptr, *parent = nullptr, *keep = nullptr, *index = nullptr, *small = nullptr;
   int2 *child = nullptr, *range = nullptr;
   float4 *nlo = nullptr, *nhi = nullptr;
   unsigned* arrived = nullptr;
   void* temp = nullptr;
   size_t temp_bytes = 0, t2 = 0;
   const int ni = std::max(n - 1, 1);
   const unsigned g = (unsigned)((n + kBlock - 1) / kBlock);
   hipError_t e = salloc(&keys, n, st);
   if (e == hipSuccess) e = salloc(&keys2, n, st);
   if (e == hipSuccess) e = salloc(&vals, n, st);
   if (e == hipSuccess) e = salloc(&vals2, n, st);
   if (e == hipSuccess) e = salloc(&parent, (size_t)2 * n, st);
   if (e == hipSuccess) e = salloc(&child, ni, st);
   if (e == hipSuccess) e = salloc(&range, ni, st);
   if (e == hipSuccess) e = salloc(&keep, ni, st);
   if (e == hipSuccess) e = salloc(&index, ni, st);
   if (e == hipSuccess) e = salloc(&nlo, ni, st);
   if (e == hipSuccess) e = salloc(&nhi, ni, st);
   if (e == hipSuccess) e = salloc(&arrived, ni, st);
   if (e == hipSuccess) e = salloc(&small, 8, st);  // centroid bounds (6), deepest leaf
   if (e == hipSuccess)
      e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys2, vals, vals2, n, 0, 63, st);
   if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, t2, keep, index, ni, st);
   if (e == hipSuccess) e = hipMallocAsync(&temp, std::max(temp_bytes, t2), st);
   if (e == hipSuccess) {
      hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, st, small, small + 6);
      hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min(g, 1024u)), dim3(kBlock), 0, st, d_lo, d_hi, n, small);
      hipLaunchKernelGGL(k_morton, dim3(g), dim3(kBlock), 0, st, d_lo, d_hi, n, small, keys, vals);
      temp_bytes = std::max(temp_bytes, t2);
      e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys2, vals, vals2, n, 0, 63, st);
   }
   int total = 0, maxd = 0;
   if (e == hipSuccess && n > 1) {
      const unsigned gi = (unsigned)((n - 1 + kBlock - 1) / kBlock);
      hipLaunchKernelGGL(k_radix_tree, dim3(gi), dim3(kBlock), 0, st, keys2, n, child, range, parent);
      e = launch_fill_u32(arrived, 0u, ni, st);
      hipLaunchKernelGGL(k_refit, dim3(g), dim3(kBlock), 0, st, child, parent, n, d_lo, d_hi, vals2, nlo, nhi,
                         arrived);
      hipLaunchKernelGGL(k_keep, dim3(gi), dim3(kBlock), 0, st, child, range, parent, n, keep, small + 6);
      size_t tb = temp_bytes;
      e = hipcub::DeviceScan::ExclusiveSum(temp, tb, keep, index, n - 1, st);
      int last[2] = {0, 0};
      if (e == hipSuccess) e = hipMemcpyAsync(&last[0], index + (n - 2), sizeof(int), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipMemcpyAsync(&last[1], keep + (n - 2), sizeof(int), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipMemcpyAsync(&maxd, small + 6, sizeof(int), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      total = last[0] + last[1];
   }
   rayon_status status = RAYON_OK;
   if (e == hipSuccess) {
      out.max_depth = maxd;
      out.num_nodes = total;
      if (total == 0) {
         out.root = pack_leaf(0, n);  // at most kLeafMax primitives: the root is one leaf
      } else {
         out.root = 0;
         e = hipMalloc((void**)&out.nodes, sizeof(Node2) * (size_t)total);
         if (e == hipSuccess)
            hipLaunchKernelGGL(k_emit, dim3((n - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, st, child, range, keep,
                               index, n, d_lo, d_hi, vals2, nlo, nhi, out.nodes);
      }
   }
   if (e == hipSuccess) e = hipMalloc((void**)&out.recs, sizeof(PrimRec) * (size_t)n);
   if (e == hipSuccess && d_normals) e = hipMalloc((void**)&out.normals, sizeof(float4) * 3 * (size_t)n);
   if (e == hipSuccess)
      hipLaunchKernelGGL(k_permute, dim3(g), dim3(kBlock), 0, st, d_recs, d_normals, vals2, n, out.recs,
                         out.normals);
   if (e == hipSuccess) {
      out.order.resize(n);
      e = hipMemcpyAsync(out.order.data(), vals2, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st);
   }
   if (e == hipSuccess) e = hipGetLastError();
   if (e == hipSuccess) e = hipStreamSynchronize(st);
   if (e != hipSuccess) status = hip_fail(e, "device BVH build");
   void* tmp[] = {keys, keys2, vals, vals2, parent, child, range, keep, index, nlo, nhi, arrived, small, temp};
   for (void* p : tmp)
      if (p) (void)hipFreeAsync(p, st);
   (void)hipStreamSynchronize(st);
   if (status != RAYON_OK) out.release();
   return status;
}

}  // namespace ry
