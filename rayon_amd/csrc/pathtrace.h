// pathtrace.h — one camera sample's path (ray_color, cuda_raytracer.cuh:858-933) with
// next-event estimation on Lambertian surfaces as an addition.
#pragma once

#include "device.h"

namespace ry {

struct PathCounts {
   int segments = 0;
   int shadow = 0;
   TravCount trav;          // filled only by the counting instantiation
   unsigned mats = 0;       // material records fetched
   unsigned light_prims = 0;
};

// Emitter sample + shadow ray for a Lambertian hit; returns the contribution before the
// path throughput is applied.
template <int STRIDE, bool CNT = false>
__device__ inline V3 nee_lambert(const JobView& J, const Surf& s, V3 color, Rng& R, int* stack, PathCounts& pc) {
   const SceneView& S = J.scene;
   float ul = R.next(), ua = R.next(), ub = R.next();
   int li = min((int)(ul * (float)S.num_lights), S.num_lights - 1);
   const int slot = S.light_slots[li];
   const PrimRec lp = S.prims[slot];
   if constexpr (CNT) ++pc.light_prims;
   V3 q, nl;
   bool one_sided;
   sample_emitter(lp, ua, ub, q, nl, one_sided);
   V3 org = s.p + 1e-4f * s.n;
   V3 wd = q - org;
   float d2 = len2(wd);
   V3 wi = (1.0f / sqrtf(d2)) * wd;
   float cs = dot(s.n, wi);
   float cl = dot(nl, -wi);
   if (!one_sided) cl = fabsf(cl);
   V3 out = mk(0.0f, 0.0f, 0.0f);
   if (cs > 0.0f && cl > 0.0f) {
      ++pc.shadow;
      if (!any_hit<STRIDE, CNT>(S, org, wd, 1e-4f, 1.0f - 1e-4f, stack, &pc.trav)) {
         const MatRec LM = S.mats[info_mat(__float_as_uint(lp.a.w))];
         if constexpr (CNT) ++pc.mats;
         float g = cs * cl / d2 * S.light_area[li] * (float)S.num_lights * (1.0f / kPiF);
         out = (g * J.light_intensity) * (color * mk(LM.m1));
      }
   }
   return out;
}

template <int STRIDE, bool CNT = false>
__device__ inline V3 trace_path(const JobView& J, V3 o, V3 d, Rng& R, int* stack, PathCounts& pc) {
   const SceneView& S = J.scene;
   const bool nee = (J.flags & RAYON_FLAG_NEE) && S.num_lights > 0;
   V3 L = mk(0.0f, 0.0f, 0.0f), T = mk(1.0f, 1.0f, 1.0f);
   bool prev_lambert = false;
   for (int bounce = 0; bounce < J.max_depth; ++bounce) {
      ++pc.segments;
      const HitMin h = closest_hit<STRIDE, CNT>(S, o, d, 0.001f, 3.402823466e38f, stack, &pc.trav);
      if (h.slot < 0) {
         L = L + T * sky(J, d);
         break;
      }
      const Surf s = surface(S, h, o, d);
      if (!(s.info & F_VISIBLE) && bounce == 0) {  // cuda_raytracer.cuh:879-883
         o = s.p + 0.01f * d;
         continue;
      }
      const MatRec M = S.mats[info_mat(s.info)];
      if constexpr (CNT) ++pc.mats;
      const V3 color = surface_color(M, s.p, s.center);
      Scatter sc;
      scatter(M, color, J, s, d, R, sc);
      if (len2(sc.emitted) > 0.0f && !(nee && prev_lambert)) L = L + T * sc.emitted;
      if (!sc.scattered) break;
      if (nee && sc.lambert) L = L + T * nee_lambert<STRIDE, CNT>(J, s, color, R, stack, pc);
      T = T * sc.atten;
      prev_lambert = sc.lambert;
      o = sc.o;
      d = sc.d;
      if (bounce > 0) {  // Russian roulette, cuda_raytracer.cuh:906-920
         float q = fminf(maxc(T), 0.95f);
         if (R.next() > q) break;
         T = (1.0f / q) * T;
      }
   }
   return L;
}

// Adds one sample's radiance to a pixel's compensated sums.
__device__ __forceinline__ void accumulate(float4& a, float4& c, V3 L) {
   neumaier(a.x, c.x, L.x);
   neumaier(a.y, c.y, L.y);
   neumaier(a.z, c.z, L.z);
   float l = luminance(L);
   neumaier(a.w, c.w, l * l);
}

// New signed per-pixel count after a pass: render_acc_kernel.cu:240-271 (REF) or the
// relative standard error of the mean luminance (STDERR).
__device__ __forceinline__ int adaptive_count(const JobView& J, float4 a_before, int n_before, float4 a, float4 c,
                                              int n) {
   if (J.flags & RAYON_FLAG_ADAPTIVE_REF) {
      if (n < J.min_adaptive) return n;
      float lb = 0.0f;
      if (n_before > 0) {
         float inv = 1.0f / (float)n_before;
         lb = 0.299f * a_before.x * inv + 0.587f * a_before.y * inv + 0.114f * a_before.z * inv;
      }
      float inv = 1.0f / (float)n;
      float la = 0.299f * a.x * inv + 0.587f * a.y * inv + 0.114f * a.z * inv;
      float rel = fabsf(la - lb) / fmaxf(la, 0.001f);
      return rel < J.adaptive_threshold ? -n : n;
   }
   if (J.flags & RAYON_FLAG_ADAPTIVE_STDERR) {
      if (n < J.min_adaptive || n < 2) return n;
      float s1 = luminance(mk(a.x + c.x, a.y + c.y, a.z + c.z));
      float s2 = a.w + c.w;
      float mean = s1 / (float)n;
      float var = fmaxf(0.0f, s2 / (float)n - mean * mean) * ((float)n / (float)(n - 1));
      float se = sqrtf(var / (float)n);
      return se / fmaxf(mean, 1e-3f) < J.adaptive_threshold ? -n : n;
   }
   return n;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
   for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
   return v;
}

}  // namespace ry
