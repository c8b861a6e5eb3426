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

// State of one path between bounces.
struct PathState {
   V3 o, d, T, L;
   bool prev_lambert;
};

// A shadow ray and the radiance it carries if the segment is unoccluded.
struct ShadowReq {
   V3 o, d, contrib;
   bool valid;
};

// Products rounded separately, so the contribution is the same value whether it is
// added at once (plain path) or by a later shadow stage (wavefront).
__device__ __forceinline__ V3 mul_rn(V3 a, V3 b) {
   return mk(__fmul_rn(a.x, b.x), __fmul_rn(a.y, b.y), __fmul_rn(a.z, b.z));
}

// Emitter sample for a Lambertian hit (spec addition): picks an emitter uniformly, a
// point uniformly on it, and forms the shadow ray and its unoccluded contribution
// T * f_r * Le * cos_s * cos_l / d^2 / pdf.
template <bool CNT>
__device__ inline void nee_request(const JobView& J, const Surf& s, V3 color, V3 T, Rng& R, ShadowReq& sh,
                                   PathCounts& pc) {
   const SceneView& S = J.scene;
   float ul = R.next(), ua = R.next(), ub = R.next();
   int li = min((int)(ul * (float)S.num_lights), S.num_lights - 1);
   const LightRec lr = gload(S.lights + li);
   const PrimRec& lp = lr.g;
   if constexpr (CNT) ++pc.light_prims;
   V3 q, nl;
   bool one_sided;
   sample_emitter(lp, ua, ub, q, nl, one_sided);
   sh.o = s.p + 1e-4f * s.n;
   sh.d = q - sh.o;
   float d2 = len2(sh.d);
   V3 wi = (1.0f / sqrtf(d2)) * sh.d;
   float cs = dot(s.n, wi);
   float cl = dot(nl, -wi);
   if (!one_sided) cl = fabsf(cl);
   sh.valid = cs > 0.0f && cl > 0.0f;
   if (sh.valid) {
      ++pc.shadow;
      float g = cs * cl / d2 * lr.e.w * (float)S.num_lights * (1.0f / kPiF);
      sh.contrib = mul_rn(T, (g * J.light_intensity) * (color * mk(lr.e)));
   }
}

// One bounce of ray_color (cuda_raytracer.cuh:869-931) given the closest hit of ps's ray:
// sky on a miss, invisible pass-through at bounce 0, emission, scattering, an optional
// shadow request (NEE on Lambertian surfaces) and Russian roulette from bounce 1.
// Returns true while the path continues with ps.o / ps.d. Shared by the megakernel and
// the wavefront shade stage so both compute identical per-sample radiance.
template <bool CNT, uint32_t FT = SF_ALL>
__device__ inline bool shade_step(const JobView& J, const HitMin& h, int bounce, PathState& ps, Rng& R,
                                  ShadowReq& sh, PathCounts& pc, ShadeField fld = ShadeField{}) {
   const SceneView& S = J.scene;
   const bool nee = (J.flags & RAYON_FLAG_NEE) && S.num_lights > 0;
   sh.valid = false;
   if (h.slot < 0) {
      ps.L = ps.L + ps.T * sky(J, ps.d);
      return false;
   }
   const Surf s = surface<FT>(S, h, ps.o, ps.d, fld);
   if (!(s.info & F_VISIBLE) && bounce == 0) {  // cuda_raytracer.cuh:879-883
      ps.o = s.p + 0.01f * ps.d;
      return true;
   }
   const MatRec M = gload(S.mats + info_mat(s.info));
   if constexpr (CNT) ++pc.mats;
   const V3 color = surface_color<FT>(M, s.p, s.center);
   Scatter sc;
   scatter<Rng, FT>(M, color, J, s, ps.d, R, sc);
   // An emitter that NEE samples (F_EMITTER) was already counted by the shadow ray of the
   // Lambertian bounce that led here; every other emitter is reached only this way.
   if (len2(sc.emitted) > 0.0f && !(nee && ps.prev_lambert && (s.info & F_EMITTER))) ps.L = ps.L + ps.T * sc.emitted;
   if (!sc.scattered) return false;
   if (nee && sc.lambert) nee_request<CNT>(J, s, color, ps.T, R, sh, pc);
   ps.T = ps.T * sc.atten;
   ps.prev_lambert = sc.lambert;
   ps.o = sc.o;
   ps.d = sc.d;
   if (bounce > 0) {  // Russian roulette, cuda_raytracer.cuh:906-920
      float q = fminf(maxc(ps.T), 0.95f);
      if (R.next() > q) return false;
      ps.T = (1.0f / q) * ps.T;
   }
   return true;
}

template <int STRIDE, bool CNT = false, int LAYOUT = L_ANY>
__device__ inline V3 trace_path(const JobView& J, V3 o, V3 d, Rng& R, int* stack, PathCounts& pc) {
   PathState ps{o, d, mk(1.0f, 1.0f, 1.0f), mk(0.0f, 0.0f, 0.0f), false};
   ShadowReq sh;
   for (int bounce = 0; bounce < J.max_depth; ++bounce) {
      ++pc.segments;
      const HitMin h =
          closest_hit<STRIDE, CNT, LAYOUT>(J.scene, ps.o, ps.d, 0.001f, 3.402823466e38f, stack, &pc.trav);
      const bool cont = shade_step<CNT>(J, h, bounce, ps, R, sh, pc);
      if (sh.valid && !any_hit<STRIDE, CNT, LAYOUT>(J.scene, sh.o, sh.d, 1e-4f, 1.0f - 1e-4f, stack, &pc.trav))
         ps.L = ps.L + sh.contrib;
      if (!cont) break;
   }
   return ps.L;
}

// Adds one sample's radiance to a pixel's compensated sums.
__device__ __forceinline__ void accumulate(float4& a, float4& c, V3 L) {
   ff_add(a.x, c.x, L.x);
   ff_add(a.y, c.y, L.y);
   ff_add(a.z, c.z, L.z);
   float l = luminance(L);
   // the square is rounded on its own (no fused multiply-add into the sum), so the
   // compensation term sees the same addend as the sum
   ff_add(a.w, c.w, __fmul_rn(l, l));
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

__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
   for (int off = 32; off > 0; off >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, off));
   return v;
}

}  // namespace ry
