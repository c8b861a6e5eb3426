// device.h — per-ray device code shared by the plain megakernel and the wavefront
// stages: RNG streams, BVH traversal with an LDS stack, primitive intersection, surface
// reconstruction, material scattering and light sampling. FP32 throughout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rayon_hip.h"
#include "layout.h"

namespace ry {

constexpr int kStack = 32;            // traversal stack entries per lane (binary tree: 31 in LDS + 1 register)
constexpr float kPiF = 3.14159265358979f;

// ------------------------------------------------------------------------------------
// small vector type
// ------------------------------------------------------------------------------------
struct V3 {
   float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 mk(const float4& a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
   return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len2(V3 a) { return dot(a, a); }
__device__ __forceinline__ V3 unit(V3 a) { return (1.0f / sqrtf(dot(a, a))) * a; }
__device__ __forceinline__ float maxc(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ __forceinline__ V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

// Scene arrays are device allocations, but their pointers are read from memory (the job
// table), so the compiler can only type them as generic addresses and emits flat loads.
// Flat loads count against the LDS counter too: every wait for a node or record also
// waits for the LDS traversal stack, and vice versa. Scene records are read with this
// load through the global address space instead (16-byte vectors where the record allows).
typedef float gvec4 __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ T gload(const T* p) {
   T v;
   if constexpr (sizeof(T) % 16 == 0 && alignof(T) >= 16) {
      constexpr int n = (int)(sizeof(T) / 16);
      const __attribute__((address_space(1))) gvec4* s = (const __attribute__((address_space(1))) gvec4*)p;
      gvec4 w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = s[k];
      __builtin_memcpy(&v, w, sizeof(T));
   } else {
      static_assert(sizeof(T) % 4 == 0, "scene records are whole 32-bit words");
      constexpr int n = (int)(sizeof(T) / 4);
      const __attribute__((address_space(1))) uint32_t* s = (const __attribute__((address_space(1))) uint32_t*)p;
      uint32_t w[n];
#pragma unroll
      for (int k = 0; k < n; ++k) w[k] = s[k];
      __builtin_memcpy(&v, w, sizeof(T));
   }
   return v;
}

// Square root by the hardware instruction alone (v_sqrt_f32, within 1 ulp). sqrtf adds a
// correction sequence for correct rounding (about 15 instructions); the SDF distance
// evaluation, which runs two square roots per primitive per marching step, uses this one.
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// sin/cos of 2*pi*u for u in [0,1]: v_sin_f32 / v_cos_f32 take revolutions.
__device__ __forceinline__ float sin2pi(float u) { return __builtin_amdgcn_sinf(u); }
__device__ __forceinline__ float cos2pi(float u) { return __builtin_amdgcn_cosf(u); }

// ------------------------------------------------------------------------------------
// RNG: the reference's LCG + output permutation (cuda_utils.cuh:24-36). REF mode keeps
// one sequential stream per pixel (seeded as cuda_utils.cu:26) and converts with
// float(r)*2^-32; the default counter mode restarts the stream per (seed, pixel, sample)
// and converts the top 24 bits, so each sample is independent of every other.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lcg_step(uint32_t s) { return s * 747796405u + 2891336453u; }
__device__ __forceinline__ uint32_t permute(uint32_t s) {
   uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
   return (w >> 22u) ^ w;
}
__device__ __forceinline__ uint32_t hash32(uint32_t v) { return permute(lcg_step(v)); }

__device__ __forceinline__ uint32_t counter_init(uint32_t seed_lo, uint32_t seed_hi, uint32_t pixel,
                                                 uint32_t sample) {
   uint32_t h = hash32(seed_lo ^ hash32(seed_hi + 0x9E3779B9u));
   h = hash32(h + pixel);
   return hash32(h ^ (sample * 0x85EBCA6Bu + 0xC2B2AE35u));
}

struct Rng {
   uint32_t s;
   bool ref;
   __device__ __forceinline__ float next() {
      s = lcg_step(s);
      uint32_t r = permute(s);
      if (ref) return (float)r * (1.0f / 4294967296.0f);
      return (float)(r >> 8) * (1.0f / 16777216.0f);
   }
};

// ------------------------------------------------------------------------------------
// traversal
// ------------------------------------------------------------------------------------
struct RayInv {
   V3 o, d, inv;
};

__device__ __forceinline__ float safe_inv(float d) {
   float a = fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
   return 1.0f / a;
}

__device__ __forceinline__ RayInv make_rayinv(V3 o, V3 d) {
   RayInv r;
   r.o = o;
   r.d = d;
   r.inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
   return r;
}

// Slab test of both children of a pair node; returns entry distances, hit flags. Each slab
// distance is (b - o) * inv: two roundings and the reciprocal's, so it is within 3 * 2^-24
// relative of the exact distance whatever the magnitudes of b and o (the form b * inv -
// o * inv is off by an ulp of o * inv, which dwarfs a short distance from a far origin).
__device__ __forceinline__ void box2(const RayInv& r, const Node2& n, float tmin, float tmax, float& tn0,
                                     float& tn1, bool& h0, bool& h1) {
   float ax0 = (n.bx.x - r.o.x) * r.inv.x, bx0 = (n.bx.y - r.o.x) * r.inv.x;
   float ay0 = (n.by.x - r.o.y) * r.inv.y, by0 = (n.by.y - r.o.y) * r.inv.y;
   float az0 = (n.bz.x - r.o.z) * r.inv.z, bz0 = (n.bz.y - r.o.z) * r.inv.z;
   float ax1 = (n.bx.z - r.o.x) * r.inv.x, bx1 = (n.bx.w - r.o.x) * r.inv.x;
   float ay1 = (n.by.z - r.o.y) * r.inv.y, by1 = (n.by.w - r.o.y) * r.inv.y;
   float az1 = (n.bz.z - r.o.z) * r.inv.z, bz1 = (n.bz.w - r.o.z) * r.inv.z;
   tn0 = fmaxf(fmaxf(fminf(ax0, bx0), fminf(ay0, by0)), fmaxf(fminf(az0, bz0), tmin));
   float tf0 = fminf(fminf(fmaxf(ax0, bx0), fmaxf(ay0, by0)), fminf(fmaxf(az0, bz0), tmax));
   tn1 = fmaxf(fmaxf(fminf(ax1, bx1), fminf(ay1, by1)), fmaxf(fminf(az1, bz1), tmin));
   float tf1 = fminf(fminf(fmaxf(ax1, bx1), fmaxf(ay1, by1)), fminf(fmaxf(az1, bz1), tmax));
   // The exit (or the current closest hit) is widened by 2^-16 relative, so neither the
   // rounding of the slab distances (3 * 2^-24) nor that of a primitive's own hit distance
   // (Moller-Trumbore's relative error grows as 1 / sin of the angle between ray and
   // triangle: 2^-16 covers angles down to about half a degree) rejects a box holding a
   // primitive whose computed hit is no farther than the closest so far. The set of
   // primitives a closest-hit traversal tests then does not depend on the order it meets
   // them in (postponed leaves make that order depend on the other rays of the wave, whose
   // composition depends on the timing of the queue claims), so neither does its result.
   h0 = tn0 <= tf0 * 1.0000153f;
   h1 = tn1 <= tf1 * 1.0000153f;
}

struct HitMin {
   float t;
   int slot;
   int orig;
   float u, v;
   int sdf;  // slot of an SDF field left for a later marching stage (-1: none)
};

// Moller-Trumbore on (v0, e1, e2), the arithmetic of hit_triangle
// (cuda_raytracer.cuh:277-304) with the inclusive [tmin, tmax] range.
__device__ __forceinline__ bool isect_tri(const PrimRec& p, V3 o, V3 d, float tmin, float tmax, float& t, float& u,
                                          float& v) {
   V3 v0 = mk(p.a), e1 = mk(p.b), e2 = mk(p.c);
   V3 h = cross(d, e2);
   float a = dot(e1, h);
   if (fabsf(a) < 1e-8f) return false;
   float f = 1.0f / a;
   V3 s = o - v0;
   u = f * dot(s, h);
   if (u < 0.0f || u > 1.0f) return false;
   V3 q = cross(s, e1);
   v = f * dot(d, q);
   if (v < 0.0f || u + v > 1.0f) return false;
   t = f * dot(e2, q);
   return t >= tmin && t <= tmax;
}

// |oc|^2 - r^2 with error-free FMA products and compensated sums (Ogita-Rump-Oishi
// Dot2): for a large sphere seen from near its surface (the r=1000 ground) the direct
// FP32 form cancels ~16 of 24 bits.
__device__ __forceinline__ float sphere_c(V3 f, float r) {
   float p = f.x * f.x;
   float s = fmaf(f.x, f.x, -p);
   float terms[3] = {f.y, f.z, r};
#pragma unroll
   for (int i = 0; i < 3; ++i) {
      const float x = terms[i];
      const float y = i < 2 ? x : -x;
      const float h = x * y;
      const float e = fmaf(x, y, -h);
      const float sum = p + h;
      const float bb = sum - p;
      const float q = (p - (sum - bb)) + (h - bb);
      p = sum;
      s += q + e;
   }
   return p + s;
}

// hit_sphere's root selection (cuda_raytracer.cuh:203-222): the nearer root if it lies in
// [tmin, tmax], else the farther one. Same roots as the reference's expressions in exact
// arithmetic, formed for FP32 accuracy: c = |oc|^2 - r^2 compensated; the discriminant
// as b^2 - a c with exact products when the origin is within ~1.4 r of the centre (large
// spheres seen from near their surface), else from the perpendicular distance
// a r^2 - |d x oc|^2 (small or distant spheres); roots q/a and c/q with
// q = -b - sign(b) sqrt(disc), which never subtracts nearly equal numbers.
__device__ __forceinline__ bool isect_sphere(V3 c, float rad, V3 o, V3 d, float tmin, float tmax, float& t) {
   const V3 oc = o - c;
   const float a = dot(d, d);
   const float hb = dot(oc, d);
   const float cc = sphere_c(oc, rad);
   float disc;
   if (cc < 0.5f * dot(oc, oc)) {
      const float h2 = hb * hb, ac = a * cc;
      disc = (h2 - ac) + (fmaf(hb, hb, -h2) - fmaf(a, cc, -ac));
   } else {
      const V3 x = cross(d, oc);
      disc = a * (rad * rad) - dot(x, x);
   }
   if (disc < 0.0f) return false;
   const float q = -hb - copysignf(sqrtf(disc), hb);
   float r0 = q / a, r1 = cc / q;
   if (q == 0.0f) r0 = r1 = 0.0f;
   const float tn = fminf(r0, r1), tf = fmaxf(r0, r1);
   float root = tn;
   if (root < tmin || tmax < root) {
      root = tf;
      if (root < tmin || tmax < root) return false;
   }
   t = root;
   return true;
}

// hit_rectangle (cuda_raytracer.cuh:235-262).
__device__ __forceinline__ bool isect_rect(V3 q, V3 u, V3 v, V3 o, V3 d, float tmin, float tmax, float& t) {
   V3 n = unit(cross(u, v));
   float denom = dot(n, d);
   if (fabsf(denom) < 1e-8f) return false;
   t = dot(n, q - o) / denom;
   if (t < tmin || t > tmax) return false;
   V3 rel = (o + t * d) - q;
   float al = dot(rel, u) / dot(u, u);
   float be = dot(rel, v) / dot(v, v);
   return !(al < 0.0f || al > 1.0f || be < 0.0f || be > 1.0f);
}

// ---- SDF field: smooth union of analytic primitives (sdf.hpp:105-143, :410-414)
// Every primitive is evaluated by the one rounded-box formula (no per-type branch, so the
// lanes of a wave stay together whatever primitive each one reaches): a sphere is a box of
// zero half-extent rounded by r; a torus is the same applied to (|xz| - R, y, 0).
__device__ __forceinline__ float sdf_prim(const SdfRec& s, V3 p) {
   const V3 o = p - mk(s.c);
   const float lxz = hw_sqrt(o.x * o.x + o.z * o.z);
   const bool torus = s.c.w >= 0.0f;
   const float ax = torus ? lxz - s.c.w : o.x, ay = o.y, az = torus ? 0.0f : o.z;
   const float qx = fabsf(ax) - s.p.x, qy = fabsf(ay) - s.p.y, qz = fabsf(az) - s.p.z;
   const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
   return hw_sqrt(mx * mx + my * my + mz * mz) + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f) - s.p.w;
}

__device__ __forceinline__ float smooth_union(float d, float d2, float k, float ik) {
   float h = fminf(fmaxf(0.5f + 0.5f * (d2 - d) * ik, 0.0f), 1.0f);
   return d2 * (1.0f - h) + d * h - k * h * (1.0f - h);
}

// The whole field: primitives combined in index order.
__device__ __forceinline__ float sdf_field(const SceneView& S, V3 p) {
   float d = sdf_prim(gload(S.sdf), p);
   const float k = S.sdf_k, ik = 1.0f / S.sdf_k;
   for (int i = 1; i < S.num_sdf; ++i) d = smooth_union(d, sdf_prim(gload(S.sdf + i), p), k, ik);
   return d;
}

// Fields of up to 64 primitives are evaluated over the primitives near the ray only. A
// primitive whose bounding sphere stays farther than `margin` from every point considered
// has a distance >= margin there; in a smooth union a term that exceeds the running value
// by at least k leaves it unchanged (h = 1), so near a surface (field value << margin) the
// subset gives the same field as all primitives unless dozens of far primitives blend with
// each other first. Marching steps are capped below the far primitives' distance bound.
constexpr int kSdfMaskMax = 64;

__device__ __forceinline__ float sdf_margin(const SceneView& S) { return fmaxf(10.0f * S.sdf_k, 1.0f); }

// Where the field's primitive records are read from: the scene's array in device memory,
// or a copy in the workgroup's LDS (a marching kernel stages the records of a pass whose
// fields all belong to one scene). Both hold the same values, so the field is the same.
struct SdfGlobalRecs {
   const SdfRec* p;
   __device__ __forceinline__ SdfRec operator()(int i) const { return gload(p + i); }
};
struct SdfLdsRecs {
   const __attribute__((address_space(3))) gvec4* p;  // LDS, two vectors per record: ds_read, not flat loads
   __device__ __forceinline__ SdfRec operator()(int i) const {
      const gvec4 w[2] = {p[2 * i], p[2 * i + 1]};
      SdfRec r;
      __builtin_memcpy(&r, w, sizeof(r));
      return r;
   }
};

// The bounding spheres, from the scene or from an LDS copy.
struct SdfGlobalSph {
   const float4* p;
   __device__ __forceinline__ float4 operator()(int i) const { return gload(p + i); }
};
struct SdfLdsSph {
   const __attribute__((address_space(3))) gvec4* p;
   __device__ __forceinline__ float4 operator()(int i) const {
      const gvec4 w = p[i];
      return make_float4(w.x, w.y, w.z, w.w);
   }
};

template <class REC>
__device__ __forceinline__ float sdf_field_masked(const SceneView& S, V3 p, uint64_t mask, REC rec) {
   const float k = S.sdf_k, ik = 1.0f / S.sdf_k;
   int i = __builtin_ctzll(mask);
   float d = sdf_prim(rec(i), p);
   for (mask &= mask - 1; mask; mask &= mask - 1) {
      i = __builtin_ctzll(mask);
      d = smooth_union(d, sdf_prim(rec(i), p), k, ik);
   }
   return d;
}

__device__ __forceinline__ float sdf_field_masked(const SceneView& S, V3 p, uint64_t mask) {
   return sdf_field_masked(S, p, mask, SdfGlobalRecs{S.sdf});
}

// Primitives within `margin` of the segment o + t*dn, t in [t0, t1]; `far` receives the
// smallest distance bound of the others. Each primitive is tested through its bounding
// sphere (S.sdf_sph: centre, radius), comparing squared distances with the sphere grown
// by the margin, so only the distance bounds of the excluded primitives take a root.
// SPH: the bounding spheres, from the scene (a float4 pointer in device memory) or from a
// copy in LDS (a pointer in address space 3; every lane reads the same sphere at once).
template <class SPH>
__device__ __forceinline__ uint64_t sdf_ray_mask(const SceneView& S, V3 o, V3 dn, float t0, float t1, float margin,
                                                 float& far, SPH sph) {
   uint64_t mask = 0;
   far = 3.4e38f;
   for (int i = 0; i < S.num_sdf; ++i) {
      const float4 b = sph(i);
      const V3 oc = mk(b) - o;
      const float tc = fminf(fmaxf(dot(oc, dn), t0), t1);
      const V3 q = tc * dn - oc;
      const float d2 = dot(q, q), r = margin + b.w;
      const bool in = d2 <= r * r;
      mask |= in ? 1ull << i : 0ull;
      far = fminf(far, in ? far : hw_sqrt(d2) - b.w);
   }
   return mask;
}

// Field at a point from the primitives within `margin` of it (all of them for large fields).
template <class REC>
__device__ __forceinline__ float sdf_field_near(const SceneView& S, V3 p, uint64_t mask, REC rec) {
   return S.num_sdf > kSdfMaskMax ? sdf_field(S, p) : sdf_field_masked(S, p, mask, rec);
}
__device__ __forceinline__ float sdf_field_near(const SceneView& S, V3 p, uint64_t mask) {
   return sdf_field_near(S, p, mask, SdfGlobalRecs{S.sdf});
}

template <class SPH>
__device__ __forceinline__ uint64_t sdf_point_mask(const SceneView& S, V3 p, SPH sph) {
   if (S.num_sdf > kSdfMaskMax) return ~0ull;
   const float margin = sdf_margin(S);
   uint64_t mask = 0;
   for (int i = 0; i < S.num_sdf; ++i) {
      const float4 b = sph(i);
      const V3 q = p - mk(b);
      const float r = margin + b.w;
      mask |= dot(q, q) <= r * r ? 1ull << i : 0ull;
   }
   return mask ? mask : 1ull;
}
__device__ __forceinline__ uint64_t sdf_point_mask(const SceneView& S, V3 p) {
   return sdf_point_mask(S, p, SdfGlobalSph{S.sdf_sph});
}

// Sphere tracing from the entry of the field's box along the unit direction
// (rayMarch, sdf.hpp:518-563), split into a set-up and a resumable run so the batched path
// can march in slices of steps and regroup the rays still marching between slices.
struct SdfMarch {
   float tw;     // distance marched along the unit direction
   float tend;   // end of the segment
   float cap;    // step cap below the distance bound of the primitives left out
   int used;     // steps taken
   uint64_t mask;
};

// Returns false when the ray cannot meet the field in [tmin, tmax].
template <class SPH>
__device__ __forceinline__ bool sdf_march_begin(const SceneView& S, const PrimRec& p, V3 o, V3 d, float tmin,
                                                float tmax, SdfMarch& m, SPH sph) {
   const float dl = sqrtf(dot(d, d));
   const V3 dn = (1.0f / dl) * d;
   // the march (start, primitive subset, step cap) depends on tmin and the box only, not
   // on tmax, so it takes the same steps whatever closer hit bounds it
   float t0 = tmin * dl, tb = 3.402823466e38f;
   float lo[3] = {p.a.x, p.a.y, p.a.z}, hi[3] = {p.b.x, p.b.y, p.b.z};
   float oo[3] = {o.x, o.y, o.z}, dd[3] = {dn.x, dn.y, dn.z};
#pragma unroll
   for (int k = 0; k < 3; ++k) {
      float inv = safe_inv(dd[k]);
      float a = (lo[k] - oo[k]) * inv, b = (hi[k] - oo[k]) * inv;
      t0 = fmaxf(t0, fminf(a, b));
      tb = fminf(tb, fmaxf(a, b));
   }
   const float t1 = fminf(tb, tmax * dl);
   if (t0 > t1) return false;
   m.mask = ~0ull;
   m.cap = 3.4e38f;
   if (S.num_sdf <= kSdfMaskMax) {
      float far;
      m.mask = sdf_ray_mask(S, o, dn, t0, tb, sdf_margin(S), far, sph);
      if (!m.mask) return false;  // every primitive stays >= margin away: no surface on the segment
      m.cap = far - 2.0f * S.sdf_k;
   }
   m.tw = t0;
   m.tend = t1;
   m.used = 0;
   return true;
}

__device__ __forceinline__ bool sdf_march_begin(const SceneView& S, const PrimRec& p, V3 o, V3 d, float tmin,
                                                float tmax, SdfMarch& m) {
   return sdf_march_begin(S, p, o, d, tmin, tmax, m, SdfGlobalSph{S.sdf_sph});
}

// Up to `steps` more steps. Returns 1 on a hit (t = ray parameter), 0 when the march ended
// without one, -1 when it needs more steps.
// REC: the primitive records of a field of at most kSdfMaskMax primitives (larger fields
// are read from the scene).
template <class REC>
__device__ __forceinline__ int sdf_march_run(const SceneView& S, V3 o, V3 d, SdfMarch& m, int steps, float& t,
                                             REC rec) {
   const float dl = sqrtf(dot(d, d));
   const float idl = 1.0f / dl;
   const V3 dn = idl * d;
   const int stop = min(S.sdf_steps, m.used + steps);
   for (; m.used < stop; ++m.used) {
      const V3 q = o + m.tw * dn;
      const float dist = S.num_sdf > kSdfMaskMax ? sdf_field(S, q) : sdf_field_masked(S, q, m.mask, rec);
      if (dist < S.sdf_eps) {
         t = m.tw * idl;
         return 1;
      }
      m.tw += fminf(dist, m.cap);
      if (m.tw > m.tend) return 0;
   }
   return m.used >= S.sdf_steps ? 0 : -1;
}

__device__ __forceinline__ int sdf_march_run(const SceneView& S, V3 o, V3 d, SdfMarch& m, int steps, float& t) {
   return sdf_march_run(S, o, d, m, steps, t, SdfGlobalRecs{S.sdf});
}

__device__ inline bool isect_sdf(const SceneView& S, const PrimRec& p, V3 o, V3 d, float tmin, float tmax, float& t) {
   SdfMarch m;
   if (!sdf_march_begin(S, p, o, d, tmin, tmax, m)) return false;
   return sdf_march_run(S, o, d, m, S.sdf_steps, t) == 1;
}

// Closest hit over the scene. Exact t ties go to the larger original primitive index,
// which is the linear-scan rule of the reference (cuda_raytracer.cuh:568-585) made
// independent of traversal order. `stack` points at this lane's first LDS entry; entry
// i lives at stack[i * STRIDE].
// Optional traversal statistics (node pairs fetched, primitive records fetched).
struct TravCount {
   unsigned nodes = 0, prims = 0;
};

// Traversal stack of the binary tree: `CAP` entries, the top one in a register and the
// others in LDS (entry i at lds[i * STRIDE]), so a pop hands over the next node at once and
// the LDS read that refills the top overlaps the node fetch. A push beyond CAP is dropped;
// the builders bound the tree depth (the most entries a traversal can hold) by the capacity
// of the kernel that traverses it, so that never happens.
template <int STRIDE, int CAP>
struct BinStack {
   int* lds;
   int n = 0;  // entries held, the top one in `top`
   int top = kSentinel;
   __device__ __forceinline__ explicit BinStack(int* p) : lds(p) {}
   __device__ __forceinline__ void push(int x) {
      if (n >= CAP) return;
      if (n > 0) lds[(n - 1) * STRIDE] = top;
      top = x;
      ++n;
   }
   __device__ __forceinline__ int pop() {
      if (n == 0) return kSentinel;
      const int x = top;
      --n;
      if (n > 0) top = lds[(n - 1) * STRIDE];
      return x;
   }
};

template <int STRIDE, bool CNT = false, bool DEFER = false, int STACK = kStack>
__device__ inline HitMin closest_hit_bin(const SceneView& S, V3 o, V3 d, float tmin, float tmax, int* stack,
                                         TravCount* tc = nullptr) {
   HitMin best;
   best.t = tmax;
   best.slot = -1;
   best.orig = -1;
   best.u = best.v = 0.0f;
   best.sdf = -1;
   int cur = S.root;
   if (cur == kSentinel) return best;
   // the scene's arrays in registers: S may live in memory the compiler cannot prove
   // unchanged, and a reload before each fetch would add a dependent load to every visit
   const Node2* const nodes = S.nodes;
   const PrimRec* const prims = S.prims;
   const RayInv r = make_rayinv(o, d);
   BinStack<STRIDE, STACK> st(stack);
   for (;;) {
      while (cur >= 0) {
         const Node2 n = gload(nodes + cur);
         if constexpr (CNT) ++tc->nodes;
         float tn0, tn1;
         bool h0, h1;
         box2(r, n, tmin, best.t, tn0, tn1, h0, h1);
         int r0 = n.meta.z ? pack_leaf(n.meta.x, n.meta.z) : n.meta.x;
         int r1 = n.meta.w ? pack_leaf(n.meta.y, n.meta.w) : n.meta.y;
         if (h0 && h1) {
            int nr = r0, fr = r1;
            if (tn1 < tn0) {
               nr = r1;
               fr = r0;
            }
            st.push(fr);
            cur = nr;
         } else if (h0) {
            cur = r0;
         } else if (h1) {
            cur = r1;
         } else {
            cur = st.pop();
         }
      }
      if (cur == kSentinel) break;
      const int packed = ~cur;
      const int first = packed >> 3, cnt = (packed & 7) + 1;
      for (int i = first; i < first + cnt; ++i) {
         const PrimRec p = gload(prims + i);
         if constexpr (CNT) ++tc->prims;
         const uint32_t info = __float_as_uint(p.a.w);
         const uint32_t kind = info_kind(info);
         float t = 0.0f, u = 0.0f, v = 0.0f;
         bool hit;
         if (kind == K_TRI)
            hit = isect_tri(p, o, d, tmin, best.t, t, u, v);
         else if (kind == K_SPHERE || kind == K_DISPLACED)
            hit = isect_sphere(mk(p.a), p.b.x, o, d, tmin, best.t, t);
         else if (kind == K_RECT)
            hit = isect_rect(mk(p.a), mk(p.b), mk(p.c), o, d, tmin, best.t, t);
         else if constexpr (DEFER) {
            best.sdf = i;
            hit = false;
         } else {
            hit = isect_sdf(S, p, o, d, tmin, best.t, t);
         }
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
      cur = st.pop();
      if (cur == kSentinel) break;
   }
   return best;
}

// Any hit in [tmin, tmax] by a primitive that is not an emitter (shadow rays).
template <int STRIDE, bool CNT = false, int STACK = kStack>
__device__ inline bool any_hit_bin(const SceneView& S, V3 o, V3 d, float tmin, float tmax, int* stack,
                                   TravCount* tc = nullptr) {
   int cur = S.root;
   if (cur == kSentinel) return false;
   const Node2* const nodes = S.nodes;  // in registers, as in closest_hit_bin
   const PrimRec* const prims = S.prims;
   const RayInv r = make_rayinv(o, d);
   BinStack<STRIDE, STACK> st(stack);
   for (;;) {
      while (cur >= 0) {
         const Node2 n = gload(nodes + cur);
         if constexpr (CNT) ++tc->nodes;
         float tn0, tn1;
         bool h0, h1;
         box2(r, n, tmin, tmax, tn0, tn1, h0, h1);
         int r0 = n.meta.z ? pack_leaf(n.meta.x, n.meta.z) : n.meta.x;
         int r1 = n.meta.w ? pack_leaf(n.meta.y, n.meta.w) : n.meta.y;
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
      }
      if (cur == kSentinel) return false;
      const int packed = ~cur;
      const int first = packed >> 3, cnt = (packed & 7) + 1;
      for (int i = first; i < first + cnt; ++i) {
         const PrimRec p = gload(prims + i);
         if constexpr (CNT) ++tc->prims;
         const uint32_t info = __float_as_uint(p.a.w);
         if (info & F_EMITTER) continue;
         const uint32_t kind = info_kind(info);
         float t, u, v;
         bool hit;
         if (kind == K_TRI)
            hit = isect_tri(p, o, d, tmin, tmax, t, u, v);
         else if (kind == K_SPHERE || kind == K_DISPLACED)
            hit = isect_sphere(mk(p.a), p.b.x, o, d, tmin, tmax, t);
         else if (kind == K_RECT)
            hit = isect_rect(mk(p.a), mk(p.b), mk(p.c), o, d, tmin, tmax, t);
         else
            hit = isect_sdf(S, p, o, d, tmin, tmax, t);
         if (hit) return true;
      }
      cur = st.pop();
      if (cur == kSentinel) return false;
   }
}

// Tests one primitive for the closest hit (shared by both tree layouts).
template <bool CNT, bool DEFER = false>
__device__ __forceinline__ void test_closest(const SceneView& S, const PrimRec* prims, int i, V3 o, V3 d, float tmin,
                                             HitMin& best, TravCount* tc) {
   const PrimRec p = gload(prims + i);
   if constexpr (CNT) ++tc->prims;
   const uint32_t info = __float_as_uint(p.a.w);
   const uint32_t kind = info_kind(info);
   float t = 0.0f, u = 0.0f, v = 0.0f;
   bool hit;
   if (kind == K_TRI)
      hit = isect_tri(p, o, d, tmin, best.t, t, u, v);
   else if (kind == K_SPHERE || kind == K_DISPLACED)
      hit = isect_sphere(mk(p.a), p.b.x, o, d, tmin, best.t, t);
   else if (kind == K_RECT)
      hit = isect_rect(mk(p.a), mk(p.b), mk(p.c), o, d, tmin, best.t, t);
   else if constexpr (DEFER) {
      best.sdf = i;
      hit = false;
   } else {
      hit = isect_sdf(S, p, o, d, tmin, best.t, t);
   }
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

template <bool CNT>
__device__ __forceinline__ bool test_any(const SceneView& S, const PrimRec* prims, int i, V3 o, V3 d, float tmin,
                                         float tmax, TravCount* tc) {
   const PrimRec p = gload(prims + i);
   if constexpr (CNT) ++tc->prims;
   const uint32_t info = __float_as_uint(p.a.w);
   if (info & F_EMITTER) return false;
   const uint32_t kind = info_kind(info);
   float t, u, v;
   if (kind == K_TRI) return isect_tri(p, o, d, tmin, tmax, t, u, v);
   if (kind == K_SPHERE || kind == K_DISPLACED) return isect_sphere(mk(p.a), p.b.x, o, d, tmin, tmax, t);
   if (kind == K_RECT) return isect_rect(mk(p.a), mk(p.b), mk(p.c), o, d, tmin, tmax, t);
   return isect_sdf(S, p, o, d, tmin, tmax, t);
}

// Empty slots of a wide node: bit 7 of a slot's meta byte (never set for a leaf, whose
// offset and count take bits 0..6, nor for an internal child, whose byte is 0). The
// multiply moves bits 7, 15, 23 and 31 of a word to bits 28..31 (the other partial
// products land on distinct lower bits, so nothing carries into them).
__device__ __forceinline__ uint32_t wide_empty_slots(uint32_t m0, uint32_t m1) {
   const uint32_t lo = ((m0 & 0x80808080u) * 0x00204081u) >> 28;
   const uint32_t hi = ((m1 & 0x80808080u) * 0x00204081u) >> 28;
   return lo | (hi << 4);
}

// Moves bit i of an 8-bit slot mask to bit i ^ perm (visiting order of the ray's octant).
__device__ __forceinline__ uint32_t wide_xor_slots(uint32_t x, uint32_t perm) {
   uint32_t y = ((x & 0x55u) << 1) | ((x >> 1) & 0x55u);
   x = (perm & 1u) ? y : x;
   y = ((x & 0x33u) << 2) | ((x >> 2) & 0x33u);
   x = (perm & 2u) ? y : x;
   y = ((x & 0x0Fu) << 4) | (x >> 4);
   return (perm & 4u) ? y : x;
}

// Tests the eight quantized child boxes of a wide node. Returns the hit internal
// children as a mask in visiting order (bit i = slot i ^ perm) and the hit leaf children
// as a slot mask. Each child plane is q * (2^e / d) + (p - o) / d; the exit distance is
// widened by 2^-16 relative plus 2^-21 of the largest plane offset (see box2) so FP32
// rounding never rejects a box the exact test accepts, nor one holding a primitive whose
// computed hit is no farther than the closest so far. Per child the test is six byte
// conversions, six multiply-adds, the entry/exit min-max and one compare into a slot mask;
// which slots are internal, which are empty and the visiting order are applied to the
// whole mask once per node.
__device__ __forceinline__ void wide_children(const WNode& n, V3 o, V3 inv, uint32_t perm, float tmin, float tmax,
                                              uint32_t& hit_int, uint32_t& hit_leaf) {
   const uint32_t em = n.e_imask;
   const float sx = __uint_as_float((em & 0xFFu) << 23);
   const float sy = __uint_as_float(((em >> 8) & 0xFFu) << 23);
   const float sz = __uint_as_float(((em >> 16) & 0xFFu) << 23);
   const float ax = sx * inv.x, ay = sy * inv.y, az = sz * inv.z;
   const float bx = (n.px - o.x) * inv.x, by = (n.py - o.y) * inv.y, bz = (n.pz - o.z) * inv.z;
   const float slack = 4.76837158e-7f * fmaxf(fmaxf(fabsf(bx), fabsf(by)), fabsf(bz));
   // The entry plane of each axis is the lo plane for a positive direction and the hi plane
   // for a negative one: pick the byte words once per node (the slots' planes then need no
   // per-slot min/max to order them). Bytes: lo x/y/z in words 0-1/2-3/4-5, hi in 6-7/8-9/10-11.
   const bool nx = inv.x < 0.0f, ny = inv.y < 0.0f, nz = inv.z < 0.0f;
   const uint32_t nx0 = nx ? n.q[6] : n.q[0], nx1 = nx ? n.q[7] : n.q[1];
   const uint32_t fx0 = nx ? n.q[0] : n.q[6], fx1 = nx ? n.q[1] : n.q[7];
   const uint32_t ny0 = ny ? n.q[8] : n.q[2], ny1 = ny ? n.q[9] : n.q[3];
   const uint32_t fy0 = ny ? n.q[2] : n.q[8], fy1 = ny ? n.q[3] : n.q[9];
   const uint32_t nz0 = nz ? n.q[10] : n.q[4], nz1 = nz ? n.q[11] : n.q[5];
   const uint32_t fz0 = nz ? n.q[4] : n.q[10], fz1 = nz ? n.q[5] : n.q[11];
   uint32_t raw = 0;
#pragma unroll
   for (int sl = 0; sl < 8; ++sl) {
      const int sh = 8 * (sl & 3);  // one byte-to-float conversion per plane (v_cvt_f32_ubyteN)
      const float qnx = (float)(((sl < 4 ? nx0 : nx1) >> sh) & 0xFFu);
      const float qfx = (float)(((sl < 4 ? fx0 : fx1) >> sh) & 0xFFu);
      const float qny = (float)(((sl < 4 ? ny0 : ny1) >> sh) & 0xFFu);
      const float qfy = (float)(((sl < 4 ? fy0 : fy1) >> sh) & 0xFFu);
      const float qnz = (float)(((sl < 4 ? nz0 : nz1) >> sh) & 0xFFu);
      const float qfz = (float)(((sl < 4 ? fz0 : fz1) >> sh) & 0xFFu);
      const float tn = fmaxf(fmaxf(fmaf(qnx, ax, bx), fmaf(qny, ay, by)), fmaxf(fmaf(qnz, az, bz), tmin));
      const float tf = fminf(fminf(fmaf(qfx, ax, bx), fmaf(qfy, ay, by)), fminf(fmaf(qfz, az, bz), tmax));
      raw |= tn <= fmaf(tf, 1.0000153f, slack) ? (1u << sl) : 0u;
   }
   // an empty slot stores lo = 255 > hi = 0 on every axis, an inverted box, and its meta
   // byte has bit 7 set: it is dropped here even where the widening would let it pass
   raw &= ~wide_empty_slots(n.meta[0], n.meta[1]);
   const uint32_t imask = em >> 24;
   hit_leaf = raw & ~imask;
   hit_int = wide_xor_slots(raw & imask, perm);
}

__device__ __forceinline__ WNode load_wnode(const WNode* nodes, uint32_t i) {
   return gload(nodes + i);
}

// Closest hit over the 8-wide quantized tree. The stack holds one entry per partly
// visited node: its first internal child and the remaining hit mask (with the node's
// internal-child mask, to locate a child among its internal siblings).
template <int STRIDE, bool CNT, bool DEFER = false>
__device__ inline HitMin closest_hit_wide(const SceneView& S, V3 o, V3 d, float tmin, float tmax, int* stack,
                                          TravCount* tc) {
   HitMin best;
   best.t = tmax;
   best.slot = -1;
   best.orig = -1;
   best.u = best.v = 0.0f;
   best.sdf = -1;
   const V3 inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
   const uint32_t perm = 7u ^ ((d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u));
   int sp = 0;
   uint32_t node = 0;
   const WNode* const wn = S.wnodes;  // in registers, as in closest_hit_bin
   const PrimRec* const prims = S.prims;
   for (;;) {
      const WNode n = load_wnode(wn, node);
      if constexpr (CNT) ++tc->nodes;
      uint32_t hit_int, hit_leaf;
      wide_children(n, o, inv, perm, tmin, best.t, hit_int, hit_leaf);
      while (hit_leaf) {
         const int sl = __ffs(hit_leaf) - 1;
         hit_leaf &= hit_leaf - 1u;
         const uint32_t m = ((sl < 4 ? n.meta[0] : n.meta[1]) >> (8 * (sl & 3))) & 0xFFu;
         const int first = (int)n.prim_base + (int)(m & 31u);
         const int cnt = (int)(m >> 5) + 1;
         for (int i = first; i < first + cnt; ++i) test_closest<CNT, DEFER>(S, prims, i, o, d, tmin, best, tc);
      }
      const uint32_t imask = n.e_imask >> 24;
      uint32_t base = n.child_base;
      if (!hit_int) {  // continue with the nearest pending child of an ancestor
         if (sp == 0) break;
         base = (uint32_t)stack[(2 * (sp - 1)) * STRIDE];
         const uint32_t bits = (uint32_t)stack[(2 * (sp - 1) + 1) * STRIDE];
         hit_int = bits & 0xFFu;
         const uint32_t im = bits >> 8;
         const int i = __ffs(hit_int) - 1;
         hit_int &= hit_int - 1u;
         const uint32_t sl = (uint32_t)i ^ perm;
         node = base + (uint32_t)__popc(im & ((1u << sl) - 1u));
         if (hit_int)
            stack[(2 * (sp - 1) + 1) * STRIDE] = (int)(hit_int | (im << 8));
         else
            --sp;
         continue;
      }
      const int i = __ffs(hit_int) - 1;
      hit_int &= hit_int - 1u;
      const uint32_t sl = (uint32_t)i ^ perm;
      node = base + (uint32_t)__popc(imask & ((1u << sl) - 1u));
      if (hit_int && sp < kStack / 2) {
         stack[(2 * sp) * STRIDE] = (int)base;
         stack[(2 * sp + 1) * STRIDE] = (int)(hit_int | (imask << 8));
         ++sp;
      }
   }
   return best;
}

template <int STRIDE, bool CNT>
__device__ inline bool any_hit_wide(const SceneView& S, V3 o, V3 d, float tmin, float tmax, int* stack,
                                    TravCount* tc) {
   const V3 inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
   const uint32_t perm = 7u ^ ((d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u));
   int sp = 0;
   uint32_t node = 0;
   const WNode* const wn = S.wnodes;  // in registers, as in closest_hit_bin
   const PrimRec* const prims = S.prims;
   for (;;) {
      const WNode n = load_wnode(wn, node);
      if constexpr (CNT) ++tc->nodes;
      uint32_t hit_int, hit_leaf;
      wide_children(n, o, inv, perm, tmin, tmax, hit_int, hit_leaf);
      while (hit_leaf) {
         const int sl = __ffs(hit_leaf) - 1;
         hit_leaf &= hit_leaf - 1u;
         const uint32_t m = ((sl < 4 ? n.meta[0] : n.meta[1]) >> (8 * (sl & 3))) & 0xFFu;
         const int first = (int)n.prim_base + (int)(m & 31u);
         const int cnt = (int)(m >> 5) + 1;
         for (int i = first; i < first + cnt; ++i)
            if (test_any<CNT>(S, prims, i, o, d, tmin, tmax, tc)) return true;
      }
      const uint32_t imask = n.e_imask >> 24;
      uint32_t base = n.child_base;
      if (!hit_int) {
         if (sp == 0) return false;
         base = (uint32_t)stack[(2 * (sp - 1)) * STRIDE];
         const uint32_t bits = (uint32_t)stack[(2 * (sp - 1) + 1) * STRIDE];
         hit_int = bits & 0xFFu;
         const uint32_t im = bits >> 8;
         const int i = __ffs(hit_int) - 1;
         hit_int &= hit_int - 1u;
         const uint32_t sl = (uint32_t)i ^ perm;
         node = base + (uint32_t)__popc(im & ((1u << sl) - 1u));
         if (hit_int)
            stack[(2 * (sp - 1) + 1) * STRIDE] = (int)(hit_int | (im << 8));
         else
            --sp;
         continue;
      }
      const int i = __ffs(hit_int) - 1;
      hit_int &= hit_int - 1u;
      const uint32_t sl = (uint32_t)i ^ perm;
      node = base + (uint32_t)__popc(imask & ((1u << sl) - 1u));
      if (hit_int && sp < kStack / 2) {
         stack[(2 * sp) * STRIDE] = (int)base;
         stack[(2 * sp + 1) * STRIDE] = (int)(hit_int | (imask << 8));
         ++sp;
      }
   }
}

// Tree layouts a kernel is compiled for: the binary pair-node tree, the 8-wide quantized
// tree, or either (chosen per ray from the scene, for batches that mix layouts).
enum Layout : int { L_BIN = 0, L_WIDE = 1, L_ANY = 2 };

// DEFER: an SDF field met by the traversal is not marched; its slot is returned in
// HitMin::sdf for a later marching stage (which marches up to the closest other hit).
// STACK: LDS entries per lane the caller provides; binary trees need at most their depth,
// the wide traversal always takes kStack.
template <int STRIDE, bool CNT = false, int LAYOUT = L_ANY, bool DEFER = false, int STACK = kStack>
__device__ inline HitMin closest_hit(const SceneView& S, V3 o, V3 d, float tmin, float tmax, int* stack,
                                     TravCount* tc = nullptr) {
   if (LAYOUT == L_WIDE || (LAYOUT == L_ANY && S.wnodes))
      return closest_hit_wide<STRIDE, CNT, DEFER>(S, o, d, tmin, tmax, stack, tc);
   return closest_hit_bin<STRIDE, CNT, DEFER, STACK>(S, o, d, tmin, tmax, stack, tc);
}

template <int STRIDE, bool CNT = false, int LAYOUT = L_ANY, int STACK = kStack>
__device__ inline bool any_hit(const SceneView& S, V3 o, V3 d, float tmin, float tmax, int* stack,
                               TravCount* tc = nullptr) {
   if (LAYOUT == L_WIDE || (LAYOUT == L_ANY && S.wnodes))
      return any_hit_wide<STRIDE, CNT>(S, o, d, tmin, tmax, stack, tc);
   return any_hit_bin<STRIDE, CNT, STACK>(S, o, d, tmin, tmax, stack, tc);
}

// ------------------------------------------------------------------------------------
// surface reconstruction for the winning primitive
// ------------------------------------------------------------------------------------
struct Surf {
   V3 p, n, tangent, center;
   bool front;
   uint32_t info;
};

__device__ __forceinline__ V3 fib_point(int i, int n) {  // shader_golf.cu:4-13
   const float ga = 2.39996323f;
   float k = (float)i + 0.5f;
   float z = 1.0f - 2.0f * (k / (float)n);
   float s = sqrtf(fmaxf(0.0f, 1.0f - z * z));
   float th = ga * k;
   return mk(cosf(th) * s, sinf(th) * s, z);
}

__device__ inline float dimple(V3 p) {  // shader_golf.cu:15-43
   V3 q = unit(p);
   float best = -1.0f;
   for (int i = 0; i < 150; ++i) best = fmaxf(best, dot(q, fib_point(i, 150)));
   best = fmaxf(fminf(best, 1.0f), -1.0f);
   float ang = acosf(best);
   if (ang < 0.24f) return -0.35f * cosf(ang / 0.24f * kPiF * 0.5f);
   return 0.0f;
}

__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
   float t = fmaxf(0.0f, fminf(1.0f, (x - e0) / (e1 - e0)));
   return t * t * (3.0f - 2.0f * t);
}

// Where the normal of an SDF hit reads the field from: the scene's arrays, or the copy of
// this field that the calling wave holds in LDS (`held`; the same values, so the same normal).
struct ShadeField {
   const __attribute__((address_space(3))) gvec4* rec = nullptr;
   const __attribute__((address_space(3))) gvec4* sph = nullptr;
   bool held = false;
};

// Central differences of the field at p (sdf.hpp:499-507).
template <class SPH, class REC>
__device__ __forceinline__ V3 sdf_gradient(const SceneView& S, V3 p, float e, SPH sph, REC rec) {
   const uint64_t m = sdf_point_mask(S, p, sph);
   return mk(sdf_field_near(S, p + mk(e, 0, 0), m, rec) - sdf_field_near(S, p - mk(e, 0, 0), m, rec),
             sdf_field_near(S, p + mk(0, e, 0), m, rec) - sdf_field_near(S, p - mk(0, e, 0), m, rec),
             sdf_field_near(S, p + mk(0, 0, e), m, rec) - sdf_field_near(S, p - mk(0, 0, e), m, rec));
}

// FT: the scene features (SceneFeature) whose code is compiled in.
template <uint32_t FT = SF_ALL>
__device__ inline Surf surface(const SceneView& S, const HitMin& h, V3 o, V3 d, ShadeField fld = ShadeField{}) {
   const PrimRec p = gload(S.prims + h.slot);
   Surf s;
   s.info = __float_as_uint(p.a.w);
   const uint32_t kind = info_kind(s.info);
   s.p = o + h.t * d;
   s.center = mk(0.0f, 0.0f, 0.0f);
   if (kind == K_TRI) {  // cuda_raytracer.cuh:306-346
      V3 e1 = mk(p.b), e2 = mk(p.c);
      V3 geo = unit(cross(e1, e2));
      s.front = dot(d, geo) < 0.0f;
      V3 sn = geo;
      const bool smooth = (s.info & F_NORMALS) != 0;
      if (smooth) {
         const float4* nn = S.normals + 3 * h.slot;
         float w = 1.0f - h.u - h.v;
         sn = unit(w * mk(gload(nn)) + h.u * mk(gload(nn + 1)) + h.v * mk(gload(nn + 2)));
         if (dot(sn, geo) < 0.0f) sn = -sn;
      }
      s.n = s.front ? sn : -sn;
      if (smooth && dot(s.n, -d) <= 0.0f) s.n = s.front ? geo : -geo;
      s.tangent = unit(e1);
   } else if (kind == K_RECT) {  // cuda_raytracer.cuh:264-270
      V3 n = unit(cross(mk(p.b), mk(p.c)));
      s.front = dot(d, n) < 0.0f;
      s.n = s.front ? n : -n;
      s.tangent = unit(mk(p.b));
   } else if ((FT & SF_SDF) && kind == K_SDF) {  // central differences, sdf.hpp:499-507
      const float e = S.sdf_eps;
      const V3 g = fld.held ? sdf_gradient(S, s.p, e, SdfLdsSph{fld.sph}, SdfLdsRecs{fld.rec})
                            : sdf_gradient(S, s.p, e, SdfGlobalSph{S.sdf_sph}, SdfGlobalRecs{S.sdf});
      V3 n = unit(g);
      s.front = dot(d, n) < 0.0f;
      s.n = s.front ? n : -n;
      V3 tg = cross(mk(0.0f, 1.0f, 0.0f), n);
      s.tangent = len2(tg) < 1e-6f ? mk(1.0f, 0.0f, 0.0f) : unit(tg);
   } else {  // sphere, cuda_raytracer.cuh:223-231
      V3 c = mk(p.a);
      float rad = p.b.x;
      s.center = c;
      V3 out = (1.0f / rad) * (s.p - c);
      s.front = dot(d, out) < 0.0f;
      s.n = s.front ? out : -out;
      V3 tg = cross(mk(0.0f, 1.0f, 0.0f), out);
      s.tangent = len2(tg) < 1e-6f ? mk(1.0f, 0.0f, 0.0f) : unit(tg);
      if ((FT & SF_DISPLACED) && kind == K_DISPLACED) {  // golf-ball displacement, shader_golf.cu:61-138
         V3 sp = s.p;
         V3 bo = unit(sp - c);
         float disp = dimple(bo);
         float dn = fminf(1.0f, fmaxf(0.0f, -disp / 0.35f));
         s.p = sp + (rad * 0.35f * (1.0f - dn)) * bo;
         V3 bn = unit(s.p - c);
         V3 nrm = bn;
         if (disp < -0.001f) {
            V3 helper = fabsf(bn.x) > 0.8f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
            V3 t1 = unit(cross(helper, bn));
            V3 t2 = cross(bn, t1);
            const float hs = 0.015f;
            float d0 = dimple(bn);
            float d1 = dimple(unit(bn + hs * t1));
            float d2 = dimple(unit(bn + hs * t2));
            V3 g = ((d1 - d0) / hs) * t1 + ((d2 - d0) / hs) * t2;
            V3 dnv = -0.2f * g;
            float ndv = fmaxf(0.0f, dot(bn, unit(-d)));
            dnv = smoothstepf(0.1f, 0.4f, ndv) * dnv;
            float l = sqrtf(len2(dnv));
            if (l > 0.4f && l > 1e-6f) dnv = (0.4f / l) * dnv;
            V3 pert = unit(bn + dnv);
            if (dot(pert, bn) < 0.0f) pert = -pert;
            if (!(pert.x == pert.x) || !(pert.y == pert.y) || !(pert.z == pert.z)) pert = bn;
            nrm = pert;
         }
         float ndotv = dot(d, nrm);
         if (!(ndotv == ndotv)) {
            nrm = bn;
            ndotv = dot(d, nrm);
         }
         s.front = ndotv < 0.0f;
         s.n = s.front ? nrm : -nrm;
         s.p = s.p + 1e-3f * s.n;
      }
   }
   return s;
}

// ------------------------------------------------------------------------------------
// materials
// ------------------------------------------------------------------------------------
__device__ __forceinline__ V3 reflect(V3 v, V3 n) { return v - (2.0f * dot(v, n)) * n; }

__device__ __forceinline__ V3 refract(V3 uv, V3 n, float eta) {  // cuda_raytracer.cuh:86-92
   float c = fminf(dot(-uv, n), 1.0f);
   V3 perp = eta * (uv + c * n);
   V3 par = (-sqrtf(fabsf(1.0f - len2(perp)))) * n;
   return perp + par;
}

__device__ __forceinline__ float schlick(float c, float ri) {  // cuda_raytracer.cuh:94-99
   float r0 = (1.0f - ri) / (1.0f + ri);
   r0 *= r0;
   float x = 1.0f - c;
   float x2 = x * x;
   return r0 + (1.0f - r0) * (x2 * x2 * x);
}

__device__ __forceinline__ void onb(V3 n, V3& u, V3& v) {  // cuda_utils.cuh:99-107
   if (fabsf(n.x) > fabsf(n.z))
      u = unit(mk(-n.y, n.x, 0.0f));
   else
      u = unit(mk(0.0f, -n.z, n.y));
   v = cross(n, u);
}

template <class RNG>
__device__ __forceinline__ V3 on_sphere(RNG& R) {  // cuda_utils.cuh:66-77
   float a = R.next();
   float z = 1.0f - 2.0f * R.next();
   float s = sqrtf(fmaxf(0.0f, 1.0f - z * z));
   return mk(s * cos2pi(a), s * sin2pi(a), z);
}

template <class RNG>
__device__ __forceinline__ V3 cosine_dir(V3 n, RNG& R) {  // cuda_raytracer.cuh:621-638
   V3 w = unit(n), u, v;
   onb(w, u, v);
   float u1 = R.next(), u2 = R.next();
   float r = sqrtf(u1);
   return (r * cos2pi(u2)) * u + (r * sin2pi(u2)) * v + sqrtf(fmaxf(0.0f, 1.0f - u1)) * w;
}

__device__ __forceinline__ float lambda_ggx(V3 w, float ax, float ay) {  // microfacet_ggx.cuh:21-28
   float wz2 = w.z * w.z;
   if (wz2 < 1e-16f) return 0.0f;
   float a2 = (ax * w.x) * (ax * w.x) + (ay * w.y) * (ay * w.y);
   return (sqrtf(1.0f + a2 / wz2) - 1.0f) * 0.5f;
}

__device__ __forceinline__ V3 sample_vndf(V3 wo, float ax, float ay, float u1, float u2) {  // :45-76
   V3 wh = unit(mk(ax * wo.x, ay * wo.y, wo.z));
   if (wh.z < 0.0f) wh = -wh;
   V3 t1 = wh.z < 0.99999f ? unit(cross(mk(0.0f, 0.0f, 1.0f), wh)) : mk(1.0f, 0.0f, 0.0f);
   V3 t2 = cross(wh, t1);
   float r = sqrtf(u1);
   float px = r * cos2pi(u2), py = r * sin2pi(u2);
   float hh = sqrtf(fmaxf(0.0f, 1.0f - px * px));
   float sm = (1.0f + wh.z) * 0.5f;
   py = (1.0f - sm) * hh + sm * py;
   float pz = sqrtf(fmaxf(0.0f, 1.0f - px * px - py * py));
   V3 nh = px * t1 + py * t2 + pz * wh;
   return unit(mk(ax * nh.x, ay * nh.y, fmaxf(1e-6f, nh.z)));
}

__device__ __forceinline__ float fr_complex1(float cos2, float sin2, float ci, float eta, float k) {  // :83-101
   float eta2 = eta * eta, k2 = k * k;
   float t0 = eta2 - k2 - sin2;
   float ab = sqrtf(fmaxf(t0 * t0 + 4.0f * eta2 * k2, 0.0f));
   float a = sqrtf(fmaxf((ab + t0) * 0.5f, 0.0f));
   float rs = (ab + cos2 - 2.0f * a * ci) / fmaxf(ab + cos2 + 2.0f * a * ci, 1e-10f);
   float rp = rs * (ab * cos2 + sin2 * sin2 - 2.0f * a * ci * sin2) /
              fmaxf(ab * cos2 + sin2 * sin2 + 2.0f * a * ci * sin2, 1e-10f);
   return (rs + rp) * 0.5f;
}

__device__ __forceinline__ V3 fr_complex(float ci, V3 eta, V3 k) {
   ci = fminf(fmaxf(ci, 0.0f), 1.0f);
   float c2 = ci * ci, s2 = 1.0f - c2;
   return mk(fr_complex1(c2, s2, ci, eta.x, k.x), fr_complex1(c2, s2, ci, eta.y, k.y),
             fr_complex1(c2, s2, ci, eta.z, k.z));
}

__device__ __forceinline__ float fresnel_dielectric(float ci, float eta) {
   float s2t = eta * eta * fmaxf(0.0f, 1.0f - ci * ci);
   if (s2t >= 1.0f) return 1.0f;
   float ct = sqrtf(1.0f - s2t);
   float rs = (eta * ci - ct) / (eta * ci + ct);
   float rp = (ci - eta * ct) / (ci + eta * ct);
   return 0.5f * (rs * rs + rp * rp);
}

struct Scatter {
   V3 o, d, atten, emitted;
   bool scattered;
   bool lambert;
};

// apply_material's type mapping (cuda_raytracer.cuh:418-485) and scatter_material
// (:613-850); RAYON_MAT_GGX_* are additions of this library. RNG is the path's stream
// (Rng), or a table of given numbers for the material probe.
template <class RNG, uint32_t FT = SF_ALL>
__device__ inline void scatter(const MatRec& M, V3 color, const JobView& J, const Surf& s, V3 dir, RNG& R,
                               Scatter& o) {
   const uint32_t mt = __float_as_uint(M.m0.w) & 0xFFu;
   const bool quirks = (J.flags & RAYON_FLAG_REF_QUIRKS) != 0;
   const V3 n = s.n;
   o.emitted = mk(0.0f, 0.0f, 0.0f);
   o.atten = mk(1.0f, 1.0f, 1.0f);
   o.scattered = false;
   o.lambert = false;
   switch (mt) {
   case RAYON_MAT_LAMBERTIAN:
   case RAYON_MAT_SDF_MATERIAL: {
      o.d = cosine_dir(n, R);
      o.o = s.p + 1e-4f * n;
      o.atten = color;
      o.scattered = true;
      o.lambert = true;
      return;
   }
   case RAYON_MAT_MIRROR: {
      V3 rf = reflect(unit(dir), n);
      o.o = s.p;
      o.d = rf;
      o.atten = color;
      o.scattered = dot(rf, n) > 0.0f;
      return;
   }
   case RAYON_MAT_METAL:
   case RAYON_MAT_ROUGH_MIRROR: {
      float rough = M.m1.w, fz = 1.0f;
      if (quirks) {
         rough = rough > 0.0f ? rough : 0.3f;
         fz = J.metal_fuzziness;
      }
      V3 sph = on_sphere(R);
      V3 pn = unit(n + (rough * fz) * sph);
      V3 rf = reflect(unit(dir), pn);
      o.o = s.p;
      o.d = rf;
      o.atten = color;
      o.scattered = dot(rf, n) > 0.0f;
      return;
   }
   case RAYON_MAT_GLASS:
   case RAYON_MAT_DIELECTRIC: {
      V3 ud = unit(dir);
      float ri = quirks ? J.glass_ior : M.m2.w;
      float ratio = s.front ? 1.0f / ri : ri;
      float ct = fminf(dot(-ud, n), 1.0f);
      float st = sqrtf(1.0f - ct * ct);
      bool tir = ratio * st > 1.0f;
      if (tir || schlick(ct, ratio) > R.next())
         o.d = reflect(ud, n);
      else
         o.d = refract(ud, n, ratio);
      o.o = s.p;
      o.scattered = true;
      return;
   }
   case RAYON_MAT_LIGHT:
      o.emitted = J.light_intensity * mk(M.m1);
      return;
   case RAYON_MAT_CONSTANT:
      o.emitted = color;
      return;
   case RAYON_MAT_SHOW_NORMALS:
      o.emitted = 0.5f * (n + mk(1.0f, 1.0f, 1.0f));
      return;
   case RAYON_MAT_ANISOTROPIC_METAL:
   case RAYON_MAT_GGX_CONDUCTOR: {
      const bool aniso = (FT & SF_ANISO) && mt == RAYON_MAT_ANISOTROPIC_METAL;
      const float rough = M.m1.w;
      float ax, ay;
      if (aniso) {
         float aspect = sqrtf(fmaxf(1e-4f, 1.0f - 0.9f * M.m3.w));
         float r2 = rough * rough;
         ax = fmaxf(1e-4f, r2 / aspect);
         ay = fmaxf(1e-4f, r2 * aspect);
      } else {
         ax = ay = fmaxf(1e-4f, rough * rough);
      }
      const V3 eta = mk(M.m2), kk = mk(M.m3);
      if (fmaxf(ax, ay) < 1e-3f) {
         V3 rf = reflect(unit(dir), n);
         o.o = s.p + 1e-4f * n;
         o.d = rf;
         float ci = fmaxf(dot(-unit(dir), n), 0.0f);
         o.atten = color * fr_complex(ci, eta, kk);
         o.scattered = dot(rf, n) > 0.0f;
         return;
      }
      V3 N = unit(n), T, B;
      if (aniso) {
         T = unit(s.tangent - dot(s.tangent, N) * N);
         B = cross(N, T);
      } else {
         onb(N, T, B);
      }
      V3 wo = unit(-dir);
      V3 wol = mk(dot(wo, T), dot(wo, B), dot(wo, N));
      if (wol.z <= 0.0f) return;
      float u1 = R.next(), u2 = R.next();
      V3 wm = sample_vndf(wol, ax, ay, u1, u2);
      V3 wil = reflect(-wol, wm);
      if (wil.z <= 0.0f) return;
      V3 F = fr_complex(fmaxf(dot(wol, wm), 0.0f), eta, kk);
      float lo = lambda_ggx(wol, ax, ay), li = lambda_ggx(wil, ax, ay);
      float w = (1.0f / (1.0f + lo + li)) / fmaxf(1.0f / (1.0f + lo), 1e-6f);
      o.o = s.p + 1e-4f * n;
      o.d = wil.x * T + wil.y * B + wil.z * N;
      o.atten = color * (w * F);
      o.scattered = true;
      return;
   }
   case RAYON_MAT_GGX_DIELECTRIC: {
      const float rough = M.m1.w;
      float a = fmaxf(1e-4f, rough * rough);
      float ri = M.m2.w;
      float eta = s.front ? 1.0f / ri : ri;
      V3 N = unit(n), T, B;
      onb(N, T, B);
      V3 wo = unit(-dir);
      V3 wol = mk(dot(wo, T), dot(wo, B), dot(wo, N));
      if (wol.z <= 0.0f) return;
      float u1 = R.next(), u2 = R.next();
      V3 wm = sample_vndf(wol, a, a, u1, u2);
      float F = fresnel_dielectric(dot(wol, wm), eta);
      bool refl = R.next() < F;
      V3 wil;
      if (refl) {
         wil = reflect(-wol, wm);
         if (wil.z <= 0.0f) return;
      } else {
         wil = refract(-wol, wm, eta);
         if (wil.z >= 0.0f) return;
      }
      float lo = lambda_ggx(wol, a, a), li = lambda_ggx(wil, a, a);
      float w = (1.0f / (1.0f + lo + li)) / fmaxf(1.0f / (1.0f + lo), 1e-6f);
      o.o = refl ? s.p + 1e-4f * n : s.p - 1e-4f * n;
      o.d = wil.x * T + wil.y * B + wil.z * N;
      o.atten = w * color;
      o.scattered = true;
      return;
   }
   case RAYON_MAT_THIN_FILM: {  // cuda_raytracer.cuh:753-809
      if constexpr (!(FT & SF_LAYERED)) return;
      V3 ud = unit(dir);
      float ci = fmaxf(fabsf(dot(-ud, n)), 0.001f);
      float n0 = M.m2.w, n1 = M.m4.y, dd = M.m4.x;
      float si = sqrtf(fmaxf(0.0f, 1.0f - ci * ci));
      float st = (n0 / n1) * si;
      float ct = st >= 1.0f ? 0.0f : sqrtf(1.0f - st * st);
      float r01 = (n0 - n1) / (n0 + n1);
      r01 *= r01;
      float x01 = 1.0f - ci;
      float R01 = r01 + (1.0f - r01) * x01 * x01 * x01 * x01 * x01;
      float r12 = (n1 - n0) / (n1 + n0);
      r12 *= r12;
      float x12 = 1.0f - ct;
      float R12 = r12 + (1.0f - r12) * x12 * x12 * x12 * x12 * x12;
      float sq = sqrtf(R01 * R12);
      float base = 4.0f * kPiF * n1 * dd * ct;
      float c0 = cosf(base / 650.0f), c1 = cosf(base / 550.0f), c2 = cosf(base / 450.0f);
      float Rr = (R01 + R12 + 2.0f * sq * c0) / fmaxf(1.0f + R01 * R12 + 2.0f * sq * c0, 1e-8f);
      float Rg = (R01 + R12 + 2.0f * sq * c1) / fmaxf(1.0f + R01 * R12 + 2.0f * sq * c1, 1e-8f);
      float Rb = (R01 + R12 + 2.0f * sq * c2) / fmaxf(1.0f + R01 * R12 + 2.0f * sq * c2, 1e-8f);
      float avg = (Rr + Rg + Rb) / 3.0f;
      if (R.next() < avg) {
         V3 rf = reflect(ud, n);
         o.o = s.p + 1e-4f * n;
         o.d = rf;
         o.atten = (1.0f / fmaxf(avg, 0.001f)) * mk(Rr, Rg, Rb);
         o.scattered = dot(rf, n) > 0.0f;
      } else {
         o.o = s.p - 1e-4f * n;
         o.d = ud;
         o.atten = (1.0f / fmaxf(1.0f - avg, 0.001f)) * mk(1.0f - Rr, 1.0f - Rg, 1.0f - Rb);
         o.scattered = true;
      }
      return;
   }
   case RAYON_MAT_CLEAR_COAT: {  // cuda_raytracer.cuh:810-846
      if constexpr (!(FT & SF_LAYERED)) return;
      V3 ud = unit(dir);
      float ct = fminf(dot(-ud, n), 1.0f);
      float ior = M.m2.w > 1.0f ? M.m2.w : 1.5f;
      float F = schlick(ct, ior);
      if (R.next() < F) {
         const float rough = M.m1.w;
         V3 pn = n;
         if (rough > 1e-3f) pn = unit(n + rough * on_sphere(R));
         V3 rf = reflect(ud, pn);
         o.o = s.p + 1e-4f * n;
         o.d = rf;
         o.scattered = dot(rf, n) > 0.0f;
      } else {
         o.d = cosine_dir(n, R);
         o.o = s.p + 1e-4f * n;
         o.atten = color;
         o.scattered = true;
      }
      return;
   }
   default:
      return;
   }
}

// apply_procedural_pattern (cuda_raytracer.cuh:377-416)
template <uint32_t FT = SF_ALL>
__device__ inline V3 surface_color(const MatRec& M, V3 p, V3 center) {
   V3 base = mk(M.m0);
   const uint32_t pat = (__float_as_uint(M.m0.w) >> 8) & 0xFFu;
   if (!(FT & SF_DOTS) || pat != RAYON_PATTERN_FIBONACCI_DOTS) return base;
   V3 dir = unit(p - center);
   int N = (int)M.m4.z;
   float best = -1.0f;
   for (int i = 0; i < N; ++i) best = fmaxf(best, dot(dir, fib_point(i, N)));
   best = fmaxf(fminf(best, 1.0f), -1.0f);
   float mask = acosf(best) < M.m4.w ? 0.0f : 1.0f;
   return mask * base + (1.0f - mask) * mk(M.m5);
}

// Uniform point on an emitter (rectangle, triangle, sphere surface).
__device__ __forceinline__ void sample_emitter(const PrimRec& p, float a, float b, V3& q, V3& nl, bool& one_sided) {
   const uint32_t kind = info_kind(__float_as_uint(p.a.w));
   one_sided = false;
   if (kind == K_RECT) {
      q = mk(p.a) + a * mk(p.b) + b * mk(p.c);
      nl = unit(cross(mk(p.b), mk(p.c)));
   } else if (kind == K_TRI) {
      float su = sqrtf(a);
      q = mk(p.a) + (su * (1.0f - b)) * mk(p.b) + (su * b) * mk(p.c);
      nl = unit(cross(mk(p.b), mk(p.c)));
   } else {
      float z = 1.0f - 2.0f * a;
      float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
      nl = mk(r * cos2pi(b), r * sin2pi(b), z);
      q = mk(p.a) + p.b.x * nl;
      one_sided = true;
   }
}

__device__ __forceinline__ V3 sky(const JobView& J, V3 d) {  // cuda_raytracer.cuh:924-929
   V3 ud = unit(d);
   float t = 0.5f * (ud.y + 1.0f);
   return J.background_intensity * mk(1.0f - 0.5f * t, 1.0f - 0.3f * t, 1.0f);
}

// Compensated add of an already rounded x into the float-float pair (sum, c): the exact
// error of sum + x (Knuth's TwoSum) joins the low part, and the pair is renormalised
// (Fast2Sum) so |c| stays within half an ulp of sum. The pair carries about 48 bits, so
// a pixel's sum over millions of samples stays within an ulp of the exact FP32-rounded
// sum; a Neumaier sum, whose FP32 compensation term grows with the number of samples,
// drifts by hundreds of ulp over 2^20 equal samples. Adds only, so no contraction can
// change it (callers round products with __fmul_rn before passing them).
__device__ __forceinline__ void ff_add(float& sum, float& c, float x) {
   const float t = sum + x;
   const float z = t - sum;
   float e = (sum - (t - z)) + (x - z);
   e += c;
   sum = t + e;
   c = e - (sum - t);
}

__device__ __forceinline__ float luminance(V3 c) { return 0.299f * c.x + 0.587f * c.y + 0.114f * c.z; }

// Camera sample: render_acc_kernel.cu:200-221.
__device__ __forceinline__ void camera_ray(const JobView& J, int x, int y, Rng& R, V3& o, V3& d) {
   float ou = R.next() - 0.5f;
   float ov = R.next() - 0.5f;
   V3 C = ld3(J.cam_c);
   V3 pc = ld3(J.p00) + ((float)x + ou) * ld3(J.du) + ((float)y + ov) * ld3(J.dv);
   o = C;
   d = pc - C;
   if ((J.flags & RAYON_FLAG_DOF) && J.dof_aperture > 0.0f) {
      V3 nd = unit(d);
      V3 focus = C + J.dof_focus * nd;
      float dx, dy;
      do {
         dx = 2.0f * R.next() - 1.0f;
         dy = 2.0f * R.next() - 1.0f;
      } while (dx * dx + dy * dy >= 1.0f);
      o = C + J.dof_aperture * (dx * ld3(J.cu) + dy * ld3(J.cv));
      d = focus - o;
   }
}

}  // namespace ry
