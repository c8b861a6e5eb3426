// oracle.cpp — FP64 CPU restatement of the reference's GPU path tracer.
//
// TEST INFRASTRUCTURE ONLY: the checker for librayon_hip.so. Loaded by the tests, the
// smoke check and the benchmark's CPU-baseline leg; never by the product.
//
// Every function cites the reference code it restates (paths relative to
// src/rayon/gpu_renderers/ unless stated). The arithmetic is FP64 on FP32 scene data;
// random numbers are the same FP32 values the device draws. Additions with no reference
// semantics (NEE, GGX conductor/dielectric, counter RNG, SDF field on the GPU path,
// standard-error adaptivity) are marked "spec".
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

// The checker computes in FP64 (real = double). The benchmark's CPU baseline is this same
// restatement built with ORACLE_REAL=float, the binned-SAH tree (ORACLE_BVH_BINNED) and
// static row chunks over the host threads (ORACLE_STATIC_ROWS); see oracle/build_cpu.sh.
#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif

namespace {

using real = ORACLE_REAL;

// min / max / clamp that compute in `real` whenever a floating-point operand is involved
// (literals are double), and in the common type of integer operands otherwise
template <class A, class B>
inline auto mx(A a, B b) {
   if constexpr (std::is_floating_point_v<A> || std::is_floating_point_v<B>)
      return std::max<real>((real)a, (real)b);
   else
      return std::max<std::common_type_t<A, B>>(a, b);
}
template <class A, class B>
inline auto mn(A a, B b) {
   if constexpr (std::is_floating_point_v<A> || std::is_floating_point_v<B>)
      return std::min<real>((real)a, (real)b);
   else
      return std::min<std::common_type_t<A, B>>(a, b);
}
template <class A, class B, class C>
inline real cl(A v, B lo, C hi) {
   return std::clamp<real>((real)v, (real)lo, (real)hi);
}

constexpr real kPi = (real)3.14159265358979323846;

// Planted mistakes for the pinning tests (oracle_set_fault): each one is a plausible
// error in a formula the tests must catch. 0 = none, the only value outside those tests.
int g_fault = 0;
enum Fault {
   FAULT_VNDF_NO_WARP = 1,        // VNDF without the hemisphere warp of the disk sample
   FAULT_VNDF_SWAP_ALPHA = 2,     // VNDF stretch with alpha_x and alpha_y swapped
   FAULT_CONDUCTOR_RP_SIGN = 3,   // conductor Fresnel: wrong sign of 2 a cos sin^2 in Rp's numerator
   FAULT_DIELECTRIC_RP = 4,       // dielectric Fresnel: Rs used in place of Rp
   FAULT_DISK_NO_REJECT = 5,      // unit-disk sampler that keeps points of the square
   FAULT_LAMBDA_HALF = 6,         // Smith Lambda without its factor 1/2
};

struct V3 {
   real x, y, z;
};
inline V3 v3(real x, real y, real z) { return {x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(real s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline real len2(V3 a) { return dot(a, a); }
inline real len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { return (1.0 / len(a)) * a; }
inline real maxc(V3 a) { return mx(a.x, mx(a.y, a.z)); }
inline V3 fv(const float* p) { return {p[0], p[1], p[2]}; }
inline V3 dv(const double* p) { return {(real)p[0], (real)p[1], (real)p[2]}; }
inline real axis(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ------------------------------------------------------------------------------------
// RNG
// ------------------------------------------------------------------------------------
// cuda_utils.cuh:28-35: LCG step then xorshift-multiply output permutation.
inline uint32_t lcg_step(uint32_t s) { return s * 747796405u + 2891336453u; }
inline uint32_t permute(uint32_t s) {
   uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
   return (w >> 22u) ^ w;
}
inline uint32_t hash32(uint32_t v) { return permute(lcg_step(v)); }

struct Rng {
   uint32_t state;
   bool ref;  // reference float conversion (cuda_utils.cuh:35) vs 24-bit (spec)
   float next() {
      state = lcg_step(state);
      uint32_t r = permute(state);
      if (ref) return (float)r * (1.0f / 4294967296.0f);  // can round to 1.0f
      return (float)(r >> 8) * (1.0f / 16777216.0f);
   }
};

// spec: counter-based stream start for (seed, pixel, sample).
uint32_t counter_init(uint64_t seed, uint32_t pixel, uint32_t sample) {
   uint32_t h = hash32((uint32_t)seed ^ hash32((uint32_t)(seed >> 32) + 0x9E3779B9u));
   h = hash32(h + pixel);
   h = hash32(h ^ (sample * 0x85EBCA6Bu + 0xC2B2AE35u));
   return h;
}

// ------------------------------------------------------------------------------------
// Scene
// ------------------------------------------------------------------------------------
enum PrimType { P_SPHERE = 0, P_RECT = 1, P_DISPLACED = 3, P_TRI = 4, P_SDF = 7 };

struct Prim {
   int type;
   int mat;
   bool visible;
   bool has_n;
   V3 a, b, c;      // sphere: center, (r,0,0) | rect: corner,u,v | tri: v0,v1,v2
   V3 n0, n1, n2;
   V3 lo, hi;
   int64_t id;
};

struct SdfPrim {
   int type;
   V3 c;
   real p[4];
};

struct Node {
   V3 lo, hi;
   int64_t left, right;  // interior
   int64_t first, count; // leaf if count > 0
};

struct Light {
   int64_t prim;
   real area;
};

}  // namespace

struct oracle_scene {
   std::vector<Prim> prims;
   std::vector<rayon_material> mats;
   std::vector<SdfPrim> sdf;
   real sdf_k = 0.1;
   int sdf_steps = 256;
   real sdf_eps = 1e-4;
   std::vector<Node> nodes;
   std::vector<int64_t> order;  // leaf order -> prim index
   std::vector<Light> lights;
};

namespace {

void prim_bounds(Prim& p) {
   switch (p.type) {
   case P_SPHERE:
   case P_DISPLACED: {
      real r = p.b.x * (p.type == P_DISPLACED ? 1.4 : 1.0);
      p.lo = p.a - v3(r, r, r);
      p.hi = p.a + v3(r, r, r);
      break;
   }
   case P_RECT: {
      V3 c[4] = {p.a, p.a + p.b, p.a + p.c, p.a + p.b + p.c};
      p.lo = p.hi = c[0];
      for (int i = 1; i < 4; ++i) {
         p.lo = v3(mn(p.lo.x, c[i].x), mn(p.lo.y, c[i].y), mn(p.lo.z, c[i].z));
         p.hi = v3(mx(p.hi.x, c[i].x), mx(p.hi.y, c[i].y), mx(p.hi.z, c[i].z));
      }
      break;
   }
   case P_TRI: {
      p.lo = v3(mn(mn(p.a.x, p.b.x), p.c.x), mn(mn(p.a.y, p.b.y), p.c.y), mn(mn(p.a.z, p.b.z), p.c.z));
      p.hi = v3(mx(mx(p.a.x, p.b.x), p.c.x), mx(mx(p.a.y, p.b.y), p.c.y), mx(mx(p.a.z, p.b.z), p.c.z));
      break;
   }
   default:
      break;
   }
   // padded outward by far more than the rounding of the bounds (FP32 builds round more)
   const real e = (sizeof(real) < 8 ? 1e-6 : 1e-9) * (1.0 + mx(len(p.lo), len(p.hi)));
   p.lo = p.lo - v3(e, e, e);
   p.hi = p.hi + v3(e, e, e);
}

// Oracle BVH: object-median split on the widest centroid axis, leaves <= 4. Independent
// of the product's binned-SAH builder; only the closest hit matters, not the tree.
int64_t build_node(oracle_scene& s, int64_t first, int64_t count) {
   Node n;
   n.lo = v3(1e300, 1e300, 1e300);
   n.hi = v3(-1e300, -1e300, -1e300);
   V3 clo = n.lo, chi = n.hi;
   for (int64_t i = first; i < first + count; ++i) {
      const Prim& p = s.prims[s.order[i]];
      n.lo = v3(mn(n.lo.x, p.lo.x), mn(n.lo.y, p.lo.y), mn(n.lo.z, p.lo.z));
      n.hi = v3(mx(n.hi.x, p.hi.x), mx(n.hi.y, p.hi.y), mx(n.hi.z, p.hi.z));
      V3 c = 0.5 * (p.lo + p.hi);
      clo = v3(mn(clo.x, c.x), mn(clo.y, c.y), mn(clo.z, c.z));
      chi = v3(mx(chi.x, c.x), mx(chi.y, c.y), mx(chi.z, c.z));
   }
   int64_t idx = (int64_t)s.nodes.size();
   s.nodes.push_back(n);
   if (count <= 4) {
      s.nodes[idx].first = first;
      s.nodes[idx].count = count;
      return idx;
   }
   V3 ext = chi - clo;
   int ax = ext.x > ext.y ? (ext.x > ext.z ? 0 : 2) : (ext.y > ext.z ? 1 : 2);
   int64_t mid = first + count / 2;
#ifdef ORACLE_BVH_BINNED
   // the GPU path's build rule instead: binned SAH over 16 centroid bins per axis, cost
   // A_L n_L + A_R n_R, the cheapest split over the three axes; the median if none splits
   {
      constexpr int B = 16;
      auto area = [](V3 lo, V3 hi) {
         V3 e = v3(mx(0.0, hi.x - lo.x), mx(0.0, hi.y - lo.y), mx(0.0, hi.z - lo.z));
         return 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
      };
      real best = 1e30;
      int best_ax = -1, best_b = -1;
      for (int a = 0; a < 3; ++a) {
         const real lo_a = axis(clo, a), ext_a = axis(chi, a) - lo_a;
         if (!(ext_a > 0)) continue;
         V3 blo[B], bhi[B];
         int64_t cnt[B] = {0};
         for (int b = 0; b < B; ++b) blo[b] = v3(1e30, 1e30, 1e30), bhi[b] = v3(-1e30, -1e30, -1e30);
         auto bin = [&](const Prim& p) {
            int b = (int)((axis(0.5 * (p.lo + p.hi), a) - lo_a) * (B / ext_a));
            return b < 0 ? 0 : (b >= B ? B - 1 : b);
         };
         for (int64_t i = first; i < first + count; ++i) {
            const Prim& p = s.prims[s.order[i]];
            const int b = bin(p);
            ++cnt[b];
            blo[b] = v3(mn(blo[b].x, p.lo.x), mn(blo[b].y, p.lo.y), mn(blo[b].z, p.lo.z));
            bhi[b] = v3(mx(bhi[b].x, p.hi.x), mx(bhi[b].y, p.hi.y), mx(bhi[b].z, p.hi.z));
         }
         real rarea[B];
         int64_t rcnt[B];
         V3 rl = v3(1e30, 1e30, 1e30), rh = v3(-1e30, -1e30, -1e30);
         int64_t rc = 0;
         for (int b = B - 1; b > 0; --b) {
            rl = v3(mn(rl.x, blo[b].x), mn(rl.y, blo[b].y), mn(rl.z, blo[b].z));
            rh = v3(mx(rh.x, bhi[b].x), mx(rh.y, bhi[b].y), mx(rh.z, bhi[b].z));
            rc += cnt[b];
            rarea[b] = area(rl, rh);
            rcnt[b] = rc;
         }
         V3 ll = v3(1e30, 1e30, 1e30), lh = v3(-1e30, -1e30, -1e30);
         int64_t lc = 0;
         for (int b = 0; b < B - 1; ++b) {
            ll = v3(mn(ll.x, blo[b].x), mn(ll.y, blo[b].y), mn(ll.z, blo[b].z));
            lh = v3(mx(lh.x, bhi[b].x), mx(lh.y, bhi[b].y), mx(lh.z, bhi[b].z));
            lc += cnt[b];
            if (lc == 0 || rcnt[b + 1] == 0) continue;
            const real cost = area(ll, lh) * (real)lc + rarea[b + 1] * (real)rcnt[b + 1];
            if (cost < best) best = cost, best_ax = a, best_b = b + 1;
         }
      }
      if (best_ax >= 0) {
         const real lo_a = axis(clo, best_ax), ext_a = axis(chi, best_ax) - lo_a;
         auto it = std::partition(s.order.begin() + first, s.order.begin() + first + count, [&](int64_t pi) {
            const Prim& p = s.prims[pi];
            int b = (int)((axis(0.5 * (p.lo + p.hi), best_ax) - lo_a) * (B / ext_a));
            b = b < 0 ? 0 : (b >= B ? B - 1 : b);
            return b < best_b;
         });
         const int64_t m = it - s.order.begin();
         if (m > first && m < first + count) {
            const int64_t l = build_node(s, first, m - first);
            const int64_t r = build_node(s, m, first + count - m);
            s.nodes[idx].left = l;
            s.nodes[idx].right = r;
            s.nodes[idx].count = 0;
            return idx;
         }
      }
   }
#endif
   std::nth_element(s.order.begin() + first, s.order.begin() + mid, s.order.begin() + first + count,
                    [&](int64_t i, int64_t j) {
                       real ci = axis(s.prims[i].lo + s.prims[i].hi, ax);
                       real cj = axis(s.prims[j].lo + s.prims[j].hi, ax);
                       return ci < cj || (ci == cj && i < j);
                    });
   int64_t l = build_node(s, first, mid - first);
   int64_t r = build_node(s, mid, first + count - mid);
   s.nodes[idx].left = l;
   s.nodes[idx].right = r;
   s.nodes[idx].count = 0;
   return idx;
}

struct Ray {
   V3 o, d;
};

struct Hit {
   real t;
   V3 p, normal, outward, tangent;
   bool front;
   int64_t prim;
};

// slab test, cuda_raytracer.cuh:154-197 (inclusive bounds)
bool hit_box(const Ray& r, V3 lo, V3 hi, real tmin, real tmax) {
   real t0 = -1e300, t1 = 1e300;
   for (int k = 0; k < 3; ++k) {
      real o = axis(r.o, k), d = axis(r.d, k);
      real inv = 1.0 / d;
      real a = (axis(lo, k) - o) * inv, b = (axis(hi, k) - o) * inv;
      if (std::isnan(a) || std::isnan(b)) {  // d == 0 and origin on a slab plane
         if (o < axis(lo, k) || o > axis(hi, k)) return false;
         continue;
      }
      if (a > b) std::swap(a, b);
      t0 = mx(t0, a);
      t1 = mn(t1, b);
   }
   return t0 <= t1 && t1 >= tmin && t0 <= tmax;
}

// hit_sphere, cuda_raytracer.cuh:203-233
bool isect_sphere(V3 c, real rad, const Ray& r, real tmin, real tmax, Hit& h) {
   V3 oc = r.o - c;
   real a = dot(r.d, r.d);
   real hb = dot(oc, r.d);
   real cc = dot(oc, oc) - rad * rad;
   real disc = hb * hb - a * cc;
   if (disc < 0) return false;
   real sq = std::sqrt(disc);
   real root = (-hb - sq) / a;
   if (root < tmin || tmax < root) {
      root = (-hb + sq) / a;
      if (root < tmin || tmax < root) return false;
   }
   h.t = root;
   h.p = r.o + root * r.d;
   h.outward = (1.0 / rad) * (h.p - c);
   h.front = dot(r.d, h.outward) < 0;
   h.normal = h.front ? h.outward : -h.outward;
   V3 tg = cross(v3(0, 1, 0), h.outward);
   if (len2(tg) < 1e-6) tg = v3(1, 0, 0);
   h.tangent = unit(tg);
   return true;
}

// hit_rectangle, cuda_raytracer.cuh:235-271
bool isect_rect(V3 q, V3 u, V3 v, const Ray& r, real tmin, real tmax, Hit& h) {
   V3 n = unit(cross(u, v));
   real denom = dot(n, r.d);
   if (std::fabs(denom) < 1e-8) return false;
   real t = dot(n, q - r.o) / denom;
   if (t < tmin || t > tmax) return false;
   V3 p = r.o + t * r.d;
   V3 rel = p - q;
   real al = dot(rel, u) / dot(u, u), be = dot(rel, v) / dot(v, v);
   if (al < 0 || al > 1 || be < 0 || be > 1) return false;
   h.t = t;
   h.p = p;
   h.front = dot(r.d, n) < 0;
   h.outward = n;
   h.normal = h.front ? n : -n;
   h.tangent = unit(u);
   return true;
}

// hit_triangle (Moller-Trumbore), cuda_raytracer.cuh:277-347
bool isect_tri(const Prim& pr, const Ray& r, real tmin, real tmax, Hit& h) {
   V3 e1 = pr.b - pr.a, e2 = pr.c - pr.a;
   V3 hh = cross(r.d, e2);
   real a = dot(e1, hh);
   if (std::fabs(a) < 1e-8) return false;
   real f = 1.0 / a;
   V3 s = r.o - pr.a;
   real u = f * dot(s, hh);
   if (u < 0 || u > 1) return false;
   V3 q = cross(s, e1);
   real v = f * dot(r.d, q);
   if (v < 0 || u + v > 1) return false;
   real t = f * dot(e2, q);
   if (t < tmin || t > tmax) return false;
   h.t = t;
   h.p = r.o + t * r.d;
   V3 geo = unit(cross(e1, e2));
   h.front = dot(r.d, geo) < 0;
   h.outward = geo;
   V3 sn = geo;
   if (pr.has_n) {
      real w = 1 - u - v;
      sn = unit(w * pr.n0 + u * pr.n1 + v * pr.n2);
      if (dot(sn, geo) < 0) sn = -sn;
   }
   h.normal = h.front ? sn : -sn;
   if (pr.has_n && dot(h.normal, -r.d) <= 0) h.normal = h.front ? geo : -geo;
   h.tangent = unit(e1);
   return true;
}

// fibonacci_point, shaders/shader_golf.cu:4-13
V3 fib_point(int i, int n) {
   const real ga = 2.39996323;
   real k = i + 0.5;
   real v = k / n;
   real phi = std::acos(1.0 - 2.0 * v);
   real theta = ga * k;
   real s = std::sin(phi);
   return v3(std::cos(theta) * s, std::sin(theta) * s, std::cos(phi));
}

// hexagonalDimplePattern / distanceToNearestDimple, shader_golf.cu:15-43
real dimple(V3 p) {
   V3 q = unit(p);
   real best = -1;
   for (int i = 0; i < 150; ++i) best = mx(best, dot(q, fib_point(i, 150)));
   best = cl(best, -1.0, 1.0);
   real ang = std::acos(best);
   if (ang < 0.24) return -0.35 * std::cos(ang / 0.24 * kPi * 0.5);
   return 0.0;
}

real smoothstep(real e0, real e1, real x) {
   real t = cl((x - e0) / (e1 - e0), 0.0, 1.0);
   return t * t * (3 - 2 * t);
}

// hit_golf_ball_sphere, shader_golf.cu:53-140
bool isect_golf(V3 c, real rad, const Ray& r, real tmin, real tmax, Hit& h) {
   if (!isect_sphere(c, rad, r, tmin, tmax, h)) return false;
   V3 sp = h.p;
   real disp = dimple(unit(sp - c));
   V3 bo = unit(sp - c);
   real dn = mn(1.0, mx(0.0, -disp / 0.35));
   real push = rad * 0.35 * (1.0 - dn);
   h.p = sp + push * bo;
   V3 bn = unit(h.p - c);
   if (disp < -0.001) {
      V3 helper = std::fabs(bn.x) > 0.8 ? v3(0, 1, 0) : v3(1, 0, 0);
      V3 t1 = unit(cross(helper, bn));
      V3 t2 = cross(bn, t1);
      const real hs = 0.015;
      real d0 = dimple(bn);
      real d1 = dimple(unit(bn + hs * t1));
      real d2 = dimple(unit(bn + hs * t2));
      V3 g = ((d1 - d0) / hs) * t1 + ((d2 - d0) / hs) * t2;
      V3 dnv = -0.2 * g;
      real ndv = mx(0.0, dot(bn, unit(-r.d)));
      dnv = smoothstep(0.1, 0.4, ndv) * dnv;
      real l = len(dnv);
      if (l > 0.4 && l > 1e-6) dnv = (0.4 / l) * dnv;
      V3 pert = unit(bn + dnv);
      if (dot(pert, bn) < 0) pert = -pert;
      if (std::isnan(pert.x) || std::isnan(pert.y) || std::isnan(pert.z)) pert = bn;
      h.normal = pert;
   } else {
      h.normal = bn;
   }
   real ndotv = dot(r.d, h.normal);
   if (std::isnan(ndotv)) {
      h.normal = bn;
      ndotv = dot(r.d, h.normal);
   }
   h.front = ndotv < 0;
   if (!h.front) h.normal = -h.normal;
   h.p = h.p + 1e-3 * h.normal;
   return true;
}

// ---- SDF field (spec; primitive formulas of cpu_shapes/sdf.hpp:105-143, smooth union :410-414)
real sdf_prim(const SdfPrim& s, V3 p) {
   V3 o = p - s.c;
   switch (s.type) {
   case RAYON_SDF_SPHERE:
      return len(o) - s.p[0];
   case RAYON_SDF_ROUNDED_BOX: {
      real rr = s.p[3];
      V3 q = v3(std::fabs(o.x) - (s.p[0] - rr), std::fabs(o.y) - (s.p[1] - rr), std::fabs(o.z) - (s.p[2] - rr));
      V3 m = v3(mx(q.x, 0.0), mx(q.y, 0.0), mx(q.z, 0.0));
      return len(m) + mn(mx(q.x, mx(q.y, q.z)), 0.0) - rr;
   }
   default: {
      real qx = std::sqrt(o.x * o.x + o.z * o.z) - s.p[0];
      return std::sqrt(qx * qx + o.y * o.y) - s.p[1];
   }
   }
}

real smooth_union(real d1, real d2, real k) {
   real h = cl(0.5 + 0.5 * (d2 - d1) / k, 0.0, 1.0);
   return d2 * (1.0 - h) + d1 * h - k * h * (1.0 - h);
}

real sdf_field(const oracle_scene& s, V3 p) {
   real d = sdf_prim(s.sdf[0], p);
   for (size_t i = 1; i < s.sdf.size(); ++i) d = smooth_union(d, sdf_prim(s.sdf[i], p), s.sdf_k);
   return d;
}

// Sphere tracing along the unit direction (rayMarch, sdf.hpp:518-563) from the entry of
// the field's bounding box; normal by central differences (estimateNormal, :499-507).
bool isect_sdf(const oracle_scene& s, const Prim& pr, const Ray& r, real tmin, real tmax, Hit& h) {
   real dl = len(r.d);
   V3 dn = (1.0 / dl) * r.d;
   real t0 = tmin * dl, t1 = tmax * dl;
   for (int k = 0; k < 3; ++k) {
      real o = axis(r.o, k), d = axis(dn, k);
      real inv = 1.0 / d;
      real a = (axis(pr.lo, k) - o) * inv, b = (axis(pr.hi, k) - o) * inv;
      if (a > b) std::swap(a, b);
      if (std::isnan(a)) continue;
      t0 = mx(t0, a);
      t1 = mn(t1, b);
   }
   if (t0 > t1) return false;
   real t = t0;
   for (int step = 0; step < s.sdf_steps; ++step) {
      V3 p = r.o + t * dn;
      real dist = sdf_field(s, p);
      if (dist < s.sdf_eps) {
         const real e = s.sdf_eps;
         V3 g = v3(sdf_field(s, p + v3(e, 0, 0)) - sdf_field(s, p - v3(e, 0, 0)),
                   sdf_field(s, p + v3(0, e, 0)) - sdf_field(s, p - v3(0, e, 0)),
                   sdf_field(s, p + v3(0, 0, e)) - sdf_field(s, p - v3(0, 0, e)));
         V3 n = unit(g);
         h.t = t / dl;
         h.p = p;
         h.outward = n;
         h.front = dot(r.d, n) < 0;
         h.normal = h.front ? n : -n;
         V3 tg = cross(v3(0, 1, 0), n);
         if (len2(tg) < 1e-6) tg = v3(1, 0, 0);
         h.tangent = unit(tg);
         return true;
      }
      t += dist;
      if (t > t1) break;
   }
   return false;
}

bool isect_prim(const oracle_scene& s, const Prim& p, const Ray& r, real tmin, real tmax, Hit& h) {
   switch (p.type) {
   case P_SPHERE:
      return isect_sphere(p.a, p.b.x, r, tmin, tmax, h);
   case P_RECT:
      return isect_rect(p.a, p.b, p.c, r, tmin, tmax, h);
   case P_TRI:
      return isect_tri(p, r, tmin, tmax, h);
   case P_DISPLACED:
      return isect_golf(p.a, p.b.x, r, tmin, tmax, h);
   case P_SDF:
      return isect_sdf(s, p, r, tmin, tmax, h);
   default:
      return false;
   }
}

// closest hit; equal t goes to the larger primitive index (linear-scan order of
// cuda_raytracer.cuh:568-585, inclusive bound :215-220).
bool closest(const oracle_scene& s, const Ray& r, real tmin, real tmax, bool brute, Hit& best) {
   bool any = false;
   best.t = tmax;
   best.prim = -1;
   Hit h;
   auto consider = [&](int64_t pi) {
      if (isect_prim(s, s.prims[pi], r, tmin, best.t, h)) {
         if (!any || h.t < best.t || (h.t == best.t && pi > best.prim)) {
            best = h;
            best.prim = pi;
            any = true;
         }
      }
   };
   if (brute || s.nodes.empty()) {
      for (int64_t i = 0; i < (int64_t)s.prims.size(); ++i) consider(i);
      return any;
   }
   std::vector<int64_t> stack;
   stack.reserve(128);
   stack.push_back(0);
   while (!stack.empty()) {
      int64_t ni = stack.back();
      stack.pop_back();
      const Node& n = s.nodes[ni];
      if (!hit_box(r, n.lo, n.hi, tmin, best.t)) continue;
      if (n.count > 0) {
         for (int64_t i = n.first; i < n.first + n.count; ++i) consider(s.order[i]);
      } else {
         stack.push_back(n.left);
         stack.push_back(n.right);
      }
   }
   return any;
}

// spec: the emitters next-event estimation samples — LIGHT-material rectangles, triangles
// and spheres. Shadow rays pass through them, and a BSDF-sampled path that reaches one
// after a Lambertian bounce does not add its emission (NEE already counted it). Every
// other emitter (golf-ball or SDF lights, CONSTANT and SHOW_NORMALS surfaces) is reached
// by BSDF sampling only, so its emission always counts and it occludes shadow rays.
bool nee_light(const oracle_scene& s, const Prim& p) {
   return p.mat >= 0 && p.mat < (int)s.mats.size() && s.mats[p.mat].type == RAYON_MAT_LIGHT &&
          (p.type == P_RECT || p.type == P_TRI || p.type == P_SPHERE);
}

bool any_hit(const oracle_scene& s, const Ray& r, real tmin, real tmax) {
   Hit h;
   auto test = [&](int64_t pi) {
      const Prim& p = s.prims[pi];
      if (nee_light(s, p)) return false;
      return isect_prim(s, p, r, tmin, tmax, h);
   };
   if (s.nodes.empty()) {
      for (int64_t i = 0; i < (int64_t)s.prims.size(); ++i)
         if (test(i)) return true;
      return false;
   }
   std::vector<int64_t> stack;
   stack.reserve(128);
   stack.push_back(0);
   while (!stack.empty()) {
      int64_t ni = stack.back();
      stack.pop_back();
      const Node& n = s.nodes[ni];
      if (!hit_box(r, n.lo, n.hi, tmin, tmax)) continue;
      if (n.count > 0) {
         for (int64_t i = n.first; i < n.first + n.count; ++i)
            if (test(s.order[i])) return true;
      } else {
         stack.push_back(n.left);
         stack.push_back(n.right);
      }
   }
   return false;
}

// ------------------------------------------------------------------------------------
// Materials
// ------------------------------------------------------------------------------------
inline V3 reflect(V3 v, V3 n) { return v - (2 * dot(v, n)) * n; }  // cuda_raytracer.cuh:84

V3 refract(V3 uv, V3 n, real eta) {  // cuda_raytracer.cuh:86-92
   real c = mn(dot(-uv, n), 1.0);
   V3 perp = eta * (uv + c * n);
   V3 par = (-std::sqrt(std::fabs(1.0 - len2(perp)))) * n;
   return perp + par;
}

real schlick(real c, real ri) {  // cuda_raytracer.cuh:94-99
   real r0 = (1 - ri) / (1 + ri);
   r0 *= r0;
   return r0 + (1 - r0) * std::pow(1 - c, 5);
}

void onb(V3 n, V3& u, V3& v) {  // cuda_utils.cuh:99-107
   if (std::fabs(n.x) > std::fabs(n.z))
      u = unit(v3(-n.y, n.x, 0));
   else
      u = unit(v3(0, -n.z, n.y));
   v = cross(n, u);
}

struct Draws {  // random numbers in consumption order
   Rng* rng;
   const float* fixed;
   int nfixed;
   int used = 0;
   real next() {
      ++used;
      if (rng) return rng->next();
      return used - 1 < nfixed ? fixed[used - 1] : 0.5;
   }
};

V3 on_sphere(Draws& R) {  // randOnUnitSphere, cuda_utils.cuh:66-77
   real th = 2 * kPi * R.next();
   real ph = std::acos(1 - 2 * R.next());
   return v3(std::sin(ph) * std::cos(th), std::sin(ph) * std::sin(th), std::cos(ph));
}

V3 cosine_dir(V3 n, Draws& R) {  // Lambert lobe, cuda_raytracer.cuh:621-638
   V3 w = unit(n), u, v;
   onb(w, u, v);
   real u1 = R.next(), u2 = R.next();
   real r = std::sqrt(u1), th = 2 * kPi * u2;
   return (r * std::cos(th)) * u + (r * std::sin(th)) * v + std::sqrt(mx(0.0, 1 - u1)) * w;
}

// microfacet_ggx.cuh:21-39
real lambda_ggx(V3 w, real ax, real ay) {
   real wz2 = w.z * w.z;
   if (wz2 < 1e-16) return 0;
   real a2 = (ax * w.x) * (ax * w.x) + (ay * w.y) * (ay * w.y);
   return (std::sqrt(1 + a2 / wz2) - 1) * (g_fault == FAULT_LAMBDA_HALF ? 1.0 : 0.5);
}
real g1_ggx(V3 w, real ax, real ay) { return 1 / (1 + lambda_ggx(w, ax, ay)); }
real g2_ggx(V3 wo, V3 wi, real ax, real ay) {
   return 1 / (1 + lambda_ggx(wo, ax, ay) + lambda_ggx(wi, ax, ay));
}

// VNDF sampling, microfacet_ggx.cuh:45-76
V3 sample_vndf(V3 wo, real ax, real ay, real u1, real u2) {
   if (g_fault == FAULT_VNDF_SWAP_ALPHA) std::swap(ax, ay);
   V3 wh = unit(v3(ax * wo.x, ay * wo.y, wo.z));
   if (wh.z < 0) wh = -wh;
   V3 t1 = wh.z < 0.99999 ? unit(cross(v3(0, 0, 1), wh)) : v3(1, 0, 0);
   V3 t2 = cross(wh, t1);
   real r = std::sqrt(u1), phi = 2 * kPi * u2;
   real px = r * std::cos(phi), py = r * std::sin(phi);
   real hh = std::sqrt(mx(0.0, 1 - px * px));
   real sm = (1 + wh.z) * 0.5;
   if (g_fault != FAULT_VNDF_NO_WARP) py = (1 - sm) * hh + sm * py;
   real pz = std::sqrt(mx(0.0, 1 - px * px - py * py));
   V3 nh = px * t1 + py * t2 + pz * wh;
   return unit(v3(ax * nh.x, ay * nh.y, mx(1e-6, nh.z)));
}

// FrComplex, microfacet_ggx.cuh:83-111
real fr_complex1(real cos2, real sin2, real ci, real eta, real k) {
   real eta2 = eta * eta, k2 = k * k;
   real t0 = eta2 - k2 - sin2;
   real ab = std::sqrt(mx(t0 * t0 + 4 * eta2 * k2, 0.0));
   real a = std::sqrt(mx((ab + t0) * 0.5, 0.0));
   real rs = (ab + cos2 - 2 * a * ci) / mx(ab + cos2 + 2 * a * ci, 1e-10);
   const real sg = g_fault == FAULT_CONDUCTOR_RP_SIGN ? -1.0 : 1.0;
   real rp = rs * (ab * cos2 + sin2 * sin2 - sg * 2 * a * ci * sin2) /
               mx(ab * cos2 + sin2 * sin2 + 2 * a * ci * sin2, 1e-10);
   return (rs + rp) * 0.5;
}
V3 fr_complex(real ci, V3 eta, V3 k) {
   ci = cl(ci, 0.0, 1.0);
   real c2 = ci * ci, s2 = 1 - c2;
   return v3(fr_complex1(c2, s2, ci, eta.x, k.x), fr_complex1(c2, s2, ci, eta.y, k.y),
             fr_complex1(c2, s2, ci, eta.z, k.z));
}

// spec: exact unpolarised dielectric Fresnel; eta = n_incident / n_transmitted.
real fresnel_dielectric(real ci, real eta) {
   real s2t = eta * eta * mx(0.0, 1 - ci * ci);
   if (s2t >= 1) return 1;
   real ct = std::sqrt(1 - s2t);
   real rs = (eta * ci - ct) / (eta * ci + ct);
   real rp = (ci - eta * ct) / (ci + eta * ct);
   if (g_fault == FAULT_DIELECTRIC_RP) rp = rs;
   return 0.5 * (rs * rs + rp * rp);
}

struct ScatterIn {
   V3 p, n, tangent, dir;
   bool front;
};

struct ScatterOut {
   Ray ray;
   V3 atten{1, 1, 1}, emitted{0, 0, 0};
   bool scattered = false;
   bool lambert = false;
};

struct Knobs {
   uint32_t flags;
   real fuzz, glass_ior, light;
};

// scatter_material, cuda_raytracer.cuh:613-850, after apply_material's type mapping
// (:418-485); GGX conductor / dielectric are spec additions.
void scatter(const rayon_material& m, V3 color, const Knobs& K, const ScatterIn& in, Draws& R, ScatterOut& o) {
   const bool quirks = (K.flags & RAYON_FLAG_REF_QUIRKS) != 0;
   const V3 n = in.n;
   o.emitted = v3(0, 0, 0);
   o.lambert = false;
   switch (m.type) {
   case RAYON_MAT_LAMBERTIAN:
   case RAYON_MAT_SDF_MATERIAL: {
      V3 d = cosine_dir(n, R);
      o.ray = {in.p + 1e-4 * n, d};
      o.atten = color;
      o.scattered = true;
      o.lambert = true;
      return;
   }
   case RAYON_MAT_MIRROR: {
      V3 rf = reflect(unit(in.dir), n);
      o.ray = {in.p, rf};
      o.atten = color;
      o.scattered = dot(rf, n) > 0;
      return;
   }
   case RAYON_MAT_METAL:
   case RAYON_MAT_ROUGH_MIRROR: {
      real rough = m.roughness;
      real fz = 1.0;
      if (quirks) {
         rough = rough > 0 ? rough : 0.3;  // cuda_raytracer.cuh:432-437
         fz = K.fuzz;                      // g_metal_fuzziness
      }
      V3 pn = unit(n + (rough * fz) * on_sphere(R));
      V3 rf = reflect(unit(in.dir), pn);
      o.ray = {in.p, rf};
      o.atten = color;
      o.scattered = dot(rf, n) > 0;
      return;
   }
   case RAYON_MAT_GLASS:
   case RAYON_MAT_DIELECTRIC: {
      V3 ud = unit(in.dir);
      real ri = quirks ? K.glass_ior : m.refractive_index;
      real ratio = in.front ? 1.0 / ri : ri;
      real ct = mn(dot(-ud, n), 1.0);
      real st = std::sqrt(1 - ct * ct);
      bool tir = ratio * st > 1;
      V3 d;
      if (tir || schlick(ct, ratio) > R.next())
         d = reflect(ud, n);
      else
         d = refract(ud, n, ratio);
      o.ray = {in.p, d};
      o.atten = v3(1, 1, 1);
      o.scattered = true;
      return;
   }
   case RAYON_MAT_LIGHT:
      o.emitted = K.light * fv(m.emission);
      o.scattered = false;
      return;
   case RAYON_MAT_CONSTANT:
      o.emitted = color;
      o.scattered = false;
      return;
   case RAYON_MAT_SHOW_NORMALS:
      o.emitted = 0.5 * (n + v3(1, 1, 1));
      o.scattered = false;
      return;
   case RAYON_MAT_ANISOTROPIC_METAL:
   case RAYON_MAT_GGX_CONDUCTOR: {
      const bool aniso = m.type == RAYON_MAT_ANISOTROPIC_METAL;
      real ax, ay;
      if (aniso) {
         real aspect = std::sqrt(mx(1e-4, 1 - 0.9 * (real)m.anisotropy));
         real r2 = (real)m.roughness * m.roughness;
         ax = mx(1e-4, r2 / aspect);
         ay = mx(1e-4, r2 * aspect);
      } else {
         ax = ay = mx(1e-4, (real)m.roughness * m.roughness);
      }
      V3 eta = fv(m.eta), kk = fv(m.k);
      if (mx(ax, ay) < 1e-3) {
         V3 rf = reflect(unit(in.dir), n);
         o.ray = {in.p + 1e-4 * n, rf};
         real ci = mx(dot(-unit(in.dir), n), 0.0);
         o.atten = color * fr_complex(ci, eta, kk);
         o.scattered = dot(rf, n) > 0;
         return;
      }
      V3 N = unit(n), T, B;
      if (aniso) {
         T = unit(in.tangent - dot(in.tangent, N) * N);
         B = cross(N, T);
      } else {
         onb(N, T, B);
      }
      V3 wo = unit(-in.dir);
      V3 wol = v3(dot(wo, T), dot(wo, B), dot(wo, N));
      if (wol.z <= 0) {
         o.scattered = false;
         return;
      }
      real u1 = R.next(), u2 = R.next();
      V3 wm = sample_vndf(wol, ax, ay, u1, u2);
      V3 wil = reflect(-wol, wm);
      if (wil.z <= 0) {
         o.scattered = false;
         return;
      }
      V3 wi = wil.x * T + wil.y * B + wil.z * N;
      V3 F = fr_complex(mx(dot(wol, wm), 0.0), eta, kk);
      real G = g2_ggx(wol, wil, ax, ay), G1 = g1_ggx(wol, ax, ay);
      o.ray = {in.p + 1e-4 * n, wi};
      o.atten = color * ((G / mx(G1, 1e-6)) * F);
      o.scattered = true;
      return;
   }
   case RAYON_MAT_GGX_DIELECTRIC: {
      real a = mx(1e-4, (real)m.roughness * m.roughness);
      real ri = m.refractive_index;
      real eta = in.front ? 1.0 / ri : ri;
      V3 N = unit(n), T, B;
      onb(N, T, B);
      V3 wo = unit(-in.dir);
      V3 wol = v3(dot(wo, T), dot(wo, B), dot(wo, N));
      if (wol.z <= 0) {
         o.scattered = false;
         return;
      }
      real u1 = R.next(), u2 = R.next();
      V3 wm = sample_vndf(wol, a, a, u1, u2);
      real com = dot(wol, wm);
      real F = fresnel_dielectric(com, eta);
      V3 wil;
      bool refl = R.next() < F;
      if (refl) {
         wil = reflect(-wol, wm);
         if (wil.z <= 0) {
            o.scattered = false;
            return;
         }
      } else {
         wil = refract(-wol, wm, eta);
         if (wil.z >= 0) {
            o.scattered = false;
            return;
         }
      }
      V3 wi = wil.x * T + wil.y * B + wil.z * N;
      real G = g2_ggx(wol, wil, a, a), G1 = g1_ggx(wol, a, a);
      o.ray = {refl ? in.p + 1e-4 * n : in.p - 1e-4 * n, wi};
      o.atten = (G / mx(G1, 1e-6)) * color;
      o.scattered = true;
      return;
   }
   case RAYON_MAT_THIN_FILM: {  // cuda_raytracer.cuh:753-809
      V3 ud = unit(in.dir);
      real ci = mx(std::fabs(dot(-ud, n)), 0.001);
      real n0 = m.refractive_index, n1 = m.film_ior, d = m.film_thickness;
      real si = std::sqrt(mx(0.0, 1 - ci * ci));
      real st = (n0 / n1) * si;
      real ct = st >= 1 ? 0 : std::sqrt(1 - st * st);
      real r01 = (n0 - n1) / (n0 + n1);
      r01 *= r01;
      real x01 = 1 - ci;
      real R01 = r01 + (1 - r01) * x01 * x01 * x01 * x01 * x01;
      real r12 = (n1 - n0) / (n1 + n0);
      r12 *= r12;
      real x12 = 1 - ct;
      real R12 = r12 + (1 - r12) * x12 * x12 * x12 * x12 * x12;
      real sq = std::sqrt(R01 * R12);
      auto airy = [&](real lam) {
         real delta = 4 * kPi * n1 * d * ct / lam;
         real c = std::cos(delta);
         return (R01 + R12 + 2 * sq * c) / mx(1 + R01 * R12 + 2 * sq * c, 1e-8);
      };
      real Rr = airy(650), Rg = airy(550), Rb = airy(450);
      real avg = (Rr + Rg + Rb) / 3;
      if (R.next() < avg) {
         V3 rf = reflect(ud, n);
         o.ray = {in.p + 1e-4 * n, rf};
         o.atten = (1 / mx(avg, 0.001)) * v3(Rr, Rg, Rb);
         o.scattered = dot(rf, n) > 0;
      } else {
         o.ray = {in.p - 1e-4 * n, ud};
         o.atten = (1 / mx(1 - avg, 0.001)) * v3(1 - Rr, 1 - Rg, 1 - Rb);
         o.scattered = true;
      }
      return;
   }
   case RAYON_MAT_CLEAR_COAT: {  // cuda_raytracer.cuh:810-846 (+ ior default of :478)
      V3 ud = unit(in.dir);
      real ct = mn(dot(-ud, n), 1.0);
      real ior = m.refractive_index > 1 ? m.refractive_index : 1.5;
      real F = schlick(ct, ior);
      if (R.next() < F) {
         V3 pn = m.roughness > 1e-3 ? unit(n + (real)m.roughness * on_sphere(R)) : n;
         V3 rf = reflect(ud, pn);
         o.ray = {in.p + 1e-4 * n, rf};
         o.atten = v3(1, 1, 1);
         o.scattered = dot(rf, n) > 0;
      } else {
         V3 d = cosine_dir(n, R);
         o.ray = {in.p + 1e-4 * n, d};
         o.atten = color;
         o.scattered = true;
      }
      return;
   }
   default:
      o.scattered = false;
      return;
   }
}

// apply_procedural_pattern, cuda_raytracer.cuh:377-416
V3 surface_color(const rayon_material& m, V3 p, V3 center) {
   V3 base = fv(m.albedo);
   if (m.pattern != RAYON_PATTERN_FIBONACCI_DOTS) return base;
   V3 dir = unit(p - center);
   int N = (int)m.pattern_param1;
   real best = -1;
   for (int i = 0; i < N; ++i) best = mx(best, dot(dir, fib_point(i, N)));
   best = cl(best, -1.0, 1.0);
   real mask = std::acos(best) < m.pattern_param2 ? 0.0 : 1.0;
   return mask * base + (1 - mask) * fv(m.pattern_color);
}

// spec: uniform point on an emitter; returns point, geometric normal, area.
void sample_light(const Prim& p, real a, real b, V3& q, V3& nl, real& area) {
   switch (p.type) {
   case P_RECT:
      q = p.a + a * p.b + b * p.c;
      nl = cross(p.b, p.c);
      area = len(nl);
      nl = unit(nl);
      return;
   case P_TRI: {
      real su = std::sqrt(a);
      q = (1 - su) * p.a + (su * (1 - b)) * p.b + (su * b) * p.c;
      nl = cross(p.b - p.a, p.c - p.a);
      area = 0.5 * len(nl);
      nl = unit(nl);
      return;
   }
   default: {  // sphere: uniform on the surface
      real z = 1 - 2 * a, r = std::sqrt(mx(0.0, 1 - z * z)), ph = 2 * kPi * b;
      nl = v3(r * std::cos(ph), r * std::sin(ph), z);
      q = p.a + p.b.x * nl;
      area = 4 * kPi * p.b.x * p.b.x;
      return;
   }
   }
}

struct Cam {
   V3 C, p00, du, dv, u, v;
};

// random_in_unit_disk (cuda_raytracer.cuh:114-122): rejection from the square [-1,1)^2,
// two draws per attempt.
void unit_disk(Draws& R, real& dx, real& dy) {
   do {
      dx = 2 * R.next() - 1;
      dy = 2 * R.next() - 1;
   } while (dx * dx + dy * dy >= 1 && g_fault != FAULT_DISK_NO_REJECT);
}

// One camera sample: render_acc_kernel.cu:198-229 + ray_color, cuda_raytracer.cuh:858-933
// (+ spec NEE on Lambertian surfaces).
V3 trace_sample(const oracle_scene& s, const rayon_job_params& P, const Cam& cam, int64_t x, int64_t y, Rng& rng,
                int64_t& segments) {
   Draws R{&rng, nullptr, 0};
   const Knobs K{P.flags, P.metal_fuzziness, P.glass_ior, P.light_intensity};
   const bool nee = (P.flags & RAYON_FLAG_NEE) && !s.lights.empty();
   real ou = R.next() - 0.5;
   real ov = R.next() - 0.5;
   V3 pc = cam.p00 + ((real)x + ou) * cam.du + ((real)y + ov) * cam.dv;
   Ray ray{cam.C, pc - cam.C};
   if ((P.flags & RAYON_FLAG_DOF) && P.dof_aperture > 0) {  // render_acc_kernel.cu:209-221
      V3 nd = unit(ray.d);
      V3 focus = cam.C + (real)P.dof_focus_distance * nd;
      real dx, dy;
      unit_disk(R, dx, dy);
      V3 off = (real)P.dof_aperture * (dx * cam.u + dy * cam.v);
      ray.o = cam.C + off;
      ray.d = focus - ray.o;
   }
   V3 L = v3(0, 0, 0), T = v3(1, 1, 1);
   bool prev_lambert = false;
   for (int bounce = 0; bounce < P.max_depth; ++bounce) {
      ++segments;
      Hit h;
      if (!closest(s, ray, 0.001, 3.402823466e38, false, h)) {
         V3 ud = unit(ray.d);
         real t = 0.5 * (ud.y + 1);
         V3 sky = (1 - t) * v3(1, 1, 1) + t * v3(0.5, 0.7, 1.0);
         L = L + (real)P.background_intensity * (T * sky);
         break;
      }
      const Prim& pr = s.prims[h.prim];
      if (!pr.visible && bounce == 0) {
         ray = {h.p + 0.01 * ray.d, ray.d};
         continue;
      }
      const rayon_material& m = s.mats[pr.mat];
      V3 center = (pr.type == P_SPHERE || pr.type == P_DISPLACED) ? pr.a : v3(0, 0, 0);
      V3 color = surface_color(m, h.p, center);
      ScatterIn in{h.p, h.normal, h.tangent, ray.d, h.front};
      ScatterOut o;
      scatter(m, color, K, in, R, o);
      if (len2(o.emitted) > 0 && !(nee && prev_lambert && nee_light(s, pr))) L = L + T * o.emitted;
      if (!o.scattered) break;
      if (nee && o.lambert) {
         real ul = R.next(), ua = R.next(), ub = R.next();
         int64_t li = mn((int64_t)(ul * (real)s.lights.size()), (int64_t)s.lights.size() - 1);
         const Prim& lp = s.prims[s.lights[li].prim];
         V3 q, nl;
         real area;
         sample_light(lp, ua, ub, q, nl, area);
         V3 org = h.p + 1e-4 * h.normal;
         V3 wd = q - org;
         real d2 = len2(wd);
         V3 wi = (1.0 / std::sqrt(d2)) * wd;
         real cs = dot(h.normal, wi);
         real cl = lp.type == P_SPHERE ? dot(nl, -wi) : std::fabs(dot(nl, -wi));
         if (cs > 0 && cl > 0) {
            ++segments;
            if (!any_hit(s, {org, wd}, 1e-4, 1 - 1e-4)) {
               V3 Le = K.light * fv(s.mats[lp.mat].emission);
               real g = cs * cl / d2 * area * (real)s.lights.size() / kPi;
               L = L + g * (T * color * Le);
            }
         }
      }
      T = T * o.atten;
      prev_lambert = o.lambert;
      ray = o.ray;
      if (bounce > 0) {  // Russian roulette, cuda_raytracer.cuh:906-920
         real q = mn(maxc(T), 0.95);
         if (R.next() > q) break;
         T = (1.0 / q) * T;
      }
   }
   return L;
}

Cam make_cam(const rayon_job_params& P, const rayon_camera* c) {
   Cam k;
   if (c) {
      k.C = dv(c->center);
      k.p00 = dv(c->pixel00);
      k.du = dv(c->delta_u);
      k.dv = dv(c->delta_v);
      k.u = dv(c->u);
      k.v = dv(c->v);
   } else {
      k.C = fv(P.camera_center);
      k.p00 = fv(P.pixel00);
      k.du = fv(P.delta_u);
      k.dv = fv(P.delta_v);
      k.u = fv(P.cam_u);
      k.v = fv(P.cam_v);
   }
   return k;
}

template <class F>
void parallel_for(int64_t n, int threads, F&& f) {
   if (threads <= 0) threads = (int)mx(1u, std::thread::hardware_concurrency());
   threads = (int)std::min<int64_t>(threads, std::max<int64_t>(1, n));
#ifdef ORACLE_STATIC_ROWS
   // the reference CPU renderer's schedule (renderer_cpu_parallel.hpp:56-63): one contiguous
   // chunk of the (row-major) pixel list per thread, the last one taking the remainder
   {
      const int64_t per = std::max<int64_t>(1, n / threads);
      std::vector<std::thread> pool;
      for (int t = 0; t < threads; ++t) {
         const int64_t a = t * per, b = t == threads - 1 ? n : std::min<int64_t>(n, a + per);
         pool.emplace_back([&f, a, b] {
            for (int64_t i = a; i < b; ++i) f(i);
         });
      }
      for (auto& t : pool) t.join();
      return;
   }
#endif
   std::atomic<int64_t> next{0};
   auto work = [&]() {
      for (;;) {
         int64_t i0 = next.fetch_add(16);
         if (i0 >= n) break;
         for (int64_t i = i0; i < mn(n, i0 + 16); ++i) f(i);
      }
   };
   if (threads == 1) {
      work();
      return;
   }
   std::vector<std::thread> pool;
   for (int t = 0; t < threads; ++t) pool.emplace_back(work);
   for (auto& t : pool) t.join();
}

}  // namespace

extern "C" {

oracle_scene* oracle_scene_create(const rayon_scene_desc* d) {
   auto* s = new oracle_scene();
   s->mats.assign(d->materials, d->materials + d->num_materials);
   int64_t id = 0;
   for (int64_t i = 0; i < d->num_geometries; ++i) {
      const rayon_geometry& g = d->geometries[i];
      Prim p{};
      p.type = g.type;
      p.mat = g.material_id;
      p.visible = g.visible != 0;
      p.has_n = g.has_normals != 0;
      if (g.type == RAYON_GEOM_SPHERE || g.type == RAYON_GEOM_DISPLACED_SPHERE) {
         p.a = fv(g.p);
         p.b = v3(g.p[3], 0, 0);
      } else {
         p.a = fv(g.p);
         p.b = fv(g.p + 3);
         p.c = fv(g.p + 6);
         p.n0 = fv(g.n);
         p.n1 = fv(g.n + 3);
         p.n2 = fv(g.n + 6);
      }
      p.id = id++;
      prim_bounds(p);
      s->prims.push_back(p);
   }
   for (int mi = 0; mi < d->num_meshes; ++mi) {
      const rayon_mesh& m = d->meshes[mi];
      for (int64_t t = 0; t < m.num_triangles; ++t) {
         Prim p{};
         p.type = P_TRI;
         p.mat = m.material_id;
         p.visible = m.visible != 0;
         const int32_t* ix = m.indices + 3 * t;
         p.a = fv(m.positions + 3 * (int64_t)ix[0]);
         p.b = fv(m.positions + 3 * (int64_t)ix[1]);
         p.c = fv(m.positions + 3 * (int64_t)ix[2]);
         if (m.normals) {
            p.has_n = true;
            p.n0 = fv(m.normals + 3 * (int64_t)ix[0]);
            p.n1 = fv(m.normals + 3 * (int64_t)ix[1]);
            p.n2 = fv(m.normals + 3 * (int64_t)ix[2]);
         }
         p.id = id++;
         prim_bounds(p);
         s->prims.push_back(p);
      }
   }
   if (d->num_sdf_prims > 0) {
      s->sdf_k = d->sdf_smooth_k;
      s->sdf_steps = d->sdf_max_steps;
      s->sdf_eps = d->sdf_eps;
      Prim p{};
      p.type = P_SDF;
      p.mat = d->sdf_material_id;
      p.visible = true;
      p.lo = v3(1e300, 1e300, 1e300);
      p.hi = v3(-1e300, -1e300, -1e300);
      for (int i = 0; i < d->num_sdf_prims; ++i) {
         const rayon_sdf_prim& q = d->sdf_prims[i];
         SdfPrim sp{q.type, fv(q.center), {q.params[0], q.params[1], q.params[2], q.params[3]}};
         s->sdf.push_back(sp);
         real ex, ey, ez;
         if (q.type == RAYON_SDF_SPHERE)
            ex = ey = ez = q.params[0];
         else if (q.type == RAYON_SDF_ROUNDED_BOX)
            ex = q.params[0], ey = q.params[1], ez = q.params[2];
         else
            ex = ez = q.params[0] + q.params[1], ey = q.params[1];
         V3 c = fv(q.center);
         p.lo = v3(std::min(p.lo.x, c.x - ex), std::min(p.lo.y, c.y - ey), std::min(p.lo.z, c.z - ez));
         p.hi = v3(std::max(p.hi.x, c.x + ex), std::max(p.hi.y, c.y + ey), std::max(p.hi.z, c.z + ez));
      }
      // the surface of an n-way smooth union lies within (n - 1) k / 4 of the primitives
      const double pad = 0.25 * d->sdf_smooth_k * std::max(1, d->num_sdf_prims - 1) + 2 * d->sdf_eps;
      p.lo = p.lo - v3(pad, pad, pad);
      p.hi = p.hi + v3(pad, pad, pad);
      p.id = id++;
      s->prims.push_back(p);
   }
   s->order.resize(s->prims.size());
   for (size_t i = 0; i < s->order.size(); ++i) s->order[i] = (int64_t)i;
   if (!s->prims.empty()) build_node(*s, 0, (int64_t)s->prims.size());
   for (size_t i = 0; i < s->prims.size(); ++i) {
      const Prim& p = s->prims[i];
      if (nee_light(*s, p)) {
         s->lights.push_back({(int64_t)i, 0});
      }
   }
   return s;
}

void oracle_scene_free(oracle_scene* s) { delete s; }
int64_t oracle_scene_num_prims(const oracle_scene* s) { return (int64_t)s->prims.size(); }
int32_t oracle_scene_num_lights(const oracle_scene* s) { return (int32_t)s->lights.size(); }

void oracle_trace_closest(const oracle_scene* s, const float* origins, const float* dirs, int64_t n, double t_min,
                          double t_max, int32_t brute, int64_t* out_prim, double* out_t) {
   parallel_for(n, 0, [&](int64_t i) {
      Ray r{fv(origins + 3 * i), fv(dirs + 3 * i)};
      Hit h;
      if (closest(*s, r, t_min, t_max, brute != 0, h)) {
         out_prim[i] = s->prims[h.prim].id;
         out_t[i] = h.t;
      } else {
         out_prim[i] = -1;
         out_t[i] = -1;
      }
   });
}

double oracle_intersect_prim(const oracle_scene* s, int64_t prim, const float* o, const float* d, double t_min,
                             double t_max) {
   if (prim < 0 || prim >= (int64_t)s->prims.size()) return -1;
   Ray r{fv(o), fv(d)};
   Hit h;
   return isect_prim(*s, s->prims[prim], r, t_min, t_max, h) ? h.t : -1.0;
}

void oracle_trace_any(const oracle_scene* s, const float* origins, const float* dirs, int64_t n, double t_min,
                      double t_max, int32_t* out_hit) {
   parallel_for(n, 0, [&](int64_t i) {
      Ray r{fv(origins + 3 * i), fv(dirs + 3 * i)};
      out_hit[i] = any_hit(*s, r, t_min, t_max) ? 1 : 0;
   });
}

uint32_t oracle_ref_rng_init(int64_t pixel_index) { return (uint32_t)(1234u + (uint32_t)pixel_index * 747796405u); }
float oracle_ref_rng_next(uint32_t* st) {
   Rng r{*st, true};
   float f = r.next();
   *st = r.state;
   return f;
}
uint32_t oracle_counter_rng_init(uint64_t seed, uint32_t pixel, uint32_t sample) {
   return counter_init(seed, pixel, sample);
}
float oracle_counter_rng_next(uint32_t* st) {
   Rng r{*st, false};
   float f = r.next();
   *st = r.state;
   return f;
}

void oracle_render(const oracle_scene* s, const rayon_job_params* P, const rayon_camera* c, const int64_t* pixels,
                   int64_t npix, int32_t sample0, int32_t nsamples, int32_t threads, double* sums,
                   uint32_t* rng_state, int64_t* segments) {
   Cam cam = make_cam(*P, c);
   const bool refrng = (P->flags & RAYON_FLAG_REF_RNG) != 0;
   std::atomic<int64_t> segs{0};
   parallel_for(npix, threads, [&](int64_t i) {
      int64_t pix = pixels ? pixels[i] : i;
      int64_t x = pix % P->width, y = pix / P->width;
      int64_t local = 0;
      Rng rng{0, refrng};
      if (refrng) rng.state = rng_state[i];
      for (int32_t k = 0; k < nsamples; ++k) {
         if (!refrng) rng.state = counter_init(P->seed, (uint32_t)pix, (uint32_t)(sample0 + k));
         V3 L = trace_sample(*s, *P, cam, x, y, rng, local);
         sums[3 * i + 0] += L.x;
         sums[3 * i + 1] += L.y;
         sums[3 * i + 2] += L.z;
      }
      if (refrng) rng_state[i] = rng.state;
      segs.fetch_add(local, std::memory_order_relaxed);
   });
   if (segments) *segments = segs.load();
}

void oracle_render_samples(const oracle_scene* s, const rayon_job_params* P, const rayon_camera* c,
                           const int64_t* pixels, int64_t npix, int32_t sample0, int32_t nsamples, int32_t threads,
                           double* out) {
   Cam cam = make_cam(*P, c);
   parallel_for(npix, threads, [&](int64_t i) {
      int64_t pix = pixels ? pixels[i] : i;
      int64_t x = pix % P->width, y = pix / P->width;
      int64_t local = 0;
      for (int32_t k = 0; k < nsamples; ++k) {
         Rng rng{counter_init(P->seed, (uint32_t)pix, (uint32_t)(sample0 + k)), false};
         V3 L = trace_sample(*s, *P, cam, x, y, rng, local);
         double* o = out + 3 * (i * nsamples + k);
         o[0] = L.x;
         o[1] = L.y;
         o[2] = L.z;
      }
   });
}

void oracle_camera_frame(const double lf[3], const double la[3], const double up[3], double vfov, int32_t W,
                         int32_t H, rayon_camera* out) {
   V3 from = dv(lf), at = dv(la), vup = dv(up);
   double f = len(from - at);
   double h = std::tan(vfov * kPi / 180.0 / 2);
   double vh = 2 * h * f, vw = vh * ((double)W / H);
   V3 w = unit(from - at), u = unit(cross(vup, w)), v = cross(w, u);
   V3 vu = vw * u, vv = vh * (-v);
   V3 du = (1.0 / W) * vu, dvv = (1.0 / H) * vv;
   V3 ul = from - f * w - 0.5 * vu - 0.5 * vv;
   V3 p00 = ul + 0.5 * (du + dvv);
   auto put = [](double* d, V3 a) {
      d[0] = a.x;
      d[1] = a.y;
      d[2] = a.z;
   };
   put(out->center, from);
   put(out->pixel00, p00);
   put(out->delta_u, du);
   put(out->delta_v, dvv);
   put(out->u, u);
   put(out->v, v);
   put(out->w, w);
}

int32_t oracle_adaptive_ref(int32_t count_before, const double ab[3], const double aa[3], int32_t added,
                            int32_t min_samples, double thr) {
   if (count_before < 0) return count_before;
   int32_t nc = count_before + added;
   if (nc < min_samples) return nc;
   double lb = 0;
   if (count_before > 0) lb = (0.299 * ab[0] + 0.587 * ab[1] + 0.114 * ab[2]) / count_before;
   double la = (0.299 * aa[0] + 0.587 * aa[1] + 0.114 * aa[2]) / nc;
   double rel = std::fabs(la - lb) / std::max(la, 0.001);
   return rel < thr ? -nc : nc;
}

int32_t oracle_adaptive_stderr(double s1, double s2, int64_t n, int32_t min_samples, double thr) {
   if (n < min_samples || n < 2) return 0;
   double mean = s1 / n;
   double var = std::max(0.0, s2 / n - mean * mean) * n / (n - 1);
   double se = std::sqrt(var / n);
   return se / std::max(mean, 1e-3) < thr ? 1 : 0;
}

uint8_t oracle_gamma_u8(double sum, int64_t n, double gamma) {
   if (n <= 0) n = 1;
   double v = std::min(std::pow(std::max(sum / n, 0.0), 1.0 / gamma), 0.999);
   return (uint8_t)(256.0 * v);
}

// Surface colour of a material at point p of a sphere centred at `center` (the dots pattern).
void oracle_surface_color(const rayon_material* m, const double p[3], const double center[3], double out[3]) {
   const V3 c = surface_color(*m, dv(p), dv(center));
   out[0] = c.x;
   out[1] = c.y;
   out[2] = c.z;
}

// Golf-ball sphere hit: t, displaced point and shading normal; returns 0 on a miss.
int32_t oracle_golf_hit(const double c[3], double rad, const double o[3], const double d[3], double* t, double p[3],
                        double n[3]) {
   Ray r{dv(o), dv(d)};
   Hit h;
   if (!isect_golf(dv(c), rad, r, 0.001, 1e300, h)) return 0;
   *t = h.t;
   p[0] = h.p.x;
   p[1] = h.p.y;
   p[2] = h.p.z;
   n[0] = h.normal.x;
   n[1] = h.normal.y;
   n[2] = h.normal.z;
   return 1;
}

int32_t oracle_scatter_probe(const rayon_material* m, uint32_t flags, float fuzz, float glass_ior, float light,
                             const double dir[3], const double n[3], int32_t front, const double tangent[3],
                             const float* rnd, int32_t nrnd, double od[3], double at[3], double em[3],
                             int32_t* used) {
   Draws R{nullptr, rnd, nrnd};
   Knobs K{flags, fuzz, glass_ior, light};
   ScatterIn in{v3(0, 0, 0), dv(n), dv(tangent), dv(dir), front != 0};
   ScatterOut o;
   scatter(*m, fv(m->albedo), K, in, R, o);
   od[0] = o.ray.d.x;
   od[1] = o.ray.d.y;
   od[2] = o.ray.d.z;
   at[0] = o.atten.x;
   at[1] = o.atten.y;
   at[2] = o.atten.z;
   em[0] = o.emitted.x;
   em[1] = o.emitted.y;
   em[2] = o.emitted.z;
   *used = R.used;
   return o.scattered ? 1 : 0;
}

void oracle_set_fault(int32_t fault) { g_fault = fault; }

void oracle_sample_vndf(const double wo[3], double ax, double ay, const double* u, int64_t n, double* out) {
   const V3 w = dv(wo);
   for (int64_t i = 0; i < n; ++i) {
      V3 m = sample_vndf(w, ax, ay, u[2 * i], u[2 * i + 1]);
      out[3 * i] = m.x;
      out[3 * i + 1] = m.y;
      out[3 * i + 2] = m.z;
   }
}

double oracle_ggx_lambda(const double w[3], double ax, double ay) { return lambda_ggx(dv(w), ax, ay); }

double oracle_fresnel_dielectric(double ci, double eta) { return fresnel_dielectric(ci, eta); }

double oracle_fresnel_conductor(double ci, double eta, double k) {
   ci = std::clamp(ci, 0.0, 1.0);
   return fr_complex1(ci * ci, 1 - ci * ci, ci, eta, k);
}

int64_t oracle_unit_disk(const float* rnd, int64_t nrnd, int64_t n, double* out_xy) {
   Draws R{nullptr, rnd, (int)std::min<int64_t>(nrnd, 0x7fffffff)};
   for (int64_t i = 0; i < n; ++i) {
      real x, y;
      unit_disk(R, x, y);
      out_xy[2 * i] = x;
      out_xy[2 * i + 1] = y;
      if (R.used > R.nfixed) return -1;
   }
   return R.used;
}

void oracle_scatter_batch(const rayon_material* m, uint32_t flags, float fuzz, float glass_ior, float light,
                          const double dir[3], const double n[3], int32_t front, const double tangent[3],
                          const float* rnd, int32_t nrnd, int64_t count, double* out_dir, double* atten,
                          int32_t* scattered) {
   const Knobs K{flags, fuzz, glass_ior, light};
   const ScatterIn in{v3(0, 0, 0), dv(n), dv(tangent), dv(dir), front != 0};
   for (int64_t i = 0; i < count; ++i) {
      Draws R{nullptr, rnd + i * nrnd, nrnd};
      ScatterOut o;
      scatter(*m, fv(m->albedo), K, in, R, o);
      out_dir[3 * i] = o.ray.d.x;
      out_dir[3 * i + 1] = o.ray.d.y;
      out_dir[3 * i + 2] = o.ray.d.z;
      atten[3 * i] = o.atten.x;
      atten[3 * i + 1] = o.atten.y;
      atten[3 * i + 2] = o.atten.z;
      scattered[i] = o.scattered ? 1 : 0;
   }
}

}  // extern "C"
