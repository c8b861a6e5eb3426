// layout.h — device-resident scene and job layouts, shared by host and device code.
//
// HBM layout (one scene):
//   prims   : PrimRec[num_prims], 48 B, leaf order (the slot order of the BVH leaves,
//             as the reference reorders geometries, scene_description.hpp:735-777)
//   normals : float4[3*num_prims] vertex normals, only read for smooth triangles
//   nodes   : Node2[num_nodes], 64 B = one cache line per visited pair of children
//   wnodes  : WNode[num_qnodes], 8-wide quantized nodes (80 B), built on request
//   mats    : MatRec[num_mats], 96 B
//   lights  : LightRec per emitter (64 B)
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RY_HD __host__ __device__
#else
#define RY_HD
struct alignas(16) float4 {
   float x, y, z, w;
};
struct alignas(16) int4 {
   int x, y, z, w;
};
#endif

namespace ry {

// Job statistics are spread over this many 8-counter slots (picked by workgroup index)
// so concurrent waves do not serialise on one address; readers sum the slots.
constexpr int kCounterSlots = 64;

// Packed reference meaning "no node" (empty scene root, empty stack).
constexpr int kSentinel = (int)0x80000000;

// primitive types held in PrimRec::a.w bits 28..31
enum PrimKind : uint32_t { K_TRI = 0, K_SPHERE = 1, K_RECT = 2, K_DISPLACED = 3, K_SDF = 4 };
enum PrimFlag : uint32_t { F_VISIBLE = 1u << 24, F_NORMALS = 1u << 25, F_EMITTER = 1u << 26 };

// Primitive record, 3 x float4:
//   triangle : a = (v0, info)  b = (v1-v0, orig_id)  c = (v2-v0, 0)
//   sphere   : a = (center, info)  b = (radius, 0, 0, orig_id)
//   rectangle: a = (corner, info)  b = (u, orig_id)  c = (v, 0)
//   sdf field: a = (bounds lo, info) b = (bounds hi, orig_id)
// info = kind << 28 | flags | material (bits 0..23)
struct alignas(16) PrimRec {
   float4 a, b, c;
};

// Shading features of a scene beyond the lean set (triangles, spheres, rectangles; Lambert,
// mirror, metal, glass, dielectric, light, constant, normals, GGX conductor and
// dielectric). Kernels instantiated without a feature leave its code out, which frees
// registers; a pass takes such a kernel only when no scene in it uses the feature. The
// code kept is the same source in every instantiation, so a sample's bits do not depend on
// which one ran it.
enum SceneFeature : uint32_t {
   SF_DISPLACED = 1u,  // golf-ball sphere
   SF_SDF = 2u,        // SDF field
   SF_DOTS = 4u,       // Fibonacci-dots pattern
   SF_LAYERED = 8u,    // thin film, clear coat
   SF_ANISO = 16u,     // anisotropic metal
   SF_ALL = 0xFFFFFFFFu,
   SF_LEAN = 0u,
};

RY_HD inline uint32_t info_kind(uint32_t info) { return info >> 28; }
RY_HD inline uint32_t info_mat(uint32_t info) { return info & 0xFFFFFFu; }

// Binary BVH stored as pair nodes: node i holds the boxes of both children of one
// interior node. Child reference: interior -> node index (>= 0) with count 0;
// leaf -> first primitive slot with count in 1..8.
struct alignas(64) Node2 {
   float4 bx;  // c0.lo.x, c0.hi.x, c1.lo.x, c1.hi.x
   float4 by;
   float4 bz;
   int4 meta;  // ref0, ref1, count0, count1
};

// Packed child reference for stacks: interior >= 0, leaf = ~(first << 3 | (count - 1)).
RY_HD inline int pack_leaf(int first, int count) { return ~((first << 3) | (count - 1)); }

// 8-wide compressed node, 80 B (five 16-byte loads): the compressed wide BVH layout of
// Ylitie, Karras and Laine (HPG 2017). Child boxes are 8-bit offsets on a per-axis
// power-of-two grid anchored at the node origin p, rounded outward (lo down, hi up), so a
// decoded child box always contains the exact one. Internal children are stored
// contiguously from child_base in slot order; the primitives of the leaf children follow
// each other from prim_base. Slot s holds the child whose centroid lies in octant s of
// the node (bit k set: negative side of axis k), so a ray visits slots in the order
// i ^ (7 ^ octant(direction)) to go roughly front to back.
struct alignas(16) WNode {
   float px, py, pz;
   uint32_t e_imask;     // ex | ey << 8 | ez << 16 (biased exponents) | imask << 24
   uint32_t child_base;  // first internal child
   uint32_t prim_base;   // first primitive slot of the leaf children
   uint32_t meta[2];     // per slot: internal -> 0; leaf -> offset (bits 0..4) | (count-1) << 5;
                         // empty -> 0x80 (its box is also inverted: lo 255, hi 0)
   uint32_t q[12];       // qlo_x[8], qlo_y[8], qlo_z[8], qhi_x[8], qhi_y[8], qhi_z[8] as bytes
};

// Material record, 6 x float4.
struct alignas(16) MatRec {
   float4 m0;  // albedo, type (bits 0..7) | pattern (bits 8..15)
   float4 m1;  // emission, roughness
   float4 m2;  // eta, refractive_index
   float4 m3;  // k, anisotropy
   float4 m4;  // film_thickness, film_ior, pattern_param1, pattern_param2
   float4 m5;  // pattern_color, 0
};

// Emitter record for next-event estimation: the emitter's primitive record (as in prims)
// and its material's emission with the emitter's area, so a light sample is one load.
struct alignas(16) LightRec {
   PrimRec g;
   float4 e;  // emission (the material's m1.xyz), area
};

// SDF primitive in rounded-box form: distance = |max(q,0)| + min(max q, 0) - rounding with
// q = |a| - half, where a = p - center, or (|xz| - R, y, 0) for a torus.
//   sphere r        : half (0,0,0), rounding r, R = -1
//   rounded box h,r : half h - r,   rounding r, R = -1
//   torus R, r      : half (0,0,0), rounding r, R
struct alignas(16) SdfRec {
   float4 c;  // center, R (torus major radius; -1 for the other kinds)
   float4 p;  // half extents, rounding
};

// Device view of one scene.
struct SceneView {
   const PrimRec* prims;
   const float4* normals;
   const Node2* nodes;
   const MatRec* mats;
   const LightRec* lights;  // NEE emitter table
   const SdfRec* sdf;
   const float4* sdf_sph;  // per SDF primitive: bounding sphere centre and radius
   const WNode* wnodes;  // 8-wide quantized tree, or NULL to traverse the binary pair nodes
   int num_lights;
   int num_sdf;
   int root;        // packed child reference of the root, or INT_MIN if empty
   int num_prims;
   float sdf_k;
   float sdf_eps;
   int sdf_steps;
   int sdf_prim;    // slot of the SDF field primitive, or -1
};

// Per-job constants of one launch.
struct JobView {
   SceneView scene;
   float4* acc;        // per pixel: sum.rgb, sum of luminance^2
   float4* comp;       // per pixel: low parts of the float-float sums in acc (rgb, luminance^2)
   int* counts;        // per pixel sample count, negative = converged
   uint32_t* rng;      // per pixel sequential stream (RAYON_FLAG_REF_RNG only)
   unsigned long long* counters;  // kCounterSlots x 8: segments, shadow segments, primary
                                  // samples, converged, node/prim/material/light fetches
   float cam_c[3], p00[3], du[3], dv[3], cu[3], cv[3];
   int width, height;
   int max_depth;
   uint32_t flags;
   uint32_t seed_lo, seed_hi;
   float light_intensity, background_intensity, metal_fuzziness, glass_ior;
   float dof_aperture, dof_focus;
   int min_adaptive;
   float adaptive_threshold;
   int sample_base;    // index of the first sample of this pass
   int spp_target;
   const int* live;    // batched passes of adaptive jobs: the pixels still sampling, or NULL
   int nlive;          // entries of `live` (-1: every pixel, in pixel order)
   int pad_;
};

}  // namespace ry
