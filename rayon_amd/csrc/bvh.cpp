// bvh.cpp — host BVH builder and primitive gathering.
//
// The reference builds a full-sweep SAH over sorted centroids in FP64
// (scenes/scene_description.hpp:793-925: per-axis std::sort, prefix/suffix bounds,
// cost 1 + (A_L/A_P) n_L + (A_R/A_P) n_R, leaves of <= 4) and then reorders the
// geometry array into leaf order (:735-777). This builder keeps the leaf rule and the
// leaf-order reorder but bins centroids (16 bins per axis) so a 1.3M-triangle mesh
// builds in well under a second on a few threads, bounds every box by outward-rounded
// FP32 values (the reference's FP64->FP32 cast can round a box inward), and forces
// object-median splits when the remaining subtree could exceed kMaxDepth levels, so the
// device's fixed LDS stack can never overflow (the reference drops pushes past 32,
// cuda_raytracer.cuh:561-564).
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

#include "internal.h"

// Cost of a primitive test relative to a wide-node visit in the surface-area collapse.
#ifndef RY_WIDE_PRIM_COST
#define RY_WIDE_PRIM_COST 0.6
#endif

namespace ry {

namespace {

inline float round_down(double x) {
   float f = (float)x;
   if ((double)f > x) f = std::nextafter(f, -INFINITY);
   return f;
}
inline float round_up(double x) {
   float f = (float)x;
   if ((double)f < x) f = std::nextafter(f, INFINITY);
   return f;
}

void prim_bounds(HostPrim& p) {
   double lo[3], hi[3];
   if (p.kind == K_SPHERE || p.kind == K_DISPLACED) {
      for (int k = 0; k < 3; ++k) {
         lo[k] = (double)p.a[k] - (double)p.b[0];
         hi[k] = (double)p.a[k] + (double)p.b[0];
      }
   } else if (p.kind == K_RECT) {
      for (int k = 0; k < 3; ++k) {
         double c[4] = {p.a[k], (double)p.a[k] + p.b[k], (double)p.a[k] + p.c[k],
                        (double)p.a[k] + p.b[k] + p.c[k]};
         lo[k] = std::min(std::min(c[0], c[1]), std::min(c[2], c[3]));
         hi[k] = std::max(std::max(c[0], c[1]), std::max(c[2], c[3]));
      }
   } else if (p.kind == K_TRI) {  // a = v0, b = v1, c = v2 (edges are formed at upload)
      for (int k = 0; k < 3; ++k) {
         lo[k] = std::min(std::min(p.a[k], p.b[k]), p.c[k]);
         hi[k] = std::max(std::max(p.a[k], p.b[k]), p.c[k]);
      }
   } else {  // SDF field: bounds precomputed into lo/hi by the gatherer
      return;
   }
   for (int k = 0; k < 3; ++k) {
      p.lo[k] = round_down(lo[k]);
      p.hi[k] = round_up(hi[k]);
   }
}

struct Box {
   float lo[3] = {INFINITY, INFINITY, INFINITY};
   float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
   void grow(const float* l, const float* h) {
      for (int k = 0; k < 3; ++k) {
         lo[k] = std::min(lo[k], l[k]);
         hi[k] = std::max(hi[k], h[k]);
      }
   }
   void grow(const Box& b) { grow(b.lo, b.hi); }
   double area() const {
      double e[3];
      for (int k = 0; k < 3; ++k) e[k] = std::max(0.0, (double)hi[k] - lo[k]);
      return 2.0 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
   }
};

struct BNode {
   Box box;
   int left = -1, right = -1;
   int first = 0, count = 0;
};

struct Builder {
   const std::vector<HostPrim>& prims;
   std::vector<float> cen;       // 3 per prim
   std::vector<int32_t> idx;
   std::vector<BNode> nodes;
   std::atomic<int> next_node{0};
   std::atomic<int> threads{1};
   std::atomic<int> max_depth{0};

   explicit Builder(const std::vector<HostPrim>& p) : prims(p) {}

   static int ceil_log2(long long n) {
      int k = 0;
      while ((1ll << k) < n) ++k;
      return k;
   }

   int alloc() { return next_node.fetch_add(1); }

   void build(int ni, int first, int count, int depth) {
      BNode& node = nodes[ni];
      Box cb;
      for (int i = first; i < first + count; ++i) {
         const HostPrim& p = prims[idx[i]];
         node.box.grow(p.lo, p.hi);
         const float* c = &cen[3 * (size_t)idx[i]];
         cb.grow(c, c);
      }
      if (count <= 4) {
         node.first = first;
         node.count = count;
         int md = max_depth.load();
         while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {
         }
         return;
      }
      int mid = -1;
      const bool force_median = depth + 1 + ceil_log2((count + 3) / 4) >= kMaxDepth;
      float ext[3];
      for (int k = 0; k < 3; ++k) ext[k] = cb.hi[k] - cb.lo[k];
      int wide = ext[0] >= ext[1] ? (ext[0] >= ext[2] ? 0 : 2) : (ext[1] >= ext[2] ? 1 : 2);
      if (!force_median && ext[wide] > 0.0f) {
         constexpr int B = 16;
         double best = INFINITY;
         int best_axis = -1, best_split = -1;
         for (int ax = 0; ax < 3; ++ax) {
            if (!(ext[ax] > 0.0f)) continue;
            Box bins[B];
            int cnt[B] = {0};
            const double scale = (double)B / ext[ax];
            for (int i = first; i < first + count; ++i) {
               const HostPrim& p = prims[idx[i]];
               int b = (int)(((double)cen[3 * (size_t)idx[i] + ax] - cb.lo[ax]) * scale);
               b = std::min(std::max(b, 0), B - 1);
               ++cnt[b];
               bins[b].grow(p.lo, p.hi);
            }
            double rarea[B];
            int rcnt[B];
            Box acc;
            int c = 0;
            for (int b = B - 1; b > 0; --b) {
               acc.grow(bins[b]);
               c += cnt[b];
               rarea[b] = acc.area();
               rcnt[b] = c;
            }
            Box lacc;
            int lc = 0;
            for (int b = 0; b < B - 1; ++b) {
               lacc.grow(bins[b]);
               lc += cnt[b];
               if (lc == 0 || rcnt[b + 1] == 0) continue;
               double cost = lacc.area() * lc + rarea[b + 1] * rcnt[b + 1];
               if (cost < best) {
                  best = cost;
                  best_axis = ax;
                  best_split = b + 1;
               }
            }
         }
         if (best_axis >= 0) {
            const double scale = (double)B / ext[best_axis];
            const float lo = cb.lo[best_axis];
            auto it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](int32_t pi) {
               int b = (int)(((double)cen[3 * (size_t)pi + best_axis] - lo) * scale);
               b = std::min(std::max(b, 0), B - 1);
               return b < best_split;
            });
            mid = (int)(it - idx.begin());
            if (mid == first || mid == first + count) mid = -1;
         }
      }
      if (mid < 0) {  // object median on the widest centroid axis
         mid = first + count / 2;
         std::nth_element(idx.begin() + first, idx.begin() + mid, idx.begin() + first + count,
                          [&](int32_t i, int32_t j) {
                             float ci = cen[3 * (size_t)i + wide], cj = cen[3 * (size_t)j + wide];
                             return ci < cj || (ci == cj && i < j);
                          });
      }
      const int l = alloc(), r = alloc();
      node.left = l;
      node.right = r;
      const int nl = mid - first, nr = first + count - mid;
      if (count > 65536 && threads.load() < 16) {
         threads.fetch_add(1);
         std::thread t([=] { build(l, first, nl, depth + 1); });
         build(r, mid, nr, depth + 1);
         t.join();
         threads.fetch_sub(1);
      } else {
         build(l, first, nl, depth + 1);
         build(r, mid, nr, depth + 1);
      }
   }
};

void fill_child(Node2& pn, int which, const BNode& c, int ref) {
   float* bx = &pn.bx.x;
   float* by = &pn.by.x;
   float* bz = &pn.bz.x;
   bx[2 * which] = c.box.lo[0];
   bx[2 * which + 1] = c.box.hi[0];
   by[2 * which] = c.box.lo[1];
   by[2 * which + 1] = c.box.hi[1];
   bz[2 * which] = c.box.lo[2];
   bz[2 * which + 1] = c.box.hi[2];
   int* meta = &pn.meta.x;
   meta[which] = c.count ? c.first : ref;
   meta[2 + which] = c.count;
}

int emit(const std::vector<BNode>& bn, int b, std::vector<Node2>& out) {
   const int pi = (int)out.size();
   out.emplace_back();
   std::memset(&out[pi], 0, sizeof(Node2));
   const BNode& node = bn[b];
   const int kids[2] = {node.left, node.right};
   for (int w = 0; w < 2; ++w) {
      const BNode& c = bn[kids[w]];
      int ref = 0;
      if (!c.count) ref = emit(bn, kids[w], out);
      fill_child(out[pi], w, c, ref);
   }
   return pi;
}


// ---- 8-wide quantized tree (compressed wide BVH of Ylitie, Karras and Laine, HPG 2017,
// collapsed greedily from the binary tree: a wide node's children are found by
// repeatedly opening its largest-area internal child until there are eight).
struct WideBuilder {
   std::vector<BNode>& bn;
   const std::vector<int32_t>& old_order;
   std::vector<WNode> out;
   std::vector<int32_t> order;
   int max_depth = 0;

   WideBuilder(std::vector<BNode>& b, const std::vector<int32_t>& o) : bn(b), old_order(o) {}

   // Which binary nodes become wide nodes: the surface-area cost of the wide tree is
   // minimised by dynamic programming over the binary tree (the collapse of Ylitie, Karras
   // and Laine, HPG 2017, with the binary leaves kept as they are). cost[b][k-1] is the
   // least cost of covering b's subtree with at most k child slots of one wide node;
   // cut[b][k-1] is 0 when b takes one slot itself (a leaf, or an internal child that
   // becomes a wide node), else the number of slots its left child's subtree gets.
   // A wide node costs kNodeCost per unit of surface area (testing its eight children), a
   // primitive kPrimCost.
   static constexpr double kNodeCost = 1.0;
   static constexpr double kPrimCost = RY_WIDE_PRIM_COST;
   std::vector<std::array<float, 8>> cost;
   std::vector<std::array<int8_t, 8>> cut;

   void plan(int b) {
      BNode& n = bn[b];
      auto& c = cost[b];
      auto& k1 = cut[b];
      if (n.count) {
         c.fill((float)(n.box.area() * n.count * kPrimCost));
         k1.fill(0);
         return;
      }
      plan(n.left);
      plan(n.right);
      const auto& cl = cost[n.left];
      const auto& cr = cost[n.right];
      float split[9];  // split[k]: both children within k slots (k >= 2)
      int8_t arg[9];
      for (int k = 2; k <= 8; ++k) {
         split[k] = INFINITY;
         arg[k] = 1;
         for (int a = 1; a < k; ++a) {
            const float v = cl[a - 1] + cr[k - a - 1];
            if (v < split[k]) {
               split[k] = v;
               arg[k] = (int8_t)a;
            }
         }
      }
      const float own = (float)(n.box.area() * kNodeCost) + split[8];  // b as a wide node
      c[0] = own;
      k1[0] = 0;
      wide_cut[b] = arg[8];
      for (int k = 2; k <= 8; ++k) {
         if (own <= split[k]) {
            c[k - 1] = own;
            k1[k - 1] = 0;
         } else {
            c[k - 1] = split[k];
            k1[k - 1] = arg[k];
         }
      }
   }
   std::vector<int8_t> wide_cut;  // slots of the left child's subtree when b is a wide node

   // The binary nodes that fill at most k slots for b's subtree, as planned.
   void cover(int b, int k, int* kids, int& nk) {
      const int a = cut[b][k - 1];
      if (a == 0) {
         kids[nk++] = b;
         return;
      }
      cover(bn[b].left, a, kids, nk);
      cover(bn[b].right, k - a, kids, nk);
   }

   static int exponent_for(double ext) {
      // smallest e with 255 * 2^e >= ext, within the normal float range
      int e = -126;
      if (ext > 0) {
         e = (int)std::ceil(std::log2(ext / 255.0));
         while (std::ldexp(255.0, e) < ext) ++e;
         while (e > -126 && std::ldexp(255.0, e - 1) >= ext) --e;
      }
      return std::min(std::max(e, -126), 127);
   }

   void emit(int wi, int b, int depth) {
      max_depth = std::max(max_depth, depth);
      int kids[8];
      int nk = 0;
      if (bn[b].count) {
         kids[nk++] = b;
      } else {
         const int a = wide_cut[b];
         cover(bn[b].left, a, kids, nk);
         cover(bn[b].right, 8 - a, kids, nk);
      }
      Box nb;
      for (int k = 0; k < nk; ++k) nb.grow(bn[kids[k]].box);
      // octant slot assignment: slot bit k set <=> centroid on the negative side of axis k
      double pc[3], d[8][3];
      for (int a = 0; a < 3; ++a) pc[a] = 0.5 * ((double)nb.lo[a] + nb.hi[a]);
      for (int k = 0; k < nk; ++k)
         for (int a = 0; a < 3; ++a) d[k][a] = 0.5 * ((double)bn[kids[k]].box.lo[a] + bn[kids[k]].box.hi[a]) - pc[a];
      struct Cand {
         double score;
         int kid, slot;
      };
      Cand cand[64];
      int nc = 0;
      for (int k = 0; k < nk; ++k)
         for (int sl = 0; sl < 8; ++sl) {
            double sc = 0;
            for (int a = 0; a < 3; ++a) sc += ((sl >> a) & 1) ? -d[k][a] : d[k][a];
            cand[nc++] = {sc, k, sl};
         }
      std::sort(cand, cand + nc, [](const Cand& x, const Cand& y) {
         return x.score > y.score || (x.score == y.score && (x.kid < y.kid || (x.kid == y.kid && x.slot < y.slot)));
      });
      int slot_kid[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
      bool placed[8] = {false, false, false, false, false, false, false, false};
      for (int c = 0; c < nc; ++c)
         if (!placed[cand[c].kid] && slot_kid[cand[c].slot] < 0) {
            placed[cand[c].kid] = true;
            slot_kid[cand[c].slot] = kids[cand[c].kid];
         }
      WNode w;
      std::memset(&w, 0, sizeof(w));
      w.px = nb.lo[0];
      w.py = nb.lo[1];
      w.pz = nb.lo[2];
      int e[3];
      double scale[3];
      for (int a = 0; a < 3; ++a) {
         e[a] = exponent_for((double)nb.hi[a] - (double)nb.lo[a]);
         scale[a] = std::ldexp(1.0, e[a]);
      }
      uint8_t q[6][8];
      uint8_t meta[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      uint32_t imask = 0;
      int n_int = 0;
      for (int sl = 0; sl < 8; ++sl) {
         const int c = slot_kid[sl];
         if (c < 0) {  // empty slot: an inverted box, and bit 7 of its meta byte set
            for (int a = 0; a < 3; ++a) {
               q[a][sl] = 255;
               q[3 + a][sl] = 0;
            }
            meta[sl] = 0x80;
            continue;
         }
         const BNode& cn = bn[c];
         for (int a = 0; a < 3; ++a) {
            const double p = (a == 0) ? w.px : (a == 1 ? w.py : w.pz);
            double lo = std::floor(((double)cn.box.lo[a] - p) / scale[a]);
            double hi = std::ceil(((double)cn.box.hi[a] - p) / scale[a]);
            q[a][sl] = (uint8_t)std::min(std::max(lo, 0.0), 255.0);
            q[3 + a][sl] = (uint8_t)std::min(std::max(hi, 0.0), 255.0);
         }
         if (!cn.count) {
            imask |= 1u << sl;
            ++n_int;
         }
      }
      w.e_imask = (uint32_t)(e[0] + 127) | ((uint32_t)(e[1] + 127) << 8) | ((uint32_t)(e[2] + 127) << 16) |
                  (imask << 24);
      w.child_base = (uint32_t)out.size();
      out.resize(out.size() + n_int);
      w.prim_base = (uint32_t)order.size();
      for (int sl = 0; sl < 8; ++sl) {
         const int c = slot_kid[sl];
         if (c < 0 || !bn[c].count) continue;
         BNode& cn = bn[c];
         const int off = (int)(order.size() - w.prim_base);
         for (int i = cn.first; i < cn.first + cn.count; ++i) order.push_back(old_order[i]);
         meta[sl] = (uint8_t)(off | ((cn.count - 1) << 5));
         cn.first = (int)(w.prim_base + off);  // the binary tree follows the new order
      }
      std::memcpy(w.meta, meta, 8);
      std::memcpy(w.q, q, 48);
      out[wi] = w;
      int k = 0;
      for (int sl = 0; sl < 8; ++sl) {
         const int c = slot_kid[sl];
         if (c < 0 || bn[c].count) continue;
         emit((int)w.child_base + k, c, depth + 1);
         ++k;
      }
   }
};

}  // namespace

void compute_bounds(HostPrim& p) { prim_bounds(p); }

bool collapse_pair_nodes(std::vector<Node2>& nodes, int nslots, std::vector<WNode>& wnodes, std::vector<int32_t>& perm,
                         int& wide_depth) {
   const int np = (int)nodes.size();
   if (np == 0) return false;  // a root leaf: nothing to collapse
   // binary nodes: pair node p is node p; the leaf children follow, one node each; a node's
   // box is the one its parent's pair node holds for it (the root's: both children's union)
   std::vector<BNode> bn((size_t)np);
   for (int p = 0; p < np; ++p) {
      const Node2& n = nodes[p];
      const int ref[2] = {n.meta.x, n.meta.y}, cnt[2] = {n.meta.z, n.meta.w};
      const float lox[2] = {n.bx.x, n.bx.z}, hix[2] = {n.bx.y, n.bx.w};
      const float loy[2] = {n.by.x, n.by.z}, hiy[2] = {n.by.y, n.by.w};
      const float loz[2] = {n.bz.x, n.bz.z}, hiz[2] = {n.bz.y, n.bz.w};
      int kid[2];
      for (int c = 0; c < 2; ++c) {
         if (cnt[c]) {
            BNode leaf;
            leaf.first = ref[c];
            leaf.count = cnt[c];
            kid[c] = (int)bn.size();
            bn.push_back(leaf);
         } else {
            if (ref[c] <= 0 || ref[c] >= np) return false;  // an internal child is some other pair node
            kid[c] = ref[c];
         }
         const float lo[3] = {lox[c], loy[c], loz[c]}, hi[3] = {hix[c], hiy[c], hiz[c]};
         bn[kid[c]].box = Box{};
         bn[kid[c]].box.grow(lo, hi);
      }
      bn[p].left = kid[0];
      bn[p].right = kid[1];
   }
   bn[0].box = Box{};
   bn[0].box.grow(bn[bn[0].left].box);
   bn[0].box.grow(bn[bn[0].right].box);
   std::vector<int32_t> ident((size_t)nslots);
   for (int i = 0; i < nslots; ++i) ident[i] = i;
   WideBuilder wb(bn, ident);
   wb.cost.resize(bn.size());
   wb.cut.resize(bn.size());
   wb.wide_cut.assign(bn.size(), 1);
   wb.plan(0);
   wb.out.resize(1);
   wb.order.reserve(nslots);
   wb.emit(0, 0, 0);
   if (wb.max_depth >= kMaxWideDepth || (int)wb.order.size() != nslots) return false;
   // the binary leaves follow the new primitive order too
   for (int p = 0; p < np; ++p) {
      Node2& n = nodes[p];
      if (n.meta.z) n.meta.x = bn[bn[p].left].first;
      if (n.meta.w) n.meta.y = bn[bn[p].right].first;
   }
   wnodes = std::move(wb.out);
   perm = std::move(wb.order);
   wide_depth = wb.max_depth;
   return true;
}

bool build_bvh(const std::vector<HostPrim>& prims, BuildResult& out, bool wide) {
   const int n = (int)prims.size();
   out.order.clear();
   out.nodes.clear();
   out.wnodes.clear();
   out.wide_depth = 0;
   if (n == 0) {
      out.root = kSentinel;
      out.max_depth = 0;
      return true;
   }
   Builder b(prims);
   b.cen.resize(3 * (size_t)n);
   b.idx.resize(n);
   for (int i = 0; i < n; ++i) {
      b.idx[i] = i;
      for (int k = 0; k < 3; ++k) b.cen[3 * (size_t)i + k] = 0.5f * (prims[i].lo[k] + prims[i].hi[k]);
   }
   b.nodes.resize(2 * (size_t)n);
   const int root = b.alloc();
   b.build(root, 0, n, 0);
   out.max_depth = b.max_depth.load();
   if (out.max_depth > kMaxDepth) return false;
   out.order = b.idx;
   if (wide) {
      // collapse into the 8-wide tree; this also reorders the primitives (each binary
      // leaf stays one contiguous block) and rewrites the binary leaves to match
      WideBuilder wb(b.nodes, b.idx);
      wb.cost.resize(b.nodes.size());
      wb.cut.resize(b.nodes.size());
      wb.wide_cut.assign(b.nodes.size(), 1);
      wb.plan(root);
      wb.out.resize(1);
      wb.order.reserve(n);
      wb.emit(0, root, 0);
      if (wb.max_depth < kMaxWideDepth && (int)wb.order.size() == n) {
         out.wnodes = std::move(wb.out);
         out.order = std::move(wb.order);
         out.wide_depth = wb.max_depth;
      } else {
         // too deep for the device stack: keep the binary layout (restore leaf order)
         return build_bvh(prims, out, false);
      }
   }
   const BNode& r = b.nodes[root];
   if (r.count) {
      out.root = pack_leaf(r.first, r.count);
   } else {
      out.nodes.reserve(n);
      emit(b.nodes, root, out.nodes);
      out.root = 0;
   }
   return true;
}

MatRec pack_material(const rayon_material& m) {
   MatRec r;
   uint32_t tp = (uint32_t)(m.type & 0xFF) | ((uint32_t)(m.pattern & 0xFF) << 8);
   float tpf;
   std::memcpy(&tpf, &tp, 4);
   r.m0 = {m.albedo[0], m.albedo[1], m.albedo[2], tpf};
   r.m1 = {m.emission[0], m.emission[1], m.emission[2], m.roughness};
   r.m2 = {m.eta[0], m.eta[1], m.eta[2], m.refractive_index};
   r.m3 = {m.k[0], m.k[1], m.k[2], m.anisotropy};
   r.m4 = {m.film_thickness, m.film_ior, m.pattern_param1, m.pattern_param2};
   r.m5 = {m.pattern_color[0], m.pattern_color[1], m.pattern_color[2], 0.0f};
   return r;
}

rayon_status gather_prims(const rayon_scene_desc* d, std::vector<HostPrim>& prims) {
   if (!d) {
      set_error("scene description is NULL");
      return RAYON_ERR_INVALID;
   }
   if (d->num_materials <= 0 || !d->materials) {
      set_error("scene needs at least one material");
      return RAYON_ERR_INVALID;
   }
   int64_t total = d->num_geometries + (d->num_sdf_prims > 0 ? 1 : 0);
   for (int i = 0; i < d->num_meshes; ++i) total += d->meshes[i].num_triangles;
   if (total >= (1ll << 27)) {
      set_error("scene has %lld primitives; at most 2^27-1 are supported", (long long)total);
      return RAYON_ERR_INVALID;
   }
   auto mat_ok = [&](int m) { return m >= 0 && m < d->num_materials; };
   // F_EMITTER marks the emitters next-event estimation samples: LIGHT-material triangles,
   // spheres and rectangles. Other emitters (golf-ball and SDF lights, CONSTANT surfaces)
   // are only reached by BSDF sampling, so they occlude shadow rays and always add
   // their emission.
   auto flags_for = [&](int mat, bool visible, bool normals, uint32_t kind) {
      uint32_t f = 0;
      if (visible) f |= F_VISIBLE;
      if (normals) f |= F_NORMALS;
      if (d->materials[mat].type == RAYON_MAT_LIGHT && (kind == K_TRI || kind == K_SPHERE || kind == K_RECT))
         f |= F_EMITTER;
      return f;
   };
   prims.clear();
   prims.reserve((size_t)total);
   int32_t orig = 0;
   for (int64_t i = 0; i < d->num_geometries; ++i) {
      const rayon_geometry& g = d->geometries[i];
      if (!mat_ok(g.material_id)) {
         set_error("geometry %lld: material id %d out of range", (long long)i, g.material_id);
         return RAYON_ERR_INVALID;
      }
      HostPrim p{};
      switch (g.type) {
      case RAYON_GEOM_SPHERE:
         p.kind = K_SPHERE;
         break;
      case RAYON_GEOM_DISPLACED_SPHERE:
         p.kind = K_DISPLACED;
         break;
      case RAYON_GEOM_RECTANGLE:
         p.kind = K_RECT;
         break;
      case RAYON_GEOM_TRIANGLE:
         p.kind = K_TRI;
         break;
      default:
         set_error("geometry %lld: unsupported type %d", (long long)i, g.type);
         return RAYON_ERR_UNSUPPORTED;
      }
      if ((p.kind == K_SPHERE || p.kind == K_DISPLACED) && !(g.p[3] > 0.0f)) {
         set_error("geometry %lld: sphere radius must be positive", (long long)i);
         return RAYON_ERR_INVALID;
      }
      p.mat = (uint32_t)g.material_id;
      p.flags = flags_for(g.material_id, g.visible != 0, p.kind == K_TRI && g.has_normals, p.kind);
      p.orig = orig++;
      std::memcpy(p.a, g.p, 12);
      std::memcpy(p.b, g.p + 3, 12);
      std::memcpy(p.c, g.p + 6, 12);
      std::memcpy(p.n, g.n, 36);
      prim_bounds(p);
      prims.push_back(p);
   }
   for (int mi = 0; mi < d->num_meshes; ++mi) {
      const rayon_mesh& m = d->meshes[mi];
      if (!mat_ok(m.material_id) || !m.positions || !m.indices) {
         set_error("mesh %d: bad material id or NULL arrays", mi);
         return RAYON_ERR_INVALID;
      }
      const uint32_t fl = flags_for(m.material_id, m.visible != 0, m.normals != nullptr, K_TRI);
      for (int64_t t = 0; t < m.num_triangles; ++t) {
         HostPrim p{};
         p.kind = K_TRI;
         p.mat = (uint32_t)m.material_id;
         p.flags = fl;
         p.orig = orig++;
         const int32_t* ix = m.indices + 3 * t;
         for (int v = 0; v < 3; ++v) {
            if (ix[v] < 0 || ix[v] >= m.num_vertices) {
               set_error("mesh %d triangle %lld: vertex index %d out of range", mi, (long long)t, ix[v]);
               return RAYON_ERR_INVALID;
            }
         }
         std::memcpy(p.a, m.positions + 3 * (size_t)ix[0], 12);
         std::memcpy(p.b, m.positions + 3 * (size_t)ix[1], 12);
         std::memcpy(p.c, m.positions + 3 * (size_t)ix[2], 12);
         if (m.normals) {
            for (int v = 0; v < 3; ++v) std::memcpy(p.n + 3 * v, m.normals + 3 * (size_t)ix[v], 12);
         }
         prim_bounds(p);
         prims.push_back(p);
      }
   }
   if (d->num_sdf_prims > 0) {
      if (!mat_ok(d->sdf_material_id) || !d->sdf_prims || !(d->sdf_smooth_k > 0.0f) || d->sdf_max_steps <= 0 ||
          !(d->sdf_eps > 0.0f)) {
         set_error("SDF field: bad material id, smooth k, step cap or epsilon");
         return RAYON_ERR_INVALID;
      }
      HostPrim p{};
      p.kind = K_SDF;
      p.mat = (uint32_t)d->sdf_material_id;
      p.flags = flags_for(d->sdf_material_id, true, false, K_SDF);
      p.orig = orig++;
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int i = 0; i < d->num_sdf_prims; ++i) {
         const rayon_sdf_prim& q = d->sdf_prims[i];
         const float* a = q.params;
         const bool ok = q.type == RAYON_SDF_SPHERE        ? a[0] > 0.0f
                         : q.type == RAYON_SDF_ROUNDED_BOX ? (a[0] > 0.0f && a[1] > 0.0f && a[2] > 0.0f &&
                                                              a[3] >= 0.0f && a[3] <= std::min({a[0], a[1], a[2]}))
                         : q.type == RAYON_SDF_TORUS       ? (a[0] >= 0.0f && a[1] > 0.0f)
                                                           : true;
         if (!ok) {
            set_error("SDF primitive %d: radius, half extents or rounding out of range", i);
            return RAYON_ERR_INVALID;
         }
         double e[3];
         if (q.type == RAYON_SDF_SPHERE) {
            e[0] = e[1] = e[2] = q.params[0];
         } else if (q.type == RAYON_SDF_ROUNDED_BOX) {
            e[0] = q.params[0];
            e[1] = q.params[1];
            e[2] = q.params[2];
         } else if (q.type == RAYON_SDF_TORUS) {
            e[0] = e[2] = (double)q.params[0] + q.params[1];
            e[1] = q.params[1];
         } else {
            set_error("SDF primitive %d: unsupported type %d", i, q.type);
            return RAYON_ERR_UNSUPPORTED;
         }
         for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], (double)q.center[k] - e[k]);
            hi[k] = std::max(hi[k], (double)q.center[k] + e[k]);
         }
      }
      // Each smooth-union step lowers the running value by at most k/4 below the smaller of
      // its two inputs, so the field of n primitives is at least (distance to the nearest
      // primitive) - (n - 1) k / 4: its surface lies within that much of the primitives' boxes.
      const double pad = 0.25 * d->sdf_smooth_k * std::max(1, d->num_sdf_prims - 1) + 2.0 * d->sdf_eps;
      for (int k = 0; k < 3; ++k) {
         p.lo[k] = round_down(lo[k] - pad);
         p.hi[k] = round_up(hi[k] + pad);
         p.a[k] = p.lo[k];
         p.b[k] = p.hi[k];
      }
      prims.push_back(p);
   }
   return RAYON_OK;
}

namespace {

bool box_inside(const float* lo, const float* hi, float bx0, float bx1, float by0, float by1, float bz0,
                float bz1) {
   return lo[0] >= bx0 && hi[0] <= bx1 && lo[1] >= by0 && hi[1] <= by1 && lo[2] >= bz0 && hi[2] <= bz1;
}

// Walks the pair-node tree: every child box must contain the boxes below it; leaves hold
// 1..8 slots; each slot is visited once.
bool check_subtree(const BuildResult& br, const std::vector<HostPrim>& prims, int ref, int count, const float* box,
                   std::vector<int>& seen, int depth, int& max_depth) {
   if (depth > max_depth) max_depth = depth;
   if (count > 0) {
      if (count > 8) return false;
      for (int s = ref; s < ref + count; ++s) {
         if (s < 0 || s >= (int)br.order.size()) return false;
         if (seen[s]++) return false;
         const HostPrim& p = prims[br.order[s]];
         if (box && !box_inside(p.lo, p.hi, box[0], box[1], box[2], box[3], box[4], box[5])) return false;
      }
      return true;
   }
   if (ref < 0 || ref >= (int)br.nodes.size()) return false;
   const Node2& n = br.nodes[ref];
   const float* bx = &n.bx.x;
   const float* by = &n.by.x;
   const float* bz = &n.bz.x;
   const int* meta = &n.meta.x;
   for (int w = 0; w < 2; ++w) {
      const float cb[6] = {bx[2 * w], bx[2 * w + 1], by[2 * w], by[2 * w + 1], bz[2 * w], bz[2 * w + 1]};
      if (box) {
         const float lo[3] = {cb[0], cb[2], cb[4]}, hi[3] = {cb[1], cb[3], cb[5]};
         if (!box_inside(lo, hi, box[0], box[1], box[2], box[3], box[4], box[5])) return false;
      }
      if (!check_subtree(br, prims, meta[w], meta[2 + w], cb, seen, depth + 1, max_depth)) return false;
   }
   return true;
}

// Walks the 8-wide tree: every decoded (quantized) child box must contain the exact
// bounds of everything below it; returns those exact bounds through lo/hi.
bool check_wide(const BuildResult& br, const std::vector<HostPrim>& prims, uint32_t ni, std::vector<int>& seen,
                int depth, int& max_depth, double lo[3], double hi[3]) {
   if (ni >= br.wnodes.size() || depth >= kMaxWideDepth) return false;
   max_depth = std::max(max_depth, depth);
   const WNode& w = br.wnodes[ni];
   const uint8_t* q = reinterpret_cast<const uint8_t*>(w.q);
   const uint8_t* meta = reinterpret_cast<const uint8_t*>(w.meta);
   const double p[3] = {w.px, w.py, w.pz};
   double sc[3];
   for (int a = 0; a < 3; ++a) sc[a] = std::ldexp(1.0, (int)((w.e_imask >> (8 * a)) & 0xFFu) - 127);
   const uint32_t imask = w.e_imask >> 24;
   for (int a = 0; a < 3; ++a) {
      lo[a] = INFINITY;
      hi[a] = -INFINITY;
   }
   int k = 0;
   for (int sl = 0; sl < 8; ++sl) {
      const bool empty = q[sl] > q[24 + sl];
      if (empty != ((meta[sl] & 0x80u) != 0)) return false;  // the two marks of an empty slot agree
      if (empty) continue;
      double clo[3], chi[3];
      for (int a = 0; a < 3; ++a) {
         clo[a] = p[a] + q[8 * a + sl] * sc[a];
         chi[a] = p[a] + q[24 + 8 * a + sl] * sc[a];
      }
      double elo[3] = {INFINITY, INFINITY, INFINITY}, ehi[3] = {-INFINITY, -INFINITY, -INFINITY};
      if ((imask >> sl) & 1u) {
         if (!check_wide(br, prims, w.child_base + k, seen, depth + 1, max_depth, elo, ehi)) return false;
         ++k;
      } else {
         const int first = (int)w.prim_base + (meta[sl] & 31);
         const int cnt = (meta[sl] >> 5) + 1;
         for (int s = first; s < first + cnt; ++s) {
            if (s < 0 || s >= (int)br.order.size() || seen[s]++) return false;
            const HostPrim& hp = prims[br.order[s]];
            for (int a = 0; a < 3; ++a) {
               elo[a] = std::min(elo[a], (double)hp.lo[a]);
               ehi[a] = std::max(ehi[a], (double)hp.hi[a]);
            }
         }
      }
      for (int a = 0; a < 3; ++a) {
         if (elo[a] < clo[a] || ehi[a] > chi[a]) return false;
         lo[a] = std::min(lo[a], elo[a]);
         hi[a] = std::max(hi[a], ehi[a]);
      }
   }
   return true;
}

// Binary tree from br.root: containment, leaf sizes, every slot exactly once; the depth of
// the deepest leaf through `depth`.
bool check_binary(const BuildResult& br, const std::vector<HostPrim>& prims, int& depth) {
   std::vector<int> seen(prims.size(), 0);
   depth = 0;
   bool ok;
   if (prims.empty()) {
      ok = true;
   } else if (br.root < 0) {
      const int packed = ~br.root;
      ok = check_subtree(br, prims, packed >> 3, (packed & 7) + 1, nullptr, seen, 0, depth);
   } else {
      ok = check_subtree(br, prims, 0, 0, nullptr, seen, 0, depth);
   }
   for (int s : seen) ok = ok && s == 1;
   return ok;
}

}  // namespace

}  // namespace ry

extern "C" rayon_status rayon_scene_check(const rayon_scene* sc, const rayon_scene_desc* d, int32_t* depth) {
   using namespace ry;
   if (!sc || !d) {
      set_error("rayon_scene_check: NULL argument");
      return RAYON_ERR_INVALID;
   }
   std::vector<HostPrim> prims;
   rayon_status st = gather_prims(d, prims);
   if (st != RAYON_OK) return st;
   if (prims.size() != sc->order.size()) {
      set_error("rayon_scene_check: the description has %zu primitives, the scene %zu", prims.size(),
                sc->order.size());
      return RAYON_ERR_INVALID;
   }
   BuildResult br;
   br.order = sc->order;
   br.root = sc->view.root;
   br.nodes.resize((size_t)sc->info.num_nodes);
   RY_HIP(hipSetDevice(sc->ctx->device));
   if (!br.nodes.empty())
      RY_HIP(hipMemcpy(br.nodes.data(), sc->view.nodes, br.nodes.size() * sizeof(Node2), hipMemcpyDeviceToHost));
   int dep = 0;
   if (!check_binary(br, prims, dep)) {
      set_error("BVH invariant violated (containment, leaf size or slot coverage)");
      return RAYON_ERR_INVALID;
   }
   if (sc->view.wnodes) {  // the 8-wide tree over the same primitive order
      br.wnodes.resize((size_t)sc->info.num_qnodes);
      RY_HIP(hipMemcpy(br.wnodes.data(), sc->view.wnodes, br.wnodes.size() * sizeof(WNode), hipMemcpyDeviceToHost));
      std::vector<int> wseen(prims.size(), 0);
      int wdepth = 0;
      double lo[3], hi[3];
      bool wok = check_wide(br, prims, 0, wseen, 0, wdepth, lo, hi) && wdepth == sc->wide_depth;
      for (int v : wseen) wok = wok && v == 1;
      if (!wok) {
         set_error("wide BVH invariant violated (quantized containment, leaf size, slot coverage or depth)");
         return RAYON_ERR_INVALID;
      }
   }
   if (depth) *depth = dep;
   return RAYON_OK;
}

extern "C" rayon_status rayon_bvh_check(const rayon_scene_desc* d, rayon_scene_info* out) {
   using namespace ry;
   std::vector<HostPrim> prims;
   rayon_status st = gather_prims(d, prims);
   if (st != RAYON_OK) return st;
   BuildResult br;
   if (!build_bvh(prims, br, true)) {
      set_error("BVH depth %d exceeds %d", br.max_depth, kMaxDepth);
      return RAYON_ERR_INVALID;
   }
   if (!br.wnodes.empty()) {
      std::vector<int> wseen(prims.size(), 0);
      int wdepth = 0;
      double lo[3], hi[3];
      bool wok = check_wide(br, prims, 0, wseen, 0, wdepth, lo, hi);
      for (int v : wseen) wok = wok && v == 1;
      if (!wok) {
         set_error("wide BVH invariant violated (quantized containment, leaf size or slot coverage)");
         return RAYON_ERR_INVALID;
      }
   }
   int depth = 0;
   if (!check_binary(br, prims, depth)) {
      set_error("BVH invariant violated (containment, leaf size or slot coverage)");
      return RAYON_ERR_INVALID;
   }
   if (out) {
      *out = rayon_scene_info{};
      out->num_prims = (int64_t)prims.size();
      out->num_nodes = (int64_t)br.nodes.size();
      out->num_qnodes = (int64_t)br.wnodes.size();
      out->max_depth = br.max_depth;
   }
   return RAYON_OK;
}
