// api.cpp — the C ABI of include/rayon_hip.h: contexts, scene upload, jobs, the plain
// render path, resolves and parity queries.
//
// Reference counterparts: renderer_cuda_device.cu:157-478 (accumulative render,
// resolve, adaptive buffers) and scene_builder_cuda.cu:148-335 (scene upload/free).
// Unlike the reference, every call is reentrant (no function-static buffers or global
// __constant__ knobs), failures return a status instead of printf-and-continue
// (renderer_cuda_device.cu:262-266, :286-296), and device memory stays owned by the
// library's handles.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "internal.h"
#include "kernels.h"
#include "wavefront.h"

namespace ry {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(g_err, sizeof(g_err), fmt, ap);
   va_end(ap);
}

rayon_status hip_fail(hipError_t e, const char* what) {
   set_error("%s failed: %s", what, hipGetErrorString(e));
   return e == hipErrorOutOfMemory ? RAYON_ERR_OOM : RAYON_ERR_HIP;
}

static rayon_status dev_alloc(std::vector<DevBuf>& bufs, size_t bytes, void** out) {
   void* p = nullptr;
   if (bytes) {
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess) return hip_fail(e, "hipMalloc");
   }
   bufs.push_back({p, bytes});
   *out = p;
   return RAYON_OK;
}

static void free_all(std::vector<DevBuf>& bufs) {
   for (auto& b : bufs)
      if (b.p) (void)hipFree(b.p);
   bufs.clear();
}

// Job buffers come from the device's stream-ordered pool: a server creates and retires
// jobs continuously, and pool allocations neither map fresh pages each time nor make a
// free wait for the whole device.
static rayon_status job_alloc(rayon_ctx* ctx, std::vector<DevBuf>& bufs, size_t bytes, void** out) {
   void* p = nullptr;
   if (bytes) {
      hipError_t e = hipMallocAsync(&p, bytes, ctx->stream);
      if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
   }
   bufs.push_back({p, bytes});
   *out = p;
   return RAYON_OK;
}

static void job_free_all(rayon_ctx* ctx, std::vector<DevBuf>& bufs) {
   for (auto& b : bufs)
      if (b.p) (void)hipFreeAsync(b.p, ctx->stream);
   bufs.clear();
}

template <class T>
static rayon_status upload(std::vector<DevBuf>& bufs, const std::vector<T>& host, T** out) {
   void* p = nullptr;
   rayon_status s = dev_alloc(bufs, host.size() * sizeof(T), &p);
   if (s != RAYON_OK) return s;
   if (!host.empty()) RY_HIP(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
   *out = (T*)p;
   return RAYON_OK;
}

static float bits_as_float(uint32_t u) {
   float f;
   std::memcpy(&f, &u, 4);
   return f;
}

static uint32_t ry_bits(float f) {
   uint32_t u;
   std::memcpy(&u, &f, 4);
   return u;
}

}  // namespace ry

using namespace ry;

void ry::fill_jobview(rayon_job* j) {
   JobView& J = j->view;
   const rayon_job_params& P = j->params;
   J.scene = j->scene->view;
   for (int k = 0; k < 3; ++k) {
      J.cam_c[k] = P.camera_center[k];
      J.p00[k] = P.pixel00[k];
      J.du[k] = P.delta_u[k];
      J.dv[k] = P.delta_v[k];
      J.cu[k] = P.cam_u[k];
      J.cv[k] = P.cam_v[k];
   }
   J.width = P.width;
   J.height = P.height;
   J.max_depth = P.max_depth;
   J.flags = P.flags;
   J.seed_lo = (uint32_t)P.seed;
   J.seed_hi = (uint32_t)(P.seed >> 32);
   J.light_intensity = P.light_intensity;
   J.background_intensity = P.background_intensity;
   J.metal_fuzziness = P.metal_fuzziness;
   J.glass_ior = P.glass_ior;
   J.dof_aperture = P.dof_aperture;
   J.dof_focus = P.dof_focus_distance;
   J.min_adaptive = P.min_adaptive_samples;
   J.adaptive_threshold = P.adaptive_threshold;
   J.spp_target = P.spp;
   J.sample_base = (int)j->samples_done;
   J.live = nullptr;
   J.nlive = -1;
}

extern "C" {

const char* rayon_last_error(void) { return ry::g_err; }
int32_t rayon_abi_version(void) { return RAYON_ABI_VERSION; }

rayon_status rayon_camera_frame(const double lf[3], const double la[3], const double up[3], double vfov,
                                int32_t W, int32_t H, rayon_camera* out) {
   if (!lf || !la || !up || !out || W <= 0 || H <= 0) {
      set_error("rayon_camera_frame: bad arguments");
      return RAYON_ERR_INVALID;
   }
   // camera_base.hpp:76-99, FP64
   double w[3], u[3], v[3];
   double f2 = 0;
   for (int k = 0; k < 3; ++k) {
      w[k] = lf[k] - la[k];
      f2 += w[k] * w[k];
   }
   const double focal = std::sqrt(f2);
   if (!(focal > 0)) {
      set_error("rayon_camera_frame: look_from equals look_at");
      return RAYON_ERR_INVALID;
   }
   for (int k = 0; k < 3; ++k) w[k] /= focal;
   const double vh = 2.0 * std::tan(vfov * M_PI / 180.0 / 2.0) * focal;
   const double vw = vh * ((double)W / H);
   double c[3] = {up[1] * w[2] - up[2] * w[1], up[2] * w[0] - up[0] * w[2], up[0] * w[1] - up[1] * w[0]};
   double cl = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
   for (int k = 0; k < 3; ++k) u[k] = c[k] / cl;
   v[0] = w[1] * u[2] - w[2] * u[1];
   v[1] = w[2] * u[0] - w[0] * u[2];
   v[2] = w[0] * u[1] - w[1] * u[0];
   for (int k = 0; k < 3; ++k) {
      const double vu = vw * u[k], vv = -vh * v[k];
      out->center[k] = lf[k];
      out->delta_u[k] = vu / W;
      out->delta_v[k] = vv / H;
      const double ul = lf[k] - focal * w[k] - vu / 2 - vv / 2;
      out->pixel00[k] = ul + 0.5 * (out->delta_u[k] + out->delta_v[k]);
      out->u[k] = u[k];
      out->v[k] = v[k];
      out->w[k] = w[k];
   }
   return RAYON_OK;
}

void rayon_job_params_default(rayon_job_params* p) {
   std::memset(p, 0, sizeof(*p));
   p->width = 640;
   p->height = 360;
   p->max_depth = 16;  // constants.hpp:21
   p->spp = 64;        // constants.hpp:20
   p->seed = 1234;     // cuda_utils.cu:26 seed of the reference stream
   p->light_intensity = 1.0f;
   p->background_intensity = 1.0f;
   p->metal_fuzziness = 0.8f;  // renderer_cuda_device.cu:20
   p->glass_ior = 1.5f;        // renderer_cuda_device.cu:21
   p->dof_aperture = 0.1f;     // renderer_cuda_device.cu:25
   p->dof_focus_distance = 10.0f;
   p->min_adaptive_samples = 32;  // renderer_cuda_device.cuh:24
   p->adaptive_threshold = 0.01f;
}

void rayon_job_params_set_camera(rayon_job_params* p, const rayon_camera* c) {
   for (int k = 0; k < 3; ++k) {
      p->camera_center[k] = (float)c->center[k];
      p->pixel00[k] = (float)c->pixel00[k];
      p->delta_u[k] = (float)c->delta_u[k];
      p->delta_v[k] = (float)c->delta_v[k];
      p->cam_u[k] = (float)c->u[k];
      p->cam_v[k] = (float)c->v[k];
   }
}

rayon_status rayon_ctx_create(int32_t device, rayon_ctx** out) {
   if (!out) {
      set_error("rayon_ctx_create: out is NULL");
      return RAYON_ERR_INVALID;
   }
   int n = 0;
   RY_HIP(hipGetDeviceCount(&n));
   if (device < 0 || device >= n) {
      set_error("rayon_ctx_create: device %d not present (%d devices)", device, n);
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipSetDevice(device));
   // the device address of the fixed staging array, once per device, with it current
   void* scratch = nullptr;
   {
      static std::mutex mu;
      static std::map<int, void*> addr;
      std::lock_guard<std::mutex> lk(mu);
      auto it = addr.find(device);
      if (it == addr.end()) {
         hipError_t e = noalloc_scratch_address(&scratch);
         if (e != hipSuccess || !scratch) return hip_fail(e, "hipGetSymbolAddress (staging array)");
         addr[device] = scratch;
      } else {
         scratch = it->second;
      }
   }
   auto* c = new rayon_ctx();
   c->device = device;
   c->scratch = scratch;
   hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
   if (e != hipSuccess) {
      delete c;
      return hip_fail(e, "hipStreamCreateWithFlags");
   }
   // keep memory freed back to the device's pool for the next jobs
   hipMemPool_t pool = nullptr;
   if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
   }
   *out = c;
   return RAYON_OK;
}

rayon_status rayon_ctx_destroy(rayon_ctx* c) {
   if (!c) return RAYON_OK;
   (void)hipSetDevice(c->device);
   if (c->serve_pv) {
      (void)hipStreamSynchronize(c->serve_pv);
      (void)hipStreamDestroy(c->serve_pv);
      c->serve_pv = nullptr;
   }
   if (c->serve_first) {
      (void)rayon_ctx_destroy(c->serve_first);
      c->serve_first = nullptr;
   }
   (void)hipStreamSynchronize(c->stream);
   wave_destroy(c);
   (void)hipStreamDestroy(c->stream);
   delete c;
   return RAYON_OK;
}

void* rayon_ctx_stream(rayon_ctx* c) { return c ? (void*)c->stream : nullptr; }

rayon_status rayon_ctx_sync(rayon_ctx* c) {
   if (!c) {
      set_error("rayon_ctx_sync: NULL context");
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipSetDevice(c->device));
   RY_HIP(hipStreamSynchronize(c->stream));
   return RAYON_OK;
}

// Record (and vertex normals, when `nrm` is given) of one primitive, as layout.h defines
// them: the packing shared by the host and the device builds.
static PrimRec pack_rec(const HostPrim& p, float4* nrm) {
   const uint32_t info = (p.kind << 28) | p.flags | p.mat;
   PrimRec r;
   r.a = {p.a[0], p.a[1], p.a[2], bits_as_float(info)};
   const float of = bits_as_float((uint32_t)p.orig);
   if (p.kind == K_TRI) {
      r.b = {p.b[0] - p.a[0], p.b[1] - p.a[1], p.b[2] - p.a[2], of};
      r.c = {p.c[0] - p.a[0], p.c[1] - p.a[1], p.c[2] - p.a[2], 0.0f};
      if (nrm && (p.flags & F_NORMALS)) {
         for (int v = 0; v < 3; ++v) nrm[v] = {p.n[3 * v], p.n[3 * v + 1], p.n[3 * v + 2], 0.0f};
      }
   } else if (p.kind == K_SPHERE || p.kind == K_DISPLACED) {
      r.b = {p.b[0], 0.0f, 0.0f, of};
      r.c = {0, 0, 0, 0};
   } else if (p.kind == K_RECT) {
      r.b = {p.b[0], p.b[1], p.b[2], of};
      r.c = {p.c[0], p.c[1], p.c[2], 0.0f};
   } else {  // SDF field: bounds
      r.a = {p.lo[0], p.lo[1], p.lo[2], bits_as_float(info)};
      r.b = {p.hi[0], p.hi[1], p.hi[2], of};
      r.c = {0, 0, 0, 0};
   }
   return r;
}

static double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
   return std::chrono::duration<double, std::milli>(b - a).count();
}

// Collapses a device-built binary tree (pair nodes at dnodes, records at dp/dn in its leaf
// order) into the 8-wide quantized layout: the nodes are read back and collapsed by the host
// builder's surface-area dynamic program, the records are permuted into the wide tree's order
// on the device, and the pair nodes are rewritten to that order (both layouts stay valid).
// The scene keeps the binary layout when the collapse does not apply (a root leaf, or a wide
// tree deeper than the traversal stack), as the host build does.
static rayon_status collapse_device_tree(rayon_ctx* ctx, std::vector<DevBuf>& bufs, int n, int64_t num_nodes,
                                         Node2* dnodes, PrimRec*& dp, float4*& dn, WNode*& dw, BuildResult& br) {
   if (num_nodes <= 0) return RAYON_OK;
   std::vector<Node2> hn((size_t)num_nodes);
   RY_HIP(hipMemcpy(hn.data(), dnodes, sizeof(Node2) * hn.size(), hipMemcpyDeviceToHost));
   std::vector<WNode> wn;
   std::vector<int32_t> perm;
   int wdepth = 0;
   if (!collapse_pair_nodes(hn, n, wn, perm, wdepth)) return RAYON_OK;
   std::vector<DevBuf> tmp;
   int* dperm = nullptr;
   rayon_status st = upload(tmp, perm, &dperm);
   PrimRec* nr = nullptr;
   float4* nn = nullptr;
   hipError_t e = hipSuccess;
   if (st == RAYON_OK) e = hipMalloc((void**)&nr, sizeof(PrimRec) * (size_t)n);
   if (st == RAYON_OK && e == hipSuccess && dn) e = hipMalloc((void**)&nn, sizeof(float4) * 3 * (size_t)n);
   if (st == RAYON_OK && e == hipSuccess) e = launch_permute_recs(dp, dn, dperm, n, nr, nn, ctx->stream);
   if (st == RAYON_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   if (st == RAYON_OK && e == hipSuccess)
      e = hipMemcpy(dnodes, hn.data(), sizeof(Node2) * hn.size(), hipMemcpyHostToDevice);
   free_all(tmp);
   if (st != RAYON_OK || e != hipSuccess) {
      if (nr) (void)hipFree(nr);
      if (nn) (void)hipFree(nn);
      return st != RAYON_OK ? st : hip_fail(e, "device tree collapse");
   }
   // the permuted records replace the binary order's in the scene's allocations
   for (DevBuf& b : bufs) {
      if (b.p == dp) {
         (void)hipFree(b.p);
         b.p = nr;
      } else if (dn && b.p == dn) {
         (void)hipFree(b.p);
         b.p = nn;
      }
   }
   dp = nr;
   dn = nn;
   if ((st = upload(bufs, wn, &dw)) != RAYON_OK) return st;
   std::vector<int32_t> order((size_t)n);
   for (int s = 0; s < n; ++s) order[s] = br.order[perm[s]];
   br.order = std::move(order);
   br.wnodes = std::move(wn);
   br.wide_depth = wdepth;
   return RAYON_OK;
}

rayon_status rayon_scene_create(rayon_ctx* ctx, const rayon_scene_desc* d, uint32_t bvh_flags, rayon_scene** out) {
   if (!ctx || !out) {
      set_error("rayon_scene_create: NULL argument");
      return RAYON_ERR_INVALID;
   }
   std::vector<HostPrim> prims;
   rayon_status st = gather_prims(d, prims);
   if (st != RAYON_OK) return st;
   using clock = std::chrono::steady_clock;
   const size_t n = prims.size();
   bool any_normals = false;
   for (auto& p : prims) any_normals |= (p.flags & F_NORMALS) != 0;
   RY_HIP(hipSetDevice(ctx->device));
   auto* sc = new rayon_scene();
   sc->ctx = ctx;
   auto fail = [&](rayon_status s) {
      free_all(sc->bufs);
      delete sc;
      return s;
   };
   BuildResult br;
   PrimRec* dp = nullptr;
   float4* dn = nullptr;
   Node2* dnodes = nullptr;
   WNode* dw = nullptr;
   int64_t num_nodes = 0;
   double build_ms = 0.0, upload_ms = 0.0;
   int32_t builder = RAYON_BVH_BUILDER_HOST;
   if ((bvh_flags & RAYON_BVH_DEVICE) && n > 0) {
      // records and boxes go up in host order; the device sorts them into slot order
      const auto u0 = clock::now();
      std::vector<PrimRec> recs(n);
      std::vector<float4> normals, lo(n), hi(n);
      if (any_normals) normals.assign(3 * n, float4{0, 0, 0, 0});
      for (size_t i = 0; i < n; ++i) {
         const HostPrim& p = prims[i];
         recs[i] = pack_rec(p, any_normals ? &normals[3 * i] : nullptr);
         lo[i] = {p.lo[0], p.lo[1], p.lo[2], 0.0f};
         hi[i] = {p.hi[0], p.hi[1], p.hi[2], 0.0f};
      }
      std::vector<DevBuf> tmp;
      PrimRec* sr = nullptr;
      float4 *sn = nullptr, *slo = nullptr, *shi = nullptr;
      if ((st = upload(tmp, recs, &sr)) != RAYON_OK || (st = upload(tmp, normals, &sn)) != RAYON_OK ||
          (st = upload(tmp, lo, &slo)) != RAYON_OK || (st = upload(tmp, hi, &shi)) != RAYON_OK) {
         free_all(tmp);
         return fail(st);
      }
      const auto u1 = clock::now();
      DeviceTree dt;
      st = build_bvh_device(ctx, slo, shi, sr, any_normals ? sn : nullptr, (int)n, dt);
      const auto u2 = clock::now();
      free_all(tmp);
      if (st != RAYON_OK) return fail(st);
      if (dt.max_depth > kMaxDepth) {
         dt.release();  // deeper than the traversal stack: the host build below instead
      } else {
         if (dt.nodes) sc->bufs.push_back({dt.nodes, sizeof(Node2) * (size_t)dt.num_nodes});
         sc->bufs.push_back({dt.recs, sizeof(PrimRec) * n});
         if (dt.normals) sc->bufs.push_back({dt.normals, sizeof(float4) * 3 * n});
         dp = dt.recs;
         dn = dt.normals;
         dnodes = dt.nodes;
         num_nodes = dt.num_nodes;
         br.order = std::move(dt.order);
         br.root = dt.root;
         br.max_depth = dt.max_depth;
         builder = RAYON_BVH_BUILDER_DEVICE;
         upload_ms = ms_between(u0, u1);
         build_ms = ms_between(u1, u2);
         if (bvh_flags & RAYON_BVH_QUANTIZED) {
            // the 8-wide collapse of the device-built tree: read back, collapsed on the host
            // (a dynamic program over the tree), the records permuted on the device
            if ((st = collapse_device_tree(ctx, sc->bufs, (int)n, num_nodes, dnodes, dp, dn, dw, br)) != RAYON_OK)
               return fail(st);
            build_ms = ms_between(u1, clock::now());
         }
      }
   }
   if (builder == RAYON_BVH_BUILDER_HOST) {
      const auto t0 = clock::now();
      if (!build_bvh(prims, br, (bvh_flags & RAYON_BVH_QUANTIZED) != 0)) {
         set_error("BVH depth %d exceeds the traversal stack (%d)", br.max_depth, kMaxDepth);
         return fail(RAYON_ERR_INVALID);
      }
      const auto t1 = clock::now();
      std::vector<PrimRec> recs(n);
      std::vector<float4> normals;
      if (any_normals) normals.assign(3 * n, float4{0, 0, 0, 0});
      for (size_t s = 0; s < n; ++s) recs[s] = pack_rec(prims[br.order[s]], any_normals ? &normals[3 * s] : nullptr);
      if ((st = upload(sc->bufs, recs, &dp)) != RAYON_OK || (st = upload(sc->bufs, normals, &dn)) != RAYON_OK ||
          (st = upload(sc->bufs, br.wnodes, &dw)) != RAYON_OK || (st = upload(sc->bufs, br.nodes, &dnodes)) != RAYON_OK)
         return fail(st);
      num_nodes = (int64_t)br.nodes.size();
      build_ms = ms_between(t0, t1);
      upload_ms = ms_between(t1, clock::now());
   }
   const auto c0 = clock::now();
   sc->orig_of_slot.resize(n);
   sc->order = br.order;
   std::vector<int> slot_of(n);
   for (size_t s = 0; s < n; ++s) {
      slot_of[br.order[s]] = (int)s;
      sc->orig_of_slot[s] = prims[br.order[s]].orig;
   }
   std::vector<MatRec> mats(d->num_materials);
   for (int i = 0; i < d->num_materials; ++i) mats[i] = pack_material(d->materials[i]);
   std::vector<LightRec> lights;
   for (size_t i = 0; i < n; ++i) {
      const HostPrim& p = prims[i];
      if (!(p.flags & F_EMITTER)) continue;
      double area = 0;
      if (p.kind == K_RECT) {
         double cx = (double)p.b[1] * p.c[2] - (double)p.b[2] * p.c[1];
         double cy = (double)p.b[2] * p.c[0] - (double)p.b[0] * p.c[2];
         double cz = (double)p.b[0] * p.c[1] - (double)p.b[1] * p.c[0];
         area = std::sqrt(cx * cx + cy * cy + cz * cz);
      } else if (p.kind == K_TRI) {
         double e1[3], e2[3];
         for (int k = 0; k < 3; ++k) {
            e1[k] = (double)p.b[k] - p.a[k];
            e2[k] = (double)p.c[k] - p.a[k];
         }
         double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                cz = e1[0] * e2[1] - e1[1] * e2[0];
         area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
      } else if (p.kind == K_SPHERE) {
         area = 4.0 * M_PI * (double)p.b[0] * p.b[0];
      } else {
         continue;
      }
      const float4 em = mats[p.mat].m1;
      lights.push_back({pack_rec(p, nullptr), {em.x, em.y, em.z, (float)area}});
   }
   std::vector<SdfRec> sdf;
   std::vector<float4> sdf_sph;
   for (int i = 0; i < d->num_sdf_prims; ++i) {
      const rayon_sdf_prim& q = d->sdf_prims[i];
      SdfRec r{{q.center[0], q.center[1], q.center[2], -1.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};
      if (q.type == RAYON_SDF_SPHERE) {
         r.p.w = q.params[0];
      } else if (q.type == RAYON_SDF_ROUNDED_BOX) {
         const float rr = q.params[3];
         r.p = {q.params[0] - rr, q.params[1] - rr, q.params[2] - rr, rr};
      } else {
         r.c.w = q.params[0];
         r.p.w = q.params[1];
      }
      sdf.push_back(r);
      // bounding sphere: |half extents| + rounding + torus major radius, rounded up
      const double rad = std::sqrt((double)r.p.x * r.p.x + (double)r.p.y * r.p.y + (double)r.p.z * r.p.z) +
                         (double)r.p.w + std::max(0.0, (double)r.c.w);
      sdf_sph.push_back({r.c.x, r.c.y, r.c.z, std::nextafter((float)rad, 3.4e38f)});
   }
   SceneView& V = sc->view;
   MatRec* dm = nullptr;
   LightRec* dl = nullptr;
   SdfRec* ds = nullptr;
   float4* dsph = nullptr;
   if ((st = upload(sc->bufs, mats, &dm)) != RAYON_OK || (st = upload(sc->bufs, lights, &dl)) != RAYON_OK ||
        (st = upload(sc->bufs, sdf, &ds)) != RAYON_OK ||
       (st = upload(sc->bufs, sdf_sph, &dsph)) != RAYON_OK)
      return fail(st);
   upload_ms += ms_between(c0, clock::now());
   V.prims = dp;
   V.normals = dn;
   V.nodes = dnodes;
   V.mats = dm;
   V.lights = dl;
   V.sdf = ds;
   V.sdf_sph = dsph;
   V.wnodes = br.wnodes.empty() ? nullptr : dw;
   V.num_lights = (int)lights.size();
   V.num_sdf = (int)sdf.size();
   V.root = br.root;
   V.num_prims = (int)n;
   V.sdf_k = d->num_sdf_prims > 0 ? d->sdf_smooth_k : 0.1f;
   V.sdf_eps = d->num_sdf_prims > 0 ? d->sdf_eps : 1e-4f;
   V.sdf_steps = d->num_sdf_prims > 0 ? d->sdf_max_steps : 0;
   V.sdf_prim = -1;
   for (size_t i = 0; i < n; ++i)
      if (prims[i].kind == K_SDF) V.sdf_prim = slot_of[i];
   // shading features beyond the lean set: kernels without them serve passes whose scenes
   // use none of them
   uint32_t feat = 0;
   for (size_t i = 0; i < n; ++i) {
      if (prims[i].kind == K_DISPLACED) feat |= SF_DISPLACED;
      if (prims[i].kind == K_SDF) feat |= SF_SDF;
   }
   for (const MatRec& m : mats) {
      const uint32_t w = ry_bits(m.m0.w), type = w & 0xFFu, pat = (w >> 8) & 0xFFu;
      if (pat == RAYON_PATTERN_FIBONACCI_DOTS) feat |= SF_DOTS;
      if (type == RAYON_MAT_THIN_FILM || type == RAYON_MAT_CLEAR_COAT) feat |= SF_LAYERED;
      if (type == RAYON_MAT_ANISOTROPIC_METAL) feat |= SF_ANISO;
   }
   sc->features = feat;
   rayon_scene_info& I = sc->info;
   I.num_prims = (int64_t)n;
   I.num_nodes = num_nodes;
   I.num_qnodes = (int64_t)br.wnodes.size();
   I.max_depth = br.max_depth;
   sc->wide_depth = br.wnodes.empty() ? 0 : br.wide_depth;
   I.wide_depth = sc->wide_depth;
   I.num_lights = V.num_lights;
   for (auto& b : sc->bufs) I.device_bytes += (int64_t)b.bytes;
   I.build_ms = build_ms;
   I.upload_ms = upload_ms;
   I.builder = builder;
   *out = sc;
   return RAYON_OK;
}

rayon_status rayon_scene_info_get(const rayon_scene* s, rayon_scene_info* out) {
   if (!s || !out) {
      set_error("rayon_scene_info_get: NULL argument");
      return RAYON_ERR_INVALID;
   }
   *out = s->info;
   return RAYON_OK;
}

rayon_status rayon_scene_prim_order(const rayon_scene* s, int64_t* out, int64_t n) {
   if (!s || !out || n != (int64_t)s->orig_of_slot.size()) {
      set_error("rayon_scene_prim_order: NULL argument or n != number of primitives");
      return RAYON_ERR_INVALID;
   }
   std::memcpy(out, s->orig_of_slot.data(), sizeof(int64_t) * (size_t)n);
   return RAYON_OK;
}

rayon_status rayon_scene_destroy(rayon_scene* s) {
   if (!s) return RAYON_OK;
   (void)hipSetDevice(s->ctx->device);
   (void)hipStreamSynchronize(s->ctx->stream);
   free_all(s->bufs);
   delete s;
   return RAYON_OK;
}


rayon_status rayon_job_reset(rayon_job* j) {
   if (!j) {
      set_error("rayon_job_reset: NULL job");
      return RAYON_ERR_INVALID;
   }
   const size_t npix = (size_t)j->params.width * j->params.height;
   hipStream_t st = j->ctx->stream;
   RY_HIP(hipSetDevice(j->ctx->device));
   // fills by kernel, not hipMemsetAsync, so a caller may capture this in a graph
   RY_HIP(launch_fill_u32((uint32_t*)j->view.acc, 0u, (long long)npix * 4, st));
   RY_HIP(launch_fill_u32((uint32_t*)j->view.comp, 0u, (long long)npix * 4, st));
   RY_HIP(launch_fill_u32((uint32_t*)j->view.counts, 0u, (long long)npix, st));
   RY_HIP(launch_fill_u32((uint32_t*)j->view.counters, 0u, 2 * 8 * kCounterSlots, st));
   if (j->view.rng) RY_HIP(launch_ref_rng_init(j->view.rng, (int)npix, st));
   j->samples_done = 0;
   j->path = -1;
   return RAYON_OK;
}

rayon_status rayon_scene_set_batch_path(rayon_scene* scene, int32_t path) {
   if (!scene || path < RAYON_PATH_AUTO || path > RAYON_PATH_MEGAKERNEL) {
      set_error("rayon_scene_set_batch_path: NULL scene or unknown path %d", (int)path);
      return RAYON_ERR_INVALID;
   }
   scene->batch_path = path;
   return RAYON_OK;
}

int32_t rayon_scene_batch_path(const rayon_scene* scene) { return scene ? ry::scene_batch_path(scene) : -1; }

rayon_status rayon_job_create(rayon_ctx* ctx, rayon_scene* sc, const rayon_job_params* p, rayon_job** out) {
   return ry::job_create(ctx, sc, p, out, true, true);
}

}  // extern "C"

// sync = false leaves the buffers' allocation and clears queued on the context's stream:
// only for callers that order every later use after that stream. clear = false leaves the
// accumulation buffers and counters uncleared: the caller clears them (job_clear_ranges)
// before the job's first pass or statistics read.
rayon_status ry::job_create(rayon_ctx* ctx, rayon_scene* sc, const rayon_job_params* p, rayon_job** out, bool sync,
                            bool clear) {
   if (!ctx || !sc || !p || !out) {
      set_error("rayon_job_create: NULL argument");
      return RAYON_ERR_INVALID;
   }
   if (p->width <= 0 || p->height <= 0 || (int64_t)p->width * p->height > (1ll << 28) || p->max_depth < 1 ||
       p->max_depth > 1024 || p->spp < 0) {
      set_error("rayon_job_create: bad resolution (%dx%d), max_depth %d or spp %d", p->width, p->height,
                p->max_depth, p->spp);
      return RAYON_ERR_INVALID;
   }
   if ((p->flags & (RAYON_FLAG_ADAPTIVE_REF | RAYON_FLAG_ADAPTIVE_STDERR)) ==
       (RAYON_FLAG_ADAPTIVE_REF | RAYON_FLAG_ADAPTIVE_STDERR)) {
      set_error("rayon_job_create: choose one adaptive test");
      return RAYON_ERR_INVALID;
   }
   (void)hipSetDevice(ctx->device);
   auto* j = new rayon_job();
   j->ctx = ctx;
   j->scene = sc;
   j->params = *p;
   const size_t npix = (size_t)p->width * p->height;
   void *acc, *comp, *counts, *rng = nullptr, *ctr;
   rayon_status st;
   if ((st = job_alloc(ctx, j->pool_bufs, npix * sizeof(float4), &acc)) != RAYON_OK ||
       (st = job_alloc(ctx, j->pool_bufs, npix * sizeof(float4), &comp)) != RAYON_OK ||
       (st = job_alloc(ctx, j->pool_bufs, npix * sizeof(int), &counts)) != RAYON_OK ||
       (st = job_alloc(ctx, j->pool_bufs, kCounterSlots * 8 * sizeof(unsigned long long), &ctr)) != RAYON_OK ||
       ((p->flags & RAYON_FLAG_REF_RNG) && (st = job_alloc(ctx, j->pool_bufs, npix * sizeof(uint32_t), &rng)) != RAYON_OK)) {
      job_free_all(ctx, j->pool_bufs);
      delete j;
      return st;
   }
   j->view.acc = (float4*)acc;
   j->view.comp = (float4*)comp;
   j->view.counts = (int*)counts;
   j->view.rng = (uint32_t*)rng;
   j->view.counters = (unsigned long long*)ctr;
   fill_jobview(j);
   st = clear ? rayon_job_reset(j) : RAYON_OK;
   // the buffers are valid for every stream (and host copies) once the clears have run
   if (st == RAYON_OK && sync && hipStreamSynchronize(ctx->stream) != hipSuccess)
      st = hip_fail(hipGetLastError(), "job set-up");
   if (st != RAYON_OK) {
      job_free_all(ctx, j->pool_bufs);
      delete j;
      return st;
   }
   *out = j;
   return RAYON_OK;
}

// The buffers rayon_job_reset clears, as (pointer, 32-bit words) ranges; returns how many
// (4, or 5 with a reference RNG stream, which is set by its own kernel instead).
int ry::job_clear_ranges(const rayon_job* j, uint32_t** p, long long* n) {
   const long long npix = (long long)j->params.width * j->params.height;
   p[0] = (uint32_t*)j->view.acc;
   n[0] = npix * 4;
   p[1] = (uint32_t*)j->view.comp;
   n[1] = npix * 4;
   p[2] = (uint32_t*)j->view.counts;
   n[2] = npix;
   p[3] = (uint32_t*)j->view.counters;
   n[3] = 2 * 8 * kCounterSlots;
   return 4;
}

extern "C" {

rayon_status rayon_job_destroy(rayon_job* j) {
   if (!j) return RAYON_OK;
   (void)hipSetDevice(j->ctx->device);
   (void)hipStreamSynchronize(j->ctx->stream);
   {
      std::lock_guard<std::mutex> lk(j->ctx->mu);
      wave_forget_job(j->ctx, j);
   }
   free_all(j->bufs);
   job_free_all(j->ctx, j->pool_bufs);
   delete j;
   return RAYON_OK;
}

rayon_status rayon_job_render_plain(rayon_ctx* ctx, rayon_job* j, int32_t samples) {
   if (!ctx || !j || samples < 0) {
      set_error("rayon_job_render_plain: bad arguments");
      return RAYON_ERR_INVALID;
   }
   int64_t left = j->params.spp - j->samples_done;
   if (samples > left) samples = (int32_t)left;
   if (samples <= 0) return RAYON_OK;
   RY_HIP(hipSetDevice(ctx->device));
   fill_jobview(j);
   RY_HIP(launch_plain(j->view, samples, ctx->stream));
   j->samples_done += samples;
   return RAYON_OK;
}

rayon_status rayon_job_resolve_linear(rayon_ctx* ctx, rayon_job* j, float* rgb) {
   if (!ctx || !j || !rgb) {
      set_error("rayon_job_resolve_linear: NULL argument");
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipSetDevice(ctx->device));
   const size_t bytes = (size_t)j->params.width * j->params.height * 3 * sizeof(float);
   float* d = nullptr;
   RY_HIP(hipMalloc(&d, bytes));
   hipError_t e = launch_resolve_linear(j->view, d, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(rgb, d, bytes, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   (void)hipFree(d);
   if (e != hipSuccess) return hip_fail(e, "resolve_linear");
   return RAYON_OK;
}

rayon_status rayon_job_resolve_u8(rayon_ctx* ctx, rayon_job* j, uint8_t* out, int32_t channels, float gamma) {
   if (!ctx || !j || !out || (channels != 3 && channels != 4) || !(gamma > 0.0f)) {
      set_error("rayon_job_resolve_u8: bad arguments");
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipSetDevice(ctx->device));
   const size_t bytes = (size_t)j->params.width * j->params.height * channels;
   uint8_t* d = nullptr;
   RY_HIP(hipMalloc(&d, bytes));
   hipError_t e = launch_resolve_u8(j->view, d, channels, gamma, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   (void)hipFree(d);
   if (e != hipSuccess) return hip_fail(e, "resolve_u8");
   return RAYON_OK;
}

rayon_status rayon_job_resolve_u8_noalloc(rayon_ctx* ctx, rayon_job* j, uint8_t* out, int32_t channels,
                                          float gamma) {
   if (!ctx || !j || !out || (channels != 3 && channels != 4) || !(gamma > 0.0f)) {
      set_error("rayon_job_resolve_u8_noalloc: bad arguments");
      return RAYON_ERR_INVALID;
   }
   if (!j->ctx || j->ctx->device != ctx->device) {
      set_error("rayon_job_resolve_u8_noalloc: the job lives on another device than the context");
      return RAYON_ERR_INVALID;
   }
   const size_t bytes = (size_t)j->params.width * j->params.height * channels;
   if (bytes > kNoAllocBytes) {
      set_error("rayon_job_resolve_u8_noalloc: %dx%d x %d channels is %zu bytes, beyond the %zu-byte staging "
                "array; use rayon_job_resolve_u8",
                j->params.width, j->params.height, channels, bytes, kNoAllocBytes);
      return RAYON_ERR_INVALID;
   }
   if (!ctx->scratch) {
      set_error("rayon_job_resolve_u8_noalloc: no staging array on device %d", ctx->device);
      return RAYON_ERR_HIP;
   }
   RY_HIP(hipSetDevice(ctx->device));
   uint8_t* d = (uint8_t*)ctx->scratch;
   hipError_t e = launch_resolve_u8(j->view, d, channels, gamma, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   if (e != hipSuccess) return hip_fail(e, "resolve_u8_noalloc");
   return RAYON_OK;
}

rayon_status rayon_job_sample_counts(rayon_ctx* ctx, rayon_job* j, int32_t* out) {
   if (!ctx || !j || !out) {
      set_error("rayon_job_sample_counts: NULL argument");
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipSetDevice(ctx->device));
   RY_HIP(hipMemcpyAsync(out, j->view.counts, (size_t)j->params.width * j->params.height * sizeof(int),
                         hipMemcpyDeviceToHost, ctx->stream));
   RY_HIP(hipStreamSynchronize(ctx->stream));
   return RAYON_OK;
}

rayon_status rayon_job_stats_get(rayon_ctx* ctx, rayon_job* j, rayon_job_stats* out) {
   if (!ctx || !j || !out) {
      set_error("rayon_job_stats_get: NULL argument");
      return RAYON_ERR_INVALID;
   }
   unsigned long long slots[8 * kCounterSlots], c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
   RY_HIP(hipSetDevice(ctx->device));
   RY_HIP(hipMemcpyAsync(slots, j->view.counters, sizeof(slots), hipMemcpyDeviceToHost, ctx->stream));
   RY_HIP(hipStreamSynchronize(ctx->stream));
   for (int s = 0; s < kCounterSlots; ++s)
      for (int k = 0; k < 8; ++k) c[k] += slots[8 * s + k];
   out->samples_done = j->samples_done;
   out->segments = (int64_t)c[0] + j->pre_counts[0];
   out->shadow_segments = (int64_t)c[1] + j->pre_counts[1];
   out->primary_samples = (int64_t)c[2] + j->pre_counts[2];
   out->converged_pixels = (int64_t)c[3] + j->pre_counts[3];
   out->pad_ = 0;
   out->node_fetches = (int64_t)c[4];
   out->prim_fetches = (int64_t)c[5];
   out->material_fetches = (int64_t)c[6];
   out->light_fetches = (int64_t)c[7];
   const int64_t npix = (int64_t)j->params.width * j->params.height;
   out->done = (j->samples_done >= j->params.spp || out->converged_pixels >= npix) ? 1 : 0;
   return RAYON_OK;
}

rayon_status rayon_trace_closest(rayon_ctx* ctx, rayon_scene* sc, const float* o, const float* d, int64_t n,
                                 float tmin, float tmax, int64_t* prim, float* t) {
   if (!ctx || !sc || !o || !d || !prim || !t || n < 0) {
      set_error("rayon_trace_closest: bad arguments");
      return RAYON_ERR_INVALID;
   }
   if (n == 0) return RAYON_OK;
   RY_HIP(hipSetDevice(ctx->device));
   float *dO = nullptr, *dD = nullptr, *dT = nullptr;
   long long* dP = nullptr;
   const size_t b3 = (size_t)n * 3 * sizeof(float);
   hipError_t e = hipMalloc(&dO, b3);
   if (e == hipSuccess) e = hipMalloc(&dD, b3);
   if (e == hipSuccess) e = hipMalloc(&dT, (size_t)n * sizeof(float));
   if (e == hipSuccess) e = hipMalloc(&dP, (size_t)n * sizeof(long long));
   if (e == hipSuccess) e = hipMemcpyAsync(dO, o, b3, hipMemcpyHostToDevice, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(dD, d, b3, hipMemcpyHostToDevice, ctx->stream);
   if (e == hipSuccess) e = launch_trace_closest(sc->view, dO, dD, n, tmin, tmax, dP, dT, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(prim, dP, (size_t)n * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(t, dT, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   (void)hipFree(dO);
   (void)hipFree(dD);
   (void)hipFree(dT);
   (void)hipFree(dP);
   if (e != hipSuccess) return hip_fail(e, "trace_closest");
   return RAYON_OK;
}

rayon_status rayon_scatter_probe(rayon_ctx* ctx, const rayon_material* m, uint32_t flags, float fuzz,
                                 float glass_ior, float light, const float* dirs, const float* normals,
                                 const float* tangents, const int32_t* front, const float* rnd, int32_t nrnd,
                                 int64_t n, float* out_dir, float* out_atten, float* out_emitted,
                                 int32_t* out_scattered, int32_t* out_used) {
   if (!ctx || !m || !dirs || !normals || !tangents || !front || (!rnd && nrnd > 0) || nrnd < 0 || n < 0 ||
       !out_dir || !out_atten || !out_emitted || !out_scattered || !out_used) {
      set_error("rayon_scatter_probe: bad arguments");
      return RAYON_ERR_INVALID;
   }
   if (n == 0) return RAYON_OK;
   RY_HIP(hipSetDevice(ctx->device));
   JobView J{};
   J.flags = flags;
   J.metal_fuzziness = fuzz;
   J.glass_ior = glass_ior;
   J.light_intensity = light;
   const MatRec M = pack_material(*m);
   const size_t b3 = (size_t)n * 3 * sizeof(float), bi = (size_t)n * sizeof(int32_t);
   const size_t br = (size_t)n * (size_t)nrnd * sizeof(float);
   // one allocation: 3 inputs, 3 outputs of 3 floats, front, scattered, used, random numbers
   char* base = nullptr;
   hipError_t e = hipMalloc(&base, 6 * b3 + 3 * bi + br + 64);
   if (e != hipSuccess) return hip_fail(e, "scatter_probe");
   float* dD = (float*)base;
   float* dN = (float*)(base + b3);
   float* dT = (float*)(base + 2 * b3);
   float* dO = (float*)(base + 3 * b3);
   float* dA = (float*)(base + 4 * b3);
   float* dE = (float*)(base + 5 * b3);
   int* dF = (int*)(base + 6 * b3);
   int* dS = (int*)(base + 6 * b3 + bi);
   int* dU = (int*)(base + 6 * b3 + 2 * bi);
   float* dR = (float*)(base + 6 * b3 + 3 * bi);
   e = hipMemcpyAsync(dD, dirs, b3, hipMemcpyHostToDevice, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(dN, normals, b3, hipMemcpyHostToDevice, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(dT, tangents, b3, hipMemcpyHostToDevice, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(dF, front, bi, hipMemcpyHostToDevice, ctx->stream);
   if (e == hipSuccess && br) e = hipMemcpyAsync(dR, rnd, br, hipMemcpyHostToDevice, ctx->stream);
   if (e == hipSuccess)
      e = launch_scatter_probe(M, J, dD, dN, dT, dF, dR, nrnd, n, dO, dA, dE, dS, dU, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(out_dir, dO, b3, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(out_atten, dA, b3, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(out_emitted, dE, b3, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(out_scattered, dS, bi, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(out_used, dU, bi, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   (void)hipFree(base);
   if (e != hipSuccess) return hip_fail(e, "scatter_probe");
   return RAYON_OK;
}

rayon_status rayon_accumulate_probe(rayon_ctx* ctx, const float* rgb, int64_t n, float out_sum[4],
                                    float out_comp[4]) {
   if (!ctx || (!rgb && n > 0) || n < 0 || !out_sum || !out_comp) {
      set_error("rayon_accumulate_probe: bad arguments");
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipSetDevice(ctx->device));
   const size_t b = (size_t)n * 3 * sizeof(float);
   char* base = nullptr;
   hipError_t e = hipMalloc(&base, b + 2 * sizeof(float4) + 64);
   if (e != hipSuccess) return hip_fail(e, "accumulate_probe");
   float4* dOut = (float4*)base;
   float* dV = (float*)(base + 2 * sizeof(float4));
   if (n) e = hipMemcpyAsync(dV, rgb, b, hipMemcpyHostToDevice, ctx->stream);
   if (e == hipSuccess) e = launch_accumulate_probe(dV, n, dOut, ctx->stream);
   float4 r[2];
   if (e == hipSuccess) e = hipMemcpyAsync(r, dOut, sizeof(r), hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   (void)hipFree(base);
   if (e != hipSuccess) return hip_fail(e, "accumulate_probe");
   out_sum[0] = r[0].x, out_sum[1] = r[0].y, out_sum[2] = r[0].z, out_sum[3] = r[0].w;
   out_comp[0] = r[1].x, out_comp[1] = r[1].y, out_comp[2] = r[1].z, out_comp[3] = r[1].w;
   return RAYON_OK;
}

rayon_status rayon_job_primary_rays(rayon_ctx* ctx, rayon_job* j, int32_t sample, float* o, float* d) {
   if (!ctx || !j || !o || !d || sample < 0) {
      set_error("rayon_job_primary_rays: bad arguments");
      return RAYON_ERR_INVALID;
   }
   RY_HIP(hipSetDevice(ctx->device));
   const size_t b3 = (size_t)j->params.width * j->params.height * 3 * sizeof(float);
   float *dO = nullptr, *dD = nullptr;
   hipError_t e = hipMalloc(&dO, b3);
   if (e == hipSuccess) e = hipMalloc(&dD, b3);
   fill_jobview(j);
   if (e == hipSuccess) e = launch_primary(j->view, sample, dO, dD, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(o, dO, b3, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipMemcpyAsync(d, dD, b3, hipMemcpyDeviceToHost, ctx->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   (void)hipFree(dO);
   (void)hipFree(dD);
   if (e != hipSuccess) return hip_fail(e, "primary_rays");
   return RAYON_OK;
}

}  // extern "C"
