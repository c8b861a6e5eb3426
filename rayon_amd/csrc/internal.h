// internal.h — host-side objects behind the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rayon_hip.h"
#include "layout.h"

namespace ry {

void set_error(const char* fmt, ...);
rayon_status hip_fail(hipError_t e, const char* what);

#define RY_HIP(call)                                               \
   do {                                                            \
      hipError_t e_ = (call);                                      \
      if (e_ != hipSuccess) return ::ry::hip_fail(e_, #call);      \
   } while (0)

// Device allocation owned by a scene, job or context.
struct DevBuf {
   void* p = nullptr;
   size_t bytes = 0;
};

// Host primitive before the BVH reorder.
struct HostPrim {
   uint32_t kind;
   uint32_t flags;
   uint32_t mat;
   int32_t orig;
   float a[3], b[3], c[3];
   float n[9];
   float lo[3], hi[3];
};

struct BuildResult {
   std::vector<int32_t> order;  // slot -> host prim index
   std::vector<Node2> nodes;
   std::vector<WNode> wnodes;   // 8-wide quantized tree (root = 0), when requested
   int root = 0;                // packed root reference
   int max_depth = 0;
   int wide_depth = 0;
};

// Binned-SAH build (16 bins, leaves <= 4 primitives, depth <= kMaxDepth).
constexpr int kMaxDepth = 32;
bool build_bvh(const std::vector<HostPrim>& prims, BuildResult& out, bool wide = false);
constexpr int kMaxWideDepth = 16;  // device stack entries of the wide traversal
// The same 8-wide collapse of a binary tree given as pair nodes (root = node 0: a device-built
// tree read back). perm receives, for each primitive slot of the
// wide tree's order, the slot it had in the binary tree; the pair nodes' leaf references are
// rewritten to the new order. False when the root is a leaf or the wide tree would be too deep.
bool collapse_pair_nodes(std::vector<Node2>& nodes, int nslots, std::vector<WNode>& wnodes, std::vector<int32_t>& perm,
                         int& wide_depth);

// Linear BVH built on the device (bvh_device.hip) over primitive boxes and records in host
// order. The outputs are plain device allocations that the scene takes over.
struct DeviceTree {
   Node2* nodes = nullptr;      // pair nodes (root = 0), or none when the root is a leaf
   PrimRec* recs = nullptr;     // records in slot (leaf) order
   float4* normals = nullptr;   // vertex normals in slot order (when given)
   std::vector<int32_t> order;  // slot -> host prim index
   int root = 0;                // packed root reference
   int max_depth = 0;
   int64_t num_nodes = 0;
   void release() {
      if (nodes) (void)hipFree(nodes);
      if (recs) (void)hipFree(recs);
      if (normals) (void)hipFree(normals);
      nodes = nullptr;
      recs = nullptr;
      normals = nullptr;
   }
};
rayon_status build_bvh_device(rayon_ctx* ctx, const float4* d_lo, const float4* d_hi, const PrimRec* d_recs,
                              const float4* d_normals, int n, DeviceTree& out);

// Flattens a description into host primitives, validates it.
rayon_status gather_prims(const rayon_scene_desc* d, std::vector<HostPrim>& prims);
MatRec pack_material(const rayon_material& m);

struct Wave;  // wavefront.hip

}  // namespace ry

struct rayon_ctx {
   int device = 0;
   hipStream_t stream = nullptr;
   // first-frame streams of the servers of this context, created by the first server and kept
   // until the context is destroyed, so a process holds the same few streams whatever number
   // of servers it runs one after another: a context whose stream takes the job and frame
   // allocations and every other first frame, and a second first-frame stream
   rayon_ctx* serve_first = nullptr;
   hipStream_t serve_pv = nullptr;
   // device address of the fixed staging array on this context's device (looked up once
   // per device at context creation)
   void* scratch = nullptr;
   std::mutex mu;
   // wavefront state (shared by all batched launches of this context)
   ry::Wave* wave = nullptr;
   // autotuner results by workload signature
   std::mutex tune_mu;
   std::map<std::vector<long long>, rayon_tune_result> tune_cache;
};

struct rayon_scene {
   rayon_ctx* ctx = nullptr;
   ry::SceneView view{};
   std::vector<ry::DevBuf> bufs;
   rayon_scene_info info{};
   int wide_depth = 0;  // levels below the root of the 8-wide quantized tree (0 without one)
   std::vector<int64_t> orig_of_slot;
   std::vector<int32_t> order;  // slot -> primitive of the gathered description
   int32_t batch_path = RAYON_PATH_AUTO;  // path of this scene's jobs in batched passes
   uint32_t features = ry::SF_ALL;        // SceneFeature bits its primitives and materials use
};

struct rayon_job {
   rayon_ctx* ctx = nullptr;
   rayon_scene* scene = nullptr;
   rayon_job_params params{};
   ry::JobView view{};
   std::vector<ry::DevBuf> bufs;       // hipMalloc
   std::vector<ry::DevBuf> pool_bufs;  // stream-ordered pool (accumulation, counts, counters)
   int64_t samples_done = 0;
   // statistics of samples rendered outside the job's device counters (a server's first
   // frame that the job keeps): segments, shadow segments, primary samples, converged pixels
   int64_t pre_counts[4] = {0, 0, 0, 0};
   int32_t path = -1;       // batched path taken since the first pass (-1: not yet chosen)
   float4* prev = nullptr;  // sums at pass start (batched passes with the reference adaptive test)
   int* live = nullptr;      // pixels still sampling (batched passes of adaptive jobs)
   unsigned* live_n = nullptr;
};

namespace ry {
void fill_jobview(rayon_job* j);
int scene_batch_path(const rayon_scene* s);
rayon_status job_create(rayon_ctx* ctx, rayon_scene* sc, const rayon_job_params* p, rayon_job** out, bool sync,
                        bool clear);
int job_clear_ranges(const rayon_job* j, uint32_t** p, long long* n);
}
