// Batched fusion of many maps (per-keyframe submaps): the per-frame fusion
// of MapStruct::Fuse (RGBDFusion/src/VoxelMap.cu:825-884) for every map of a
// batch in one launch per stage, captured once in a hipGraph.
//
// A plan holds a schedule of steps; at step s, map i fuses its depth frame
// (s, i) at pose (s, i), or idles.  The descriptors of every step live in
// device memory and the kernels read the current step from a device counter
// that the graph's last node advances, so replaying the same graph walks the
// whole schedule with no host work per step.  Maps of uneven size and
// resolution share each launch: allocation maps workgroups to (map, tile) by
// a per-step prefix table; visibility and integration give every map the
// same number of workgroups (a grid dimension) and stride over its own
// block lists, whose lengths only the device knows.
//
// The launch shapes (allocation tile, workgroups per map, integration
// workgroup size) come from a small autotuner that times the captured graph
// on the plan's first step for each candidate.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "yas_fusion_dev.hpp"
#include "yas_internal.hpp"

using namespace yas;

MapView yas_map_view(yas_map* map);

namespace {

struct BatchItem {
  MapView m;
  Intr K;
  const uint16_t* depth;
  size_t pitch;
  int W, H;
  float depth_scale;
  int active;
};

__global__ void k_bset_step(int* cur, int s) {
  if (threadIdx.x == 0) *cur = s;
}

__global__ void k_bnext_step(int* cur) {
  if (threadIdx.x == 0) *cur = *cur + 1;
}

// Pose pair of every active map of the current step; clears its visible
// counter.
__global__ void k_bprep(const BatchItem* items, const double* poses, PosePair* pp, const int* cur, int n) {
  const int s = *cur;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const YAS_CONSTANT BatchItem& it = ((const YAS_CONSTANT BatchItem*)items)[(size_t)s * n + i];
  if (!it.active) return;
  PosePair p;
  pose_pair_from(poses + ((size_t)s * n + i) * 16, p);
  pp[i] = p;
  it.m.ctr[kCtrVisible] = 0;
}

// Allocation: workgroup b of the step belongs to the map i with
// start[i] <= b < start[i+1]; tiles of TW x (256 / TW) pixels.
template <int TW>
__global__ void __launch_bounds__(256) k_balloc(const BatchItem* items, const PosePair* pp, const int* wg_start,
                                                const int* cur, int n) {
  const int s = *cur;
  const int* ws = wg_start + (size_t)s * (n + 1);
  const int b = blockIdx.x;
  if (b >= ws[n]) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ws[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const YAS_CONSTANT BatchItem& it = ((const YAS_CONSTANT BatchItem*)items)[(size_t)s * n + lo];
  constexpr int TH = 256 / TW;
  const int t = b - ws[lo];
  const int tiles_x = (it.W + TW - 1) / TW;
  __shared__ int walked;
  if (threadIdx.x == 0) walked = 0;
  __syncthreads();
  alloc_walk(it.m, DepthU16{it.depth, it.pitch, it.depth_scale}, it.W, it.H, it.K, pp[lo].T,
             (t % tiles_x) * TW + (int)(threadIdx.x % TW), (t / tiles_x) * TH + (int)(threadIdx.x / TW), &walked);
  __syncthreads();
  if (threadIdx.x == 0 && walked) atomicAdd(&it.m.ctr[kCtrWalk], walked);
}

__global__ void __launch_bounds__(256) k_bvisible(const BatchItem* items, const PosePair* pp, const int* cur, int n) {
  const int s = *cur;
  const int i = blockIdx.y;
  const YAS_CONSTANT BatchItem& it = ((const YAS_CONSTANT BatchItem*)items)[(size_t)s * n + i];
  if (!it.active) return;
  visible_scan(it.m, pp[i].Tinv, it.K, it.W, it.H, blockIdx.x * blockDim.x, gridDim.x * blockDim.x);
}

template <class VT>
__global__ void __launch_bounds__(256) k_bintegrate(const BatchItem* items, const PosePair* pp, const int* cur, int n) {
  const int s = *cur;
  const int i = blockIdx.y;
  const YAS_CONSTANT BatchItem& it = ((const YAS_CONSTANT BatchItem*)items)[(size_t)s * n + i];
  if (!it.active) return;
  integrate_waves<VT>(it.m, pp[i].Tinv, it.K, DepthU16{it.depth, it.pitch, it.depth_scale}, it.W, it.H,
                      blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), gridDim.x * (blockDim.x >> 6));
}

typedef std::tuple<int, int, long> TuneKey;  // maps, voxel format, pixels of step 0
std::mutex g_tune_mutex;
std::map<TuneKey, yas_batch_config> g_tuned;

}  // namespace

struct yas_batch {
  yas_ctx* ctx;
  hipStream_t stream;  // own stream: graphs cannot be captured on the null stream
  hipEvent_t ev_in, ev_out, t0, t1;
  int n;
  std::vector<yas_map*> maps;
  int format;
  int steps;
  std::vector<BatchItem> host_items;  // [steps][n]
  BatchItem* items;
  double* poses;
  PosePair* pp;
  int* wg_start;
  int* cur;
  yas_batch_config cfg;
  int alloc_grid;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

static void drop_graph(yas_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Per-step allocation workgroup prefixes for the configured tile shape.
static int upload_wg_start(yas_batch* b) {
  const int TW = b->cfg.alloc_tile_w, TH = 256 / TW;
  std::vector<int> ws((size_t)b->steps * (b->n + 1));
  b->alloc_grid = 1;
  for (int s = 0; s < b->steps; ++s) {
    int acc = 0;
    for (int i = 0; i < b->n; ++i) {
      ws[(size_t)s * (b->n + 1) + i] = acc;
      const BatchItem& it = b->host_items[(size_t)s * b->n + i];
      if (it.active) acc += ((it.W + TW - 1) / TW) * ((it.H + TH - 1) / TH);
    }
    ws[(size_t)s * (b->n + 1) + b->n] = acc;
    b->alloc_grid = std::max(b->alloc_grid, acc);
  }
  YAS_HIP(hipMemcpy(b->wg_start, ws.data(), sizeof(int) * ws.size(), hipMemcpyHostToDevice));
  return YAS_OK;
}

static int launch_step(yas_batch* b, hipStream_t s) {
  const int n = b->n;
  hipLaunchKernelGGL(k_bprep, dim3((n + 63) / 64), dim3(64), 0, s, (const BatchItem*)b->items, (const double*)b->poses,
                     b->pp, (const int*)b->cur, n);
  switch (b->cfg.alloc_tile_w) {
    case 64:
      hipLaunchKernelGGL(k_balloc<64>, dim3(b->alloc_grid), dim3(256), 0, s, (const BatchItem*)b->items,
                         (const PosePair*)b->pp, (const int*)b->wg_start, (const int*)b->cur, n);
      break;
    case 32:
      hipLaunchKernelGGL(k_balloc<32>, dim3(b->alloc_grid), dim3(256), 0, s, (const BatchItem*)b->items,
                         (const PosePair*)b->pp, (const int*)b->wg_start, (const int*)b->cur, n);
      break;
    default:
      hipLaunchKernelGGL(k_balloc<16>, dim3(b->alloc_grid), dim3(256), 0, s, (const BatchItem*)b->items,
                         (const PosePair*)b->pp, (const int*)b->wg_start, (const int*)b->cur, n);
  }
  hipLaunchKernelGGL(k_bvisible, dim3(b->cfg.visible_groups, n), dim3(256), 0, s, (const BatchItem*)b->items,
                     (const PosePair*)b->pp, (const int*)b->cur, n);
  const dim3 ig(b->cfg.integrate_groups, n), it(b->cfg.integrate_threads);
  switch (b->format) {
    case YAS_VOXEL_I16_U8:
      hipLaunchKernelGGL(k_bintegrate<VoxI16>, ig, it, 0, s, (const BatchItem*)b->items, (const PosePair*)b->pp,
                         (const int*)b->cur, n);
      break;
    case YAS_VOXEL_F16_U16:
      hipLaunchKernelGGL(k_bintegrate<VoxF16>, ig, it, 0, s, (const BatchItem*)b->items, (const PosePair*)b->pp,
                         (const int*)b->cur, n);
      break;
    default:
      hipLaunchKernelGGL(k_bintegrate<VoxF32>, ig, it, 0, s, (const BatchItem*)b->items, (const PosePair*)b->pp,
                         (const int*)b->cur, n);
  }
  hipLaunchKernelGGL(k_bnext_step, dim3(1), dim3(64), 0, s, b->cur);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

static int capture(yas_batch* b) {
  drop_graph(b);
  int rc = upload_wg_start(b);
  if (rc) return rc;
  YAS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  rc = launch_step(b, b->stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  YAS_HIP(e);
  b->graph = g;
  YAS_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return YAS_OK;
}

static bool config_valid(const yas_batch_config& c) {
  return (c.alloc_tile_w == 16 || c.alloc_tile_w == 32 || c.alloc_tile_w == 64) && c.visible_groups >= 1 &&
         c.visible_groups <= 65535 && (c.integrate_threads == 64 || c.integrate_threads == 128 ||
                                       c.integrate_threads == 256) &&
         c.integrate_groups >= 1 && c.integrate_groups <= 65535;
}

// Replay the graph `count` times from step `first` on the batch stream,
// ordered after the context stream's earlier work.
static int replay(yas_batch* b, int first, int count) {
  YAS_HIP(hipEventRecord(b->ev_in, b->ctx->stream));
  YAS_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  hipLaunchKernelGGL(k_bset_step, dim3(1), dim3(64), 0, b->stream, b->cur, first);
  YAS_LAUNCH_CHECK();
  for (int k = 0; k < count; ++k) YAS_HIP(hipGraphLaunch(b->exec, b->stream));
  YAS_HIP(hipEventRecord(b->ev_out, b->stream));
  YAS_HIP(hipStreamWaitEvent(b->ctx->stream, b->ev_out, 0));
  return YAS_OK;
}

extern "C" {

int yas_batch_create(yas_ctx* ctx, int num_maps, yas_map* const* maps, yas_batch** out) {
  YAS_REQUIRE(ctx && maps && out && num_maps > 0, "yas_batch_create: bad argument");
  int fmt = -1;
  for (int i = 0; i < num_maps; ++i) {
    YAS_REQUIRE(maps[i], "yas_batch_create: null map");
    YAS_REQUIRE_ACTIVE(maps[i]);
    if (fmt < 0) fmt = maps[i]->P.voxel_format;
    YAS_REQUIRE(maps[i]->P.voxel_format == fmt, "yas_batch_create: maps must share one voxel format");
    for (int j = 0; j < i; ++j) YAS_REQUIRE(maps[j] != maps[i], "yas_batch_create: a map appears twice");
  }
  YAS_HIP(hipSetDevice(ctx->device));
  yas_batch* b = new yas_batch();
  b->ctx = yas_ctx_retain(ctx);
  b->n = num_maps;
  b->maps.assign(maps, maps + num_maps);
  b->format = fmt;
  b->cfg = yas_batch_config{16, 64, 256, 512, 0.0f};
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->t0);
  if (e == hipSuccess) e = hipEventCreate(&b->t1);
  if (e == hipSuccess) e = hipMalloc((void**)&b->pp, sizeof(PosePair) * num_maps);
  if (e == hipSuccess) e = hipMalloc((void**)&b->cur, sizeof(int));
  if (e != hipSuccess) {
    yas_set_error("yas_batch_create: %s", hipGetErrorString(e));
    yas_batch_destroy(b);
    return YAS_E_HIP;
  }
  *out = b;
  return YAS_OK;
}

int yas_batch_destroy(yas_batch* b) {
  if (!b) return YAS_OK;
  hipSetDevice(b->ctx->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_graph(b);
  hipFree(b->items);
  hipFree(b->poses);
  hipFree(b->pp);
  hipFree(b->wg_start);
  hipFree(b->cur);
  if (b->ev_in) hipEventDestroy(b->ev_in);
  if (b->ev_out) hipEventDestroy(b->ev_out);
  if (b->t0) hipEventDestroy(b->t0);
  if (b->t1) hipEventDestroy(b->t1);
  if (b->stream) hipStreamDestroy(b->stream);
  yas_ctx_release(b->ctx);
  delete b;
  return YAS_OK;
}

int yas_batch_plan(yas_batch* b, int steps, const uint16_t* const* depth, const size_t* pitch, const int* W,
                   const int* H, float depth_scale, const double* T) {
  YAS_REQUIRE(b && steps > 0 && depth && pitch && W && H && T && depth_scale > 0.f, "yas_batch_plan: bad argument");
  YAS_HIP(hipSetDevice(b->ctx->device));
  YAS_HIP(hipStreamSynchronize(b->stream));
  const int n = b->n;
  std::vector<BatchItem> items((size_t)steps * n);
  for (int i = 0; i < n; ++i) YAS_REQUIRE_ACTIVE(b->maps[i]);
  for (int s = 0; s < steps; ++s)
    for (int i = 0; i < n; ++i) {
      const size_t k = (size_t)s * n + i;
      BatchItem& it = items[k];
      std::memset(&it, 0, sizeof(it));
      if (!depth[k]) continue;
      YAS_REQUIRE(W[k] > 0 && H[k] > 0 && pitch[k] >= (size_t)W[k] * 2, "yas_batch_plan: bad image size");
      it.m = yas_map_view(b->maps[i]);
      it.K = b->maps[i]->K;
      it.depth = depth[k];
      it.pitch = pitch[k];
      it.W = W[k];
      it.H = H[k];
      it.depth_scale = (float)(1.0 / (double)depth_scale);  // as the single-map u16 path
      it.active = 1;
    }
  drop_graph(b);
  hipFree(b->items);
  hipFree(b->poses);
  hipFree(b->wg_start);
  b->items = nullptr;
  b->poses = nullptr;
  b->wg_start = nullptr;
  YAS_HIP(hipMalloc((void**)&b->items, sizeof(BatchItem) * items.size()));
  YAS_HIP(hipMalloc((void**)&b->poses, sizeof(double) * 16 * items.size()));
  YAS_HIP(hipMalloc((void**)&b->wg_start, sizeof(int) * (size_t)steps * (n + 1)));
  YAS_HIP(hipMemcpy(b->items, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice));
  YAS_HIP(hipMemcpy(b->poses, T, sizeof(double) * 16 * items.size(), hipMemcpyHostToDevice));
  b->host_items.swap(items);
  b->steps = steps;
  // a configuration tuned earlier for this shape of work
  long px = 0;
  for (int i = 0; i < n; ++i)
    if (b->host_items[i].active) px += (long)b->host_items[i].W * b->host_items[i].H;
  {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto f = g_tuned.find(TuneKey(n, b->format, px));
    if (f != g_tuned.end()) b->cfg = f->second;
  }
  return capture(b);
}

int yas_batch_set_config(yas_batch* b, const yas_batch_config* c) {
  YAS_REQUIRE(b && c && config_valid(*c), "yas_batch_set_config: bad configuration");
  b->cfg = *c;
  if (b->steps) return capture(b);
  return YAS_OK;
}

int yas_batch_get_config(yas_batch* b, yas_batch_config* c) {
  YAS_REQUIRE(b && c, "null argument");
  *c = b->cfg;
  return YAS_OK;
}

int yas_batch_run(yas_batch* b, int first_step, int count) {
  YAS_REQUIRE(b && b->exec, "yas_batch_run: no plan");
  YAS_REQUIRE(first_step >= 0 && count >= 0 && first_step + count <= b->steps, "yas_batch_run: steps out of range");
  for (int i = 0; i < b->n; ++i) {
    YAS_REQUIRE_ACTIVE(b->maps[i]);
    // a map reallocated since the plan (reserve, merge, re-activation) holds new arrays
    MapView v = yas_map_view(b->maps[i]);
    for (int s = first_step; s < first_step + count; ++s) {
      const BatchItem& it = b->host_items[(size_t)s * b->n + i];
      YAS_REQUIRE(!it.active || (it.m.slots == v.slots && it.m.voxels == v.voxels),
                  "yas_batch_run: a map was reallocated after yas_batch_plan; plan again");
    }
  }
  YAS_HIP(hipSetDevice(b->ctx->device));
  return replay(b, first_step, count);
}

int yas_batch_autotune(yas_batch* b, int reset_maps, yas_batch_config* best_out) {
  YAS_REQUIRE(b && b->exec, "yas_batch_autotune: no plan");
  YAS_HIP(hipSetDevice(b->ctx->device));
  const int tiles[] = {16, 32, 64};
  const int vgroups[] = {16, 64, 256};
  const int ithreads[] = {64, 256};
  const int igroups[] = {128, 512, 2048};
  yas_batch_config best = b->cfg;
  auto time_cfg = [&](const yas_batch_config& c, float* ms) -> int {
    b->cfg = c;
    int rc = capture(b);
    if (rc) return rc;
    float tot = 0.0f;
    for (int rep = 0; rep < 4; ++rep) {
      hipLaunchKernelGGL(k_bset_step, dim3(1), dim3(64), 0, b->stream, b->cur, 0);
      YAS_HIP(hipEventRecord(b->t0, b->stream));
      YAS_HIP(hipGraphLaunch(b->exec, b->stream));
      YAS_HIP(hipEventRecord(b->t1, b->stream));
      YAS_HIP(hipEventSynchronize(b->t1));
      float t = 0.0f;
      YAS_HIP(hipEventElapsedTime(&t, b->t0, b->t1));
      if (rep > 0) tot += t;  // the first run allocates; time the steady state
    }
    *ms = tot / 3.0f;
    return YAS_OK;
  };
  YAS_HIP(hipEventRecord(b->ev_in, b->ctx->stream));
  YAS_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  float best_ms = 0.0f;
  int rc = time_cfg(best, &best_ms);
  if (rc) return rc;
  // coordinate descent over the three stages' shapes
  for (int stage = 0; stage < 3; ++stage) {
    std::vector<yas_batch_config> cand;
    if (stage == 0)
      for (int t : tiles) cand.push_back(yas_batch_config{t, best.visible_groups, best.integrate_threads, best.integrate_groups, 0.f});
    if (stage == 1)
      for (int v : vgroups) cand.push_back(yas_batch_config{best.alloc_tile_w, v, best.integrate_threads, best.integrate_groups, 0.f});
    if (stage == 2)
      for (int th : ithreads)
        for (int g : igroups) cand.push_back(yas_batch_config{best.alloc_tile_w, best.visible_groups, th, g, 0.f});
    for (const auto& c : cand) {
      float ms = 0.0f;
      if ((rc = time_cfg(c, &ms))) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  }
  best.step_ms = best_ms;
  b->cfg = best;
  if ((rc = capture(b))) return rc;
  {
    long px = 0;
    for (int i = 0; i < b->n; ++i)
      if (b->host_items[i].active) px += (long)b->host_items[i].W * b->host_items[i].H;
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    g_tuned[TuneKey(b->n, b->format, px)] = best;
  }
  YAS_HIP(hipStreamSynchronize(b->stream));
  if (reset_maps)
    for (int i = 0; i < b->n; ++i)
      if ((rc = yas_map_reset(b->maps[i]))) return rc;
  if (best_out) *best_out = best;
  return YAS_OK;
}

}  // extern "C"
