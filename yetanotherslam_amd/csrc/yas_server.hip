// Serving engine: many RGB-D streams (sensors) tracked and fused
// concurrently, each with its own volume, tracker and pose resident in HBM.
// Each frame is the dense half of Tracking::TrackRGBD
// (FullSystem/src/Tracking.cpp:139-197) for its stream, as in yas_pipeline.
//
// Continuous batching: submitted frames wait in per-stream queues; each pump
// takes the oldest waiting frame of every stream (one per stream, since a
// stream's frames depend on each other through its pose and reference
// images) and launches them as one batch.  A batch is one replay of a
// captured hipGraph of about 75 kernels whose grids have one z-slice per batch
// slot; every kernel reads its slot's stream from a small table, and every
// stream keeps the workgroup counts of the single-stream pipeline, so a
// stream's results are bit-identical alone and in any mix.  Batches queue
// back to back on one HIP stream (in-order execution carries each stream's
// dependencies), so frames that arrive while a batch runs are admitted to
// the next one.  Streams are admitted while the HBM budget holds their
// buffers, and release them on removal.
//
// Per stream the reference and current image pyramids keep fixed buffers:
// after tracking, the current frame's intensity and inverse-depth pyramids
// are copied into the reference ones (the single pipeline swaps pointers;
// the values are the same).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "yas_fusion_dev.hpp"
#include "yas_internal.hpp"
#include "yas_raycast_dev.hpp"
#include "yas_server_dev.hpp"
#include "yas_tracker_dev.hpp"

using namespace yas;

int yas_pipeline_fill(yas_pipeline* p, StreamDev& d);
yas_map* yas_pipeline_map_of(yas_pipeline* p);

namespace {

constexpr int kMaxSlots = 64;     // frames per batch
constexpr int kInFlight = 3;      // batches queued on the device at once
constexpr int kResultStride = 20; // doubles per result record: pose[16], tracked, pad

struct Slot {
  int stream;
  int first;  // the stream's first frame: initialise, fuse at the start pose
  const uint16_t* depth_in;
  size_t dpitch;
  const uint8_t* rgb_in;
  size_t rpitch;
  double* out;  // result record: pose[16], tracked
};
struct Table {
  int count;
  int pad;
  Slot slot[kMaxSlots];
};

#define SLOT_OR_RETURN SLOT_FROM_OR_RETURN(0)
// Slot slot0 + blockIdx.z (the Gauss-Newton kernels can run on a sub-range
// of the batch's slots).  The table and the stream descriptors are written
// by the host only, so the kernels read them as constant memory (scalar
// loads); the device pointers loaded from there are then known to be
// global, which gives global rather than flat loads and stores.
#define SLOT_FROM_OR_RETURN(slot0)                                                        \
  const YAS_CONSTANT Table& ctab = *(const YAS_CONSTANT Table*)tab;                       \
  if ((int)blockIdx.z + (slot0) >= ctab.count) return;                                    \
  const YAS_CONSTANT Slot& sl = ctab.slot[blockIdx.z + (slot0)];                          \
  const YAS_CONSTANT StreamDev& S = ((const YAS_CONSTANT StreamDev*)streams)[sl.stream]

// ------------------------------------------------------------------ kernels
__global__ void __launch_bounds__(256) k_s_convert(const StreamDev* streams, const Table* tab) {
  SLOT_OR_RETURN;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= S.W || y >= S.H) return;
  const uint16_t d = ((const uint16_t*)((const char*)sl.depth_in + (size_t)y * sl.dpitch))[x];
  S.depth[(size_t)y * S.W + x] = (float)d * S.depth_scale_inv;
  rgb_to_gray_at(sl.rgb_in, sl.rpitch, S.gray, (size_t)S.W * 4, S.W, S.H, S.gray_mode, blockIdx.x, blockIdx.y);
}

// Pose pair of T_last (ray cast at the last pose, or fusion at the new one);
// clears the visible counter.
__global__ void k_s_pose(const StreamDev* streams, const Table* tab, int skip_first) {
  SLOT_OR_RETURN;
  if (threadIdx.x != 0 || (skip_first && sl.first)) return;
  PosePair p;
  pose_pair_from(S.T_last, p);
  *S.pose = p;
  S.m.ctr[kCtrVisible] = 0;
}

__global__ void __launch_bounds__(256) k_s_visible(const StreamDev* streams, const Table* tab, int fusion) {
  SLOT_OR_RETURN;
  if (!fusion && sl.first) return;
  if (fusion && !sl.first && *S.gate == 0) return;
  if ((int)blockIdx.x >= S.vis_groups) return;
  visible_scan(S.m, S.pose->Tinv, S.K, S.W, S.H, blockIdx.x * blockDim.x, S.vis_groups * blockDim.x);
}

// The batch's ray-march dispatch order: the z-range cells give each 16x16
// tile of every image a class by the largest depth range of its 2x2 cells,
// and the march takes the classes in turn, longest ranges first, so that
// the longest rays do not start last.  The classes are lists of
// (slot << 16 | tile); counts[c] is the length of class c.
constexpr int kMarchClasses = 4;
struct MarchOrder {
  int* tiles;   // [kMarchClasses][cap]
  int* counts;  // [kMarchClasses]
  int cap;      // kMaxSlots * tiles of the largest image
};

__global__ void __launch_bounds__(256) k_s_zrange_init(const StreamDev* streams, const Table* tab, MarchOrder mo) {
  if (blockIdx.x == 0 && blockIdx.z == 0 && threadIdx.x < kMarchClasses) mo.counts[threadIdx.x] = 0;
  SLOT_OR_RETURN;
  if (sl.first) return;
  const int n = S.cols * S.rows, g = (n + 255) / 256;
  if ((int)blockIdx.x >= g) return;
  zrange_init_at(S.zmin, S.zmax, n, blockIdx.x, g);
}

__global__ void __launch_bounds__(256) k_s_zrange(const StreamDev* streams, const Table* tab) {
  SLOT_OR_RETURN;
  if (sl.first || (int)blockIdx.x >= S.zr_groups) return;
  zrange_at<true>(S.m, S.pose, S.K, S.cols, S.rows, S.zmin, S.zmax, blockIdx.x, S.zr_groups);
}

// LDS-privatised z-range of every slot (grids of at most kZrangeLdsCells
// cells); gridDim.x workgroups share each stream's visible list.
__global__ void __launch_bounds__(256) k_s_zrange_lds(const StreamDev* streams, const Table* tab) {
  SLOT_OR_RETURN;
  if (sl.first) return;
  extern __shared__ unsigned lz[];
  zrange_lds_at(S.m, S.pose, S.K, S.cols, S.rows, S.zmin, S.zmax, lz, blockIdx.x, gridDim.x);
}

// Workgroups per stream of the batched LDS z-range.
constexpr int kZrangeLdsGroupsPerSlot = 64;

__global__ void __launch_bounds__(256) k_s_march_order(const StreamDev* streams, const Table* tab, MarchOrder mo) {
  SLOT_OR_RETURN;
  if (sl.first) return;
  const int tx = (S.W + 15) >> 4, T = tx * ((S.H + 15) >> 4);
  const int t = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (t >= T) return;
  const float* zmin = (const float*)S.zmin;
  const float* zmax = (const float*)S.zmax;
  const float eps = 1.1920928955078125e-07f;
  float span = 0.0f;
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 2; ++i) {
      const int x = (t % tx) * 2 + i, y = (t / tx) * 2 + j;
      if (x < S.cols && y < S.rows) {
        const float zn = zmin[y * S.cols + x], zf = zmax[y * S.cols + x];
        if (zn >= eps && zf > zn) span = fmaxf(span, zf - zn);
      }
    }
  const float v = span * S.vsinv;  // voxels
  const int c = v >= 96.0f ? 0 : v >= 32.0f ? 1 : v >= 8.0f ? 2 : 3;
  // one atomic per class per wavefront
  const unsigned long long below = lane_mask_lt();
  int pos = 0;
  for (int k = 0; k < kMarchClasses; ++k) {
    const unsigned long long mask = __ballot(c == k);
    if (mask == 0) continue;
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(&mo.counts[k], __popcll(mask));
    base = __shfl(base, leader, 64);
    if (c == k) pos = base + __popcll(mask & below);
  }
  mo.tiles[c * mo.cap + pos] = ((int)blockIdx.z << 16) | t;
}

template <class VT>
__global__ void __launch_bounds__(256) k_s_march(const StreamDev* streams, const Table* tab, MarchOrder mo) {
  int g = (int)(blockIdx.x + blockIdx.z * gridDim.x);
  int item = -1;
  for (int c = 0; c < kMarchClasses; ++c) {
    const int n = mo.counts[c];
    if (g < n) {
      // within a class, XCD bands again: consecutive entries (neighbouring
      // tiles, as the order kernel appends them) share an XCD's L2
      item = mo.tiles[c * mo.cap + xcd_tile(g, n)];
      break;
    }
    g -= n;
  }
  if (item < 0) return;
  const int slot = item >> 16, tile = item & 0xffff;
  const YAS_CONSTANT Table& ctab = *(const YAS_CONSTANT Table*)tab;
  const YAS_CONSTANT Slot& sl = ctab.slot[slot];
  const YAS_CONSTANT StreamDev& S = ((const YAS_CONSTANT StreamDev*)streams)[sl.stream];
  march_tile_at<VT>(S.m, S.pose, S.K, S.W, S.H, S.cols, (const float*)S.zmin, (const float*)S.zmax, S.vmap,
                    S.vsinv, S.step_scale, tile);
}

// Tracker inputs, two launches per pyramid level (the per-pixel functions
// of the separate image kernels, so the same values): the pointwise
// products of level l from level l - 1 (or the frame at level 0) -- model
// vertices (above level 0), intensity, inverse depth, vertices -- then the
// neighbourhood products of level l -- gradients and the model's and the
// frame's ICP points.  A stream's first frame has no model.
__global__ void __launch_bounds__(256) k_s_frame_level(const StreamDev* streams, const Table* tab, int l) {
  SLOT_OR_RETURN;
  if (l >= S.levels) return;
  const Level& a = S.lv[l];
  const TrackBufs& B = S.tb[l];
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= a.w || y >= a.h) return;
  const size_t k = (size_t)y * a.w + x;
  float in, iv;
  if (l == 0) {  // level 0's model vertices are the ray cast's own map
    in = ld(S.gray, (size_t)S.W * 4, x, y);
    iv = invdepth_of(ld(S.depth, (size_t)S.W * 4, x, y));
  } else {
    const Level& p = S.lv[l - 1];
    in = pyrdown_gauss_px(S.tb[l - 1].cur_int, (size_t)p.w * 4, p.w, p.h, x, y);
    iv = pyrdown_depth_px(S.cur_inv[l - 1], (size_t)p.w * 4, a.w, a.h, x, y);
    if (!sl.first) ((float4*)B.ref_v)[k] = pyrdown_vmap_px(S.tb[l - 1].ref_v, (size_t)p.w * 16, a.w, a.h, x, y);
  }
  ((float*)B.cur_int)[k] = in;
  S.cur_inv[l][k] = iv;
  ((float4*)B.cur_v)[k] = vertex_of(iv, x, y, a.invfx, a.invfy, a.cx, a.cy, 3.0f);
}
__global__ void __launch_bounds__(256) k_s_frame_level_nb(const StreamDev* streams, const Table* tab, int l) {
  SLOT_OR_RETURN;
  if (l >= S.levels) return;
  const Level& a = S.lv[l];
  const TrackBufs& B = S.tb[l];
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= a.w || y >= a.h) return;
  const size_t k = (size_t)y * a.w + x;
  const size_t p16 = (size_t)a.w * 16;
  if (!sl.first) store_icp_point((float4*)B.ref_p, a.w * a.h, (int)k, icp_point_px(B.ref_v, p16, a.w, a.h, x, y));
  float gxv, gyv;
  gradient_px(B.cur_int, (size_t)a.w * 4, a.w, a.h, x, y, gxv, gyv);
  ((float*)B.gx)[k] = gxv;
  ((float*)B.gy)[k] = gyv;
  store_icp_point((float4*)B.cur_p, a.w * a.h, (int)k, icp_point_px(B.cur_v, p16, a.w, a.h, x, y));
}
// After tracking: the current frame's intensity and inverse depth become the
// reference of the next frame (copies; the single pipeline swaps pointers).
// The current frame's intensity and inverse-depth pyramids become the
// reference ones (SwapFrameBuffer).  After a NaN update the reference keeps
// its buffers (RGBDTracking.cu:187-191 returns before the swap).
__global__ void __launch_bounds__(256) k_s_keep(const StreamDev* streams, const Table* tab, int l) {
  SLOT_OR_RETURN;
  if (l >= S.levels) return;
  if (!sl.first && !S.st->good) return;
  const Level& a = S.lv[l];
  const TrackBufs& B = S.tb[l];
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x < a.w && y < a.h) {
    const size_t k = (size_t)y * a.w + x;
    ((float*)B.ref_int)[k] = B.cur_int[k];
    ((float*)B.ref_inv)[k] = S.cur_inv[l][k];
  }
}

__global__ void k_s_track_init(const StreamDev* streams, const Table* tab, int slot0) {
  SLOT_FROM_OR_RETURN(slot0);
  if (threadIdx.x != 0 || sl.first) return;
  track_init_identity_at(S.st);
}

// At least six wavefronts per SIMD: at the batch's full size this pass is
// bound by HBM bandwidth, and the register cap keeps enough reads in flight.
template <bool ICP, bool RGB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_s_pass1(const StreamDev* streams, const Table* tab, int l, int slot0) {
  SLOT_FROM_OR_RETURN(slot0);
  if (sl.first || (int)blockIdx.x >= S.groups[l]) return;
  pass1_at<ICP, RGB, true>(S.lv[l], S.tb[l], S.st, S.part_icp, S.part_res, S.angle_th, S.dist_th, blockIdx.x,
                           S.groups[l]);
}

// The Huber threshold of a linearisation, once per stream (one workgroup
// each) rather than in every workgroup of pass 2: at the batch's size that
// fold was about a third of pass 2's instructions.
__global__ void __launch_bounds__(256) k_s_threshold(const StreamDev* streams, const Table* tab, int l, int slot0) {
  SLOT_FROM_OR_RETURN(slot0);
  if (sl.first) return;
  __shared__ float st_lds[3 * 4];
  huber_threshold(S.st, S.part_res, S.groups[l], st_lds, true);
}

__global__ void __launch_bounds__(256) k_s_pass2(const StreamDev* streams, const Table* tab, int l, int slot0) {
  SLOT_FROM_OR_RETURN(slot0);
  if (sl.first || (int)blockIdx.x >= S.groups[l]) return;
  pass2_at<true, true>(S.lv[l], S.tb[l], S.st, S.part_res, S.groups[l], S.part_rgb, blockIdx.x, S.groups[l]);
}

__global__ void __launch_bounds__(256) k_s_update(const StreamDev* streams, const Table* tab, int l, int icp, int rgb,
                                                  int mode, int first_of_level, float wrh, float wrb, float wih,
                                                  float wib, int slot0) {
  SLOT_FROM_OR_RETURN(slot0);
  if (sl.first) return;
  update_at(S.st, S.part_icp, icp ? S.groups[l] : 0, S.part_rgb, rgb ? S.groups[l] : 0, mode, first_of_level, 1,
            l == 0 ? 1 : 0, wrh, wrb, wih, wib);
}

__global__ void k_s_post(const StreamDev* streams, const Table* tab) {
  SLOT_OR_RETURN;
  if (threadIdx.x != 0) return;
  int tracked = 0;
  if (sl.first) {
    record_first_at(S.T_last, sl.out, &tracked);
    *S.gate = 1;
  } else {
    post_track_at(S.st->last_ok, &S.st->good, S.T_last, sl.out, &tracked, S.gate);
  }
  sl.out[16] = (double)tracked;
}

__global__ void __launch_bounds__(256) k_s_alloc(const StreamDev* streams, const Table* tab) {
  SLOT_OR_RETURN;
  if (!sl.first && *S.gate == 0) return;
  if ((int)blockIdx.x * 16 >= S.W || (int)blockIdx.y * 16 >= S.H) return;
  __shared__ int walked;
  if (threadIdx.x == 0) walked = 0;
  __syncthreads();
  alloc_walk(S.m, DepthF32{S.depth, (size_t)S.W * 4}, S.W, S.H, S.K, S.pose->T, blockIdx.x * 16 + (threadIdx.x & 15),
             blockIdx.y * 16 + (threadIdx.x >> 4), &walked);
  __syncthreads();
  if (threadIdx.x == 0 && walked) atomicAdd(&S.m.ctr[kCtrWalk], walked);
}

// At least six wavefronts per SIMD, as k_integrate.
template <class VT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_s_integrate(const StreamDev* streams, const Table* tab) {
  SLOT_OR_RETURN;
  if ((int)blockIdx.x >= S.int_groups) return;
  integrate_waves<VT>(S.m, S.pose->Tinv, S.K, DepthF32{S.depth, (size_t)S.W * 4}, S.W, S.H,
                      integrate_wave(S.m, blockIdx.x, S.int_groups), S.int_groups * (blockDim.x >> 6));
}

inline dim3 g64x4(int W, int H, int z) { return dim3((W + 63) / 64, (H + 3) / 4, z); }

struct Request {
  long long id;
  const uint16_t* depth;
  size_t dpitch;
  const uint8_t* rgb;
  size_t rpitch;
};

// A stream slot: free; pending (opened, waiting for HBM to admit it, no
// device memory yet); active (its pipeline resident); or hibernated (its
// volume paged out to pinned host memory to make room for others; the rest
// of its pipeline stays resident).
enum StreamState : int { kFree = 0, kPending = 1, kActive = 2, kHibernated = 3 };

struct HostStream {
  yas_pipeline* pipe = nullptr;
  int state = kFree;
  bool closing = false;   // released once its waiting and in-flight frames are done
  size_t bytes = 0;       // device bytes held now
  size_t full_bytes = 0;  // device bytes while active
  size_t map_bytes = 0;   // of which the volume's (released while hibernated)
  long long frames = 0;   // frames launched so far
  std::deque<Request> waiting;
  int in_flight = 0;
  long long order = 0;      // open order: pending streams are admitted first come, first served
  double last_use_ms = 0;   // host time of its last submit or launch (for idle eviction)
  // what it was opened with (a pending stream is created on admission)
  int W = 0, H = 0;
  float K[4] = {0, 0, 0, 0};
  yas_map_params mp{};
  float depth_scale = 1000.f;
  int gray_mode = 0;
};

inline double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Batch {
  Table* staged = nullptr;   // pinned
  double* results = nullptr; // pinned [kMaxSlots * kResultStride]
  hipEvent_t done = nullptr;
  bool busy = false;
  int count = 0;
  int stream_of[kMaxSlots];
  long long req_of[kMaxSlots];
};

// Device bytes of a volume: hash table, block pool and block lists.
size_t map_bytes(const yas_map_params& mp) {
  const size_t vb = mp.voxel_format == YAS_VOXEL_F32_F32 ? 8 : 4;
  return (size_t)mp.num_buckets * 16 + (size_t)mp.num_blocks * (vb * 512 + 8 + 4 + 4);
}

// Device bytes held by one stream (its map, ray caster, tracker and inputs).
size_t stream_bytes(int W, int H, const yas_map_params& mp, int levels) {
  const size_t vb = mp.voxel_format == YAS_VOXEL_F32_F32 ? 8 : 4;
  size_t b = map_bytes(mp);
  b += (size_t)W * H * 16 + (size_t)(W / 8) * (H / 8) * 8;  // vertex map, z-ranges
  size_t px = 0;
  for (int l = 0; l < levels; ++l) px += (size_t)(W >> l) * (H >> l);
  b += px * (6 * 4 + 5 * 16);                                 // tracker pyramids
  b += 256 * 68 * 4 + 1024;                                    // partials, state
  b += (size_t)W * H * 8 + 4096 + 1024 * (16 * 8 + 4 + 32);    // inputs, pose records
  return b;
}

}  // namespace

struct yas_server {
  yas_ctx* ctx;      // own stream (graphs cannot be captured on the null stream)
  int device;
  yas_tracker_params tp;
  int voxel_format;
  int maxW, maxH;
  size_t budget, used;
  int max_streams;
  int max_batch;      // frames per batch (<= kMaxSlots)
  int max_in_flight;  // batches queued on the device at once (<= kInFlight)
  int min_ahead;      // frames that must wait before a batch queues behind a running one
  int gn_group;       // slots per Gauss-Newton sub-range (0: the whole batch)
  std::vector<HostStream> streams;  // slot per stream id
  double evict_idle_ms;  // < 0: never hibernate; else idle streams older than this may be paged out
  long long next_order;
  long long hibernations, reactivations;
  StreamDev* dev_streams;
  Table* dev_table;
  double* dev_results;  // [kInFlight][kMaxSlots * kResultStride]
  int* march_tiles;     // ray-march dispatch order, [kMarchClasses][kMaxSlots * tiles]
  int* march_counts;    // [kMarchClasses]
  Batch batches[kInFlight];
  int next_batch;
  std::map<int, hipGraphExec_t> graphs;  // batch-size class -> graph
  std::vector<hipGraph_t> graph_defs;
  std::deque<yas_result> done;
  std::deque<int> launch_order;  // batches in flight, oldest first
  long long next_request;
  int rr;  // stream the next batch starts looking at
  long long frames_done, batches_done, slots_done;
};

static int launch_batch_kernels(yas_server* sv, hipStream_t s, int nz) {
  const StreamDev* S = sv->dev_streams;
  const Table* T = sv->dev_table;
  const int W = sv->maxW, H = sv->maxH;
  const int L = sv->tp.levels;
  const bool icp = sv->tp.mode != YAS_TRACK_RGB, rgb = sv->tp.mode != YAS_TRACK_DEPTH;
  auto level_grid = [&](int l) { return g64x4(W >> l, H >> l, nz); };
  // inputs
  hipLaunchKernelGGL(k_s_convert, g64x4(W, H, nz), dim3(256), 0, s, S, T);
  // ray cast at the last pose
  hipLaunchKernelGGL(k_s_pose, dim3(1, 1, nz), dim3(64), 0, s, S, T, 1);
  hipLaunchKernelGGL(k_s_visible, dim3(4096, 1, nz), dim3(256), 0, s, S, T, 0);
  const int tiles = ((W + 15) / 16) * ((H + 15) / 16);
  const MarchOrder mo{sv->march_tiles, sv->march_counts, kMaxSlots * tiles};
  hipLaunchKernelGGL(k_s_zrange_init, dim3(((W / 8) * (H / 8) + 255) / 256, 1, nz), dim3(256), 0, s, S, T, mo);
  const int ncell = (W / 8) * (H / 8);
  if (ncell <= kZrangeLdsCells)
    hipLaunchKernelGGL(k_s_zrange_lds, dim3(kZrangeLdsGroupsPerSlot, 1, nz), dim3(256), 2 * ncell * sizeof(unsigned),
                       s, S, T);
  else
    hipLaunchKernelGGL(k_s_zrange, dim3(2048, 1, nz), dim3(256), 0, s, S, T);
  hipLaunchKernelGGL(k_s_march_order, dim3((tiles + 255) / 256, 1, nz), dim3(256), 0, s, S, T, mo);
  const dim3 mg(tiles, 1, nz);  // one workgroup per tile of every image, in the order's sequence
  switch (sv->voxel_format) {
    case YAS_VOXEL_I16_U8: hipLaunchKernelGGL(k_s_march<VoxI16>, mg, dim3(256), 0, s, S, T, mo); break;
    case YAS_VOXEL_F16_U16: hipLaunchKernelGGL(k_s_march<VoxF16>, mg, dim3(256), 0, s, S, T, mo); break;
    default: hipLaunchKernelGGL(k_s_march<VoxF32>, mg, dim3(256), 0, s, S, T, mo);
  }
  // tracker inputs: reference model pyramid, current intensity and depth pyramids
  for (int l = 0; l < L; ++l) hipLaunchKernelGGL(k_s_frame_level, level_grid(l), dim3(256), 0, s, S, T, l);
  for (int l = 0; l < L; ++l) hipLaunchKernelGGL(k_s_frame_level_nb, level_grid(l), dim3(256), 0, s, S, T, l);
  // Gauss-Newton, coarse to fine; optionally on sub-ranges of the slots in
  // turn, so one sub-range's pyramids stay in the last-level cache across
  // the iterations
  const int gg = sv->gn_group > 0 ? std::min(sv->gn_group, nz) : nz;
  for (int s0 = 0; s0 < nz; s0 += gg) {
    const int gz = std::min(gg, nz - s0);
    hipLaunchKernelGGL(k_s_track_init, dim3(1, 1, gz), dim3(64), 0, s, S, T, s0);
    for (int l = L - 1; l >= 0; --l) {
      const int G = std::max(1, std::min(256, ((W >> l) * (H >> l) + 1023) / 1024));
      for (int it = 0; it < sv->tp.iters[l]; ++it) {
        if (icp && rgb) hipLaunchKernelGGL((k_s_pass1<true, true>), dim3(G, 1, gz), dim3(256), 0, s, S, T, l, s0);
        else if (icp) hipLaunchKernelGGL((k_s_pass1<true, false>), dim3(G, 1, gz), dim3(256), 0, s, S, T, l, s0);
        else hipLaunchKernelGGL((k_s_pass1<false, true>), dim3(G, 1, gz), dim3(256), 0, s, S, T, l, s0);
        if (rgb) {
          hipLaunchKernelGGL(k_s_threshold, dim3(1, 1, gz), dim3(256), 0, s, S, T, l, s0);
          hipLaunchKernelGGL(k_s_pass2, dim3(G, 1, gz), dim3(256), 0, s, S, T, l, s0);
        }
        hipLaunchKernelGGL(k_s_update, dim3(1, 1, gz), dim3(256), 0, s, S, T, l, icp ? 1 : 0, rgb ? 1 : 0,
                           sv->tp.mode, it == 0 ? 1 : 0, sv->tp.w_rgb_h, sv->tp.w_rgb_b, sv->tp.w_icp_h,
                           sv->tp.w_icp_b, s0);
      }
    }
  }
  // gate, pose chain, result record
  hipLaunchKernelGGL(k_s_post, dim3(1, 1, nz), dim3(64), 0, s, S, T);
  // the current pyramids become the reference of the next frame
  for (int l = 0; l < L; ++l) hipLaunchKernelGGL(k_s_keep, level_grid(l), dim3(256), 0, s, S, T, l);
  // fusion at the new pose
  hipLaunchKernelGGL(k_s_pose, dim3(1, 1, nz), dim3(64), 0, s, S, T, 0);
  hipLaunchKernelGGL(k_s_alloc, dim3((W + 15) / 16, (H + 15) / 16, nz), dim3(256), 0, s, S, T);
  hipLaunchKernelGGL(k_s_visible, dim3(4096, 1, nz), dim3(256), 0, s, S, T, 1);
  switch (sv->voxel_format) {
    case YAS_VOXEL_I16_U8: hipLaunchKernelGGL(k_s_integrate<VoxI16>, dim3(2048, 1, nz), dim3(256), 0, s, S, T); break;
    case YAS_VOXEL_F16_U16: hipLaunchKernelGGL(k_s_integrate<VoxF16>, dim3(2048, 1, nz), dim3(256), 0, s, S, T); break;
    default: hipLaunchKernelGGL(k_s_integrate<VoxF32>, dim3(2048, 1, nz), dim3(256), 0, s, S, T);
  }
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

static int graph_for(yas_server* sv, int count, hipGraphExec_t* out) {
  int nz = 1;
  while (nz < count) nz <<= 1;
  auto f = sv->graphs.find(nz);
  if (f != sv->graphs.end()) {
    *out = f->second;
    return YAS_OK;
  }
  hipStream_t s = sv->ctx->stream;
  YAS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = launch_batch_kernels(sv, s, nz);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  YAS_HIP(e);
  hipGraphExec_t ex = nullptr;
  YAS_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
  sv->graph_defs.push_back(g);
  sv->graphs[nz] = ex;
  *out = ex;
  return YAS_OK;
}

// Completed batches -> results queue, in launch order (batches finish in
// that order on the server's stream; the oldest unfinished one stops the
// scan, so results of a stream always come out in submission order).
static int collect(yas_server* sv, bool wait) {
  while (!sv->launch_order.empty()) {
    Batch& b = sv->batches[sv->launch_order.front()];
    if (wait) YAS_HIP(hipEventSynchronize(b.done));
    hipError_t q = hipEventQuery(b.done);
    if (q == hipErrorNotReady) break;
    YAS_HIP(q);
    sv->launch_order.pop_front();
    for (int i = 0; i < b.count; ++i) {
      yas_result r;
      r.stream = b.stream_of[i];
      r.request = b.req_of[i];
      const double* rec = b.results + (size_t)i * kResultStride;
      std::memcpy(r.pose, rec, sizeof(r.pose));
      r.tracked = rec[16] != 0.0 ? 1 : 0;
      sv->done.push_back(r);
      sv->streams[r.stream].in_flight--;
    }
    sv->frames_done += b.count;
    sv->batches_done += 1;
    b.busy = false;
  }
  return YAS_OK;
}

extern "C" {

int yas_server_create(yas_ctx* ctx, int max_width, int max_height, int max_streams, const yas_tracker_params* tp,
                      int voxel_format, size_t hbm_budget_bytes, yas_server** out) {
  YAS_REQUIRE(ctx && tp && out, "yas_server_create: null argument");
  YAS_REQUIRE(max_width >= 8 && max_height >= 8 && max_streams >= 1, "yas_server_create: bad sizes");
  YAS_REQUIRE((long long)((max_width + 15) / 16) * ((max_height + 15) / 16) < 65536,
              "yas_server_create: images of at most 65535 16x16 tiles");
  YAS_REQUIRE(tp->levels >= 1 && tp->levels <= kTrackMaxLevels, "yas_server_create: levels must be 1..5");
  YAS_REQUIRE(voxel_format >= 0 && voxel_format <= 2, "yas_server_create: unknown voxel format");
  yas_server* sv = new yas_server();
  sv->device = ctx->device;
  sv->tp = *tp;
  sv->voxel_format = voxel_format;
  sv->maxW = max_width;
  sv->maxH = max_height;
  sv->budget = hbm_budget_bytes;
  sv->used = 0;
  sv->max_streams = max_streams;
  sv->max_batch = kMaxSlots;
  // Measured with 32 mixed-resolution streams under Poisson load: two in
  // flight, the second only once 16 frames wait, gives the lowest p99
  // latency from 3k to 6.5k frames/s offered (p99 8.6 ms at 4.4k against
  // 11.0 ms with three always queued), and the same saturation rate.
  sv->max_in_flight = 2;
  sv->min_ahead = 16;
  sv->gn_group = 0;
  sv->streams.resize(max_streams);
  sv->next_batch = 0;
  sv->next_request = 0;
  sv->rr = 0;
  sv->frames_done = sv->batches_done = sv->slots_done = 0;
  sv->evict_idle_ms = -1.0;
  sv->next_order = 0;
  sv->hibernations = sv->reactivations = 0;
  int rc = yas_ctx_create(ctx->device, nullptr, &sv->ctx);
  if (rc) {
    delete sv;
    return rc;
  }
  hipError_t e = hipMalloc((void**)&sv->dev_streams, sizeof(StreamDev) * max_streams);
  if (e == hipSuccess) e = hipMalloc((void**)&sv->dev_table, sizeof(Table));
  if (e == hipSuccess) e = hipMalloc((void**)&sv->dev_results, sizeof(double) * kInFlight * kMaxSlots * kResultStride);
  const size_t max_tiles = (size_t)((max_width + 15) / 16) * ((max_height + 15) / 16);
  if (e == hipSuccess) e = hipMalloc((void**)&sv->march_tiles, sizeof(int) * kMarchClasses * kMaxSlots * max_tiles);
  if (e == hipSuccess) e = hipMalloc((void**)&sv->march_counts, sizeof(int) * kMarchClasses);
  for (int k = 0; e == hipSuccess && k < kInFlight; ++k) {
    e = hipHostMalloc((void**)&sv->batches[k].staged, sizeof(Table), hipHostMallocDefault);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&sv->batches[k].results, sizeof(double) * kMaxSlots * kResultStride, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->batches[k].done, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    yas_set_error("yas_server_create: %s", hipGetErrorString(e));
    yas_server_destroy(sv);
    return YAS_E_HIP;
  }
  *out = sv;
  return YAS_OK;
}

int yas_server_destroy(yas_server* sv) {
  if (!sv) return YAS_OK;
  hipSetDevice(sv->device);
  if (sv->ctx) hipStreamSynchronize(sv->ctx->stream);
  for (auto& kv : sv->graphs) hipGraphExecDestroy(kv.second);
  for (hipGraph_t g : sv->graph_defs) hipGraphDestroy(g);
  for (auto& hs : sv->streams)
    if (hs.pipe) yas_pipeline_destroy(hs.pipe);
  for (auto& b : sv->batches) {
    if (b.staged) hipHostFree(b.staged);
    if (b.results) hipHostFree(b.results);
    if (b.done) hipEventDestroy(b.done);
  }
  hipFree(sv->dev_streams);
  hipFree(sv->dev_table);
  hipFree(sv->dev_results);
  hipFree(sv->march_tiles);
  hipFree(sv->march_counts);
  if (sv->ctx) yas_ctx_destroy(sv->ctx);
  delete sv;
  return YAS_OK;
}

// ------------------------------------------------------------------ admission
// Streams are admitted by HBM budget.  A stream that does not fit waits
// (pending) and is admitted, first come first served, as memory frees up:
// when a stream is closed or removed, or, with eviction enabled, when idle
// streams are hibernated (their volumes paged out to pinned host memory,
// VoxelMap.cu:894-954).  A hibernated stream that receives a frame is
// re-activated as soon as its volume fits again.

static int refresh_descriptor(yas_server* sv, int id) {
  StreamDev d;
  int rc = yas_pipeline_fill(sv->streams[id].pipe, d);
  if (rc) return rc;
  YAS_HIP(hipMemcpyAsync(sv->dev_streams + id, &d, sizeof(d), hipMemcpyHostToDevice, sv->ctx->stream));
  YAS_HIP(hipStreamSynchronize(sv->ctx->stream));
  return YAS_OK;
}

static int activate(yas_server* sv, int id) {
  HostStream& hs = sv->streams[id];
  yas_pipeline* p = nullptr;
  int rc = yas_pipeline_create(sv->ctx, hs.W, hs.H, hs.K, &hs.mp, &sv->tp, hs.depth_scale, hs.gray_mode, &p);
  if (rc) return rc;
  hs.pipe = p;
  if ((rc = refresh_descriptor(sv, id))) {
    yas_pipeline_destroy(p);
    hs.pipe = nullptr;
    return rc;
  }
  hs.state = kActive;
  hs.bytes = hs.full_bytes;
  hs.frames = 0;
  hs.in_flight = 0;
  hs.last_use_ms = now_ms();
  sv->used += hs.full_bytes;
  return YAS_OK;
}

static void release(yas_server* sv, int id) {
  HostStream& hs = sv->streams[id];
  if (hs.pipe) yas_pipeline_destroy(hs.pipe);
  sv->used -= hs.bytes;
  hs = HostStream();
}

static bool idle_for_eviction(const yas_server* sv, const HostStream& hs, double now) {
  return hs.state == kActive && !hs.closing && hs.waiting.empty() && hs.in_flight == 0 &&
         now - hs.last_use_ms >= sv->evict_idle_ms;
}

// Make `need` bytes fit under the budget by hibernating idle streams (least
// recently used first), if eviction is on and enough of them are idle.
static int make_room(yas_server* sv, size_t need, bool* fits) {
  *fits = sv->used + need <= sv->budget;
  if (*fits || sv->evict_idle_ms < 0) return YAS_OK;
  const double now = now_ms();
  std::vector<int> cand;
  size_t freeable = 0;
  for (int i = 0; i < sv->max_streams; ++i)
    if (idle_for_eviction(sv, sv->streams[i], now)) {
      cand.push_back(i);
      freeable += sv->streams[i].map_bytes;
    }
  if (sv->used - freeable + need > sv->budget) return YAS_OK;  // not even all of them would do
  std::sort(cand.begin(), cand.end(),
            [&](int a, int b) { return sv->streams[a].last_use_ms < sv->streams[b].last_use_ms; });
  for (int i : cand) {
    if (sv->used + need <= sv->budget) break;
    HostStream& hs = sv->streams[i];
    int rc = yas_map_hibernate(yas_pipeline_map_of(hs.pipe));
    if (rc) return rc;
    hs.state = kHibernated;
    hs.bytes -= hs.map_bytes;
    sv->used -= hs.map_bytes;
    sv->hibernations++;
  }
  *fits = sv->used + need <= sv->budget;
  return YAS_OK;
}

// Release finished closing streams, re-activate hibernated streams that have
// frames waiting (oldest request first), then admit pending streams in open
// order while they fit.
static int admit(yas_server* sv) {
  for (int i = 0; i < sv->max_streams; ++i) {
    HostStream& hs = sv->streams[i];
    if (hs.state != kFree && hs.closing && hs.waiting.empty() && hs.in_flight == 0) release(sv, i);
  }
  std::vector<int> wake;
  for (int i = 0; i < sv->max_streams; ++i)
    if (sv->streams[i].state == kHibernated && !sv->streams[i].waiting.empty()) wake.push_back(i);
  std::sort(wake.begin(), wake.end(),
            [&](int a, int b) { return sv->streams[a].waiting.front().id < sv->streams[b].waiting.front().id; });
  for (int i : wake) {
    HostStream& hs = sv->streams[i];
    bool fits = false;
    int rc = make_room(sv, hs.map_bytes, &fits);
    if (rc) return rc;
    if (!fits) continue;
    if ((rc = yas_map_reactivate(yas_pipeline_map_of(hs.pipe)))) return rc;
    hs.state = kActive;
    hs.bytes += hs.map_bytes;
    sv->used += hs.map_bytes;
    sv->reactivations++;
    if ((rc = refresh_descriptor(sv, i))) return rc;
  }
  std::vector<int> pend;
  for (int i = 0; i < sv->max_streams; ++i)
    if (sv->streams[i].state == kPending) pend.push_back(i);
  std::sort(pend.begin(), pend.end(), [&](int a, int b) { return sv->streams[a].order < sv->streams[b].order; });
  for (int i : pend) {
    bool fits = false;
    int rc = make_room(sv, sv->streams[i].full_bytes, &fits);
    if (rc) return rc;
    if (!fits) break;
    if ((rc = activate(sv, i))) return rc;
  }
  return YAS_OK;
}

static int open_slot(yas_server* sv, int W, int H, const float K[4], const yas_map_params* mp, float depth_scale,
                     int gray_mode, int* id_out) {
  YAS_REQUIRE(W <= sv->maxW && H <= sv->maxH, "yas_server: image larger than the server's maximum");
  YAS_REQUIRE(mp->voxel_format == sv->voxel_format, "yas_server: voxel format differs from the server's");
  YAS_REQUIRE(depth_scale > 0.f, "yas_server: depth_scale must be > 0");
  int id = -1;
  for (int i = 0; i < sv->max_streams; ++i)
    if (sv->streams[i].state == kFree) {
      id = i;
      break;
    }
  if (id < 0) {
    yas_set_error("yas_server: all %d stream slots are taken", sv->max_streams);
    return YAS_E_CAPACITY;
  }
  HostStream& hs = sv->streams[id];
  hs = HostStream();
  hs.W = W;
  hs.H = H;
  std::memcpy(hs.K, K, sizeof(hs.K));
  hs.mp = *mp;
  hs.depth_scale = depth_scale;
  hs.gray_mode = gray_mode;
  hs.full_bytes = stream_bytes(W, H, *mp, sv->tp.levels);
  hs.map_bytes = map_bytes(*mp);
  hs.order = sv->next_order++;
  *id_out = id;
  return YAS_OK;
}

int yas_server_add_stream(yas_server* sv, int W, int H, const float K[4], const yas_map_params* mp, float depth_scale,
                          int gray_mode, int* stream_id) {
  YAS_REQUIRE(sv && K && mp && stream_id, "yas_server_add_stream: null argument");
  YAS_HIP(hipSetDevice(sv->device));
  int id = -1;
  int rc = open_slot(sv, W, H, K, mp, depth_scale, gray_mode, &id);
  if (rc) return rc;
  HostStream& hs = sv->streams[id];
  bool fits = false;
  if ((rc = make_room(sv, hs.full_bytes, &fits))) {
    hs = HostStream();
    return rc;
  }
  if (!fits) {
    yas_set_error("yas_server_add_stream: %zu bytes would exceed the HBM budget (%zu of %zu in use)", hs.full_bytes,
                  sv->used, sv->budget);
    hs = HostStream();
    return YAS_E_CAPACITY;
  }
  if ((rc = activate(sv, id))) {
    hs = HostStream();
    return rc;
  }
  *stream_id = id;
  return YAS_OK;
}

int yas_server_open_stream(yas_server* sv, int W, int H, const float K[4], const yas_map_params* mp,
                           float depth_scale, int gray_mode, int* stream_id) {
  YAS_REQUIRE(sv && K && mp && stream_id, "yas_server_open_stream: null argument");
  YAS_HIP(hipSetDevice(sv->device));
  int id = -1;
  int rc = open_slot(sv, W, H, K, mp, depth_scale, gray_mode, &id);
  if (rc) return rc;
  sv->streams[id].state = kPending;
  *stream_id = id;
  return admit(sv);
}

int yas_server_close_stream(yas_server* sv, int stream_id) {
  YAS_REQUIRE(sv && stream_id >= 0 && stream_id < sv->max_streams && sv->streams[stream_id].state != kFree,
              "yas_server_close_stream: no such stream");
  YAS_HIP(hipSetDevice(sv->device));
  HostStream& hs = sv->streams[stream_id];
  if (hs.state == kPending) hs.waiting.clear();  // never admitted: its frames are dropped
  hs.closing = true;
  int rc = collect(sv, false);
  return rc ? rc : admit(sv);
}

int yas_server_remove_stream(yas_server* sv, int stream_id) {
  YAS_REQUIRE(sv && stream_id >= 0 && stream_id < sv->max_streams && sv->streams[stream_id].state != kFree,
              "yas_server_remove_stream: no such stream");
  YAS_HIP(hipSetDevice(sv->device));
  HostStream& hs = sv->streams[stream_id];
  YAS_REQUIRE(hs.waiting.empty(), "yas_server_remove_stream: the stream still has frames waiting");
  while (hs.in_flight > 0) {
    int rc = collect(sv, true);
    if (rc) return rc;
  }
  release(sv, stream_id);
  return admit(sv);
}

int yas_server_set_eviction(yas_server* sv, double idle_ms) {
  YAS_REQUIRE(sv, "null server");
  sv->evict_idle_ms = idle_ms;
  return YAS_OK;
}

int yas_server_stream_state(yas_server* sv, int stream_id, int* state) {
  YAS_REQUIRE(sv && state && stream_id >= 0 && stream_id < sv->max_streams, "yas_server_stream_state: bad argument");
  *state = sv->streams[stream_id].state;
  return YAS_OK;
}

int yas_server_stream_map(yas_server* sv, int stream_id, yas_map** map) {
  YAS_REQUIRE(sv && map && stream_id >= 0 && stream_id < sv->max_streams && sv->streams[stream_id].pipe,
              "yas_server_stream_map: no such stream");
  *map = yas_pipeline_map_of(sv->streams[stream_id].pipe);
  return YAS_OK;
}

int yas_server_set_max_batch(yas_server* sv, int max_batch) {
  YAS_REQUIRE(sv && max_batch >= 1 && max_batch <= kMaxSlots, "yas_server_set_max_batch: 1..64");
  sv->max_batch = max_batch;
  return YAS_OK;
}

int yas_server_set_max_in_flight(yas_server* sv, int batches, int min_ahead) {
  YAS_REQUIRE(sv && batches >= 1 && batches <= kInFlight && min_ahead >= 1,
              "yas_server_set_max_in_flight: 1..3 batches, min_ahead >= 1");
  sv->max_in_flight = batches;
  sv->min_ahead = min_ahead;
  return YAS_OK;
}

int yas_server_set_gn_group(yas_server* sv, int slots) {
  YAS_REQUIRE(sv && slots >= 0 && slots <= kMaxSlots, "yas_server_set_gn_group: 0..64");
  if (slots == sv->gn_group) return YAS_OK;
  YAS_HIP(hipStreamSynchronize(sv->ctx->stream));
  for (auto& kv : sv->graphs) hipGraphExecDestroy(kv.second);
  for (hipGraph_t g : sv->graph_defs) hipGraphDestroy(g);
  sv->graphs.clear();
  sv->graph_defs.clear();
  sv->gn_group = slots;
  return YAS_OK;
}

int yas_server_prepare(yas_server* sv) {
  YAS_REQUIRE(sv, "null server");
  YAS_HIP(hipSetDevice(sv->device));
  for (int n = 1;; n <<= 1) {
    hipGraphExec_t ex = nullptr;
    int rc = graph_for(sv, n, &ex);
    if (rc) return rc;
    if (n >= sv->max_batch) break;
  }
  return YAS_OK;
}

int yas_server_stream_pipeline(yas_server* sv, int stream_id, yas_pipeline** pipe) {
  YAS_REQUIRE(sv && pipe && stream_id >= 0 && stream_id < sv->max_streams && sv->streams[stream_id].pipe,
              "yas_server_stream_pipeline: no such stream");
  *pipe = sv->streams[stream_id].pipe;
  return YAS_OK;
}

int yas_server_submit(yas_server* sv, int stream_id, const uint16_t* depth, size_t dpitch, const uint8_t* rgb,
                      size_t rpitch, long long* request_id) {
  YAS_REQUIRE(sv && depth && rgb && stream_id >= 0 && stream_id < sv->max_streams &&
                  sv->streams[stream_id].state != kFree && !sv->streams[stream_id].closing,
              "yas_server_submit: no such open stream");
  Request r{sv->next_request++, depth, dpitch, rgb, rpitch};
  HostStream& hs = sv->streams[stream_id];
  hs.waiting.push_back(r);
  hs.last_use_ms = now_ms();
  if (request_id) *request_id = r.id;
  return YAS_OK;
}

int yas_server_pump(yas_server* sv, int* launched) {
  YAS_REQUIRE(sv, "null server");
  YAS_HIP(hipSetDevice(sv->device));
  int n_launched = 0;
  int rc = collect(sv, false);
  if (!rc) rc = admit(sv);
  if (rc) return rc;
  for (;;) {
    Batch& b = sv->batches[sv->next_batch];
    if (b.busy) break;  // kInFlight batches queued already
    int busy = 0;
    for (int k = 0; k < kInFlight; ++k) busy += sv->batches[k].busy ? 1 : 0;
    if (busy >= sv->max_in_flight) break;
    // size of the next batch, and its graph, before any frame leaves its queue
    int want = 0;
    for (int i = 0; i < sv->max_streams && want < sv->max_batch; ++i)
      if (sv->streams[i].state == kActive && !sv->streams[i].waiting.empty()) ++want;
    if (want == 0 || (busy > 0 && want < std::min(sv->min_ahead, sv->max_batch))) break;
    hipGraphExec_t ex = nullptr;
    if ((rc = graph_for(sv, want, &ex))) return rc;
    int count = 0;
    const double now = now_ms();
    // round-robin start, so a batch cap does not favour low stream ids
    for (int j = 0; j < sv->max_streams && count < sv->max_batch; ++j) {
      const int i = (sv->rr + j) % sv->max_streams;
      HostStream& hs = sv->streams[i];
      if (hs.state != kActive || hs.waiting.empty()) continue;
      const Request r = hs.waiting.front();
      hs.waiting.pop_front();
      Slot& s = b.staged->slot[count];
      s.stream = i;
      s.first = hs.frames == 0 ? 1 : 0;
      s.depth_in = r.depth;
      s.dpitch = r.dpitch;
      s.rgb_in = r.rgb;
      s.rpitch = r.rpitch;
      s.out = sv->dev_results + ((size_t)sv->next_batch * kMaxSlots + count) * kResultStride;
      b.stream_of[count] = i;
      b.req_of[count] = r.id;
      hs.frames++;
      hs.in_flight++;
      hs.last_use_ms = now;
      ++count;
    }
    sv->rr = (b.stream_of[count - 1] + 1) % sv->max_streams;
    b.staged->count = count;
    b.count = count;
    hipStream_t s = sv->ctx->stream;
    YAS_HIP(hipMemcpyAsync(sv->dev_table, b.staged, sizeof(int) * 2 + sizeof(Slot) * count, hipMemcpyHostToDevice, s));
    YAS_HIP(hipGraphLaunch(ex, s));
    YAS_HIP(hipMemcpyAsync(b.results, sv->dev_results + (size_t)sv->next_batch * kMaxSlots * kResultStride,
                           sizeof(double) * kResultStride * count, hipMemcpyDeviceToHost, s));
    YAS_HIP(hipEventRecord(b.done, s));
    b.busy = true;
    sv->launch_order.push_back(sv->next_batch);
    sv->slots_done += count;
    sv->next_batch = (sv->next_batch + 1) % kInFlight;
    ++n_launched;
  }
  if (launched) *launched = n_launched;
  return YAS_OK;
}

int yas_server_poll(yas_server* sv, yas_result* out, int capacity, int* count) {
  YAS_REQUIRE(sv && count && (out || capacity == 0), "yas_server_poll: bad argument");
  int rc = collect(sv, false);
  if (rc) return rc;
  int n = 0;
  while (n < capacity && !sv->done.empty()) {
    out[n++] = sv->done.front();
    sv->done.pop_front();
  }
  *count = n;
  return YAS_OK;
}

int yas_server_drain(yas_server* sv) {
  YAS_REQUIRE(sv, "null server");
  for (;;) {
    int launched = 0;
    int rc = yas_server_pump(sv, &launched);
    if (rc) return rc;
    bool waiting = false, stuck = false;
    for (auto& hs : sv->streams) {
      if (hs.waiting.empty()) continue;
      if (hs.state == kActive) waiting = true;
      else stuck = true;
    }
    bool busy = false;
    for (auto& b : sv->batches) busy = busy || b.busy;
    if (!waiting && !busy) {
      if (!stuck) return YAS_OK;
      if ((rc = admit(sv))) return rc;
      bool admitted = false;
      for (auto& hs : sv->streams) admitted = admitted || (hs.state == kActive && !hs.waiting.empty());
      if (!admitted) {
        yas_set_error("yas_server_drain: frames wait for streams that do not fit the HBM budget");
        return YAS_E_CAPACITY;
      }
      continue;
    }
    if (!launched) {
      if ((rc = collect(sv, true))) return rc;
    }
  }
}

int yas_server_get_stats(yas_server* sv, yas_server_stats* st) {
  YAS_REQUIRE(sv && st, "null argument");
  int n = 0, pending = 0, hib = 0;
  for (auto& hs : sv->streams) {
    n += hs.state == kActive || hs.state == kHibernated;
    pending += hs.state == kPending;
    hib += hs.state == kHibernated;
  }
  st->streams = n;
  st->bytes_in_use = sv->used;
  st->bytes_budget = sv->budget;
  st->frames_done = sv->frames_done;
  st->batches_done = sv->batches_done;
  st->streams_pending = pending;
  st->streams_hibernated = hib;
  st->hibernations = sv->hibernations;
  st->reactivations = sv->reactivations;
  return YAS_OK;
}

}  // extern "C"
