// The dense half of the reference's per-frame tracking step
// (FullSystem/src/Tracking.cpp:139-197, StereoInitialization + TrackRGBD) on
// one volume, with every pose kept in device memory: ray cast at the last
// pose, track against the ray-cast model, gate |t| > 0.1 m, compose
// T_cur = T_last * DT^-1 and fuse at T_cur.  A frame enqueues kernels only;
// the host never waits on the device between frames.
#include <algorithm>
#include <cstring>
#include <vector>

#include "yas_internal.hpp"
#include "yas_raycast_dev.hpp"
#include "yas_server_dev.hpp"

using namespace yas;

int yas_map_fuse_prepared(yas_map* map, const float* depth, size_t pitch, int W, int H, const int* gate);
int yas_map_alloc_stage(yas_map* map, const float* depth, size_t pitch, int W, int H, const int* gate,
                        const hipEvent_t* ev);
int yas_map_visible_stage(yas_map* map, int W, int H, const int* gate, const hipEvent_t* ev);
int yas_map_integrate_stage(yas_map* map, const float* depth, size_t pitch, int W, int H, const hipEvent_t* ev,
                            int* stats);
MapView yas_map_view(yas_map* map);
PosePair* yas_map_pose_dev_mut(yas_map* map);
int yas_raycast_prepared(yas_raycaster* rc, yas_map* map, hipEvent_t before_march, const hipEvent_t* march_ev);
float4* yas_raycaster_vmap_dev(yas_raycaster* rc);
void yas_raycaster_zrange_dev(yas_raycaster* rc, unsigned** zmin, unsigned** zmax, int* ncell);
TrackState* yas_tracker_state_dev(yas_tracker* t);
int yas_tracker_solve_async(yas_tracker* t);
const double* yas_tracker_result_dev(yas_tracker* t);
const int* yas_tracker_good_dev(yas_tracker* t);
int yas_tracker_set_tracking_image_dev(yas_tracker* t, const float* g, size_t pitch);
int yas_tracker_prepare_frame_dev(yas_tracker* t, const float* vmap, size_t vpitch, const float* g, size_t gpitch,
                                  const float* d, size_t dpitch);
int yas_tracker_set_tracking_depth_dev(yas_tracker* t, const float* d, size_t pitch);
int yas_tracker_unswap_copies(yas_tracker* t, RefRestore& r);

namespace {

// The start of a frame in one launch of 64x4-pixel workgroups: depth in
// metres and the gray image; from the first workgroup the map's float pose
// pair of T_last with its visible counter cleared (for the ray cast, or the
// first frame's fusion) and the tracker's state reset to identity; from the
// first ncell threads the ray caster's z-range cells reset.
struct FrameBegin {
  const uint16_t* depth_in;
  size_t dpitch;
  const uint8_t* rgb;
  size_t cpitch;
  float* depth;  // pitch fpitch
  float* gray;   // pitch fpitch
  size_t fpitch;
  float scale;
  int W, H, gray_mode;
  const double* T_last;
  PosePair* pose;
  int* ctr;
  TrackState* st;
  unsigned* zmin;
  unsigned* zmax;
  int ncell;
};

__global__ void __launch_bounds__(256) k_frame_begin(FrameBegin a) {
  const int gid = (blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid < a.ncell) {
    a.zmin[gid] = __float_as_uint(kZrangeFar);
    a.zmax[gid] = 0u;
  }
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    if (threadIdx.x == 0) {
      PosePair p;
      pose_pair_from(a.T_last, p);
      *a.pose = p;
      a.ctr[kCtrVisible] = 0;
    }
    track_init_identity_at(a.st);
  }
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x < a.W && y < a.H) {
    const uint16_t d = ((const uint16_t*)((const char*)a.depth_in + (size_t)y * a.dpitch))[x];
    ((float*)((char*)a.depth + (size_t)y * a.fpitch))[x] = (float)d * a.scale;
  }
  rgb_to_gray_at(a.rgb, a.cpitch, a.gray, a.fpitch, a.W, a.H, a.gray_mode, blockIdx.x, blockIdx.y);
}

// The frame's outcome (one thread), which then writes the map's pose pair
// of the (possibly new) T_last for the fusion and clears its visible
// counter, after saving the ray cast's visible count; and after a NaN
// update the reference pyramids put back as they were before the host-side
// swap (all threads).
__global__ void __launch_bounds__(256) k_post_track(const double* DT, const int* good, double* T_last,
                                                    double* pose_out, int* tracked, int* gate, int* ctr,
                                                    int* ray_visible_out, PosePair* pose, RefRestore rr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *ray_visible_out = ctr[kCtrVisible];
    post_track_at(DT, good, T_last, pose_out, tracked, gate);
    PosePair p;
    pose_pair_from(T_last, p);
    *pose = p;
    ctr[kCtrVisible] = 0;
  }
  if (*good) return;
  const int stride = gridDim.x * blockDim.x;
  for (int c = 0; c < rr.count; ++c)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < rr.n[c]; i += stride) rr.dst[c][i] = rr.src[c][i];
}

__global__ void k_record_first(const double* T_last, double* pose_out, int* tracked) {
  if (threadIdx.x != 0) return;
  record_first_at(T_last, pose_out, tracked);
}

__global__ void k_set_identity(double* T) {
  if (threadIdx.x < 16) T[threadIdx.x] = (threadIdx.x % 5 == 0) ? 1.0 : 0.0;
}

}  // namespace

// Events of a timed frame: its start and end (ev[0], ev[kStages]); on the
// frames sampled for detail also the stage boundaries in between and a
// start/stop pair recorded by each single-kernel stage's own launch.
constexpr int kStages = 8;  // conversion, visibility+z-range, march, tracker inputs, solve, alloc, visible, integrate
constexpr int kTimedKernels = 4;  // k_march, k_alloc, k_visible, k_integrate
constexpr int kEvents = kStages + 1 + 2 * kTimedKernels;

struct yas_pipeline {
  yas_ctx* ctx;
  bool timing;
  std::vector<std::vector<hipEvent_t>> pending;  // kEvents events per timed frame
  std::vector<bool> pending_detail;              // whether that frame was sampled for detail
  std::vector<std::vector<hipEvent_t>> spare;
  int detail_every;        // every n-th timed frame records the stage and kernel events
  long long timed_count;   // frames timed since timing was enabled
  int sampled;             // frames whose stage and kernel times are in stage_ms / kernel_ms
  double stage_ms[kStages];
  double kernel_ms[kTimedKernels];
  std::vector<double> frame_ms;
  int* frame_stats;  // [cap][8]: allocated, visible (fuse), flags, walk steps, visible (ray cast)
  int W, H;
  float depth_scale;
  int gray_mode;
  yas_map* map;
  yas_raycaster* rc;
  yas_tracker* trk;
  float* depth;
  float* gray;
  double* T_last;
  double* poses;
  int* tracked;
  int* gate;
  int frames;
  int cap;
};

static int grow(yas_pipeline* p) {
  int ncap = p->cap ? p->cap * 2 : 1024;
  double* np = nullptr;
  int* nt = nullptr;
  YAS_HIP(hipMalloc((void**)&np, sizeof(double) * 16 * (size_t)ncap));
  YAS_HIP(hipMalloc((void**)&nt, sizeof(int) * (size_t)ncap));
  if (p->cap) {
    YAS_HIP(hipMemcpyAsync(np, p->poses, sizeof(double) * 16 * p->cap, hipMemcpyDeviceToDevice, p->ctx->stream));
    YAS_HIP(hipMemcpyAsync(nt, p->tracked, sizeof(int) * p->cap, hipMemcpyDeviceToDevice, p->ctx->stream));
    YAS_HIP(hipStreamSynchronize(p->ctx->stream));
    hipFree(p->poses);
    hipFree(p->tracked);
  }
  int* ns = nullptr;
  YAS_HIP(hipMalloc((void**)&ns, sizeof(int) * 8 * (size_t)ncap));
  YAS_HIP(hipMemsetAsync(ns, 0, sizeof(int) * 8 * (size_t)ncap, p->ctx->stream));
  if (p->cap) {
    YAS_HIP(hipMemcpyAsync(ns, p->frame_stats, sizeof(int) * 8 * p->cap, hipMemcpyDeviceToDevice, p->ctx->stream));
    YAS_HIP(hipStreamSynchronize(p->ctx->stream));
    hipFree(p->frame_stats);
  }
  p->frame_stats = ns;
  p->poses = np;
  p->tracked = nt;
  p->cap = ncap;
  return YAS_OK;
}

int yas_tracker_fill(yas_tracker* t, StreamDev& d);
int yas_raycaster_fill(yas_raycaster* rc, const MapView& v, StreamDev& d);
const PosePair* yas_map_pose_dev(yas_map* map);

// The device description of this pipeline's stream (serving engine).
int yas_pipeline_fill(yas_pipeline* p, StreamDev& d) {
  std::memset(&d, 0, sizeof(d));
  d.W = p->W;
  d.H = p->H;
  d.depth_scale_inv = (float)(1.0 / (double)p->depth_scale);
  d.gray_mode = p->gray_mode;
  d.K = p->map->K;
  d.m = yas_map_view(p->map);
  d.pose = (PosePair*)yas_map_pose_dev(p->map);
  long nb = p->map->P.num_blocks;
  d.vis_groups = (int)std::min<long>(std::max<long>((nb + 255) / 256, 1), 4096);
  d.int_groups = (int)std::min<long>(std::max<long>((nb + 3) / 4, 1), 2048);
  d.depth = p->depth;
  d.gray = p->gray;
  d.T_last = p->T_last;
  d.gate = p->gate;
  int rc = yas_raycaster_fill(p->rc, d.m, d);
  if (!rc) rc = yas_tracker_fill(p->trk, d);
  // the batched tracker reads the level-0 model vertices straight from the
  // ray cast instead of a copy (the single pipeline swaps its buffers, so it
  // keeps the copy)
  d.tb[0].ref_v = d.vmap;
  return rc;
}
yas_map* yas_pipeline_map_of(yas_pipeline* p) { return p->map; }
yas_tracker* yas_pipeline_tracker_of(yas_pipeline* p) { return p->trk; }

extern "C" {

int yas_pipeline_create(yas_ctx* ctx, int W, int H, const float K[4], const yas_map_params* mp,
                        const yas_tracker_params* tp, float depth_scale, int gray_mode, yas_pipeline** out) {
  YAS_REQUIRE(ctx && K && mp && tp && out, "yas_pipeline_create: null argument");
  YAS_REQUIRE(depth_scale > 0.f, "yas_pipeline_create: depth_scale must be > 0");
  YAS_HIP(hipSetDevice(ctx->device));
  yas_pipeline* p = new yas_pipeline();
  p->ctx = yas_ctx_retain(ctx);
  p->timing = false;
  p->detail_every = 1;
  p->timed_count = 0;
  p->sampled = 0;
  std::memset(p->stage_ms, 0, sizeof(p->stage_ms));
  std::memset(p->kernel_ms, 0, sizeof(p->kernel_ms));
  p->map = nullptr; p->rc = nullptr; p->trk = nullptr;
  p->depth = p->gray = nullptr;
  p->T_last = p->poses = nullptr;
  p->tracked = p->gate = p->frame_stats = nullptr;
  p->frames = p->cap = 0;
  p->W = W;
  p->H = H;
  p->depth_scale = depth_scale;
  p->gray_mode = gray_mode;
  double Kd[4] = {K[0], K[1], K[2], K[3]};
  int rc = yas_map_create(ctx, mp, K, &p->map);
  if (!rc) rc = yas_raycaster_create(ctx, W, H, K, &p->rc);
  if (!rc) rc = yas_tracker_create(ctx, W, H, Kd, tp, &p->trk);
  if (rc) {
    yas_pipeline_destroy(p);
    return rc;
  }
  hipError_t e = hipMalloc((void**)&p->depth, sizeof(float) * W * H);
  if (e == hipSuccess) e = hipMalloc((void**)&p->gray, sizeof(float) * W * H);
  if (e == hipSuccess) e = hipMalloc((void**)&p->T_last, sizeof(double) * 16);
  if (e == hipSuccess) e = hipMalloc((void**)&p->gate, sizeof(int));
  if (e != hipSuccess) {
    yas_set_error("yas_pipeline_create: %s", hipGetErrorString(e));
    yas_pipeline_destroy(p);
    return YAS_E_HIP;
  }
  if ((rc = grow(p))) {
    yas_pipeline_destroy(p);
    return rc;
  }
  hipLaunchKernelGGL(k_set_identity, dim3(1), dim3(64), 0, ctx->stream, p->T_last);
  YAS_LAUNCH_CHECK();
  *out = p;
  return YAS_OK;
}

int yas_pipeline_destroy(yas_pipeline* p) {
  if (!p) return YAS_OK;
  hipSetDevice(p->ctx->device);
  hipStreamSynchronize(p->ctx->stream);
  yas_map_destroy(p->map);
  yas_raycaster_destroy(p->rc);
  yas_tracker_destroy(p->trk);
  hipFree(p->depth);
  hipFree(p->gray);
  hipFree(p->T_last);
  hipFree(p->poses);
  hipFree(p->tracked);
  hipFree(p->gate);
  hipFree(p->frame_stats);
  for (auto* v : {&p->pending, &p->spare})
    for (auto& ev : *v)
      for (hipEvent_t e : ev) hipEventDestroy(e);
  yas_ctx_release(p->ctx);
  delete p;
  return YAS_OK;
}

int yas_pipeline_step(yas_pipeline* p, const uint16_t* depth, size_t dpitch, const uint8_t* rgb, size_t cpitch) {
  YAS_REQUIRE(p && depth && rgb, "yas_pipeline_step: null argument");
  YAS_HIP(hipSetDevice(p->ctx->device));
  int rc;
  if (p->frames >= p->cap && (rc = grow(p))) return rc;
  const int W = p->W, H = p->H;
  const size_t fp = (size_t)W * sizeof(float);
  hipStream_t s = p->ctx->stream;
  // events of this frame (timed frames after the first only)
  hipEvent_t ev[kEvents] = {};
  const bool timed = p->timing && p->frames > 0;
  bool detail = false;
  if (timed) {
    detail = p->timed_count % p->detail_every == 0;
    p->timed_count++;
    std::vector<hipEvent_t> set;
    if (!p->spare.empty()) {
      set = std::move(p->spare.back());
      p->spare.pop_back();
    } else {
      set.resize(kEvents);
      for (auto& e : set) YAS_HIP(hipEventCreate(&e));
    }
    for (int i = 0; i < kEvents; ++i) ev[i] = set[i];
    p->pending.push_back(std::move(set));
    p->pending_detail.push_back(detail);
  }
  // start/stop pairs of the kernels timed by their own launch
  const hipEvent_t* kev = detail ? ev + kStages + 1 : nullptr;
  auto kpair = [&](int k) -> const hipEvent_t* { return kev ? kev + 2 * k : nullptr; };
  auto mark = [&](int i) -> int {
    if (timed && (detail || i == 0 || i == kStages)) YAS_HIP(hipEventRecord(ev[i], s));
    return YAS_OK;
  };
  if ((rc = mark(0))) return rc;
  {
    FrameBegin a;
    a.depth_in = depth;
    a.dpitch = dpitch;
    a.rgb = rgb;
    a.cpitch = cpitch;
    a.depth = p->depth;
    a.gray = p->gray;
    a.fpitch = fp;
    a.scale = (float)(1.0 / (double)p->depth_scale);
    a.W = W;
    a.H = H;
    a.gray_mode = p->gray_mode;
    a.T_last = p->T_last;
    a.pose = yas_map_pose_dev_mut(p->map);
    a.ctr = yas_map_view(p->map).ctr;
    a.st = yas_tracker_state_dev(p->trk);
    yas_raycaster_zrange_dev(p->rc, &a.zmin, &a.zmax, &a.ncell);
    hipLaunchKernelGGL(k_frame_begin, dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, s, a);
    YAS_LAUNCH_CHECK();
  }
  if ((rc = mark(1))) return rc;
  double* pose_out = p->poses + 16 * (size_t)p->frames;
  int* tracked_out = p->tracked + p->frames;
  const int* gate = nullptr;
  if (p->frames == 0) {
    // StereoInitialization (Tracking.cpp:139-155); the map's pose pair is
    // already that of T_last
    if ((rc = yas_tracker_set_reference_image(p->trk, p->gray, fp))) return rc;
    if ((rc = yas_tracker_set_reference_depth(p->trk, p->depth, fp))) return rc;
    hipLaunchKernelGGL(k_record_first, dim3(1), dim3(64), 0, s, (const double*)p->T_last, pose_out, tracked_out);
    YAS_LAUNCH_CHECK();
  } else {
    // TrackRGBD (Tracking.cpp:162-197)
    if ((rc = yas_raycast_prepared(p->rc, p->map, detail ? ev[2] : nullptr, kpair(0)))) return rc;
    if ((rc = mark(3))) return rc;
    if ((rc = yas_tracker_prepare_frame_dev(p->trk, (const float*)yas_raycaster_vmap_dev(p->rc), (size_t)W * 16,
                                            p->gray, fp, p->depth, fp)))
      return rc;
    if ((rc = mark(4))) return rc;
    if ((rc = yas_tracker_solve_async(p->trk))) return rc;
    yas_tracker_swap_frame_buffer(p->trk);
    RefRestore rr;
    yas_tracker_unswap_copies(p->trk, rr);
    hipLaunchKernelGGL(k_post_track, dim3(32), dim3(256), 0, s, yas_tracker_result_dev(p->trk),
                       yas_tracker_good_dev(p->trk), p->T_last, pose_out, tracked_out, p->gate,
                       yas_map_view(p->map).ctr, p->frame_stats + 8 * (size_t)p->frames + 4,
                       yas_map_pose_dev_mut(p->map), rr);
    YAS_LAUNCH_CHECK();
    gate = p->gate;
  }
  if ((rc = mark(5))) return rc;
  if ((rc = yas_map_alloc_stage(p->map, p->depth, fp, W, H, gate, kpair(1)))) return rc;
  if ((rc = mark(6))) return rc;
  if ((rc = yas_map_visible_stage(p->map, W, H, gate, kpair(2)))) return rc;
  if ((rc = mark(7))) return rc;
  if ((rc = yas_map_integrate_stage(p->map, p->depth, fp, W, H, kpair(3), p->frame_stats + 8 * (size_t)p->frames)))
    return rc;
  if ((rc = mark(8))) return rc;
  p->frames++;
  return YAS_OK;
}

int yas_pipeline_pose(yas_pipeline* p, int frame, double T[16], int* tracked) {
  YAS_REQUIRE(p && T, "null argument");
  YAS_REQUIRE(frame >= 0 && frame < p->frames, "yas_pipeline_pose: no such frame");
  YAS_HIP(hipSetDevice(p->ctx->device));
  YAS_HIP(hipStreamSynchronize(p->ctx->stream));
  YAS_HIP(hipMemcpy(T, p->poses + 16 * (size_t)frame, sizeof(double) * 16, hipMemcpyDeviceToHost));
  if (tracked) YAS_HIP(hipMemcpy(tracked, p->tracked + frame, sizeof(int), hipMemcpyDeviceToHost));
  return YAS_OK;
}

int yas_pipeline_enable_timing(yas_pipeline* p, int detail_every) {
  YAS_REQUIRE(p && detail_every >= 0, "yas_pipeline_enable_timing: bad argument");
  p->timing = detail_every != 0;
  if (detail_every) p->detail_every = detail_every;
  p->timed_count = 0;
  return YAS_OK;
}

int yas_pipeline_timing(yas_pipeline* p, double stage_ms[8], double* frame_ms, int capacity, int* num_frames) {
  YAS_REQUIRE(p, "null pipeline");
  YAS_HIP(hipStreamSynchronize(p->ctx->stream));
  for (size_t f = 0; f < p->pending.size(); ++f) {
    auto& set = p->pending[f];
    if (p->pending_detail[f]) {
      for (int i = 0; i < kStages; ++i) {
        float ms = 0.0f;
        YAS_HIP(hipEventElapsedTime(&ms, set[i], set[i + 1]));
        p->stage_ms[i] += ms;
      }
      for (int k = 0; k < kTimedKernels; ++k) {
        float ms = 0.0f;
        YAS_HIP(hipEventElapsedTime(&ms, set[kStages + 1 + 2 * k], set[kStages + 2 + 2 * k]));
        p->kernel_ms[k] += ms;
      }
      p->sampled++;
    }
    float whole = 0.0f;
    YAS_HIP(hipEventElapsedTime(&whole, set[0], set[kStages]));
    p->frame_ms.push_back(whole);
    p->spare.push_back(std::move(set));
  }
  p->pending.clear();
  p->pending_detail.clear();
  if (stage_ms) std::memcpy(stage_ms, p->stage_ms, sizeof(p->stage_ms));
  int n = (int)p->frame_ms.size();
  if (frame_ms)
    for (int i = 0; i < n && i < capacity; ++i) frame_ms[i] = p->frame_ms[i];
  if (num_frames) *num_frames = n;
  return YAS_OK;
}

int yas_pipeline_reset_timing(yas_pipeline* p) {
  YAS_REQUIRE(p, "null pipeline");
  double tmp[8];
  int rc = yas_pipeline_timing(p, tmp, nullptr, 0, nullptr);
  if (rc) return rc;
  std::memset(p->stage_ms, 0, sizeof(p->stage_ms));
  std::memset(p->kernel_ms, 0, sizeof(p->kernel_ms));
  p->sampled = 0;
  p->frame_ms.clear();
  return YAS_OK;
}

int yas_pipeline_kernel_timing(yas_pipeline* p, double kernel_ms[4], int* frames_sampled) {
  YAS_REQUIRE(p && kernel_ms, "null argument");
  int rc = yas_pipeline_timing(p, nullptr, nullptr, 0, nullptr);
  if (rc) return rc;
  std::memcpy(kernel_ms, p->kernel_ms, sizeof(p->kernel_ms));
  if (frames_sampled) *frames_sampled = p->sampled;
  return YAS_OK;
}

int yas_pipeline_frame_stats(yas_pipeline* p, int first, int count, int* out) {
  YAS_REQUIRE(p && out, "null argument");
  YAS_REQUIRE(first >= 0 && count >= 0 && first + count <= p->frames, "yas_pipeline_frame_stats: bad frame range");
  YAS_HIP(hipStreamSynchronize(p->ctx->stream));
  if (count) YAS_HIP(hipMemcpy(out, p->frame_stats + 8 * (size_t)first, sizeof(int) * 8 * count, hipMemcpyDeviceToHost));
  return YAS_OK;
}

int yas_pipeline_num_frames(yas_pipeline* p, int* n) {
  YAS_REQUIRE(p && n, "null argument");
  *n = p->frames;
  return YAS_OK;
}
int yas_pipeline_map(yas_pipeline* p, yas_map** m) {
  YAS_REQUIRE(p && m, "null argument");
  *m = p->map;
  return YAS_OK;
}
int yas_pipeline_raycaster(yas_pipeline* p, yas_raycaster** r) {
  YAS_REQUIRE(p && r, "null argument");
  *r = p->rc;
  return YAS_OK;
}
int yas_pipeline_tracker(yas_pipeline* p, yas_tracker** t) {
  YAS_REQUIRE(p && t, "null argument");
  *t = p->trk;
  return YAS_OK;
}

}  // extern "C"
