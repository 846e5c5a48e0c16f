// Dense RGB-D tracker: the reference's RGBDTracking
// (RGBDFusion/src/RGBDTracking.cu, RGBDFusion/include/TrackingUtils.h) and the
// image operations it uses (RGBDFusion/src/ImageProc.cu).
//
// MI355X design: the whole Gauss-Newton loop stays on the device, with no
// host round trip and no grid-wide synchronisation inside a kernel (on the
// multi-XCD gfx950 an agent-scope fence writes back the L2, so kernel
// boundaries are the cheaper barrier).  One iteration is:
//   pass 1   grid-stride over the level's pixels: ICP point-to-plane rows
//            (28 upper-triangle products + count) and photometric residuals
//            with per-thread Welford statistics; each workgroup reduces its
//            values (transpose butterfly across the 64 lanes, LDS across
//            waves) into one row of partials.
//   pass 2   (photometric) every workgroup folds pass 1's statistics rows in
//            the same fixed order, so all agree on mean, sigma and the Huber
//            threshold, then reduces the weighted photometric rows.
//   update   one workgroup folds the partial rows in a fixed order, solves
//            the 6x6 system in FP64 (the reference's Eigen LDLT, restated),
//            applies T <- exp(xi) T in FP64 and keeps the best pose.
// A full solve takes each update but the last at the start of the next
// pass 1 instead: every workgroup of that pass folds the same rows and takes
// the same step, one launch fewer per iteration with the same values.
// Results are deterministic run to run.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "yas_internal.hpp"
#include "yas_tracker_dev.hpp"
#include "yas_server_dev.hpp"

using namespace yas;

namespace {

__global__ void __launch_bounds__(256) k_copy_f32(const float* src, size_t sp, float* dst, int W, int H) {
  copy_f32_at(src, sp, dst, W, H, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_rgb_to_gray(const uint8_t* rgb, size_t sp, float* dst, size_t dp, int W,
                                                     int H, int mode) {
  rgb_to_gray_at(rgb, sp, dst, dp, W, H, mode, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_pyrdown_gauss(const float* src, size_t sp, int sw, int sh, float* dst,
                                                       size_t dp, int dw, int dh) {
  pyrdown_gauss_at(src, sp, sw, sh, dst, dp, dw, dh, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_gradient(const float* s, size_t sp, float* gx, float* gy, size_t dp, int W,
                                                  int H) {
  gradient_at(s, sp, gx, gy, dp, W, H, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_invdepth(const float* z, size_t sp, float* o, size_t dp, int W, int H) {
  invdepth_at(z, sp, o, dp, W, H, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_pyrdown_depth(const float* s, size_t sp, float* o, size_t dp, int W, int H) {
  pyrdown_depth_at(s, sp, o, dp, W, H, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_pyrdown_vmap(const float4* s, size_t sp, float4* o, size_t dp, int W, int H) {
  pyrdown_vmap_at(s, sp, o, dp, W, H, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_copy_f4(const float4* s, size_t sp, float4* o, int W, int H) {
  copy_f4_at(s, sp, o, W, H, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_vertex(const float* inv, size_t sp, float4* v, size_t dp, int W, int H,
                                                float invfx, float invfy, float cx, float cy, float cut) {
  vertex_at(inv, sp, v, dp, W, H, invfx, invfy, cx, cy, cut, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_normal(const float4* v, size_t sp, float4* n, size_t dp, int W, int H) {
  normal_at(v, sp, n, dp, W, H, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_icp_point(const float4* v, float4* o, int W, int H) {
  icp_point_at(v, (size_t)W * 16, o, W, H, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_vmap_to_depth(const float4* v, size_t sp, float* d, size_t dp, int W, int H) {
  vmap_to_depth_at(v, sp, d, dp, W, H, blockIdx.x, blockIdx.y);
}

template <bool ICP, bool RGB>
__global__ void __launch_bounds__(256) k_pass1(Level L, TrackBufs B, const TrackState* st, float* part_icp,
                                               float* part_res, float angle_th, float dist_th) {
  pass1_at<ICP, RGB, true>(L, B, st, part_icp, part_res, angle_th, dist_th, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) k_pass2(Level L, TrackBufs B, TrackState* st, const float* part_res, int g1,
                                               float* part_rgb) {
  pass2_at<true>(L, B, st, part_res, g1, part_rgb, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(256) k_update(TrackState* st, const float* part_icp, int gi, const float* part_rgb,
                                                int gr, int mode, int first_of_level, int do_step, int record,
                                                float wrh, float wrb, float wih, float wib) {
  update_at(st, part_icp, gi, part_rgb, gr, mode, first_of_level, do_step, record, wrh, wrb, wih, wib);
}
// The last step of a solve, from the state buffer the last pass used into the tracker's state.
__global__ void __launch_bounds__(256) k_update_into(const TrackState* st_in, TrackState* st_out, StepArgs a) {
  gn_update(st_in, st_out, true, a.part_icp, a.gi, a.part_rgb, a.gr, a.mode, a.first_of_level, 1, a.record, a.wrh,
            a.wrb, a.wih, a.wib, nullptr);
}
// Step of the previous linearisation, then pass 1 at the new pose.
template <bool ICP, bool RGB>
__global__ void __launch_bounds__(256) k_step_pass1(Level L, TrackBufs B, StepArgs a, float* part_icp,
                                                    float* part_res, float angle_th, float dist_th) {
  pass1_at<ICP, RGB, true, true>(L, B, nullptr, part_icp, part_res, angle_th, dist_th, blockIdx.x, gridDim.x, &a);
}
// A frame's tracker inputs in levels + 1 launches instead of seven per
// level: the pointwise products of level l from level l - 1 (or from the
// frame at level 0): intensity, inverse depth and vertices of the current
// frame, and the model's vertices; and the neighbourhood products of level
// l: gradients and both ICP point maps (vertex and normal), which need the
// products of level l but nothing of level l + 1, so they share a launch
// with those.  Each pixel gets the value the separate kernels give it.
struct FrameLevel {
  int l, w, h, pw, ph;  // level, its size, the size of level l - 1
  float invfx, invfy, cx, cy;
  const float* gray;   // level 0 sources
  size_t gp;
  const float* depth;
  size_t dp;
  const float4* vmap;
  size_t vp;
  const float* int_prev;  // level l - 1
  const float* inv_prev;
  const float4* refv_prev;
  float* cur_int;  // level l
  float* cur_inv;
  float4* cur_v;
  float4* ref_v;
  float* gx;
  float* gy;
  float4* cur_p;  // ICP point maps
  float4* ref_p;
};
__device__ inline void frame_level_px(const FrameLevel& a, int x, int y) {
  if (x >= a.w || y >= a.h) return;
  const size_t k = (size_t)y * a.w + x;
  float in, iv;
  float4 rv;
  if (a.l == 0) {
    in = ld(a.gray, a.gp, x, y);
    iv = invdepth_of(ld(a.depth, a.dp, x, y));
    rv = ((const float4*)((const char*)a.vmap + (size_t)y * a.vp))[x];
  } else {
    in = pyrdown_gauss_px(a.int_prev, (size_t)a.pw * 4, a.pw, a.ph, x, y);
    iv = pyrdown_depth_px(a.inv_prev, (size_t)a.pw * 4, a.w, a.h, x, y);
    rv = pyrdown_vmap_px(a.refv_prev, (size_t)a.pw * 16, a.w, a.h, x, y);
  }
  a.cur_int[k] = in;
  a.cur_inv[k] = iv;
  a.cur_v[k] = vertex_of(iv, x, y, a.invfx, a.invfy, a.cx, a.cy, 3.0f);
  a.ref_v[k] = rv;
}
__device__ inline void frame_level_nb_px(const FrameLevel& a, int x, int y) {
  if (x >= a.w || y >= a.h) return;
  const size_t k = (size_t)y * a.w + x;
  float gxv, gyv;
  gradient_px(a.cur_int, (size_t)a.w * 4, a.w, a.h, x, y, gxv, gyv);
  a.gx[k] = gxv;
  a.gy[k] = gyv;
  store_icp_point(a.cur_p, a.w * a.h, (int)k, icp_point_px(a.cur_v, (size_t)a.w * 16, a.w, a.h, x, y));
  store_icp_point(a.ref_p, a.w * a.h, (int)k, icp_point_px(a.ref_v, (size_t)a.w * 16, a.w, a.h, x, y));
}
__global__ void __launch_bounds__(256) k_frame_level(FrameLevel a) {
  frame_level_px(a, blockIdx.x * 64 + (threadIdx.x & 63), blockIdx.y * 4 + (threadIdx.x >> 6));
}
__global__ void __launch_bounds__(256) k_frame_level_nb(FrameLevel a) {
  frame_level_nb_px(a, blockIdx.x * 64 + (threadIdx.x & 63), blockIdx.y * 4 + (threadIdx.x >> 6));
}
// Level l's pointwise products and level l - 1's neighbourhood products,
// which do not depend on each other, in one launch: the first `nprod`
// workgroups take the products' 64x4 tiles, the others the neighbourhood
// products'.
__global__ void __launch_bounds__(256) k_frame_level_pair(FrameLevel prod, FrameLevel nb, int nprod) {
  const int b = blockIdx.x;
  if (b < nprod) {
    const int tx = (prod.w + 63) >> 6;
    frame_level_px(prod, (b % tx) * 64 + (threadIdx.x & 63), (b / tx) * 4 + (threadIdx.x >> 6));
  } else {
    const int tx = (nb.w + 63) >> 6, c = b - nprod;
    frame_level_nb_px(nb, (c % tx) * 64 + (threadIdx.x & 63), (c / tx) * 4 + (threadIdx.x >> 6));
  }
}
__global__ void k_track_init(TrackState* st, Mat44 T) { track_init_at(st, T.m); }
__global__ void k_track_init_identity(TrackState* st) { track_init_identity_at(st); }

inline dim3 grid64x4(int W, int H) { return dim3((W + 63) / 64, (H + 3) / 4); }

}  // namespace

// ------------------------------------------------------------------ host
struct yas_tracker {
  yas_ctx* ctx;
  int W, H;
  yas_tracker_params P;
  Level lv[kMaxLevels];
  float *cur_int[kMaxLevels], *ref_int[kMaxLevels], *gx[kMaxLevels], *gy[kMaxLevels];
  float *cur_inv[kMaxLevels], *ref_inv[kMaxLevels];
  float4 *cur_v[kMaxLevels], *ref_v[kMaxLevels];
  float4 *cur_p[kMaxLevels], *ref_p[kMaxLevels];  // ICP point maps (what the ICP term reads)
  float4 *cur_n[kMaxLevels], *ref_n[kMaxLevels];    // normal maps, formed on request (yas_tracker_level)
  float3* resid[kMaxLevels];
  float* part_icp;  // [kRedGroups][32]
  float* part_icp2;  // the other of the pair a solve alternates between
  float* part_res;  // [kRedGroups][4]
  float* part_rgb;  // [kRedGroups][32]
  TrackState* st;
  TrackState* st2;  // a solve's steps alternate between st and st2
  TrackState* host_st;
  float angle_th, dist_th;
  float* raw_depth;  // depth uploaded by the host-image setters (W*H, lazily allocated)
  float* raw_image;  // intensity uploaded by the host-image setters
};

static int image_pyramid(yas_tracker* t, const float* g, size_t pitch, float** pyr, bool grads) {
  hipStream_t s = t->ctx->stream;
  hipLaunchKernelGGL(k_copy_f32, grid64x4(t->W, t->H), dim3(256), 0, s, g, pitch, pyr[0], t->W, t->H);
  for (int l = 1; l < t->P.levels; ++l) {
    const Level& a = t->lv[l - 1];
    const Level& b = t->lv[l];
    hipLaunchKernelGGL(k_pyrdown_gauss, dim3((b.w + 15) / 16, (b.h + 15) / 16), dim3(256), 0, s, pyr[l - 1],
                       (size_t)a.w * 4, a.w, a.h, pyr[l], (size_t)b.w * 4, b.w, b.h);
  }
  if (grads)
    for (int l = 0; l < t->P.levels; ++l) {
      const Level& a = t->lv[l];
      hipLaunchKernelGGL(k_gradient, grid64x4(a.w, a.h), dim3(256), 0, s, pyr[l], (size_t)a.w * 4, t->gx[l], t->gy[l],
                         (size_t)a.w * 4, a.w, a.h);
    }
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

static int depth_pyramid(yas_tracker* t, const float* d, size_t pitch, float** inv, float4** v, float4** pt) {
  hipStream_t s = t->ctx->stream;
  for (int l = 0; l < t->P.levels; ++l) {
    const Level& a = t->lv[l];
    if (l == 0)
      hipLaunchKernelGGL(k_invdepth, grid64x4(a.w, a.h), dim3(256), 0, s, d, pitch, inv[0], (size_t)a.w * 4, a.w, a.h);
    else
      hipLaunchKernelGGL(k_pyrdown_depth, grid64x4(a.w, a.h), dim3(256), 0, s, inv[l - 1], (size_t)t->lv[l - 1].w * 4,
                         inv[l], (size_t)a.w * 4, a.w, a.h);
    hipLaunchKernelGGL(k_vertex, grid64x4(a.w, a.h), dim3(256), 0, s, inv[l], (size_t)a.w * 4, v[l], (size_t)a.w * 16,
                       a.w, a.h, a.invfx, a.invfy, a.cx, a.cy, 3.0f);
    hipLaunchKernelGGL(k_icp_point, grid64x4(a.w, a.h), dim3(256), 0, s, v[l], pt[l], a.w, a.h);
  }
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

static int model_pyramid(yas_tracker* t) {
  hipStream_t s = t->ctx->stream;
  for (int l = 0; l < t->P.levels; ++l) {
    const Level& a = t->lv[l];
    if (l > 0)
      hipLaunchKernelGGL(k_pyrdown_vmap, grid64x4(a.w, a.h), dim3(256), 0, s, t->ref_v[l - 1],
                         (size_t)t->lv[l - 1].w * 16, t->ref_v[l], (size_t)a.w * 16, a.w, a.h);
    hipLaunchKernelGGL(k_icp_point, grid64x4(a.w, a.h), dim3(256), 0, s, t->ref_v[l], t->ref_p[l], a.w, a.h);
  }
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

static int groups_for(const Level& a) {
  int n = (a.w * a.h + 1023) / 1024;  // about 4 pixels per thread
  return n < 1 ? 1 : (n > kRedGroups ? kRedGroups : n);
}

static TrackBufs bufs(yas_tracker* t, int l) {
  TrackBufs B;
  B.ref_v = t->ref_v[l]; B.ref_p = t->ref_p[l]; B.cur_v = t->cur_v[l]; B.cur_p = t->cur_p[l];
  B.ref_inv = t->ref_inv[l]; B.ref_int = t->ref_int[l]; B.cur_int = t->cur_int[l];
  B.gx = t->gx[l]; B.gy = t->gy[l]; B.resid = t->resid[l];
  return B;
}

// One linearisation at level l (icp / rgb terms as requested) followed by the
// fold and, when do_step, one Gauss-Newton step.
static int launch_iteration(yas_tracker* t, int l, bool icp, bool rgb, bool do_step, bool first_of_level) {
  const Level& a = t->lv[l];
  hipStream_t s = t->ctx->stream;
  const int G = groups_for(a);
  TrackBufs B = bufs(t, l);
  if (icp && rgb)
    hipLaunchKernelGGL((k_pass1<true, true>), dim3(G), dim3(256), 0, s, a, B, (const TrackState*)t->st, t->part_icp,
                       t->part_res, t->angle_th, t->dist_th);
  else if (icp)
    hipLaunchKernelGGL((k_pass1<true, false>), dim3(G), dim3(256), 0, s, a, B, (const TrackState*)t->st, t->part_icp,
                       t->part_res, t->angle_th, t->dist_th);
  else
    hipLaunchKernelGGL((k_pass1<false, true>), dim3(G), dim3(256), 0, s, a, B, (const TrackState*)t->st, t->part_icp,
                       t->part_res, t->angle_th, t->dist_th);
  if (rgb)
    hipLaunchKernelGGL(k_pass2, dim3(G), dim3(256), 0, s, a, B, t->st, (const float*)t->part_res, G, t->part_rgb);
  hipLaunchKernelGGL(k_update, dim3(1), dim3(256), 0, s, t->st, (const float*)t->part_icp, icp ? G : 0,
                     (const float*)t->part_rgb, rgb ? G : 0, t->P.mode, first_of_level ? 1 : 0, do_step ? 1 : 0,
                     l == 0 ? 1 : 0, t->P.w_rgb_h, t->P.w_rgb_b, t->P.w_icp_h, t->P.w_icp_b);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

// Coarse-to-fine Gauss-Newton from the pose already in the device state:
// the same linearisations and steps as launch_iteration in sequence, with
// each step (but the last) taken at the start of the next pass 1, which
// saves one launch per iteration.  Step i reads the state buffer of step
// i - 1 and writes the other one, and the ICP partial rows alternate the
// same way, since that pass 1 reads the previous rows while writing its own.
int yas_tracker_solve_async(yas_tracker* t) {
  const bool icp = t->P.mode != YAS_TRACK_RGB, rgb = t->P.mode != YAS_TRACK_DEPTH;
  hipStream_t s = t->ctx->stream;
  TrackState* S[2] = {t->st, t->st2};
  float* PI[2] = {t->part_icp, t->part_icp2};
  StepArgs a = {};
  int i = 0;
  for (int l = t->P.levels - 1; l >= 0; --l)
    for (int it = 0; it < t->P.iters[l]; ++it, ++i) {
      const Level& lv = t->lv[l];
      const int G = groups_for(lv);
      const TrackBufs B = bufs(t, l);
      TrackState* cur = S[i & 1];
      float* pic = PI[i & 1];
      if (i == 0) {
        if (icp && rgb)
          hipLaunchKernelGGL((k_pass1<true, true>), dim3(G), dim3(256), 0, s, lv, B, (const TrackState*)cur, pic,
                             t->part_res, t->angle_th, t->dist_th);
        else if (icp)
          hipLaunchKernelGGL((k_pass1<true, false>), dim3(G), dim3(256), 0, s, lv, B, (const TrackState*)cur, pic,
                             t->part_res, t->angle_th, t->dist_th);
        else
          hipLaunchKernelGGL((k_pass1<false, true>), dim3(G), dim3(256), 0, s, lv, B, (const TrackState*)cur, pic,
                             t->part_res, t->angle_th, t->dist_th);
      } else {
        a.st_in = S[(i - 1) & 1];
        a.st_out = cur;
        if (icp && rgb)
          hipLaunchKernelGGL((k_step_pass1<true, true>), dim3(G), dim3(256), 0, s, lv, B, a, pic, t->part_res,
                             t->angle_th, t->dist_th);
        else if (icp)
          hipLaunchKernelGGL((k_step_pass1<true, false>), dim3(G), dim3(256), 0, s, lv, B, a, pic, t->part_res,
                             t->angle_th, t->dist_th);
        else
          hipLaunchKernelGGL((k_step_pass1<false, true>), dim3(G), dim3(256), 0, s, lv, B, a, pic, t->part_res,
                             t->angle_th, t->dist_th);
      }
      if (rgb)
        hipLaunchKernelGGL(k_pass2, dim3(G), dim3(256), 0, s, lv, B, cur, (const float*)t->part_res, G, t->part_rgb);
      // the step of this linearisation, taken by the next launch
      a.part_icp = pic;
      a.gi = icp ? G : 0;
      a.part_rgb = t->part_rgb;
      a.gr = rgb ? G : 0;
      a.mode = t->P.mode;
      a.first_of_level = it == 0 ? 1 : 0;
      a.record = l == 0 ? 1 : 0;
      a.wrh = t->P.w_rgb_h; a.wrb = t->P.w_rgb_b; a.wih = t->P.w_icp_h; a.wib = t->P.w_icp_b;
    }
  if (i > 0) hipLaunchKernelGGL(k_update_into, dim3(1), dim3(256), 0, s, (const TrackState*)S[(i - 1) & 1], t->st, a);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

int yas_tracker_init_identity_async(yas_tracker* t) {
  hipLaunchKernelGGL(k_track_init_identity, dim3(1), dim3(64), 0, t->ctx->stream, t->st);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

// Buffers and constants of this tracker, for the serving engine's batched
// kernels (which keep the buffer roles fixed and copy the current frame's
// pyramids into the reference ones instead of swapping them).
int yas_tracker_fill(yas_tracker* t, StreamDev& d) {
  YAS_REQUIRE(t->P.levels <= kTrackMaxLevels, "tracker levels");
  d.levels = t->P.levels;
  for (int l = 0; l < t->P.levels; ++l) {
    d.lv[l] = t->lv[l];
    d.tb[l] = bufs(t, l);
    d.cur_inv[l] = t->cur_inv[l];
    d.groups[l] = groups_for(t->lv[l]);
  }
  d.part_icp = t->part_icp;
  d.part_res = t->part_res;
  d.part_rgb = t->part_rgb;
  d.st = t->st;
  d.angle_th = t->angle_th;
  d.dist_th = t->dist_th;
  return YAS_OK;
}
const yas_tracker_params& yas_tracker_params_of(yas_tracker* t) { return t->P; }

// Device pointers of the solver's result (camera pose change and status).
const double* yas_tracker_result_dev(yas_tracker* t) { return t->st->last_ok; }
const int* yas_tracker_good_dev(yas_tracker* t) { return &t->st->good; }
// The state a solve starts from (the pipeline resets it in its first launch).
TrackState* yas_tracker_state_dev(yas_tracker* t) { return t->st; }

int yas_tracker_set_tracking_image_dev(yas_tracker* t, const float* g, size_t pitch) {
  return image_pyramid(t, g, pitch, t->cur_int, true);
}
int yas_tracker_set_tracking_depth_dev(yas_tracker* t, const float* d, size_t pitch) {
  return depth_pyramid(t, d, pitch, t->cur_inv, t->cur_v, t->cur_p);
}

// The pipeline's per-frame tracker inputs: the model from a ray-cast vertex
// map and the current frame from its intensity and depth, as
// set_reference_model + set_tracking_image + set_tracking_depth give them.
int yas_tracker_prepare_frame_dev(yas_tracker* t, const float* vmap, size_t vpitch, const float* g, size_t gpitch,
                                  const float* d, size_t dpitch) {
  YAS_REQUIRE(t && vmap && g && d && vpitch >= (size_t)t->W * 16, "yas_tracker_prepare_frame_dev: bad argument");
  hipStream_t s = t->ctx->stream;
  FrameLevel a[kMaxLevels];
  for (int l = 0; l < t->P.levels; ++l) {
    FrameLevel& f = a[l];
    const Level& v = t->lv[l];
    f.l = l;
    f.w = v.w;
    f.h = v.h;
    f.pw = l > 0 ? t->lv[l - 1].w : 0;
    f.ph = l > 0 ? t->lv[l - 1].h : 0;
    f.invfx = v.invfx; f.invfy = v.invfy; f.cx = v.cx; f.cy = v.cy;
    f.gray = g; f.gp = gpitch;
    f.depth = d; f.dp = dpitch;
    f.vmap = (const float4*)vmap; f.vp = vpitch;
    f.int_prev = l > 0 ? t->cur_int[l - 1] : nullptr;
    f.inv_prev = l > 0 ? t->cur_inv[l - 1] : nullptr;
    f.refv_prev = l > 0 ? t->ref_v[l - 1] : nullptr;
    f.cur_int = t->cur_int[l]; f.cur_inv = t->cur_inv[l]; f.cur_v = t->cur_v[l]; f.ref_v = t->ref_v[l];
    f.gx = t->gx[l]; f.gy = t->gy[l]; f.cur_p = t->cur_p[l]; f.ref_p = t->ref_p[l];
  }
  // level 0's products; then level l's products beside level l - 1's
  // neighbourhood products; then the coarsest level's neighbourhood products
  hipLaunchKernelGGL(k_frame_level, grid64x4(a[0].w, a[0].h), dim3(256), 0, s, a[0]);
  for (int l = 1; l < t->P.levels; ++l) {
    const dim3 gp = grid64x4(a[l].w, a[l].h), gn = grid64x4(a[l - 1].w, a[l - 1].h);
    const int nprod = (int)(gp.x * gp.y);
    hipLaunchKernelGGL(k_frame_level_pair, dim3(nprod + gn.x * gn.y), dim3(256), 0, s, a[l], a[l - 1], nprod);
  }
  hipLaunchKernelGGL(k_frame_level_nb, grid64x4(a[t->P.levels - 1].w, a[t->P.levels - 1].h), dim3(256), 0, s,
                     a[t->P.levels - 1]);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

int yas_rgb_to_gray_async(yas_ctx* ctx, const uint8_t* rgb, size_t sp, float* dst, size_t dp, int W, int H, int mode) {
  hipLaunchKernelGGL(k_rgb_to_gray, grid64x4(W, H), dim3(256), 0, ctx->stream, rgb, sp, dst, dp, W, H, mode);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

extern "C" {

int yas_tracker_create(yas_ctx* ctx, int W, int H, const double K[4], const yas_tracker_params* P, yas_tracker** out) {
  YAS_REQUIRE(ctx && K && P && out, "yas_tracker_create: null argument");
  YAS_REQUIRE(P->levels >= 1 && P->levels <= kMaxLevels, "yas_tracker_create: levels must be 1..5");
  YAS_REQUIRE(P->mode >= 0 && P->mode <= 2, "yas_tracker_create: unknown mode");
  YAS_REQUIRE(W > 0 && H > 0 && W % (1 << (P->levels - 1)) == 0 && H % (1 << (P->levels - 1)) == 0,
              "yas_tracker_create: width and height must be divisible by 2^(levels-1)");
  YAS_REQUIRE((W >> (P->levels - 1)) >= 4 && (H >> (P->levels - 1)) >= 4,
              "yas_tracker_create: the coarsest level must be at least 4x4");
  YAS_HIP(hipSetDevice(ctx->device));
  yas_tracker* t = new yas_tracker();
  std::memset(t, 0, sizeof(*t));
  t->ctx = yas_ctx_retain(ctx);
  t->W = W;
  t->H = H;
  t->P = *P;
  // RGBDTracking.cu:432-433: sin(20 deg) computed in float, 1 cm.
  t->angle_th = (float)std::sin(20.f * 3.14159254f / 180.f);
  t->dist_th = 0.01f;
  hipError_t e = hipSuccess;
  auto alloc = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  for (int l = 0; l < P->levels; ++l) {
    Level& a = t->lv[l];
    double s = (double)(1 << l);
    a.Kd[0] = K[0] / s; a.Kd[1] = K[1] / s; a.Kd[2] = K[2] / s; a.Kd[3] = K[3] / s;
    a.w = W >> l;
    a.h = H >> l;
    a.fx = (float)a.Kd[0]; a.fy = (float)a.Kd[1]; a.cx = (float)a.Kd[2]; a.cy = (float)a.Kd[3];
    a.invfx = (float)(1.0 / a.Kd[0]);
    a.invfy = (float)(1.0 / a.Kd[1]);
    a.Kinv[0] = 1.0 / a.Kd[0];
    a.Kinv[1] = 1.0 / a.Kd[1];
    a.Kinv[2] = -a.Kd[2] / a.Kd[0];
    a.Kinv[3] = -a.Kd[3] / a.Kd[1];
    size_t n = (size_t)a.w * a.h;
    for (float** p : {&t->cur_int[l], &t->ref_int[l], &t->gx[l], &t->gy[l], &t->cur_inv[l], &t->ref_inv[l]})
      alloc((void**)p, n * sizeof(float));
    for (float4** p : {&t->cur_v[l], &t->cur_n[l], &t->ref_v[l], &t->ref_n[l]}) alloc((void**)p, n * sizeof(float4));
    alloc((void**)&t->resid[l], n * sizeof(float3));
    for (float4** p : {&t->cur_p[l], &t->ref_p[l]}) alloc((void**)p, n * sizeof(IcpPoint));
  }
  alloc((void**)&t->part_icp, sizeof(float) * 32 * kRedGroups);
  alloc((void**)&t->part_icp2, sizeof(float) * 32 * kRedGroups);
  alloc((void**)&t->part_res, sizeof(float) * 4 * kRedGroups);
  alloc((void**)&t->part_rgb, sizeof(float) * 32 * kRedGroups);
  alloc((void**)&t->st, sizeof(TrackState));
  alloc((void**)&t->st2, sizeof(TrackState));
  if (e == hipSuccess) e = hipHostMalloc((void**)&t->host_st, sizeof(TrackState), hipHostMallocDefault);
  if (e != hipSuccess) {
    yas_set_error("yas_tracker_create: %s", hipGetErrorString(e));
    yas_tracker_destroy(t);
    return YAS_E_HIP;
  }
  // zeroed images, zeroed state (tickets armed)
  for (int l = 0; l < P->levels; ++l) {
    size_t n = (size_t)t->lv[l].w * t->lv[l].h;
    for (float* p : {t->cur_int[l], t->ref_int[l], t->gx[l], t->gy[l], t->cur_inv[l], t->ref_inv[l]})
      YAS_HIP(hipMemsetAsync(p, 0, n * sizeof(float), ctx->stream));
    for (float4* p : {t->cur_v[l], t->cur_n[l], t->ref_v[l], t->ref_n[l]})
      YAS_HIP(hipMemsetAsync(p, 0, n * sizeof(float4), ctx->stream));
    YAS_HIP(hipMemsetAsync(t->resid[l], 0, n * sizeof(float3), ctx->stream));
    for (float4* p : {t->cur_p[l], t->ref_p[l]}) YAS_HIP(hipMemsetAsync(p, 0, n * sizeof(IcpPoint), ctx->stream));
  }
  YAS_HIP(hipMemsetAsync(t->st, 0, sizeof(TrackState), ctx->stream));
  YAS_HIP(hipMemsetAsync(t->st2, 0, sizeof(TrackState), ctx->stream));
  int rc = yas_tracker_init_identity_async(t);
  if (rc) {
    yas_tracker_destroy(t);
    return rc;
  }
  *out = t;
  return YAS_OK;
}

int yas_tracker_destroy(yas_tracker* t) {
  if (!t) return YAS_OK;
  hipSetDevice(t->ctx->device);
  hipStreamSynchronize(t->ctx->stream);
  for (int l = 0; l < kMaxLevels; ++l) {
    for (float* p : {t->cur_int[l], t->ref_int[l], t->gx[l], t->gy[l], t->cur_inv[l], t->ref_inv[l]}) hipFree(p);
    for (float4* p : {t->cur_v[l], t->cur_n[l], t->ref_v[l], t->ref_n[l]}) hipFree(p);
    hipFree(t->resid[l]);
    for (float4* p : {t->cur_p[l], t->ref_p[l]}) hipFree(p);
  }
  hipFree(t->part_icp);
  hipFree(t->part_icp2);
  hipFree(t->part_res);
  hipFree(t->part_rgb);
  hipFree(t->st);
  hipFree(t->st2);
  hipFree(t->raw_depth);
  hipFree(t->raw_image);
  if (t->host_st) hipHostFree(t->host_st);
  yas_ctx_release(t->ctx);
  delete t;
  return YAS_OK;
}

int yas_tracker_set_reference_image(yas_tracker* t, const float* g, size_t pitch) {
  YAS_REQUIRE(t && g && pitch >= (size_t)t->W * 4, "yas_tracker_set_reference_image: bad argument");
  return image_pyramid(t, g, pitch, t->ref_int, false);
}
int yas_tracker_set_tracking_image(yas_tracker* t, const float* g, size_t pitch) {
  YAS_REQUIRE(t && g && pitch >= (size_t)t->W * 4, "yas_tracker_set_tracking_image: bad argument");
  return image_pyramid(t, g, pitch, t->cur_int, true);
}
int yas_tracker_set_reference_depth(yas_tracker* t, const float* d, size_t pitch) {
  YAS_REQUIRE(t && d && pitch >= (size_t)t->W * 4, "yas_tracker_set_reference_depth: bad argument");
  return depth_pyramid(t, d, pitch, t->ref_inv, t->ref_v, t->ref_p);
}
int yas_tracker_set_tracking_depth(yas_tracker* t, const float* d, size_t pitch) {
  YAS_REQUIRE(t && d && pitch >= (size_t)t->W * 4, "yas_tracker_set_tracking_depth: bad argument");
  return depth_pyramid(t, d, pitch, t->cur_inv, t->cur_v, t->cur_p);
}
// Host images (RGBDTracking.h:17-21 take cv::Mat): the reference converts an
// 8-bit gray image to float on the host and uploads it (RGBDTracking.cu:68-
// 71,104-107), and uploads float depth into one raw-depth buffer
// (mGpuBufferRawDepth, RGBDTracking.cu:86-87,122-123) that
// GetReferenceDepth returns.  The copy waits for the stream, so the host
// buffer may be reused when the call returns.
static int upload_host(yas_tracker* t, float** dst, const void* src, size_t pitch, bool gray_u8) {
  YAS_HIP(hipSetDevice(t->ctx->device));
  const size_t row = (size_t)t->W * sizeof(float);
  if (!*dst) YAS_HIP(hipMalloc((void**)dst, row * t->H));
  std::vector<float> conv;
  const void* from = src;
  size_t from_pitch = pitch;
  if (gray_u8) {
    conv.resize((size_t)t->W * t->H);
    for (int y = 0; y < t->H; ++y) {
      const uint8_t* r = (const uint8_t*)src + (size_t)y * pitch;
      for (int x = 0; x < t->W; ++x) conv[(size_t)y * t->W + x] = (float)r[x];
    }
    from = conv.data();
    from_pitch = row;
  }
  YAS_HIP(hipMemcpy2DAsync(*dst, row, from, from_pitch, row, t->H, hipMemcpyHostToDevice, t->ctx->stream));
  YAS_HIP(hipStreamSynchronize(t->ctx->stream));
  return YAS_OK;
}

int yas_tracker_set_reference_image_host(yas_tracker* t, const uint8_t* g, size_t pitch) {
  YAS_REQUIRE(t && g && pitch >= (size_t)t->W, "yas_tracker_set_reference_image_host: bad argument");
  int rc = upload_host(t, &t->raw_image, g, pitch, true);
  return rc ? rc : image_pyramid(t, t->raw_image, (size_t)t->W * 4, t->ref_int, false);
}
int yas_tracker_set_tracking_image_host(yas_tracker* t, const uint8_t* g, size_t pitch) {
  YAS_REQUIRE(t && g && pitch >= (size_t)t->W, "yas_tracker_set_tracking_image_host: bad argument");
  int rc = upload_host(t, &t->raw_image, g, pitch, true);
  return rc ? rc : image_pyramid(t, t->raw_image, (size_t)t->W * 4, t->cur_int, true);
}
int yas_tracker_set_reference_depth_host(yas_tracker* t, const float* d, size_t pitch) {
  YAS_REQUIRE(t && d && pitch >= (size_t)t->W * 4, "yas_tracker_set_reference_depth_host: bad argument");
  int rc = upload_host(t, &t->raw_depth, d, pitch, false);
  return rc ? rc : depth_pyramid(t, t->raw_depth, (size_t)t->W * 4, t->ref_inv, t->ref_v, t->ref_p);
}
int yas_tracker_set_tracking_depth_host(yas_tracker* t, const float* d, size_t pitch) {
  YAS_REQUIRE(t && d && pitch >= (size_t)t->W * 4, "yas_tracker_set_tracking_depth_host: bad argument");
  int rc = upload_host(t, &t->raw_depth, d, pitch, false);
  return rc ? rc : depth_pyramid(t, t->raw_depth, (size_t)t->W * 4, t->cur_inv, t->cur_v, t->cur_p);
}

// GetReferenceDepth (RGBDTracking.cu:547-550) returns the raw-depth buffer,
// i.e. the depth most recently uploaded from the host (NULL before any).
int yas_tracker_raw_depth(yas_tracker* t, float** depth, size_t* pitch) {
  YAS_REQUIRE(t && depth && pitch, "yas_tracker_raw_depth: null argument");
  *depth = t->raw_depth;
  *pitch = (size_t)t->W * sizeof(float);
  return YAS_OK;
}

// WriteDebugImages (RGBDTracking.cu:557-568): the level-0 current and
// reference intensity and the two gradient images, as 8-bit binary PGM
// files curr_image.pgm, last_image.pgm, gx.pgm, gy.pgm in `dir` (values
// rounded to nearest and clamped to [0, 255]).
int yas_tracker_write_debug_images(yas_tracker* t, const char* dir) {
  YAS_REQUIRE(t && dir, "yas_tracker_write_debug_images: null argument");
  YAS_HIP(hipSetDevice(t->ctx->device));
  YAS_HIP(hipStreamSynchronize(t->ctx->stream));
  const size_t n = (size_t)t->W * t->H;
  const struct {
    const char* name;
    const float* img;
  } out[4] = {{"curr_image.pgm", t->cur_int[0]}, {"last_image.pgm", t->ref_int[0]}, {"gx.pgm", t->gx[0]},
              {"gy.pgm", t->gy[0]}};
  std::vector<float> f(n);
  std::vector<unsigned char> b(n);
  for (const auto& o : out) {
    YAS_HIP(hipMemcpy(f.data(), o.img, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
      float v = std::nearbyint(f[i]);
      b[i] = (unsigned char)(v != v ? 0.f : std::min(255.f, std::max(0.f, v)));
    }
    std::string path = std::string(dir) + "/" + o.name;
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
      yas_set_error("yas_tracker_write_debug_images: cannot open %s", path.c_str());
      return YAS_E_INVALID;
    }
    std::fprintf(fp, "P5\n%d %d\n255\n", t->W, t->H);
    bool ok = std::fwrite(b.data(), 1, n, fp) == n;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) {
      yas_set_error("yas_tracker_write_debug_images: write to %s failed", path.c_str());
      return YAS_E_INVALID;
    }
  }
  return YAS_OK;
}

int yas_tracker_set_reference_model(yas_tracker* t, const float* vmap, size_t pitch) {
  YAS_REQUIRE(t && vmap && pitch >= (size_t)t->W * 16, "yas_tracker_set_reference_model: bad argument");
  hipLaunchKernelGGL(k_copy_f4, grid64x4(t->W, t->H), dim3(256), 0, t->ctx->stream, (const float4*)vmap, pitch,
                     t->ref_v[0], t->W, t->H);
  YAS_LAUNCH_CHECK();
  return model_pyramid(t);
}

}  // extern "C"

// After yas_tracker_swap_frame_buffer: the copies that restore the previous
// reference intensity and inverse depth (the buffers now named current).
int yas_tracker_unswap_copies(yas_tracker* t, RefRestore& r) {
  r.count = 0;
  for (int l = 0; l < t->P.levels; ++l) {
    const int n = t->lv[l].w * t->lv[l].h;
    r.dst[r.count] = t->ref_int[l]; r.src[r.count] = t->cur_int[l]; r.n[r.count++] = n;
    r.dst[r.count] = t->ref_inv[l]; r.src[r.count] = t->cur_inv[l]; r.n[r.count++] = n;
  }
  return YAS_OK;
}

extern "C" {

int yas_tracker_swap_frame_buffer(yas_tracker* t) {
  YAS_REQUIRE(t, "null tracker");
  for (int l = 0; l < kMaxLevels; ++l) {
    std::swap(t->ref_v[l], t->cur_v[l]);
    std::swap(t->ref_p[l], t->cur_p[l]);
    std::swap(t->ref_inv[l], t->cur_inv[l]);
    std::swap(t->ref_int[l], t->cur_int[l]);
  }
  return YAS_OK;
}

int yas_tracker_get_transform(yas_tracker* t, const double init[16], int swap, double out[16], yas_track_stats* stats) {
  YAS_REQUIRE(t && out, "yas_tracker_get_transform: null argument");
  YAS_HIP(hipSetDevice(t->ctx->device));
  Mat44 M;
  if (init) std::memcpy(M.m, init, sizeof(M.m));
  else
    for (int i = 0; i < 16; ++i) M.m[i] = (i % 5 == 0) ? 1.0 : 0.0;
  hipLaunchKernelGGL(k_track_init, dim3(1), dim3(64), 0, t->ctx->stream, t->st, M);
  YAS_LAUNCH_CHECK();
  int rc = yas_tracker_solve_async(t);
  if (rc) return rc;
  YAS_HIP(hipMemcpyAsync(t->host_st, t->st, sizeof(TrackState), hipMemcpyDeviceToHost, t->ctx->stream));
  YAS_HIP(hipStreamSynchronize(t->ctx->stream));
  const TrackState& h = *t->host_st;
  // RGBDTracking.cu:187-191: a NaN update returns before the buffer swap
  if (swap && h.good) yas_tracker_swap_frame_buffer(t);
  if (h.good) std::memcpy(out, h.last_ok, sizeof(h.last_ok));
  else
    for (int i = 0; i < 16; ++i) out[i] = (i % 5 == 0) ? 1.0 : 0.0;
  if (stats) {
    stats->good = h.good;
    stats->icp_inliers = h.icp_inliers;
    stats->rgb_inliers = h.rgb_inliers;
    stats->last_error = h.last_err;
  }
  return h.good ? YAS_OK : YAS_E_DIVERGED;
}

int yas_tracker_is_tracking_good(yas_tracker* t, int* good) {
  YAS_REQUIRE(t && good, "null argument");
  YAS_HIP(hipMemcpyAsync(t->host_st, t->st, sizeof(TrackState), hipMemcpyDeviceToHost, t->ctx->stream));
  YAS_HIP(hipStreamSynchronize(t->ctx->stream));
  *good = t->host_st->good;
  return YAS_OK;
}

int yas_tracker_step_sums(yas_tracker* t, int level, const double T[16], int kind, float out[29], float stats[4]) {
  YAS_REQUIRE(t && T && out, "yas_tracker_step_sums: null argument");
  YAS_REQUIRE(level >= 0 && level < t->P.levels, "yas_tracker_step_sums: bad level");
  YAS_HIP(hipSetDevice(t->ctx->device));
  Mat44 M;
  std::memcpy(M.m, T, sizeof(M.m));
  hipLaunchKernelGGL(k_track_init, dim3(1), dim3(64), 0, t->ctx->stream, t->st, M);
  int rc = launch_iteration(t, level, kind == 0, kind != 0, false, false);
  if (rc) return rc;
  YAS_HIP(hipMemcpyAsync(t->host_st, t->st, sizeof(TrackState), hipMemcpyDeviceToHost, t->ctx->stream));
  YAS_HIP(hipStreamSynchronize(t->ctx->stream));
  const float* s = kind == 0 ? t->host_st->sums_icp : t->host_st->sums_rgb;
  std::memcpy(out, s, sizeof(float) * 29);
  if (stats) std::memcpy(stats, t->host_st->rgb_stats, sizeof(float) * 4);
  return YAS_OK;
}

namespace {
// One wavefront per system: the Gauss-Newton step's 6x6 solve as the
// tracker's kernels run it (ldlt_solve6_wave), for yas_tracker_solve6.
__global__ void __launch_bounds__(64) k_solve6(const float* H, const float* b, double* x) {
  __shared__ float sH[36], sb[6];
  __shared__ double sx[6];
  const size_t s = blockIdx.x;
  const int lane = threadIdx.x;
  if (lane < 36) sH[lane] = H[s * 36 + lane];
  if (lane < 6) sb[lane] = b[s * 6 + lane];
  wave_lds_sync();
  double xi[6];
  ldlt_solve6_wave(sH, sb, sx, xi);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) x[s * 6 + i] = xi[i];
  }
}
}  // namespace

int yas_tracker_solve6(yas_ctx* ctx, const float* H, const float* b, double* x, int n) {
  YAS_REQUIRE(ctx && H && b && x, "yas_tracker_solve6: null argument");
  YAS_REQUIRE(n >= 1 && n <= (1 << 20), "yas_tracker_solve6: n must be 1..2^20");
  YAS_HIP(hipSetDevice(ctx->device));
  const size_t nh = (size_t)n * 36 * sizeof(float), nb = (size_t)n * 6 * sizeof(float), nx = (size_t)n * 6 * sizeof(double);
  char* buf = nullptr;
  YAS_HIP(hipMalloc((void**)&buf, nh + nb + nx));
  float* dH = (float*)buf;
  float* db = (float*)(buf + nh);
  double* dx = (double*)(buf + nh + nb);
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemcpyAsync(dH, H, nh, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(db, b, nb, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_solve6, dim3(n), dim3(64), 0, s, dH, db, dx);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(x, dx, nx, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  else hipStreamSynchronize(s);
  hipFree(buf);
  YAS_HIP(e);
  return YAS_OK;
}

int yas_tracker_level(yas_tracker* t, int which, int level, void** ptr, size_t* pitch, int* w, int* h) {
  YAS_REQUIRE(t && ptr, "null argument");
  YAS_REQUIRE(level >= 0 && level < t->P.levels, "bad level");
  const Level& a = t->lv[level];
  void* one[6] = {t->cur_int[level], t->ref_int[level], t->gx[level], t->gy[level], t->cur_inv[level], t->ref_inv[level]};
  void* vec[5] = {t->cur_v[level], t->cur_n[level], t->ref_v[level], t->ref_n[level], t->resid[level]};
  YAS_REQUIRE(which >= 0 && which < 11, "bad buffer id");
  if (which == 7 || which == 9) {
    // The passes read vertex and normal as ICP points; the float4 normal
    // map is formed here from the vertex map, as the pyramid kernels would.
    YAS_HIP(hipSetDevice(t->ctx->device));
    hipLaunchKernelGGL(k_normal, grid64x4(a.w, a.h), dim3(256), 0, t->ctx->stream,
                       which == 7 ? t->cur_v[level] : t->ref_v[level], (size_t)a.w * 16,
                       which == 7 ? t->cur_n[level] : t->ref_n[level], (size_t)a.w * 16, a.w, a.h);
    YAS_LAUNCH_CHECK();
    YAS_HIP(hipStreamSynchronize(t->ctx->stream));
  }
  *ptr = which < 6 ? one[which] : vec[which - 6];
  if (pitch) *pitch = (size_t)a.w * (which < 6 ? 4 : which < 10 ? 16 : 12);
  if (w) *w = a.w;
  if (h) *h = a.h;
  return YAS_OK;
}

int yas_rgb_to_intensity(yas_ctx* ctx, const uint8_t* rgb, size_t sp, float* dst, size_t dp, int W, int H, int mode) {
  YAS_REQUIRE(ctx && rgb && dst && W > 0 && H > 0, "yas_rgb_to_intensity: bad argument");
  return yas_rgb_to_gray_async(ctx, rgb, sp, dst, dp, W, H, mode);
}
int yas_depth_to_invdepth(yas_ctx* ctx, const float* d, size_t sp, float* inv, size_t dp, int W, int H) {
  YAS_REQUIRE(ctx && d && inv && W > 0 && H > 0, "bad argument");
  hipLaunchKernelGGL(k_invdepth, grid64x4(W, H), dim3(256), 0, ctx->stream, d, sp, inv, dp, W, H);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_pyrdown_depth(yas_ctx* ctx, const float* s, size_t sp, float* d, size_t dp, int W, int H) {
  YAS_REQUIRE(ctx && s && d && W > 0 && H > 0, "bad argument");
  hipLaunchKernelGGL(k_pyrdown_depth, grid64x4(W, H), dim3(256), 0, ctx->stream, s, sp, d, dp, W, H);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_pyrdown_gauss(yas_ctx* ctx, const float* s, size_t sp, int sw, int sh, float* d, size_t dp) {
  YAS_REQUIRE(ctx && s && d && sw >= 4 && sh >= 4, "yas_pyrdown_gauss: source must be at least 4x4");
  int dw = sw / 2, dh = sh / 2;
  hipLaunchKernelGGL(k_pyrdown_gauss, dim3((dw + 15) / 16, (dh + 15) / 16), dim3(256), 0, ctx->stream, s, sp, sw, sh, d,
                     dp, dw, dh);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_compute_vertex_map(yas_ctx* ctx, const float* inv, size_t sp, float* v, size_t dp, int W, int H, const float K[4],
                           float cutoff) {
  YAS_REQUIRE(ctx && inv && v && K && W > 0 && H > 0, "bad argument");
  Intr k = make_intr(K);
  hipLaunchKernelGGL(k_vertex, grid64x4(W, H), dim3(256), 0, ctx->stream, inv, sp, (float4*)v, dp, W, H, k.invfx,
                     k.invfy, k.cx, k.cy, cutoff);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_compute_normal_map(yas_ctx* ctx, const float* v, size_t sp, float* n, size_t dp, int W, int H) {
  YAS_REQUIRE(ctx && v && n && W > 0 && H > 0, "bad argument");
  hipLaunchKernelGGL(k_normal, grid64x4(W, H), dim3(256), 0, ctx->stream, (const float4*)v, sp, (float4*)n, dp, W, H);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_image_gradient(yas_ctx* ctx, const float* s, size_t sp, float* gx, float* gy, size_t dp, int W, int H) {
  YAS_REQUIRE(ctx && s && gx && gy && W > 0 && H > 0, "bad argument");
  hipLaunchKernelGGL(k_gradient, grid64x4(W, H), dim3(256), 0, ctx->stream, s, sp, gx, gy, dp, W, H);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_vmap_to_depth(yas_ctx* ctx, const float* v, size_t sp, float* d, size_t dp, int W, int H) {
  YAS_REQUIRE(ctx && v && d && W > 0 && H > 0, "bad argument");
  hipLaunchKernelGGL(k_vmap_to_depth, grid64x4(W, H), dim3(256), 0, ctx->stream, (const float4*)v, sp, d, dp, W, H);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

}  // extern "C"
