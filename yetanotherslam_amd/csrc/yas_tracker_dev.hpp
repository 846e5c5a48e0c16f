// Device code of the dense tracker (RGBDFusion/src/RGBDTracking.cu,
// RGBDFusion/include/TrackingUtils.h) and its image operations
// (RGBDFusion/src/ImageProc.cu), shared by the single-tracker kernels and
// the kernels that run many trackers in one launch.  Every body takes its
// block coordinates (bx, by) and, for the reductions, the number of
// workgroups nb of its own grid, so a tracker's arithmetic and summation
// order are the same whichever launch runs it.
#pragma once

#include <cfloat>
#include <cmath>

#include "yas_internal.hpp"

namespace yas {

constexpr int kRedGroups = 256;  // workgroups of a reduction pass
constexpr int kMaxLevels = 5;

struct Level {
  int w, h;
  float fx, fy, cx, cy, invfx, invfy;
  double Kd[4];
  double Kinv[4];  // K^-1 terms in double: 1/fx, 1/fy, -cx/fx, -cy/fy (computed on the host)
};

struct TrackState {
  double est[16];
  double last_ok[16];
  float sums_rgb[32];
  float sums_icp[32];
  float rgb_stats[4];  // sum r, count, mean, Huber threshold
  float last_err;
  int good;
  int icp_inliers;
  int rgb_inliers;
};

// ------------------------------------------------------------------ reductions
// Wave-wide sums of 32 values by a transpose butterfly: at each step a lane
// keeps half of its values and trades the other half with its partner, so
// 32 values need 16+8+4+2+1+1 shuffles.  On return, lanes 2i and 2i+1 hold
// the total of value i.  The step size is a template constant so every
// register index is static.
template <int S>
__device__ inline void butterfly_step(float (&v)[32], int lane) {
  const bool hi = (lane & (2 * S)) != 0;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const float send = hi ? v[i] : v[i + S];
    const float keep = hi ? v[i + S] : v[i];
    v[i] = keep + __shfl_xor(send, 2 * S, 64);
  }
}
__device__ inline float wave_sum32(float (&v)[32]) {
  const int lane = threadIdx.x & 63;
  butterfly_step<16>(v, lane);
  butterfly_step<8>(v, lane);
  butterfly_step<4>(v, lane);
  butterfly_step<2>(v, lane);
  butterfly_step<1>(v, lane);
  return v[0] + __shfl_xor(v[0], 1, 64);
}

// Workgroup sums of 32 values; thread t < 32 returns the total of value t.
__device__ inline float block_sum32(float (&v)[32], float* lds /* [nwaves * 32] */) {
  const float w = wave_sum32(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if ((lane & 1) == 0) lds[wid * 32 + (lane >> 1)] = w;
  __syncthreads();
  float t = 0.0f;
  if (threadIdx.x < 32) {
    const int nw = blockDim.x >> 6;
    t = lds[threadIdx.x];
    for (int k = 1; k < nw; ++k) t += lds[k * 32 + threadIdx.x];
  }
  __syncthreads();
  return t;
}

// Chan et al.'s pairwise update of (count, mean, M2) statistics.
__device__ inline void chan_merge(float& n, float& mean, float& m2, float nb, float meanb, float m2b) {
  const float nn = n + nb;
  if (nb == 0.0f) return;
  if (n == 0.0f) {
    n = nb; mean = meanb; m2 = m2b;
    return;
  }
  const float d = meanb - mean;
  mean = mean + d * (nb / nn);
  m2 = m2 + m2b + d * d * (n * nb / nn);
  n = nn;
}

// Wavefront merge of per-thread statistics: every lane gets the wavefront's.
__device__ inline void wave_stats(float& n, float& mean, float& m2) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    float nb = __shfl_xor(n, o, 64), mb = __shfl_xor(mean, o, 64), qb = __shfl_xor(m2, o, 64);
    // merge in a lane-independent order so both partners get identical values
    if ((threadIdx.x & o) == 0) {
      chan_merge(n, mean, m2, nb, mb, qb);
    } else {
      float n0 = nb, m0 = mb, q0 = qb;
      chan_merge(n0, m0, q0, n, mean, m2);
      n = n0; mean = m0; m2 = q0;
    }
  }
}

// Workgroup merge of per-thread statistics; every thread gets the result,
// merging the wavefronts' in the same order itself (one barrier).  `st` is
// read after the barrier, so nothing in the kernel may write it afterwards.
__device__ inline void block_stats(float& n, float& mean, float& m2, float* st /* [3 * nwaves] */) {
  wave_stats(n, mean, m2);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) {
    st[3 * wid] = n; st[3 * wid + 1] = mean; st[3 * wid + 2] = m2;
  }
  __syncthreads();
  float a = st[0], b = st[1], c = st[2];
  for (int k = 1; k < nw; ++k) chan_merge(a, b, c, st[3 * k], st[3 * k + 1], st[3 * k + 2]);
  n = a; mean = b; m2 = c;
}

// block_sum32 of two arrays in one LDS round: their wavefront sums overlap
// and one pair of barriers serves both.  Thread t < 32 returns value t's
// total of `a`, thread 32 + t that of `b`, each summed as block_sum32 does.
__device__ inline float block_sum32x2(float (&a)[32], float (&b)[32], float* lds /* [2 * nwaves * 32] */) {
  const float wa = wave_sum32(a);
  const float wb = wave_sum32(b);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((lane & 1) == 0) {
    lds[wid * 32 + (lane >> 1)] = wa;
    lds[(nw + wid) * 32 + (lane >> 1)] = wb;
  }
  __syncthreads();
  float t = 0.0f;
  if (threadIdx.x < 64) {
    const int base = threadIdx.x < 32 ? 0 : nw * 32, i = threadIdx.x & 31;
    t = lds[base + i];
    for (int k = 1; k < nw; ++k) t += lds[base + k * 32 + i];
  }
  __syncthreads();
  return t;
}

// The end of a pass 1 with both terms: block_sum32 of the ICP values and
// block_stats of the residual statistics in one LDS round.  Thread t < 32
// returns ICP value t's total; thread 32 gets the merged statistics in (n,
// mean, m2).  The same orders as the two separate reductions.  No barrier
// follows, so the caller must not use `lds` again.
__device__ inline float block_end_pass1(float (&acc)[32], float& n, float& mean, float& m2,
                                        float* lds /* [nwaves * 35] */) {
  const float w = wave_sum32(acc);
  wave_stats(n, mean, m2);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  float* st = lds + nw * 32;
  if ((lane & 1) == 0) lds[wid * 32 + (lane >> 1)] = w;
  if (lane == 0) {
    st[3 * wid] = n; st[3 * wid + 1] = mean; st[3 * wid + 2] = m2;
  }
  __syncthreads();
  float t = 0.0f;
  if (threadIdx.x < 32) {
    t = lds[threadIdx.x];
    for (int k = 1; k < nw; ++k) t += lds[k * 32 + threadIdx.x];
  } else if (threadIdx.x == 32) {
    float a = st[0], b = st[1], c = st[2];
    for (int k = 1; k < nw; ++k) chan_merge(a, b, c, st[3 * k], st[3 * k + 1], st[3 * k + 2]);
    n = a; mean = b; m2 = c;
  }
  return t;
}

// A value every lane holds alike (one LDS word read by all), moved into a
// scalar register so that the pixel loops spend no vector register on it.
__device__ inline float wave_uniform(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

__device__ inline void load_pose_f(const double* T, float* lds_pose /* 12 */) {
  if (threadIdx.x < 12) {
    int i = threadIdx.x / 4, j = threadIdx.x % 4;
    int k = j < 3 ? 3 * i + j : 9 + i;
    lds_pose[k] = (float)T[4 * i + j];
  }
}

// ------------------------------------------------------------------ image ops
__device__ inline float ld(const float* p, size_t pitch, int x, int y) {
  return ((const float*)((const char*)p + (size_t)y * pitch))[x];
}

__device__ inline void copy_f32_at(const float* src, size_t sp, float* dst, int W, int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x < W && y < H) dst[(size_t)y * W + x] = ld(src, sp, x, y);
}

// mode 0: 0.299R + 0.587G + 0.114B (float); mode 1: 8-bit fixed point.
__device__ inline void rgb_to_gray_at(const uint8_t* rgb, size_t sp, float* dst, size_t dp, int W,
                                                     int H, int mode, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  const uint8_t* p = rgb + (size_t)y * sp + 3 * x;
  int r = p[0], g = p[1], b = p[2];
  float v = mode == 0 ? 0.299f * (float)r + 0.587f * (float)g + 0.114f * (float)b
                      : (float)((r * 4899 + g * 9617 + b * 1868 + (1 << 13)) >> 14);
  ((float*)((char*)dst + (size_t)y * dp))[x] = v;
}

// OpenCV pyrDown restated: binomial [1 4 6 4 1]/16 per axis (vertical sums of
// the 5 source columns first, then horizontal), BORDER_REFLECT_101.
__device__ inline int reflect101(int i, int n) {
  if (i < 0) i = -i;
  if (i >= n) i = 2 * n - 2 - i;
  return i;
}
__device__ __attribute__((always_inline)) inline float pyrdown_gauss_px(const float* src, size_t sp, int sw, int sh,
                                                                       int x, int y) {
  const float k[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
  int rows[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) rows[i] = reflect101(2 * y + i - 2, sh);
  float col[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    int sx = reflect101(2 * x + j - 2, sw);
    float acc = k[0] * ld(src, sp, sx, rows[0]);
#pragma unroll
    for (int i = 1; i < 5; ++i) acc = acc + k[i] * ld(src, sp, sx, rows[i]);
    col[j] = acc;
  }
  float acc = k[0] * col[0];
#pragma unroll
  for (int j = 1; j < 5; ++j) acc = acc + k[j] * col[j];
  return acc;
}
__device__ __attribute__((always_inline)) inline void pyrdown_gauss_at(const float* src, size_t sp, int sw, int sh,
                                                                       float* dst, size_t dp, int dw, int dh, int bx,
                                                                       int by) {
  int x = bx * 16 + (threadIdx.x & 15), y = by * 16 + (threadIdx.x >> 4);
  if (x >= dw || y >= dh) return;
  ((float*)((char*)dst + (size_t)y * dp))[x] = pyrdown_gauss_px(src, sp, sw, sh, x, y);
}

// ImageProc.cu:7-25.
__device__ inline void gradient_px(const float* s, size_t sp, int W, int H, int x, int y, float& a, float& b) {
  a = 0.0f;
  b = 0.0f;
  if (!(x < 1 || y < 1 || x > W - 2 || y > H - 2)) {
    a = (ld(s, sp, x + 1, y) - ld(s, sp, x - 1, y)) * 0.5f;
    b = (ld(s, sp, x, y + 1) - ld(s, sp, x, y - 1)) * 0.5f;
  }
}
__device__ inline void gradient_at(const float* s, size_t sp, float* gx, float* gy, size_t dp, int W,
                                                  int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  float a, b;
  gradient_px(s, sp, W, H, x, y, a, b);
  ((float*)((char*)gx + (size_t)y * dp))[x] = a;
  ((float*)((char*)gy + (size_t)y * dp))[x] = b;
}

// ImageProc.cu:141-153.
__device__ inline float invdepth_of(float v) { return (v > 0.2f && v < 8.0f) ? 1.0f / v : 0.0f; }
__device__ inline void invdepth_at(const float* z, size_t sp, float* o, size_t dp, int W, int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  ((float*)((char*)o + (size_t)y * dp))[x] = invdepth_of(ld(z, sp, x, y));
}

// ImageProc.cu:166-175, last row/column (left unwritten there) set to 0.
__device__ inline float pyrdown_depth_px(const float* s, size_t sp, int W, int H, int x, int y) {
  return (x < W - 1 && y < H - 1) ? ld(s, sp, 2 * x, 2 * y) : 0.0f;
}
__device__ inline void pyrdown_depth_at(const float* s, size_t sp, float* o, size_t dp, int W, int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  ((float*)((char*)o + (size_t)y * dp))[x] = pyrdown_depth_px(s, sp, W, H, x, y);
}

// Whole-float4 subsampling of the reference model (departure from the
// reference's float kernel run over float4 data, RGBDTracking.cu:146).
__device__ inline float4 pyrdown_vmap_px(const float4* s, size_t sp, int W, int H, int x, int y) {
  float4 v = make_float4(0.f, 0.f, 0.f, -1.f);
  if (x < W - 1 && y < H - 1) v = ((const float4*)((const char*)s + (size_t)(2 * y) * sp))[2 * x];
  return v;
}
__device__ inline void pyrdown_vmap_at(const float4* s, size_t sp, float4* o, size_t dp, int W, int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  ((float4*)((char*)o + (size_t)y * dp))[x] = pyrdown_vmap_px(s, sp, W, H, x, y);
}

__device__ inline void copy_f4_at(const float4* s, size_t sp, float4* o, int W, int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x < W && y < H) o[(size_t)y * W + x] = ((const float4*)((const char*)s + (size_t)y * sp))[x];
}

// ImageProc.cu:189-209.
__device__ inline float4 vertex_of(float iz, int x, int y, float invfx, float invfy, float cx, float cy, float cut) {
  float z = 1.0f / iz;
  float4 o = make_float4(0.f, 0.f, 0.f, -1.f);
  if (iz > 0.0f && z < cut) o = make_float4(z * ((float)x - cx) * invfx, z * ((float)y - cy) * invfy, z, 1.0f);
  return o;
}
__device__ inline void vertex_at(const float* inv, size_t sp, float4* v, size_t dp, int W, int H,
                                                float invfx, float invfy, float cx, float cy, float cut, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  ((float4*)((char*)v + (size_t)y * dp))[x] = vertex_of(ld(inv, sp, x, y), x, y, invfx, invfy, cx, cy, cut);
}

// ImageProc.cu:222-249.
__device__ inline float4 normal_px(const float4* v, size_t sp, int W, int H, int x, int y) {
  float4 o = make_float4(0.f, 0.f, 0.f, -1.f);
  if (x != W - 1 && y != H - 1) {
    const float4 a = ((const float4*)((const char*)v + (size_t)y * sp))[x];
    const float4 b = ((const float4*)((const char*)v + (size_t)y * sp))[x + 1];
    const float4 c = ((const float4*)((const char*)v + (size_t)(y + 1) * sp))[x];
    if (a.w > 0.f && b.w > 0.f && c.w > 0.f) {
      float p0 = c.x - a.x, p1 = c.y - a.y, p2 = c.z - a.z;
      float q0 = b.x - a.x, q1 = b.y - a.y, q2 = b.z - a.z;
      float r0 = p1 * q2 - p2 * q1, r1 = p2 * q0 - p0 * q2, r2 = p0 * q1 - p1 * q0;
      float n2 = r0 * r0 + r1 * r1 + r2 * r2;
      if (n2 > 0.0f) {
        float nn = sqrtf(n2);
        r0 = r0 / nn; r1 = r1 / nn; r2 = r2 / nn;
      }
      o = make_float4(r0, r1, r2, 1.0f);
    }
  }
  return o;
}
__device__ inline void normal_at(const float4* v, size_t sp, float4* n, size_t dp, int W, int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  ((float4*)((char*)n + (size_t)y * dp))[x] = normal_px(v, sp, W, H, x, y);
}

// What the ICP association reads of a pixel: its vertex and normal, 24
// bytes instead of two float4.  A normal exists (w = 1) only where the
// pixel's own vertex exists (normal_px tests it), and the association needs
// both, so a pixel without a normal stores a NaN x instead of the two w
// flags: the reference side skips it, and on the current side the NaN
// distance fails the 1 cm gate, as the w tests would.
struct IcpPoint {
  float vx, vy, vz, nx, ny, nz;
};
// A level's ICP point map (n pixels) is two planes, so every access is one
// aligned 16-byte and one aligned 8-byte access: (vx, vy, vz, nx) as n
// float4, then (ny, nz) as n float2.
__device__ inline IcpPoint load_icp_point(const float4* m, int n, int k) {
  const float4 a = m[k];
  const float2 b = reinterpret_cast<const float2*>(m + n)[k];
  IcpPoint p;
  p.vx = a.x; p.vy = a.y; p.vz = a.z; p.nx = a.w; p.ny = b.x; p.nz = b.y;
  return p;
}
__device__ inline void store_icp_point(float4* m, int n, int k, const IcpPoint& p) {
  m[k] = make_float4(p.vx, p.vy, p.vz, p.nx);
  reinterpret_cast<float2*>(m + n)[k] = make_float2(p.ny, p.nz);
}
__device__ inline IcpPoint icp_point_px(const float4* v, size_t sp, int W, int H, int x, int y) {
  const float4 a = ((const float4*)((const char*)v + (size_t)y * sp))[x];
  const float4 n = normal_px(v, sp, W, H, x, y);
  IcpPoint p;
  p.vx = n.w > 0.0f ? a.x : __builtin_nanf("");
  p.vy = a.y; p.vz = a.z;
  p.nx = n.x; p.ny = n.y; p.nz = n.z;
  return p;
}
__device__ inline void icp_point_at(const float4* v, size_t sp, float4* o, int W, int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  store_icp_point(o, W * H, y * W + x, icp_point_px(v, sp, W, H, x, y));
}

// ImageProc.cu:262-275.
__device__ inline void vmap_to_depth_at(const float4* v, size_t sp, float* d, size_t dp, int W, int H, int bx, int by) {
  int x = bx * 64 + (threadIdx.x & 63), y = by * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  float4 a = ((const float4*)((const char*)v + (size_t)y * sp))[x];
  ((float*)((char*)d + (size_t)y * dp))[x] = a.w > 0.f ? a.z : 0.0f;
}

// ------------------------------------------------------------------ per-pixel terms
struct TrackBufs {
  const float4* ref_v;
  const float4* ref_p;  // ICP point maps (two planes, load_icp_point)
  const float4* cur_v;
  const float4* cur_p;
  const float* ref_inv;
  const float* ref_int;
  const float* cur_int;
  const float* gx;
  const float* gy;
  float3* resid;  // photometric rows (r, gx, gy); r is NaN where a pixel has none
};

// RGBDTracking.cu:316-372: projective association of reference pixel k into
// the current frame, gated at sin(20 deg) normal angle and 1 cm; on a match
// the point-to-plane row [n, v x n, n.(v_c - v)] is accumulated.  Only
// pixels with a vertex and a normal on both sides can match (IcpPoint).
__device__ inline void icp_pixel(const Level& L, const TrackBufs& B, const PoseF& T, float angle_th, float dist_th,
                                 const IcpPoint& r, float (&acc)[32]) {
  if (!(r.vx == r.vx)) return;
  float vx, vy, vz;
  xform(T, r.vx, r.vy, r.vz, vx, vy, vz);
  const float invz = 1.0f / vz;
  const int u = floor_sat(L.fx * vx * invz + L.cx + 0.5f);
  const int v = floor_sat(L.fy * vy * invz + L.cy + 0.5f);
  if (u < 0 || v < 0 || u >= L.w || v >= L.h) return;
  const IcpPoint q = load_icp_point(B.cur_p, L.w * L.h, v * L.w + u);
  float nx, ny, nz;
  rot3(T, r.nx, r.ny, r.nz, nx, ny, nz);
  const float dx = vx - q.vx, dy = vy - q.vy, dz = vz - q.vz;
  const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
  const float c0 = q.ny * nz - q.nz * ny, c1 = q.nz * nx - q.nx * nz, c2 = q.nx * ny - q.ny * nx;
  const float ang = sqrtf(c0 * c0 + c1 * c1 + c2 * c2);
  if (!(ang < angle_th && dist <= dist_th)) return;  // a NaN distance (no current point) fails here
  float row[7];
  row[0] = nx; row[1] = ny; row[2] = nz;
  row[3] = vy * nz - vz * ny;
  row[4] = vz * nx - vx * nz;
  row[5] = vx * ny - vy * nx;
  row[6] = nx * (q.vx - vx) + ny * (q.vy - vy) + nz * (q.vz - vz);
  int c = 0;
#pragma unroll
  for (int i = 0; i < 7; ++i)
#pragma unroll
    for (int j = i; j < 7; ++j) acc[c++] += row[i] * row[j];
  acc[28] += 1.0f;
}

// Warp parameters of ImageProc.cu:42-78: RK = R K^-1 (formed in double), t.
// Kinv holds K^-1's terms (Level::Kinv), so no division is on this path.
__device__ inline void load_warp(const double* T, const double* Kinv, float* lds /* 12 */) {
  if (threadIdx.x >= 16 && threadIdx.x < 28) {
    int q = threadIdx.x - 16;
    int i = q / 4, j = q % 4;
    if (j < 3) {
      double Ki[9] = {Kinv[0], 0.0, Kinv[2], 0.0, Kinv[1], Kinv[3], 0.0, 0.0, 1.0};
      lds[3 * i + j] = (float)(T[4 * i] * Ki[j] + T[4 * i + 1] * Ki[3 + j] + T[4 * i + 2] * Ki[6 + j]);
    } else {
      lds[9 + i] = (float)T[4 * i + 3];
    }
  }
}

__device__ inline float bilerp(const float* m, int w, float x, float y) {
  int u = (int)floorf(x), v = (int)floorf(y);
  float cx = x - (float)u, cy = y - (float)v;
  const float* r0 = m + (size_t)v * w;
  const float* r1 = r0 + w;
  return (r0[u] * (1.0f - cx) + r0[u + 1] * cx) * (1.0f - cy) + (r1[u] * (1.0f - cx) + r1[u + 1] * cx) * cy;
}

// ImageProc.cu:42-61: reference pixel k at its inverse depth, moved by the
// warp (RK, t); w = -1 where the inverse depth is 0.  Both passes evaluate it
// (the same operations, so the same floats) instead of storing it.
__device__ inline float4 warp_ref(const Level& L, const float* wp, int k, float iz) {
  const int y = k / L.w, x = k - y * L.w;
  float4 p = make_float4(0.f, 0.f, 0.f, -1.f);
  if (iz > 0.0f) {
    float z = 1.0f / iz;
    float px = (wp[0] * (float)x + wp[1] * (float)y + wp[2]) * z + wp[9];
    float py = (wp[3] * (float)x + wp[4] * (float)y + wp[5]) * z + wp[10];
    float pz = (wp[6] * (float)x + wp[7] * (float)y + wp[8]) * z + wp[11];
    p = make_float4(px, py, pz, 1.0f);
  }
  return p;
}

// TrackingUtils.h:73-151 with the warp of ImageProc.cu:42-61 fused in; the
// residual statistics are accumulated with Welford's update.
__device__ inline void rgb_pixel(const Level& L, const TrackBufs& B, const float* wp, int k, float iz, float ref,
                                 float& n, float& mean, float& m2) {
  const float4 p = warp_ref(L, wp, k, iz);
  float r = 0.0f, gx = 0.0f, gy = 0.0f;
  bool ok = false;
  if (p.w > 0.0f) {
    float u = L.fx * p.x / p.z + L.cx;
    float v = L.fy * p.y / p.z + L.cy;
    if (u > 0.0f && v > 0.0f && u < (float)(L.w - 1) && v < (float)(L.h - 1)) {
      float cur = bilerp(B.cur_int, L.w, u, v);
      r = cur - ref;
      gx = bilerp(B.gx, L.w, u, v);
      gy = bilerp(B.gy, L.w, u, v);
      ok = (gx * gx + gy * gy) > 0.0f && isfinite(r);
    }
  }
  B.resid[k] = ok ? make_float3(r, gx, gy) : make_float3(__builtin_nanf(""), 0.0f, 0.0f);
  if (ok) {
    n += 1.0f;
    float d = r - mean;
    mean = mean + d / n;
    m2 = m2 + d * (r - mean);
  }
}

// Pass 1: ICP rows and/or photometric residuals of one linearisation.
// Per-pixel reference values that pass 1 reads at pixel k before any
// dependent (gathered) read.
struct RefPixel {
  IcpPoint p;       // model vertex and normal (ICP)
  float iz, inten;  // inverse depth and intensity (photometric)
};
template <bool ICP, bool RGB>
__device__ inline RefPixel load_ref_pixel(const TrackBufs& B, int n, int k) {
  RefPixel r;
  if (ICP) r.p = load_icp_point(B.ref_p, n, k);
  if (RGB) {
    r.iz = B.ref_inv[k];
    r.inten = B.ref_int[k];
  }
  return r;
}

// PREFETCH: the next pixel's reference values are loaded before the
// current pixel's gathers, so their latencies overlap (more registers; the
// values and the order of the sums are the same).
// The Gauss-Newton step of the previous linearisation, taken by a pass 1
// before it linearises (StepArgs): every workgroup folds the same partial
// rows and takes the same step, so all get the same pose; workgroup 0 also
// stores the new state.  The step reads st_in and writes st_out, a
// different buffer, because other workgroups may still be reading st_in.
struct StepArgs {
  const TrackState* st_in;
  TrackState* st_out;
  const float* part_icp;
  int gi;
  const float* part_rgb;
  int gr;
  int mode, first_of_level, record;
  float wrh, wrb, wih, wib;
};
__device__ __attribute__((always_inline)) inline void gn_update(const TrackState* st_in, TrackState* st_out,
                                                                bool write, const float* part_icp, int gi,
                                                                const float* part_rgb, int gr, int mode,
                                                                int first_of_level, int do_step, int record, float wrh,
                                                                float wrb, float wih, float wib, double* est_out);

template <bool ICP, bool RGB, bool PREFETCH = false, bool STEP = false>
__device__ inline void pass1_at(const Level& L, const TrackBufs& B, const TrackState* st, float* part_icp,
                                float* part_res, float angle_th, float dist_th, int bx, int nb,
                                const StepArgs* sa = nullptr) {
  __shared__ float pose[12];
  __shared__ float wp[12];
  __shared__ float lds[4 * 35];
  __shared__ double s_est[STEP ? 16 : 1];
  const int N = L.w * L.h;
  const int stride = nb * blockDim.x;
  int k = bx * blockDim.x + threadIdx.x;
  RefPixel nxt = {};
  if (PREFETCH && k < N) nxt = load_ref_pixel<ICP, RGB>(B, N, k);  // in flight while the pose loads
  const double* est;
  if (STEP) {
    gn_update(sa->st_in, sa->st_out, bx == 0, sa->part_icp, sa->gi, sa->part_rgb, sa->gr, sa->mode,
              sa->first_of_level, 1, sa->record, sa->wrh, sa->wrb, sa->wih, sa->wib, s_est);
    __syncthreads();
    est = s_est;
  } else {
    est = st->est;
  }
  load_pose_f(est, pose);
  if (RGB) load_warp(est, L.Kinv, wp);
  __syncthreads();
  PoseF T;
#pragma unroll
  for (int i = 0; i < 9; ++i) T.r[i] = wave_uniform(pose[i]);
  T.t[0] = wave_uniform(pose[9]); T.t[1] = wave_uniform(pose[10]); T.t[2] = wave_uniform(pose[11]);
  float wu[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) wu[i] = RGB ? wave_uniform(wp[i]) : 0.0f;
  float acc[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) acc[i] = 0.0f;
  float n = 0.0f, mean = 0.0f, m2 = 0.0f;
  for (; k < N; k += stride) {
    RefPixel cur;
    if (PREFETCH) {
      cur = nxt;
      if (k + stride < N) nxt = load_ref_pixel<ICP, RGB>(B, N, k + stride);
    } else {
      cur = load_ref_pixel<ICP, RGB>(B, N, k);
    }
    if (ICP) icp_pixel(L, B, T, angle_th, dist_th, cur.p, acc);
    if (RGB) rgb_pixel(L, B, wu, k, cur.iz, cur.inten, n, mean, m2);
  }
  if (ICP && RGB) {
    const float t = block_end_pass1(acc, n, mean, m2, lds);
    if (threadIdx.x < 32) part_icp[bx * 32 + threadIdx.x] = t;
    if (threadIdx.x == 32) {
      part_res[bx * 4] = n;
      part_res[bx * 4 + 1] = mean;
      part_res[bx * 4 + 2] = m2;
    }
  } else if (ICP) {
    float t = block_sum32(acc, lds);
    if (threadIdx.x < 32) part_icp[bx * 32 + threadIdx.x] = t;
  } else {
    block_stats(n, mean, m2, lds + 4 * 32);
    if (threadIdx.x == 0) {
      part_res[bx * 4] = n;
      part_res[bx * 4 + 1] = mean;
      part_res[bx * 4 + 2] = m2;
    }
  }
}

// Pass 2 (TrackingUtils.h:153-220, RGBDTracking.cu:262-290): fold the
// residual statistics of pass 1 (identically in every workgroup), derive
// sigma and the Huber threshold 4.685 sigma, and reduce the weighted rows.
// PREFETCH: each pixel's residual and inverse depth are loaded one pixel
// ahead (the first pixel's while the statistics are folded).
// The residual statistics of pass 1 folded in a fixed order (every thread
// gets the same values), with sigma's Huber threshold 4.685 sigma; thread 0
// of workgroup 0 records {sum r, count, mean, threshold} when `record`.
__device__ inline float huber_threshold(TrackState* st, const float* part_res, int g1, float* st_lds, bool record) {
  float n = 0.0f, mean = 0.0f, m2 = 0.0f;
  for (int g = threadIdx.x; g < g1; g += blockDim.x)
    chan_merge(n, mean, m2, part_res[g * 4], part_res[g * 4 + 1], part_res[g * 4 + 2]);
  block_stats(n, mean, m2, st_lds);
  const float sigma = sqrtf(m2 / (n - 1.0f));
  const float th = (float)(4.685 * (double)sigma);
  if (record && threadIdx.x == 0) {
    st->rgb_stats[0] = n * mean;
    st->rgb_stats[1] = n;
    st->rgb_stats[2] = mean;
    st->rgb_stats[3] = th;
  }
  return th;
}

// PRESTATS: the threshold was folded by a launch of its own (one workgroup
// per tracker, huber_threshold) and is read from the state, instead of
// every workgroup folding it again.
template <bool PREFETCH = false, bool PRESTATS = false>
__device__ inline void pass2_at(const Level& L, const TrackBufs& B, TrackState* st, const float* part_res, int g1,
                                float* part_rgb, int bx, int nb) {
  __shared__ float lds[4 * 35];
  __shared__ float wp[12];
  const int N = L.w * L.h;
  const int stride = nb * blockDim.x;
  int k = bx * blockDim.x + threadIdx.x;
  float3 e_next = make_float3(0.f, 0.f, 0.f);
  float iz_next = 0.0f;
  if (PREFETCH && k < N) {
    e_next = B.resid[k];
    iz_next = B.ref_inv[k];
  }
  load_warp(st->est, L.Kinv, wp);
  float th;
  if (PRESTATS) {
    th = st->rgb_stats[3];
    __syncthreads();  // the warp parameters in LDS
  } else {
    // apart from the region block_sum32 uses below
    th = huber_threshold(st, part_res, g1, lds + 4 * 32, bx == 0);
  }
  float wu[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) wu[i] = wave_uniform(wp[i]);
  float acc[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) acc[i] = 0.0f;
  for (; k < N; k += stride) {
    float3 e;
    float iz;
    if (PREFETCH) {
      e = e_next;
      iz = iz_next;
      if (k + stride < N) {
        e_next = B.resid[k + stride];
        iz_next = B.ref_inv[k + stride];
      }
    } else {
      e = B.resid[k];
      iz = B.ref_inv[k];
    }
    if (!(e.x == e.x)) continue;  // no residual at this pixel
    const float4 p = warp_ref(L, wu, k, iz);
    const float zinv = 1.0f / p.z;
    float wt = 1.0f;
    if (fabsf(e.x) > th) wt = th / fabsf(e.x);
    float row[7];
    row[0] = e.y * L.fx * zinv;
    row[1] = e.z * L.fy * zinv;
    row[2] = -(row[0] * p.x + row[1] * p.y) * zinv;
    row[3] = row[2] * p.y - e.z * L.fy;
    row[4] = -row[2] * p.x + e.y * L.fx;
    row[5] = -row[0] * p.y + row[1] * p.x;
    row[6] = -e.x;
    int c = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i)
#pragma unroll
      for (int j = i; j < 7; ++j) acc[c++] += wt * row[i] * row[j];
    acc[28] += 1.0f;
  }
  float t = block_sum32(acc, lds);
  if (threadIdx.x < 32) part_rgb[bx * 32 + threadIdx.x] = t;
}

// ------------------------------------------------------------------ GN update
// Swap by selects on both values, so the run-time pivot never becomes an
// array index (which would move the array to scratch memory).
__host__ __device__ inline void swap_if(bool p, double& a, double& b) {
  const double x = a, z = b;
  a = p ? z : x;
  b = p ? x : z;
}
// Eigen 3.3's LDLT restated (Eigen/src/Cholesky/LDLT.h, unblocked: diagonal
// pivoting, left-looking, pseudo-inverse of D), solving H x = b for the
// symmetric 6x6 H, with its pivots found first.  Eigen's pivot k is the
// largest |diagonal| among positions k..5 of the matrix as permuted so far
// (the first such position on ties), and when step k searches them those
// diagonal entries are still the input's: the factorisation changes A(k,k)
// and the column below it at step k only.  So the transpositions follow
// from the input diagonal alone, and the factorisation is the unpivoted one
// of P H P^T, the same operations on the same values as Eigen's.  H and b
// are read through the permutation, and `scratch` (6 doubles) undoes it; on
// the device all three live in LDS.  The unrolled factorisation and solves
// then use static register indices only.
__host__ __device__ inline void ldlt_solve6_presorted(const float* H, const float* b, double* scratch,
                                                      double (&x)[6]) {
  constexpr int n = 6;
  double d[n];
  int p[n];
#pragma unroll
  for (int i = 0; i < n; ++i) {
    d[i] = (double)H[i * n + i];
    p[i] = i;
  }
  bool stop = false;
#pragma unroll
  for (int k = 0; k < n; ++k) {
    int big = k;
    double bv = fabs(d[k]);
#pragma unroll
    for (int i = k + 1; i < n; ++i)
      if (fabs(d[i]) > bv) {
        bv = fabs(d[i]);
        big = i;
      }
#pragma unroll
    for (int i = k + 1; i < n; ++i) {
      const bool sw = i == big;
      swap_if(sw, d[k], d[i]);
      const int pk = p[k], pi = p[i];
      p[k] = sw ? pi : pk;
      p[i] = sw ? pk : pi;
    }
    if (k == 0 && !(fabs(d[0]) > 0.0)) stop = true;  // no pivot at all (then big was 0: nothing moved)
  }
  if (stop) {
#pragma unroll
    for (int i = 0; i < n; ++i) p[i] = i;
  }
  double A[n * n], y[n];
#pragma unroll
  for (int i = 0; i < n; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) A[i * n + j] = (double)H[p[i] * n + p[j]];
    y[i] = (double)b[p[i]];
  }
  if (!stop) {
#pragma unroll
    for (int k = 0; k < n; ++k) {
      if (k > 0) {
        double temp[n];
#pragma unroll
        for (int j = 0; j < k; ++j) temp[j] = A[j * n + j] * A[k * n + j];
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < k; ++j) s += A[k * n + j] * temp[j];
        A[k * n + k] -= s;
#pragma unroll
        for (int i = k + 1; i < n; ++i) {
          double r = 0.0;
#pragma unroll
          for (int j = 0; j < k; ++j) r += A[i * n + j] * temp[j];
          A[i * n + k] -= r;
        }
      }
      const double akk = A[k * n + k];
      if (fabs(akk) > 0.0) {
#pragma unroll
        for (int i = k + 1; i < n; ++i) A[i * n + k] /= akk;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) y[i] -= A[i * n + j] * y[j];
#pragma unroll
  for (int i = 0; i < n; ++i) {
    const double dd = A[i * n + i];
    y[i] = fabs(dd) > DBL_MIN ? y[i] / dd : 0.0;
  }
#pragma unroll
  for (int i = n - 1; i >= 0; --i)
#pragma unroll
    for (int j = i + 1; j < n; ++j) y[i] -= A[j * n + i] * y[j];
#pragma unroll
  for (int i = 0; i < n; ++i) scratch[p[i]] = y[i];
#pragma unroll
  for (int i = 0; i < n; ++i) x[i] = scratch[i];
}

// Lanes of one wavefront exchanging values through LDS or registers: the
// fences keep the compiler from moving LDS accesses across this point (a
// wavefront's own LDS accesses complete in order).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Lane `src`'s value (src a constant or wavefront-uniform).
__device__ inline double lane_value(double v, int src) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// ldlt_solve6_presorted run by a whole wavefront (all 64 lanes active,
// every lane gets x): lane r < 6 holds row r of the permuted matrix, and at
// factorisation step k the rows r >= k update their column k at once, each
// by the same operations in the same order as the serial loops (the
// products temp[j] = D_j L_kj from lane j's diagonal and lane k's row, the
// sums from j = 0 up, the division by D_k), so x has the same bits.  The
// pivots and the back substitution stay serial (the latter's sums run over
// j ascending, which needs the finished y_j in that order); lanes 6..63
// mirror row 5.  H and b are read from LDS; `scratch` (6 doubles, LDS)
// undoes the permutation.
__device__ inline void ldlt_solve6_wave(const float* H, const float* b, double* scratch, double (&x)[6]) {
  constexpr int n = 6;
  const int lane = threadIdx.x & 63;
  const int r = lane < n ? lane : n - 1;
  double d[n];
  int p[n];
#pragma unroll
  for (int i = 0; i < n; ++i) {
    d[i] = (double)H[i * n + i];
    p[i] = i;
  }
  bool stop = false;
#pragma unroll
  for (int k = 0; k < n; ++k) {
    int big = k;
    double bv = fabs(d[k]);
#pragma unroll
    for (int i = k + 1; i < n; ++i)
      if (fabs(d[i]) > bv) {
        bv = fabs(d[i]);
        big = i;
      }
#pragma unroll
    for (int i = k + 1; i < n; ++i) {
      const bool sw = i == big;
      swap_if(sw, d[k], d[i]);
      const int pk = p[k], pi = p[i];
      p[k] = sw ? pi : pk;
      p[i] = sw ? pk : pi;
    }
    if (k == 0 && !(fabs(d[0]) > 0.0)) stop = true;
  }
  if (stop) {
#pragma unroll
    for (int i = 0; i < n; ++i) p[i] = i;
  }
  int pr = p[0];
#pragma unroll
  for (int j = 1; j < n; ++j) pr = r == j ? p[j] : pr;
  double a[n];  // row r: a[j] = A(r, j) for j <= r
#pragma unroll
  for (int j = 0; j < n; ++j) a[j] = (double)H[pr * n + p[j]];
  double y = (double)b[pr];
  if (!stop) {
#pragma unroll
    for (int k = 0; k < n; ++k) {
      if (k > 0) {
        double temp[n];
#pragma unroll
        for (int j = 0; j < k; ++j) temp[j] = lane_value(a[j], j) * lane_value(a[j], k);
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < k; ++j) s += a[j] * temp[j];
        if (r >= k) a[k] -= s;
      }
      const double akk = lane_value(a[k], k);
      if (fabs(akk) > 0.0 && r > k) a[k] /= akk;
    }
  }
  // forward substitution: y_r -= A(r, j) y_j for j = 0 .. r-1, in that order
#pragma unroll
  for (int j = 0; j < n - 1; ++j) {
    const double yj = lane_value(y, j);
    if (r > j) y -= a[j] * yj;
  }
  double dd = a[0];
#pragma unroll
  for (int j = 1; j < n; ++j) dd = r == j ? a[j] : dd;
  y = fabs(dd) > DBL_MIN ? y / dd : 0.0;
  double yu[n];
#pragma unroll
  for (int i = 0; i < n; ++i) yu[i] = lane_value(y, i);
#pragma unroll
  for (int i = n - 1; i >= 0; --i)
#pragma unroll
    for (int j = i + 1; j < n; ++j) yu[i] -= lane_value(a[i], j) * yu[j];
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < n; ++i) scratch[p[i]] = yu[i];
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < n; ++i) x[i] = scratch[i];
}

// SE(3) exponential, twist (rho, omega), Rodrigues' formula.
__device__ inline void se3_exp(const double* xi, double* T) {
  const double* w = xi + 3;
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = sqrt(th2);
  double A, B, C;
  if (th < 1e-8) {
    A = 1.0 - th2 / 6.0;
    B = 0.5 - th2 / 24.0;
    C = 1.0 / 6.0 - th2 / 120.0;
  } else {
    A = sin(th) / th;
    B = (1.0 - cos(th)) / th2;
    C = (th - sin(th)) / (th2 * th);
  }
  double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  double W2[9], Rm[9], V[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W2[3 * i + j] = W[3 * i] * W[j] + W[3 * i + 1] * W[3 + j] + W[3 * i + 2] * W[6 + j];
  for (int i = 0; i < 9; ++i) {
    double I = (i % 4 == 0) ? 1.0 : 0.0;
    Rm[i] = I + A * W[i] + B * W2[i];
    V[i] = I + B * W[i] + C * W2[i];
  }
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) T[4 * i + j] = Rm[3 * i + j];
    T[4 * i + 3] = V[3 * i] * xi[0] + V[3 * i + 1] * xi[1] + V[3 * i + 2] * xi[2];
  }
  T[12] = T[13] = T[14] = 0.0;
  T[15] = 1.0;
}

// RankUpdateHessian<6,7> (TrackingUtils.h:45-58).
__device__ inline void unpack29(const float* s, float* H, float* b) {
  int k = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 7; ++j) {
      float v = s[k++];
      if (j == 6) b[i] = v;
      else H[6 * i + j] = H[6 * j + i] = v;
    }
}

// Wavefront 0's part of gn_update (all 64 lanes, which leave together):
// inlier counts and, when do_step and the state is good, the step from the
// pose s_est.  Writes the state's fields to st and the stepped pose to
// est_out (each when not null).  Each value is formed by the operations of
// the serial form, in its order: H and b one element per lane
// (RankUpdateHessian's packing, the RGB term weighted first), the solve by
// ldlt_solve6_wave, the exponential in every lane alike, and E * est one
// element per lane for lanes 0..15.  lds_H (36), lds_b (6), lds_x (6) and
// lds_E (16) are LDS.
__device__ __attribute__((always_inline)) inline void gn_step_wave0(const float (*sums)[32], const double* s_est,
                                                                    int s_good, float s_last_err, TrackState* st,
                                                                    double* est_out, int gi, int gr, int mode,
                                                                    int first_of_level, int do_step, int record,
                                                                    float wrh, float wrb, float wih, float wib,
                                                                    float* lds_H, float* lds_b, double* lds_x,
                                                                    double* lds_E) {
  const int lane = threadIdx.x & 63;
  if (st && lane == 0 && record && gi) st->icp_inliers = (int)sums[0][28];
  if (st && lane == 0 && record && gr) st->rgb_inliers = (int)sums[1][28];
  if (!do_step) return;
  float last_err = s_last_err;
  if (first_of_level) {
    last_err = FLT_MAX;
    if (st && lane == 0) st->last_err = FLT_MAX;
  }
  if (!s_good) return;
  if (lane < 42) {
    // element (i, j), i <= j, of the packed upper rows [H | b] (7 - i entries in row i)
    const bool isb = lane >= 36;
    const int e0 = lane / 6, e1 = lane - 6 * e0;
    const int i = isb ? lane - 36 : (e0 < e1 ? e0 : e1);
    const int j = isb ? 6 : (e0 < e1 ? e1 : e0);
    const int k = 7 * i - (i * (i - 1)) / 2 + (j - i);
    float v = 0.0f;
    if (mode != YAS_TRACK_DEPTH) {
      const float w = mode == YAS_TRACK_RGBD ? (isb ? wrb : wrh) : 1.0f;
      v = v + w * sums[1][k];
    }
    if (mode != YAS_TRACK_RGB) {
      const float w = mode == YAS_TRACK_RGBD ? (isb ? wib : wih) : 1.0f;
      v = v + w * sums[0][k];
    }
    if (isb) lds_b[i] = v;
    else lds_H[lane] = v;
  }
  float rsum = 0.0f, nres = 0.0f;
  if (mode != YAS_TRACK_DEPTH) {
    rsum = sums[1][27];
    nres = sums[1][28];
  }
  if (mode != YAS_TRACK_RGB) rsum = sums[0][27];
  const float err = sqrtf(rsum) / (nres + 6.0f);
  wave_lds_sync();
  double xi[6];
  ldlt_solve6_wave(lds_H, lds_b, lds_x, xi);
  if (isnan(xi[0])) {
    if (st && lane == 0) st->good = 0;
    return;
  }
  double E[16];
  se3_exp(xi, E);
  if (lane < 16) {
#pragma unroll
    for (int q = 0; q < 16; ++q) lds_E[q] = E[q];
  }
  wave_lds_sync();
  if (lane < 16) {
    const int i = lane >> 2, j = lane & 3;
    const double t = lds_E[4 * i] * s_est[j] + lds_E[4 * i + 1] * s_est[4 + j] + lds_E[4 * i + 2] * s_est[8 + j] +
                     lds_E[4 * i + 3] * s_est[12 + j];
    if (est_out) est_out[lane] = t;
    if (st) {
      st->est[lane] = t;
      if (err < last_err) st->last_ok[lane] = t;
    }
  }
  if (st && lane == 0 && err < last_err) st->last_err = err;
}

// Fold the partial rows of one linearisation and, when `do_step`, take one
// Gauss-Newton step of RGBDTracking.cu:161-204.  One workgroup.  The state
// is read from st_in; when `write`, st_out receives it with the fold's and
// the step's changes (st_out may be st_in).  When est_out (LDS) is given,
// wavefront 0 leaves the pose after the step there.
__device__ __attribute__((always_inline)) inline void gn_update(const TrackState* st_in, TrackState* st_out,
                                                                bool write, const float* part_icp, int gi,
                                                                const float* part_rgb, int gr, int mode,
                                                                int first_of_level, int do_step, int record, float wrh,
                                                                float wrb, float wih, float wib, double* est_out) {
  __shared__ float lds[2 * 4 * 32];
  __shared__ float sums[2][32];
  __shared__ double s_est[16];
  __shared__ int s_good;
  __shared__ float s_last_err;
  __shared__ float s_H[36], s_b[6];  // wavefront 0's step
  __shared__ double s_x[6], s_E[16];
  // the state the step reads, loaded alongside the partial rows
  if (threadIdx.x < 16) s_est[threadIdx.x] = st_in->est[threadIdx.x];
  if (threadIdx.x == 16) s_good = st_in->good;
  if (threadIdx.x == 17) s_last_err = st_in->last_err;
  // both terms' rows are read in one pass, so their loads overlap
  float vi[32], vr[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) vi[i] = vr[i] = 0.0f;
  const int g = gi > gr ? gi : gr;
  for (int r = threadIdx.x; r < g; r += blockDim.x) {
    if (r < gi) {
#pragma unroll
      for (int i = 0; i < 32; ++i) vi[i] += part_icp[r * 32 + i];
    }
    if (r < gr) {
#pragma unroll
      for (int i = 0; i < 32; ++i) vr[i] += part_rgb[r * 32 + i];
    }
  }
  if (gi && gr) {
    const float t = block_sum32x2(vi, vr, lds);
    if (threadIdx.x < 64) sums[threadIdx.x >> 5][threadIdx.x & 31] = t;
  } else if (gi) {
    const float t = block_sum32(vi, lds);
    if (threadIdx.x < 32) sums[0][threadIdx.x] = t;
  } else if (gr) {
    const float t = block_sum32(vr, lds);
    if (threadIdx.x < 32) sums[1][threadIdx.x] = t;
  }
  if (write && st_out != st_in) {
    // the fields this update leaves alone carry over
    const int nw = (int)(sizeof(TrackState) / 4);
    for (int w = threadIdx.x; w < nw; w += blockDim.x) ((int*)st_out)[w] = ((const int*)st_in)[w];
  }
  __syncthreads();
  if (write && threadIdx.x < 29) {
    if (gi) st_out->sums_icp[threadIdx.x] = sums[0][threadIdx.x];
    if (gr) st_out->sums_rgb[threadIdx.x] = sums[1][threadIdx.x];
  }
  if (threadIdx.x >= 64) return;
  if (est_out && threadIdx.x < 16) est_out[threadIdx.x] = s_est[threadIdx.x];  // the pose where no step is taken
  gn_step_wave0(sums, s_est, s_good, s_last_err, write ? st_out : nullptr, est_out, gi, gr, mode, first_of_level,
                do_step, record, wrh, wrb, wih, wib, s_H, s_b, s_x, s_E);
}
__device__ inline void update_at(TrackState* st, const float* part_icp, int gi, const float* part_rgb, int gr,
                                 int mode, int first_of_level, int do_step, int record, float wrh, float wrb, float wih,
                                 float wib) {
  gn_update(st, st, true, part_icp, gi, part_rgb, gr, mode, first_of_level, do_step, record, wrh, wrb, wih, wib,
            nullptr);
}

__device__ inline void track_init_at(TrackState* st, const double* T) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < 16; ++i) st->est[i] = st->last_ok[i] = T[i];
  st->good = 1;
  st->last_err = FLT_MAX;
}
__device__ inline void track_init_identity_at(TrackState* st) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < 16; ++i) st->est[i] = st->last_ok[i] = (i % 5 == 0) ? 1.0 : 0.0;
  st->good = 1;
  st->last_err = FLT_MAX;
}

}  // namespace yas
