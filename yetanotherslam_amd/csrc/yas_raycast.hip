// Ray casting of a TSDF volume into a camera-frame vertex map: the
// reference's RayTraceEngine (RGBDFusion/src/RayTraceEngine.cu:11-454).
//
// Two kernels per ray cast, after the visibility pass:
//   k_zrange  one thread per visible block projects its 8 corners onto the
//             1/8-resolution grid and min/max-reduces the block's depth range
//             into every covered cell.  Depths are positive floats, whose bit
//             patterns order like the values, so unsigned atomicMin/atomicMax
//             are exact replacements for the reference's CAS loops
//             (VoxelStructUtils.h:150-171).  The reference's split of each box
//             into 16x16 tiles (RayTraceEngine.cu:146-214) only distributed
//             work; the per-cell result is the same min/max over all covering
//             blocks, without its 100000-tile cap.
//   k_march   one thread per pixel marches from the cell's zmin to zmax in
//             voxel units (RayTraceEngine.cu:331-410).  The block looked up last
//             is cached in registers, so consecutive samples inside one 8^3
//             block cost no hash probe.
#include <cstring>

#include "yas_internal.hpp"

using namespace yas;

MapView yas_map_view(yas_map* map);
const PosePair* yas_map_pose_dev(yas_map* map);
yas_ctx* yas_map_ctx(yas_map* map);
int yas_map_visible_async(yas_map* map, const double T[16], int W, int H);
int yas_map_visible_async_dev(yas_map* map, const double* T_dev, int W, int H);

struct yas_raycaster {
  yas_ctx* ctx;
  int W, H, cols, rows;
  Intr K;
  float4* vmap;
  unsigned* zmin;
  unsigned* zmax;
  int* host_ctr;
  yas_map* last_map;
};

__global__ void __launch_bounds__(256) k_zrange_init(unsigned* zmin, unsigned* zmax, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    zmin[i] = __float_as_uint(100.0f);
    zmax[i] = 0u;
  }
}

// RayTraceEngine.cu:80-144 (createRenderingBlock) + 184-214 (fill).
__global__ void __launch_bounds__(256) k_zrange(MapView m, const PosePair* pose, Intr K, int cols, int rows,
                                                unsigned* zmin, unsigned* zmax) {
  const int nv = m.ctr[kCtrVisible];
  const int nvis = nv < m.num_blocks ? nv : m.num_blocks;
  const PoseF Ti = pose->Tinv;
  const float scale = m.voxel_size * 8.0f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nvis; i += gridDim.x * blockDim.x) {
    int bx, by, bz;
    unpack_key(m.block_keys[m.visible[i]], bx, by, bz);
    short ul0 = (short)cols, ul1 = (short)rows, lr0 = -1, lr1 = -1;
    float zr0 = m.depth_max, zr1 = m.depth_min;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      float x = (float)(bx + (c & 1)) * scale;
      float y = (float)(by + ((c >> 1) & 1)) * scale;
      float z = (float)(bz + ((c >> 2) & 1)) * scale;
      float qx, qy, qz;
      xform(Ti, x, y, z, qx, qy, qz);
      float pu = (K.fx * qx / qz + K.cx) / 8.0f;
      float pv = (K.fy * qy / qz + K.cy) / 8.0f;
      if ((float)ul0 > floorf(pu)) ul0 = (short)floor_sat(pu);
      if ((float)lr0 < ceilf(pu)) lr0 = (short)ceil_sat(pu);
      if ((float)ul1 > floorf(pv)) ul1 = (short)floor_sat(pv);
      if ((float)lr1 < ceilf(pv)) lr1 = (short)ceil_sat(pv);
      if (zr0 > qz) zr0 = qz;
      if (zr1 < qz) zr1 = qz;
    }
    if (ul0 < 0) ul0 = 0;
    if (ul1 < 0) ul1 = 0;
    if (lr0 >= cols) lr0 = (short)(cols - 1);
    if (lr1 >= rows) lr1 = (short)(rows - 1);
    if (ul0 > lr0 || ul1 > lr1) continue;
    if (zr0 < m.depth_min) zr0 = m.depth_min;
    if (zr1 < m.depth_min) continue;
    const unsigned a = __float_as_uint(zr0), b = __float_as_uint(zr1);
    for (int y = ul1; y <= lr1; ++y)
      for (int x = ul0; x <= lr0; ++x) {
        atomicMin(&zmin[y * cols + x], a);
        atomicMax(&zmax[y * cols + x], b);
      }
  }
}

// Voxel reads with a one-entry block cache.
template <class VT>
struct VoxelReader {
  const MapView& m;
  const typename VT::T* vox;
  int cbx, cby, cbz, cblk, cflags;
  __device__ VoxelReader(const MapView& mv)
      : m(mv), vox((const typename VT::T*)mv.voxels), cbx(INT32_MIN), cby(0), cbz(0), cblk(-1), cflags(0) {}
  // RayTraceEngine.cu:279-292: nearest read at floor(p), valid iff observed.
  // Blocks summarised as all-unobserved or all-free answer without a voxel
  // load (the values they would return are exactly (1, invalid) / (1, valid)).
  __device__ inline float read(float x, float y, float z, bool& valid) {
    int vx = floor_sat(x), vy = floor_sat(y), vz = floor_sat(z);
    int bx = vox_to_block(vx), by = vox_to_block(vy), bz = vox_to_block(vz);
    if (bx != cbx || by != cby || bz != cbz) {
      cbx = bx; cby = by; cbz = bz;
      cflags = 0;
      cblk = find_block_xyz(m, bx, by, bz, &cflags);
    }
    if (cblk >= 0 && cflags == kBlockFree) {
      valid = true;
      return 1.0f;
    }
    if (cblk >= 0 && cflags != kBlockUnobserved) {
      typename VT::T v = vox[(size_t)cblk * 512 + (vox_to_local(vz) * 8 + vox_to_local(vy)) * 8 + vox_to_local(vx)];
      if (VT::observed(v)) {
        valid = true;
        return VT::sdf(v);
      }
    }
    valid = false;
    return 1.0f;
  }
  // RayTraceEngine.cu:294-329: trilinear read, valid iff all 8 reads are.
  __device__ inline float interp(float x, float y, float z, bool& valid) {
    float fx = x - floorf(x), fy = y - floorf(y), fz = z - floorf(z);
    bool vp, v;
    float s0 = read(x, y, z, vp);
    float s1 = read(x + 1.0f, y, z, v); vp &= v;
    float r0 = (1.0f - fx) * s0 + fx * s1;
    s0 = read(x, y + 1.0f, z, v); vp &= v;
    s1 = read(x + 1.0f, y + 1.0f, z, v); vp &= v;
    float r1 = (1.0f - fx) * s0 + fx * s1;
    float r2 = (1.0f - fy) * r0 + fy * r1;
    s0 = read(x, y, z + 1.0f, v); vp &= v;
    s1 = read(x + 1.0f, y, z + 1.0f, v); vp &= v;
    r0 = (1.0f - fx) * s0 + fx * s1;
    s0 = read(x, y + 1.0f, z + 1.0f, v); vp &= v;
    s1 = read(x + 1.0f, y + 1.0f, z + 1.0f, v); vp &= v;
    r1 = (1.0f - fx) * s0 + fx * s1;
    float r3 = (1.0f - fy) * r0 + fy * r1;
    valid = vp;
    return (1.0f - fz) * r2 + fz * r3;
  }
};

// RayTraceEngine.cu:331-410.  Departure: pixels the reference leaves
// unwritten get w = -1 (x = NaN, as the reference writes first).
template <class VT>
__global__ void __launch_bounds__(256) k_march(MapView m, const PosePair* pose, Intr K, int W, int H, int cols,
                                               const float* zmin, const float* zmax, float4* vmap, float vsinv,
                                               float step_scale) {
  const int tiles_x = (W + 15) >> 4;
  const int tile = xcd_tile(blockIdx.x, tiles_x * ((H + 15) >> 4));
  const int x = (tile % tiles_x) * 16 + (threadIdx.x & 15);
  const int y = (tile / tiles_x) * 16 + (threadIdx.x >> 4);
  if (x >= W || y >= H) return;
  float4 out = make_float4(__int_as_float(0x7fffffff), 0.0f, 0.0f, -1.0f);
  const float zn = zmin[(y >> 3) * cols + (x >> 3)];
  const float zf = zmax[(y >> 3) * cols + (x >> 3)];
  const float eps = 1.1920928955078125e-07f;
  if (!(zn < eps || zf < eps || zn != zn || zf != zf)) {
    const PoseF T = pose->T;
    float pnx = K.invfx * ((float)x - K.cx) * zn, pny = K.invfy * ((float)y - K.cy) * zn;
    float dist_s = sqrtf(pnx * pnx + pny * pny + zn * zn) * vsinv;
    float bsx, bsy, bsz;
    xform(T, pnx, pny, zn, bsx, bsy, bsz);
    bsx = bsx * vsinv; bsy = bsy * vsinv; bsz = bsz * vsinv;
    float pfx = K.invfx * ((float)x - K.cx) * zf, pfy = K.invfy * ((float)y - K.cy) * zf;
    float dist_e = sqrtf(pfx * pfx + pfy * pfy + zf * zf) * vsinv;
    float bex, bey, bez;
    xform(T, pfx, pfy, zf, bex, bey, bez);
    bex = bex * vsinv; bey = bey * vsinv; bez = bez * vsinv;
    float dx = bex - bsx, dy = bey - bsy, dz = bez - bsz;
    float n2 = dx * dx + dy * dy + dz * dz;
    if (n2 > 0.0f) {
      float nn = sqrtf(n2);
      dx = dx / nn; dy = dy / nn; dz = dz / nn;
    }
    VoxelReader<VT> rd(m);
    float rx = bsx, ry = bsy, rz = bsz;
    float sdf = 1.0f, last;
    bool valid = false, abandoned = false;
    while (dist_s < dist_e) {
      last = sdf;
      sdf = rd.read(rx, ry, rz, valid);
      if (sdf <= 0.5f && sdf >= -0.5f) sdf = rd.interp(rx, ry, rz, valid);
      if (sdf <= 0.0f) break;
      if (sdf >= 0.0f && last < 0.0f) {
        abandoned = true;
        break;
      }
      float step = valid ? fmaxf(sdf * step_scale, 1.0f) : 1.0f;
      rx = rx + step * dx; ry = ry + step * dy; rz = rz + step * dz;
      dist_s = dist_s + step;
    }
    if (!abandoned && sdf <= 0.0f) {
      float step = sdf * step_scale;
      rx = rx + step * dx; ry = ry + step * dy; rz = rz + step * dz;
      sdf = rd.interp(rx, ry, rz, valid);
      step = sdf * step_scale;
      rx = rx + step * dx; ry = ry + step * dy; rz = rz + step * dz;
      if (valid) {
        const PoseF Ti = pose->Tinv;
        float cx, cy, cz;
        xform(Ti, rx / vsinv, ry / vsinv, rz / vsinv, cx, cy, cz);
        out = make_float4(cx, cy, cz, 1.0f);
      }
    }
  }
  vmap[(size_t)y * W + x] = out;
}

static int launch_raycast(yas_raycaster* rc, yas_map* map, hipEvent_t before_march = nullptr) {
  hipStream_t s = rc->ctx->stream;
  MapView v = yas_map_view(map);
  const PosePair* pose = yas_map_pose_dev(map);
  const int ncell = rc->cols * rc->rows;
  hipLaunchKernelGGL(k_zrange_init, dim3((ncell + 255) / 256), dim3(256), 0, s, rc->zmin, rc->zmax, ncell);
  int g = (v.num_blocks + 255) / 256;
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_zrange, dim3(g), dim3(256), 0, s, v, pose, rc->K, rc->cols, rc->rows, rc->zmin, rc->zmax);
  YAS_LAUNCH_CHECK();
  if (before_march) YAS_HIP(hipEventRecord(before_march, s));
  const float vsinv = (float)(1.0 / (double)v.voxel_size);
  const float step_scale = v.trunc / v.voxel_size;
  dim3 grid(((rc->W + 15) / 16) * ((rc->H + 15) / 16));
  switch (v.format) {
    case YAS_VOXEL_I16_U8:
      hipLaunchKernelGGL(k_march<VoxI16>, grid, dim3(256), 0, s, v, pose, rc->K, rc->W, rc->H, rc->cols,
                         (const float*)rc->zmin, (const float*)rc->zmax, rc->vmap, vsinv, step_scale);
      break;
    case YAS_VOXEL_F16_U16:
      hipLaunchKernelGGL(k_march<VoxF16>, grid, dim3(256), 0, s, v, pose, rc->K, rc->W, rc->H, rc->cols,
                         (const float*)rc->zmin, (const float*)rc->zmax, rc->vmap, vsinv, step_scale);
      break;
    default:
      hipLaunchKernelGGL(k_march<VoxF32>, grid, dim3(256), 0, s, v, pose, rc->K, rc->W, rc->H, rc->cols,
                         (const float*)rc->zmin, (const float*)rc->zmax, rc->vmap, vsinv, step_scale);
  }
  YAS_LAUNCH_CHECK();
  rc->last_map = map;
  return YAS_OK;
}

// Ray cast with the pose held in device memory (pipeline).
// `before_march` (optional) is recorded between the z-range and march kernels.
int yas_raycast_dev(yas_raycaster* rc, yas_map* map, const double* T_dev, hipEvent_t before_march) {
  int r = yas_map_visible_async_dev(map, T_dev, rc->W, rc->H);
  if (r) return r;
  return launch_raycast(rc, map, before_march);
}

float4* yas_raycaster_vmap_dev(yas_raycaster* rc) { return rc->vmap; }

extern "C" {

int yas_raycaster_create(yas_ctx* ctx, int W, int H, const float K[4], yas_raycaster** out) {
  YAS_REQUIRE(ctx && K && out, "yas_raycaster_create: null argument");
  YAS_REQUIRE(W >= 8 && H >= 8, "yas_raycaster_create: image too small");
  YAS_REQUIRE(W % 8 == 0 && H % 8 == 0, "yas_raycaster_create: width and height must be multiples of 8");
  YAS_HIP(hipSetDevice(ctx->device));
  yas_raycaster* rc = new yas_raycaster();
  rc->ctx = ctx;
  rc->W = W;
  rc->H = H;
  rc->cols = W / 8;
  rc->rows = H / 8;
  rc->K = make_intr(K);
  hipError_t e = hipMalloc((void**)&rc->vmap, sizeof(float4) * (size_t)W * H);
  if (e == hipSuccess) e = hipMalloc((void**)&rc->zmin, sizeof(unsigned) * rc->cols * rc->rows);
  if (e == hipSuccess) e = hipMalloc((void**)&rc->zmax, sizeof(unsigned) * rc->cols * rc->rows);
  if (e == hipSuccess) e = hipHostMalloc((void**)&rc->host_ctr, sizeof(int) * kCtrCount, hipHostMallocDefault);
  if (e != hipSuccess) {
    yas_set_error("yas_raycaster_create: %s", hipGetErrorString(e));
    yas_raycaster_destroy(rc);
    return YAS_E_HIP;
  }
  // nothing traced yet: every pixel invalid
  YAS_HIP(hipMemsetAsync(rc->vmap, 0, sizeof(float4) * (size_t)W * H, ctx->stream));
  *out = rc;
  return YAS_OK;
}

int yas_raycaster_destroy(yas_raycaster* rc) {
  if (!rc) return YAS_OK;
  hipSetDevice(rc->ctx->device);
  hipStreamSynchronize(rc->ctx->stream);
  hipFree(rc->vmap);
  hipFree(rc->zmin);
  hipFree(rc->zmax);
  if (rc->host_ctr) hipHostFree(rc->host_ctr);
  delete rc;
  return YAS_OK;
}

int yas_raycast(yas_raycaster* rc, yas_map* map, const double T[16]) {
  YAS_REQUIRE(rc && map && T, "yas_raycast: null argument");
  YAS_REQUIRE_ACTIVE(map);
  YAS_HIP(hipSetDevice(rc->ctx->device));
  YAS_REQUIRE(yas_map_ctx(map)->stream == rc->ctx->stream, "yas_raycast: map and ray caster must share a stream");
  int r = yas_map_visible_async(map, T, rc->W, rc->H);
  if (r) return r;
  return launch_raycast(rc, map);
}

int yas_raycaster_vmap(yas_raycaster* rc, float** vmap, size_t* pitch) {
  YAS_REQUIRE(rc && vmap, "null argument");
  *vmap = (float*)rc->vmap;
  if (pitch) *pitch = sizeof(float4) * (size_t)rc->W;
  return YAS_OK;
}

int yas_raycaster_zrange(yas_raycaster* rc, float** zmin, float** zmax, int* cols, int* rows) {
  YAS_REQUIRE(rc, "null argument");
  if (zmin) *zmin = (float*)rc->zmin;
  if (zmax) *zmax = (float*)rc->zmax;
  if (cols) *cols = rc->cols;
  if (rows) *rows = rc->rows;
  return YAS_OK;
}

int yas_raycaster_num_visible(yas_raycaster* rc, int* count) {
  YAS_REQUIRE(rc && count, "null argument");
  *count = 0;
  if (!rc->last_map) return YAS_OK;
  MapView v = yas_map_view(rc->last_map);
  YAS_HIP(hipMemcpyAsync(rc->host_ctr, v.ctr, sizeof(int) * kCtrCount, hipMemcpyDeviceToHost, rc->ctx->stream));
  YAS_HIP(hipStreamSynchronize(rc->ctx->stream));
  *count = rc->host_ctr[kCtrVisible];
  return YAS_OK;
}

}  // extern "C"
