// Ray casting of a TSDF volume into a camera-frame vertex map: the
// reference's RayTraceEngine (RGBDFusion/src/RayTraceEngine.cu:11-454).
//
// The kernels' bodies are in yas_raycast_dev.hpp.
#include <cstring>
#include <mutex>

#include "yas_internal.hpp"
#include "yas_raycast_dev.hpp"
#include "yas_server_dev.hpp"

using namespace yas;

MapView yas_map_view(yas_map* map);
const PosePair* yas_map_pose_dev(yas_map* map);
yas_ctx* yas_map_ctx(yas_map* map);
int yas_map_visible_async(yas_map* map, const double T[16], int W, int H);
int yas_map_visible_async_dev(yas_map* map, const double* T_dev, int W, int H);
int yas_map_visible_prepared(yas_map* map, int W, int H);

struct yas_raycaster {
  yas_ctx* ctx;
  int W, H, cols, rows;
  Intr K;
  float4* vmap;
  unsigned* zmin;
  unsigned* zmax;
  int* host_ctr;
  yas_map* last_map;
  float* depth;  // VMapToDepth of the last vertex map (lazily allocated)
};

__global__ void __launch_bounds__(256) k_zrange_init(unsigned* zmin, unsigned* zmax, int n) {
  zrange_init_at(zmin, zmax, n, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(256) k_zrange(MapView m, const PosePair* pose, Intr K, int cols, int rows,
                                                unsigned* zmin, unsigned* zmax) {
  zrange_at(m, pose, K, cols, rows, zmin, zmax, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(256) k_zrange_lds(MapView m, const PosePair* pose, Intr K, int cols, int rows,
                                                    unsigned* zmin, unsigned* zmax) {
  extern __shared__ unsigned lz[];
  zrange_lds_at(m, pose, K, cols, rows, zmin, zmax, lz, blockIdx.x, gridDim.x);
}

// Workgroups of the LDS-privatised z-range: one per CU.
constexpr int kZrangeLdsGroups = 256;

template <class VT>
__global__ void __launch_bounds__(256) k_march(MapView m, const PosePair* pose, Intr K, int W, int H, int cols,
                                               const float* zmin, const float* zmax, float4* vmap, float vsinv,
                                               float step_scale) {
  march_at<VT>(m, pose, K, W, H, cols, zmin, zmax, vmap, vsinv, step_scale, blockIdx.x);
}

// z-range tiles then the march (RayTraceEngine.cu:412-450) after the map's
// visibility pass: z-range maps zmin/zmax of (W/8)*(H/8) cells, vertex map
// `vmap` dense W*H float4.  `cells_reset`: the caller has already reset the
// z-range cells on this stream.
static int launch_raycast_to(hipStream_t s, yas_map* map, const Intr& K, int W, int H, unsigned* zmin,
                             unsigned* zmax, float4* vmap, hipEvent_t before_march,
                             const hipEvent_t* march_ev = nullptr, bool cells_reset = false) {
  MapView v = yas_map_view(map);
  const PosePair* pose = yas_map_pose_dev(map);
  const int cols = W / 8, rows = H / 8, ncell = cols * rows;
  if (!cells_reset)
    hipLaunchKernelGGL(k_zrange_init, dim3((ncell + 255) / 256), dim3(256), 0, s, zmin, zmax, ncell);
  if (ncell <= kZrangeLdsCells) {
    hipLaunchKernelGGL(k_zrange_lds, dim3(kZrangeLdsGroups), dim3(256), 2 * ncell * sizeof(unsigned), s, v, pose, K,
                       cols, rows, zmin, zmax);
  } else {
    int g = (v.num_blocks + 255) / 256;
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_zrange, dim3(g), dim3(256), 0, s, v, pose, K, cols, rows, zmin, zmax);
  }
  YAS_LAUNCH_CHECK();
  if (before_march) YAS_HIP(hipEventRecord(before_march, s));
  const float vsinv = (float)(1.0 / (double)v.voxel_size);
  const float step_scale = v.trunc / v.voxel_size;
  dim3 grid(((W + 15) / 16) * ((H + 15) / 16));
  const float* zn = (const float*)zmin;
  const float* zf = (const float*)zmax;
  switch (v.format) {
    case YAS_VOXEL_I16_U8:
      launch_timed(k_march<VoxI16>, grid, dim3(256), s, march_ev, v, pose, K, W, H, cols, zn, zf, vmap, vsinv,
                   step_scale);
      break;
    case YAS_VOXEL_F16_U16:
      launch_timed(k_march<VoxF16>, grid, dim3(256), s, march_ev, v, pose, K, W, H, cols, zn, zf, vmap, vsinv,
                   step_scale);
      break;
    default:
      launch_timed(k_march<VoxF32>, grid, dim3(256), s, march_ev, v, pose, K, W, H, cols, zn, zf, vmap, vsinv,
                   step_scale);
  }
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

static int launch_raycast(yas_raycaster* rc, yas_map* map, hipEvent_t before_march = nullptr,
                          const hipEvent_t* march_ev = nullptr, bool cells_reset = false) {
  int r = launch_raycast_to(rc->ctx->stream, map, rc->K, rc->W, rc->H, rc->zmin, rc->zmax, rc->vmap, before_march,
                            march_ev, cells_reset);
  if (r) return r;
  rc->last_map = map;
  return YAS_OK;
}

// ------------------------------------------------------------------ ray cast without allocation
// yas_raycast_into keeps its only scratch, the two z-range maps, in this
// file-scope device array: zmin in the first kScratchCells words, zmax in
// the next (256 KiB, held on every device the library runs on until the
// process ends).  The kernels get its device address, looked up once per
// device with that device current; the array's name itself is a host-side
// stand-in and is never passed to a kernel.
constexpr int kScratchCells = 32768;
static __device__ __align__(256) unsigned scratch[2 * kScratchCells];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned* g_scratch_dev[kMaxDevices];

static int scratch_on_device(int device, unsigned** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  YAS_REQUIRE(device >= 0 && device < kMaxDevices, "yas_raycast_into: device index out of range");
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      yas_set_error("yas_raycast_into: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return YAS_E_HIP;
    }
    g_scratch_dev[device] = (unsigned*)p;
  }
  *out = g_scratch_dev[device];
  return YAS_OK;
}

// Ray cast with the pose held in device memory (pipeline).
// `before_march` (optional) is recorded between the z-range and march kernels.
int yas_raycast_dev(yas_raycaster* rc, yas_map* map, const double* T_dev, hipEvent_t before_march,
                    const hipEvent_t* march_ev) {
  int r = yas_map_visible_async_dev(map, T_dev, rc->W, rc->H);
  if (r) return r;
  return launch_raycast(rc, map, before_march, march_ev);
}

// Ray cast whose set-up the caller has already enqueued on this stream: the
// map's pose pair written, its visible counter cleared and this ray caster's
// z-range cells reset (the pipeline does all three in its first launch).
int yas_raycast_prepared(yas_raycaster* rc, yas_map* map, hipEvent_t before_march, const hipEvent_t* march_ev) {
  int r = yas_map_visible_prepared(map, rc->W, rc->H);
  if (r) return r;
  return launch_raycast(rc, map, before_march, march_ev, true);
}

float4* yas_raycaster_vmap_dev(yas_raycaster* rc) { return rc->vmap; }
void yas_raycaster_zrange_dev(yas_raycaster* rc, unsigned** zmin, unsigned** zmax, int* ncell) {
  *zmin = rc->zmin;
  *zmax = rc->zmax;
  *ncell = rc->cols * rc->rows;
}

// Buffers and launch constants of this ray caster for the serving engine.
int yas_raycaster_fill(yas_raycaster* rc, const MapView& v, StreamDev& d) {
  d.cols = rc->cols;
  d.rows = rc->rows;
  d.zmin = rc->zmin;
  d.zmax = rc->zmax;
  d.vmap = rc->vmap;
  d.vsinv = (float)(1.0 / (double)v.voxel_size);
  d.step_scale = v.trunc / v.voxel_size;
  int g = (v.num_blocks + 255) / 256;
  d.zr_groups = g > 2048 ? 2048 : g;
  return YAS_OK;
}

extern "C" {

int yas_raycaster_create(yas_ctx* ctx, int W, int H, const float K[4], yas_raycaster** out) {
  YAS_REQUIRE(ctx && K && out, "yas_raycaster_create: null argument");
  YAS_REQUIRE(W >= 8 && H >= 8, "yas_raycaster_create: image too small");
  YAS_REQUIRE(W % 8 == 0 && H % 8 == 0, "yas_raycaster_create: width and height must be multiples of 8");
  YAS_HIP(hipSetDevice(ctx->device));
  yas_raycaster* rc = new yas_raycaster();
  rc->ctx = yas_ctx_retain(ctx);
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
  hipFree(rc->depth);
  if (rc->host_ctr) hipHostFree(rc->host_ctr);
  yas_ctx_release(rc->ctx);
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

int yas_raycast_into(yas_map* map, const double T[16], const float K[4], int W, int H, float* vmap, size_t pitch) {
  YAS_REQUIRE(map && T && K && vmap, "yas_raycast_into: null argument");
  YAS_REQUIRE(W >= 8 && H >= 8 && W % 8 == 0 && H % 8 == 0,
              "yas_raycast_into: width and height must be multiples of 8");
  YAS_REQUIRE((long long)(W / 8) * (H / 8) <= kScratchCells,
              "yas_raycast_into: image exceeds the 32768 z-range cells of the scratch array "
              "((width/8)*(height/8); e.g. 1280x960 fits, 2048x1152 does not)");
  YAS_REQUIRE(pitch == sizeof(float4) * (size_t)W, "yas_raycast_into: the vertex map must be dense (pitch = 16 * width)");
  YAS_REQUIRE_ACTIVE(map);
  yas_ctx* ctx = yas_map_ctx(map);
  YAS_HIP(hipSetDevice(ctx->device));
  unsigned* zr = nullptr;
  int r = scratch_on_device(ctx->device, &zr);
  if (r) return r;
  if ((r = yas_map_visible_async(map, T, W, H))) return r;
  return launch_raycast_to(ctx->stream, map, make_intr(K), W, H, zr, zr + kScratchCells, (float4*)vmap, nullptr);
}

int yas_raycaster_vmap(yas_raycaster* rc, float** vmap, size_t* pitch) {
  YAS_REQUIRE(rc && vmap, "null argument");
  *vmap = (float*)rc->vmap;
  if (pitch) *pitch = sizeof(float4) * (size_t)rc->W;
  return YAS_OK;
}

// GetRayTracingResultDepth (VoxelMap.cu:981-987): VMapToDepth
// (ImageProc.cu:262-285) of the last vertex map into the ray caster's depth
// image (overwritten by the next call; the reference returns a new GpuMat).
// With a host buffer the image is also copied there and the call waits.
int yas_raycaster_depth(yas_raycaster* rc, float** depth, size_t* pitch, float* host, size_t host_pitch) {
  YAS_REQUIRE(rc, "yas_raycaster_depth: null ray caster");
  YAS_REQUIRE(!host || host_pitch >= sizeof(float) * (size_t)rc->W, "yas_raycaster_depth: host pitch too small");
  YAS_HIP(hipSetDevice(rc->ctx->device));
  const size_t row = sizeof(float) * (size_t)rc->W;
  if (!rc->depth) YAS_HIP(hipMalloc((void**)&rc->depth, row * rc->H));
  int r = yas_vmap_to_depth(rc->ctx, (const float*)rc->vmap, sizeof(float4) * (size_t)rc->W, rc->depth, row, rc->W,
                            rc->H);
  if (r) return r;
  if (host) {
    YAS_HIP(hipMemcpy2DAsync(host, host_pitch, rc->depth, row, row, rc->H, hipMemcpyDeviceToHost, rc->ctx->stream));
    YAS_HIP(hipStreamSynchronize(rc->ctx->stream));
  }
  if (depth) *depth = rc->depth;
  if (pitch) *pitch = row;
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
