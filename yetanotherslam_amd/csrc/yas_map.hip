// Voxel-hashed TSDF volume on the device: MapStruct's allocation, visibility
// and integration (RGBDFusion/src/VoxelMap.cu:24-58, 517-884) re-designed for
// gfx950.
//
// Layout in HBM:
//   slots[num_buckets]       16 B: u64 packed block coordinate (~0 = free)
//                            and the i32 pool index of its block
//   block_keys[num_blocks]   u64 key of each pool block (dense list of the
//                            allocated blocks; visibility scans this list
//                            instead of the whole table)
//   voxels[num_blocks*512]   4 or 8 bytes per voxel, index z*64 + y*8 + x, so
//                            one z-slice of a block is 64 consecutive voxels:
//                            one coalesced 256/512-byte access per wavefront.
//
// The hash is open addressing with linear probing from the reference hash's
// home slot.  Insertion is lock-free: a 64-bit compare-and-swap claims a free
// slot for a key; within a wavefront, lanes carrying the same block key are
// reduced to one leader with ballots before any memory traffic, and new
// blocks take pool indices from one atomic per wavefront.  The resulting set
// of allocated blocks does not depend on the order of insertion.
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "yas_internal.hpp"
#include "yas_fusion_dev.hpp"

using namespace yas;

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

void yas_set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

extern "C" {

const char* yas_version(void) { return "yas-mi355x 0.1.0 (gfx950)"; }
const char* yas_last_error(void) { return g_last_error.c_str(); }

int yas_ctx_create(int device, void* stream, yas_ctx** out) {
  YAS_REQUIRE(out != nullptr, "yas_ctx_create: out is NULL");
  int n = 0;
  YAS_HIP(hipGetDeviceCount(&n));
  YAS_REQUIRE(device >= 0 && device < n, "yas_ctx_create: no such device");
  YAS_HIP(hipSetDevice(device));
  yas_ctx* c = new yas_ctx();
  c->device = device;
  c->refs.store(1);
  if (stream) {
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      yas_set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
      return YAS_E_HIP;
    }
    c->own_stream = true;
  }
  *out = c;
  return YAS_OK;
}

int yas_ctx_create_null_stream(int device, yas_ctx** out) {
  YAS_REQUIRE(out != nullptr, "yas_ctx_create_null_stream: out is NULL");
  int n = 0;
  YAS_HIP(hipGetDeviceCount(&n));
  YAS_REQUIRE(device >= 0 && device < n, "yas_ctx_create_null_stream: no such device");
  YAS_HIP(hipSetDevice(device));
  yas_ctx* c = new yas_ctx();
  c->device = device;
  c->refs.store(1);
  c->stream = nullptr;  // the device's null stream
  c->own_stream = false;
  *out = c;
  return YAS_OK;
}

// Drops the creator's reference; the context is freed once the objects
// built on it are destroyed too.
int yas_ctx_destroy(yas_ctx* ctx) {
  yas_ctx_release(ctx);
  return YAS_OK;
}

int yas_ctx_synchronize(yas_ctx* ctx) {
  YAS_REQUIRE(ctx, "null context");
  YAS_HIP(hipStreamSynchronize(ctx->stream));
  return YAS_OK;
}

void* yas_ctx_stream(yas_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

}  // extern "C"

yas_ctx* yas_ctx_retain(yas_ctx* c) {
  if (c) c->refs.fetch_add(1);
  return c;
}

void yas_ctx_release(yas_ctx* c) {
  if (!c || c->refs.fetch_sub(1) != 1) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  if (c->own_stream) hipStreamDestroy(c->stream);
  delete c;
}

// ------------------------------------------------------------------ kernels

__global__ void __launch_bounds__(256) k_reset_hash(HashSlot* slots, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    HashSlot e;
    e.key = kEmptyKey;
    e.val = -1;
    e.flags = 0;
    slots[i] = e;
  }
}

__global__ void __launch_bounds__(256) k_zero_u32(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Writes the float pose pair of a by-value double pose and clears the
// visible-block counter for the call that follows.
__global__ void k_pose_prep(Mat44 T, PosePair* out, int* ctr) {
  if (threadIdx.x == 0) {
    PosePair p;
    pose_pair_from(T.m, p);
    *out = p;
    if (ctr) ctr[kCtrVisible] = 0;
  }
}

// Same from a device-resident double pose (used by the pipeline).
__global__ void k_pose_prep_dev(const double* T, PosePair* out, int* ctr) {
  if (threadIdx.x == 0) {
    PosePair p;
    pose_pair_from(T, p);
    *out = p;
    if (ctr) ctr[kCtrVisible] = 0;
  }
}

__global__ void __launch_bounds__(256) k_depth_u16(const uint16_t* src, size_t sp, float* dst, size_t dp, int W,
                                                     int H, float scale) {
  int x = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  uint16_t d = ((const uint16_t*)((const char*)src + y * sp))[x];
  ((float*)((char*)dst + y * dp))[x] = (float)d * scale;
}

// K3 (VoxelMap.cu:517-652), 16x16 pixels per workgroup: a wavefront covers
// a 16x4 patch whose rays mostly share blocks.
__global__ void __launch_bounds__(256) k_alloc(MapView m, const float* depth, size_t pitch, int W, int H, Intr K,
                                               const PosePair* pose, const int* gate) {
  if (gate && *gate == 0) return;
  __shared__ int walked;
  if (threadIdx.x == 0) walked = 0;
  __syncthreads();
  alloc_walk(m, DepthF32{depth, pitch}, W, H, K, pose->T, blockIdx.x * 16 + (threadIdx.x & 15),
             blockIdx.y * 16 + (threadIdx.x >> 4), &walked);
  __syncthreads();
  if (threadIdx.x == 0 && walked) atomicAdd(&m.ctr[kCtrWalk], walked);
}

// K4 (VoxelMap.cu:654-717).
__global__ void __launch_bounds__(256) k_visible(MapView m, const PosePair* pose, Intr K, int W, int H,
                                                 const int* gate) {
  if (gate && *gate == 0) return;
  visible_scan(m, pose->Tinv, K, W, H, blockIdx.x * blockDim.x, gridDim.x * blockDim.x);
}

// K5 (VoxelMap.cu:719-786).  `stats` (optional) receives the map's first
// four counters (allocated, visible, flags, walk steps), final once
// allocation and visibility have run.
template <class VT>
// At least six wavefronts per SIMD.  With the block index made
// wavefront-uniform and 32-bit depth offsets every storage's kernel fits
// in 64 registers (eight wavefronts, no spills); asking for eight with
// 64-bit depth addresses spilled the float storage's kernel and slowed it.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_integrate(MapView m, const PosePair* pose, Intr K, const float* depth, size_t pitch, int W, int H, int* stats) {
  if (stats && blockIdx.x == 0 && threadIdx.x < 4) stats[threadIdx.x] = m.ctr[threadIdx.x];
  integrate_waves<VT>(m, pose->Tinv, K, DepthF32{depth, pitch}, W, H, integrate_wave(m, blockIdx.x, gridDim.x),
                      gridDim.x * (blockDim.x >> 6));
}

// ------------------------------------------------------------------ host side

static int grid_for(long n, int per, int cap) {
  long g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

static int launch_integrate(yas_map* map, const float* depth, size_t pitch, int W, int H,
                            const hipEvent_t* ev = nullptr, int* stats = nullptr) {
  hipStream_t s = map->ctx->stream;
  // enough wavefronts for one visible block each up to 32768 visible blocks
  // (those past the visible count return at once)
  dim3 grid(grid_for(map->P.num_blocks, 4, 8192));
  const PosePair* pose = map->pose_dev;
  switch (map->P.voxel_format) {
    case YAS_VOXEL_I16_U8:
      launch_timed(k_integrate<VoxI16>, grid, dim3(256), s, ev, map->v, pose, map->K, depth, pitch, W, H, stats);
      break;
    case YAS_VOXEL_F16_U16:
      launch_timed(k_integrate<VoxF16>, grid, dim3(256), s, ev, map->v, pose, map->K, depth, pitch, W, H, stats);
      break;
    default:
      launch_timed(k_integrate<VoxF32>, grid, dim3(256), s, ev, map->v, pose, map->K, depth, pitch, W, H, stats);
  }
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

// Allocation + visibility + integration with the pose already in pose_dev.
// The three fusion stages with the pose already in pose_dev.  A non-null
// `gate` pointing at 0 turns allocation and visibility into no-ops on the
// device (integration then sees no visible blocks).
int yas_map_alloc_stage(yas_map* map, const float* depth, size_t pitch, int W, int H, const int* gate,
                        const hipEvent_t* ev = nullptr) {
  YAS_REQUIRE(depth_span_fits(pitch, W, H), "depth image spans 4 GiB or more");
  dim3 g((W + 15) / 16, (H + 15) / 16);
  launch_timed(k_alloc, g, dim3(256), map->ctx->stream, ev, map->v, depth, pitch, W, H, map->K,
               (const PosePair*)map->pose_dev, gate);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_map_visible_stage(yas_map* map, int W, int H, const int* gate, const hipEvent_t* ev = nullptr) {
  launch_timed(k_visible, dim3(grid_for(map->P.num_blocks, 256, 4096)), dim3(256), map->ctx->stream, ev, map->v,
               (const PosePair*)map->pose_dev, map->K, W, H, gate);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_map_integrate_stage(yas_map* map, const float* depth, size_t pitch, int W, int H,
                            const hipEvent_t* ev = nullptr, int* stats = nullptr) {
  YAS_REQUIRE(depth_span_fits(pitch, W, H), "depth image spans 4 GiB or more");
  return launch_integrate(map, depth, pitch, W, H, ev, stats);
}
int yas_map_fuse_prepared(yas_map* map, const float* depth, size_t pitch, int W, int H, const int* gate) {
  int rc = yas_map_alloc_stage(map, depth, pitch, W, H, gate);
  if (!rc) rc = yas_map_visible_stage(map, W, H, gate);
  if (!rc) rc = yas_map_integrate_stage(map, depth, pitch, W, H);
  return rc;
}

int yas_map_prepare_pose_dev(yas_map* map, const double* T_dev) {
  hipLaunchKernelGGL(k_pose_prep_dev, dim3(1), dim3(64), 0, map->ctx->stream, T_dev, map->pose_dev, map->v.ctr);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

static int prepare_pose(yas_map* map, const double T[16], bool reset_visible) {
  Mat44 M;
  std::memcpy(M.m, T, sizeof(M.m));
  hipLaunchKernelGGL(k_pose_prep, dim3(1), dim3(64), 0, map->ctx->stream, M, map->pose_dev,
                     reset_visible ? map->v.ctr : (int*)nullptr);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

MapView yas_map_view(yas_map* map) { return map->v; }
const PosePair* yas_map_pose_dev(yas_map* map) { return map->pose_dev; }
PosePair* yas_map_pose_dev_mut(yas_map* map) { return map->pose_dev; }
Intr yas_map_intr(yas_map* map) { return map->K; }
yas_ctx* yas_map_ctx(yas_map* map) { return map->ctx; }

// Visibility only (used by the ray caster): clears and refills the visible list.
int yas_map_visible_async(yas_map* map, const double T[16], int W, int H) {
  int rc = prepare_pose(map, T, true);
  if (rc) return rc;
  hipLaunchKernelGGL(k_visible, dim3(grid_for(map->P.num_blocks, 256, 4096)), dim3(256), 0, map->ctx->stream,
                     map->v, (const PosePair*)map->pose_dev, map->K, W, H, (const int*)nullptr);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
// Visibility with the pose pair already in pose_dev and the visible counter
// already cleared.
int yas_map_visible_prepared(yas_map* map, int W, int H) {
  hipLaunchKernelGGL(k_visible, dim3(grid_for(map->P.num_blocks, 256, 4096)), dim3(256), 0, map->ctx->stream,
                     map->v, (const PosePair*)map->pose_dev, map->K, W, H, (const int*)nullptr);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}
int yas_map_visible_async_dev(yas_map* map, const double* T_dev, int W, int H) {
  int rc = yas_map_prepare_pose_dev(map, T_dev);
  if (rc) return rc;
  hipLaunchKernelGGL(k_visible, dim3(grid_for(map->P.num_blocks, 256, 4096)), dim3(256), 0, map->ctx->stream,
                     map->v, (const PosePair*)map->pose_dev, map->K, W, H, (const int*)nullptr);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

int yas_convert_depth_u16(yas_ctx* ctx, const uint16_t* src, size_t sp, float* dst, size_t dp, int W, int H,
                          float depth_scale) {
  float s = (float)(1.0 / (double)depth_scale);
  hipLaunchKernelGGL(k_depth_u16, dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, ctx->stream, src, sp, dst, dp, W, H, s);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

extern "C" {

int yas_map_create(yas_ctx* ctx, const yas_map_params* P, const float K[4], yas_map** out) {
  YAS_REQUIRE(ctx && P && K && out, "yas_map_create: null argument");
  YAS_REQUIRE(P->num_buckets > 0 && P->num_blocks > 0, "yas_map_create: sizes must be positive");
  YAS_REQUIRE(P->voxel_size > 0.f && P->trunc_dist > 0.f, "yas_map_create: voxel_size and trunc_dist must be > 0");
  YAS_REQUIRE(P->voxel_format >= 0 && P->voxel_format <= 2, "yas_map_create: unknown voxel format");
  YAS_REQUIRE(P->max_weight >= 1 && P->max_weight <= (P->voxel_format == YAS_VOXEL_I16_U8 ? 255 : 65535),
              "yas_map_create: max_weight out of range for the voxel format");
  YAS_HIP(hipSetDevice(ctx->device));
  yas_map* m = new yas_map();
  m->ctx = yas_ctx_retain(ctx);
  m->P = *P;
  m->K = make_intr(K);
  for (int i = 0; i < 16; ++i) m->pose[i] = (i % 5 == 0) ? 1.0 : 0.0;
  MapView& v = m->v;
  std::memset(&v, 0, sizeof(v));
  v.num_buckets = P->num_buckets;
  v.num_blocks = P->num_blocks;
  v.voxel_size = P->voxel_size;
  v.trunc = P->trunc_dist;
  v.depth_min = P->depth_min;
  v.depth_max = P->depth_max;
  v.max_weight = P->max_weight;
  v.format = P->voxel_format;
  size_t vb = (size_t)voxel_bytes(P->voxel_format);
  hipError_t e = hipSuccess;
  auto alloc = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  alloc((void**)&v.slots, sizeof(HashSlot) * P->num_buckets);
  alloc((void**)&v.block_keys, sizeof(unsigned long long) * P->num_blocks);
  alloc((void**)&v.block_slot, sizeof(int) * P->num_blocks);
  alloc((void**)&v.voxels, vb * 512 * (size_t)P->num_blocks);
  alloc((void**)&v.visible, sizeof(int) * P->num_blocks);
  alloc((void**)&v.ctr, sizeof(int) * kCtrCount);
  alloc((void**)&m->pose_dev, sizeof(PosePair));
  if (e == hipSuccess) e = hipHostMalloc((void**)&m->host_ctr, sizeof(int) * kCtrCount, hipHostMallocDefault);
  if (e != hipSuccess) {
    yas_set_error("yas_map_create: allocation failed: %s", hipGetErrorString(e));
    yas_map_destroy(m);
    return YAS_E_HIP;
  }
  int rc = yas_map_reset(m);
  if (rc) {
    yas_map_destroy(m);
    return rc;
  }
  *out = m;
  return YAS_OK;
}

int yas_map_reset(yas_map* m) {
  YAS_REQUIRE(m, "null map");
  YAS_REQUIRE_ACTIVE(m);
  hipStream_t s = m->ctx->stream;
  hipLaunchKernelGGL(k_reset_hash, dim3(grid_for(m->P.num_buckets, 256, 8192)), dim3(256), 0, s, m->v.slots,
                     m->P.num_buckets);
  YAS_LAUNCH_CHECK();
  size_t words = (size_t)voxel_bytes(m->P.voxel_format) * 512 * (size_t)m->P.num_blocks / 4;
  hipLaunchKernelGGL(k_zero_u32, dim3(grid_for((long)(words / 4), 256, 16384)), dim3(256), 0, s, (uint32_t*)m->v.voxels,
                     words);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, (uint32_t*)m->v.ctr, (size_t)kCtrCount);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

int yas_map_destroy(yas_map* m) {
  if (!m) return YAS_OK;
  hipSetDevice(m->ctx->device);
  hipStreamSynchronize(m->ctx->stream);
  hipFree(m->v.slots);
  hipFree(m->v.block_keys);
  hipFree(m->v.block_slot);
  hipFree(m->v.voxels);
  hipFree(m->v.visible);
  hipFree(m->v.ctr);
  hipFree(m->pose_dev);
  hipFree(m->depth_f32);
  if (m->host_ctr) hipHostFree(m->host_ctr);
  if (m->hib_keys) hipHostFree(m->hib_keys);
  if (m->hib_flags) hipHostFree(m->hib_flags);
  if (m->hib_voxels) hipHostFree(m->hib_voxels);
  yas_ctx_release(m->ctx);
  delete m;
  return YAS_OK;
}

int yas_map_set_pose(yas_map* m, const double T[16]) {
  YAS_REQUIRE(m && T, "null argument");
  std::lock_guard<std::mutex> lk(m->pose_mutex);
  std::memcpy(m->pose, T, sizeof(m->pose));
  return YAS_OK;
}

int yas_map_get_pose(yas_map* m, double T[16]) {
  YAS_REQUIRE(m && T, "null argument");
  std::lock_guard<std::mutex> lk(m->pose_mutex);
  std::memcpy(T, m->pose, sizeof(m->pose));
  return YAS_OK;
}

int yas_map_fuse(yas_map* m, const float* depth, size_t pitch, int W, int H, const double T[16]) {
  YAS_REQUIRE(m && depth && T, "yas_map_fuse: null argument");
  YAS_REQUIRE_ACTIVE(m);
  YAS_REQUIRE(W > 0 && H > 0 && pitch >= (size_t)W * sizeof(float), "yas_map_fuse: bad image size");
  YAS_HIP(hipSetDevice(m->ctx->device));
  int rc = prepare_pose(m, T, true);
  if (rc) return rc;
  return yas_map_fuse_prepared(m, depth, pitch, W, H, nullptr);
}

int yas_map_fuse_u16(yas_map* m, const uint16_t* depth, size_t pitch, int W, int H, float depth_scale,
                     const double T[16]) {
  YAS_REQUIRE(m && depth && T, "yas_map_fuse_u16: null argument");
  YAS_REQUIRE_ACTIVE(m);
  YAS_REQUIRE(W > 0 && H > 0 && pitch >= (size_t)W * sizeof(uint16_t), "yas_map_fuse_u16: bad image size");
  YAS_REQUIRE(depth_scale > 0.f, "yas_map_fuse_u16: depth_scale must be > 0");
  YAS_HIP(hipSetDevice(m->ctx->device));
  size_t need = (size_t)W * H * sizeof(float);
  if (m->depth_f32_bytes < need) {
    YAS_HIP(hipStreamSynchronize(m->ctx->stream));
    hipFree(m->depth_f32);
    m->depth_f32 = nullptr;
    m->depth_f32_bytes = 0;
    YAS_HIP(hipMalloc((void**)&m->depth_f32, need));
    m->depth_f32_bytes = need;
  }
  int rc = yas_convert_depth_u16(m->ctx, depth, pitch, m->depth_f32, (size_t)W * sizeof(float), W, H, depth_scale);
  if (rc) return rc;
  return yas_map_fuse(m, m->depth_f32, (size_t)W * sizeof(float), W, H, T);
}

int yas_map_check_visible_blocks(yas_map* m, int W, int H, const double T[16], int* count) {
  YAS_REQUIRE(m && T && count, "null argument");
  YAS_REQUIRE_ACTIVE(m);
  YAS_HIP(hipSetDevice(m->ctx->device));
  int rc = yas_map_visible_async(m, T, W, H);
  if (rc) return rc;
  YAS_HIP(hipMemcpyAsync(m->host_ctr, m->v.ctr, sizeof(int) * kCtrCount, hipMemcpyDeviceToHost, m->ctx->stream));
  YAS_HIP(hipStreamSynchronize(m->ctx->stream));
  *count = m->host_ctr[kCtrVisible];
  return YAS_OK;
}

// The blocks of the last visibility pass (the reference's visibleTable,
// VoxelMap.h:125, which its engines read directly): keys in list order.
int yas_map_visible_blocks(yas_map* m, int32_t* keys, int capacity, int* count) {
  YAS_REQUIRE(m && count && (keys || capacity == 0), "yas_map_visible_blocks: null argument");
  YAS_REQUIRE_ACTIVE(m);
  YAS_HIP(hipSetDevice(m->ctx->device));
  YAS_HIP(hipMemcpyAsync(m->host_ctr, m->v.ctr, sizeof(int) * kCtrCount, hipMemcpyDeviceToHost, m->ctx->stream));
  YAS_HIP(hipStreamSynchronize(m->ctx->stream));
  const int nv = m->host_ctr[kCtrVisible];
  *count = nv;
  const int n = nv < capacity ? nv : capacity;
  if (n > 0) {
    std::vector<int> idx(n);
    std::vector<unsigned long long> k((size_t)m->P.num_blocks);
    YAS_HIP(hipMemcpy(idx.data(), m->v.visible, sizeof(int) * n, hipMemcpyDeviceToHost));
    YAS_HIP(hipMemcpy(k.data(), m->v.block_keys, sizeof(unsigned long long) * m->P.num_blocks, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) unpack_key(k[idx[i]], keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]);
  }
  return nv > capacity ? YAS_E_CAPACITY : YAS_OK;
}

int yas_map_get_stats(yas_map* m, yas_map_stats* st) {
  YAS_REQUIRE(m && st, "null argument");
  YAS_HIP(hipSetDevice(m->ctx->device));
  YAS_HIP(hipMemcpyAsync(m->host_ctr, m->v.ctr, sizeof(int) * kCtrCount, hipMemcpyDeviceToHost, m->ctx->stream));
  YAS_HIP(hipStreamSynchronize(m->ctx->stream));
  int a = m->host_ctr[kCtrAlloc];
  st->num_allocated = a < m->P.num_blocks ? a : m->P.num_blocks;
  st->num_visible = m->host_ctr[kCtrVisible];
  st->flags = m->host_ctr[kCtrFlags];
  st->reserved = m->host_ctr[kCtrWalk];
  return YAS_OK;
}

int yas_map_voxel_bytes(yas_map* m) { return m ? voxel_bytes(m->P.voxel_format) : 0; }

int yas_map_get_params(yas_map* m, yas_map_params* params, float K[4]) {
  YAS_REQUIRE(m && params, "null argument");
  *params = m->P;
  if (K) {
    K[0] = m->K.fx;
    K[1] = m->K.fy;
    K[2] = m->K.cx;
    K[3] = m->K.cy;
  }
  return YAS_OK;
}

int yas_map_export(yas_map* m, int32_t* keys, void* voxels, int capacity, int* count) {
  YAS_REQUIRE(m && count, "null argument");
  YAS_REQUIRE_ACTIVE(m);
  yas_map_stats st;
  int rc = yas_map_get_stats(m, &st);
  if (rc) return rc;
  *count = st.num_allocated;
  int n = st.num_allocated < capacity ? st.num_allocated : capacity;
  if (n <= 0) return YAS_OK;
  if (keys) {
    std::vector<unsigned long long> k(n);
    YAS_HIP(hipMemcpy(k.data(), m->v.block_keys, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) unpack_key(k[i], keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]);
  }
  if (voxels) {
    size_t bytes = (size_t)voxel_bytes(m->P.voxel_format) * 512 * (size_t)n;
    YAS_HIP(hipMemcpy(voxels, m->v.voxels, bytes, hipMemcpyDeviceToHost));
  }
  return st.num_allocated > capacity ? YAS_E_CAPACITY : YAS_OK;
}

int yas_depth_u16_to_f32(yas_ctx* ctx, const uint16_t* src, size_t sp, float* dst, size_t dp, int W, int H,
                         float depth_scale) {
  YAS_REQUIRE(ctx && src && dst && W > 0 && H > 0 && depth_scale > 0.f, "yas_depth_u16_to_f32: bad argument");
  return yas_convert_depth_u16(ctx, src, sp, dst, dp, W, H, depth_scale);
}

}  // extern "C"
