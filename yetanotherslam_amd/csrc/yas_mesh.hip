// Marching-cubes mesh extraction: the reference's MeshEngine
// (RGBDFusion/src/MeshEngine.cu:7-339, tables RGBDFusion/utils/TriangleTable.h).
//
// MI355X design: one 512-thread workgroup per allocated block, one thread per
// voxel.  The +1 corners of a block's voxels live in up to seven neighbour
// blocks; their pool indices are resolved once per workgroup (seven hash
// probes) into LDS instead of eight probes per voxel.  Extraction is two
// passes with a device prefix scan in between, so every block's triangles
// are contiguous and ordered by voxel index, then table order: the output is
// deterministic, unlike the reference's atomic triangle counter
// (MeshEngine.cu:246).
#include <cstring>
#include <vector>

#include "yas_internal.hpp"

using namespace yas;

MapView yas_map_view(yas_map* map);
yas_ctx* yas_map_ctx(yas_map* map);
int yas_map_get_stats(yas_map* m, yas_map_stats* st);

namespace {

// Triangulation table: bits 0-3 = triangle count, then 4 bits per edge.
__constant__ uint64_t kTriTable[256] = {
#include "mc_tables.inc"
};

// MeshEngine.cu:76-112 corner order and MeshEngine.cu:150-224 edge geometry.
__constant__ int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__constant__ int kEdgeA[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
__constant__ int kEdgeB[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};

// MeshEngine.cu:114-123 (thresholds compared in double: the literal is a double).
__device__ inline float edge_t(float v1, float v2) {
  if (fabs((double)(0.0f - v1)) < 1e-6) return 0.0f;
  if (fabs((double)(0.0f - v2)) < 1e-6) return 1.0f;
  if (fabs((double)(v1 - v2)) < 1e-6) return 0.0f;
  return (0.0f - v1) / (v2 - v1);
}

__device__ inline void edge_offset(int e, float t, float& ox, float& oy, float& oz) {
  switch (e) {
    case 0: ox = t; oy = 0.f; oz = 0.f; break;
    case 1: ox = 1.f; oy = t; oz = 0.f; break;
    case 2: ox = 1.f - t; oy = 1.f; oz = 0.f; break;
    case 3: ox = 0.f; oy = 1.f - t; oz = 0.f; break;
    case 4: ox = t; oy = 0.f; oz = 1.f; break;
    case 5: ox = 1.f; oy = t; oz = 1.f; break;
    case 6: ox = 1.f - t; oy = 1.f; oz = 1.f; break;
    case 7: ox = 0.f; oy = 1.f - t; oz = 1.f; break;
    case 8: ox = 0.f; oy = 0.f; oz = t; break;
    case 9: ox = 1.f; oy = 0.f; oz = t; break;
    case 10: ox = 1.f; oy = 1.f; oz = t; break;
    default: ox = 0.f; oy = 1.f; oz = t; break;
  }
}

// Cube index of this thread's voxel (-1 if a corner is unobserved) and the
// 8 corner values.  nb[k] = pool index of the block at offset
// (k&1, k>>1&1, k>>2&1), -1 if absent.
template <class VT>
__device__ inline int voxel_case(const typename VT::T* vox, const int* nb, int lx, int ly, int lz, float* s) {
  int cube = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    int x = lx + kCorner[c][0], y = ly + kCorner[c][1], z = lz + kCorner[c][2];
    int k = (x >> 3) | ((y >> 3) << 1) | ((z >> 3) << 2);
    int blk = nb[k];
    if (blk < 0) return -1;
    typename VT::T v = vox[(size_t)blk * 512 + ((z & 7) * 8 + (y & 7)) * 8 + (x & 7)];
    if (!VT::observed(v)) return -1;
    s[c] = VT::sdf(v);
    if (s[c] < 0.0f) cube |= 1 << c;
  }
  return cube;
}

__device__ inline void resolve_neighbours(const MapView& m, int blk, int* nb, int& bx, int& by, int& bz) {
  unpack_key(m.block_keys[blk], bx, by, bz);
  if (threadIdx.x < 8) {
    int k = threadIdx.x;
    nb[k] = k == 0 ? blk : find_block_xyz(m, bx + (k & 1), by + ((k >> 1) & 1), bz + ((k >> 2) & 1));
  }
  __syncthreads();
}

// Pass 1: triangles per block (and optional cube indices).
template <class VT>
__global__ void __launch_bounds__(512) k_mc_count(MapView m, int nblocks, int* counts, int16_t* cases) {
  __shared__ int nb[8];
  __shared__ int wsum[8];
  const int blk = blockIdx.x;
  if (blk >= nblocks) return;
  int bx, by, bz;
  resolve_neighbours(m, blk, nb, bx, by, bz);
  const int l = threadIdx.x, lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
  float s[8];
  int cube = voxel_case<VT>((const typename VT::T*)m.voxels, nb, lx, ly, lz, s);
  if (cases) cases[(size_t)blk * 512 + l] = (int16_t)cube;
  int nt = cube > 0 ? (int)(kTriTable[cube] & 0xF) : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nt += __shfl_xor(nt, o, 64);
  if ((l & 63) == 0) wsum[l >> 6] = nt;
  __syncthreads();
  if (l == 0) {
    int t = 0;
    for (int w = 0; w < 8; ++w) t += wsum[w];
    counts[blk] = t;
  }
}

// Exclusive scan of the block counts: 1024-thread single workgroup, chunks
// of 1024 with a running carry.  `total` receives the sum.
__global__ void __launch_bounds__(1024) k_scan(const int* in, int* out, int n, int* total) {
  __shared__ int buf[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < n ? in[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int t = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) out[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Pass 2: emit.  Triangle slot = block offset + exclusive prefix (in voxel
// index order) of the per-voxel counts.
template <class VT>
__global__ void __launch_bounds__(512) k_mc_emit(MapView m, int nblocks, const int* offsets, float* tris,
                                                 float* normals, int cap) {
  __shared__ int nb[8];
  __shared__ int wsum[8];
  const int blk = blockIdx.x;
  if (blk >= nblocks) return;
  int bx, by, bz;
  resolve_neighbours(m, blk, nb, bx, by, bz);
  const int l = threadIdx.x, lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
  float s[8];
  int cube = voxel_case<VT>((const typename VT::T*)m.voxels, nb, lx, ly, lz, s);
  uint64_t word = cube > 0 ? kTriTable[cube] : 0ull;
  int nt = (int)(word & 0xF);
  // inclusive scan within the wavefront
  int inc = nt;
  const int lane = l & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) wsum[l >> 6] = inc;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < (l >> 6); ++w) before += wsum[w];
  if (nt == 0) return;
  int slot = offsets[blk] + before + inc - nt;
  const float vs = m.voxel_size;
  float px = (float)(bx * 8 + lx), py = (float)(by * 8 + ly), pz = (float)(bz * 8 + lz);
  for (int i = 0; i < nt; ++i, ++slot) {
    if (slot >= cap) return;
    float a[3], b[3], c[3];
    float* dst[3] = {a, b, c};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      int e = (int)((word >> (4 + 12 * i + 4 * k)) & 0xF);
      float t = edge_t(s[kEdgeA[e]], s[kEdgeB[e]]);
      float ox, oy, oz;
      edge_offset(e, t, ox, oy, oz);
      dst[k][0] = (px + ox) * vs;
      dst[k][1] = (py + oy) * vs;
      dst[k][2] = (pz + oz) * vs;
    }
    float u0 = b[0] - a[0], u1 = b[1] - a[1], u2 = b[2] - a[2];
    float w0 = c[0] - a[0], w1 = c[1] - a[1], w2 = c[2] - a[2];
    float n0 = u1 * w2 - u2 * w1, n1 = u2 * w0 - u0 * w2, n2 = u0 * w1 - u1 * w0;
    float q = n0 * n0 + n1 * n1 + n2 * n2;
    if (q > 0.0f) {
      float nn = sqrtf(q);
      n0 = n0 / nn; n1 = n1 / nn; n2 = n2 / nn;
    }
    float* T = tris + (size_t)slot * 9;
    T[0] = a[0]; T[1] = a[1]; T[2] = a[2];
    T[3] = b[0]; T[4] = b[1]; T[5] = b[2];
    T[6] = c[0]; T[7] = c[1]; T[8] = c[2];
    if (normals) {
      float* N = normals + (size_t)slot * 9;
      N[0] = N[3] = N[6] = n0;
      N[1] = N[4] = N[7] = n1;
      N[2] = N[5] = N[8] = n2;
    }
  }
}

}  // namespace

struct yas_mesher {
  yas_ctx* ctx;
  int max_tri;
  float* tris;
  float* normals;
  int* counts;
  int* offsets;
  int* total;
  int16_t* cases;
  int block_cap;
  int num_blocks;
  int num_tri;
  yas_map* map;
};

static int ensure_block_cap(yas_mesher* me, int n) {
  if (n <= me->block_cap) return YAS_OK;
  hipFree(me->counts);
  hipFree(me->offsets);
  hipFree(me->cases);
  me->counts = me->offsets = nullptr;
  me->cases = nullptr;
  me->block_cap = 0;
  YAS_HIP(hipMalloc((void**)&me->counts, sizeof(int) * n));
  YAS_HIP(hipMalloc((void**)&me->offsets, sizeof(int) * n));
  YAS_HIP(hipMalloc((void**)&me->cases, sizeof(int16_t) * 512 * (size_t)n));
  me->block_cap = n;
  return YAS_OK;
}

extern "C" {

int yas_mesher_create(yas_ctx* ctx, int max_triangles, yas_mesher** out) {
  YAS_REQUIRE(ctx && out && max_triangles > 0, "yas_mesher_create: bad argument");
  YAS_HIP(hipSetDevice(ctx->device));
  yas_mesher* me = new yas_mesher();
  std::memset(me, 0, sizeof(*me));
  me->ctx = ctx;
  me->max_tri = max_triangles;
  hipError_t e = hipMalloc((void**)&me->tris, sizeof(float) * 9 * (size_t)max_triangles);
  if (e == hipSuccess) e = hipMalloc((void**)&me->normals, sizeof(float) * 9 * (size_t)max_triangles);
  if (e == hipSuccess) e = hipMalloc((void**)&me->total, sizeof(int));
  if (e != hipSuccess) {
    yas_set_error("yas_mesher_create: %s", hipGetErrorString(e));
    yas_mesher_destroy(me);
    return YAS_E_HIP;
  }
  *out = me;
  return YAS_OK;
}

int yas_mesher_destroy(yas_mesher* me) {
  if (!me) return YAS_OK;
  hipSetDevice(me->ctx->device);
  hipStreamSynchronize(me->ctx->stream);
  hipFree(me->tris);
  hipFree(me->normals);
  hipFree(me->counts);
  hipFree(me->offsets);
  hipFree(me->total);
  hipFree(me->cases);
  delete me;
  return YAS_OK;
}

int yas_mesh_extract(yas_mesher* me, yas_map* map, int* num_triangles) {
  YAS_REQUIRE(me && map, "yas_mesh_extract: null argument");
  YAS_REQUIRE(yas_map_ctx(map)->stream == me->ctx->stream, "yas_mesh_extract: map and mesher must share a stream");
  YAS_HIP(hipSetDevice(me->ctx->device));
  yas_map_stats st;
  int rc = yas_map_get_stats(map, &st);
  if (rc) return rc;
  const int n = st.num_allocated;
  me->num_blocks = n;
  me->num_tri = 0;
  me->map = map;
  if (num_triangles) *num_triangles = 0;
  if (n == 0) return YAS_OK;
  if ((rc = ensure_block_cap(me, n))) return rc;
  MapView v = yas_map_view(map);
  hipStream_t s = me->ctx->stream;
  switch (v.format) {
    case YAS_VOXEL_I16_U8:
      hipLaunchKernelGGL(k_mc_count<VoxI16>, dim3(n), dim3(512), 0, s, v, n, me->counts, me->cases);
      break;
    case YAS_VOXEL_F16_U16:
      hipLaunchKernelGGL(k_mc_count<VoxF16>, dim3(n), dim3(512), 0, s, v, n, me->counts, me->cases);
      break;
    default:
      hipLaunchKernelGGL(k_mc_count<VoxF32>, dim3(n), dim3(512), 0, s, v, n, me->counts, me->cases);
  }
  YAS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (const int*)me->counts, me->offsets, n, me->total);
  YAS_LAUNCH_CHECK();
  switch (v.format) {
    case YAS_VOXEL_I16_U8:
      hipLaunchKernelGGL(k_mc_emit<VoxI16>, dim3(n), dim3(512), 0, s, v, n, (const int*)me->offsets, me->tris,
                         me->normals, me->max_tri);
      break;
    case YAS_VOXEL_F16_U16:
      hipLaunchKernelGGL(k_mc_emit<VoxF16>, dim3(n), dim3(512), 0, s, v, n, (const int*)me->offsets, me->tris,
                         me->normals, me->max_tri);
      break;
    default:
      hipLaunchKernelGGL(k_mc_emit<VoxF32>, dim3(n), dim3(512), 0, s, v, n, (const int*)me->offsets, me->tris,
                         me->normals, me->max_tri);
  }
  YAS_LAUNCH_CHECK();
  int total = 0;
  YAS_HIP(hipMemcpyAsync(&total, me->total, sizeof(int), hipMemcpyDeviceToHost, s));
  YAS_HIP(hipStreamSynchronize(s));
  me->num_tri = total;
  if (num_triangles) *num_triangles = total;
  if (total > me->max_tri) {
    yas_set_error("yas_mesh_extract: %d triangles exceed the capacity %d", total, me->max_tri);
    return YAS_E_CAPACITY;
  }
  return YAS_OK;
}

int yas_mesher_download(yas_mesher* me, float* vertices, float* normals, int capacity) {
  YAS_REQUIRE(me, "null mesher");
  int n = me->num_tri < me->max_tri ? me->num_tri : me->max_tri;
  if (n > capacity) n = capacity;
  if (n <= 0) return YAS_OK;
  if (vertices) YAS_HIP(hipMemcpy(vertices, me->tris, sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToHost));
  if (normals) YAS_HIP(hipMemcpy(normals, me->normals, sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToHost));
  return YAS_OK;
}

int yas_mesher_blocks(yas_mesher* me, int32_t* keys, int32_t* first, int32_t* count, int16_t* cases, int capacity,
                      int* num_blocks) {
  YAS_REQUIRE(me, "null mesher");
  if (num_blocks) *num_blocks = me->num_blocks;
  int n = me->num_blocks < capacity ? me->num_blocks : capacity;
  if (n <= 0 || !me->map) return YAS_OK;
  YAS_HIP(hipStreamSynchronize(me->ctx->stream));
  if (keys) {
    MapView v = yas_map_view(me->map);
    std::vector<unsigned long long> k(n);
    YAS_HIP(hipMemcpy(k.data(), v.block_keys, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) unpack_key(k[i], keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]);
  }
  if (first) YAS_HIP(hipMemcpy(first, me->offsets, sizeof(int) * n, hipMemcpyDeviceToHost));
  if (count) YAS_HIP(hipMemcpy(count, me->counts, sizeof(int) * n, hipMemcpyDeviceToHost));
  if (cases) YAS_HIP(hipMemcpy(cases, me->cases, sizeof(int16_t) * 512 * (size_t)n, hipMemcpyDeviceToHost));
  return YAS_OK;
}

}  // extern "C"
