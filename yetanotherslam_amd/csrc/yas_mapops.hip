// Volume management around the per-frame path: map-to-map fusion
// (MapStruct::Fuse(MapStruct*), RGBDFusion/src/VoxelMap.cu:400-449), growth
// (MapStruct::Reserve, VoxelMap.cu:485-515), device <-> host paging
// (Hibernate / ReActivate, VoxelMap.cu:894-954), a versioned on-disk format
// (the reference's SaveToFile / ReadFromFile are empty, VoxelMap.cu:886-892)
// and a PLY writer for the indexed mesh.
//
// The open-addressed table is rebuilt from the dense block list wherever the
// reference copies its bucket/chain arrays: growth, re-activation and
// loading move only the allocated blocks (keys, summary flags, voxels), not
// the whole table or pool.
#include <cstdio>
#include <cstring>
#include <vector>

#include "yas_internal.hpp"

using namespace yas;

int yas_map_get_stats(yas_map* m, yas_map_stats* st);

namespace {

__global__ void __launch_bounds__(256) k_fill_slots(HashSlot* slots, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    HashSlot e;
    e.key = kEmptyKey;
    e.val = -1;
    e.flags = 0;
    slots[i] = e;
  }
}

__global__ void __launch_bounds__(256) k_zero_words(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Gather each pool block's summary flags from its hash slot.
__global__ void __launch_bounds__(256) k_gather_flags(const HashSlot* slots, const int* block_slot, int n, int* flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = slots[block_slot[i]].flags;
}

// Insert pool blocks [0, n) (distinct keys) into an empty table: slot =
// first free slot probing from the reference hash's home slot.
__global__ void __launch_bounds__(256) k_rehash(MapView m, const int* flags, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long key = m.block_keys[i];
  int x, y, z;
  unpack_key(key, x, y, z);
  int slot = ref_hash(x, y, z, m.num_buckets);
  const int lim = m.num_buckets < kMaxProbe ? m.num_buckets : kMaxProbe;
  for (int p = 0; p < lim; ++p) {
    unsigned long long old = atomicCAS(&m.slots[slot].key, kEmptyKey, key);
    if (old == kEmptyKey) {
      m.slots[slot].val = i;
      m.slots[slot].flags = flags ? flags[i] : 0;
      m.block_slot[i] = slot;
      return;
    }
    slot = slot + 1 == m.num_buckets ? 0 : slot + 1;
  }
  atomicOr(&m.ctr[kCtrFlags], kFlagProbeLimit);
}

__global__ void k_set_alloc(int* ctr, int n) {
  if (threadIdx.x == 0) ctr[kCtrAlloc] = n;
}

// Step 1 of the merge (VoxelMap.cu:228-259, every voxel visited; the
// reference's 1-D index over an 8x8 launch visits one column in eight):
// one wavefront per source block, lanes (x, y), the 8 z-slices in turn.
// Each observed voxel, moved into the destination frame, allocates its
// block there.
template <class VT>
__global__ void __launch_bounds__(256) k_merge_alloc(MapView dst, MapView src, PoseF Tsd, int nsrc) {
  const int lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  const typename VT::T* sv = (const typename VT::T*)src.voxels;
  const float vs = dst.voxel_size;
  for (int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nsrc; w += nwaves) {
    int bx, by, bz;
    unpack_key(src.block_keys[w], bx, by, bz);
    for (int lz = 0; lz < 8; ++lz) {
      const typename VT::T v = sv[(size_t)w * 512 + lz * 64 + lane];
      bool has = VT::observed(v);
      uint64_t key = 0;
      int home = 0;
      if (has) {
        const float px = (float)(bx * 8 + (lane & 7)) * vs;
        const float py = (float)(by * 8 + (lane >> 3)) * vs;
        const float pz = (float)(bz * 8 + lz) * vs;
        float qx, qy, qz;
        xform(Tsd, px, py, pz, qx, qy, qz);
        const int kx = vox_to_block(floor_sat(qx / vs)), ky = vox_to_block(floor_sat(qy / vs)),
                  kz = vox_to_block(floor_sat(qz / vs));
        if (key_in_range(kx, ky, kz)) {
          key = pack_key(kx, ky, kz);
          home = ref_hash(kx, ky, kz, dst.num_buckets);
        } else {
          has = false;
          atomicOr(&dst.ctr[kCtrFlags], kFlagKeyRange);
        }
      }
      wave_insert(dst, key, has, home);
    }
  }
}

// Voxel reads of the source map with a one-entry block cache; ReadSDF
// (VoxelMap.cu:290-310): observed -> (sdf, weight), else invalid.
template <class VT>
struct MergeReader {
  const MapView& m;
  int cbx, cby, cbz, cblk;
  __device__ explicit MergeReader(const MapView& mv) : m(mv), cbx(INT32_MIN), cby(0), cbz(0), cblk(-1) {}
  __device__ inline bool read(int vx, int vy, int vz, float& sdf, int& wt) {
    const int bx = vox_to_block(vx), by = vox_to_block(vy), bz = vox_to_block(vz);
    if (bx != cbx || by != cby || bz != cbz) {
      cbx = bx; cby = by; cbz = bz;
      cblk = find_block_xyz(m, bx, by, bz);
    }
    if (cblk < 0) return false;
    const typename VT::T v =
        ((const typename VT::T*)m.voxels)[(size_t)cblk * 512 + (vox_to_local(vz) * 8 + vox_to_local(vy)) * 8 +
                                          vox_to_local(vx)];
    if (!VT::observed(v)) return false;
    sdf = VT::sdf(v);
    wt = VT::weight(v);
    return true;
  }
};

// Step 2 (VoxelMap.cu:360-378 with InterpolateSDF :312-358): every voxel of
// every destination block samples the source trilinearly at its position
// moved into the source frame; with all 8 corners observed it merges the
// sample with the first corner's weight.  The block's summary flags are
// re-derived afterwards, as after integration.
template <class VT>
__global__ void __launch_bounds__(256) k_merge_fuse(MapView dst, MapView src, PoseF Tds) {
  const int na = dst.ctr[kCtrAlloc];
  const int nblk = na < dst.num_blocks ? na : dst.num_blocks;
  const int lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  typename VT::T* dv = (typename VT::T*)dst.voxels;
  const float vs = dst.voxel_size;
  MergeReader<VT> rd(src);
  for (int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nblk; w += nwaves) {
    int bx, by, bz;
    unpack_key(dst.block_keys[w], bx, by, bz);
    bool unobserved = true, free_space = true;
    for (int lz = 0; lz < 8; ++lz) {
      typename VT::T* p = dv + (size_t)w * 512 + lz * 64 + lane;
      typename VT::T cur = *p;
      const float px = (float)(bx * 8 + (lane & 7)) * vs;
      const float py = (float)(by * 8 + (lane >> 3)) * vs;
      const float pz = (float)(bz * 8 + lz) * vs;
      float qx, qy, qz;
      xform(Tds, px, py, pz, qx, qy, qz);
      const float fx = qx / vs, fy = qy / vs, fz = qz / vs;
      const int x0 = floor_sat(fx), y0 = floor_sat(fy), z0 = floor_sat(fz);
      const float x = fx - (float)x0, y = fy - (float)y0, z = fz - (float)z0;
      float s[8];
      int wt0 = 0, wt = 0;
      bool ok = true;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (ok) ok = rd.read(x0 + (c & 1), y0 + ((c >> 1) & 1), z0 + ((c >> 2) & 1), s[c], wt);
        if (c == 0) wt0 = wt;
      }
      if (ok) {
        float s2 = s[0] * (1.0f - x) + s[1] * x;
        const float s3 = s2 * (1.0f - y) + (s[2] * (1.0f - x) + s[3] * x) * y;
        s2 = s[4] * (1.0f - x) + s[5] * x;
        const float sdf = s3 * (1.0f - z) + (s2 * (1.0f - y) + (s[6] * (1.0f - x) + s[7] * x) * y) * z;
        cur = VT::merge(cur, sdf, wt0, dst.max_weight);
        *p = cur;
      }
      const bool obs = VT::observed(cur);
      unobserved = unobserved && !obs;
      free_space = free_space && obs && VT::sdf(cur) == 1.0f;
    }
    const unsigned long long un = __ballot(unobserved), fr = __ballot(free_space);
    if (lane == 0) {
      const int f = (un == ~0ull ? kBlockUnobserved : 0) | (fr == ~0ull ? kBlockFree : 0);
      dst.slots[dst.block_slot[w]].flags = f;
    }
  }
}

int grid_of(long n, int per, int cap) {
  long g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

int allocated_blocks(yas_map* m, int* n) {
  yas_map_stats st;
  int rc = yas_map_get_stats(m, &st);
  if (rc) return rc;
  *n = st.num_allocated < m->P.num_blocks ? st.num_allocated : m->P.num_blocks;
  return YAS_OK;
}

// Allocate the device arrays of a map for its current parameters (slots,
// block list, pool, visible list); nothing is initialised.
int alloc_device_arrays(yas_map* m, MapView& v) {
  const size_t vb = (size_t)voxel_bytes(m->P.voxel_format);
  hipError_t e = hipSuccess;
  auto alloc = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  v.slots = nullptr;
  v.block_keys = nullptr;
  v.block_slot = nullptr;
  v.voxels = nullptr;
  v.visible = nullptr;
  alloc((void**)&v.slots, sizeof(HashSlot) * m->P.num_buckets);
  alloc((void**)&v.block_keys, sizeof(unsigned long long) * m->P.num_blocks);
  alloc((void**)&v.block_slot, sizeof(int) * m->P.num_blocks);
  alloc((void**)&v.voxels, vb * 512 * (size_t)m->P.num_blocks);
  alloc((void**)&v.visible, sizeof(int) * m->P.num_blocks);
  if (e != hipSuccess) {
    hipFree(v.slots);
    hipFree(v.block_keys);
    hipFree(v.block_slot);
    hipFree(v.voxels);
    hipFree(v.visible);
    yas_set_error("map allocation failed: %s", hipGetErrorString(e));
    return YAS_E_HIP;
  }
  return YAS_OK;
}

void free_device_arrays(MapView& v) {
  hipFree(v.slots);
  hipFree(v.block_keys);
  hipFree(v.block_slot);
  hipFree(v.voxels);
  hipFree(v.visible);
  v.slots = nullptr;
  v.block_keys = nullptr;
  v.block_slot = nullptr;
  v.voxels = nullptr;
  v.visible = nullptr;
}

// With the device arrays of `v` allocated and blocks [0, n) of keys/voxels
// in place: zero the rest of the pool, clear the table, rebuild it from the
// keys (with `flags_dev`, or none), set the allocation counter.
int rebuild(yas_map* m, MapView& v, const int* flags_dev, int n) {
  hipStream_t s = m->ctx->stream;
  const size_t vb = (size_t)voxel_bytes(m->P.voxel_format);
  const size_t tail_words = (size_t)(m->P.num_blocks - n) * 512 * vb / 4;
  if (tail_words) {
    hipLaunchKernelGGL(k_zero_words, dim3(grid_of((long)(tail_words / 4), 256, 16384)), dim3(256), 0, s,
                       (uint32_t*)((char*)v.voxels + (size_t)n * 512 * vb), tail_words);
    YAS_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_fill_slots, dim3(grid_of(m->P.num_buckets, 256, 8192)), dim3(256), 0, s, v.slots,
                     m->P.num_buckets);
  YAS_LAUNCH_CHECK();
  if (n > 0) {
    hipLaunchKernelGGL(k_rehash, dim3((n + 255) / 256), dim3(256), 0, s, v, flags_dev, n);
    YAS_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_set_alloc, dim3(1), dim3(64), 0, s, v.ctr, n);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

// PLY (binary little endian) with float xyz vertices and int32 triangles.
int write_ply(const char* path, const float* verts, int nv, const int32_t* tris, int nt) {
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    yas_set_error("cannot open %s for writing", path);
    return YAS_E_INVALID;
  }
  std::fprintf(f,
               "ply\nformat binary_little_endian 1.0\ncomment yas marching cubes\nelement vertex %d\n"
               "property float x\nproperty float y\nproperty float z\nelement face %d\n"
               "property list uchar int vertex_indices\nend_header\n",
               nv, nt);
  bool ok = std::fwrite(verts, sizeof(float) * 3, (size_t)nv, f) == (size_t)nv;
  std::vector<unsigned char> rec(13);
  for (int i = 0; ok && i < nt; ++i) {
    rec[0] = 3;
    std::memcpy(&rec[1], tris + 3 * (size_t)i, 12);
    ok = std::fwrite(rec.data(), 1, 13, f) == 13;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    yas_set_error("write to %s failed", path);
    return YAS_E_INVALID;
  }
  return YAS_OK;
}

// On-disk map: a fixed header, then per allocated block its key (3 x int32)
// and summary flags (int32), then the voxels of those blocks in pool order
// (the storage type's little-endian bytes).
struct MapFileHeader {
  char magic[8];  // "YASMAP\0\1"
  uint32_t version;
  int32_t voxel_format;
  int32_t num_buckets, num_blocks, num_allocated, max_weight;
  float voxel_size, trunc_dist, depth_min, depth_max;
  float K[4];
  double pose[16];
  uint32_t reserved[8];
};
constexpr char kMagic[8] = {'Y', 'A', 'S', 'M', 'A', 'P', 0, 1};

}  // namespace

int yas_mesher_indexed_host(yas_mesher* me, std::vector<float>& verts, std::vector<int32_t>& tris);

extern "C" {

int yas_map_reserve(yas_map* m, int num_buckets, int num_blocks) {
  YAS_REQUIRE(m, "yas_map_reserve: null map");
  YAS_REQUIRE_ACTIVE(m);
  YAS_REQUIRE(num_buckets >= 0 && num_blocks >= 0, "yas_map_reserve: negative size");
  if (num_buckets <= m->P.num_buckets && num_blocks <= m->P.num_blocks) return YAS_OK;
  YAS_HIP(hipSetDevice(m->ctx->device));
  int n = 0;
  int rc = allocated_blocks(m, &n);
  if (rc) return rc;
  hipStream_t s = m->ctx->stream;
  MapView old = m->v;
  int* flags = nullptr;
  YAS_HIP(hipMalloc((void**)&flags, sizeof(int) * (size_t)(n > 0 ? n : 1)));
  if (n > 0) {
    hipLaunchKernelGGL(k_gather_flags, dim3((n + 255) / 256), dim3(256), 0, s, (const HashSlot*)old.slots,
                       (const int*)old.block_slot, n, flags);
    YAS_LAUNCH_CHECK();
  }
  const yas_map_params oldP = m->P;
  if (num_buckets > m->P.num_buckets) m->P.num_buckets = num_buckets;
  if (num_blocks > m->P.num_blocks) m->P.num_blocks = num_blocks;
  MapView nv = old;
  if ((rc = alloc_device_arrays(m, nv))) {
    m->P = oldP;
    hipFree(flags);
    return rc;
  }
  nv.num_buckets = m->P.num_buckets;
  nv.num_blocks = m->P.num_blocks;
  const size_t vb = (size_t)voxel_bytes(m->P.voxel_format);
  if (n > 0) {
    YAS_HIP(hipMemcpyAsync(nv.block_keys, old.block_keys, sizeof(unsigned long long) * n, hipMemcpyDeviceToDevice, s));
    YAS_HIP(hipMemcpyAsync(nv.voxels, old.voxels, vb * 512 * (size_t)n, hipMemcpyDeviceToDevice, s));
  }
  if ((rc = rebuild(m, nv, flags, n)) == YAS_OK && hipStreamSynchronize(s) != hipSuccess) {
    yas_set_error("yas_map_reserve: %s", hipGetErrorString(hipGetLastError()));
    rc = YAS_E_HIP;
  }
  hipFree(flags);
  if (rc) {
    free_device_arrays(nv);
    m->P = oldP;
    return rc;
  }
  free_device_arrays(old);
  m->v = nv;
  return YAS_OK;
}

int yas_map_fuse_map(yas_map* dst, yas_map* src) {
  YAS_REQUIRE(dst && src && dst != src, "yas_map_fuse_map: null or identical maps");
  YAS_REQUIRE_ACTIVE(dst);
  YAS_REQUIRE_ACTIVE(src);
  YAS_REQUIRE(dst->ctx->stream == src->ctx->stream, "yas_map_fuse_map: maps must share a stream");
  YAS_REQUIRE(dst->P.voxel_format == src->P.voxel_format, "yas_map_fuse_map: voxel formats differ");
  YAS_REQUIRE(dst->P.voxel_size == src->P.voxel_size, "yas_map_fuse_map: voxel sizes differ");
  YAS_HIP(hipSetDevice(dst->ctx->device));
  int nsrc = 0;
  int rc = allocated_blocks(src, &nsrc);
  if (rc) return rc;
  // VoxelMap.cu:412-417: grow by the source's capacities first.
  if ((rc = yas_map_reserve(dst, dst->P.num_buckets + src->P.num_buckets, dst->P.num_blocks + src->P.num_blocks)))
    return rc;
  double Td[16], Ts[16], inv[16], Tsd[16], Tds[16];
  {
    std::lock_guard<std::mutex> a(dst->pose_mutex);
    std::memcpy(Td, dst->pose, sizeof(Td));
  }
  {
    std::lock_guard<std::mutex> b(src->pose_mutex);
    std::memcpy(Ts, src->pose, sizeof(Ts));
  }
  inv44(Td, inv);
  mul44(inv, Ts, Tsd);  // source frame -> destination frame
  inv44(Ts, inv);
  mul44(inv, Td, Tds);  // destination frame -> source frame
  PosePair psd, pds;
  pose_pair_from(Tsd, psd);
  pose_pair_from(Tds, pds);
  hipStream_t s = dst->ctx->stream;
  if (nsrc > 0) {
    const int g = grid_of(nsrc, 4, 4096);
    switch (dst->P.voxel_format) {
      case YAS_VOXEL_I16_U8:
        hipLaunchKernelGGL(k_merge_alloc<VoxI16>, dim3(g), dim3(256), 0, s, dst->v, src->v, psd.T, nsrc);
        break;
      case YAS_VOXEL_F16_U16:
        hipLaunchKernelGGL(k_merge_alloc<VoxF16>, dim3(g), dim3(256), 0, s, dst->v, src->v, psd.T, nsrc);
        break;
      default:
        hipLaunchKernelGGL(k_merge_alloc<VoxF32>, dim3(g), dim3(256), 0, s, dst->v, src->v, psd.T, nsrc);
    }
    YAS_LAUNCH_CHECK();
  }
  const int g = grid_of(dst->P.num_blocks, 4, 4096);
  switch (dst->P.voxel_format) {
    case YAS_VOXEL_I16_U8:
      hipLaunchKernelGGL(k_merge_fuse<VoxI16>, dim3(g), dim3(256), 0, s, dst->v, src->v, pds.T);
      break;
    case YAS_VOXEL_F16_U16:
      hipLaunchKernelGGL(k_merge_fuse<VoxF16>, dim3(g), dim3(256), 0, s, dst->v, src->v, pds.T);
      break;
    default:
      hipLaunchKernelGGL(k_merge_fuse<VoxF32>, dim3(g), dim3(256), 0, s, dst->v, src->v, pds.T);
  }
  YAS_LAUNCH_CHECK();
  yas_map_stats st;
  if ((rc = yas_map_get_stats(dst, &st))) return rc;
  if (st.flags & kFlagPoolOverflow) {
    yas_set_error("yas_map_fuse_map: block pool exhausted (%d blocks)", dst->P.num_blocks);
    return YAS_E_OUT_OF_BLOCKS;
  }
  if (st.flags & (kFlagProbeLimit | kFlagKeyRange)) {
    yas_set_error("yas_map_fuse_map: hash table full or block key out of range");
    return YAS_E_HASH_FULL;
  }
  return YAS_OK;
}

int yas_map_hibernate(yas_map* m) {
  YAS_REQUIRE(m, "yas_map_hibernate: null map");
  if (m->hibernating) return YAS_OK;
  YAS_HIP(hipSetDevice(m->ctx->device));
  int n = 0;
  int rc = allocated_blocks(m, &n);
  if (rc) return rc;
  hipStream_t s = m->ctx->stream;
  const size_t vb = (size_t)voxel_bytes(m->P.voxel_format);
  const size_t nn = (size_t)(n > 0 ? n : 1);
  int* flags = nullptr;
  YAS_HIP(hipMalloc((void**)&flags, sizeof(int) * nn));
  hipError_t e = hipHostMalloc((void**)&m->hib_keys, sizeof(unsigned long long) * nn, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&m->hib_flags, sizeof(int) * nn, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&m->hib_voxels, vb * 512 * nn, hipHostMallocDefault);
  if (e != hipSuccess) {
    hipFree(flags);
    hipHostFree(m->hib_keys);
    hipHostFree(m->hib_flags);
    hipHostFree(m->hib_voxels);
    m->hib_keys = nullptr;
    m->hib_flags = nullptr;
    m->hib_voxels = nullptr;
    yas_set_error("yas_map_hibernate: host allocation failed: %s", hipGetErrorString(e));
    return YAS_E_HIP;
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_gather_flags, dim3((n + 255) / 256), dim3(256), 0, s, (const HashSlot*)m->v.slots,
                       (const int*)m->v.block_slot, n, flags);
    YAS_LAUNCH_CHECK();
    YAS_HIP(hipMemcpyAsync(m->hib_keys, m->v.block_keys, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, s));
    YAS_HIP(hipMemcpyAsync(m->hib_flags, flags, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    YAS_HIP(hipMemcpyAsync(m->hib_voxels, m->v.voxels, vb * 512 * (size_t)n, hipMemcpyDeviceToHost, s));
  }
  YAS_HIP(hipStreamSynchronize(s));
  hipFree(flags);
  free_device_arrays(m->v);
  m->hib_blocks = n;
  m->hibernating = true;
  return YAS_OK;
}

int yas_map_reactivate(yas_map* m) {
  YAS_REQUIRE(m, "yas_map_reactivate: null map");
  if (!m->hibernating) return YAS_OK;
  YAS_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  const int n = m->hib_blocks;
  const size_t vb = (size_t)voxel_bytes(m->P.voxel_format);
  MapView nv = m->v;
  int rc = alloc_device_arrays(m, nv);
  if (rc) return rc;
  int* flags = nullptr;
  YAS_HIP(hipMalloc((void**)&flags, sizeof(int) * (size_t)(n > 0 ? n : 1)));
  if (n > 0) {
    YAS_HIP(hipMemcpyAsync(nv.block_keys, m->hib_keys, sizeof(unsigned long long) * n, hipMemcpyHostToDevice, s));
    YAS_HIP(hipMemcpyAsync(flags, m->hib_flags, sizeof(int) * n, hipMemcpyHostToDevice, s));
    YAS_HIP(hipMemcpyAsync(nv.voxels, m->hib_voxels, vb * 512 * (size_t)n, hipMemcpyHostToDevice, s));
  }
  if ((rc = rebuild(m, nv, flags, n)) == YAS_OK && hipStreamSynchronize(s) != hipSuccess) {
    yas_set_error("yas_map_reactivate: %s", hipGetErrorString(hipGetLastError()));
    rc = YAS_E_HIP;
  }
  hipFree(flags);
  if (rc) {
    free_device_arrays(nv);  // still hibernating; the host copy is intact
    return rc;
  }
  hipHostFree(m->hib_keys);
  hipHostFree(m->hib_flags);
  hipHostFree(m->hib_voxels);
  m->hib_keys = nullptr;
  m->hib_flags = nullptr;
  m->hib_voxels = nullptr;
  m->v = nv;
  m->hibernating = false;
  return YAS_OK;
}

int yas_map_is_hibernating(yas_map* m, int* hibernating) {
  YAS_REQUIRE(m && hibernating, "null argument");
  *hibernating = m->hibernating ? 1 : 0;
  return YAS_OK;
}

int yas_map_save(yas_map* m, const char* path) {
  YAS_REQUIRE(m && path, "yas_map_save: null argument");
  YAS_REQUIRE_ACTIVE(m);
  YAS_HIP(hipSetDevice(m->ctx->device));
  int n = 0;
  int rc = allocated_blocks(m, &n);
  if (rc) return rc;
  hipStream_t s = m->ctx->stream;
  const size_t vb = (size_t)voxel_bytes(m->P.voxel_format);
  std::vector<unsigned long long> keys(n);
  std::vector<int> flags(n);
  std::vector<unsigned char> vox(vb * 512 * (size_t)n);
  if (n > 0) {
    int* fd = nullptr;
    YAS_HIP(hipMalloc((void**)&fd, sizeof(int) * n));
    hipLaunchKernelGGL(k_gather_flags, dim3((n + 255) / 256), dim3(256), 0, s, (const HashSlot*)m->v.slots,
                       (const int*)m->v.block_slot, n, fd);
    YAS_LAUNCH_CHECK();
    YAS_HIP(hipStreamSynchronize(s));
    YAS_HIP(hipMemcpy(keys.data(), m->v.block_keys, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
    YAS_HIP(hipMemcpy(flags.data(), fd, sizeof(int) * n, hipMemcpyDeviceToHost));
    YAS_HIP(hipMemcpy(vox.data(), m->v.voxels, vox.size(), hipMemcpyDeviceToHost));
    hipFree(fd);
  }
  MapFileHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kMagic, 8);
  h.version = 1;
  h.voxel_format = m->P.voxel_format;
  h.num_buckets = m->P.num_buckets;
  h.num_blocks = m->P.num_blocks;
  h.num_allocated = n;
  h.max_weight = m->P.max_weight;
  h.voxel_size = m->P.voxel_size;
  h.trunc_dist = m->P.trunc_dist;
  h.depth_min = m->P.depth_min;
  h.depth_max = m->P.depth_max;
  h.K[0] = m->K.fx;
  h.K[1] = m->K.fy;
  h.K[2] = m->K.cx;
  h.K[3] = m->K.cy;
  {
    std::lock_guard<std::mutex> lk(m->pose_mutex);
    std::memcpy(h.pose, m->pose, sizeof(h.pose));
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    yas_set_error("yas_map_save: cannot open %s", path);
    return YAS_E_INVALID;
  }
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
  for (int i = 0; ok && i < n; ++i) {
    int32_t rec[4];
    unpack_key(keys[i], rec[0], rec[1], rec[2]);
    rec[3] = flags[i];
    ok = std::fwrite(rec, sizeof(rec), 1, f) == 1;
  }
  if (ok && n > 0) ok = std::fwrite(vox.data(), 1, vox.size(), f) == vox.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    yas_set_error("yas_map_save: write to %s failed", path);
    return YAS_E_INVALID;
  }
  return YAS_OK;
}

int yas_map_load(yas_ctx* ctx, const char* path, yas_map** out) {
  YAS_REQUIRE(ctx && path && out, "yas_map_load: null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    yas_set_error("yas_map_load: cannot open %s", path);
    return YAS_E_INVALID;
  }
  MapFileHeader h;
  bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kMagic, 8) == 0 && h.version == 1 &&
            h.num_allocated >= 0 && h.num_allocated <= h.num_blocks;
  if (!ok) {
    std::fclose(f);
    yas_set_error("yas_map_load: %s is not a version-1 map file", path);
    return YAS_E_INVALID;
  }
  const int n = h.num_allocated;
  const size_t vb = (size_t)voxel_bytes(h.voxel_format);
  std::vector<int32_t> rec((size_t)n * 4);
  std::vector<unsigned char> vox(vb * 512 * (size_t)n);
  ok = (n == 0 || std::fread(rec.data(), sizeof(int32_t) * 4, (size_t)n, f) == (size_t)n) &&
       (n == 0 || std::fread(vox.data(), 1, vox.size(), f) == vox.size());
  std::fclose(f);
  if (!ok) {
    yas_set_error("yas_map_load: %s is truncated", path);
    return YAS_E_INVALID;
  }
  yas_map_params P{h.num_buckets, h.num_blocks, h.voxel_size, h.trunc_dist, h.voxel_format, h.max_weight,
                   h.depth_min, h.depth_max};
  yas_map* m = nullptr;
  int rc = yas_map_create(ctx, &P, h.K, &m);
  if (rc) return rc;
  std::vector<unsigned long long> keys(n);
  std::vector<int> flags(n);
  for (int i = 0; i < n; ++i) {
    if (!key_in_range(rec[4 * i], rec[4 * i + 1], rec[4 * i + 2])) {
      yas_map_destroy(m);
      yas_set_error("yas_map_load: block key out of range in %s", path);
      return YAS_E_INVALID;
    }
    keys[i] = pack_key(rec[4 * i], rec[4 * i + 1], rec[4 * i + 2]);
    flags[i] = rec[4 * i + 3];
  }
  int* fd = nullptr;
  hipStream_t s = ctx->stream;
  hipError_t e = hipMalloc((void**)&fd, sizeof(int) * (size_t)(n > 0 ? n : 1));
  // The uploads go on the map's stream, after the clearing kernels that
  // yas_map_create queued there (a copy on the null stream is not ordered
  // with a non-blocking stream and could be overwritten by them).  The host
  // vectors stay alive until the synchronisation below.
  if (e == hipSuccess && n > 0) {
    e = hipMemcpyAsync(m->v.block_keys, keys.data(), sizeof(unsigned long long) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(fd, flags.data(), sizeof(int) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(m->v.voxels, vox.data(), vox.size(), hipMemcpyHostToDevice, s);
  }
  if (e != hipSuccess) {
    hipFree(fd);
    yas_map_destroy(m);
    yas_set_error("yas_map_load: %s", hipGetErrorString(e));
    return YAS_E_HIP;
  }
  rc = rebuild(m, m->v, fd, n);
  if (rc == YAS_OK && hipStreamSynchronize(s) != hipSuccess) rc = YAS_E_HIP;
  hipFree(fd);
  if (rc) {
    yas_map_destroy(m);
    return rc;
  }
  yas_map_set_pose(m, h.pose);
  *out = m;
  return YAS_OK;
}

int yas_mesher_write_ply(yas_mesher* me, const char* path) {
  YAS_REQUIRE(me && path, "yas_mesher_write_ply: null argument");
  std::vector<float> verts;
  std::vector<int32_t> tris;
  int rc = yas_mesher_indexed_host(me, verts, tris);
  if (rc) return rc;
  return write_ply(path, verts.data(), (int)(verts.size() / 3), tris.data(), (int)(tris.size() / 3));
}

}  // extern "C"
