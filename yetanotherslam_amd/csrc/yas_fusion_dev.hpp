// Device bodies of the fusion stages (allocation, visibility, integration),
// shared by the single-map kernels and the batched kernels that run many
// maps in one launch.  Arithmetic follows the conventions of yas_internal.hpp.
#pragma once

#include "yas_internal.hpp"

namespace yas {

// Depth images: float metres, or uint16 raw units converted as
// (float)raw * scale (FullSystem/src/System.cpp:52, the same conversion as
// the standalone u16 -> f32 kernel, so both give identical floats).
// DepthF32 addresses a pixel by a 32-bit byte offset from the image's
// (wavefront-uniform) base, which the loads take as a scalar base plus one
// vector register; the host launchers refuse images whose last pixel lies
// 4 GiB or more past the base (depth_span_fits).
struct DepthF32 {
  const float* p;
  size_t pitch;
  __device__ inline float at(int x, int y) const {
    return *(const float*)((const char*)p + ((unsigned)y * (unsigned)pitch + (unsigned)x * 4u));
  }
};
inline bool depth_span_fits(size_t pitch, int W, int H) {
  return pitch < (1ull << 32) && (unsigned long long)(H - 1) * pitch + (unsigned long long)W * 4u <= 0xFFFFFFFFull;
}
struct DepthU16 {
  const uint16_t* p;
  size_t pitch;
  float scale;
  __device__ inline float at(int x, int y) const {
    return (float)((const uint16_t*)((const char*)p + (size_t)y * pitch))[x] * scale;
  }
};

// Pixel ray -> block (VoxelStructUtils.h:97-110, 28-49).
__device__ inline void pixel_block(int x, int y, float z, const Intr& K, const PoseF& T, float vs, int& bx, int& by,
                                   int& bz) {
  float px = K.invfx * ((float)x - K.cx) * z;
  float py = K.invfy * ((float)y - K.cy) * z;
  float wx, wy, wz;
  xform(T, px, py, z, wx, wy, wz);
  bx = vox_to_block(floor_sat(wx / vs));
  by = vox_to_block(floor_sat(wy / vs));
  bz = vox_to_block(floor_sat(wz / vs));
}

// K3 (VoxelMap.cu:517-652) for pixel (x, y): every block on the integer 3D
// Bresenham walk between the blocks of depth-mu/2 and depth+mu/2 along the
// pixel ray.  All lanes of the wavefront must call it (the insertion is a
// wavefront operation); lanes outside the image pass no blocks.  The walk
// length is added to *walked (LDS).
template <class D>
__device__ inline void alloc_walk(const MapView& m, const D& depth, int W, int H, const Intr& K, const PoseF& T, int x,
                                  int y, int* walked) {
  const float vs = m.voxel_size;
  int px = 0, py = 0, pz = 0;          // current block
  int dax = 0, day = 0, daz = 0;       // unit step along the dominant axis
  int d1x = 0, d1y = 0, d1z = 0;       // unit step along the first other axis
  int d2x = 0, d2y = 0, d2z = 0;       // unit step along the second other axis
  int Ea = 0, E1 = 0, E2 = 0, err1 = 0, err2 = 0;
  int remaining = 0;
  if (x < W && y < H) {
    float d = depth.at(x, y);
    const float half = m.trunc * 0.5f;
    if (!(d != d || d < m.depth_min || d > m.depth_max)) {
      float dn = fmaxf(m.depth_min, d - half);
      float df = fminf(m.depth_max, d + half);
      if (dn < df) {
        int ex, ey, ez;
        pixel_block(x, y, dn, K, T, vs, px, py, pz);
        pixel_block(x, y, df, K, T, vs, ex, ey, ez);
        int dx = ex - px, dy = ey - py, dz = ez - pz;
        int ix = dx < 0 ? -1 : 1, iy = dy < 0 ? -1 : 1, iz = dz < 0 ? -1 : 1;
        int ax = abs(dx), ay = abs(dy), az = abs(dz);
        if (ax >= ay && ax >= az) {
          dax = ix; d1y = iy; d2z = iz;
          Ea = ax << 1; E1 = ay << 1; E2 = az << 1;
          remaining = ax + 1;
        } else if (ay >= ax && ay >= az) {
          day = iy; d1x = ix; d2z = iz;
          Ea = ay << 1; E1 = ax << 1; E2 = az << 1;
          remaining = ay + 1;
        } else {
          // the reference's z branch walks y first, then x (VoxelMap.cu:628-647)
          daz = iz; d1y = iy; d2x = ix;
          Ea = az << 1; E1 = ay << 1; E2 = ax << 1;
          remaining = az + 1;
        }
        err1 = E1 - 1;
        err2 = E2 - 1;
      }
    }
  }
  for (;;) {
    bool has = remaining > 0;
    unsigned long long act = __ballot(has);
    if (act == 0ull) break;
    bool ok = has && key_in_range(px, py, pz);
    if (has && !ok) atomicOr(&m.ctr[kCtrFlags], kFlagKeyRange);
    uint64_t key = ok ? pack_key(px, py, pz) : 0ull;
    int home = ok ? ref_hash(px, py, pz, m.num_buckets) : 0;
    if (lane_id() == __ffsll(act) - 1) atomicAdd(walked, __popcll(act));
    wave_insert(m, key, ok, home);
    if (has) {
      --remaining;
      if (remaining > 0) {
        if (err1 > 0) { px += d1x; py += d1y; pz += d1z; err1 -= Ea; }
        if (err2 > 0) { px += d2x; py += d2y; pz += d2z; err2 -= Ea; }
        err1 += E1;
        err2 += E2;
        px += dax; py += day; pz += daz;
      }
    }
  }
}

// VoxelStructUtils.h:112-147.
__device__ inline bool block_visible(int bx, int by, int bz, const PoseF& Ti, const Intr& K, int W, int H,
                                     float vs, float dmin, float dmax) {
  const float scale = vs * 8.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    float x = (float)(bx + (c & 1)) * scale;
    float y = (float)(by + ((c >> 1) & 1)) * scale;
    float z = (float)(bz + ((c >> 2) & 1)) * scale;
    float qx, qy, qz;
    xform(Ti, x, y, z, qx, qy, qz);
    float u = K.fx * qx / qz + K.cx;
    float v = K.fy * qy / qz + K.cy;
    if (!(u < 0.0f || v < 0.0f || u > (float)(W - 1) || v > (float)(H - 1) || qz < dmin || qz > dmax)) return true;
  }
  return false;
}

// K4 (VoxelMap.cu:654-717): compact the visible blocks of the pool into
// `visible` with one atomic per wavefront.  The workgroup covers pool
// indices base + threadIdx.x for base = start, start + stride, ... (start
// and stride uniform over the workgroup, so wavefronts stay whole).
__device__ inline void visible_scan(const MapView& m, const PoseF& Ti, const Intr& K, int W, int H, int start,
                                    int stride) {
  const int n0 = m.ctr[kCtrAlloc];
  const int n = n0 < m.num_blocks ? n0 : m.num_blocks;
  for (int base = start; base < n; base += stride) {
    int i = base + threadIdx.x;
    bool vis = false;
    if (i < n) {
      int bx, by, bz;
      unpack_key(m.block_keys[i], bx, by, bz);
      vis = block_visible(bx, by, bz, Ti, K, W, H, m.voxel_size, m.depth_min, m.depth_max);
    }
    unsigned long long mask = __ballot(vis);
    if (mask) {
      int first = __ffsll(mask) - 1;
      int b = 0;
      if (lane_id() == first) b = atomicAdd(&m.ctr[kCtrVisible], __popcll(mask));
      b = __shfl(b, first);
      const int slot = b + __popcll(mask & lane_mask_lt());
      if (vis && slot < m.num_blocks) m.visible[slot] = i;
    }
  }
}

// The first visible block of this thread's wavefront in integrate_waves,
// for workgroup b of `groups` (the wavefronts then stride by groups times
// the wavefronts per workgroup).  When every visible block gets a wavefront
// of its own, XCD j (which runs workgroups b with b % 8 == j) takes the j-th
// eighth of the visible list: its blocks, listed in allocation order,
// project near one another, so each XCD's L2 holds part of the depth image
// rather than all of it.  Idle wavefronts get an index past the list.
__device__ inline int integrate_wave(const MapView& m, int b, int groups) {
  const int wpg = blockDim.x >> 6;
  const int nv0 = m.ctr[kCtrVisible];
  const int nv = nv0 < m.num_blocks ? nv0 : m.num_blocks;
  const int per = ((nv + wpg - 1) / wpg + 7) >> 3;  // workgroups with a block, per XCD
  int g = b;
  if (per * 8 <= groups) {
    const int j = b & 7, k = b >> 3;
    g = k < per ? j * per + k : groups;
  }
  return __builtin_amdgcn_readfirstlane(g * wpg + (int)(threadIdx.x >> 6));
}

// K5 (VoxelMap.cu:719-786): one wavefront per visible block, lane = (x, y),
// looping over the 8 z-slices; each slice is 64 consecutive voxels.  After
// the update the wavefront re-derives the block's summary flags from the
// final voxel values and stores them in the block's hash slot.  Wavefront
// `wave` of `nwaves` takes visible blocks wave, wave + nwaves, ...
template <class VT, class D>
__device__ inline void integrate_waves(const MapView& m, const PoseF& Ti, const Intr& K, const D& depth, int W, int H,
                                       int wave, int nwaves) {
  const int nv = m.ctr[kCtrVisible];
  const int nvis = nv < m.num_blocks ? nv : m.num_blocks;
  const int lane = threadIdx.x & 63;
  const int lx = lane & 7, ly = lane >> 3;
  const float vs = m.voxel_size, mu = m.trunc;
  typename VT::T* vox = (typename VT::T*)m.voxels;
  // the next block's index and key are loaded while this one is updated
  int nblk = wave < nvis ? m.visible[wave] : 0;
  unsigned long long nkey = wave < nvis ? m.block_keys[nblk] : 0ull;
  for (int w = wave; w < nvis; w += nwaves) {
    const int blk = nblk;
    const unsigned long long key = nkey;
    if (w + nwaves < nvis) {
      nblk = m.visible[w + nwaves];
      nkey = m.block_keys[nblk];
    }
    typename VT::T* bv = vox + (size_t)blk * 512;
    // the block's 8 slices are loaded together, then each pixel's depth is
    // read (at a clamped pixel, used only when in the image) before any update
    typename VT::T cur[8];
#pragma unroll
    for (int lz = 0; lz < 8; ++lz) cur[lz] = bv[lz * 64 + lane];
    int bx, by, bz;
    unpack_key(key, bx, by, bz);
    const float wx = (float)(bx * 8 + lx) * vs;
    const float wy = (float)(by * 8 + ly) * vs;
    float pzs[8], ds[8];
    bool in[8];
#pragma unroll
    for (int lz = 0; lz < 8; ++lz) {
      const float wz = (float)(bz * 8 + lz) * vs;
      float px, py, pz;
      xform(Ti, wx, wy, wz, px, py, pz);
      const int u = floor_sat(K.fx * px / pz + K.cx + 0.5f);
      const int v = floor_sat(K.fy * py / pz + K.cy + 0.5f);
      in[lz] = !(u < 0 || v < 0 || u > W - 1 || v > H - 1);
      ds[lz] = depth.at(min(max(u, 0), W - 1), min(max(v, 0), H - 1));
      pzs[lz] = pz;
    }
    bool unobserved = true, free_space = true;
#pragma unroll
    for (int lz = 0; lz < 8; ++lz) {
      if (in[lz]) {
        const float d = ds[lz];
        if (!(d != d || d > m.depth_max || d < m.depth_min)) {
          float sdf = d - pzs[lz];
          if (!(sdf < -mu)) {
            sdf = fminf(1.0f, sdf / mu);
            cur[lz] = VT::update(cur[lz], sdf, m.max_weight);
            bv[lz * 64 + lane] = cur[lz];
          }
        }
      }
      const bool obs = VT::observed(cur[lz]);
      unobserved = unobserved && !obs;
      free_space = free_space && obs && VT::sdf(cur[lz]) == 1.0f;
    }
    const unsigned long long un = __ballot(unobserved), fr = __ballot(free_space);
    if (lane == 0) {
      const int f = (un == ~0ull ? kBlockUnobserved : 0) | (fr == ~0ull ? kBlockFree : 0);
      m.slots[m.block_slot[blk]].flags = f;
    }
  }
}

}  // namespace yas
