// Device code of the ray caster (RGBDFusion/src/RayTraceEngine.cu:58-410),
// shared by the single-map kernels and the kernels that ray-cast many maps in
// one launch.  Bodies take their block index (and grid size for the strided
// loops) explicitly.
//
//   zrange   one thread per visible block projects its 8 corners onto the
//            1/8-resolution grid and min/max-reduces the block's depth range
//            into every covered cell.  Depths are positive floats, whose bit
//            patterns order like the values, so unsigned atomicMin/atomicMax
//            are exact replacements for the reference's CAS loops
//            (VoxelStructUtils.h:150-171).  The reference's split of each box
//            into 16x16 tiles (RayTraceEngine.cu:146-214) only distributed
//            work; the per-cell result is the same min/max over all covering
//            blocks, without its 100000-tile cap.
//   march    one thread per pixel marches from the cell's zmin to zmax in
//            voxel units (RayTraceEngine.cu:331-410).  The block looked up last
//            is cached in registers, so consecutive samples inside one 8^3
//            block cost no hash probe.
#pragma once

#include "yas_internal.hpp"

namespace yas {

// Cells hold this zmin until a block covers them.
constexpr float kZrangeFar = 100.0f;

__device__ inline void zrange_init_at(unsigned* zmin, unsigned* zmax, int n, int bx, int nb) {
  for (int i = bx * blockDim.x + threadIdx.x; i < n; i += nb * blockDim.x) {
    zmin[i] = __float_as_uint(kZrangeFar);
    zmax[i] = 0u;
  }
}

// RayTraceEngine.cu:80-144 (createRenderingBlock).  One block's cells and
// depth range: the 8 corners projected onto the 1/8 grid, clipped to the
// image; false when the block covers no cell or lies wholly in front of
// depth_min.  `a`, `b` are the float bits of the range.
struct CellBox {
  int x0, y0, x1, y1;
  unsigned a, b;
};

__device__ inline bool block_cells(const MapView& m, const PoseF& Ti, const Intr& K, int cols, int rows,
                                   unsigned long long key, CellBox& cb) {
  const float scale = m.voxel_size * 8.0f;
  int bx, by, bz;
  unpack_key(key, bx, by, bz);
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
  if (ul0 > lr0 || ul1 > lr1) return false;
  if (zr0 < m.depth_min) zr0 = m.depth_min;
  if (zr1 < m.depth_min) return false;
  cb.x0 = ul0;
  cb.y0 = ul1;
  cb.x1 = lr0;
  cb.y1 = lr1;
  cb.a = __float_as_uint(zr0);
  cb.b = __float_as_uint(zr1);
  return true;
}

// LDS-privatised form for grids of at most kZrangeLdsCells cells: workgroup
// `bx0` of `nb` takes a contiguous slice of the visible list, reduces it
// into its own copy of the grid in LDS (`lz`: zmin then zmax, 2*cols*rows
// words), then merges every cell it touched into the global grid with one
// atomic per cell and bound.  Neighbouring entries of the visible list were
// allocated together and cover the same cells, so most updates meet in LDS.
// Min and max are order-free, so the grid equals zrange_at's bit for bit.
constexpr int kZrangeLdsCells = 8192;

__device__ inline void zrange_lds_at(const MapView& m, const PosePair* pose, const Intr& K, int cols, int rows,
                                     unsigned* zmin, unsigned* zmax, unsigned* lz, int bx0, int nb) {
  const int ncell = cols * rows;
  const unsigned far = __float_as_uint(kZrangeFar);
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    lz[c] = far;
    lz[ncell + c] = 0u;
  }
  __syncthreads();
  const int nv = m.ctr[kCtrVisible];
  const int nvis = nv < m.num_blocks ? nv : m.num_blocks;
  const int per = (nvis + nb - 1) / nb;
  const int i0 = bx0 * per;
  const int i1 = i0 + per < nvis ? i0 + per : nvis;
  const PoseF Ti = pose->Tinv;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    CellBox cb;
    if (!block_cells(m, Ti, K, cols, rows, m.block_keys[m.visible[i]], cb)) continue;
    for (int y = cb.y0; y <= cb.y1; ++y)
      for (int x = cb.x0; x <= cb.x1; ++x) {
        atomicMin(&lz[y * cols + x], cb.a);
        atomicMax(&lz[ncell + y * cols + x], cb.b);
      }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const unsigned a = lz[c], b = lz[ncell + c];
    if (a < far) atomicMin(&zmin[c], a);
    if (b != 0u) atomicMax(&zmax[c], b);
  }
}

// RayTraceEngine.cu:184-214 (fill), straight into the global grid.
// kReadFirst: read a cell before each atomic and skip the atomic when the
// read already shows it would change nothing (a cell's zmin only falls and
// its zmax only rises, so a stale read errs towards doing the atomic).  That
// pays where many maps update cells at once (the batched launches).
template <bool kReadFirst = false>
__device__ inline void zrange_at(const MapView& m, const PosePair* pose, const Intr& K, int cols, int rows,
                                 unsigned* zmin, unsigned* zmax, int bx0, int nb) {
  const int nv = m.ctr[kCtrVisible];
  const int nvis = nv < m.num_blocks ? nv : m.num_blocks;
  const PoseF Ti = pose->Tinv;
  for (int i = bx0 * blockDim.x + threadIdx.x; i < nvis; i += nb * blockDim.x) {
    CellBox cb;
    if (!block_cells(m, Ti, K, cols, rows, m.block_keys[m.visible[i]], cb)) continue;
    const int ul0 = cb.x0, ul1 = cb.y0, lr0 = cb.x1, lr1 = cb.y1;
    const unsigned a = cb.a, b = cb.b;
    for (int y = ul1; y <= lr1; ++y)
      for (int x = ul0; x <= lr0; ++x) {
        unsigned* pn = &zmin[y * cols + x];
        unsigned* pf = &zmax[y * cols + x];
        if (!kReadFirst || a < __hip_atomic_load(pn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
          atomicMin(pn, a);
        if (!kReadFirst || b > __hip_atomic_load(pf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
          atomicMax(pf, b);
      }
  }
}

// Voxel coordinate of a sample: floor, saturated into the int range, NaN to
// 0.  A ray whose end points lie within +-2^28 voxels stays within +-2^30
// (it advances at most its length plus one step past its end), where a
// plain conversion of the floor gives the same integer.
template <bool kInRange>
__device__ inline int voxel_of(float x) {
  return kInRange ? (int)floorf(x) : floor_sat(x);
}

// Voxel reads with a one-entry block cache.
template <class VT, bool kInRange = false>
struct VoxelReader {
  const MapView& m;
  const typename VT::T* vox;
  int cbx, cby, cbz, cblk, cflags;
  int rx0 = 0, ry0 = 0, rz0 = 0;  // voxel of the last read()
  float rs = 1.0f;                // its answer
  bool rv = false;
  __device__ VoxelReader(const MapView& mv)
      : m(mv), vox((const typename VT::T*)mv.voxels), cbx(INT32_MIN), cby(0), cbz(0), cblk(-1), cflags(0) {}
  // RayTraceEngine.cu:279-292: nearest read at floor(p), valid iff observed.
  // Blocks summarised as all-unobserved or all-free answer without a voxel
  // load (the values they would return are exactly (1, invalid) / (1, valid)).
  // The voxel and its answer are kept for interp_after_read.
  __device__ inline float read(float x, float y, float z, bool& valid) {
    rx0 = voxel_of<kInRange>(x); ry0 = voxel_of<kInRange>(y); rz0 = voxel_of<kInRange>(z);
    rs = read_at(block_of(rx0), block_of(ry0), block_of(rz0),
                 (vox_to_local(rz0) * 8 + vox_to_local(ry0)) * 8 + vox_to_local(rx0), valid);
    rv = valid;
    return rs;
  }
  // RayTraceEngine.cu:294-329: trilinear read, valid iff all 8 reads are.
  // The corners are floor(x) and floor(x + 1) per axis, as the reference
  // rounds them; each is converted once and the 8 reads run in the
  // reference's order.
  __device__ inline float interp(float x, float y, float z, bool& valid) {
    const int x0 = voxel_of<kInRange>(x), y0 = voxel_of<kInRange>(y), z0 = voxel_of<kInRange>(z);
    bool v0;
    const float s0 = read_at(block_of(x0), block_of(y0), block_of(z0),
                             (vox_to_local(z0) * 8 + vox_to_local(y0)) * 8 + vox_to_local(x0), v0);
    return interp_rest(x, y, z, x0, y0, z0, s0, v0, valid);
  }
  // The same at the point of the last read(): its voxel is the first corner,
  // read just before with the same answer and leaving the cache on its block.
  __device__ inline float interp_after_read(float x, float y, float z, bool& valid) {
    return interp_rest(x, y, z, rx0, ry0, rz0, rs, rv, valid);
  }
  // How a block answers a corner read: its pool index when the voxel must
  // be loaded, else -1 for (1, valid) (all free) or -2 for (1, invalid)
  // (absent or all unobserved) -- read_at's cases.
  __device__ static inline int corner_code(int blk, int flags) {
    if (blk < 0 || flags == kBlockUnobserved) return -2;
    return flags == kBlockFree ? -1 : blk;
  }
  // The 7 remaining corners: corner c is (c&1 ? x1 : x0, c&2 ? y1 : y0,
  // c&4 ? z1 : z0).  When at most one axis's two corners straddle a block
  // face (nearly every call), the corners lie in corner 0's block, which is
  // cached, and at most one other, looked up once; then the voxel loads of
  // all 7 corners are issued before any is used.  Read one by one through
  // the one-entry cache they made a chain of dependent loads, with a lookup
  // per read when the corners alternate between two blocks.  Corners
  // spread over more blocks take the reads in turn.  The values and the
  // arithmetic order are those of 8 reads in the reference's order.
  __device__ inline float interp_rest(float x, float y, float z, int x0, int y0, int z0, float s0, bool vp,
                                      bool& valid) {
    const float fx = x - floorf(x), fy = y - floorf(y), fz = z - floorf(z);
    const int x1 = voxel_of<kInRange>(x + 1.0f), y1 = voxel_of<kInRange>(y + 1.0f), z1 = voxel_of<kInRange>(z + 1.0f);
    const int bx0 = block_of(x0), bx1 = block_of(x1), by0 = block_of(y0), by1 = block_of(y1);
    const int bz0 = block_of(z0), bz1 = block_of(z1);
    const int lx0 = vox_to_local(x0), lx1 = vox_to_local(x1);
    const int mask = (int)(bx1 != bx0) | ((int)(by1 != by0) << 1) | ((int)(bz1 != bz0) << 2);
    float sv[8];
    sv[0] = s0;
    if ((mask & (mask - 1)) == 0) {
      const int code_a = corner_code(cblk, cflags);
      int code_b = code_a;
      if (mask) {
        int fb = 0;
        const int ib = find_block_xyz(m, (mask & 1) ? bx1 : bx0, (mask & 2) ? by1 : by0, (mask & 4) ? bz1 : bz0, &fb);
        code_b = corner_code(ib, fb);
      }
      const int ly[2] = {vox_to_local(y0), vox_to_local(y1)};
      const int lz[2] = {vox_to_local(z0), vox_to_local(z1)};
      typename VT::T raw[8];
#pragma unroll
      for (int c = 1; c < 8; ++c) {
        const int code = (c & mask) ? code_b : code_a;
        const size_t at =
            code >= 0 ? (size_t)code * 512 + (lz[c >> 2] * 8 + ly[(c >> 1) & 1]) * 8 + ((c & 1) ? lx1 : lx0) : 0;
        raw[c] = vox[at];
      }
#pragma unroll
      for (int c = 1; c < 8; ++c) {
        const int code = (c & mask) ? code_b : code_a;
        bool v = code == -1;
        sv[c] = 1.0f;
        if (code >= 0 && VT::observed(raw[c])) {
          v = true;
          sv[c] = VT::sdf(raw[c]);
        }
        vp &= v;
      }
    } else {
      const int r00 = (vox_to_local(z0) * 8 + vox_to_local(y0)) * 8, r01 = (vox_to_local(z0) * 8 + vox_to_local(y1)) * 8;
      const int r10 = (vox_to_local(z1) * 8 + vox_to_local(y0)) * 8, r11 = (vox_to_local(z1) * 8 + vox_to_local(y1)) * 8;
      bool v;
      sv[1] = read_at(bx1, by0, bz0, r00 + lx1, v); vp &= v;
      sv[2] = read_at(bx0, by1, bz0, r01 + lx0, v); vp &= v;
      sv[3] = read_at(bx1, by1, bz0, r01 + lx1, v); vp &= v;
      sv[4] = read_at(bx0, by0, bz1, r10 + lx0, v); vp &= v;
      sv[5] = read_at(bx1, by0, bz1, r10 + lx1, v); vp &= v;
      sv[6] = read_at(bx0, by1, bz1, r11 + lx0, v); vp &= v;
      sv[7] = read_at(bx1, by1, bz1, r11 + lx1, v); vp &= v;
    }
    float r0 = (1.0f - fx) * sv[0] + fx * sv[1];
    float r1 = (1.0f - fx) * sv[2] + fx * sv[3];
    const float r2 = (1.0f - fy) * r0 + fy * r1;
    r0 = (1.0f - fx) * sv[4] + fx * sv[5];
    r1 = (1.0f - fx) * sv[6] + fx * sv[7];
    const float r3 = (1.0f - fy) * r0 + fy * r1;
    valid = vp;
    return (1.0f - fz) * r2 + fz * r3;
  }
  // Block of a voxel coordinate: floor division by 8.  It equals
  // vox_to_block except below INT32_MIN + 7 (a saturated coordinate), where
  // both lie outside the key range and the lookup finds nothing.
  __device__ static inline int block_of(int v) { return v >> 3; }
  // The cached block answers every read alike: absent, or summarised as
  // all-unobserved or all-free (read_at's early returns).
  __device__ inline bool cached_uniform() const {
    return cblk < 0 || cflags == kBlockFree || cflags == kBlockUnobserved;
  }
  // One comparison, no short-circuit branches.
  __device__ inline bool in_cached_block(float x, float y, float z) const {
    return ((block_of(voxel_of<kInRange>(x)) ^ cbx) | (block_of(voxel_of<kInRange>(y)) ^ cby) |
            (block_of(voxel_of<kInRange>(z)) ^ cbz)) == 0;
  }
  // Read of local voxel `l` of block (bx, by, bz) through the one-entry cache.
  __device__ inline float read_at(int bx, int by, int bz, int l, bool& valid) {
    if ((bx ^ cbx) | (by ^ cby) | (bz ^ cbz)) {
      cbx = bx; cby = by; cbz = bz;
      cflags = 0;
      cblk = find_block_xyz(m, bx, by, bz, &cflags);
    }
    if (cblk >= 0 && cflags == kBlockFree) {
      valid = true;
      return 1.0f;
    }
    if (cblk >= 0 && cflags != kBlockUnobserved) {
      typename VT::T v = vox[(size_t)cblk * 512 + l];
      if (VT::observed(v)) {
        valid = true;
        return VT::sdf(v);
      }
    }
    valid = false;
    return 1.0f;
  }
};

// The sample loop of RayTraceEngine.cu:367-384 and the two refinement steps
// of :386-400 after a zero crossing; on a valid hit, (rx, ry, rz) is the
// surface point in voxel units.
template <class VT, bool kInRange>
__device__ inline bool march_ray(const MapView& m, float& rx, float& ry, float& rz, float dx, float dy, float dz,
                                 float dist_s, float dist_e, float step_scale) {
  VoxelReader<VT, kInRange> rd(m);
  float sdf = 1.0f, last;
  bool valid = false;
  while (dist_s < dist_e) {
    last = sdf;
    sdf = rd.read(rx, ry, rz, valid);
    const bool near = sdf <= 0.5f && sdf >= -0.5f;
    if (near) sdf = rd.interp_after_read(rx, ry, rz, valid);
    if (sdf <= 0.0f) break;
    if (sdf >= 0.0f && last < 0.0f) return false;  // abandoned
    float step = valid ? fmaxf(sdf * step_scale, 1.0f) : 1.0f;
    const float sdx = step * dx, sdy = step * dy, sdz = step * dz;
    rx = rx + sdx; ry = ry + sdy; rz = rz + sdz;
    dist_s = dist_s + step;
    // When this sample was answered by a block that answers every read with
    // the same (1, valid) (no trilinear read, so the cache still holds the
    // sample's block), each further sample inside it would repeat this
    // iteration exactly (sdf 1, last 1, the same step): take those steps
    // without the reads until the ray leaves the block or its range.
    if (!near && rd.cached_uniform()) {
      bool more = (dist_s < dist_e) & rd.in_cached_block(rx, ry, rz);
      while (more) {
        rx = rx + sdx; ry = ry + sdy; rz = rz + sdz;
        dist_s = dist_s + step;
        more = (dist_s < dist_e) & rd.in_cached_block(rx, ry, rz);
      }
    }
  }
  if (!(sdf <= 0.0f)) return false;
  float step = sdf * step_scale;
  rx = rx + step * dx; ry = ry + step * dy; rz = rz + step * dz;
  sdf = rd.interp(rx, ry, rz, valid);
  step = sdf * step_scale;
  rx = rx + step * dx; ry = ry + step * dy; rz = rz + step * dz;
  return valid;
}

// RayTraceEngine.cu:331-410.  Departure: pixels the reference leaves
// unwritten get w = -1 (x = NaN, as the reference writes first).
// Tile `tile` (raster order) of the image's (W/16)*(H/16) tiles of 16x16.
template <class VT>
__device__ inline void march_tile_at(const MapView& m, const PosePair* pose, const Intr& K, int W, int H, int cols,
                                     const float* zmin, const float* zmax, float4* vmap, float vsinv,
                                     float step_scale, int tile) {
  // A 16x16 tile per workgroup; each wavefront takes one 8x8 z-range cell,
  // so its rays share a depth range and run for similar numbers of steps.
  const int tiles_x = (W + 15) >> 4;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int x = (tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
  const int y = (tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
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
    const float lim = 268435456.0f;  // 2^28 voxels
    const bool in_range = fabsf(bsx) < lim && fabsf(bsy) < lim && fabsf(bsz) < lim && fabsf(bex) < lim &&
                          fabsf(bey) < lim && fabsf(bez) < lim && step_scale < 1048576.0f;
    float rx = bsx, ry = bsy, rz = bsz;
    const bool hit = __all(in_range) ? march_ray<VT, true>(m, rx, ry, rz, dx, dy, dz, dist_s, dist_e, step_scale)
                                     : march_ray<VT, false>(m, rx, ry, rz, dx, dy, dz, dist_s, dist_e, step_scale);
    if (hit) {
      const PoseF Ti = pose->Tinv;
      float cx, cy, cz;
      xform(Ti, rx / vsinv, ry / vsinv, rz / vsinv, cx, cy, cz);
      out = make_float4(cx, cy, cz, 1.0f);
    }
  }
  vmap[(size_t)y * W + x] = out;
}

// Workgroup b in launch order: the tiles dealt to the XCDs in bands (xcd_tile).
template <class VT>
__device__ inline void march_at(const MapView& m, const PosePair* pose, const Intr& K, int W, int H, int cols,
                                const float* zmin, const float* zmax, float4* vmap, float vsinv, float step_scale, int b) {
  march_tile_at<VT>(m, pose, K, W, H, cols, zmin, zmax, vmap, vsinv, step_scale,
                    xcd_tile(b, ((W + 15) >> 4) * ((H + 15) >> 4)));
}

}  // namespace yas
