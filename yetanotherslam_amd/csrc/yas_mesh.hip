// Marching-cubes mesh extraction: the reference's MeshEngine
// (RGBDFusion/src/MeshEngine.cu:7-339, tables RGBDFusion/utils/TriangleTable.h).
//
// MI355X design: one wavefront (64 threads) per allocated block, eight
// voxels per thread.  The +1 corners of a block's voxels live in up to seven
// neighbour blocks; their pool indices are resolved once per workgroup
// (seven hash probes) instead of eight probes per voxel, and the block with
// its +1 halo is staged in LDS, so each voxel's eight corners are LDS reads.
// Extraction is a count
// pass, a three-phase device scan of the per-block counts and an emit pass,
// so every block's output is contiguous and ordered by voxel index, then
// table order: the output is deterministic, unlike the reference's atomic
// triangle counter (MeshEngine.cu:246).
//
// Two output forms share the count pass:
//   * the reference's triangle soup (9 positions + 9 face-normal floats per
//     triangle, MeshEngine.cu:246-264), bit-identical to the scalar
//     restatement;
//   * an indexed mesh (float32 xyz vertices, int32 triangle corners).  Each
//     voxel owns the three cube edges leaving its corner along +x, +y, +z;
//     an owned edge with a sign change between two observed end voxels gets
//     one vertex, placed by the reference's interpolation evaluated from the
//     lower end.  Triangles of a cube reference the owners' vertices, so a
//     vertex is shared by every triangle that touches its edge.

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

// Owner of each cube edge: the voxel offset (from the cube's corner 0) of
// the edge's lower end and the edge's axis (0 = x, 1 = y, 2 = z).
__constant__ int kOwner[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {1, 0, 1, 1},
                                  {0, 1, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};

// A workgroup of kMcThreads threads takes one block; thread t owns voxels
// l = r * kMcThreads + t (r < kMcRounds), l = z*64 + y*8 + x, so the owned
// voxels of a round are consecutive and the rounds in order are the voxel
// order.  The block's 512 voxels and its +1 halo (217 voxels of up to seven
// neighbour blocks) are staged in LDS as a 9x9x9 cube of (sdf, observed),
// index (z*9 + y)*9 + x, once per block: the corner reads of the cubes then
// come from LDS (eight per voxel), and a workgroup of one wavefront lets
// eight times as many blocks wait on memory at once as one thread per voxel
// did (measured against 128 threads: 5 % faster, against 512: 1.7 times).
constexpr int kMcThreads = 64;
constexpr int kMcRounds = 512 / kMcThreads;
constexpr int kMcCube = 9 * 9 * 9;
constexpr int kMcHalo = kMcCube - 512;

// One word per voxel: the sdf, or kMcAbsent's bits where the voxel is not
// observed (absent block or weight 0).  An observed NaN sdf is staged as the
// canonical quiet NaN, so no observed value has those bits.
constexpr unsigned kMcAbsent = 0xFFFFFFFFu;
struct McCube {
  float v[kMcCube];
};
__device__ inline float mc_staged(bool observed, float sdf) {
  return !observed ? __uint_as_float(kMcAbsent) : sdf != sdf ? __uint_as_float(0x7FC00000u) : sdf;
}

// Pool indices of the block and its neighbours at offset (k&1, k>>1&1,
// k>>2&1) in nb[k] (-1 if absent), and the staged cube.  The block's own
// voxels are loaded while threads 0-7 probe the neighbours.  With `vinfo`,
// the count pass's per-voxel vertex info of the same 9x9x9 voxels is staged
// in `cvi` too (0 in absent blocks).
// The neighbours come from `nbr_in` (the count pass's record, 8 per block)
// when given, else from the hash, and are recorded in `nbr_out` when given.
template <class VT>
__device__ inline void stage_cube(const MapView& m, int blk, int* nb, McCube& c, int& bx, int& by, int& bz,
                                  const int* nbr_in, int* nbr_out, const uint16_t* vinfo = nullptr,
                                  uint16_t* cvi = nullptr) {
  const typename VT::T* vox = (const typename VT::T*)m.voxels;
  const int t = threadIdx.x;
  typename VT::T own[kMcRounds];
#pragma unroll
  for (int r = 0; r < kMcRounds; ++r) own[r] = vox[(size_t)blk * 512 + r * kMcThreads + t];
  if (vinfo) {
#pragma unroll
    for (int r = 0; r < kMcRounds; ++r) {
      const int l = r * kMcThreads + t;
      cvi[((l >> 6) * 9 + ((l >> 3) & 7)) * 9 + (l & 7)] = vinfo[(size_t)blk * 512 + l];
    }
  }
  unpack_key(m.block_keys[blk], bx, by, bz);
  if (t < 8) {
    const int b = nbr_in ? nbr_in[(size_t)blk * 8 + t]
                         : t == 0 ? blk : find_block_xyz(m, bx + (t & 1), by + ((t >> 1) & 1), bz + ((t >> 2) & 1));
    nb[t] = b;
    if (nbr_out) nbr_out[(size_t)blk * 8 + t] = b;
  }
#pragma unroll
  for (int r = 0; r < kMcRounds; ++r) {
    const int l = r * kMcThreads + t, x = l & 7, y = (l >> 3) & 7, z = l >> 6;
    const int i = (z * 9 + y) * 9 + x;
    c.v[i] = mc_staged(VT::observed(own[r]), VT::sdf(own[r]));
  }
  __syncthreads();
  for (int h = t; h < kMcHalo; h += kMcThreads) {
    // the x = 8 face (81), then y = 8 with x < 8 (72), then z = 8 with x, y < 8 (64)
    int x, y, z;
    if (h < 81) {
      x = 8; y = h % 9; z = h / 9;
    } else if (h < 153) {
      x = (h - 81) & 7; y = 8; z = (h - 81) >> 3;
    } else {
      x = (h - 153) & 7; y = (h - 153) >> 3; z = 8;
    }
    const int b = nb[(x >> 3) | ((y >> 3) << 1) | ((z >> 3) << 2)];
    const size_t at = (size_t)b * 512 + ((z & 7) * 8 + (y & 7)) * 8 + (x & 7);
    float v = __uint_as_float(kMcAbsent);
    uint16_t vi = 0;
    if (b >= 0) {
      const typename VT::T w = vox[at];
      if (vinfo) vi = vinfo[at];
      v = mc_staged(VT::observed(w), VT::sdf(w));
    }
    const int i = (z * 9 + y) * 9 + x;
    c.v[i] = v;
    if (vinfo) cvi[i] = vi;
  }
  __syncthreads();
}

// The 8 corner values of voxel (lx, ly, lz) and a bit mask of the corners
// that are observed (in an allocated block with weight > 0), from the cube.
__device__ inline int read_corners(const McCube& c, int lx, int ly, int lz, float* s) {
  int valid = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = ((lz + kCorner[k][2]) * 9 + (ly + kCorner[k][1])) * 9 + (lx + kCorner[k][0]);
    s[k] = c.v[i];
    valid |= (int)(__float_as_uint(s[k]) != kMcAbsent) << k;
  }
  return valid;
}

// Cube index (MeshEngine.cu:76-112): -1 unless all 8 corners are observed.
__device__ inline int cube_index(int valid, const float* s) {
  if (valid != 0xFF) return -1;
  int cube = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if (s[c] < 0.0f) cube |= 1 << c;
  return cube;
}

// Owned edges of the voxel (bit a = edge along axis a from corner 0 to
// corner 1, 3 or 4) that carry a vertex.
__device__ inline int owned_edges(int valid, const float* s) {
  const int far[3] = {1, 3, 4};
  int m = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int c = far[a];
    if ((valid & 1) && (valid & (1 << c)) && ((s[0] < 0.0f) != (s[c] < 0.0f))) m |= 1 << a;
  }
  return m;
}

// Exclusive prefix of v over the threads of a workgroup (thread order), and
// the workgroup total.  `wsum` holds one int per wavefront.
// Exclusive prefix of v over a one-wavefront workgroup (the marching-cubes
// passes run kMcThreads = 64 threads); `total` receives the sum.  No LDS
// word and no barrier.
static_assert(kMcThreads == 64, "the marching-cubes passes run one wavefront per block");
__device__ inline int wave_excl_scan(int v, int& total) {
  const int lane = threadIdx.x & 63;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  total = __builtin_amdgcn_readlane(inc, 63);
  return inc - v;
}
// Triangle counts of the 256 cases (kTriTable's low nibbles) into LDS, for
// the count pass to read per voxel instead of the table in memory (the
// emission passes need the whole table word); published by stage_cube's
// first barrier.
__device__ inline void stage_tri_counts(unsigned char* nt) {
  for (int i = threadIdx.x; i < 256; i += kMcThreads) nt[i] = (unsigned char)(kTriTable[i] & 0xF);
}

__device__ inline int block_excl_scan(int v, int* wsum, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int before = 0;
  total = 0;
  for (int i = 0; i < nw; ++i) {
    const int x = wsum[i];
    before += i < w ? x : 0;
    total += x;
  }
  __syncthreads();
  return before + inc - v;
}

// Count pass: per block, triangles and owned vertices; per voxel, the cube
// index and (local vertex prefix << 3 | owned edge mask).  One scan a round
// of the packed counts (triangles << 16 | vertices: a block has at most 2560
// and 1536).
template <class VT>
__global__ void __launch_bounds__(kMcThreads) k_mc_count(MapView m, int nblocks, int2* counts, int16_t* cases,
                                                         uint16_t* vinfo, int* nbr) {
  __shared__ int nb[8];
  __shared__ unsigned char tri_n[256];
  __shared__ McCube c;
  const int blk = blockIdx.x;
  if (blk >= nblocks) return;
  int bx, by, bz;
  stage_tri_counts(tri_n);
  stage_cube<VT>(m, blk, nb, c, bx, by, bz, nullptr, nbr);
  int tri = 0, vert = 0;  // totals of the rounds so far
#pragma unroll 1
  for (int r = 0; r < kMcRounds; ++r) {
    const int l = r * kMcThreads + threadIdx.x, lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    float s[8];
    const int valid = read_corners(c, lx, ly, lz, s);
    const int cube = cube_index(valid, s);
    const int om = owned_edges(valid, s);
    cases[(size_t)blk * 512 + l] = (int16_t)cube;
    const int nt = cube > 0 ? (int)tri_n[cube] : 0;
    int total;
    const int pre = wave_excl_scan((nt << 16) | __popc(om), total);
    vinfo[(size_t)blk * 512 + l] = (uint16_t)(((vert + (pre & 0xFFFF)) << 3) | om);
    tri += total >> 16;
    vert += total & 0xFFFF;
  }
  if (threadIdx.x == 0) counts[blk] = make_int2(tri, vert);
}

// Three-phase exclusive scan of int2 counts: 1024-element chunks scanned in
// parallel, the chunk totals scanned by one workgroup, then added back.
__global__ void __launch_bounds__(1024) k_scan_local(const int2* in, int2* out, int n, int2* chunk) {
  __shared__ int wsum[16];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const int2 v = i < n ? in[i] : make_int2(0, 0);
  int tx, ty;
  const int ex = block_excl_scan(v.x, wsum, tx);
  const int ey = block_excl_scan(v.y, wsum, ty);
  if (i < n) out[i] = make_int2(ex, ey);
  if (threadIdx.x == 0) chunk[blockIdx.x] = make_int2(tx, ty);
}

__global__ void __launch_bounds__(1024) k_scan_chunks(int2* chunk, int nchunks, int2* total) {
  __shared__ int wsum[16];
  __shared__ int carry[2];
  if (threadIdx.x == 0) carry[0] = carry[1] = 0;
  __syncthreads();
  for (int base = 0; base < nchunks; base += 1024) {
    const int i = base + threadIdx.x;
    const int2 v = i < nchunks ? chunk[i] : make_int2(0, 0);
    int tx, ty;
    const int ex = block_excl_scan(v.x, wsum, tx);
    const int ey = block_excl_scan(v.y, wsum, ty);
    if (i < nchunks) chunk[i] = make_int2(carry[0] + ex, carry[1] + ey);
    __syncthreads();
    if (threadIdx.x == 0) {
      carry[0] += tx;
      carry[1] += ty;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = make_int2(carry[0], carry[1]);
}

__global__ void __launch_bounds__(1024) k_scan_add(int2* out, const int2* chunk, int n) {
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i >= n) return;
  const int2 c = chunk[blockIdx.x], v = out[i];
  out[i] = make_int2(v.x + c.x, v.y + c.y);
}

// Vertex on cube edge e of the voxel at integer position p (voxel units),
// MeshEngine.cu:114-224: corner A + t * (edge direction), scaled by vs.
__device__ inline void edge_vertex(int e, const float* s, float px, float py, float pz, float vs, float* out) {
  float t = edge_t(s[kEdgeA[e]], s[kEdgeB[e]]);
  float ox, oy, oz;
  edge_offset(e, t, ox, oy, oz);
  out[0] = (px + ox) * vs;
  out[1] = (py + oy) * vs;
  out[2] = (pz + oz) * vs;
}

// Emit, reference form: triangle slot = block offset + exclusive prefix (in
// voxel index order) of the per-voxel counts.
template <class VT>
__global__ void __launch_bounds__(kMcThreads) k_mc_emit(MapView m, int nblocks, const int2* offsets, float* tris,
                                                        float* normals, int cap, const int* nbr) {
  __shared__ int nb[8];
  __shared__ McCube c;
  const int blk = blockIdx.x;
  if (blk >= nblocks) return;
  int bx, by, bz;
  stage_cube<VT>(m, blk, nb, c, bx, by, bz, nbr, nullptr);
  const float vs = m.voxel_size;
  int base = offsets[blk].x;
#pragma unroll 1
  for (int r = 0; r < kMcRounds; ++r) {
    const int l = r * kMcThreads + threadIdx.x, lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    float s[8];
    const int cube = cube_index(read_corners(c, lx, ly, lz, s), s);
    const uint64_t word = cube > 0 ? kTriTable[cube] : 0ull;
    const int nt = (int)(word & 0xF);
    int total;
    int slot = base + wave_excl_scan(nt, total);
    base += total;
    const float px = (float)(bx * 8 + lx), py = (float)(by * 8 + ly), pz = (float)(bz * 8 + lz);
    for (int i = 0; i < nt && slot < cap; ++i, ++slot) {
      float a[3], b[3], cc[3];
      edge_vertex((int)((word >> (4 + 12 * i)) & 0xF), s, px, py, pz, vs, a);
      edge_vertex((int)((word >> (8 + 12 * i)) & 0xF), s, px, py, pz, vs, b);
      edge_vertex((int)((word >> (12 + 12 * i)) & 0xF), s, px, py, pz, vs, cc);
      float u0 = b[0] - a[0], u1 = b[1] - a[1], u2 = b[2] - a[2];
      float w0 = cc[0] - a[0], w1 = cc[1] - a[1], w2 = cc[2] - a[2];
      float n0 = u1 * w2 - u2 * w1, n1 = u2 * w0 - u0 * w2, n2 = u0 * w1 - u1 * w0;
      float q = n0 * n0 + n1 * n1 + n2 * n2;
      if (q > 0.0f) {
        float nn = sqrtf(q);
        n0 = n0 / nn; n1 = n1 / nn; n2 = n2 / nn;
      }
      float* T = tris + (size_t)slot * 9;
      T[0] = a[0]; T[1] = a[1]; T[2] = a[2];
      T[3] = b[0]; T[4] = b[1]; T[5] = b[2];
      T[6] = cc[0]; T[7] = cc[1]; T[8] = cc[2];
      if (normals) {
        float* N = normals + (size_t)slot * 9;
        N[0] = N[3] = N[6] = n0;
        N[1] = N[4] = N[7] = n1;
        N[2] = N[5] = N[8] = n2;
      }
    }
  }
}

// Emit, indexed form: the voxel's owned vertices, then its triangles as
// indices of the owners' vertices.
template <class VT>
__global__ void __launch_bounds__(kMcThreads) k_mc_emit_indexed(MapView m, int nblocks, const int2* offsets,
                                                                const uint16_t* vinfo, float* verts, int vcap,
                                                                int* tris, int tcap, const int* nbr) {
  __shared__ int nb[8];
  __shared__ int vbase[8];
  __shared__ McCube c;
  __shared__ uint16_t cvi[kMcCube];
  const int blk = blockIdx.x;
  if (blk >= nblocks) return;
  int bx, by, bz;
  stage_cube<VT>(m, blk, nb, c, bx, by, bz, nbr, nullptr, vinfo, cvi);
  if (threadIdx.x < 8) vbase[threadIdx.x] = nb[threadIdx.x] >= 0 ? offsets[nb[threadIdx.x]].y : 0;
  __syncthreads();
  const float vs = m.voxel_size;
  int base = offsets[blk].x;
  const int far[3] = {1, 3, 4};
#pragma unroll 1
  for (int r = 0; r < kMcRounds; ++r) {
    const int l = r * kMcThreads + threadIdx.x, lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    float s[8];
    const int valid = read_corners(c, lx, ly, lz, s);
    const int cube = cube_index(valid, s);
    const uint64_t word = cube > 0 ? kTriTable[cube] : 0ull;
    const int nt = (int)(word & 0xF);
    int total;
    int slot = base + wave_excl_scan(nt, total);
    base += total;
    const float px = (float)(bx * 8 + lx), py = (float)(by * 8 + ly), pz = (float)(bz * 8 + lz);
    // owned vertices, interpolated from this voxel towards +x, +y, +z
    const int own = cvi[(lz * 9 + ly) * 9 + lx];
    int vid = vbase[0] + (own >> 3);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!(own & (1 << a))) continue;
      const float t = edge_t(s[0], s[far[a]]);
      if (vid < vcap) {
        float* V = verts + (size_t)vid * 3;
        V[0] = (px + (a == 0 ? t : 0.0f)) * vs;
        V[1] = (py + (a == 1 ? t : 0.0f)) * vs;
        V[2] = (pz + (a == 2 ? t : 0.0f)) * vs;
      }
      ++vid;
    }
    for (int i = 0; i < nt && slot < tcap; ++i, ++slot) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int e = (int)((word >> (4 + 12 * i + 4 * k)) & 0xF);
        const int x = lx + kOwner[e][0], y = ly + kOwner[e][1], z = lz + kOwner[e][2];
        const int nbk = (x >> 3) | ((y >> 3) << 1) | ((z >> 3) << 2);
        // the owner's block is present: every corner of an emitting cube is observed
        const int oi = cvi[(z * 9 + y) * 9 + x];
        tris[(size_t)slot * 3 + k] = vbase[nbk] + (oi >> 3) + __popc((oi & 7) & ((1 << kOwner[e][3]) - 1));
      }
    }
  }
}

}  // namespace

struct yas_mesher {
  yas_ctx* ctx;
  int max_tri;
  float* tris;      // soup: 9 floats per triangle; indexed: 3 ints per triangle
  float* normals;   // soup: 9 floats per triangle; indexed: 3 floats per vertex
  int2* counts;     // per block {triangles, vertices}
  int2* offsets;    // exclusive prefix of counts
  int2* chunk;      // scan chunk totals
  int2* total;      // {triangles, vertices}
  int16_t* cases;   // per voxel cube index
  uint16_t* vinfo;  // per voxel (local vertex prefix << 3 | owned edges)
  int* nbr;         // per block, pool indices of it and its 7 +x/+y/+z neighbours (-1 absent)
  int block_cap;
  int num_blocks;
  int num_tri;
  int num_vert;
  int indexed;
  yas_map* map;
};

static int ensure_block_cap(yas_mesher* me, int n) {
  if (n <= me->block_cap) return YAS_OK;
  hipFree(me->counts);
  hipFree(me->offsets);
  hipFree(me->chunk);
  hipFree(me->cases);
  hipFree(me->vinfo);
  hipFree(me->nbr);
  me->nbr = nullptr;
  me->counts = me->offsets = me->chunk = nullptr;
  me->cases = nullptr;
  me->vinfo = nullptr;
  me->block_cap = 0;
  YAS_HIP(hipMalloc((void**)&me->counts, sizeof(int2) * n));
  YAS_HIP(hipMalloc((void**)&me->offsets, sizeof(int2) * n));
  YAS_HIP(hipMalloc((void**)&me->chunk, sizeof(int2) * ((n + 1023) / 1024)));
  YAS_HIP(hipMalloc((void**)&me->cases, sizeof(int16_t) * 512 * (size_t)n));
  YAS_HIP(hipMalloc((void**)&me->vinfo, sizeof(uint16_t) * 512 * (size_t)n));
  YAS_HIP(hipMalloc((void**)&me->nbr, sizeof(int) * 8 * (size_t)n));
  me->block_cap = n;
  return YAS_OK;
}

// Count + scan over every allocated block of `map`; fills me->num_blocks and
// leaves the totals on the device.
static int count_and_scan(yas_mesher* me, yas_map* map, MapView& v) {
  YAS_REQUIRE(me && map, "yas_mesh_extract: null argument");
  YAS_REQUIRE_ACTIVE(map);
  YAS_REQUIRE(yas_map_ctx(map)->stream == me->ctx->stream, "yas_mesh_extract: map and mesher must share a stream");
  YAS_HIP(hipSetDevice(me->ctx->device));
  yas_map_stats st;
  int rc = yas_map_get_stats(map, &st);
  if (rc) return rc;
  const int n = st.num_allocated;
  me->num_blocks = n;
  me->num_tri = me->num_vert = 0;
  me->map = map;
  if (n == 0) return YAS_OK;
  if ((rc = ensure_block_cap(me, n))) return rc;
  v = yas_map_view(map);
  hipStream_t s = me->ctx->stream;
  switch (v.format) {
    case YAS_VOXEL_I16_U8:
      hipLaunchKernelGGL(k_mc_count<VoxI16>, dim3(n), dim3(kMcThreads), 0, s, v, n, me->counts, me->cases, me->vinfo,
                         me->nbr);
      break;
    case YAS_VOXEL_F16_U16:
      hipLaunchKernelGGL(k_mc_count<VoxF16>, dim3(n), dim3(kMcThreads), 0, s, v, n, me->counts, me->cases, me->vinfo,
                         me->nbr);
      break;
    default:
      hipLaunchKernelGGL(k_mc_count<VoxF32>, dim3(n), dim3(kMcThreads), 0, s, v, n, me->counts, me->cases, me->vinfo,
                         me->nbr);
  }
  YAS_LAUNCH_CHECK();
  const int nchunks = (n + 1023) / 1024;
  hipLaunchKernelGGL(k_scan_local, dim3(nchunks), dim3(1024), 0, s, (const int2*)me->counts, me->offsets, n, me->chunk);
  YAS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(1024), 0, s, me->chunk, nchunks, me->total);
  YAS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_add, dim3(nchunks), dim3(1024), 0, s, me->offsets, (const int2*)me->chunk, n);
  YAS_LAUNCH_CHECK();
  return YAS_OK;
}

static int read_totals(yas_mesher* me) {
  int2 t;
  YAS_HIP(hipMemcpyAsync(&t, me->total, sizeof(int2), hipMemcpyDeviceToHost, me->ctx->stream));
  YAS_HIP(hipStreamSynchronize(me->ctx->stream));
  me->num_tri = t.x;
  me->num_vert = t.y;
  return YAS_OK;
}

// Host copies of the last indexed extraction (for file writers).
int yas_mesher_indexed_host(yas_mesher* me, std::vector<float>& verts, std::vector<int32_t>& tris) {
  YAS_REQUIRE(me && me->indexed, "no indexed extraction to write");
  const int nv = me->num_vert < me->max_tri * 3 ? me->num_vert : me->max_tri * 3;
  const int nt = me->num_tri < me->max_tri ? me->num_tri : me->max_tri;
  verts.resize((size_t)nv * 3);
  tris.resize((size_t)nt * 3);
  if (nv > 0) YAS_HIP(hipMemcpy(verts.data(), me->normals, sizeof(float) * 3 * (size_t)nv, hipMemcpyDeviceToHost));
  if (nt > 0) YAS_HIP(hipMemcpy(tris.data(), me->tris, sizeof(int32_t) * 3 * (size_t)nt, hipMemcpyDeviceToHost));
  return YAS_OK;
}

extern "C" {

int yas_mesher_create(yas_ctx* ctx, int max_triangles, yas_mesher** out) {
  YAS_REQUIRE(ctx && out && max_triangles > 0, "yas_mesher_create: bad argument");
  YAS_HIP(hipSetDevice(ctx->device));
  yas_mesher* me = new yas_mesher();
  std::memset(me, 0, sizeof(*me));
  me->ctx = yas_ctx_retain(ctx);
  me->max_tri = max_triangles;
  hipError_t e = hipMalloc((void**)&me->tris, sizeof(float) * 9 * (size_t)max_triangles);
  if (e == hipSuccess) e = hipMalloc((void**)&me->normals, sizeof(float) * 9 * (size_t)max_triangles);
  if (e == hipSuccess) e = hipMalloc((void**)&me->total, sizeof(int2));
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
  hipFree(me->chunk);
  hipFree(me->total);
  hipFree(me->cases);
  hipFree(me->vinfo);
  hipFree(me->nbr);
  yas_ctx_release(me->ctx);
  delete me;
  return YAS_OK;
}

int yas_mesh_extract(yas_mesher* me, yas_map* map, int* num_triangles) {
  if (num_triangles) *num_triangles = 0;
  MapView v;
  int rc = count_and_scan(me, map, v);
  if (rc) return rc;
  me->indexed = 0;
  const int n = me->num_blocks;
  if (n == 0) return YAS_OK;
  hipStream_t s = me->ctx->stream;
  switch (v.format) {
    case YAS_VOXEL_I16_U8:
      hipLaunchKernelGGL(k_mc_emit<VoxI16>, dim3(n), dim3(kMcThreads), 0, s, v, n, (const int2*)me->offsets, me->tris,
                         me->normals, me->max_tri, (const int*)me->nbr);
      break;
    case YAS_VOXEL_F16_U16:
      hipLaunchKernelGGL(k_mc_emit<VoxF16>, dim3(n), dim3(kMcThreads), 0, s, v, n, (const int2*)me->offsets, me->tris,
                         me->normals, me->max_tri, (const int*)me->nbr);
      break;
    default:
      hipLaunchKernelGGL(k_mc_emit<VoxF32>, dim3(n), dim3(kMcThreads), 0, s, v, n, (const int2*)me->offsets, me->tris,
                         me->normals, me->max_tri, (const int*)me->nbr);
  }
  YAS_LAUNCH_CHECK();
  if ((rc = read_totals(me))) return rc;
  if (num_triangles) *num_triangles = me->num_tri;
  if (me->num_tri > me->max_tri) {
    yas_set_error("yas_mesh_extract: %d triangles exceed the capacity %d", me->num_tri, me->max_tri);
    return YAS_E_CAPACITY;
  }
  return YAS_OK;
}

int yas_mesh_extract_indexed(yas_mesher* me, yas_map* map, int* num_vertices, int* num_triangles) {
  if (num_vertices) *num_vertices = 0;
  if (num_triangles) *num_triangles = 0;
  MapView v;
  int rc = count_and_scan(me, map, v);
  if (rc) return rc;
  me->indexed = 1;
  const int n = me->num_blocks;
  if (n == 0) return YAS_OK;
  hipStream_t s = me->ctx->stream;
  // triangle indices in the soup buffer (3 of its 9 floats per triangle),
  // vertices in the normal buffer (capacity max_tri * 3 vertices)
  int* tri_idx = (int*)me->tris;
  const int vcap = me->max_tri * 3;
  switch (v.format) {
    case YAS_VOXEL_I16_U8:
      hipLaunchKernelGGL(k_mc_emit_indexed<VoxI16>, dim3(n), dim3(kMcThreads), 0, s, v, n, (const int2*)me->offsets,
                         (const uint16_t*)me->vinfo, me->normals, vcap, tri_idx, me->max_tri,
                         (const int*)me->nbr);
      break;
    case YAS_VOXEL_F16_U16:
      hipLaunchKernelGGL(k_mc_emit_indexed<VoxF16>, dim3(n), dim3(kMcThreads), 0, s, v, n, (const int2*)me->offsets,
                         (const uint16_t*)me->vinfo, me->normals, vcap, tri_idx, me->max_tri,
                         (const int*)me->nbr);
      break;
    default:
      hipLaunchKernelGGL(k_mc_emit_indexed<VoxF32>, dim3(n), dim3(kMcThreads), 0, s, v, n, (const int2*)me->offsets,
                         (const uint16_t*)me->vinfo, me->normals, vcap, tri_idx, me->max_tri,
                         (const int*)me->nbr);
  }
  YAS_LAUNCH_CHECK();
  if ((rc = read_totals(me))) return rc;
  if (num_vertices) *num_vertices = me->num_vert;
  if (num_triangles) *num_triangles = me->num_tri;
  if (me->num_tri > me->max_tri || me->num_vert > vcap) {
    yas_set_error("yas_mesh_extract_indexed: %d triangles / %d vertices exceed the capacity %d / %d", me->num_tri,
                  me->num_vert, me->max_tri, vcap);
    return YAS_E_CAPACITY;
  }
  return YAS_OK;
}

int yas_mesher_download(yas_mesher* me, float* vertices, float* normals, int capacity) {
  YAS_REQUIRE(me, "null mesher");
  YAS_REQUIRE(!me->indexed, "yas_mesher_download: the last extraction was indexed");
  int n = me->num_tri < me->max_tri ? me->num_tri : me->max_tri;
  if (n > capacity) n = capacity;
  if (n <= 0) return YAS_OK;
  if (vertices) YAS_HIP(hipMemcpy(vertices, me->tris, sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToHost));
  if (normals) YAS_HIP(hipMemcpy(normals, me->normals, sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToHost));
  return YAS_OK;
}

int yas_mesher_download_indexed(yas_mesher* me, float* vertices, int32_t* triangles, int vertex_capacity,
                                int triangle_capacity) {
  YAS_REQUIRE(me, "null mesher");
  YAS_REQUIRE(me->indexed, "yas_mesher_download_indexed: the last extraction was not indexed");
  int nv = me->num_vert < me->max_tri * 3 ? me->num_vert : me->max_tri * 3;
  int nt = me->num_tri < me->max_tri ? me->num_tri : me->max_tri;
  if (nv > vertex_capacity) nv = vertex_capacity;
  if (nt > triangle_capacity) nt = triangle_capacity;
  if (vertices && nv > 0) YAS_HIP(hipMemcpy(vertices, me->normals, sizeof(float) * 3 * (size_t)nv, hipMemcpyDeviceToHost));
  if (triangles && nt > 0) YAS_HIP(hipMemcpy(triangles, me->tris, sizeof(int32_t) * 3 * (size_t)nt, hipMemcpyDeviceToHost));
  return YAS_OK;
}

int yas_mesher_indexed_device(yas_mesher* me, const float** vertices, const int32_t** triangles) {
  YAS_REQUIRE(me && me->indexed, "yas_mesher_indexed_device: no indexed extraction");
  if (vertices) *vertices = me->normals;
  if (triangles) *triangles = (const int32_t*)me->tris;
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
  if (first || count) {
    std::vector<int2> o(n), c(n);
    YAS_HIP(hipMemcpy(o.data(), me->offsets, sizeof(int2) * n, hipMemcpyDeviceToHost));
    YAS_HIP(hipMemcpy(c.data(), me->counts, sizeof(int2) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
      if (first) first[i] = o[i].x;
      if (count) count[i] = c[i].x;
    }
  }
  if (cases) YAS_HIP(hipMemcpy(cases, me->cases, sizeof(int16_t) * 512 * (size_t)n, hipMemcpyDeviceToHost));
  return YAS_OK;
}

}  // extern "C"
