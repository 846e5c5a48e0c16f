// ORACLE - test infrastructure only.
//
// A scalar C++ restatement of the reference's RGBDFusion dense path, used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker and the CPU baseline.  Nothing in the product links or calls it.
//
// It is compiled with g++ -O2 -ffp-contract=off (no fast-math, no FMA
// contraction) so that every float operation below is one IEEE operation in
// the order written.  A second build with -fopenmp (liboracle_omp.so) runs
// the per-pixel and per-block loops on all host cores; it exists only to time
// the CPU baseline on every core (its pixel sums add in another order), and
// the tests use the single-threaded build.  The HIP kernels evaluate the same operations in the same
// order, which is what makes block sets, TSDF values and marching-cubes cases
// bit-exact between the two.  Sums over pixels (J^T J, residual statistics)
// are accumulated here in double and rounded to the working type once, which
// makes this the "exact-sum" restatement; GPU tree reductions land within
// rounding of it.
//
// Everything is templated on the real type R: R=float is the parity oracle,
// R=double the FP64 accuracy oracle (same code, every operation in double).
//
// Reference files followed (paths relative to the reference root):
//   RGBDFusion/utils/VoxelStructUtils.h:6-317   hashing, coordinates, visibility
//   RGBDFusion/src/VoxelMap.cu:517-884          allocation, visibility, integration
//   RGBDFusion/src/RayTraceEngine.cu:58-450     z-range tiles and ray marching
//   RGBDFusion/src/MeshEngine.cu:7-339          marching cubes
//   RGBDFusion/src/ImageProc.cu:7-78,141-285    image pyramids, vertex/normal maps
//   RGBDFusion/include/TrackingUtils.h:45-220,578-607  photometric terms
//   RGBDFusion/src/RGBDTracking.cu:5-545        tracker, ICP step, GN loop
//   FullSystem/src/Tracking.cpp:139-197         the per-frame call sequence
//
// Parity pins: see tests/test_oracle_*.py (closed forms, brute force, KATs,
// table checksums, numpy/scipy golden vectors).  Third-party arithmetic the
// reference delegates to OpenCV (pyrDown, cvtColor, convertTo), Eigen (LDLT)
// and Sophus (SE3 exp/compose) is restated from their documented algorithms;
// where no reference test pins it the test module says "parity unpinned".
#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- helpers

template <class R> inline bool is_nan(R v) { return v != v; }

// float -> int rounding down with saturation (NaN -> 0); the behaviour of the
// GPU's __float2int_rd / cvt.rmi used at VoxelMap.cu:755-756.
template <class R> inline int floor_sat(R f) {
  if (is_nan(f)) return 0;
  double d = std::floor((double)f);
  if (d <= -2147483648.0) return INT_MIN;
  if (d >= 2147483647.0) return INT_MAX;
  return (int)d;
}
template <class R> inline int ceil_sat(R f) {
  if (is_nan(f)) return 0;
  double d = std::ceil((double)f);
  if (d <= -2147483648.0) return INT_MIN;
  if (d >= 2147483647.0) return INT_MAX;
  return (int)d;
}
// C++ (int)/(short) conversion of a float: truncation toward zero, saturated.
template <class R> inline int trunc_sat(R f) {
  if (is_nan(f)) return 0;
  double d = std::trunc((double)f);
  if (d <= -2147483648.0) return INT_MIN;
  if (d >= 2147483647.0) return INT_MAX;
  return (int)d;
}

// IEEE binary16 round-to-nearest-even from binary32.
inline uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t exp = (x >> 23) & 0xFFu;
  uint32_t man = x & 0x7FFFFFu;
  if (exp == 0xFF) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u | (man >> 13) : 0u));
  int e = (int)exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7C00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    man |= 0x800000u;
    int shift = 14 - e;
    uint32_t h = man >> shift;
    uint32_t rem = man & ((1u << shift) - 1u);
    uint32_t halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)e << 10) | (man >> 13);
  uint32_t rem = man & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
  return (uint16_t)(sign | h);
}
inline float f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1Fu;
  uint32_t man = h & 0x3FFu;
  uint32_t x;
  if (exp == 0) {
    if (man == 0) {
      x = sign;
    } else {
      int e = -1;
      do { man <<= 1; ++e; } while (!(man & 0x400u));
      x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3FFu) << 13);
    }
  } else if (exp == 31) {
    x = sign | 0x7F800000u | (man << 13);
  } else {
    x = sign | ((exp - 15 + 127) << 23) | (man << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

// VoxelStructUtils.h:16-21 - spatial hash with int32 wrapping multiplies.
inline int ref_hash(int x, int y, int z, int n) {
  uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
  int32_t r = (int32_t)h;
  r %= n;
  return r < 0 ? r + n : r;
}
// VoxelStructUtils.h:39-49 - block of a voxel (floor division by 8).
inline int vox_to_block(int v) {
  if (v < 0) v -= 7;
  return v / 8;
}
// VoxelStructUtils.h:56-70 - voxel position inside its block.
inline int vox_to_local(int v) {
  int l = v % 8;
  return l < 0 ? l + 8 : l;
}

struct I3 { int x, y, z; };
inline uint64_t pack_key(int x, int y, int z) {
  return ((uint64_t)(uint32_t)(x & 0x1FFFFF)) | ((uint64_t)(uint32_t)(y & 0x1FFFFF) << 21) |
         ((uint64_t)(uint32_t)(z & 0x1FFFFF) << 42);
}
inline int sext21(uint64_t v) {
  int x = (int)(v & 0x1FFFFF);
  return (x & 0x100000) ? x - 0x200000 : x;
}
inline I3 unpack_key(uint64_t k) { return {sext21(k), sext21(k >> 21), sext21(k >> 42)}; }
inline bool key_in_range(int x, int y, int z) {
  return x >= -(1 << 20) && x < (1 << 20) && y >= -(1 << 20) && y < (1 << 20) && z >= -(1 << 20) &&
         z < (1 << 20);
}

template <class R> struct Pose { R r[9]; R t[3]; };

// p' = R p + t, every row summed left to right.
template <class R> inline void xform(const Pose<R>& T, const R p[3], R o[3]) {
  for (int i = 0; i < 3; ++i) o[i] = T.r[3 * i] * p[0] + T.r[3 * i + 1] * p[1] + T.r[3 * i + 2] * p[2] + T.t[i];
}
template <class R> inline void rot(const Pose<R>& T, const R p[3], R o[3]) {
  for (int i = 0; i < 3; ++i) o[i] = T.r[3 * i] * p[0] + T.r[3 * i + 1] * p[1] + T.r[3 * i + 2] * p[2];
}

// 4x4 row-major double -> working-type pose and its inverse.  The inverse is
// formed in double (R^T, -R^T t with each sum left to right) and then rounded.
template <class R> inline Pose<R> pose_from(const double T[16]) {
  Pose<R> p;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) p.r[3 * i + j] = (R)T[4 * i + j];
    p.t[i] = (R)T[4 * i + 3];
  }
  return p;
}
inline void inverse_d(const double T[16], double o[16]) {
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) o[4 * i + j] = T[4 * j + i];
    o[4 * i + 3] = -(T[i] * T[3] + T[4 + i] * T[7] + T[8 + i] * T[11]);
  }
  o[12] = o[13] = o[14] = 0.0;
  o[15] = 1.0;
}
inline void mul_d(const double A[16], const double B[16], double C[16]) {
  double t[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) t[4 * i + j] = A[4 * i] * B[j] + A[4 * i + 1] * B[4 + j] + A[4 * i + 2] * B[8 + j] + A[4 * i + 3] * B[12 + j];
  std::memcpy(C, t, sizeof(t));
}
inline void identity_d(double T[16]) {
  for (int i = 0; i < 16; ++i) T[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

template <class R> struct Intr {
  R fx, fy, cx, cy, invfx, invfy;
};
// Intrinsics as the reference derives them: float copies of K and
// float(1.0 / (double)fx) for the inverses (VoxelMap.cu:827-832).
template <class R> inline Intr<R> intr_from(double fx, double fy, double cx, double cy) {
  Intr<R> k;
  k.fx = (R)fx; k.fy = (R)fy; k.cx = (R)cx; k.cy = (R)cy;
  k.invfx = (R)(1.0 / (double)k.fx);
  k.invfy = (R)(1.0 / (double)k.fy);
  return k;
}

// ---------------------------------------------------------------- voxels
// Storage variants of one update rule (VoxelMap.cu:764-783,
// VoxelStructUtils.h:6-14).
struct VRef { int16_t s; uint8_t w; uint8_t pad; };
struct VF32 { float s; float w; };
struct VF16 { uint16_t s; uint16_t w; };
struct VF64 { double s; double w; };

template <class V> struct Vox;
template <> struct Vox<VRef> {
  typedef float R;
  static float sdf(const VRef& v) { return (float)v.s / 32767.0f; }
  static float wt(const VRef& v) { return (float)v.w; }
  // the voxel whose decoded values are (s, w), for values read back from
  // the same storage (the nearest int16 of s * 32767 is the stored one)
  static VRef make(double s, double w) {
    VRef v;
    v.s = (int16_t)std::lround(s * 32767.0);
    v.w = (uint8_t)w;
    v.pad = 0;
    return v;
  }
  static int16_t pack(float f) {
    int i = trunc_sat(f * 32767.0f);
    return (int16_t)std::max(-32768, std::min(32767, i));
  }
  static void update(VRef& v, float sdf, int /*maxw*/) {
    float old = (float)v.s / 32767.0f;
    int w = v.w;
    if (w == 0) {
      v.s = pack(sdf);
      v.w = 1;
      return;
    }
    v.s = pack((old * (float)w + sdf) / (float)(w + 1));
    v.w = (uint8_t)std::min(255, w + 1);
  }
  // VoxelMap.cu:373-375: (sdf*w + sdf'*w') / (w + w'), weight capped at 255.
  static void merge(VRef& v, float sdf, int wo, int /*maxw*/) {
    float old = (float)v.s / 32767.0f;
    int w = v.w;
    v.s = pack((old * (float)w + sdf * (float)wo) / (float)(w + wo));
    v.w = (uint8_t)std::min(255, w + wo);
  }
};
template <> struct Vox<VF32> {
  typedef float R;
  static float sdf(const VF32& v) { return v.s; }
  static float wt(const VF32& v) { return v.w; }
  static VF32 make(double s, double w) { return VF32{(float)s, (float)w}; }
  static void update(VF32& v, float sdf, int maxw) {
    float w = v.w;
    if (w == 0.0f) {
      v.s = sdf;
      v.w = 1.0f;
      return;
    }
    v.s = (v.s * w + sdf) / (w + 1.0f);
    v.w = std::min((float)maxw, w + 1.0f);
  }
  static void merge(VF32& v, float sdf, int wo, int maxw) {
    float w = v.w;
    v.s = (v.s * w + sdf * (float)wo) / (w + (float)wo);
    v.w = std::min((float)maxw, w + (float)wo);
  }
};
template <> struct Vox<VF16> {
  typedef float R;
  static float sdf(const VF16& v) { return f16_to_f32(v.s); }
  static float wt(const VF16& v) { return (float)v.w; }
  static VF16 make(double s, double w) { return VF16{f32_to_f16((float)s), (uint16_t)w}; }
  static void update(VF16& v, float sdf, int maxw) {
    int w = v.w;
    if (w == 0) {
      v.s = f32_to_f16(sdf);
      v.w = 1;
      return;
    }
    float old = f16_to_f32(v.s);
    v.s = f32_to_f16((old * (float)w + sdf) / (float)(w + 1));
    v.w = (uint16_t)std::min(maxw, w + 1);
  }
  static void merge(VF16& v, float sdf, int wo, int maxw) {
    int w = v.w;
    float old = f16_to_f32(v.s);
    v.s = f32_to_f16((old * (float)w + sdf * (float)wo) / (float)(w + wo));
    v.w = (uint16_t)std::min(maxw, w + wo);
  }
};
template <> struct Vox<VF64> {
  typedef double R;
  static double sdf(const VF64& v) { return v.s; }
  static double wt(const VF64& v) { return v.w; }
  static VF64 make(double s, double w) { return VF64{s, w}; }
  static void update(VF64& v, double sdf, int maxw) {
    double w = v.w;
    if (w == 0.0) {
      v.s = sdf;
      v.w = 1.0;
      return;
    }
    v.s = (v.s * w + sdf) / (w + 1.0);
    v.w = std::min((double)maxw, w + 1.0);
  }
  static void merge(VF64& v, double sdf, int wo, int maxw) {
    double w = v.w;
    v.s = (v.s * w + sdf * (double)wo) / (w + (double)wo);
    v.w = std::min((double)maxw, w + (double)wo);
  }
};

// ---------------------------------------------------------------- marching cubes
static const uint64_t* decoded_tri_table();

// MeshEngine.cu:114-123 - the edge crossing parameter (thresholds compared
// in double, as the reference's 1e-6 literal is a double).
template <class R> inline R mc_edge_t(R v1, R v2) {
  if (std::fabs((double)(0 - v1)) < 1e-6) return 0;
  if (std::fabs((double)(0 - v2)) < 1e-6) return 1;
  if (std::fabs((double)(v1 - v2)) < 1e-6) return 0;
  return (0 - v1) / (v2 - v1);
}

// MeshEngine.cu:125-269 for one cube whose corner 0 is the voxel p (integer
// voxel coordinates) with corner values s in Bourke order
// (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1).
// Returns the cube index; writes its triangles (9 values each, metres) and
// face normals normalize((b-a) x (c-a)), and the triangle count.
template <class R> inline int mc_cube(const int p[3], const R s[8], R vs, R* tris, R* normals, int* ntri) {
  *ntri = 0;
  int cube = 0;
  for (int c = 0; c < 8; ++c)
    if (s[c] < 0) cube |= 1 << c;
  if (cube <= 0) return cube;
  const uint64_t word = decoded_tri_table()[cube];
  const int nt = (int)(word & 0xF);
  if (nt == 0) return cube;
  const R pf[3] = {(R)p[0], (R)p[1], (R)p[2]};
  static const int ea[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
  static const int eb[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};
  R vert[12][3];
  uint32_t used = 0;
  for (int i = 0; i < 3 * nt; ++i) used |= 1u << ((word >> (4 + 4 * i)) & 0xF);
  for (int e = 0; e < 12; ++e) {
    if (!(used & (1u << e))) continue;
    R t = mc_edge_t(s[ea[e]], s[eb[e]]);
    R o[3];
    switch (e) {
      case 0: o[0] = t; o[1] = 0; o[2] = 0; break;
      case 1: o[0] = 1; o[1] = t; o[2] = 0; break;
      case 2: o[0] = 1 - t; o[1] = 1; o[2] = 0; break;
      case 3: o[0] = 0; o[1] = 1 - t; o[2] = 0; break;
      case 4: o[0] = t; o[1] = 0; o[2] = 1; break;
      case 5: o[0] = 1; o[1] = t; o[2] = 1; break;
      case 6: o[0] = 1 - t; o[1] = 1; o[2] = 1; break;
      case 7: o[0] = 0; o[1] = 1 - t; o[2] = 1; break;
      case 8: o[0] = 0; o[1] = 0; o[2] = t; break;
      case 9: o[0] = 1; o[1] = 0; o[2] = t; break;
      case 10: o[0] = 1; o[1] = 1; o[2] = t; break;
      default: o[0] = 0; o[1] = 1; o[2] = t; break;
    }
    for (int k = 0; k < 3; ++k) vert[e][k] = pf[k] + o[k];
  }
  for (int i = 0; i < nt; ++i) {
    R a[3], b[3], c[3];
    int e0 = (word >> (4 + 12 * i)) & 0xF, e1 = (word >> (8 + 12 * i)) & 0xF, e2 = (word >> (12 + 12 * i)) & 0xF;
    for (int k = 0; k < 3; ++k) {
      a[k] = vert[e0][k] * vs;
      b[k] = vert[e1][k] * vs;
      c[k] = vert[e2][k] * vs;
    }
    R u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    R w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    R n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    R n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (n2 > 0) {
      R nn = std::sqrt(n2);
      for (int k = 0; k < 3; ++k) n[k] = n[k] / nn;
    }
    for (int k = 0; k < 3; ++k) {
      tris[9 * i + k] = a[k];
      tris[9 * i + 3 + k] = b[k];
      tris[9 * i + 6 + k] = c[k];
      normals[9 * i + k] = normals[9 * i + 3 + k] = normals[9 * i + 6 + k] = n[k];
    }
  }
  *ntri = nt;
  return cube;
}

// ---------------------------------------------------------------- map
struct MapBase {
  virtual ~MapBase() {}
  virtual int fuse(const float* depth, int W, int H, const double K[4], const double T[16]) = 0;
  virtual int allocate(const float* depth, int W, int H, const double K[4], const double T[16]) = 0;
  virtual int visible(int W, int H, const double K[4], const double T[16], int32_t* keys_out, int cap) = 0;
  virtual int num_blocks() const = 0;
  virtual int export_blocks(int32_t* keys, double* sdf, double* wt, int cap) const = 0;
  virtual void import_blocks(const int32_t* keys, const double* sdf, const double* wt, int n) = 0;
  virtual int raycast(int W, int H, const double K[4], const double T[16], double* vmap, double* zmin,
                      double* zmax) = 0;
  virtual long mesh(int32_t* block_keys, int32_t* tri_counts, int16_t* cases, double* tris, double* normals,
                    long tri_cap, int block_cap) = 0;
  virtual int precision() const = 0;
  virtual int fuse_map(MapBase* src, const double Tdst[16], const double Tsrc[16]) = 0;
};

template <class V> struct Map : MapBase {
  typedef typename Vox<V>::R R;
  typedef std::array<V, 512> Block;
  R vs, mu, dmin, dmax;
  int maxw;
  std::unordered_map<uint64_t, Block> blocks;
  int64_t walk_steps = 0;  // total Bresenham visits of the last allocation
  bool out_of_range = false;

  Map(double voxel, double trunc, int maxw_, double dmin_, double dmax_)
      : vs((R)voxel), mu((R)trunc), dmin((R)dmin_), dmax((R)dmax_), maxw(maxw_) {}

  int precision() const override { return sizeof(R) == 4 ? 32 : 64; }

  // ------------------------------------------------------------ map merge
  // MapStruct::Fuse(MapStruct*) (VoxelMap.cu:400-449) with the defect of
  // CreateBlockFunctor (VoxelMap.cu:233: a 1-D index over an 8x8 launch,
  // which visits one voxel column in eight) fixed: every observed source
  // voxel is visited.  Step 1 (VoxelMap.cu:228-259): each observed voxel of
  // src, moved into this map's frame, allocates its block here.  Step 2
  // (VoxelMap.cu:360-378): every voxel here samples src trilinearly at its
  // position moved into src's frame and, when the 8 samples are observed,
  // merges it with the weight of the first sample.
  int fuse_map(MapBase* base, const double Tdst[16], const double Tsrc[16]) override {
    Map<V>* src = dynamic_cast<Map<V>*>(base);
    if (!src) return -1;
    double inv[16], Tsd[16], Tds[16];
    inverse_d(Tdst, inv);
    mul_d(inv, Tsrc, Tsd);  // src frame -> this frame
    inverse_d(Tsrc, inv);
    mul_d(inv, Tdst, Tds);  // this frame -> src frame
    const Pose<R> Psd = pose_from<R>(Tsd), Pds = pose_from<R>(Tds);
    int created = 0;
    for (const auto& kv : src->blocks) {
      const I3 B = unpack_key(kv.first);
      for (int l = 0; l < 512; ++l) {
        if (Vox<V>::wt(kv.second[l]) == 0) continue;
        const R p[3] = {(R)(B.x * 8 + (l & 7)) * vs, (R)(B.y * 8 + ((l >> 3) & 7)) * vs, (R)(B.z * 8 + (l >> 6)) * vs};
        R q[3];
        xform(Psd, p, q);
        insert(vox_to_block(floor_sat(q[0] / vs)), vox_to_block(floor_sat(q[1] / vs)), vox_to_block(floor_sat(q[2] / vs)),
               created);
      }
    }
    for (auto& kv : blocks) {
      const I3 B = unpack_key(kv.first);
      for (int l = 0; l < 512; ++l) {
        const R p[3] = {(R)(B.x * 8 + (l & 7)) * vs, (R)(B.y * 8 + ((l >> 3) & 7)) * vs, (R)(B.z * 8 + (l >> 6)) * vs};
        R q[3];
        xform(Pds, p, q);
        int wo = 0;
        R sdf;
        if (!src->interp_merge(q, sdf, wo)) continue;
        Vox<V>::merge(kv.second[l], sdf, wo, maxw);
      }
    }
    return created;
  }

  // MapStructFusionFunctor::InterpolateSDF / ReadSDF (VoxelMap.cu:290-358):
  // trilinear in the reference's evaluation order; *wt gets the weight of
  // the first corner.  False (and no value) unless all 8 corners are observed.
  bool interp_merge(const R q[3], R& out, int& wt) const {
    R vp[3];
    int v0[3];
    for (int i = 0; i < 3; ++i) {
      vp[i] = q[i] / vs;
      v0[i] = floor_sat(vp[i]);
    }
    const R x = vp[0] - (R)v0[0], y = vp[1] - (R)v0[1], z = vp[2] - (R)v0[2];
    R s[8];
    for (int c = 0; c < 8; ++c) {
      const V* v = find_voxel(v0[0] + (c & 1), v0[1] + ((c >> 1) & 1), v0[2] + ((c >> 2) & 1));
      if (!v || Vox<V>::wt(*v) == 0) return false;
      if (c == 0) wt = (int)Vox<V>::wt(*v);
      s[c] = Vox<V>::sdf(*v);
    }
    R s2 = s[0] * (1 - x) + s[1] * x;
    R s3 = s2 * (1 - y) + (s[2] * (1 - x) + s[3] * x) * y;
    s2 = s[4] * (1 - x) + s[5] * x;
    out = s3 * (1 - z) + (s2 * (1 - y) + (s[6] * (1 - x) + s[7] * x) * y) * z;
    return true;
  }

  const V* find_voxel(int vx, int vy, int vz) const {
    auto it = blocks.find(pack_key(vox_to_block(vx), vox_to_block(vy), vox_to_block(vz)));
    if (it == blocks.end()) return nullptr;
    return &it->second[(vox_to_local(vz) * 8 + vox_to_local(vy)) * 8 + vox_to_local(vx)];
  }

  void insert(int x, int y, int z, int& created) {
    ++walk_steps;
    if (!key_in_range(x, y, z)) {
      out_of_range = true;
      return;
    }
    uint64_t k = pack_key(x, y, z);
    if (blocks.find(k) == blocks.end()) {
      Block b;
      std::memset(b.data(), 0, sizeof(Block));
      blocks.emplace(k, b);
      ++created;
    }
  }

  // VoxelStructUtils.h:97-110 then :28-49 - pixel (x,y) at depth z to block.
  I3 block_of(int x, int y, R z, const Intr<R>& K, const Pose<R>& T) const {
    R pc[3] = {K.invfx * ((R)x - K.cx) * z, K.invfy * ((R)y - K.cy) * z, z};
    R pw[3];
    xform(T, pc, pw);
    int v[3];
    for (int i = 0; i < 3; ++i) v[i] = floor_sat(pw[i] / vs);
    return {vox_to_block(v[0]), vox_to_block(v[1]), vox_to_block(v[2])};
  }

  // VoxelMap.cu:517-652 - every block on the 3D Bresenham walk from the
  // block at depth-mu/2 to the block at depth+mu/2 along pixel (x, y)'s ray,
  // passed to emit(bx, by, bz) in walk order.
  template <class F>
  void walk_pixel(const float* depth, int W, int x, int y, const Intr<R>& K, const Pose<R>& T, R half,
                  F&& emit) const {
    R d = (R)depth[(size_t)y * W + x];
    if (is_nan(d) || d < dmin || d > dmax) return;
    R dn = std::max(dmin, d - half);
    R df = std::min(dmax, d + half);
    if (dn >= df) return;
    I3 b = block_of(x, y, dn, K, T);
    I3 e = block_of(x, y, df, K, T);
    int dir[3] = {e.x - b.x, e.y - b.y, e.z - b.z};
    int inc[3] = {dir[0] < 0 ? -1 : 1, dir[1] < 0 ? -1 : 1, dir[2] < 0 ? -1 : 1};
    int ad[3] = {std::abs(dir[0]), std::abs(dir[1]), std::abs(dir[2])};
    int e2[3] = {ad[0] << 1, ad[1] << 1, ad[2] << 1};
    int p[3] = {b.x, b.y, b.z};
    // Dominant axis a, the two others walked in the reference's order
    // (the z branch updates y before x, VoxelMap.cu:628-647).
    int a, o1, o2;
    if (ad[0] >= ad[1] && ad[0] >= ad[2]) { a = 0; o1 = 1; o2 = 2; }
    else if (ad[1] >= ad[0] && ad[1] >= ad[2]) { a = 1; o1 = 0; o2 = 2; }
    else { a = 2; o1 = 1; o2 = 0; }
    int err1 = e2[o1] - 1, err2 = e2[o2] - 1;
    emit(p[0], p[1], p[2]);
    for (int i = 0; i < ad[a]; ++i) {
      if (err1 > 0) { p[o1] += inc[o1]; err1 -= e2[a]; }
      if (err2 > 0) { p[o2] += inc[o2]; err2 -= e2[a]; }
      err1 += e2[o1];
      err2 += e2[o2];
      p[a] += inc[a];
      emit(p[0], p[1], p[2]);
    }
  }

  // VoxelMap.cu:517-652 - allocate every block on every pixel's walk.
  int allocate_impl(const float* depth, int W, int H, const Intr<R>& K, const Pose<R>& T) {
    int created = 0;
    walk_steps = 0;
    const R half = mu * (R)0.5;
#ifdef _OPENMP
    // Multi-threaded timing build: the walks run in parallel and their blocks
    // are inserted afterwards (the allocated set does not depend on order).
    std::vector<std::vector<I3>> walked(omp_get_max_threads());
#pragma omp parallel
    {
      std::vector<I3>& mine = walked[omp_get_thread_num()];
#pragma omp for schedule(static)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
          walk_pixel(depth, W, x, y, K, T, half, [&](int bx, int by, int bz) { mine.push_back({bx, by, bz}); });
    }
    for (const auto& v : walked)
      for (const I3& b : v) insert(b.x, b.y, b.z, created);
#else
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x)
        walk_pixel(depth, W, x, y, K, T, half, [&](int bx, int by, int bz) { insert(bx, by, bz, created); });
#endif
    return created;
  }

  // VoxelStructUtils.h:112-147 - a block is visible if any of its 8 corners
  // projects into [0,W-1]x[0,H-1] with depth in [dmin,dmax].
  bool block_visible(const I3& B, const Pose<R>& Tinv, const Intr<R>& K, int W, int H) const {
    const R scale = vs * (R)8;
    for (int c = 0; c < 8; ++c) {
      R p[3] = {(R)(B.x + (c & 1)) * scale, (R)(B.y + ((c >> 1) & 1)) * scale, (R)(B.z + ((c >> 2) & 1)) * scale};
      R q[3];
      xform(Tinv, p, q);
      R u = K.fx * q[0] / q[2] + K.cx;
      R v = K.fy * q[1] / q[2] + K.cy;
      if (!(u < 0 || v < 0 || u > (R)(W - 1) || v > (R)(H - 1) || q[2] < dmin || q[2] > dmax)) return true;
    }
    return false;
  }

  std::vector<uint64_t> visible_impl(int W, int H, const Intr<R>& K, const Pose<R>& Tinv) const {
    std::vector<uint64_t> keys;
    keys.reserve(blocks.size());
    for (const auto& kv : blocks) keys.push_back(kv.first);
    std::vector<char> vis(keys.size());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)keys.size(); ++i) vis[i] = block_visible(unpack_key(keys[i]), Tinv, K, W, H);
    std::vector<uint64_t> out;
    for (size_t i = 0; i < keys.size(); ++i)
      if (vis[i]) out.push_back(keys[i]);
    std::sort(out.begin(), out.end());
    return out;
  }

  // VoxelMap.cu:719-786 - projective TSDF update of every voxel of every
  // visible block.
  void integrate_impl(const float* depth, int W, int H, const Intr<R>& K, const Pose<R>& Tinv,
                      const std::vector<uint64_t>& vis) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long iv = 0; iv < (long)vis.size(); ++iv) {
      const uint64_t key = vis[iv];
      Block& blk = blocks.find(key)->second;
      I3 B = unpack_key(key);
      for (int lz = 0; lz < 8; ++lz)
        for (int ly = 0; ly < 8; ++ly)
          for (int lx = 0; lx < 8; ++lx) {
            R pw[3] = {(R)(B.x * 8 + lx) * vs, (R)(B.y * 8 + ly) * vs, (R)(B.z * 8 + lz) * vs};
            R p[3];
            xform(Tinv, pw, p);
            int u = floor_sat(K.fx * p[0] / p[2] + K.cx + (R)0.5);
            int v = floor_sat(K.fy * p[1] / p[2] + K.cy + (R)0.5);
            if (u < 0 || v < 0 || u > W - 1 || v > H - 1) continue;
            R d = (R)depth[(size_t)v * W + u];
            if (is_nan(d) || d > dmax || d < dmin) continue;
            R sdf = d - p[2];
            if (sdf < -mu) continue;
            sdf = std::min((R)1, sdf / mu);
            Vox<V>::update(blk[(lz * 8 + ly) * 8 + lx], sdf, maxw);
          }
    }
  }

  int fuse(const float* depth, int W, int H, const double Kd[4], const double T[16]) override {
    Intr<R> K = intr_from<R>(Kd[0], Kd[1], Kd[2], Kd[3]);
    Pose<R> Tf = pose_from<R>(T);
    double Ti[16];
    inverse_d(T, Ti);
    Pose<R> Tinv = pose_from<R>(Ti);
    int created = allocate_impl(depth, W, H, K, Tf);
    std::vector<uint64_t> vis = visible_impl(W, H, K, Tinv);
    integrate_impl(depth, W, H, K, Tinv, vis);
    return created;
  }
  int allocate(const float* depth, int W, int H, const double Kd[4], const double T[16]) override {
    Intr<R> K = intr_from<R>(Kd[0], Kd[1], Kd[2], Kd[3]);
    return allocate_impl(depth, W, H, K, pose_from<R>(T));
  }
  int visible(int W, int H, const double Kd[4], const double T[16], int32_t* keys, int cap) override {
    Intr<R> K = intr_from<R>(Kd[0], Kd[1], Kd[2], Kd[3]);
    double Ti[16];
    inverse_d(T, Ti);
    std::vector<uint64_t> vis = visible_impl(W, H, K, pose_from<R>(Ti));
    int n = (int)vis.size();
    for (int i = 0; i < n && i < cap; ++i) {
      I3 b = unpack_key(vis[i]);
      keys[3 * i] = b.x; keys[3 * i + 1] = b.y; keys[3 * i + 2] = b.z;
    }
    return n;
  }
  int num_blocks() const override { return (int)blocks.size(); }

  // Blocks sorted by (x, y, z); voxel data decoded to double (exact).
  int export_blocks(int32_t* keys, double* sdf, double* wt, int cap) const override {
    std::vector<std::pair<std::array<int, 3>, const Block*>> v;
    for (const auto& kv : blocks) {
      I3 b = unpack_key(kv.first);
      v.push_back({{b.x, b.y, b.z}, &kv.second});
    }
    std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    int n = (int)v.size();
    for (int i = 0; i < n && i < cap; ++i) {
      for (int k = 0; k < 3; ++k) keys[3 * i + k] = v[i].first[k];
      if (sdf)
        for (int j = 0; j < 512; ++j) sdf[(size_t)i * 512 + j] = (double)Vox<V>::sdf((*v[i].second)[j]);
      if (wt)
        for (int j = 0; j < 512; ++j) wt[(size_t)i * 512 + j] = (double)Vox<V>::wt((*v[i].second)[j]);
    }
    return n;
  }

  // Test set-up: blocks given by keys and decoded voxel values (as
  // export_blocks gives them), replacing any block of the same key.  Lets
  // the read-only functions (ray cast, marching cubes) run on a volume
  // built elsewhere.
  void import_blocks(const int32_t* keys, const double* sdf, const double* wt, int n) override {
    for (int i = 0; i < n; ++i) {
      Block& b = blocks[pack_key(keys[3 * i], keys[3 * i + 1], keys[3 * i + 2])];
      for (int j = 0; j < 512; ++j) b[j] = Vox<V>::make(sdf[(size_t)i * 512 + j], wt[(size_t)i * 512 + j]);
    }
  }

  // ------------------------------------------------------------ raycast
  // RayTraceEngine.cu:58-264 - per visible block, its projected bounding box
  // on the 1/8-resolution grid and its corner depth range, min/max-reduced
  // into the per-cell [zmin, zmax] maps (initialised to 100 / 0).
  void zrange_impl(const std::vector<uint64_t>& vis, int W, int H, const Intr<R>& K, const Pose<R>& Tinv,
                   std::vector<R>& zmin, std::vector<R>& zmax) const {
    const int cols = W / 8, rows = H / 8;
    zmin.assign((size_t)cols * rows, (R)100);
    zmax.assign((size_t)cols * rows, (R)0);
    const R scale = vs * (R)8;
    for (uint64_t key : vis) {
      I3 B = unpack_key(key);
      int16_t ul[2] = {(int16_t)cols, (int16_t)rows}, lr[2] = {-1, -1};
      R zr[2] = {dmax, dmin};
      for (int c = 0; c < 8; ++c) {
        R p[3] = {(R)(B.x + (c & 1)) * scale, (R)(B.y + ((c >> 1) & 1)) * scale, (R)(B.z + ((c >> 2) & 1)) * scale};
        R q[3];
        xform(Tinv, p, q);
        R pu = (K.fx * q[0] / q[2] + K.cx) / (R)8;
        R pv = (K.fy * q[1] / q[2] + K.cy) / (R)8;
        R fu = std::floor(pu), fv = std::floor(pv), cu = std::ceil(pu), cv = std::ceil(pv);
        if ((R)ul[0] > fu) ul[0] = (int16_t)floor_sat(pu);
        if ((R)lr[0] < cu) lr[0] = (int16_t)ceil_sat(pu);
        if ((R)ul[1] > fv) ul[1] = (int16_t)floor_sat(pv);
        if ((R)lr[1] < cv) lr[1] = (int16_t)ceil_sat(pv);
        if (zr[0] > q[2]) zr[0] = q[2];
        if (zr[1] < q[2]) zr[1] = q[2];
      }
      if (ul[0] < 0) ul[0] = 0;
      if (ul[1] < 0) ul[1] = 0;
      if (lr[0] >= cols) lr[0] = (int16_t)(cols - 1);
      if (lr[1] >= rows) lr[1] = (int16_t)(rows - 1);
      if (ul[0] > lr[0] || ul[1] > lr[1]) continue;
      if (zr[0] < dmin) zr[0] = dmin;
      if (zr[1] < dmin) continue;
      for (int y = ul[1]; y <= lr[1]; ++y)
        for (int x = ul[0]; x <= lr[0]; ++x) {
          R& a = zmin[(size_t)y * cols + x];
          R& b = zmax[(size_t)y * cols + x];
          if (zr[0] < a) a = zr[0];
          if (zr[1] > b) b = zr[1];
        }
    }
  }

  // RayTraceEngine.cu:279-292 - nearest-voxel read at floor(p) (voxel units).
  R read_sdf(const R p[3], bool& valid) const {
    const V* v = find_voxel(floor_sat(std::floor(p[0])), floor_sat(std::floor(p[1])), floor_sat(std::floor(p[2])));
    if (v && Vox<V>::wt(*v) != 0) {
      valid = true;
      return Vox<V>::sdf(*v);
    }
    valid = false;
    return (R)1;
  }
  // RayTraceEngine.cu:294-329 - trilinear read, valid iff all 8 reads are.
  R read_sdf_interp(const R p[3], bool& valid) const {
    R f[3] = {p[0] - std::floor(p[0]), p[1] - std::floor(p[1]), p[2] - std::floor(p[2])};
    bool vp, v;
    R s0, s1, r0, r1, r2, r3;
    auto rd = [&](R ox, R oy, R oz, bool& ok) {
      R q[3] = {p[0] + ox, p[1] + oy, p[2] + oz};
      return read_sdf(q, ok);
    };
    s0 = read_sdf(p, vp);
    s1 = rd(1, 0, 0, v); vp &= v;
    r0 = ((R)1 - f[0]) * s0 + f[0] * s1;
    s0 = rd(0, 1, 0, v); vp &= v;
    s1 = rd(1, 1, 0, v); vp &= v;
    r1 = ((R)1 - f[0]) * s0 + f[0] * s1;
    r2 = ((R)1 - f[1]) * r0 + f[1] * r1;
    s0 = rd(0, 0, 1, v); vp &= v;
    s1 = rd(1, 0, 1, v); vp &= v;
    r0 = ((R)1 - f[0]) * s0 + f[0] * s1;
    s0 = rd(0, 1, 1, v); vp &= v;
    s1 = rd(1, 1, 1, v); vp &= v;
    r1 = ((R)1 - f[0]) * s0 + f[0] * s1;
    r3 = ((R)1 - f[1]) * r0 + f[1] * r1;
    valid = vp;
    return ((R)1 - f[2]) * r2 + f[2] * r3;
  }

  // RayTraceEngine.cu:331-410 - march each pixel's ray from its cell's zmin
  // to zmax in voxel units.  Departure: pixels the reference leaves unwritten
  // (no z-range, or a positive read after a negative one) get w = -1.
  void march(int x, int y, const R* zmin, const R* zmax, int W, const Intr<R>& K, const Pose<R>& T,
             const Pose<R>& Tinv, R out[4]) const {
    const int cols = W / 8;
    const R vsinv = (R)(1.0 / (double)vs);
    const R step_scale = mu / vs;
    out[0] = std::numeric_limits<R>::quiet_NaN();
    out[1] = out[2] = 0;
    out[3] = (R)-1;
    int u = x / 8, v = y / 8;
    R zn = zmin[(size_t)v * cols + u], zf = zmax[(size_t)v * cols + u];
    const R eps = (R)std::numeric_limits<float>::epsilon();
    if (zn < eps || zf < eps || is_nan(zn) || is_nan(zf)) return;
    R pn[3] = {K.invfx * ((R)x - K.cx) * zn, K.invfy * ((R)y - K.cy) * zn, zn};
    R dist_s = std::sqrt(pn[0] * pn[0] + pn[1] * pn[1] + pn[2] * pn[2]) * vsinv;
    R bs[3], be[3], t[3];
    xform(T, pn, t);
    for (int i = 0; i < 3; ++i) bs[i] = t[i] * vsinv;
    R pf[3] = {K.invfx * ((R)x - K.cx) * zf, K.invfy * ((R)y - K.cy) * zf, zf};
    R dist_e = std::sqrt(pf[0] * pf[0] + pf[1] * pf[1] + pf[2] * pf[2]) * vsinv;
    xform(T, pf, t);
    for (int i = 0; i < 3; ++i) be[i] = t[i] * vsinv;
    R dir[3] = {be[0] - bs[0], be[1] - bs[1], be[2] - bs[2]};
    R n2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
    if (n2 > 0) {
      R n = std::sqrt(n2);
      for (int i = 0; i < 3; ++i) dir[i] = dir[i] / n;
    }
    R res[3] = {bs[0], bs[1], bs[2]};
    R sdf = 1, last;
    bool valid = false;
    while (dist_s < dist_e) {
      last = sdf;
      sdf = read_sdf(res, valid);
      if (sdf <= (R)0.5 && sdf >= (R)-0.5) sdf = read_sdf_interp(res, valid);
      if (sdf <= 0) break;
      if (sdf >= 0 && last < 0) return;
      R step = valid ? std::max(sdf * step_scale, (R)1) : (R)1;
      for (int i = 0; i < 3; ++i) res[i] = res[i] + step * dir[i];
      dist_s = dist_s + step;
    }
    if (sdf <= 0) {
      R step = sdf * step_scale;
      for (int i = 0; i < 3; ++i) res[i] = res[i] + step * dir[i];
      sdf = read_sdf_interp(res, valid);
      step = sdf * step_scale;
      for (int i = 0; i < 3; ++i) res[i] = res[i] + step * dir[i];
      if (valid) {
        R pm[3] = {res[0] / vsinv, res[1] / vsinv, res[2] / vsinv};
        R pc[3];
        xform(Tinv, pm, pc);
        out[0] = pc[0]; out[1] = pc[1]; out[2] = pc[2]; out[3] = 1;
      }
    }
  }

  // The z-range grid has (W/8)·(H/8) cells and RayTraceEngine.cu:340-344
  // reads cell (x/8, y/8) for every pixel, past the grid when W or H is not
  // a multiple of 8; such sizes are refused (-1), as the library does.
  int raycast(int W, int H, const double Kd[4], const double T[16], double* vmap, double* zmin_o,
              double* zmax_o) override {
    if (W < 8 || H < 8 || W % 8 != 0 || H % 8 != 0) return -1;
    Intr<R> K = intr_from<R>(Kd[0], Kd[1], Kd[2], Kd[3]);
    Pose<R> Tf = pose_from<R>(T);
    double Ti[16];
    inverse_d(T, Ti);
    Pose<R> Tinv = pose_from<R>(Ti);
    std::vector<uint64_t> vis = visible_impl(W, H, K, Tinv);
    std::vector<R> zmin, zmax;
    zrange_impl(vis, W, H, K, Tinv, zmin, zmax);
    if (zmin_o)
      for (size_t i = 0; i < zmin.size(); ++i) { zmin_o[i] = zmin[i]; zmax_o[i] = zmax[i]; }
    int hits = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : hits)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        R o[4];
        march(x, y, zmin.data(), zmax.data(), W, K, Tf, Tinv, o);
        for (int c = 0; c < 4; ++c) vmap[((size_t)y * W + x) * 4 + c] = (double)o[c];
        hits += o[3] > 0;
      }
    return hits;
  }

  // ------------------------------------------------------------ mesh
  // MeshEngine.cu:43-269 - per allocated block, per voxel: the 8 corners in
  // Bourke order must all be observed (w > 0); then mc_cube.  Output order:
  // blocks sorted by key, voxels by local index z*64+y*8+x, triangles in
  // table order.
  long mesh(int32_t* bkeys, int32_t* tcounts, int16_t* cases, double* tris, double* normals, long tri_cap,
            int block_cap) override {
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    std::vector<std::array<int, 3>> keys;
    for (const auto& kv : blocks) {
      I3 b = unpack_key(kv.first);
      keys.push_back({b.x, b.y, b.z});
    }
    std::sort(keys.begin(), keys.end());
    long ntri = 0;
    int bi = 0;
    R t9[5 * 9], n9[5 * 9];
    for (const auto& bk : keys) {
      int cnt = 0;
      for (int l = 0; l < 512; ++l) {
        int lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
        int p[3] = {bk[0] * 8 + lx, bk[1] * 8 + ly, bk[2] * 8 + lz};
        R s[8];
        bool ok = true;
        for (int c = 0; c < 8 && ok; ++c) {
          const V* v = find_voxel(p[0] + off[c][0], p[1] + off[c][1], p[2] + off[c][2]);
          if (!v || Vox<V>::wt(*v) == 0) ok = false;
          else s[c] = Vox<V>::sdf(*v);
        }
        int cube = -1, nt = 0;
        if (ok) cube = mc_cube<R>(p, s, vs, t9, n9, &nt);
        if (cases && bi < block_cap) cases[(size_t)bi * 512 + l] = (int16_t)cube;
        for (int i = 0; i < nt; ++i, ++ntri) {
          if (ntri < tri_cap)
            for (int k = 0; k < 9; ++k) {
              tris[ntri * 9 + k] = (double)t9[9 * i + k];
              if (normals) normals[ntri * 9 + k] = (double)n9[9 * i + k];
            }
        }
        cnt += nt;
      }
      if (bi < block_cap) {
        if (bkeys) { bkeys[3 * bi] = bk[0]; bkeys[3 * bi + 1] = bk[1]; bkeys[3 * bi + 2] = bk[2]; }
        if (tcounts) tcounts[bi] = cnt;
      }
      ++bi;
    }
    return ntri;
  }
};


// The triangulation table, decoded once from its hex-string form into
// (count | edges << 4) words.
static const char* kTriHex[256] = {
#include "mc_tables_oracle.inc"
};
static uint64_t g_tri_words[256];
static bool g_tri_init = false;
const uint64_t* decoded_tri_table() {
  if (!g_tri_init) {
    for (int i = 0; i < 256; ++i) {
      const char* s = kTriHex[i];
      int n = (int)std::strlen(s);
      uint64_t w = (uint64_t)(n / 3);
      for (int j = 0; j < n; ++j) {
        int d = s[j] <= '9' ? s[j] - '0' : s[j] - 'a' + 10;
        w |= (uint64_t)d << (4 + 4 * j);
      }
      g_tri_words[i] = w;
    }
    g_tri_init = true;
  }
  return g_tri_words;
}

// ---------------------------------------------------------------- images
template <class R> struct Img {
  int w = 0, h = 0;
  std::vector<R> d;
  void resize(int W, int H) { w = W; h = H; d.assign((size_t)W * H, (R)0); }
  R& at(int x, int y) { return d[(size_t)y * w + x]; }
  R at(int x, int y) const { return d[(size_t)y * w + x]; }
};
template <class R> struct Img4 {
  int w = 0, h = 0;
  std::vector<std::array<R, 4>> d;
  void resize(int W, int H) {
    w = W; h = H;
    d.assign((size_t)W * H, std::array<R, 4>{0, 0, 0, -1});
  }
  std::array<R, 4>& at(int x, int y) { return d[(size_t)y * w + x]; }
  const std::array<R, 4>& at(int x, int y) const { return d[(size_t)y * w + x]; }
};

// ImageProc.cu:141-153 - inverse depth, 0 outside (0.2, 8) m.
template <class R> void depth_to_invdepth(const Img<R>& z, Img<R>& o) {
  o.resize(z.w, z.h);
  for (size_t i = 0; i < z.d.size(); ++i) {
    R v = z.d[i];
    o.d[i] = (v > (R)0.2 && v < (R)8) ? (R)1 / v : (R)0;
  }
}
// ImageProc.cu:166-175 - dst(y,x) = src(2y,2x); the last row and column,
// which the reference leaves unwritten, are set to 0 (no depth).
template <class R> void pyr_down_depth(const Img<R>& s, Img<R>& o, int W, int H) {
  o.resize(W, H);
  for (int y = 0; y < H - 1; ++y)
    for (int x = 0; x < W - 1; ++x) o.at(x, y) = s.at(2 * x, 2 * y);
}
// Departure from RGBDTracking.cu:146 (a float kernel run over float4 data):
// the reference-model pyramid subsamples whole float4 pixels, last row and
// column invalid like the depth pyramid.
template <class R> void pyr_down_vmap(const Img4<R>& s, Img4<R>& o, int W, int H) {
  o.resize(W, H);
  for (int y = 0; y < H - 1; ++y)
    for (int x = 0; x < W - 1; ++x) o.at(x, y) = s.at(2 * x, 2 * y);
}
// ImageProc.cu:189-209 - camera-frame vertex from inverse depth, valid iff
// invz > 0 and z < cutoff; invalid pixels are (0,0,0,-1).
template <class R> void vertex_map(const Img<R>& inv, Img4<R>& o, R invfx, R invfy, R cx, R cy, R cut) {
  o.resize(inv.w, inv.h);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < inv.h; ++y)
    for (int x = 0; x < inv.w; ++x) {
      R iz = inv.at(x, y);
      R z = (R)1 / iz;
      if (iz > 0 && z < cut) o.at(x, y) = {z * ((R)x - cx) * invfx, z * ((R)y - cy) * invfy, z, (R)1};
      else o.at(x, y) = {0, 0, 0, (R)-1};
    }
}
// ImageProc.cu:222-249 - forward-difference normal
// normalize((v(y+1,x)-v) x (v(y,x+1)-v)); last row/column invalid.
template <class R> void normal_map(const Img4<R>& v, Img4<R>& o) {
  o.resize(v.w, v.h);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < v.h; ++y)
    for (int x = 0; x < v.w; ++x) {
      auto& n = o.at(x, y);
      n = {0, 0, 0, (R)-1};
      if (x == v.w - 1 || y == v.h - 1) continue;
      const auto& a = v.at(x, y);
      const auto& b = v.at(x + 1, y);
      const auto& c = v.at(x, y + 1);
      if (!(a[3] > 0 && b[3] > 0 && c[3] > 0)) continue;
      R p[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      R q[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      R r[3] = {p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]};
      R n2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      if (n2 > 0) {
        R nn = std::sqrt(n2);
        for (int k = 0; k < 3; ++k) r[k] = r[k] / nn;
      }
      n = {r[0], r[1], r[2], (R)1};
    }
}
// ImageProc.cu:7-25 - central differences, 0 on the border.
template <class R> void gradient(const Img<R>& s, Img<R>& gx, Img<R>& gy) {
  gx.resize(s.w, s.h);
  gy.resize(s.w, s.h);
#pragma omp parallel for schedule(static)
  for (int y = 1; y < s.h - 1; ++y)
    for (int x = 1; x < s.w - 1; ++x) {
      gx.at(x, y) = (s.at(x + 1, y) - s.at(x - 1, y)) * (R)0.5;
      gy.at(x, y) = (s.at(x, y + 1) - s.at(x, y - 1)) * (R)0.5;
    }
}
// OpenCV pyrDown (used at RGBDTracking.cu:76,112): 5x5 binomial filter
// [1 4 6 4 1]/16 in each direction, BORDER_REFLECT_101, keep even pixels.
// Order: vertical weighted sum of each of the 5 columns, then horizontal.
inline int reflect101(int i, int n) {
  if (i < 0) i = -i;
  if (i >= n) i = 2 * n - 2 - i;
  return i;
}
template <class R> void pyr_down_gauss(const Img<R>& s, Img<R>& o, int W, int H) {
  o.resize(W, H);
  const R k[5] = {(R)0.0625, (R)0.25, (R)0.375, (R)0.25, (R)0.0625};
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      R col[5];
      for (int j = 0; j < 5; ++j) {
        int sx = reflect101(2 * x + j - 2, s.w);
        R acc = k[0] * s.at(sx, reflect101(2 * y - 2, s.h));
        for (int i = 1; i < 5; ++i) acc = acc + k[i] * s.at(sx, reflect101(2 * y + i - 2, s.h));
        col[j] = acc;
      }
      R acc = k[0] * col[0];
      for (int j = 1; j < 5; ++j) acc = acc + k[j] * col[j];
      o.at(x, y) = acc;
    }
}

// TrackingUtils.h:60-71 - bilinear interpolation.
template <class R> inline R bilerp(const Img<R>& m, R x, R y) {
  int u = (int)std::floor(x), v = (int)std::floor(y);
  R cx = x - (R)u, cy = y - (R)v;
  return (m.at(u, v) * ((R)1 - cx) + m.at(u + 1, v) * cx) * ((R)1 - cy) +
         (m.at(u, v + 1) * ((R)1 - cx) + m.at(u + 1, v + 1) * cx) * cy;
}

// ---------------------------------------------------------------- SE(3)
// Exponential of a twist xi = (rho, omega) (translation first, the tangent
// order of Sophus::SE3::exp used at RGBDTracking.cu:195), by Rodrigues'
// formula: R = I + A W + B W^2, t = (I + B W + C W^2) rho with
// A = sin(th)/th, B = (1-cos th)/th^2, C = (th - sin th)/th^3.
inline void se3_exp(const double xi[6], double T[16]) {
  const double* w = xi + 3;
  double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  double th = std::sqrt(th2);
  double A, B, C;
  if (th < 1e-8) {
    A = 1.0 - th2 / 6.0;
    B = 0.5 - th2 / 24.0;
    C = 1.0 / 6.0 - th2 / 120.0;
  } else {
    A = std::sin(th) / th;
    B = (1.0 - std::cos(th)) / th2;
    C = (th - std::sin(th)) / (th2 * th);
  }
  double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  double W2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W2[3 * i + j] = W[3 * i] * W[j] + W[3 * i + 1] * W[3 + j] + W[3 * i + 2] * W[6 + j];
  double Rm[9], V[9];
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

// Eigen 3.3 LDLT (used at RGBDTracking.cu:185): in-place lower LDL^T with
// diagonal pivoting (largest remaining |diagonal|), then solve with the
// pseudo-inverse of D (|d| <= DBL_MIN -> component set to 0).
inline void ldlt_solve6(const double Ain[36], const double bin[6], double x[6]) {
  const int n = 6;
  double A[36];
  std::memcpy(A, Ain, sizeof(A));
  int tr[6];
  bool zero = false;
  for (int k = 0; k < n; ++k) {
    int big = k;
    double bv = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + i]) > bv) { bv = std::fabs(A[i * n + i]); big = i; }
    tr[k] = big;
    if (big != k) {
      for (int j = 0; j < k; ++j) std::swap(A[k * n + j], A[big * n + j]);
      for (int i = big + 1; i < n; ++i) std::swap(A[i * n + k], A[i * n + big]);
      std::swap(A[k * n + k], A[big * n + big]);
      for (int i = k + 1; i < big; ++i) {
        double t = A[i * n + k];
        A[i * n + k] = A[big * n + i];
        A[big * n + i] = t;
      }
    }
    double temp[6];
    if (k > 0) {
      for (int j = 0; j < k; ++j) temp[j] = A[j * n + j] * A[k * n + j];
      double s = 0;
      for (int j = 0; j < k; ++j) s += A[k * n + j] * temp[j];
      A[k * n + k] -= s;
      for (int i = k + 1; i < n; ++i) {
        double r = 0;
        for (int j = 0; j < k; ++j) r += A[i * n + j] * temp[j];
        A[i * n + k] -= r;
      }
    }
    double akk = A[k * n + k];
    bool valid = std::fabs(akk) > 0.0;
    if (k == 0 && !valid) {
      for (int j = 0; j < n; ++j) tr[j] = j;
      zero = true;
      break;
    }
    if (valid)
      for (int i = k + 1; i < n; ++i) A[i * n + k] /= akk;
  }
  (void)zero;
  double y[6];
  std::memcpy(y, bin, sizeof(y));
  for (int k = 0; k < n; ++k) std::swap(y[k], y[tr[k]]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) y[i] -= A[i * n + j] * y[j];
  const double tol = std::numeric_limits<double>::min();
  for (int i = 0; i < n; ++i) {
    double d = A[i * n + i];
    if (std::fabs(d) > tol) y[i] /= d;
    else y[i] = 0.0;
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = i + 1; j < n; ++j) y[i] -= A[j * n + i] * y[j];
  for (int k = n - 1; k >= 0; --k) std::swap(y[k], y[tr[k]]);
  std::memcpy(x, y, sizeof(y));
}

// TrackingUtils.h:45-58 - unpack the 28 upper-triangle products of
// [J | r] into the 6x6 H (row-major) and the 6-vector b.
template <class R> inline void unpack29(const R* s, R H[36], R b[6]) {
  int k = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 7; ++j) {
      R v = s[k++];
      if (j == 6) b[i] = v;
      else H[6 * i + j] = H[6 * j + i] = v;
    }
}

// ---------------------------------------------------------------- tracker
enum { MODE_RGB = 0, MODE_DEPTH = 1, MODE_RGBD = 2 };

struct TrackerBase {
  virtual ~TrackerBase() {}
  virtual void set_ref_image(const float* g) = 0;
  virtual void set_ref_depth(const float* d) = 0;
  virtual void set_cur_image(const float* g) = 0;
  virtual void set_cur_depth(const float* d) = 0;
  virtual void set_ref_model(const float* vmap4) = 0;
  virtual void set_ref_model_d(const double* vmap4) = 0;
  virtual void step_sums(int lvl, const double T[16], int kind, double out[29], double stats[4]) = 0;
  virtual int get_transform(const double init[16], int swap, double out[16]) = 0;
  virtual void swap_buffers() = 0;
  virtual int get_level(int which, int lvl, double* out) = 0;
};

template <class R> struct Tracker : TrackerBase {
  int W, H, L, mode;
  int iters[8];
  double Kd[8][4];
  R fx[8], fy[8], cx[8], cy[8], invfx[8], invfy[8];
  int w[8], h[8];
  float wt_rgb_h, wt_rgb_b, wt_icp_h, wt_icp_b;
  std::vector<Img<R>> cur_int, ref_int, gxs, gys, cur_inv, ref_inv;
  std::vector<Img4<R>> cur_v, cur_n, ref_v, ref_n, warped, resid;

  Tracker(int W_, int H_, const double K[4], int L_, const int* it, int mode_, const float wts[4])
      : W(W_), H(H_), L(L_), mode(mode_) {
    wt_rgb_h = wts[0]; wt_rgb_b = wts[1]; wt_icp_h = wts[2]; wt_icp_b = wts[3];
    for (int l = 0; l < L; ++l) {
      iters[l] = it[l];
      double s = (double)(1 << l);
      Kd[l][0] = K[0] / s; Kd[l][1] = K[1] / s; Kd[l][2] = K[2] / s; Kd[l][3] = K[3] / s;
      fx[l] = (R)Kd[l][0]; fy[l] = (R)Kd[l][1]; cx[l] = (R)Kd[l][2]; cy[l] = (R)Kd[l][3];
      invfx[l] = (R)(1.0 / Kd[l][0]);
      invfy[l] = (R)(1.0 / Kd[l][1]);
      w[l] = W >> l;
      h[l] = H >> l;
    }
    for (auto* v : {&cur_int, &ref_int, &gxs, &gys, &cur_inv, &ref_inv}) v->resize(L);
    for (auto* v : {&cur_v, &cur_n, &ref_v, &ref_n, &warped, &resid}) v->resize(L);
    for (int l = 0; l < L; ++l) {
      for (auto* v : {&cur_int, &ref_int, &gxs, &gys, &cur_inv, &ref_inv}) (*v)[l].resize(w[l], h[l]);
      for (auto* v : {&cur_v, &cur_n, &ref_v, &ref_n, &warped, &resid}) (*v)[l].resize(w[l], h[l]);
    }
  }

  void image_pyramid(const float* g, std::vector<Img<R>>& pyr, bool grads) {
    for (int i = 0; i < W * H; ++i) pyr[0].d[i] = (R)g[i];
    for (int l = 1; l < L; ++l) pyr_down_gauss(pyr[l - 1], pyr[l], w[l], h[l]);
    if (grads)
      for (int l = 0; l < L; ++l) gradient(pyr[l], gxs[l], gys[l]);
  }
  // RGBDTracking.cu:80-100 / 118-138.
  void depth_pyramid(const float* d, std::vector<Img<R>>& inv, std::vector<Img4<R>>& v, std::vector<Img4<R>>& n) {
    Img<R> z;
    z.resize(W, H);
    for (int i = 0; i < W * H; ++i) z.d[i] = (R)d[i];
    for (int l = 0; l < L; ++l) {
      if (l == 0) depth_to_invdepth(z, inv[0]);
      else pyr_down_depth(inv[l - 1], inv[l], w[l], h[l]);
      vertex_map(inv[l], v[l], invfx[l], invfy[l], cx[l], cy[l], (R)3);
      normal_map(v[l], n[l]);
    }
  }
  void set_ref_image(const float* g) override { image_pyramid(g, ref_int, false); }
  void set_cur_image(const float* g) override { image_pyramid(g, cur_int, true); }
  void set_ref_depth(const float* d) override { depth_pyramid(d, ref_inv, ref_v, ref_n); }
  void set_cur_depth(const float* d) override { depth_pyramid(d, cur_inv, cur_v, cur_n); }
  // RGBDTracking.cu:140-150.
  void model_pyramid() {
    for (int l = 0; l < L; ++l) {
      if (l > 0) pyr_down_vmap(ref_v[l - 1], ref_v[l], w[l], h[l]);
      normal_map(ref_v[l], ref_n[l]);
    }
  }
  void set_ref_model(const float* vm) override {
    for (int i = 0; i < W * H; ++i)
      for (int c = 0; c < 4; ++c) ref_v[0].d[i][c] = (R)vm[4 * i + c];
    model_pyramid();
  }
  void set_ref_model_d(const double* vm) override {
    for (int i = 0; i < W * H; ++i)
      for (int c = 0; c < 4; ++c) ref_v[0].d[i][c] = (R)vm[4 * i + c];
    model_pyramid();
  }
  void swap_buffers() override {
    std::swap(ref_v, cur_v);
    std::swap(ref_n, cur_n);
    std::swap(ref_inv, cur_inv);
    std::swap(ref_int, cur_int);
  }

  // RGBDTracking.cu:305-406 - ICP: per reference pixel, projective
  // association into the current frame and the point-to-plane row
  // [n, v x n, n.(v_c - v)]; 28 upper-triangle products + match count.
  void icp_sums(int l, const double Td[16], double s[29]) {
    Pose<R> T = pose_from<R>(Td);
    const R angle_th = (R)std::sin(20.f * 3.14159254f / 180.f);
    const R dist_th = (R)0.01;
    for (int i = 0; i < 29; ++i) s[i] = 0;
    const int cols = w[l], rows = h[l];
#pragma omp parallel for schedule(static) reduction(+ : s[:29])
    for (int y = 0; y < rows; ++y)
      for (int x = 0; x < cols; ++x) {
        R row[7] = {0, 0, 0, 0, 0, 0, 0};
        bool found = false;
        const auto& vl = ref_v[l].at(x, y);
        if (vl[3] >= 0) {
          R p[3] = {vl[0], vl[1], vl[2]}, v[3];
          xform(T, p, v);
          R invz = (R)1 / v[2];
          int u = floor_sat(fx[l] * v[0] * invz + cx[l] + (R)0.5);
          int vv = floor_sat(fy[l] * v[1] * invz + cy[l] + (R)0.5);
          if (!(u < 0 || vv < 0 || u >= cols || vv >= rows)) {
            const auto& vc = cur_v[l].at(u, vv);
            if (vc[3] >= 0) {
              const auto& nl = ref_n[l].at(x, y);
              R nn[3] = {nl[0], nl[1], nl[2]}, n[3];
              rot(T, nn, n);
              const auto& nc = cur_n[l].at(u, vv);
              R dv[3] = {v[0] - vc[0], v[1] - vc[1], v[2] - vc[2]};
              R dist = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
              R cr[3] = {nc[1] * n[2] - nc[2] * n[1], nc[2] * n[0] - nc[0] * n[2], nc[0] * n[1] - nc[1] * n[0]};
              R ang = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
              found = ang < angle_th && dist <= dist_th && nl[3] > 0 && nc[3] > 0;
              if (found) {
                row[0] = n[0]; row[1] = n[1]; row[2] = n[2];
                row[3] = v[1] * n[2] - v[2] * n[1];
                row[4] = v[2] * n[0] - v[0] * n[2];
                row[5] = v[0] * n[1] - v[1] * n[0];
                row[6] = n[0] * (vc[0] - v[0]) + n[1] * (vc[1] - v[1]) + n[2] * (vc[2] - v[2]);
              }
            }
          }
        }
        if (!found) continue;
        int k = 0;
        for (int i = 0; i < 7; ++i)
          for (int j = i; j < 7; ++j) s[k++] += (double)(row[i] * row[j]);
        s[28] += 1.0;
      }
  }

  // ImageProc.cu:42-61 - reference pixels with inverse depth, warped into
  // the current frame: p = (R K^-1)(x,y,1) z + t.
  void transform_ref(int l, const double Td[16]) {
    double Ki[9] = {1.0 / Kd[l][0], 0, -Kd[l][2] / Kd[l][0], 0, 1.0 / Kd[l][1], -Kd[l][3] / Kd[l][1], 0, 0, 1};
    R RK[9], t[3];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j)
        RK[3 * i + j] = (R)(Td[4 * i] * Ki[j] + Td[4 * i + 1] * Ki[3 + j] + Td[4 * i + 2] * Ki[6 + j]);
      t[i] = (R)Td[4 * i + 3];
    }
    const Img<R>& inv = ref_inv[l];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h[l]; ++y)
      for (int x = 0; x < w[l]; ++x) {
        R iz = inv.at(x, y);
        if (iz > 0) {
          R z = (R)1 / iz;
          R p[3];
          for (int i = 0; i < 3; ++i) p[i] = (RK[3 * i] * (R)x + RK[3 * i + 1] * (R)y + RK[3 * i + 2]) * z + t[i];
          warped[l].at(x, y) = {p[0], p[1], p[2], (R)1};
        } else {
          warped[l].at(x, y) = {0, 0, 0, (R)-1};
        }
      }
  }

  // TrackingUtils.h:73-220,578-607 and RGBDTracking.cu:225-292 -
  // photometric residuals, their spread, and Huber-weighted J^T J.
  void rgb_sums(int l, const double Td[16], double s[29], double stats[4]) {
    transform_ref(l, Td);
    const int cols = w[l], rows = h[l];
    double sr = 0, cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : sr, cnt)
    for (int y = 0; y < rows; ++y)
      for (int x = 0; x < cols; ++x) {
        const auto& p = warped[l].at(x, y);
        R r = 0, gx = 0, gy = 0;
        bool ok = false;
        if (p[3] > 0) {
          R u = fx[l] * p[0] / p[2] + cx[l];
          R v = fy[l] * p[1] / p[2] + cy[l];
          if (u > 0 && v > 0 && u < (R)(cols - 1) && v < (R)(rows - 1)) {
            R ref = ref_int[l].at(x, y);
            R cur = bilerp(cur_int[l], u, v);
            r = cur - ref;
            gx = bilerp(gxs[l], u, v);
            gy = bilerp(gys[l], u, v);
            ok = (gx * gx + gy * gy) > 0 && std::isfinite((double)r);
          }
        }
        resid[l].at(x, y) = {r, gx, gy, (R)(ok ? 0.5 : -0.5)};
        if (ok) { sr += (double)r; cnt += 1; }
      }
    R rsum = (R)sr, n = (R)cnt;
    R mean = rsum / n;
    double sv = 0;
#pragma omp parallel for schedule(static) reduction(+ : sv)
    for (int y = 0; y < rows; ++y)
      for (int x = 0; x < cols; ++x) {
        const auto& e = resid[l].at(x, y);
        if (e[3] > 0) sv += (double)((e[0] - mean) * (e[0] - mean));
      }
    R sigma = std::sqrt((R)sv / (n - (R)1));
    R th = (R)(4.685 * (double)sigma);
    for (int i = 0; i < 29; ++i) s[i] = 0;
#pragma omp parallel for schedule(static) reduction(+ : s[:29])
    for (int y = 0; y < rows; ++y)
      for (int x = 0; x < cols; ++x) {
        const auto& e = resid[l].at(x, y);
        if (!(e[3] > 0)) continue;
        const auto& p = warped[l].at(x, y);
        R zinv = (R)1 / p[2];
        R wt = 1;
        if (std::fabs(e[0]) > th) wt = th / std::fabs(e[0]);
        R row[7];
        row[0] = e[1] * fx[l] * zinv;
        row[1] = e[2] * fy[l] * zinv;
        row[2] = -(row[0] * p[0] + row[1] * p[1]) * zinv;
        row[3] = row[2] * p[1] - e[2] * fy[l];
        row[4] = -row[2] * p[0] + e[1] * fx[l];
        row[5] = -row[0] * p[1] + row[1] * p[0];
        row[6] = -e[0];
        int k = 0;
        for (int i = 0; i < 7; ++i)
          for (int j = i; j < 7; ++j) s[k++] += (double)(wt * row[i] * row[j]);
        s[28] += 1.0;
      }
    if (stats) {
      stats[0] = (double)rsum; stats[1] = (double)n; stats[2] = (double)mean; stats[3] = (double)th;
    }
  }

  void step_sums(int lvl, const double T[16], int kind, double out[29], double stats[4]) override {
    if (kind == 0) icp_sums(lvl, T, out);
    else rgb_sums(lvl, T, out, stats);
    // Round the exact sums to the working type once.
    for (int i = 0; i < 29; ++i) out[i] = (double)(R)out[i];
  }

  // RGBDTracking.cu:152-214 - coarse-to-fine Gauss-Newton; each iteration
  // solves H xi = b in double, applies T <- exp(xi) T, and remembers the
  // updated pose whenever the pre-update error improved on the level's best.
  int get_transform(const double init[16], int swap, double out[16]) override {
    double est[16], last_ok[16];
    std::memcpy(est, init, sizeof(est));
    std::memcpy(last_ok, init, sizeof(est));
    int good = 1;
    for (int l = L - 1; l >= 0 && good; --l) {
      float last_err = std::numeric_limits<float>::max();
      for (int it = 0; it < iters[l]; ++it) {
        double si[29], sr[29];
        R H[36] = {0}, b[6] = {0};
        R rsum = 0, nres = 0;
        if (mode != MODE_DEPTH) {
          rgb_sums(l, est, sr, nullptr);
          R s[29], Hr[36], br[6];
          for (int i = 0; i < 29; ++i) s[i] = (R)sr[i];
          unpack29(s, Hr, br);
          R wh = mode == MODE_RGBD ? (R)wt_rgb_h : (R)1, wb = mode == MODE_RGBD ? (R)wt_rgb_b : (R)1;
          for (int i = 0; i < 36; ++i) H[i] = H[i] + wh * Hr[i];
          for (int i = 0; i < 6; ++i) b[i] = b[i] + wb * br[i];
          rsum = s[27];
          nres = s[28];
        }
        if (mode != MODE_RGB) {
          icp_sums(l, est, si);
          R s[29], Hi[36], bi[6];
          for (int i = 0; i < 29; ++i) s[i] = (R)si[i];
          unpack29(s, Hi, bi);
          R wh = mode == MODE_RGBD ? (R)wt_icp_h : (R)1, wb = mode == MODE_RGBD ? (R)wt_icp_b : (R)1;
          for (int i = 0; i < 36; ++i) H[i] = H[i] + wh * Hi[i];
          for (int i = 0; i < 6; ++i) b[i] = b[i] + wb * bi[i];
          rsum = s[27];
        }
        // error = sqrt(residualSum) / (numResidual + 6); in depth-only mode
        // the reference never sets numResidual, taken here as 0.
        float err = (float)(std::sqrt(rsum) / (nres + (R)6));
        double Hd[36], bd[6], xi[6];
        for (int i = 0; i < 36; ++i) Hd[i] = (double)H[i];
        for (int i = 0; i < 6; ++i) bd[i] = (double)b[i];
        ldlt_solve6(Hd, bd, xi);
        if (std::isnan(xi[0])) {
          good = 0;
          break;
        }
        double E[16];
        se3_exp(xi, E);
        mul_d(E, est, est);
        if (err < last_err) {
          std::memcpy(last_ok, est, sizeof(est));
          last_err = err;
        }
      }
    }
    // RGBDTracking.cu:187-191: a NaN update returns the identity before the
    // buffer swap
    if (swap && good) swap_buffers();
    if (!good) identity_d(last_ok);
    std::memcpy(out, last_ok, sizeof(last_ok));
    return good;
  }

  // which: 0 cur intensity, 1 ref intensity, 2 gx, 3 gy, 4 cur invdepth,
  // 5 ref invdepth (1 channel); 6 cur vmap, 7 cur nmap, 8 ref vmap,
  // 9 ref nmap, 10 warped, 11 residual (4 channels).
  int get_level(int which, int l, double* out) override {
    if (l < 0 || l >= L) return -1;
    const Img<R>* one[6] = {&cur_int[l], &ref_int[l], &gxs[l], &gys[l], &cur_inv[l], &ref_inv[l]};
    const Img4<R>* four[6] = {&cur_v[l], &cur_n[l], &ref_v[l], &ref_n[l], &warped[l], &resid[l]};
    if (which < 6) {
      for (size_t i = 0; i < one[which]->d.size(); ++i) out[i] = (double)one[which]->d[i];
      return 1;
    }
    if (which < 12) {
      const Img4<R>* m = four[which - 6];
      for (size_t i = 0; i < m->d.size(); ++i)
        for (int c = 0; c < 4; ++c) out[4 * i + c] = (double)m->d[i][c];
      return 4;
    }
    return -1;
  }
};

// System.cpp:52 (depth.convertTo(CV_32F, 1/DepthScale)) restated:
// z = float(d) * float(1/scale).
inline void depth_u16_to_f32(const uint16_t* d, float* z, size_t n, double scale) {
  float s = (float)(1.0 / scale);
  for (size_t i = 0; i < n; ++i) z[i] = (float)d[i] * s;
}
// Intensity: mode 0 = FP32 luma 0.299R + 0.587G + 0.114B; mode 1 = OpenCV's
// 8-bit RGB2GRAY fixed-point rounding (System.cpp:47) as float.
inline void rgb_to_intensity(const uint8_t* rgb, float* g, size_t n, int mode) {
  for (size_t i = 0; i < n; ++i) {
    int r = rgb[3 * i], gg = rgb[3 * i + 1], b = rgb[3 * i + 2];
    if (mode == 0) g[i] = 0.299f * (float)r + 0.587f * (float)gg + 0.114f * (float)b;
    else g[i] = (float)((r * 4899 + gg * 9617 + b * 1868 + (1 << 13)) >> 14);
  }
}

// ---------------------------------------------------------------- pipeline
// FullSystem/src/Tracking.cpp:139-197 replayed on one volume: frame 0 is
// fused at the identity; each later frame ray-casts the volume at the last
// pose, tracks the new frame against it, rejects |t| > 0.1 m, and fuses at
// T_w_cur = T_w_last * DT^-1.
struct Pipeline {
  std::unique_ptr<MapBase> map;
  std::unique_ptr<TrackerBase> trk;
  int W, H, prec, gray_mode;
  double K[4], depth_scale;
  double T_last[16];
  int frames = 0;
  std::vector<float> z, g;
  std::vector<double> vmap;
  double t_raycast = 0, t_track = 0, t_fuse = 0, t_prep = 0;
};

}  // namespace orc

using namespace orc;

extern "C" {

// voxfmt: 0 = int16+u8 (reference), 1 = f32+f32, 2 = f16+u16; precision 64
// ignores voxfmt (double sdf + double weight).
void* orc_map_new(int precision, int voxfmt, double voxel, double trunc, int maxw, double dmin, double dmax) {
  if (precision == 64) return new Map<VF64>(voxel, trunc, maxw, dmin, dmax);
  if (voxfmt == 0) return new Map<VRef>(voxel, trunc, maxw, dmin, dmax);
  if (voxfmt == 2) return new Map<VF16>(voxel, trunc, maxw, dmin, dmax);
  return new Map<VF32>(voxel, trunc, maxw, dmin, dmax);
}
void orc_map_free(void* m) { delete (MapBase*)m; }
int orc_map_fuse(void* m, const float* depth, int W, int H, const double K[4], const double T[16]) {
  return ((MapBase*)m)->fuse(depth, W, H, K, T);
}
int orc_map_allocate(void* m, const float* depth, int W, int H, const double K[4], const double T[16]) {
  return ((MapBase*)m)->allocate(depth, W, H, K, T);
}
int orc_map_visible(void* m, int W, int H, const double K[4], const double T[16], int32_t* keys, int cap) {
  return ((MapBase*)m)->visible(W, H, K, T, keys, cap);
}
int orc_map_num_blocks(void* m) { return ((MapBase*)m)->num_blocks(); }
int orc_map_fuse_map(void* dst, void* src, const double Tdst[16], const double Tsrc[16]) {
  return ((MapBase*)dst)->fuse_map((MapBase*)src, Tdst, Tsrc);
}
int orc_map_export(void* m, int32_t* keys, double* sdf, double* wt, int cap) {
  return ((MapBase*)m)->export_blocks(keys, sdf, wt, cap);
}
void orc_map_import(void* m, const int32_t* keys, const double* sdf, const double* wt, int n) {
  ((MapBase*)m)->import_blocks(keys, sdf, wt, n);
}
int orc_map_raycast(void* m, int W, int H, const double K[4], const double T[16], double* vmap, double* zmin,
                    double* zmax) {
  return ((MapBase*)m)->raycast(W, H, K, T, vmap, zmin, zmax);
}
long orc_map_mesh(void* m, int32_t* bkeys, int32_t* tcounts, int16_t* cases, double* tris, double* normals,
                  long tri_cap, int block_cap) {
  return ((MapBase*)m)->mesh(bkeys, tcounts, cases, tris, normals, tri_cap, block_cap);
}
// Threads the loops of this build use (1 without OpenMP).
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
int orc_ref_hash(int x, int y, int z, int n) { return ref_hash(x, y, z, n); }
int orc_vox_to_block(int v) { return vox_to_block(v); }
int orc_vox_to_local(int v) { return vox_to_local(v); }
int16_t orc_pack_sdf(float f) { return Vox<VRef>::pack(f); }
uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
const uint64_t* orc_mc_table(void) { return decoded_tri_table(); }
// One marching cube (precision 32 or 64): returns the cube index, writes up
// to 5 triangles (45 values) and normals, *ntri their count.
int orc_mc_cube(int precision, const int p[3], const double s[8], double vs, double* tris, double* normals, int* ntri) {
  if (precision == 64) return mc_cube<double>(p, s, vs, tris, normals, ntri);
  float sf[8], tf[45], nf[45];
  for (int i = 0; i < 8; ++i) sf[i] = (float)s[i];
  int c = mc_cube<float>(p, sf, (float)vs, tf, nf, ntri);
  for (int i = 0; i < 9 * *ntri; ++i) { tris[i] = tf[i]; normals[i] = nf[i]; }
  return c;
}

void orc_se3_exp(const double xi[6], double T[16]) { se3_exp(xi, T); }
void orc_ldlt_solve6(const double A[36], const double b[6], double x[6]) { ldlt_solve6(A, b, x); }

void orc_image_op(int op, int precision, const float* src, int W, int H, double* out, const double p[4]) {
  // op 0: pyrdown_gauss (out W/2 x H/2), 1: gradient (out gx then gy),
  // 2: depth->invdepth, 3: vertex map from invdepth (p = invfx,invfy,cx,cy),
  // 4: normal map from a 4-channel vertex map in src.
  auto run = [&](auto tag) {
    typedef decltype(tag) R;
    Img<R> s;
    s.resize(W, H);
    if (op != 4)
      for (int i = 0; i < W * H; ++i) s.d[i] = (R)src[i];
    if (op == 0) {
      Img<R> o;
      pyr_down_gauss(s, o, W / 2, H / 2);
      for (size_t i = 0; i < o.d.size(); ++i) out[i] = (double)o.d[i];
    } else if (op == 1) {
      Img<R> gx, gy;
      gradient(s, gx, gy);
      for (int i = 0; i < W * H; ++i) { out[i] = (double)gx.d[i]; out[W * H + i] = (double)gy.d[i]; }
    } else if (op == 2) {
      Img<R> o;
      depth_to_invdepth(s, o);
      for (int i = 0; i < W * H; ++i) out[i] = (double)o.d[i];
    } else if (op == 3) {
      Img4<R> o;
      vertex_map(s, o, (R)p[0], (R)p[1], (R)p[2], (R)p[3], (R)3);
      for (int i = 0; i < W * H; ++i)
        for (int c = 0; c < 4; ++c) out[4 * i + c] = (double)o.d[i][c];
    } else if (op == 4) {
      Img4<R> v, o;
      v.resize(W, H);
      for (int i = 0; i < W * H; ++i)
        for (int c = 0; c < 4; ++c) v.d[i][c] = (R)src[4 * i + c];
      normal_map(v, o);
      for (int i = 0; i < W * H; ++i)
        for (int c = 0; c < 4; ++c) out[4 * i + c] = (double)o.d[i][c];
    }
  };
  if (precision == 64) run(0.0);
  else run(0.0f);
}

void* orc_trk_new(int precision, int W, int H, const double K[4], int L, const int* iters, int mode,
                  const float wts[4]) {
  if (precision == 64) return new Tracker<double>(W, H, K, L, iters, mode, wts);
  return new Tracker<float>(W, H, K, L, iters, mode, wts);
}
void orc_trk_free(void* t) { delete (TrackerBase*)t; }
void orc_trk_set_ref_image(void* t, const float* g) { ((TrackerBase*)t)->set_ref_image(g); }
void orc_trk_set_ref_depth(void* t, const float* d) { ((TrackerBase*)t)->set_ref_depth(d); }
void orc_trk_set_cur_image(void* t, const float* g) { ((TrackerBase*)t)->set_cur_image(g); }
void orc_trk_set_cur_depth(void* t, const float* d) { ((TrackerBase*)t)->set_cur_depth(d); }
void orc_trk_set_ref_model(void* t, const float* v) { ((TrackerBase*)t)->set_ref_model(v); }
void orc_trk_step_sums(void* t, int lvl, const double T[16], int kind, double out[29], double stats[4]) {
  ((TrackerBase*)t)->step_sums(lvl, T, kind, out, stats);
}
int orc_trk_get_transform(void* t, const double init[16], int swap, double out[16]) {
  return ((TrackerBase*)t)->get_transform(init, swap, out);
}
void orc_trk_swap(void* t) { ((TrackerBase*)t)->swap_buffers(); }
int orc_trk_get_level(void* t, int which, int lvl, double* out) { return ((TrackerBase*)t)->get_level(which, lvl, out); }

void orc_depth_u16_to_f32(const uint16_t* d, float* z, long n, double scale) { depth_u16_to_f32(d, z, (size_t)n, scale); }
void orc_rgb_to_intensity(const uint8_t* rgb, float* g, long n, int mode) { rgb_to_intensity(rgb, g, (size_t)n, mode); }

// Pipeline: map params (voxel, trunc, maxw, voxfmt), tracker params.
void* orc_pipe_new(int precision, int voxfmt, double voxel, double trunc, int maxw, int W, int H, const double K[4],
                   int L, const int* iters, int mode, const float wts[4], double depth_scale, int gray_mode) {
  if (W < 8 || H < 8 || W % 8 != 0 || H % 8 != 0) return nullptr;  // the ray cast's z-range grid
  Pipeline* p = new Pipeline();
  p->map.reset((MapBase*)orc_map_new(precision, voxfmt, voxel, trunc, maxw, 0.1, 3.0));
  p->trk.reset((TrackerBase*)orc_trk_new(precision, W, H, K, L, iters, mode, wts));
  p->W = W; p->H = H; p->prec = precision; p->gray_mode = gray_mode;
  std::memcpy(p->K, K, sizeof(p->K));
  p->depth_scale = depth_scale;
  identity_d(p->T_last);
  p->z.resize((size_t)W * H);
  p->g.resize((size_t)W * H);
  p->vmap.resize((size_t)W * H * 4);
  return p;
}
void orc_pipe_free(void* p) { delete (Pipeline*)p; }
void* orc_pipe_map(void* p) { return ((Pipeline*)p)->map.get(); }
void* orc_pipe_tracker(void* p) { return ((Pipeline*)p)->trk.get(); }

// One frame; returns 1 if tracked and fused, 0 if rejected (|t| > 0.1 m).
// T_out = camera-to-map pose of this frame.
int orc_pipe_step(void* pp, const uint16_t* depth, const uint8_t* rgb, double T_out[16], double times_ms[4]) {
  Pipeline* p = (Pipeline*)pp;
  typedef std::chrono::steady_clock clk;
  auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  const size_t n = (size_t)p->W * p->H;
  auto t0 = clk::now();
  depth_u16_to_f32(depth, p->z.data(), n, p->depth_scale);
  rgb_to_intensity(rgb, p->g.data(), n, p->gray_mode);
  auto t1 = clk::now();
  int ok = 1;
  double tr = 0, tt = 0, tf = 0;
  if (p->frames == 0) {
    p->trk->set_ref_image(p->g.data());
    p->trk->set_ref_depth(p->z.data());
    auto a = clk::now();
    p->map->fuse(p->z.data(), p->W, p->H, p->K, p->T_last);
    tf = ms(a, clk::now());
  } else {
    auto a = clk::now();
    p->map->raycast(p->W, p->H, p->K, p->T_last, p->vmap.data(), nullptr, nullptr);
    p->trk->set_ref_model_d(p->vmap.data());
    auto b = clk::now();
    p->trk->set_cur_image(p->g.data());
    p->trk->set_cur_depth(p->z.data());
    double I[16], DT[16];
    identity_d(I);
    int good = p->trk->get_transform(I, 1, DT);
    auto c = clk::now();
    // Tracking.cpp:178-187 gates on |t| alone: the identity that a NaN
    // update returns is accepted and fused at T_last.
    (void)good;
    double tn = std::sqrt(DT[3] * DT[3] + DT[7] * DT[7] + DT[11] * DT[11]);
    if (tn > 0.1) {
      ok = 0;
    } else {
      double Di[16];
      inverse_d(DT, Di);
      mul_d(p->T_last, Di, p->T_last);
      p->map->fuse(p->z.data(), p->W, p->H, p->K, p->T_last);
    }
    tr = ms(a, b);
    tt = ms(b, c);
    tf = ms(c, clk::now());
  }
  p->frames++;
  if (T_out) std::memcpy(T_out, p->T_last, sizeof(p->T_last));
  if (times_ms) {
    times_ms[0] = ms(t0, t1);
    times_ms[1] = tr;
    times_ms[2] = tt;
    times_ms[3] = tf;
  }
  return ok;
}

}  // extern "C"
