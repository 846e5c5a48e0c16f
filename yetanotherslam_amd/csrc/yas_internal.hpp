// Internal definitions shared by the HIP translation units of libyas.
//
// Arithmetic conventions (kept identical to the reference's kernels and
// evaluated in the order written; the library is compiled with
// -ffp-contract=off and IEEE division/sqrt, so each expression is one
// rounding per operation):
//   * rigid transform p' = R p + t, each row summed left to right;
//   * projection u = fx * x / z + cx ((fx*x)/z then + cx);
//   * float->int "round down" conversions saturate and map NaN to 0.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/yas.h"

// Qualifier of descriptor tables that only the host writes: kernels read them
// through the constant address space (scalar loads), and device pointers
// loaded from there are taken as global pointers by the compiler.  The host
// pass of a HIP source sees the kernels too and has no such address space.
#if defined(__HIP_DEVICE_COMPILE__)
#define YAS_CONSTANT __attribute__((address_space(4)))
#else
#define YAS_CONSTANT
#endif

// ------------------------------------------------------------------ errors
void yas_set_error(const char* fmt, ...);

#define YAS_HIP(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) {                                                                        \
      yas_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);     \
      return YAS_E_HIP;                                                                            \
    }                                                                                              \
  } while (0)

#define YAS_LAUNCH_CHECK()                                                                         \
  do {                                                                                             \
    hipError_t e_ = hipGetLastError();                                                             \
    if (e_ != hipSuccess) {                                                                        \
      yas_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
      return YAS_E_HIP;                                                                            \
    }                                                                                              \
  } while (0)

#define YAS_REQUIRE(cond, msg)                                                                     \
  do {                                                                                             \
    if (!(cond)) {                                                                                 \
      yas_set_error("%s", msg);                                                                    \
      return YAS_E_INVALID;                                                                        \
    }                                                                                              \
  } while (0)

// A kernel launch on stream s.  Given a pair of events, the launch records
// them at the start and the end of the kernel's own execution
// (hipExtLaunchKernel), which times the kernel without the dispatch gaps of
// events recorded around it.
template <typename F, typename... Args>
inline void launch_timed(F kernel, dim3 grid, dim3 block, hipStream_t s, const hipEvent_t* ev, Args... args) {
  if (ev) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev[0], ev[1], 0, args...);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

struct yas_ctx {
  int device;
  hipStream_t stream;
  bool own_stream;
  std::atomic<int> refs;  // the creator's reference plus one per object built on it
};

// Every object built on a context holds a reference to it, so the context
// (and its stream) outlives them whatever order they are destroyed in.
yas_ctx* yas_ctx_retain(yas_ctx* c);
void yas_ctx_release(yas_ctx* c);

namespace yas {

constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kMaxProbe = 65536;
constexpr int kBlockThreads = 256;

enum : int { kFlagPoolOverflow = 1, kFlagProbeLimit = 2, kFlagKeyRange = 4 };
// Map counters in device memory.
enum : int { kCtrAlloc = 0, kCtrVisible = 1, kCtrFlags = 2, kCtrWalk = 3, kCtrCount = 8 };

// ------------------------------------------------------------------ keys
// Block coordinates are packed as three 21-bit two's-complement fields.
__host__ __device__ inline bool key_in_range(int x, int y, int z) {
  return x >= -(1 << 20) && x < (1 << 20) && y >= -(1 << 20) && y < (1 << 20) && z >= -(1 << 20) &&
         z < (1 << 20);
}
__host__ __device__ inline uint64_t pack_key(int x, int y, int z) {
  return ((uint64_t)(uint32_t)(x & 0x1FFFFF)) | ((uint64_t)(uint32_t)(y & 0x1FFFFF) << 21) |
         ((uint64_t)(uint32_t)(z & 0x1FFFFF) << 42);
}
__host__ __device__ inline int sext21(uint64_t v) {
  int x = (int)(v & 0x1FFFFF);
  return (x & 0x100000) ? x - 0x200000 : x;
}
__host__ __device__ inline void unpack_key(uint64_t k, int& x, int& y, int& z) {
  x = sext21(k);
  y = sext21(k >> 21);
  z = sext21(k >> 42);
}
// Reference spatial hash (VoxelStructUtils.h:16-21): int32 wrapping
// multiplies, xor, remainder fixed into [0, n).  It picks the home slot.
// For a power-of-two n that remainder is the low bits of the hash.
__host__ __device__ inline int ref_hash(int x, int y, int z, int n) {
  uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
  if ((n & (n - 1)) == 0) return (int)(h & (uint32_t)(n - 1));
  int32_t r = (int32_t)h;
  r %= n;
  return r < 0 ? r + n : r;
}
// VoxelStructUtils.h:39-49 / :56-70.
__host__ __device__ inline int vox_to_block(int v) {
  if (v < 0) v -= 7;
  return v / 8;
}
__host__ __device__ inline int vox_to_local(int v) {
  int l = v % 8;
  return l < 0 ? l + 8 : l;
}

// ------------------------------------------------------------------ conversions
// Branch-free: floor, clamp into the int range, convert; 2^31 and above give
// INT32_MAX, NaN gives 0.
__device__ inline int floor_sat(float f) {
  const float g = floorf(f);
  int i = (int)fminf(fmaxf(g, -2147483648.0f), 2147483520.0f);
  i = g >= 2147483648.0f ? INT32_MAX : i;
  return f != f ? 0 : i;
}
__device__ inline int ceil_sat(float f) {
  if (f != f) return 0;
  f = ceilf(f);
  if (f <= -2147483648.0f) return INT32_MIN;
  if (f >= 2147483648.0f) return INT32_MAX;
  return (int)f;
}
__device__ inline int trunc_sat(float f) {
  if (f != f) return 0;
  f = truncf(f);
  if (f <= -2147483648.0f) return INT32_MIN;
  if (f >= 2147483648.0f) return INT32_MAX;
  return (int)f;
}

// ------------------------------------------------------------------ poses
struct PoseF {
  float r[9];
  float t[3];
};
struct PosePair {
  PoseF T;     // camera -> map
  PoseF Tinv;  // map -> camera
};
struct Mat44 {
  double m[16];
};

__host__ __device__ inline void xform(const PoseF& T, float x, float y, float z, float& ox, float& oy,
                                      float& oz) {
  ox = T.r[0] * x + T.r[1] * y + T.r[2] * z + T.t[0];
  oy = T.r[3] * x + T.r[4] * y + T.r[5] * z + T.t[1];
  oz = T.r[6] * x + T.r[7] * y + T.r[8] * z + T.t[2];
}
__host__ __device__ inline void rot3(const PoseF& T, float x, float y, float z, float& ox, float& oy,
                                     float& oz) {
  ox = T.r[0] * x + T.r[1] * y + T.r[2] * z;
  oy = T.r[3] * x + T.r[4] * y + T.r[5] * z;
  oz = T.r[6] * x + T.r[7] * y + T.r[8] * z;
}
// Double pose -> float pose, and its inverse formed in double
// (R^T, -(R^T t) with each sum left to right) before rounding.
__host__ __device__ inline void pose_pair_from(const double* T, PosePair& p) {
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      p.T.r[3 * i + j] = (float)T[4 * i + j];
      p.Tinv.r[3 * i + j] = (float)T[4 * j + i];
    }
    p.T.t[i] = (float)T[4 * i + 3];
    p.Tinv.t[i] = (float)(-(T[i] * T[3] + T[4 + i] * T[7] + T[8 + i] * T[11]));
  }
}
__host__ __device__ inline void mul44(const double* A, const double* B, double* C) {
  double t[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      t[4 * i + j] = A[4 * i] * B[j] + A[4 * i + 1] * B[4 + j] + A[4 * i + 2] * B[8 + j] + A[4 * i + 3] * B[12 + j];
  for (int i = 0; i < 16; ++i) C[i] = t[i];
}
__host__ __device__ inline void inv44(const double* T, double* o) {
  double t[16];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) t[4 * i + j] = T[4 * j + i];
    t[4 * i + 3] = -(T[i] * T[3] + T[4 + i] * T[7] + T[8 + i] * T[11]);
  }
  t[12] = t[13] = t[14] = 0.0;
  t[15] = 1.0;
  for (int i = 0; i < 16; ++i) o[i] = t[i];
}

struct Intr {
  float fx, fy, cx, cy, invfx, invfy;
};
// float copies of K and float(1/(double)f), as the reference derives them
// (VoxelMap.cu:827-832, RayTraceEngine.cu:14-19).
inline Intr make_intr(const float K[4]) {
  Intr k;
  k.fx = K[0];
  k.fy = K[1];
  k.cx = K[2];
  k.cy = K[3];
  k.invfx = (float)(1.0 / (double)k.fx);
  k.invfy = (float)(1.0 / (double)k.fy);
  return k;
}

// ------------------------------------------------------------------ voxels
// Storage variants of the one update rule of VoxelMap.cu:764-783:
// first observation -> (sdf, 1); else ((old*w + sdf)/(w+1), min(cap, w+1)).
struct VoxI16 {  // int16 sdf (x32767, truncation) | uint8 weight << 16
  typedef uint32_t T;
  static constexpr int kFormat = YAS_VOXEL_I16_U8;
  __device__ static inline float sdf(T v) { return (float)(int16_t)(v & 0xFFFFu) / 32767.0f; }
  __device__ static inline bool observed(T v) { return ((v >> 16) & 0xFFu) != 0; }
  __device__ static inline uint32_t pack(float f) {
    int i = trunc_sat(f * 32767.0f);
    i = i < -32768 ? -32768 : (i > 32767 ? 32767 : i);
    return (uint32_t)(uint16_t)(int16_t)i;
  }
  __device__ static inline T update(T v, float s, int) {
    int w = (int)((v >> 16) & 0xFFu);
    if (w == 0) return pack(s) | (1u << 16);
    float old = (float)(int16_t)(v & 0xFFFFu) / 32767.0f;
    uint32_t ns = pack((old * (float)w + s) / (float)(w + 1));
    int nw = w + 1 < 255 ? w + 1 : 255;
    return ns | ((uint32_t)nw << 16);
  }
  __device__ static inline int weight(T v) { return (int)((v >> 16) & 0xFFu); }
  // Map-to-map merge (VoxelMap.cu:373-375): weighted mean, weight capped at 255.
  __device__ static inline T merge(T v, float s, int wo, int) {
    const int w = weight(v);
    const float old = sdf(v);
    const uint32_t ns = pack((old * (float)w + s * (float)wo) / (float)(w + wo));
    const int nw = w + wo < 255 ? w + wo : 255;
    return ns | ((uint32_t)nw << 16);
  }
};
struct VoxF32 {  // float sdf, float weight
  typedef float2 T;
  static constexpr int kFormat = YAS_VOXEL_F32_F32;
  __device__ static inline float sdf(T v) { return v.x; }
  __device__ static inline bool observed(T v) { return v.y != 0.0f; }
  __device__ static inline T update(T v, float s, int maxw) {
    float w = v.y;
    if (w == 0.0f) return make_float2(s, 1.0f);
    float ns = (v.x * w + s) / (w + 1.0f);
    return make_float2(ns, fminf((float)maxw, w + 1.0f));
  }
  __device__ static inline int weight(T v) { return (int)v.y; }
  __device__ static inline T merge(T v, float s, int wo, int maxw) {
    const float w = v.y;
    const float ns = (v.x * w + s * (float)wo) / (w + (float)wo);
    return make_float2(ns, fminf((float)maxw, w + (float)wo));
  }
};
struct VoxF16 {  // half sdf | uint16 weight << 16
  typedef uint32_t T;
  static constexpr int kFormat = YAS_VOXEL_F16_U16;
  __device__ static inline float h2f(uint32_t h) {
    _Float16 x = __builtin_bit_cast(_Float16, (uint16_t)(h & 0xFFFFu));
    return (float)x;
  }
  __device__ static inline uint32_t f2h(float f) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)f); }
  __device__ static inline float sdf(T v) { return h2f(v); }
  __device__ static inline bool observed(T v) { return (v >> 16) != 0; }
  __device__ static inline T update(T v, float s, int maxw) {
    int w = (int)(v >> 16);
    if (w == 0) return f2h(s) | (1u << 16);
    float old = h2f(v);
    uint32_t ns = f2h((old * (float)w + s) / (float)(w + 1));
    int nw = w + 1 < maxw ? w + 1 : maxw;
    return ns | ((uint32_t)nw << 16);
  }
  __device__ static inline int weight(T v) { return (int)(v >> 16); }
  __device__ static inline T merge(T v, float s, int wo, int maxw) {
    const int w = weight(v);
    const float old = h2f(v);
    const uint32_t ns = f2h((old * (float)w + s * (float)wo) / (float)(w + wo));
    const int nw = w + wo < maxw ? w + wo : maxw;
    return ns | ((uint32_t)nw << 16);
  }
};

inline int voxel_bytes(int fmt) { return fmt == YAS_VOXEL_F32_F32 ? 8 : 4; }

// ------------------------------------------------------------------ map view
// Device view of one volume: open-addressed hash (keys/vals), the block
// pool (block_keys/voxels) and counters.
// One hash slot: the packed block key and its pool index side by side, so a
// probe is one 16-byte load.
struct alignas(16) HashSlot {
  unsigned long long key;  // kEmptyKey when free
  int val;                 // pool index, -1 if the pool overflowed
  int flags;               // kBlock* summary of the block's voxels
};
// Exact summaries of a block's 512 voxels, maintained by allocation (a new
// block is all unobserved) and integration (recomputed after every update).
enum : int { kBlockUnobserved = 1, kBlockFree = 2 };  // all w == 0 / all w > 0 with sdf == 1

struct MapView {
  HashSlot* slots;                 // [num_buckets]
  int* block_slot;                 // [num_blocks] hash slot of each pool block
  unsigned long long* block_keys;  // [num_blocks]
  void* voxels;                    // [num_blocks * 512] of the storage type
  int* visible;                    // [num_blocks] pool indices of visible blocks
  int* ctr;                        // counters (kCtr*)
  int num_buckets;
  int num_blocks;
  float voxel_size;
  float trunc;
  float depth_min, depth_max;
  int max_weight;
  int format;
};

// Linear-probe lookup from the reference hash's home slot: the pool index
// (-1 if absent) and the block's summary flags.
__device__ inline int find_block(const MapView& m, uint64_t key, int home, int* flags = nullptr) {
  int slot = home;
  const int lim = m.num_buckets < kMaxProbe ? m.num_buckets : kMaxProbe;
  for (int p = 0; p < lim; ++p) {
    // One 16-byte load.  The empty asm statement takes the pool index and
    // flags as register operands right after it, so the compiler cannot
    // defer their half of the slot to a second, dependent load on a match.
    uint4 e = reinterpret_cast<const uint4*>(m.slots)[slot];
    asm volatile("" : "+v"(e.z), "+v"(e.w));
    const unsigned long long k = ((unsigned long long)e.y << 32) | e.x;
    if (k == key) {
      if (flags) *flags = (int)e.w;
      return (int)e.z;
    }
    if (k == kEmptyKey) return -1;
    slot = slot + 1 == m.num_buckets ? 0 : slot + 1;
  }
  return -1;
}
__device__ inline int find_block_xyz(const MapView& m, int bx, int by, int bz, int* flags = nullptr) {
  if (!key_in_range(bx, by, bz)) return -1;
  return find_block(m, pack_key(bx, by, bz), ref_hash(bx, by, bz, m.num_buckets), flags);
}

// ------------------------------------------------------------------ frame outcome
// Tracking.cpp:178-187: reject |t| > 0.1 m, otherwise
// T_last <- T_last * DT^-1.  After a NaN update the reference's GetTransform
// returns the identity (RGBDTracking.cu:187-191), which passes the gate:
// the frame is accepted and fused at T_last * I^-1 = T_last.  Records the
// frame's pose and outcome and sets the gate that lets the fusion kernels
// run.  One thread.
__device__ inline void post_track_at(const double* DT, const int* good, double* T_last, double* pose_out,
                                     int* tracked, int* gate) {
  const bool converged = *good != 0;
  double tn = converged ? sqrt(DT[3] * DT[3] + DT[7] * DT[7] + DT[11] * DT[11]) : 0.0;
  int ok = !(tn > 0.1);
  if (ok && converged) {
    double Di[16];
    inv44(DT, Di);
    mul44(T_last, Di, T_last);
  }
  for (int i = 0; i < 16; ++i) pose_out[i] = T_last[i];
  *tracked = ok;
  *gate = ok;
}
// Copies that undo SwapFrameBuffer's exchange of the reference and current
// intensity and inverse-depth pyramids (dst <- src, n floats each), for a
// frame whose update was NaN: the reference returns before the swap
// (RGBDTracking.cu:187-191), and the asynchronous pipeline, which swaps
// before the outcome is known, puts them back on the device.
struct RefRestore {
  float* dst[10];
  const float* src[10];
  int n[10];
  int count;
};

// StereoInitialization (Tracking.cpp:139-155): the first frame keeps T_last.
__device__ inline void record_first_at(const double* T_last, double* pose_out, int* tracked) {
  for (int i = 0; i < 16; ++i) pose_out[i] = T_last[i];
  *tracked = 1;
}

// ------------------------------------------------------------------ XCD mapping
// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8).  This bijection from launch order to tile index gives XCD j one
// contiguous run of tiles, so neighbouring image tiles share that XCD's L2.
__device__ inline int xcd_tile(int b, int T) {
  const int j = b & 7, k = b >> 3;
  const int q = T >> 3, r = T & 7;
  return j * q + (j < r ? j : r) + k;
}

// ------------------------------------------------------------------ wave64
__device__ inline int lane_id() { return __lane_id(); }
__device__ inline unsigned long long lane_mask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Wave-wide sum of one float (butterfly over the 64 lanes).
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Insert the keys held by the lanes with `has`, one CAS per distinct key of
// the wavefront; new blocks get consecutive pool indices from one atomic.
__device__ inline void wave_insert(const MapView& m, uint64_t key, bool has, int home) {
  bool pending = has;
  bool leader = false;
  for (;;) {
    unsigned long long mask = __ballot(pending);
    if (mask == 0ull) break;
    int l = __ffsll((unsigned long long)mask) - 1;
    unsigned long long lk = __shfl((unsigned long long)key, l);
    if (lane_id() == l) leader = true;
    if (pending && key == lk) pending = false;
  }
  bool won = false;
  int slot = home;
  if (leader) {
    const int lim = m.num_buckets < kMaxProbe ? m.num_buckets : kMaxProbe;
    int p = 0;
    for (; p < lim; ++p) {
      unsigned long long k = __hip_atomic_load(&m.slots[slot].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (k == key) break;
      if (k == kEmptyKey) {
        unsigned long long old = atomicCAS(&m.slots[slot].key, kEmptyKey, (unsigned long long)key);
        if (old == kEmptyKey) {
          won = true;
          break;
        }
        if (old == key) break;
      }
      slot = slot + 1 == m.num_buckets ? 0 : slot + 1;
    }
    if (p == lim) atomicOr(&m.ctr[kCtrFlags], kFlagProbeLimit);
  }
  unsigned long long wm = __ballot(won);
  if (wm) {
    int first = __ffsll(wm) - 1;
    int base = 0;
    if (lane_id() == first) base = atomicAdd(&m.ctr[kCtrAlloc], __popcll(wm));
    base = __shfl(base, first);
    if (won) {
      int idx = base + __popcll(wm & lane_mask_lt());
      if (idx < m.num_blocks) {
        m.block_keys[idx] = key;
        m.block_slot[idx] = slot;
        m.slots[slot].flags = kBlockUnobserved;
        m.slots[slot].val = idx;
      } else {
        m.slots[slot].val = -1;
        atomicOr(&m.ctr[kCtrFlags], kFlagPoolOverflow);
      }
    }
  }
}

}  // namespace yas

// One volume (MapStruct, VoxelMap.h:32-144): parameters, pose, device view.
// While hibernating, the hash table, block list and pool live in pinned host
// memory and the device view's big arrays are released.
struct yas_map {
  yas_ctx* ctx;
  yas_map_params P;
  yas::Intr K;
  double pose[16];
  std::mutex pose_mutex;
  yas::MapView v;
  yas::PosePair* pose_dev;  // pose of the current call
  float* depth_f32;         // scratch for uint16 depth (lazily sized)
  size_t depth_f32_bytes;
  int* host_ctr;            // pinned mirror of the counters
  bool hibernating;
  int hib_blocks;                 // allocated blocks at hibernation
  unsigned long long* hib_keys;   // pinned [hib_blocks]
  int* hib_flags;                 // pinned [hib_blocks]
  void* hib_voxels;               // pinned [hib_blocks * 512] voxels
};

#define YAS_REQUIRE_ACTIVE(map)                                                                    \
  do {                                                                                             \
    if ((map) && (map)->hibernating) {                                                             \
      yas_set_error("%s: the map is hibernating (call yas_map_reactivate)", __func__);             \
      return YAS_E_INVALID;                                                                        \
    }                                                                                              \
  } while (0)
