// yas_rgbdfusion.hpp - C++ facade with the reference's dense-fusion class
// interface (RGBDFusion/include/VoxelMap.h, RayTraceEngine.h, MeshEngine.h,
// RGBDTracking.h) over the C ABI of yas.h.  Header-only; link libyas.so.
//
// The reference passes cv::cuda::GpuMat images and Sophus::SE3d poses; OpenCV
// and Sophus are not part of this build, so the facade takes the plain
// equivalents below.  A GpuMat maps onto DeviceImage{mat.data, mat.step,
// mat.cols, mat.rows}; an SE3d onto Pose::from(T.matrix().data()) after a
// transpose to row-major (Eigen is column-major).  Errors throw yas::Error.
#pragma once

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "yas.h"

namespace yas {

struct Error : std::runtime_error {
  int code;
  Error(const char* what, int c) : std::runtime_error(std::string(what) + ": " + yas_last_error()), code(c) {}
};

inline void check(int rc, const char* what) {
  if (rc != YAS_OK) throw Error(what, rc);
}

// A device image: pointer, row pitch in bytes, size.
struct DeviceImage {
  void* data = nullptr;
  size_t pitch = 0;
  int cols = 0, rows = 0;
};

// A host image (the reference's cv::Mat): pointer, row step in bytes, size.
// Gray images are 8-bit, depth images float metres.
struct HostImage {
  const void* data = nullptr;
  size_t step = 0;
  int cols = 0, rows = 0;
};

// A rigid transform as a row-major 4x4 double matrix, with the two
// Sophus::SE3d operations the reference's callers use (Tracking.cpp:164-186).
struct Pose {
  double m[16];
  Pose() {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0 : 0.0;
  }
  static Pose from(const double* row_major) {
    Pose p;
    std::memcpy(p.m, row_major, sizeof(p.m));
    return p;
  }
  // [R t]^-1 = [R^T  -R^T t], each sum left to right.
  Pose inverse() const {
    Pose o;
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) o.m[4 * i + j] = m[4 * j + i];
      o.m[4 * i + 3] = -(m[i] * m[3] + m[4 + i] * m[7] + m[8 + i] * m[11]);
    }
    return o;
  }
  Pose operator*(const Pose& b) const {
    Pose o;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        o.m[4 * i + j] = m[4 * i] * b.m[j] + m[4 * i + 1] * b.m[4 + j] + m[4 * i + 2] * b.m[8 + j] +
                         m[4 * i + 3] * b.m[12 + j];
    return o;
  }
  double translation_norm() const { return std::sqrt(m[3] * m[3] + m[7] * m[7] + m[11] * m[11]); }
};

class Context {
 public:
  explicit Context(int device = 0, void* stream = nullptr) { check(yas_ctx_create(device, stream, &h_), "yas_ctx_create"); }
  ~Context() { yas_ctx_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  yas_ctx* get() const { return h_; }
  void Synchronize() { check(yas_ctx_synchronize(h_), "yas_ctx_synchronize"); }

 private:
  yas_ctx* h_ = nullptr;
};

class RayTraceEngine;
class MeshEngine;

// MapStruct (VoxelMap.h:32-144).
class MapStruct {
 public:
  MapStruct(Context& ctx, const float K[4]) : ctx_(&ctx) { std::memcpy(K_, K, sizeof(K_)); }
  ~MapStruct() { Release(); }
  MapStruct(const MapStruct&) = delete;
  MapStruct& operator=(const MapStruct&) = delete;

  // create(hashTableSize, bucketSize, voxelBlockSize, voxelSize, truncationDist)
  // (VoxelMap.cu:60-81): the open-addressed table uses hashTableSize slots.
  void create(int hashTableSize, int /*bucketSize*/, int voxelBlockSize, float voxelSize, float truncationDist,
              int voxelFormat = YAS_VOXEL_I16_U8, int maxWeight = 255) {
    Release();
    yas_map_params p{hashTableSize, voxelBlockSize, voxelSize, truncationDist, voxelFormat, maxWeight, 0.1f, 3.0f};
    check(yas_map_create(ctx_->get(), &p, K_, &h_), "yas_map_create");
  }
  void Reset() { check(yas_map_reset(h_), "yas_map_reset"); }
  void Release() {
    if (h_) yas_map_destroy(h_);
    h_ = nullptr;
  }
  bool Empty() const { return h_ == nullptr; }
  void SetPose(const Pose& T) { check(yas_map_set_pose(h_, T.m), "yas_map_set_pose"); }
  Pose GetPose() const {
    Pose T;
    check(yas_map_get_pose(h_, T.m), "yas_map_get_pose");
    return T;
  }
  // The public pose member the reference's callers read and assign directly
  // (Tracking.cpp:164 `Tmw = map->mTcw`, :716 `map->mTcw = pKF->GetPose()`):
  // both go through GetPose / SetPose, so the map keeps one pose.
  struct PoseField {
    MapStruct* map;
    operator Pose() const { return map->GetPose(); }
    PoseField& operator=(const Pose& T) {
      map->SetPose(T);
      return *this;
    }
    Pose inverse() const { return map->GetPose().inverse(); }
  };
  PoseField mTcw{this};
  // Fuse(GpuMat depth, SE3d Tcm) (VoxelMap.cu:825-884); depth in float metres.
  void Fuse(const DeviceImage& depth, const Pose& Tcm) {
    check(yas_map_fuse(h_, (const float*)depth.data, depth.pitch, depth.cols, depth.rows, Tcm.m), "yas_map_fuse");
  }
  // CheckNumVisibleBlocks(cols, rows, Tcm) (VoxelMap.cu:788-823).
  unsigned CheckNumVisibleBlocks(int cols, int rows, const Pose& Tcm) {
    int n = 0;
    check(yas_map_check_visible_blocks(h_, cols, rows, Tcm.m, &n), "yas_map_check_visible_blocks");
    return (unsigned)n;
  }
  unsigned GetNumVisibleBlocks() {
    yas_map_stats s;
    check(yas_map_get_stats(h_, &s), "yas_map_get_stats");
    return (unsigned)s.num_visible;
  }
  // RayTrace(Tcm) / GetRayTracingResult() (VoxelMap.cu:968-979).
  void SetRayTraceEngine(RayTraceEngine* e) { rt_ = e; }
  void SetMeshEngine(MeshEngine* e) { me_ = e; }
  inline void RayTrace(const Pose& Tcm);
  inline DeviceImage GetRayTracingResult();
  // GetRayTracingResultDepth() (VoxelMap.cu:981-987): the last ray cast's z
  // (0 where invalid), a device image overwritten by the next call; the
  // second form also copies it into `host` (W*H floats), as Tracking.cpp:622
  // does with cv::Mat(GpuMat).
  inline DeviceImage GetRayTracingResultDepth();
  inline void GetRayTracingResultDepth(std::vector<float>& host);
  // GenerateMesh() (VoxelMap.cu:143-150): host copies of vertices/normals.
  inline void GenerateMesh();
  void DeleteMesh() {
    point_.clear();
    normal_.clear();
    N = 0;
  }
  // Fuse(MapStruct*) (VoxelMap.cu:400-449): merge `other` at the relative
  // pose of the two maps, then release it; an empty map takes other's volume.
  void Fuse(MapStruct* other) {
    if (!other || other->Empty()) return;
    if (Empty()) {
      Swap(other);
      return;
    }
    DeleteMesh();
    check(yas_map_fuse_map(h_, other->h_), "yas_map_fuse_map");
    other->Release();
  }
  // Swap(MapStruct*) (VoxelMap.cu:83-102).
  void Swap(MapStruct* other) {
    std::swap(h_, other->h_);
    std::swap(K_, other->K_);
    point_.swap(other->point_);
    normal_.swap(other->normal_);
    std::swap(N, other->N);
  }
  // Reserve(hashTableSize, bucketSize, voxelBlockSize) (VoxelMap.cu:485-515).
  void Reserve(int hashTableSize, int /*bucketSize*/, int voxelBlockSize) {
    check(yas_map_reserve(h_, hashTableSize, voxelBlockSize), "yas_map_reserve");
  }
  // Hibernate() / ReActivate() (VoxelMap.cu:894-954).
  void Hibernate() {
    check(yas_map_hibernate(h_), "yas_map_hibernate");
    mbInHibernation = true;
  }
  void ReActivate() {
    check(yas_map_reactivate(h_), "yas_map_reactivate");
    mbInHibernation = false;
  }
  // SaveToFile / ReadFromFile (VoxelMap.cu:886-892).
  void SaveToFile(const std::string& path) { check(yas_map_save(h_, path.c_str()), "yas_map_save"); }
  void ReadFromFile(const std::string& path) {
    yas_map* m = nullptr;
    check(yas_map_load(ctx_->get(), path.c_str(), &m), "yas_map_load");
    Release();
    h_ = m;
    yas_map_params p;
    check(yas_map_get_params(h_, &p, K_), "yas_map_get_params");
  }
  void SetActiveFlag(bool flag) { mbActive = flag; }
  bool isActive() const { return mbActive; }
  bool mbActive = true;
  bool mbInHibernation = false;

  const float* mplPoint() const { return point_.data(); }
  const float* mplNormal() const { return normal_.data(); }
  int N = 0;  // floats in mplPoint (9 per triangle), as the reference's N

  yas_map* handle() const { return h_; }

 private:
  friend class MeshEngine;
  Context* ctx_;
  float K_[4];
  yas_map* h_ = nullptr;
  RayTraceEngine* rt_ = nullptr;
  MeshEngine* me_ = nullptr;
  std::vector<float> point_, normal_;
};

// RayTraceEngine (RayTraceEngine.h:11-48).
class RayTraceEngine {
 public:
  RayTraceEngine(Context& ctx, int w, int h, const float K[4]) : w_(w), h_(h) {
    check(yas_raycaster_create(ctx.get(), w, h, K, &rc_), "yas_raycaster_create");
  }
  ~RayTraceEngine() { yas_raycaster_destroy(rc_); }
  RayTraceEngine(const RayTraceEngine&) = delete;
  RayTraceEngine& operator=(const RayTraceEngine&) = delete;
  void RayTrace(MapStruct* map, const Pose& Tcm) { check(yas_raycast(rc_, map->handle(), Tcm.m), "yas_raycast"); }
  DeviceImage GetVMap() {
    DeviceImage im;
    float* p = nullptr;
    check(yas_raycaster_vmap(rc_, &p, &im.pitch), "yas_raycaster_vmap");
    im.data = p;
    im.cols = w_;
    im.rows = h_;
    return im;
  }
  unsigned GetNumVisibleBlock() {
    int n = 0;
    check(yas_raycaster_num_visible(rc_, &n), "yas_raycaster_num_visible");
    return (unsigned)n;
  }
  DeviceImage GetDepth(float* host = nullptr) {
    DeviceImage im;
    float* p = nullptr;
    check(yas_raycaster_depth(rc_, &p, &im.pitch, host, sizeof(float) * (size_t)w_), "yas_raycaster_depth");
    im.data = p;
    im.cols = w_;
    im.rows = h_;
    return im;
  }
  int width() const { return w_; }
  int height() const { return h_; }

 private:
  int w_, h_;
  yas_raycaster* rc_ = nullptr;
};

// MeshEngine (MeshEngine.h:8-24).
class MeshEngine {
 public:
  MeshEngine(Context& ctx, int maxNumTriangle) { check(yas_mesher_create(ctx.get(), maxNumTriangle, &me_), "yas_mesher_create"); }
  ~MeshEngine() { yas_mesher_destroy(me_); }
  MeshEngine(const MeshEngine&) = delete;
  MeshEngine& operator=(const MeshEngine&) = delete;
  void Meshify(MapStruct* map) {
    int n = 0;
    int rc = yas_mesh_extract(me_, map->handle(), &n);
    if (rc != YAS_OK && rc != YAS_E_CAPACITY) check(rc, "yas_mesh_extract");
    map->point_.resize((size_t)n * 9);
    map->normal_.resize((size_t)n * 9);
    check(yas_mesher_download(me_, map->point_.data(), map->normal_.data(), n), "yas_mesher_download");
    map->N = n * 9;
  }

 private:
  yas_mesher* me_ = nullptr;
};

inline void MapStruct::RayTrace(const Pose& Tcm) {
  if (rt_) rt_->RayTrace(this, Tcm);
}
inline DeviceImage MapStruct::GetRayTracingResult() { return rt_ ? rt_->GetVMap() : DeviceImage(); }
inline DeviceImage MapStruct::GetRayTracingResultDepth() { return rt_ ? rt_->GetDepth() : DeviceImage(); }
inline void MapStruct::GetRayTracingResultDepth(std::vector<float>& host) {
  if (!rt_) return;
  host.resize((size_t)rt_->width() * rt_->height());
  rt_->GetDepth(host.data());
}
inline void MapStruct::GenerateMesh() {
  if (me_) me_->Meshify(this);
}

// RGBDTracking (RGBDTracking.h:9-107).  Images may be device memory
// (DeviceImage) or host memory (HostImage, the reference's cv::Mat: 8-bit
// gray, float depth), which the setters upload as the reference does.
class RGBDTracking {
 public:
  RGBDTracking(Context& ctx, int w, int h, const double K[4], bool bRGB, bool bIcp, int levels = 5,
               const int* iters = nullptr) {
    yas_tracker_params p;
    p.levels = levels;
    const int def[5] = {10, 5, 3, 3, 3};
    for (int i = 0; i < 5; ++i) p.iters[i] = iters ? iters[i] : def[i];
    p.mode = (bRGB && bIcp) ? YAS_TRACK_RGBD : (bRGB ? YAS_TRACK_RGB : YAS_TRACK_DEPTH);
    p.w_rgb_h = 1.0f;
    p.w_rgb_b = 1.0f;
    p.w_icp_h = 100.0f;
    p.w_icp_b = 10.0f;
    check(yas_tracker_create(ctx.get(), w, h, K, &p, &t_), "yas_tracker_create");
    w_ = w;
    h_ = h;
  }
  ~RGBDTracking() { yas_tracker_destroy(t_); }
  RGBDTracking(const RGBDTracking&) = delete;
  RGBDTracking& operator=(const RGBDTracking&) = delete;
  void SetReferenceImage(const DeviceImage& g) {
    check(yas_tracker_set_reference_image(t_, (const float*)g.data, g.pitch), "SetReferenceImage");
  }
  void SetReferenceDepth(const DeviceImage& d) {
    check(yas_tracker_set_reference_depth(t_, (const float*)d.data, d.pitch), "SetReferenceDepth");
  }
  void SetTrackingImage(const DeviceImage& g) {
    check(yas_tracker_set_tracking_image(t_, (const float*)g.data, g.pitch), "SetTrackingImage");
  }
  void SetTrackingDepth(const DeviceImage& d) {
    check(yas_tracker_set_tracking_depth(t_, (const float*)d.data, d.pitch), "SetTrackingDepth");
  }
  void SetReferenceImage(const HostImage& g) {
    check(yas_tracker_set_reference_image_host(t_, (const uint8_t*)g.data, g.step), "SetReferenceImage");
  }
  void SetReferenceDepth(const HostImage& d) {
    check(yas_tracker_set_reference_depth_host(t_, (const float*)d.data, d.step), "SetReferenceDepth");
  }
  void SetTrackingImage(const HostImage& g) {
    check(yas_tracker_set_tracking_image_host(t_, (const uint8_t*)g.data, g.step), "SetTrackingImage");
  }
  void SetTrackingDepth(const HostImage& d) {
    check(yas_tracker_set_tracking_depth_host(t_, (const float*)d.data, d.step), "SetTrackingDepth");
  }
  // GetReferenceDepth(lvl) (RGBDTracking.cu:547-550): the raw depth most
  // recently uploaded from the host, whatever the level (as the reference).
  DeviceImage GetReferenceDepth(int /*lvl*/ = 0) const {
    DeviceImage im;
    float* p = nullptr;
    check(yas_tracker_raw_depth(t_, &p, &im.pitch), "GetReferenceDepth");
    im.data = p;
    im.cols = w_;
    im.rows = h_;
    return im;
  }
  // WriteDebugImages() (RGBDTracking.cu:557-568): curr_image, last_image, gx
  // and gy of level 0 as PGM files in `dir`.
  void WriteDebugImages(const std::string& dir = ".") {
    check(yas_tracker_write_debug_images(t_, dir.c_str()), "WriteDebugImages");
  }
  void SetReferenceModel(const DeviceImage& vmap) {
    check(yas_tracker_set_reference_model(t_, (const float*)vmap.data, vmap.pitch), "SetReferenceModel");
  }
  Pose GetTransform(const Pose& init, bool bSwapBuffer) {
    Pose out;
    yas_track_stats st;
    int rc = yas_tracker_get_transform(t_, init.m, bSwapBuffer ? 1 : 0, out.m, &st);
    if (rc != YAS_OK && rc != YAS_E_DIVERGED) check(rc, "GetTransform");
    good_ = st.good != 0;
    return out;
  }
  void SwapFrameBuffer() { check(yas_tracker_swap_frame_buffer(t_), "SwapFrameBuffer"); }
  bool IsTrackingGood() const { return good_; }

 private:
  yas_tracker* t_ = nullptr;
  bool good_ = true;
  int w_ = 0, h_ = 0;
};

}  // namespace yas
