// The reference-shaped C++ facade driven the way FullSystem's Tracking
// drives RGBDFusion (Tracking.cpp:139-197): StereoInitialization fuses the
// first frame; every later frame ray-casts the map at the last pose, tracks
// against the ray-cast model with host images (8-bit gray and float depth,
// the reference's cv::Mat), rejects |t| > 0.1 m (writing the debug images as
// the reference does) and otherwise composes T_w_cur = T_w_last * DT^-1 and
// fuses.  The map's pose is read and written through mTcw.
//
//   facade_replay FRAMES LEVELS OUTDIR
//     FRAMES: int32 W, H, n; double K[4]; then per frame W*H uint8 gray and
//             W*H float depth (metres)
//     prints "frame t tracked m0 .. m15" per frame (camera -> world),
//     then "depth_max_abs_diff", "raw_depth_ok", "triangles" lines.
//   facade_replay (no arguments)
//     two identical frames of a synthetic wavy surface; exits 0 if the
//     tracked motion is below 2 mm.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/yas_rgbdfusion.hpp"

static int replay_file(const char* path, int levels, const char* outdir) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return 10;
  int32_t hdr[3];
  double Kd[4];
  if (std::fread(hdr, sizeof(hdr), 1, f) != 1 || std::fread(Kd, sizeof(Kd), 1, f) != 1) return 11;
  const int W = hdr[0], H = hdr[1], n = hdr[2];
  const float K[4] = {(float)Kd[0], (float)Kd[1], (float)Kd[2], (float)Kd[3]};
  std::vector<uint8_t> gray((size_t)W * H);
  std::vector<float> depth((size_t)W * H);

  yas::Context ctx(0);
  yas::RayTraceEngine rt(ctx, W, H, K);
  yas::MeshEngine me(ctx, 4000000);
  yas::RGBDTracking trk(ctx, W, H, Kd, true, true, levels);  // 5 levels: {10,5,3,3,3}
  yas::MapStruct map(ctx, K);
  map.SetRayTraceEngine(&rt);
  map.SetMeshEngine(&me);
  map.create(1 << 18, 0, 1 << 15, 0.01f, 0.03f);  // reference storage: int16 sdf, uint8 weight
  float* dev_depth = nullptr;
  if (hipMalloc(&dev_depth, sizeof(float) * W * H) != hipSuccess) return 12;
  yas::DeviceImage ddepth{dev_depth, sizeof(float) * W, W, H};

  yas::Pose Tw_last;  // mLastFrame.mTcw
  double depth_diff = 0.0;
  bool raw_ok = true;
  for (int t = 0; t < n; ++t) {
    if (std::fread(gray.data(), 1, gray.size(), f) != gray.size()) return 13;
    if (std::fread(depth.data(), sizeof(float), depth.size(), f) != depth.size()) return 14;
    yas::HostImage hg{gray.data(), (size_t)W, W, H};
    yas::HostImage hd{depth.data(), sizeof(float) * W, W, H};
    if (hipMemcpy(dev_depth, depth.data(), sizeof(float) * W * H, hipMemcpyHostToDevice) != hipSuccess) return 15;
    int tracked = 1;
    if (t == 0) {
      // StereoInitialization (Tracking.cpp:139-155)
      map.mTcw = yas::Pose();
      trk.SetReferenceImage(hg);
      trk.SetReferenceDepth(hd);
      map.Fuse(ddepth, Tw_last);
    } else {
      // TrackRGBD (Tracking.cpp:162-197)
      yas::Pose Tmw = map.mTcw;
      yas::Pose Tcm = Tmw.inverse() * Tw_last;
      map.RayTrace(Tcm);
      trk.SetReferenceModel(map.GetRayTracingResult());
      trk.SetTrackingImage(hg);
      trk.SetTrackingDepth(hd);
      // the raw-depth buffer holds the depth just uploaded
      std::vector<float> back((size_t)W * H);
      yas::DeviceImage rd = trk.GetReferenceDepth();
      if (!rd.data || hipMemcpy(back.data(), rd.data, back.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
          std::memcmp(back.data(), depth.data(), back.size() * 4) != 0)
        raw_ok = false;
      yas::Pose DT = trk.GetTransform(yas::Pose(), true);
      if (DT.translation_norm() > 0.1) {
        trk.WriteDebugImages(outdir);
        tracked = 0;
      } else {
        Tw_last = Tw_last * DT.inverse();
        map.Fuse(ddepth, Tmw.inverse() * Tw_last);
      }
      // GetRayTracingResultDepth (Tracking.cpp:622): z of the ray cast
      std::vector<float> rdepth;
      map.GetRayTracingResultDepth(rdepth);
      std::vector<float> vm((size_t)W * H * 4);
      yas::DeviceImage v = map.GetRayTracingResult();
      if (hipMemcpy(vm.data(), v.data, vm.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return 16;
      for (size_t i = 0; i < rdepth.size(); ++i) {
        const float want = vm[4 * i + 3] > 0 ? vm[4 * i + 2] : 0.f;
        depth_diff = std::fmax(depth_diff, std::fabs((double)rdepth[i] - want));
      }
    }
    std::printf("frame %d %d", t, tracked);
    for (int i = 0; i < 16; ++i) std::printf(" %.17g", Tw_last.m[i]);
    std::printf("\n");
  }
  std::fclose(f);
  map.GenerateMesh();
  std::printf("depth_max_abs_diff %.9g\nraw_depth_ok %d\ntriangles %d\n", depth_diff, (int)raw_ok, map.N / 9);
  hipFree(dev_depth);
  return 0;
}

static int wavy_surface() {
  const int W = 160, H = 120;
  const float K[4] = {150.f, 150.f, 79.5f, 59.5f};
  const double Kd[4] = {150.0, 150.0, 79.5, 59.5};
  yas::Context ctx(0);
  yas::MapStruct map(ctx, K);
  yas::RayTraceEngine rt(ctx, W, H, K);
  yas::MeshEngine me(ctx, 200000);
  yas::RGBDTracking trk(ctx, W, H, Kd, false, true, 3);
  map.SetRayTraceEngine(&rt);
  map.SetMeshEngine(&me);
  map.create(1 << 14, 0, 1 << 12, 0.01f, 0.03f);
  std::vector<float> depth(W * H);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      depth[y * W + x] = 1.2f + 0.08f * std::sin(6.2831853f * 2.0f * x / W) * std::cos(6.2831853f * 1.5f * y / H);
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * W * H) != hipSuccess) return 1;
  if (hipMemcpy(d, depth.data(), sizeof(float) * W * H, hipMemcpyHostToDevice) != hipSuccess) return 1;
  yas::DeviceImage img{d, sizeof(float) * W, W, H};
  yas::Pose I;
  trk.SetReferenceDepth(img);
  map.Fuse(img, I);
  map.RayTrace(I);
  trk.SetReferenceModel(map.GetRayTracingResult());
  trk.SetTrackingDepth(img);
  yas::Pose DT = trk.GetTransform(yas::Pose(), true);
  map.Fuse(img, DT);
  // a second submap of the same view, merged in (System::FuseAllMapStruct)
  yas::MapStruct sub(ctx, K);
  sub.create(1 << 14, 0, 1 << 12, 0.01f, 0.03f);
  sub.Fuse(img, DT);
  map.Fuse(&sub);
  if (!sub.Empty()) return 3;
  map.GenerateMesh();
  std::printf("tracked=%d t=(%.2e %.2e %.2e) visible=%u triangles=%d\n", (int)trk.IsTrackingGood(), DT.m[3], DT.m[7],
              DT.m[11], map.CheckNumVisibleBlocks(W, H, I), map.N / 9);
  hipFree(d);
  // both frames are the same image: the model seen through the TSDF is
  // within a voxel of it, so the tracked motion is far below a centimetre
  return (trk.IsTrackingGood() && DT.translation_norm() < 2e-3 && map.N > 0) ? 0 : 2;
}

int main(int argc, char** argv) {
  try {
    if (argc >= 4) return replay_file(argv[1], std::atoi(argv[2]), argv[3]);
    return wavy_surface();
  } catch (const yas::Error& e) {
    std::printf("error: %s (code %d)\n", e.what(), e.code);
    return 20;
  }
}
