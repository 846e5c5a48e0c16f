// Compiles the reference-shaped C++ facade the way FullSystem's Tracking
// would use it (Tracking.cpp:139-197); run on a GPU it replays two frames of
// TrackRGBD on a synthetic wavy surface (a plane would leave point-to-plane
// ICP without constraint along the plane) and prints the tracked pose.
#include <cmath>
#include <cstdio>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/yas_rgbdfusion.hpp"

int main() {
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
  map.create(1 << 14, 0, 1 << 12, 0.01f, 0.03f);  // create + Reset
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
  map.Fuse(img, I);                       // StereoInitialization
  map.RayTrace(I);                        // TrackRGBD
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
  const double tn = std::sqrt(DT.m[3] * DT.m[3] + DT.m[7] * DT.m[7] + DT.m[11] * DT.m[11]);
  return (trk.IsTrackingGood() && tn < 2e-3 && map.N > 0) ? 0 : 2;
}
