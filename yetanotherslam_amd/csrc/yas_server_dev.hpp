// Device-side description of one stream of the serving engine: every buffer
// and constant the dense per-frame step of one RGB-D stream touches (its
// volume, ray-cast buffers, tracker pyramids and state, pose chain), so a
// batched kernel can run any stream's step from its index alone.
#pragma once

#include "yas_internal.hpp"
#include "yas_tracker_dev.hpp"

namespace yas {

constexpr int kTrackMaxLevels = 5;

struct StreamDev {
  int W, H, cols, rows;
  float depth_scale_inv;  // metres per raw depth unit, (float)(1 / (double)scale)
  int gray_mode;
  Intr K;
  // volume
  MapView m;
  PosePair* pose;
  int vis_groups;  // workgroups of the visibility scan
  int int_groups;  // workgroups of the integration (4 wavefronts each)
  // ray cast
  unsigned* zmin;
  unsigned* zmax;
  float4* vmap;
  float vsinv, step_scale;
  int zr_groups;
  // frame inputs after conversion
  float* depth;
  float* gray;
  // tracker
  int levels;
  Level lv[kTrackMaxLevels];
  TrackBufs tb[kTrackMaxLevels];           // per-level buffers of the passes
  float* cur_inv[kTrackMaxLevels];         // inverse depth of the current frame
  int groups[kTrackMaxLevels];
  float* part_icp;
  float* part_res;
  float* part_rgb;
  TrackState* st;
  float angle_th, dist_th;
  // pose chain
  double* T_last;
  int* gate;
};

}  // namespace yas
