/*
 * yas.h - C ABI of the MI355X-native RGBDFusion dense path.
 *
 * Every entry point replaces one method of the reference's dense-fusion
 * classes (RGBDFusion/include/VoxelMap.h, RayTraceEngine.h, MeshEngine.h,
 * RGBDTracking.h, ImageProc.h); the citation above each declaration names it.
 * Signatures use plain pointers and sizes only: image arguments are device
 * pointers with a row pitch in bytes (the reference passes cv::cuda::GpuMat),
 * poses are 4x4 row-major double matrices (the reference passes Sophus::SE3d),
 * intrinsics are {fx, fy, cx, cy}.
 *
 * Conventions kept from the reference:
 *   - A pose "T_map_cam" maps camera coordinates into the volume's frame
 *     (the reference's mTcw / Tcm hold this transform despite their names).
 *   - Tracker results map the reference (last) frame into the current one;
 *     the caller forms T_cur = T_last * DT^-1 (FullSystem/src/Tracking.cpp:186).
 *   - Vertex maps are float4 per pixel, w = 1 valid / -1 invalid.
 *
 * Work is issued on the context's HIP stream.  Calls that return host values
 * (counts, poses, status) synchronise that stream; the others are
 * asynchronous.  Errors are returned as yas_status codes, never by exiting
 * (the reference prints and calls exit(0), RGBDFusion/utils/CudaUtils.h:15-25).
 * Handles are externally synchronised: one thread at a time per handle.
 */
#ifndef YAS_H
#define YAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  YAS_OK = 0,
  YAS_E_INVALID = 1,        /* bad argument or handle */
  YAS_E_HIP = 2,            /* HIP runtime error; see yas_last_error() */
  YAS_E_OUT_OF_BLOCKS = 3,  /* voxel-block pool exhausted; blocks that fit were allocated */
  YAS_E_HASH_FULL = 4,      /* probe limit reached or block key out of range */
  YAS_E_DIVERGED = 5,       /* tracker produced a NaN update (RGBDTracking.cu:187-191) */
  YAS_E_CAPACITY = 6,       /* caller buffer or engine capacity too small; count reported */
  YAS_E_REJECTED = 7        /* frame rejected by the tracking gate (Tracking.cpp:178-184) */
} yas_status;

/* Voxel storage variants of the one TSDF update rule (VoxelMap.cu:764-783). */
typedef enum {
  YAS_VOXEL_I16_U8 = 0,   /* reference: int16 sdf * 32767, uint8 weight (VoxelMap.h:26-30) */
  YAS_VOXEL_F32_F32 = 1,  /* float sdf, float weight */
  YAS_VOXEL_F16_U16 = 2   /* half sdf, uint16 weight */
} yas_voxel_format;

typedef enum { YAS_TRACK_RGB = 0, YAS_TRACK_DEPTH = 1, YAS_TRACK_RGBD = 2 } yas_track_mode;

typedef struct yas_ctx yas_ctx;
typedef struct yas_map yas_map;
typedef struct yas_raycaster yas_raycaster;
typedef struct yas_mesher yas_mesher;
typedef struct yas_tracker yas_tracker;
typedef struct yas_pipeline yas_pipeline;

typedef struct {
  int num_buckets;      /* hash slots (reference: bucketSize/hashTableSize) */
  int num_blocks;       /* voxel-block pool capacity (reference: voxelBlockSize) */
  float voxel_size;     /* metres */
  float trunc_dist;     /* metres (mu) */
  int voxel_format;     /* yas_voxel_format */
  int max_weight;       /* weight cap (reference: 255) */
  float depth_min;      /* metres (reference: 0.1) */
  float depth_max;      /* metres (reference: 3.0) */
} yas_map_params;

typedef struct {
  int num_allocated;    /* blocks in the pool */
  int num_visible;      /* visible blocks of the last fuse / visibility check */
  int flags;            /* bit0 pool overflow, bit1 probe limit, bit2 key out of range */
  int reserved;
} yas_map_stats;

typedef struct {
  int levels;           /* pyramid levels, 1..5 (reference NUM_PYR 5, RGBDTracking.h:7) */
  int iters[5];         /* Gauss-Newton iterations per level (reference {10,5,3,3,3}) */
  int mode;             /* yas_track_mode */
  float w_rgb_h, w_rgb_b, w_icp_h, w_icp_b; /* H = w_rgb_h H_rgb + w_icp_h H_icp, b likewise
                                               (reference RGBD: 1, 1, 100, 10; RGBDTracking.cu:521-545) */
} yas_tracker_params;

typedef struct {
  int good;             /* 0 after a NaN update (IsTrackingGood) */
  int icp_inliers;      /* last ICP match count at level 0 */
  int rgb_inliers;      /* last photometric residual count at level 0 */
  float last_error;
} yas_track_stats;

/* ------------------------------------------------------------------ core */
const char* yas_version(void);
/* Last error text of this thread (set whenever a call returns != YAS_OK). */
const char* yas_last_error(void);

/* Context: device + HIP stream.  stream == NULL creates a library-owned
 * non-blocking stream; otherwise the caller's hipStream_t is used (not
 * destroyed).  The reference launches every kernel on the legacy default
 * stream (no stream argument in RGBDFusion/src/*.cu), which
 * yas_ctx_create_null_stream binds to: work is then ordered with every other
 * user of the device's null stream (e.g. torch's default stream).
 * Objects built on a context hold a reference to it: yas_ctx_destroy drops
 * the creator's reference, and the context is freed after its last object,
 * so destroy order does not matter. */
int yas_ctx_create(int device, void* stream, yas_ctx** out);
int yas_ctx_create_null_stream(int device, yas_ctx** out);
int yas_ctx_destroy(yas_ctx* ctx);
int yas_ctx_synchronize(yas_ctx* ctx);
void* yas_ctx_stream(yas_ctx* ctx);

/* ------------------------------------------------------------------ map
 * MapStruct (RGBDFusion/include/VoxelMap.h:32-144). */

/* MapStruct(K) + create(...) + Reset() (VoxelMap.cu:46-81, VoxelMap.h:35). */
int yas_map_create(yas_ctx* ctx, const yas_map_params* params, const float K[4], yas_map** out);
/* Reset() (VoxelMap.cu:46-58): empty hash, pool and voxels. */
int yas_map_reset(yas_map* map);
/* Release() (VoxelMap.cu:103-136). */
int yas_map_destroy(yas_map* map);
/* SetPose / GetPose (VoxelMap.cu:12-22). */
int yas_map_set_pose(yas_map* map, const double T_world_map[16]);
int yas_map_get_pose(yas_map* map, double T_world_map[16]);
/* Fuse(GpuMat depth, SE3d Tcm) (VoxelMap.cu:825-884): block allocation along
 * each ray (K3), visibility (K4), TSDF integration (K5).  depth is float
 * metres, device memory, `pitch` bytes per row.  Asynchronous; capacity
 * problems are reported by yas_map_get_stats flags. */
int yas_map_fuse(yas_map* map, const float* depth, size_t pitch, int width, int height,
                 const double T_map_cam[16]);
/* Same, with uint16 depth in units of 1/depth_scale metres converted on the
 * device (System.cpp:52 depth.convertTo followed by Fuse). */
int yas_map_fuse_u16(yas_map* map, const uint16_t* depth, size_t pitch, int width, int height,
                     float depth_scale, const double T_map_cam[16]);
/* CheckNumVisibleBlocks(cols, rows, Tcm) (VoxelMap.cu:788-823). Synchronous. */
int yas_map_check_visible_blocks(yas_map* map, int width, int height, const double T_map_cam[16],
                                 int* count);
/* Keys (int32 triples) of the blocks the last visibility pass found visible
 * (fuse, yas_map_check_visible_blocks or a ray cast): the reference's
 * visibleTable (VoxelMap.h:125), which its engines read directly
 * (RayTraceEngine.cu:240-246).  At most `capacity`; *count receives the
 * number visible. Synchronous. */
int yas_map_visible_blocks(yas_map* map, int32_t* keys, int capacity, int* count);
/* Counters and capacity flags (GetNumVisibleBlocks, VoxelMap.cu:201-206). Synchronous. */
int yas_map_get_stats(yas_map* map, yas_map_stats* stats);
/* Copy every allocated block to host memory in pool order: keys as int32
 * triples (block coordinates), voxels in the map's storage format
 * (512 per block, index z*64 + y*8 + x).  *count receives the number of
 * blocks; at most `capacity` are copied.  Synchronous.  No reference
 * counterpart beyond Hibernate's device-to-host copy (VoxelMap.cu:894-923). */
int yas_map_export(yas_map* map, int32_t* keys, void* voxels, int capacity, int* count);
/* Parameters and intrinsics {fx, fy, cx, cy} the map was created with (K may
 * be NULL); the sizes include any growth by yas_map_reserve. */
int yas_map_get_params(yas_map* map, yas_map_params* params, float K[4]);
/* Bytes per voxel of the map's storage format. */
int yas_map_voxel_bytes(yas_map* map);

/* ------------------------------------------------------------------ map management
 * Submaps and paging around the per-frame path (VoxelMap.h:49-70). */
/* Fuse(MapStruct*) (VoxelMap.cu:400-449): re-sample src's TSDF into dst at
 * the relative pose of the two maps' poses (yas_map_set_pose).  dst first
 * grows by src's capacities (Reserve, VoxelMap.cu:412-417); then every
 * observed src voxel allocates its block in dst, and every dst voxel whose 8
 * trilinear src samples are observed takes the weighted mean with the first
 * sample's weight.  Same voxel size and storage required; src is not
 * modified (the reference releases it: the facade does).  Synchronous. */
int yas_map_fuse_map(yas_map* dst, yas_map* src);
/* Reserve(h, b, v) (VoxelMap.cu:485-515): grow the hash table to num_buckets
 * slots and the block pool to num_blocks blocks (no-op when both already
 * fit); allocated blocks keep their pool index and contents. Synchronous. */
int yas_map_reserve(yas_map* map, int num_buckets, int num_blocks);
/* Hibernate() / ReActivate() (VoxelMap.cu:894-954): move the allocated
 * blocks (keys, summary flags, voxels) to pinned host memory and release the
 * table, pool and lists on the device / bring them back and rebuild the
 * table.  While hibernating, calls that need the volume return
 * YAS_E_INVALID; pose and counters stay available. Synchronous. */
int yas_map_hibernate(yas_map* map);
int yas_map_reactivate(yas_map* map);
int yas_map_is_hibernating(yas_map* map, int* hibernating);
/* SaveToFile / ReadFromFile (VoxelMap.cu:886-892, empty in the reference):
 * a versioned binary file with the map's parameters, intrinsics, pose and
 * its allocated blocks (key, summary flags, voxels).  yas_map_load creates a
 * new map on ctx. Synchronous. */
int yas_map_save(yas_map* map, const char* path);
int yas_map_load(yas_ctx* ctx, const char* path, yas_map** out);

/* ------------------------------------------------------------------ batched fusion
 * Many maps (per-keyframe submaps, Tracking.cpp:703-716) fused together: the
 * per-frame fusion of MapStruct::Fuse (VoxelMap.cu:825-884) for every map of
 * the batch in one launch per stage, captured once in a hipGraph and replayed
 * step by step.  No counterpart class in the reference. */
typedef struct yas_batch yas_batch;
typedef struct {
  int alloc_tile_w;       /* allocation tile width: 16, 32 or 64 (x 256/w rows) */
  int visible_groups;     /* visibility workgroups (256 threads) per map */
  int integrate_threads;  /* integration workgroup size: 64, 128 or 256 */
  int integrate_groups;   /* integration workgroups per map */
  float step_ms;          /* autotuner: time of one step with this shape */
} yas_batch_config;
/* A batch over num_maps distinct maps of one voxel format and one context. */
int yas_batch_create(yas_ctx* ctx, int num_maps, yas_map* const* maps, yas_batch** out);
int yas_batch_destroy(yas_batch* b);
/* Schedule of `steps` steps.  Entry k = s * num_maps + i: map i fuses the
 * uint16 depth image depth[k] (device pointer, pitch[k] bytes, W[k] x H[k],
 * metres = raw / depth_scale) at pose T + 16 k (row-major camera -> map), or
 * idles at step s when depth[k] is NULL.  Depth images must stay valid while
 * the plan runs.  Captures the step graph (with a shape tuned earlier for the
 * same kind of work, if any).  Synchronous. */
int yas_batch_plan(yas_batch* b, int steps, const uint16_t* const* depth, const size_t* pitch, const int* W,
                   const int* H, float depth_scale, const double* T);
/* Enqueue steps [first_step, first_step + count) on the context's stream
 * (one graph launch per step).  A map reallocated after the plan (reserve,
 * merge, re-activation) makes this fail: plan again.  Asynchronous. */
int yas_batch_run(yas_batch* b, int first_step, int count);
/* Time the step graph on step 0 for candidate launch shapes (coordinate
 * descent: allocation tile, then visibility, then integration) and keep the
 * fastest, also for later plans of the same kind of work in this process.
 * The trial runs fuse step 0 several times: with reset_maps != 0 the maps
 * are reset afterwards.  Synchronous. */
int yas_batch_autotune(yas_batch* b, int reset_maps, yas_batch_config* best);
int yas_batch_set_config(yas_batch* b, const yas_batch_config* config);
int yas_batch_get_config(yas_batch* b, yas_batch_config* config);

/* ------------------------------------------------------------------ serving
 * Many RGB-D streams tracked and fused concurrently (no counterpart in the
 * reference, which runs one Tracking thread per System).  Every frame is the
 * dense step of yas_pipeline_step for its stream.  Continuous batching: the
 * oldest waiting frame of each stream joins the next batch; a batch is one
 * replay of a captured graph over all its streams, and batches queue back to
 * back on the server's own HIP stream.  A stream's results are identical
 * alone and in any mix.  Streams are admitted while their device buffers fit
 * the HBM budget; others wait and are admitted as memory frees up.  A server
 * is driven from one host thread. */
typedef struct yas_server yas_server;
typedef struct {
  int stream;          /* stream id */
  int tracked;         /* 1 if the frame was tracked (or was the stream's first) */
  long long request;   /* id returned by yas_server_submit */
  double pose[16];     /* row-major camera -> map pose of the frame */
} yas_result;
typedef struct {
  int streams;              /* admitted streams (active or hibernated) */
  long long frames_done;
  long long batches_done;
  size_t bytes_in_use;
  size_t bytes_budget;
  int streams_pending;      /* opened, waiting for HBM to be admitted */
  int streams_hibernated;   /* volume paged out to host memory */
  long long hibernations;   /* volumes paged out so far */
  long long reactivations;  /* volumes paged back in so far */
} yas_server_stats;
/* All streams share the tracker parameters and voxel format; images are at
 * most max_width x max_height. */
int yas_server_create(yas_ctx* ctx, int max_width, int max_height, int max_streams, const yas_tracker_params* tp,
                      int voxel_format, size_t hbm_budget_bytes, yas_server** out);
int yas_server_destroy(yas_server* sv);
/* Frames per batch, 1..64 (default 64).  Smaller batches keep the
 * tracker's per-level images of the batch resident in the last-level cache
 * across Gauss-Newton iterations. */
int yas_server_set_max_batch(yas_server* sv, int max_batch);
/* Batches queued on the device at once, 1..3 (default 2), and the frames
 * that must be waiting before a batch is queued behind one that is already
 * on the device (default 16, at most the batch cap).  Frames that arrive
 * while a batch runs gather into the next one, so holding batches back makes
 * them larger, paying their fixed launch cost over more frames; queueing
 * ahead keeps the device busy between batches.  No reference counterpart:
 * serving is this library's own form. */
int yas_server_set_max_in_flight(yas_server* sv, int batches, int min_ahead);
/* Run the Gauss-Newton iterations on sub-ranges of `slots` batch slots in
 * turn (0, the default: the whole batch at once).  Results are unchanged. */
int yas_server_set_gn_group(yas_server* sv, int slots);
/* Capture now the batch graphs of every batch size up to the batch cap
 * (one per power of two), which are otherwise captured when a batch of that
 * size is first launched, so that no batch waits for a capture.  Call it
 * after the last yas_server_set_gn_group; it returns once they are built. */
int yas_server_prepare(yas_server* sv);
/* Admit a stream now (its volume, ray caster, tracker and pose chain are
 * allocated); YAS_E_CAPACITY when its buffers would exceed the budget (after
 * any eviction, below) or all stream slots are taken. Synchronous. */
int yas_server_add_stream(yas_server* sv, int W, int H, const float K[4], const yas_map_params* mp, float depth_scale,
                          int gray_mode, int* stream_id);
/* Open a stream that is admitted as soon as its buffers fit the budget:
 * now if they do, otherwise it waits (pending) and is admitted, first come
 * first served, when streams close or are removed, or when idle streams are
 * hibernated (yas_server_set_eviction).  Frames may be submitted to a
 * pending stream; they wait for its admission.  YAS_E_CAPACITY only when all
 * stream slots are taken. */
int yas_server_open_stream(yas_server* sv, int W, int H, const float K[4], const yas_map_params* mp,
                           float depth_scale, int gray_mode, int* stream_id);
/* Close a stream: it takes no more frames, finishes those submitted (a
 * pending stream's are dropped) and is then released, freeing its memory
 * for pending streams.  Never blocks. */
int yas_server_close_stream(yas_server* sv, int stream_id);
/* Release a stream once its frames in flight are done (none may be waiting). */
int yas_server_remove_stream(yas_server* sv, int stream_id);
/* Eviction: with idle_ms >= 0, a stream that has been idle (no frame waiting
 * or in flight) for at least idle_ms may be hibernated, its volume paged out
 * to pinned host memory as MapStruct::Hibernate does (VoxelMap.cu:894-923),
 * least recently used first, when a pending or re-activating stream needs
 * the memory; its next frame pages it back in (ReActivate,
 * VoxelMap.cu:925-954) once it fits.  Results are unchanged by paging.
 * idle_ms < 0 (the default) never hibernates. */
int yas_server_set_eviction(yas_server* sv, double idle_ms);
/* 0 free, 1 pending, 2 active, 3 hibernated. */
int yas_server_stream_state(yas_server* sv, int stream_id, int* state);
/* The stream's volume (for meshing or export between batches). */
int yas_server_stream_map(yas_server* sv, int stream_id, yas_map** map);
/* The stream's pipeline (its map, ray caster and tracker) for inspection
 * between batches; its step and timing calls must not be used. */
int yas_server_stream_pipeline(yas_server* sv, int stream_id, yas_pipeline** pipe);
/* Queue a frame: uint16 depth and RGB8 colour in device memory, ready when
 * submitted and valid until its result is polled. */
int yas_server_submit(yas_server* sv, int stream_id, const uint16_t* depth, size_t dpitch, const uint8_t* rgb,
                      size_t rpitch, long long* request_id);
/* Launch batches from the waiting frames while fewer than three batches are
 * queued; collect finished ones.  Never blocks on the device. */
int yas_server_pump(yas_server* sv, int* batches_launched);
/* Results of finished frames, oldest first. Never blocks. */
int yas_server_poll(yas_server* sv, yas_result* out, int capacity, int* count);
/* Pump until every submitted frame is finished; YAS_E_CAPACITY if frames
 * wait for streams that cannot be admitted. */
int yas_server_drain(yas_server* sv);
int yas_server_get_stats(yas_server* sv, yas_server_stats* stats);

/* ------------------------------------------------------------------ raycast
 * RayTraceEngine (RGBDFusion/include/RayTraceEngine.h:11-48). */
int yas_raycaster_create(yas_ctx* ctx, int width, int height, const float K[4], yas_raycaster** out);
int yas_raycaster_destroy(yas_raycaster* rc);
/* RayTrace(MapStruct*, SE3d) (RayTraceEngine.cu:412-450): the vertex map of
 * the volume seen from T_map_cam, in the camera frame. Asynchronous. */
int yas_raycast(yas_raycaster* rc, yas_map* map, const double T_map_cam[16]);
/* The ray cast of yas_raycast into caller memory, allocating nothing: no
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops and
 * harnesses that forbid hipMalloc.  The map (its visible-block list and
 * pose) and the dense W*H float4 vertex map `vmap` (pitch = 16*W) are the
 * caller's; intrinsics K = {fx, fy, cx, cy} as yas_raycaster_create.  The
 * only scratch, the two 1/8-resolution z-range maps, is a static device
 * array of 32768 cells per map (256 KiB per device, held until the process
 * ends): (W/8)*(H/8) > 32768 is rejected with YAS_E_INVALID.  Every call on a
 * device shares that array, so calls that overlap on one GPU (on different
 * streams) race.  Asynchronous on the map's stream; results are those of
 * yas_raycast with a ray caster of the same size and intrinsics. */
int yas_raycast_into(yas_map* map, const double T_map_cam[16], const float K[4], int width, int height,
                     float* vmap, size_t pitch);
/* GetVMap() (RayTraceEngine.cu:452-454): device alias of the last result,
 * float4 per pixel, overwritten by the next yas_raycast. */
int yas_raycaster_vmap(yas_raycaster* rc, float** vmap, size_t* pitch);
/* MapStruct::GetRayTracingResultDepth() (VoxelMap.cu:981-987, called at
 * Tracking.cpp:622): VMapToDepth (ImageProc.cu:262-285) of the last vertex
 * map, z where w > 0 and 0 elsewhere, into the ray caster's depth image
 * (device, float W*H, overwritten by the next call; the reference allocates a
 * new GpuMat).  depth/pitch may be NULL.  If host != NULL the image is also
 * copied there (row pitch host_pitch bytes) and the call waits for it. */
int yas_raycaster_depth(yas_raycaster* rc, float** depth, size_t* pitch, float* host, size_t host_pitch);
/* The 1/8-resolution [zmin, zmax] maps of the last ray cast (device, float). */
int yas_raycaster_zrange(yas_raycaster* rc, float** zmin, float** zmax, int* cols, int* rows);
/* GetNumVisibleBlock() (RayTraceEngine.cu:38-43). Synchronous. */
int yas_raycaster_num_visible(yas_raycaster* rc, int* count);

/* ------------------------------------------------------------------ mesh
 * MeshEngine (RGBDFusion/include/MeshEngine.h:8-24). */
int yas_mesher_create(yas_ctx* ctx, int max_triangles, yas_mesher** out);
int yas_mesher_destroy(yas_mesher* me);
/* Meshify(MapStruct*) (MeshEngine.cu:293-339): marching cubes over every
 * allocated block into a triangle soup on the device.  Triangles of a block
 * are contiguous, in voxel-index then table order.  *num_triangles gets the
 * full count even when it exceeds the capacity (then YAS_E_CAPACITY).
 * Synchronous. */
int yas_mesh_extract(yas_mesher* me, yas_map* map, int* num_triangles);
/* Download the last extraction: 9 floats (3 vertices) per triangle for
 * vertices and for face normals (either may be NULL). Synchronous. */
int yas_mesher_download(yas_mesher* me, float* vertices, float* normals, int capacity);
/* Marching cubes as an indexed mesh (no counterpart in the reference, whose
 * Meshify emits a soup): float32 xyz vertices and int32 triangle corners.
 * Each vertex sits on one cube edge and is shared by every triangle that
 * uses the edge; a vertex on an edge whose neighbouring cubes are all
 * incomplete is emitted but unreferenced.  Triangles come in the same order
 * as yas_mesh_extract's; their corners match its positions to a few ulps of
 * the coordinate (the interpolation runs along +axis, not along the cube's
 * edge direction).  Capacity: max_triangles
 * triangles and 3 * max_triangles vertices; the counts are returned even when
 * they exceed it (then YAS_E_CAPACITY).  The mesher holds one extraction at
 * a time.  Synchronous. */
int yas_mesh_extract_indexed(yas_mesher* me, yas_map* map, int* num_vertices, int* num_triangles);
/* Download the last indexed extraction (3 floats per vertex, 3 int32 per
 * triangle); either pointer may be NULL. Synchronous. */
int yas_mesher_download_indexed(yas_mesher* me, float* vertices, int32_t* triangles, int vertex_capacity,
                                int triangle_capacity);
/* Write the last indexed extraction as a binary little-endian PLY (float
 * x, y, z vertices; uchar-counted int32 triangle lists). Synchronous. */
int yas_mesher_write_ply(yas_mesher* me, const char* path);
/* Device pointers of the last indexed extraction (valid until the next
 * extraction or yas_mesher_destroy). */
int yas_mesher_indexed_device(yas_mesher* me, const float** vertices, const int32_t** triangles);
/* Per-block detail of the last extraction, in pool order: block keys
 * (int32 triples), first triangle index, triangle count and, when `cases`
 * is not NULL, the cube index of every voxel (-1 where a corner is
 * unobserved; 512 per block).  Synchronous. */
int yas_mesher_blocks(yas_mesher* me, int32_t* keys, int32_t* first, int32_t* count, int16_t* cases,
                      int capacity, int* num_blocks);

/* ------------------------------------------------------------------ tracker
 * RGBDTracking (RGBDFusion/include/RGBDTracking.h:9-107). */
/* RGBDTracking(w, h, K, bRGB, bIcp) (RGBDTracking.cu:5-64). */
int yas_tracker_create(yas_ctx* ctx, int width, int height, const double K[4],
                       const yas_tracker_params* params, yas_tracker** out);
int yas_tracker_destroy(yas_tracker* trk);
/* SetReferenceImage / SetTrackingImage (RGBDTracking.cu:66-78,102-116):
 * float intensity, device memory. */
int yas_tracker_set_reference_image(yas_tracker* trk, const float* gray, size_t pitch);
int yas_tracker_set_tracking_image(yas_tracker* trk, const float* gray, size_t pitch);
/* SetReferenceDepth / SetTrackingDepth (RGBDTracking.cu:80-100,118-138):
 * float depth in metres, device memory. */
int yas_tracker_set_reference_depth(yas_tracker* trk, const float* depth, size_t pitch);
int yas_tracker_set_tracking_depth(yas_tracker* trk, const float* depth, size_t pitch);
/* The same setters from host images, as the reference's cv::Mat overloads
 * (RGBDTracking.h:17-21): an 8-bit gray image is converted to float on the
 * host and uploaded (RGBDTracking.cu:68-71,104-107); float depth in metres is
 * uploaded into the tracker's raw-depth buffer (mGpuBufferRawDepth,
 * RGBDTracking.cu:86-87,122-123).  Each call waits for its upload. */
int yas_tracker_set_reference_image_host(yas_tracker* trk, const uint8_t* gray, size_t pitch);
int yas_tracker_set_tracking_image_host(yas_tracker* trk, const uint8_t* gray, size_t pitch);
int yas_tracker_set_reference_depth_host(yas_tracker* trk, const float* depth, size_t pitch);
int yas_tracker_set_tracking_depth_host(yas_tracker* trk, const float* depth, size_t pitch);
/* GetReferenceDepth() (RGBDTracking.cu:547-550): the raw-depth buffer, i.e.
 * the depth most recently uploaded by a host-image setter (device, float;
 * NULL before the first). */
int yas_tracker_raw_depth(yas_tracker* trk, float** depth, size_t* pitch);
/* WriteDebugImages() (RGBDTracking.cu:557-568, called at Tracking.cpp:182):
 * level-0 current and reference intensity and the x/y gradients as 8-bit
 * binary PGM files curr_image.pgm, last_image.pgm, gx.pgm, gy.pgm in `dir`
 * (rounded to nearest, clamped to [0, 255]). Synchronous. */
int yas_tracker_write_debug_images(yas_tracker* trk, const char* dir);
/* SetReferenceModel(GpuMat vmap) (RGBDTracking.cu:140-150). */
int yas_tracker_set_reference_model(yas_tracker* trk, const float* vmap, size_t pitch);
/* GetTransform(init, bSwapBuffer) (RGBDTracking.cu:152-214). Synchronous. */
int yas_tracker_get_transform(yas_tracker* trk, const double init[16], int swap_buffers,
                              double out[16], yas_track_stats* stats);
/* SwapFrameBuffer() (RGBDTracking.cu:294-303). */
int yas_tracker_swap_frame_buffer(yas_tracker* trk);
/* IsTrackingGood() (RGBDTracking.cu:553-556). */
int yas_tracker_is_tracking_good(yas_tracker* trk, int* good);
/* One linearisation (ComputeSingleStepDepth / ComputeSingleStepRGB,
 * RGBDTracking.cu:408-441 / 225-292) at `level` and pose T: the 28
 * upper-triangle entries of [J|r]^T [J|r] and the residual count.
 * kind 0 = ICP, 1 = photometric.  stats (may be NULL) receives the
 * photometric {sum r, count, mean, Huber threshold}. Synchronous. */
int yas_tracker_step_sums(yas_tracker* trk, int level, const double T[16], int kind, float out[29],
                          float stats[4]);
/* The Gauss-Newton step's solve of H x = b (the LDLT of RGBDTracking.cu:185,
 * Eigen's diagonal pivoting) for n systems (1 <= n <= 2^20), run
 * by the same device code as the tracker's steps (test access): H is n
 * row-major 6x6 float matrices, b n float 6-vectors, x receives n double
 * 6-vectors (host arrays). Synchronous. */
int yas_tracker_solve6(yas_ctx* ctx, const float* H, const float* b, double* x, int n);
/* Device pointer and pitch of one pyramid image (test access):
 * which = 0 current intensity, 1 reference intensity, 2 gx, 3 gy,
 * 4 current inverse depth, 5 reference inverse depth (float);
 * 6 current vmap, 7 current nmap, 8 reference vmap, 9 reference nmap,
 * 10 photometric residual rows (r, gx, gy) (float3; r is NaN where the
 * pixel has no residual).  The normal maps are formed from the vertex maps
 * by this call, which then waits for the context's stream. */
int yas_tracker_level(yas_tracker* trk, int which, int level, void** ptr, size_t* pitch, int* width,
                      int* height);

/* ------------------------------------------------------------------ image ops
 * Free functions of RGBDFusion/include/ImageProc.h:8-24 plus the host-side
 * conversions of FullSystem/src/System.cpp:43-52, on the context's stream. */
int yas_depth_u16_to_f32(yas_ctx* ctx, const uint16_t* src, size_t src_pitch, float* dst, size_t dst_pitch,
                         int width, int height, float depth_scale);
/* mode 0: 0.299R + 0.587G + 0.114B in float; mode 1: 8-bit fixed-point luma (OpenCV RGB2GRAY). */
int yas_rgb_to_intensity(yas_ctx* ctx, const uint8_t* rgb, size_t src_pitch, float* dst, size_t dst_pitch,
                         int width, int height, int mode);
int yas_depth_to_invdepth(yas_ctx* ctx, const float* depth, size_t src_pitch, float* inv, size_t dst_pitch,
                          int width, int height);                                  /* ImageProc.cu:141-164 */
int yas_pyrdown_depth(yas_ctx* ctx, const float* src, size_t src_pitch, float* dst, size_t dst_pitch,
                      int dst_width, int dst_height);                             /* ImageProc.cu:166-187 */
int yas_pyrdown_gauss(yas_ctx* ctx, const float* src, size_t src_pitch, int src_width, int src_height,
                      float* dst, size_t dst_pitch);                              /* cv::cuda::pyrDown */
int yas_compute_vertex_map(yas_ctx* ctx, const float* invdepth, size_t src_pitch, float* vmap,
                           size_t dst_pitch, int width, int height, const float K[4],
                           float cutoff);                                         /* ImageProc.cu:189-220 */
int yas_compute_normal_map(yas_ctx* ctx, const float* vmap, size_t src_pitch, float* nmap, size_t dst_pitch,
                           int width, int height);                                /* ImageProc.cu:222-259 */
int yas_image_gradient(yas_ctx* ctx, const float* src, size_t src_pitch, float* gx, float* gy,
                       size_t dst_pitch, int width, int height);                  /* ImageProc.cu:7-40 */
int yas_vmap_to_depth(yas_ctx* ctx, const float* vmap, size_t src_pitch, float* depth, size_t dst_pitch,
                      int width, int height);                                     /* ImageProc.cu:262-285 */

/* ------------------------------------------------------------------ pipeline
 * The dense half of Tracking::TrackRGBD / StereoInitialization
 * (FullSystem/src/Tracking.cpp:139-197) on one volume, with every pose kept
 * on the device so that a frame issues no host synchronisation. */
int yas_pipeline_create(yas_ctx* ctx, int width, int height, const float K[4], const yas_map_params* map,
                        const yas_tracker_params* tracker, float depth_scale, int gray_mode,
                        yas_pipeline** out);
int yas_pipeline_destroy(yas_pipeline* pl);
/* One frame from device-resident uint16 depth and packed RGB8: frame 0 is
 * fused at the identity; later frames ray-cast at the last pose, track, gate
 * |t| > 0.1 m, and fuse at T_last * DT^-1.  Asynchronous. */
int yas_pipeline_step(yas_pipeline* pl, const uint16_t* depth, size_t depth_pitch, const uint8_t* rgb,
                      size_t rgb_pitch);
/* Pose of frame `frame` (camera to map) and whether it was tracked
 * (1) or rejected (0).  Synchronous. */
int yas_pipeline_pose(yas_pipeline* pl, int frame, double T[16], int* tracked);
int yas_pipeline_num_frames(yas_pipeline* pl, int* n);
/* Device timing with HIP events on the context's stream (the reference only
 * brackets TrackRGBD with std::chrono, Examples/tum_rgbd.cpp:56-60).  With
 * detail_every = n > 0, every frame after the first records its start and
 * end, and every n-th of them (the first included) also its stage
 * boundaries and the kernel events below; 0 stops timing.
 * yas_pipeline_timing synchronises and accumulates them: stage_ms[8] = sums
 * over the sampled frames of {input conversion, visibility + z-range, ray
 * march, tracker inputs (model + frame pyramids), Gauss-Newton solve, block
 * allocation, block visibility, TSDF integration}, frame_ms = whole-frame
 * time of each timed frame (at most `capacity`), *num_frames = frames timed
 * so far. */
int yas_pipeline_enable_timing(yas_pipeline* pl, int detail_every);
int yas_pipeline_timing(yas_pipeline* pl, double stage_ms[8], double* frame_ms, int capacity, int* num_frames);
int yas_pipeline_reset_timing(yas_pipeline* pl);
/* Total duration, over the frames sampled for detail, of each of the four
 * single-kernel stages' kernels {ray march, block allocation, block
 * visibility, TSDF integration}, from events that each launch records at
 * the start and the end of the kernel's own execution (hipExtLaunchKernel),
 * and the number of sampled frames. Synchronous. */
int yas_pipeline_kernel_timing(yas_pipeline* pl, double kernel_ms[4], int* frames_sampled);
/* Counters of frames [first, first+count): 8 ints per frame {allocated
 * blocks, visible blocks at fusion, capacity flags, Bresenham walk steps
 * (cumulative), visible blocks at the ray cast, 0, 0, 0}. Synchronous. */
int yas_pipeline_frame_stats(yas_pipeline* pl, int first, int count, int* out);
int yas_pipeline_map(yas_pipeline* pl, yas_map** map);
int yas_pipeline_raycaster(yas_pipeline* pl, yas_raycaster** rc);
int yas_pipeline_tracker(yas_pipeline* pl, yas_tracker** trk);

#ifdef __cplusplus
}
#endif

#endif /* YAS_H */
