"""The benchmark's workloads, shared by bench.py and the tests that check
them, so that what is timed is what is tested.

* dense: one stream of the synthetic room at 640x480, voxel 5 mm,
  truncation 4 voxels, 2^22 buckets, a 2^19-block pool, ICP + photometric
  tracking (intensity 0.299R+0.587G+0.114B in float32, photometric weight
  0.1), three pyramid levels of 10/5/3 Gauss-Newton iterations.
* batching: 64 per-keyframe submaps of 5-60 frames (seeded uniform) at
  640x480 or 320x240 (seeded), 2^18 buckets each.
* serving: 32 streams, each a segment of the room sequence from its own
  seeded start frame, at 640x480 or 320x240 (seeded), 2^20 buckets and a
  2^16-block pool per stream.
* a volume of the size marching cubes is quoted on (200k-400k blocks at
  5 mm; the room holds about 100k): a wavy wall fused from a 7 x 6 grid of
  camera positions, about 267k blocks.
"""
import math

import numpy as np

W, H = 640, 480
VOXEL = 0.005
TRUNC = 4 * VOXEL
BUCKETS = 1 << 22
BLOCKS = 1 << 19
LEVELS, ITERS = 3, (10, 5, 3)
WEIGHTS = (0.1, 0.1, 1.0, 1.0)   # H = 0.1 H_rgb + H_icp, b = 0.1 b_rgb + b_icp
THETA0 = math.pi / 4

BATCH_MAPS, BATCH_SEED = 64, 1
BATCH_BUCKETS = 1 << 18
SERVE_STREAMS, SERVE_SEED = 32, 2
SERVE_BUCKETS, SERVE_BLOCKS = 1 << 20, 1 << 16


WALL_GRID, WALL_STEP = (7, 6), (2.4, 1.8)   # camera positions and their spacing (m)


def wavy_wall_depth(K, cx, cy, width=W, height=H, iters=12):
    """Depth (float32 metres) of the wall z = 2 + 0.15 sin(2 pi x / 0.7)
    cos(2 pi y / 0.9) m seen by a camera at (cx, cy, 0) looking along +z, by
    fixed-point iteration along each pixel's ray."""
    u, v = np.meshgrid(np.arange(width, dtype=np.float64), np.arange(height, dtype=np.float64))
    dx, dy = (u - K[2]) / K[0], (v - K[3]) / K[1]
    t = np.full((height, width), 2.0)
    for _ in range(iters):
        t = 2.0 + 0.15 * np.sin(2 * np.pi * (cx + t * dx) / 0.7) * np.cos(2 * np.pi * (cy + t * dy) / 0.9)
    return t.astype(np.float32)


def wall_frames(K):
    """(depth, camera-to-world pose) of every camera position of the wall."""
    out = []
    for i in range(WALL_GRID[0]):
        for j in range(WALL_GRID[1]):
            T = np.eye(4)
            T[0, 3], T[1, 3] = WALL_STEP[0] * i, WALL_STEP[1] * j
            out.append((wavy_wall_depth(K, T[0, 3], T[1, 3]), T))
    return out


def synth_params(width=W, height=H):
    from . import synth
    return synth.params(width=width, height=height, theta0=THETA0)


def dense_map_params(voxel_format):
    from . import rgbdfusion as rf
    return rf.map_params(BUCKETS, BLOCKS, VOXEL, TRUNC, voxel_format)


def tracker_params():
    from . import rgbdfusion as rf
    return rf.tracker_params(LEVELS, ITERS, rf.TRACK_RGBD, WEIGHTS)


class BatchSchedule:
    """Submap i fuses frames start[i] .. start[i] + nfr[i] - 1 of the room
    sequence at its resolution (hi[i]: 640x480, else 320x240), at the true
    poses relative to its first frame; step s holds each submap's s-th frame
    (None once its frames are used up)."""

    def __init__(self, nseq, n_maps=BATCH_MAPS, seed=BATCH_SEED):
        rng = np.random.default_rng(seed)
        self.n = n_maps
        self.nfr = rng.integers(5, 61, n_maps)
        self.hi = rng.integers(0, 2, n_maps).astype(bool)
        self.start = rng.integers(0, nseq - self.nfr)
        self.steps = int(self.nfr.max())

    def frame_index(self, s, i):
        return None if s >= self.nfr[i] else int(self.start[i] + s)

    def poses(self):
        """poses[s][i]: camera -> submap i's frame at step s."""
        from . import synth
        p = synth_params()
        out = []
        for s in range(self.steps):
            row = []
            for i in range(self.n):
                t = self.frame_index(s, i)
                row.append(np.eye(4) if t is None else
                           np.linalg.inv(synth.pose(p, int(self.start[i]))) @ synth.pose(p, t))
            out.append(row)
        return out

    def map_blocks(self, i):
        return 1 << 17 if self.hi[i] else 1 << 16

    def intrinsics(self, i):
        from . import synth
        return synth.intrinsics(synth_params(*((W, H) if self.hi[i] else (W // 2, H // 2))))


def serving_resolutions(n_streams=SERVE_STREAMS, seed=SERVE_SEED):
    """(rng, hi): hi[i] is True where stream i runs at 640x480 in the mixed
    configuration; the rng continues with the per-run draws."""
    rng = np.random.default_rng(seed)
    hi = rng.integers(0, 2, n_streams).astype(bool)
    return rng, hi


def serving_starts(rng, nseq, frames, n_streams=SERVE_STREAMS):
    """Each stream's first frame in the room sequence for a run of `frames`."""
    return rng.integers(0, nseq - frames, n_streams)
