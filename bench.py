#!/usr/bin/env python3
"""Benchmark of the dense RGB-D path: frames/s of (ray cast + track + fuse)
at 640x480 with 5 mm voxels, p50/p99 frame latency, roofline of the
dominant kernel and the scalar oracle timed on the host as the CPU baseline.

Workload (one step = one frame of the dense half of Tracking::TrackRGBD):
the synthetic room sequence at 640x480, voxel 5 mm, truncation 4 voxels,
2^22 hash buckets, float32 sdf + float32 weight, ICP + photometric tracker
(intensity 0.299R+0.587G+0.114B in float32, photometric weight 0.1), three
pyramid levels with 10/5/3 Gauss-Newton iterations.  All frames are uploaded
to HBM before the timed region.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-mesh] [--no-batch] [--no-serving]
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

W, H = 640, 480
VOXEL = 0.005
TRUNC = 4 * VOXEL
BUCKETS = 1 << 22
BLOCKS = 1 << 19
LEVELS, ITERS = 3, (10, 5, 3)
WEIGHTS = (0.1, 0.1, 1.0, 1.0)   # H = 0.1 H_rgb + H_icp, b = 0.1 b_rgb + b_icp
THETA0 = math.pi / 4
HBM_PEAK_GBS = 8000.0
STAGES = ["convert", "visibility+zrange", "raymarch", "tracker_inputs", "gauss_newton", "alloc", "visible",
          "integrate"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen_frames(p, n):
    from yetanotherslam_amd import synth
    depth = np.zeros((n, H, W), np.uint16)
    rgb = np.zeros((n, H, W, 3), np.uint8)
    for t in range(n):
        d, c, _ = synth.frame(p, t)
        depth[t], rgb[t] = d, c
    return depth, rgb


def cpu_baseline(p, K, depth, rgb, budget_s=20.0, max_frames=40):
    """The scalar oracle (single thread) on the first frames of the same
    sequence: frame 0 initialises the volume and is not timed."""
    from oracle import pyoracle as po
    op = po.OraclePipeline(W, H, K, VOXEL, TRUNC, LEVELS, ITERS, po.MODE_RGBD, WEIGHTS)
    op.step(depth[0], rgb[0])
    t0 = time.perf_counter()
    n = 0
    for t in range(1, min(max_frames, len(depth))):
        op.step(depth[t], rgb[t])
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return n / dt, n, dt


# HBM bytes per launch of the main loop's kernels on this workload, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a plain copy of known
# size (profiles/dense_640x480_5mm_hbm_traffic.md).  The profiler cannot run
# inside this script, so the measured figures are carried here.
TRAFFIC_BYTES = {"k_march": 55.6e6, "k_integrate": 111.7e6, "k_alloc": 4.5e6}


def walk_steps(stats):
    """Bresenham block visits over the timed frames, from the cumulative
    uint32 counter (frame 0's visits are estimated by the mean of the rest)."""
    c = stats[:, 3].astype(np.int64) & 0xFFFFFFFF
    d = np.diff(c) % (1 << 32)
    return float(d.sum() * len(c) / max(len(d), 1))


def rooflines(stats, stage_ms, steps):
    """Achieved algorithmic HBM bandwidth of the single-kernel stages (bytes
    per launch as defined for each kernel below, divided by the kernel's
    average duration measured with HIP events around its launches)."""
    sv, se, sd = 8, 16, 4            # voxel bytes (f32 sdf + f32 weight), hash slot bytes, depth bytes
    nvis_fuse = stats[:, 1].astype(np.float64)
    nvis_ray = stats[:, 4].astype(np.float64)
    cells = (W // 8) * (H // 8)
    defs = {
        # every voxel of every visible block read and written, the depth
        # image, and the block's list entry + key + slot index
        "k_integrate": (7, float(np.sum(nvis_fuse * (512 * 2 * sv + 4 + 8 + 4) + W * H * sd))),
        # compulsory ray-cast traffic: each visible block's voxels and hash
        # slot once, the vertex map written, the two z-range maps read
        "k_march": (2, float(np.sum(nvis_ray * (512 * sv + se) + W * H * 16 + 2 * cells * 4))),
        # the depth image read and one hash-slot probe per Bresenham visit
        # (the walk counter is cumulative and wraps at 2^32)
        "k_alloc": (5, float(steps * W * H * sd + walk_steps(stats) * se)),
    }
    out = {}
    for name, (i, nbytes) in defs.items():
        t = stage_ms[i] / 1e3
        ach = nbytes / t / 1e9 if t > 0 else 0.0
        out[name] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_BYTES.get(name),
                     "algorithmic_bytes_per_launch": round(nbytes / steps),
                     "avg_launch_ms": round(stage_ms[i] / steps, 5)}
    return out


def mesh_bench(pl, dev_depth, K, nframes, reps=3):
    """Marching cubes over the fused 5 mm volume (float32 sdf, the timed
    run's own map) and over the same frames fused at the tracked poses into a
    float16 sdf + uint16 weight map: indexed output (float32 vertices, int32
    triangles), and the reference's triangle soup beside it.  Wall time of
    the synchronous extraction call (kernels + one 8-byte readback), median
    of `reps` after one untimed call."""
    import ctypes
    import torch
    from yetanotherslam_amd import rgbdfusion as rf

    f16 = rf.MapStruct(np.asarray(K, np.float32), BUCKETS, BLOCKS, VOXEL, TRUNC, voxel_format=rf.VOXEL_F16_U16)
    for t in range(nframes):
        T, ok = pl.pose(t)
        if ok:
            f16.FuseU16(dev_depth[t], T)
    torch.cuda.synchronize()
    me = rf.MeshEngine(8_000_000)
    nv, nt, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    out = {}
    for name, h, sv in (("f32", pl.map_handle(), 8), ("f16u16", f16.h, 4)):
        ms = {}
        for kind in ("indexed", "soup"):
            ts = []
            for _ in range(reps + 1):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if kind == "indexed":
                    rf.native.call("yas_mesh_extract_indexed", me.h, h, ctypes.byref(nv), ctypes.byref(nt))
                else:
                    rf.native.call("yas_mesh_extract", me.h, h, ctypes.byref(nt))
                ts.append(time.perf_counter() - t0)
            ms[kind] = float(np.median(ts[1:])) * 1e3
        rf.native.call("yas_mesher_blocks", me.h, None, None, None, None, 0, ctypes.byref(nb))
        out[name] = {"blocks": nb.value, "triangles": nt.value, "vertices": nv.value,
                     "ms_indexed": round(ms["indexed"], 3), "ms_soup": round(ms["soup"], 3),
                     "voxel_read_GBps_indexed": round(2 * nb.value * 512 * sv / (ms["indexed"] * 1e-3) / 1e9, 1)}
    return out


def batch_bench(dev_depth, nseq, n_maps=64, seed=1):
    """Batching form: 64 per-keyframe submaps of 5-60 frames each (seeded
    uniform), resolution 320x240 or 640x480, voxel 5 mm, 2^18 buckets and a
    2^17- (640x480) or 2^16-block (320x240) pool each, f32 storage.  Submap i fuses frames
    start_i .. start_i + n_i - 1 of the room sequence (at its resolution) at
    the true poses relative to its first frame.  Timed twice over the same
    schedule: one yas_map_fuse_u16 call per (submap, frame), and the batched
    graph (one launch per stage per step, shapes from the autotuner)."""
    import torch
    from yetanotherslam_amd import rgbdfusion as rf
    from yetanotherslam_amd import synth

    rng = np.random.default_rng(seed)
    nfr = rng.integers(5, 61, n_maps)
    hi = rng.integers(0, 2, n_maps).astype(bool)
    start = rng.integers(0, nseq - nfr)
    p640 = synth.params(width=W, height=H, theta0=THETA0)
    p320 = synth.params(width=W // 2, height=H // 2, theta0=THETA0)
    d320 = torch.from_numpy(np.stack([synth.frame(p320, t)[0] for t in range(nseq)]).view(np.int16)).cuda()
    K = {True: synth.intrinsics(p640).astype(np.float32), False: synth.intrinsics(p320).astype(np.float32)}
    steps = int(nfr.max())
    depth, poses = [], []
    for s in range(steps):
        drow, prow = [], []
        for i in range(n_maps):
            if s >= nfr[i]:
                drow.append(None)
                prow.append(np.eye(4))
                continue
            t = int(start[i] + s)
            drow.append(dev_depth[t] if hi[i] else d320[t])
            prow.append(np.linalg.inv(synth.pose(p640, int(start[i]))) @ synth.pose(p640, t))
        depth.append(drow)
        poses.append(prow)
    maps = [rf.MapStruct(K[bool(hi[i])], 1 << 18, 1 << 17 if hi[i] else 1 << 16, VOXEL, TRUNC) for i in range(n_maps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        for i in range(n_maps):
            if depth[s][i] is not None:
                maps[i].FuseU16(depth[s][i], poses[s][i])
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    blocks_plain = [m.stats().num_allocated for m in maps]
    for m in maps:
        m.Reset()
    b = rf.FusionBatch(maps)
    b.plan(depth, poses)
    cfg = b.autotune(reset_maps=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    blocks_batch = [m.stats().num_allocated for m in maps]
    total = int(nfr.sum())
    out = {"submaps": n_maps, "frames": total, "steps": steps, "frames_640x480": int(nfr[hi].sum()),
           "frames_320x240": int(nfr[~hi].sum()), "blocks": int(sum(blocks_batch)),
           "plain_frames_per_s": round(total / t_plain, 1), "batched_frames_per_s": round(total / t_batch, 1),
           "speedup": round(t_plain / t_batch, 2), "same_block_counts": blocks_plain == blocks_batch,
           "autotuned": cfg}
    b.close()
    for m in maps:
        m.close()
    return out


def serving_bench(dev_depth, dev_rgb, nseq, n_streams=32, seed=2, frames_sat=60, duration_s=4.0, runs=None):
    """Serving form: 32 streams, each a segment of the room sequence starting
    at its own seeded frame (so its own start angle) at 640x480 or 320x240
    (seeded), voxel 5 mm, 2^20 buckets and a 2^16-block pool per stream, f32
    storage (and, for one saturation run, the precision form's float16 sdf +
    uint16 weight), the tracker of the main workload.  Each server captures
    its batch graphs when it is set up, before any timing.
      saturation: every stream's first `frames_sat` frames queued as fast as
                  the host submits them (pumping after each round of
                  streams); aggregate frames/s from first submit to last
                  result.
      poisson:    frames arrive at 15-30 Hz per stream (seeded rate,
                  exponential gaps) for `duration_s`; per-frame latency from
                  the scheduled arrival to the result's collection."""
    import torch
    from yetanotherslam_amd import rgbdfusion as rf
    from yetanotherslam_amd import synth

    rng = np.random.default_rng(seed)
    hi = rng.integers(0, 2, n_streams).astype(bool)
    p640 = synth.params(width=W, height=H, theta0=THETA0)
    p320 = synth.params(width=W // 2, height=H // 2, theta0=THETA0)
    lo_frames = [synth.frame(p320, t) for t in range(nseq)]
    d320 = torch.from_numpy(np.stack([f[0] for f in lo_frames]).view(np.int16)).cuda()
    c320 = torch.from_numpy(np.stack([f[1] for f in lo_frames])).cuda()
    K640, K320 = synth.intrinsics(p640), synth.intrinsics(p320)
    tp = rf.tracker_params(LEVELS, ITERS, rf.TRACK_RGBD, WEIGHTS)
    torch.cuda.synchronize()

    def frame(i, k):
        return (dev_depth[k], dev_rgb[k]) if hi[i] else (d320[k], c320[k])

    def new_server(fmt=rf.VOXEL_F32_F32):
        m = rf.map_params(1 << 20, 1 << 16, VOXEL, TRUNC, fmt)
        sv = rf.Server(W, H, n_streams, tp, voxel_format=fmt, hbm_budget=200 << 30)
        ids = [sv.add_stream(W, H, K640, m) if hi[i] else sv.add_stream(W // 2, H // 2, K320, m)
               for i in range(n_streams)]
        sv.prepare()  # batch graphs captured at set-up, as a long-running server would
        return sv, ids

    out = {"streams": n_streams, "streams_640x480": int(hi.sum()), "streams_320x240": int((~hi).sum())}
    mixed = hi.copy()
    # saturation: all streams at 640x480, then the seeded resolution mix
    start = rng.integers(0, nseq - frames_sat, n_streams)
    for name, hires, fmt in (("saturation_640x480", np.ones(n_streams, bool), rf.VOXEL_F32_F32),
                             ("saturation_640x480_f16u16", np.ones(n_streams, bool), rf.VOXEL_F16_U16),
                             ("saturation", mixed, rf.VOXEL_F32_F32)):
        if runs and name not in runs:
            continue
        hi[:] = hires
        sv, ids = new_server(fmt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(frames_sat):
            for i, sid in enumerate(ids):
                sv.submit(sid, *frame(i, int(start[i]) + k))
            sv.pump()  # batches start while later frames are still being queued
        sv.drain()
        t_sat = time.perf_counter() - t0
        res = sv.poll(n_streams * frames_sat + 1)
        st = sv.stats()
        out[name] = {"frames": len(res), "frames_per_s": round(len(res) / t_sat, 1),
                     "tracked": int(sum(r["tracked"] for r in res)),
                     "mean_batch": round(st["frames_done"] / max(st["batches_done"], 1), 1),
                     "hbm_bytes_in_use": int(st["bytes_in_use"])}
        sv.close()
    hi[:] = mixed
    if runs and "poisson" not in runs:
        return out
    # Poisson arrivals
    rate = rng.uniform(15.0, 30.0, n_streams)
    arrivals = []
    for i in range(n_streams):
        t, k = float(rng.exponential(1.0 / rate[i])), 0
        while t < duration_s and k < nseq:
            arrivals.append((t, i, k))
            t += float(rng.exponential(1.0 / rate[i]))
            k += 1
    arrivals.sort()
    start = rng.integers(0, max(1, nseq - max(a[2] for a in arrivals) - 1), n_streams)
    sv, ids = new_server()
    torch.cuda.synchronize()
    due = {}
    lat = []
    nxt = 0
    t0 = time.perf_counter()
    while len(lat) < len(arrivals):
        now = time.perf_counter() - t0
        while nxt < len(arrivals) and arrivals[nxt][0] <= now:
            ta, i, k = arrivals[nxt]
            due[sv.submit(ids[i], *frame(i, int(start[i]) + k))] = ta
            nxt += 1
        sv.pump()
        got = sv.poll()
        now = time.perf_counter() - t0
        for r in got:
            lat.append(now - due.pop(r["request"]))
        if not got and nxt < len(arrivals) and arrivals[nxt][0] > now:
            time.sleep(min(2e-4, arrivals[nxt][0] - now))
    t_end = time.perf_counter() - t0
    st = sv.stats()
    lat_ms = np.array(lat) * 1e3
    out["poisson"] = {"frames": len(lat), "offered_frames_per_s": round(len(arrivals) / duration_s, 1),
                      "frames_per_s": round(len(lat) / t_end, 1),
                      "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                                     "p99": round(float(np.percentile(lat_ms, 99)), 3),
                                     "max": round(float(lat_ms.max()), 3)},
                      "mean_batch": round(st["frames_done"] / max(st["batches_done"], 1), 1),
                      "rates_hz": [round(float(rate.min()), 1), round(float(rate.max()), 1)]}
    sv.close()
    return out


def copy_probe():
    """A plain device copy of known size (512 MiB read + 512 MiB written),
    used to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 128 << 20
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": n * 4, "launches": 5}), flush=True)


def main():
    # Python's cyclic collector stays off for the run: with this process's
    # heap a full collection takes tens of milliseconds, and one landing in a
    # timed region (a serving run's submission loop, say) would be measured
    # as device time.  Reference counting still frees everything this
    # script allocates.
    gc.disable()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=290)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-probe", action="store_true", help="run only the counter-calibration copy")
    ap.add_argument("--no-mesh", action="store_true", help="skip the marching-cubes measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-submap measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serving-runs", default="",
                    help="comma-separated subset of saturation_640x480, saturation_640x480_f16u16, saturation, poisson")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the dense path is a per-stream, per-volume workload "
              "(one process per GPU); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from yetanotherslam_amd import rgbdfusion as rf
    from yetanotherslam_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    p = synth.params(width=W, height=H, theta0=THETA0)
    K = synth.intrinsics(p)
    n = args.warmup + args.steps + 1
    npool = max(n, 200)  # the batching and serving forms draw segments from this sequence too
    log("generating %d frames" % npool)
    depth, rgb = gen_frames(p, npool)
    dev_depth = torch.from_numpy(depth.view(np.int16)).cuda()
    dev_rgb = torch.from_numpy(rgb).cuda()
    torch.cuda.synchronize()

    ctx = rf.default_context()
    mp = rf.map_params(BUCKETS, BLOCKS, VOXEL, TRUNC, rf.VOXEL_F32_F32)
    tp = rf.tracker_params(LEVELS, ITERS, rf.TRACK_RGBD, WEIGHTS)
    pl = rf.DensePipeline(W, H, K, mp, tp, ctx=ctx)
    lib = rf.native

    log("warmup %d frames" % (args.warmup + 1))
    for t in range(args.warmup + 1):
        pl.step(dev_depth[t], dev_rgb[t])
    torch.cuda.synchronize()

    lib.call("yas_pipeline_enable_timing", pl.h, 1)
    torch.cuda.synchronize()
    log("timing %d frames" % args.steps)
    t0 = time.perf_counter()
    for t in range(args.warmup + 1, n):
        pl.step(dev_depth[t], dev_rgb[t])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    stage = np.zeros(8)
    frame_ms = np.zeros(args.steps)
    nt = rf.ctypes.c_int()
    lib.call("yas_pipeline_timing", pl.h, stage.ctypes.data, frame_ms.ctypes.data, args.steps, rf.ctypes.byref(nt))
    stats = np.zeros((args.steps, 8), np.int32)
    lib.call("yas_pipeline_frame_stats", pl.h, args.warmup + 1, args.steps, stats.ctypes.data)
    tracked = sum(pl.pose(t)[1] for t in range(args.warmup + 1, n))

    fps = args.steps / wall
    per_stage = {k: round(v / args.steps, 4) for k, v in zip(STAGES, stage)}
    rl = rooflines(stats, stage, args.steps)
    dominant = max(rl, key=lambda k: rl[k]["avg_launch_ms"])
    log("stages ms/frame:", per_stage)
    log("visible blocks/frame: fuse %.0f, ray cast %.0f; allocated at end %d; tracked %d/%d" % (
        stats[:, 1].mean(), stats[:, 4].mean(), stats[-1, 0], tracked, args.steps))

    result = {
        "metric": "frames/s (integrate+raycast+ICP) 640x480@5mm; p50/p99 frame latency",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "room+spheres RGB-D sequence, 640x480, voxel 5 mm, truncation 4 voxels, 2^22 buckets, "
                        "f32 sdf/weight, ICP+photometric (w=0.1), 3 levels 10/5/3",
            "frames": args.steps, "width": W, "height": H, "voxel_m": VOXEL, "trunc_m": TRUNC,
            "buckets": BUCKETS, "block_pool": BLOCKS, "parallelism": "single stream, one GPU",
        },
        "latency_ms": {"p50": round(float(np.percentile(frame_ms, 50)), 4),
                       "p99": round(float(np.percentile(frame_ms, 99)), 4),
                       "mean_device": round(float(frame_ms.mean()), 4)},
        "stage_ms_per_frame": per_stage,
        "tracked_frames": int(tracked),
        "roofline": dict(rl[dominant], kernel=dominant),
        "roofline_by_kernel": rl,
    }
    if not args.no_serving:
        log("serving form: 32 streams")
        runs = set(args.serving_runs.split(",")) if args.serving_runs else None
        result["serving"] = serving_bench(dev_depth, dev_rgb, npool, runs=runs)
        for key in ("saturation_640x480", "saturation_640x480_f16u16", "saturation"):
            if key not in result["serving"]:
                continue
            sat = result["serving"][key]["frames_per_s"]
            result["serving"][key]["vs_plain_single_stream"] = round(sat / fps, 2)
        log("serving:", result["serving"])
    if not args.no_batch:
        log("batching form: 64 submaps")
        result["batching"] = batch_bench(dev_depth, npool)
        log("batching:", result["batching"])
    if not args.no_mesh:
        log("marching cubes over the fused volume")
        result["mesh"] = mesh_bench(pl, dev_depth, K, n)
        log("mesh:", result["mesh"])
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread, <= %.0f s)" % args.cpu_budget)
        v, nf, dt = cpu_baseline(p, K, depth, rgb, args.cpu_budget)
        result["cpu_baseline"] = {"value": round(v, 4), "unit": "frames/s", "cores": 1, "kind": "port",
                                  "sample": "frames 1..%d of the same sequence after frame 0 fused (%.1f s)" % (nf, dt)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
