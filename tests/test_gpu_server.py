"""GPU tests of the serving engine: a stream's results are bit-identical
alone, in a mix of streams of other resolutions, and to the single-stream
pipeline; streams are admitted by the HBM budget; continuous batching keeps
each stream's frames in order."""
import math

import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from yetanotherslam_amd import native
from yetanotherslam_amd import rgbdfusion as rf
from yetanotherslam_amd import synth

from helpers import assert_bits_equal, to_dev

pytestmark = pytest.mark.gpu

VOX, MU = 0.008, 0.024
TP = dict(levels=3, iters=(10, 5, 3), mode=rf.TRACK_RGBD, weights=(0.1, 0.1, 1.0, 1.0))


def stream_frames(W, H, theta0, n):
    p = synth.params(width=W, height=H, theta0=theta0)
    fr = [synth.frame(p, t) for t in range(n)]
    return p, synth.intrinsics(p), [(to_dev(d), to_dev(c)) for d, c, _ in fr]


def map_params():
    return rf.map_params(1 << 16, 1 << 14, VOX, MU, rf.VOXEL_F32_F32)


def tracker_params():
    return rf.tracker_params(TP["levels"], TP["iters"], TP["mode"], TP["weights"])


def run_single(p, K, frames):
    pl = rf.DensePipeline(p.width, p.height, K, map_params(), tracker_params())
    for d, c in frames:
        pl.step(d, c)
    torch.cuda.synchronize()
    poses = [pl.pose(t) for t in range(len(frames))]
    return poses, pl.export_map_sorted(rf.VOXEL_F32_F32)


def run_server(streams, n, interleave=True, prepare=False):
    """streams: list of (p, K, frames).  Submits frame t of every stream
    before frame t+1 of any, pumping as it goes.  prepare: capture the batch
    graphs before the first submit instead of at each size's first batch."""
    sv = rf.Server(320, 240, 8, tracker_params())
    ids = [sv.add_stream(p.width, p.height, K, map_params()) for p, K, _ in streams]
    if prepare:
        sv.prepare()
    torch.cuda.synchronize()
    results = {i: [] for i in ids}
    for t in range(n):
        for sid, (_, _, fr) in zip(ids, streams):
            sv.submit(sid, *fr[t])
        if interleave:
            sv.pump()
            for r in sv.poll():
                results[r["stream"]].append(r)
    sv.drain()
    for r in sv.poll():
        results[r["stream"]].append(r)
    maps = [sv.stream_map(sid).export_sorted() for sid in ids]
    return sv, ids, results, maps


def test_stream_is_identical_alone_in_a_mix_and_in_the_pipeline():
    n = 5
    a = stream_frames(320, 240, math.pi / 4, n)
    b = stream_frames(160, 120, math.pi / 3, n)
    c = stream_frames(320, 240, 1.2, n)
    single_poses, single_map = run_single(*a)
    _, _, alone, alone_maps = run_server([a], n)
    sv, ids, mix, mix_maps = run_server([b, a, c], n, prepare=True)
    ra_alone = alone[0]
    ra_mix = mix[ids[1]]
    assert len(ra_alone) == len(ra_mix) == n
    # requests of a stream come back in submission order (also after the
    # ring of in-flight batches wraps around)
    for rs in (ra_alone, ra_mix):
        assert [r["request"] for r in rs] == sorted(r["request"] for r in rs)
    for t in range(n):
        Ts, ok = single_poses[t]
        assert ra_alone[t]["tracked"] == ok and ra_mix[t]["tracked"] == ok
        np.testing.assert_array_equal(ra_alone[t]["pose"], Ts)
        np.testing.assert_array_equal(ra_mix[t]["pose"], Ts)
    for arr, ref in zip(mix_maps[1], single_map):
        np.testing.assert_array_equal(arr, ref)
    for arr, ref in zip(alone_maps[0], single_map):
        np.testing.assert_array_equal(arr, ref)
    # the other streams of the mix match their own single pipelines too
    for sid, s in ((ids[0], b), (ids[2], c)):
        poses, m = run_single(s[0], s[1], s[2])
        for t in range(n):
            np.testing.assert_array_equal(mix[sid][t]["pose"], poses[t][0])
    st = sv.stats()
    assert st["frames_done"] == 3 * n and st["batches_done"] >= n


def test_admission_by_hbm_budget():
    p, K, _ = stream_frames(160, 120, 0.5, 1)
    probe = rf.Server(160, 120, 4, tracker_params(), hbm_budget=1 << 40)
    probe.add_stream(160, 120, K, map_params())
    one = probe.stats()["bytes_in_use"]
    probe.close()
    sv = rf.Server(160, 120, 4, tracker_params(), hbm_budget=int(2.5 * one))
    s0 = sv.add_stream(160, 120, K, map_params())
    s1 = sv.add_stream(160, 120, K, map_params())
    with pytest.raises(native.YasError) as e:
        sv.add_stream(160, 120, K, map_params())
    assert e.value.code == 6  # YAS_E_CAPACITY
    sv.remove_stream(s0)
    s2 = sv.add_stream(160, 120, K, map_params())
    assert sv.stats()["streams"] == 2 and s2 != s1


def test_frames_arriving_during_a_batch_join_later_batches():
    """Continuous batching: a stream that submits several frames at once
    gets them in consecutive batches, in order; a stream that joins later is
    batched with the others' later frames; everything equals running each
    stream alone."""
    n = 4
    a = stream_frames(160, 120, 0.7, n)
    b = stream_frames(160, 120, 1.9, n)
    sv = rf.Server(160, 120, 4, tracker_params())
    ia = sv.add_stream(160, 120, a[1], map_params())
    torch.cuda.synchronize()
    for t in range(n):
        sv.submit(ia, *a[2][t])
    sv.pump()
    ib = sv.add_stream(160, 120, b[1], map_params())
    torch.cuda.synchronize()
    for t in range(n):
        sv.submit(ib, *b[2][t])
    sv.drain()
    res = sv.poll()
    by = {ia: [r for r in res if r["stream"] == ia], ib: [r for r in res if r["stream"] == ib]}
    for sid, s in ((ia, a), (ib, b)):
        assert [r["request"] for r in by[sid]] == sorted(r["request"] for r in by[sid]) and len(by[sid]) == n
        poses, _ = run_single(s[0], s[1], s[2])
        for t in range(n):
            np.testing.assert_array_equal(by[sid][t]["pose"], poses[t][0])


def results_by_stream(res):
    out = {}
    for r in res:
        out.setdefault(r["stream"], []).append(r)
    for v in out.values():
        v.sort(key=lambda r: r["request"])
    return out


def check_against_single(rs, s, n=None):
    """Results rs of a stream equal its own single pipeline's first n frames."""
    n = len(rs) if n is None else n
    poses, m = run_single(s[0], s[1], s[2][:n])
    assert len(rs) == n
    for t in range(n):
        assert rs[t]["tracked"] == poses[t][1]
        np.testing.assert_array_equal(rs[t]["pose"], poses[t][0])
    return m


def test_closing_a_stream_mid_run_leaves_the_others_identical():
    """One of three streams is closed after its third frame while the others
    keep going: the closed stream's three results and the others' five equal
    each stream run alone in a single pipeline, volumes included."""
    n = 5
    a = stream_frames(320, 240, 0.4, n)
    b = stream_frames(160, 120, 1.3, n)
    c = stream_frames(320, 240, 2.2, n)
    sv = rf.Server(320, 240, 4, tracker_params())
    ids = [sv.open_stream(s[0].width, s[0].height, s[1], map_params()) for s in (a, b, c)]
    sv.prepare()
    res = []
    for t in range(n):
        for sid, s in zip(ids, (a, b, c)):
            if sid == ids[1] and t >= 3:
                continue
            sv.submit(sid, *s[2][t])
        if t == 2:
            sv.close_stream(ids[1])  # its three frames still run
        sv.pump()
        res += sv.poll()
    sv.drain()
    res += sv.poll()
    by = results_by_stream(res)
    assert sv.stream_state(ids[1]) == "free"
    check_against_single(by[ids[1]], b, 3)
    for sid, s in ((ids[0], a), (ids[2], c)):
        m = check_against_single(by[sid], s)
        for arr, ref in zip(sv.stream_map(sid).export_sorted(), m):
            np.testing.assert_array_equal(arr, ref)
    sv.close()


def stream_budget(K):
    probe = rf.Server(160, 120, 4, tracker_params(), hbm_budget=1 << 40)
    probe.add_stream(160, 120, K, map_params())
    one = probe.stats()["bytes_in_use"]
    probe.close()
    return one


def test_streams_over_budget_wait_and_are_admitted_as_memory_frees():
    """Budget for two streams, three opened: the third waits (pending) with
    its frames queued, is admitted when the first closes, and its results
    equal its own single pipeline's, as do the others'."""
    n = 4
    s3 = [stream_frames(160, 120, th, n) for th in (0.3, 1.5, 2.7)]
    one = stream_budget(s3[0][1])
    sv = rf.Server(160, 120, 4, tracker_params(), hbm_budget=int(2.5 * one))
    ids = [sv.open_stream(160, 120, s[1], map_params()) for s in s3]
    assert [sv.stream_state(i) for i in ids] == ["active", "active", "pending"]
    assert sv.stats()["streams_pending"] == 1
    for t in range(n):
        for sid, s in zip(ids, s3):
            sv.submit(sid, *s[2][t])
    res = []
    for _ in range(3):
        sv.pump()
        res += sv.poll()
    assert sv.stream_state(ids[2]) == "pending"
    sv.close_stream(ids[0])  # its frames finish, then its memory goes to the third stream
    sv.drain()
    res += sv.poll()
    assert sv.stream_state(ids[0]) == "free" and sv.stream_state(ids[2]) == "active"
    # results come back in completion order: the first stream's last frame
    # finished before the third stream's first
    order = [r["stream"] for r in res]
    assert max(i for i, s in enumerate(order) if s == ids[0]) < order.index(ids[2])
    by = results_by_stream(res)
    for sid, s in zip(ids, s3):
        check_against_single(by[sid], s)
    sv.close()


def test_idle_streams_are_hibernated_to_admit_others_and_resume_exactly():
    """Budget for two streams with eviction on: opening a third hibernates
    the least recently used idle stream (its volume paged out to host
    memory); that stream's next frame pages it back in, hibernating the other
    idle one.  Every stream's results equal its own single pipeline's, so
    paging changes nothing."""
    n = 4
    s3 = [stream_frames(160, 120, th, n) for th in (0.6, 1.8, 3.0)]
    one = stream_budget(s3[0][1])
    sv = rf.Server(160, 120, 4, tracker_params(), hbm_budget=int(2.5 * one))
    sv.set_eviction(0.0)
    a, b = (sv.open_stream(160, 120, s[1], map_params()) for s in s3[:2])
    res = []
    for t in range(2):
        sv.submit(a, *s3[0][2][t])
        sv.submit(b, *s3[1][2][t])
        sv.drain()
        res += sv.poll()
    c = sv.open_stream(160, 120, s3[2][1], map_params())
    st = sv.stats()
    assert sv.stream_state(c) == "active" and st["hibernations"] == 1 and st["streams_hibernated"] == 1
    hib = a if sv.stream_state(a) == "hibernated" else b
    for t in range(n):
        sv.submit(c, *s3[2][2][t])
    sv.drain()
    res += sv.poll()
    for t in range(2, n):
        sv.submit(a, *s3[0][2][t])
        sv.submit(b, *s3[1][2][t])
    sv.drain()
    res += sv.poll()
    st = sv.stats()
    assert st["reactivations"] >= 1 and st["hibernations"] >= 2
    by = results_by_stream(res)
    for sid, s in zip((a, b, c), s3):
        m = check_against_single(by[sid], s)
        if sv.stream_state(sid) == "active":
            for arr, ref in zip(sv.stream_map(sid).export_sorted(), m):
                np.testing.assert_array_equal(arr, ref)
    print("hibernated first:", hib, "stats:", st)
    sv.close()


@pytest.mark.parametrize("in_flight,ahead", [(1, 1), (2, 16), (3, 1)], ids=["one", "two_ahead16", "three"])
def test_queueing_policy_changes_batches_not_results(in_flight, ahead):
    """Batches held back (one in flight), queued ahead only when enough frames
    wait (the default), or always queued: the batches differ, every stream's
    poses equal its own pipeline's.  Three streams submit four frames each
    with a pump after each round."""
    n = 4
    ss = [stream_frames(160, 120, th, n) for th in (0.3, 1.1, 2.4)]
    sv = rf.Server(160, 120, 4, tracker_params())
    sv.set_max_in_flight(in_flight, ahead)
    ids = [sv.add_stream(160, 120, s[1], map_params()) for s in ss]
    torch.cuda.synchronize()
    for t in range(n):
        for sid, s in zip(ids, ss):
            sv.submit(sid, *s[2][t])
        sv.pump()
    sv.drain()
    by = results_by_stream(sv.poll())
    st = sv.stats()
    sv.close()
    assert st["frames_done"] == 3 * n
    for sid, s in zip(ids, ss):
        check_against_single(by[sid], s)
    bad = rf.Server(160, 120, 4, tracker_params())
    with pytest.raises(Exception):
        bad.set_max_in_flight(4, 1)
    bad.close()
