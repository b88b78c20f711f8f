"""GPU batch decode: one launch sequence over frames of mixed size, profile,
slice width, scan, alpha depth and quantisation matrices, under every launch
form (grid or persistent work queue, IDCT packing, workgroup size, hipGraph
replay). Each frame must decode to exactly the oracle's samples, and a damaged
frame in the batch must report the oracle's error without disturbing the
others (the first error in picture order, src/decode.rs:288-289).
"""

import itertools

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu

# (width, height, chroma, depth, profile, log2_slice_mbs, interlace, alpha type, loaded matrices)
CASES = [
    (720, 486, 2, 10, 3, 1, 1, 0, False),   # 720x486i HQ, 2-MB slices, top field first
    (1280, 720, 2, 10, 1, 2, 0, 0, False),  # 720p LT, 4-MB slices
    (1920, 1080, 2, 10, 0, 0, 0, 0, False),  # 1080 Proxy, 1-MB slices
    (1024, 540, 3, 12, 4, 3, 0, 1, True),   # 4444, 8-bit alpha, loaded matrices
    (333, 75, 3, 12, 5, 3, 2, 2, False),    # odd size, bottom field first, 16-bit alpha
    (96, 32, 2, 10, 2, 3, 0, 0, False),     # tiny
    (4096, 2160, 3, 12, 4, 3, 0, 1, True),  # the batch's 4096x2160 4444 kind: 8-bit alpha, loaded matrices
]


def make_case(i, case):
    w, h, chroma, depth, profile, log2, interlace, atype, matrices = case
    rng = np.random.default_rng(100 + i)
    f = O.test_frame(w, h, chroma, depth, int(rng.integers(2, 40)), 50 + i)
    f.interlace = interlace
    if atype:
        mx = (1 << depth) - 1
        yy, xx = np.mgrid[0:h, 0:w]
        d = np.hypot((xx - w * 0.4) / (0.3 * w), (yy - h * 0.5) / (0.3 * h))
        f.alpha = (np.clip((1.0 - d) * min(w, h) * 0.3 / 16.0, 0, 1) * mx).round().astype(np.uint16).ravel()
    luma = chroma_m = None
    if matrices:
        luma = [int(v) for v in rng.integers(4, 33, 64)]
        chroma_m = [int(v) for v in rng.integers(4, 33, 64)]
    cfg = O.config(profile, log2_slice_mbs=log2, alpha=atype or 2, luma=luma, chroma=chroma_m)
    return O.encode(f, cfg)


@pytest.fixture(scope="module")
def batch():
    packets = [make_case(i, c) for i, c in enumerate(CASES)]
    # A damaged copy of the 720p frame: a slice-size table entry made huge.
    bad = bytearray(packets[1])
    hsize = int.from_bytes(bad[8:10], "big")
    pic = 8 + hsize
    phs = bad[pic] >> 3
    bad[pic + phs + 6] = 0xFF
    bad[pic + phs + 7] = 0xF0
    packets.append(bytes(bad))
    want = [O.decode_result(p) for p in packets]
    bufs, jobs = [], []
    for p in packets:
        h = P.FrameHeader.parse(p)
        planes = P.layout(h.width, h.height, h.chroma)
        n = planes[2][0] + planes[2][1] * planes[2][2]
        dp = P.DeviceBuffer(len(p) + 128)
        dp.upload(np.frombuffer(p, np.uint8))
        dd = P.DeviceBuffer(n * 2 + 64)
        da = P.DeviceBuffer(h.width * h.height * 2 + 64) if h.alpha else None
        bufs.append((dp, dd, da, n, h))
        jobs.append((p, dp.ptr, dd.ptr, da.ptr if da else 0))
    return packets, want, bufs, P.DecodeJobs(jobs)


def check_batch(errs, want, bufs):
    for i, (e, w, (dp, dd, da, n, h)) in enumerate(zip(errs, want, bufs)):
        if w[0] != "Ok":
            assert e is not None and (e.kind, e.msg) == (w[0], w[1]), (i, e, w)
            continue
        assert e is None, (i, e)
        got = dd.download(n * 2, dtype=np.uint16)
        np.testing.assert_array_equal(got, w[1].data, err_msg=f"frame {i} samples")
        if w[1].alpha is not None:
            np.testing.assert_array_equal(da.download(h.width * h.height * 2, dtype=np.uint16), w[1].alpha,
                                          err_msg=f"frame {i} alpha")


TUNINGS = [P.DecodeTuning(pers, chunk, blocks, wg, graph)
           for pers, chunk, blocks, wg, graph in itertools.product((0, 1), (1, 3), (8, 40, 96), (0, 64, 256, 768), (0, 1))
           if pers or chunk == 1]


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: f"p{t.persistent}c{t.chunk}b{t.idct_blocks}w{t.entropy_wg}g{t.graph}")
def test_mixed_batch_every_tuning(batch, tuning):
    packets, want, bufs, jobs = batch
    print("case: mixed batch", tuning, flush=True)
    dec = P.Decoder()
    for dp, dd, da, n, h in bufs:  # stale output must not survive
        dd.upload(np.full(n, 0xABCD, np.uint16))
    check_batch(dec.decode_device_batch(jobs, tuning), want, bufs)


def test_graph_replays_with_eager_work_between(batch):
    """A captured graph replayed several times, with other batches run eagerly
    between the replays, gives the same samples each time."""
    packets, want, bufs, jobs = batch
    dec = P.Decoder()
    g = P.DecodeTuning(1, 2, 64, 256, 1)
    eager = P.DecodeTuning(0, 1, 96, 256, 0)
    single = P.DecodeJobs([(packets[0], bufs[0][0].ptr, bufs[0][1].ptr, 0)])
    for rep in range(4):
        print("case: graph replay", rep, flush=True)
        for dp, dd, da, n, h in bufs:
            dd.upload(np.zeros(n, np.uint16))
        check_batch(dec.decode_device_batch(jobs, g), want, bufs)
        assert dec.decode_device_batch(single, eager) == [None]


@pytest.mark.parametrize("device_planned", (False, True), ids=("host-planned", "device-planned"))
def test_two_callers_at_once(batch, device_planned):
    """Two threads calling the batch entry point at once (each call takes a
    free per-call stream, so both batches are in flight together), one of
    them replaying a captured graph, each into planes of its own: every
    result of every call equals the oracle's. Device-planned: the packets'
    headers are read on the device (each stream's workspace plans them)."""
    import threading
    packets, want, bufs, jobs = batch
    if device_planned:
        jobs = P.DecodeJobs([(len(p), dp.ptr, dd.ptr, da.ptr if da else 0)
                             for p, (dp, dd, da, n, h) in zip(packets, bufs)])
    bufs2, jobs2 = [], []
    for p, (dp, dd, da, n, h) in zip(packets, bufs):
        dd2 = P.DeviceBuffer(n * 2 + 64)
        da2 = P.DeviceBuffer(h.width * h.height * 2 + 64) if h.alpha else None
        bufs2.append((dp, dd2, da2, n, h))
        jobs2.append((len(p) if device_planned else p, dp.ptr, dd2.ptr, da2.ptr if da2 else 0))
    jobs2 = P.DecodeJobs(jobs2)
    dec = P.Decoder()
    failed = []

    def run(j, b, tuning, tag):
        try:
            for rep in range(6):
                for dp, dd, da, n, h in b:
                    dd.upload(np.full(n, 0xABCD, np.uint16))
                check_batch(dec.decode_device_batch(j, tuning), want, b)
        except Exception as e:  # reported after the join
            failed.append((tag, e))
    print("case: two callers at once, device planned" if device_planned else "case: two callers at once", flush=True)
    ts = [threading.Thread(target=run, args=(jobs, bufs, P.DecodeTuning(), "eager")),
          threading.Thread(target=run, args=(jobs2, bufs2, P.DecodeTuning(0, 1, 96, 256, 1), "graph"))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not failed, failed
    assert all(v >= 0 for v in P.last_batch_timing())


def test_autotune_then_batch(batch):
    packets, want, bufs, jobs = batch
    dec = P.Decoder()
    print("case: autotune", flush=True)
    best, us = dec.autotune(jobs, reps=1)
    assert us > 0
    assert P.get_decode_tuning() == best
    check_batch(dec.decode_device_batch(jobs), want, bufs)
    P.set_decode_tuning(P.DecodeTuning())


@pytest.mark.parametrize("graph", (0, 1))
def test_device_only_packets(batch, graph):
    """Jobs without a host copy of their packet: the batch is planned on the
    device (headers parsed there). Beside them, jobs with a host copy, and packets cut short
    (inside the frame header, inside a picture header, inside the slices),
    whose errors must be the oracle's."""
    packets, want, bufs, _ = batch
    print("case: device-only packets, graph", graph, flush=True)
    cut = packets[3]
    hsize = int.from_bytes(cut[8:10], "big")
    # The interlaced frame with its second picture's picture_size made huge,
    # and a progressive one with a picture_header_size of zero.
    second = bytearray(packets[0])
    pos2 = 8 + int.from_bytes(second[8:10], "big")
    pos2 += int.from_bytes(second[pos2 + 1:pos2 + 5], "big")
    second[pos2 + 1:pos2 + 5] = b"\xff\xff\xff\xf0"
    zero = bytearray(packets[2])
    zero[8 + int.from_bytes(zero[8:10], "big")] &= 7
    short = [cut[:20], cut[:8 + hsize + 5], cut[:len(cut) // 2], bytes(second), bytes(zero)]
    extra = []
    for p in short:
        dp = P.DeviceBuffer(len(p) + 128)
        dp.upload(np.frombuffer(p, np.uint8))
        dd = P.DeviceBuffer(16 << 20)  # a 1920x1080 4:2:2 frame's planes take 8.3 MB
        da = P.DeviceBuffer(4 << 20)
        extra.append((dp, dd, da))
    jobs = [(len(p) if i % 2 == 0 else p, dp.ptr, dd.ptr, da.ptr if da else 0)
            for i, (p, (dp, dd, da, n, h)) in enumerate(zip(packets, bufs))]
    jobs += [(len(p), dp.ptr, dd.ptr, da.ptr) for p, (dp, dd, da) in zip(short, extra)]
    for dp, dd, da, n, h in bufs:
        dd.upload(np.full(n, 0xABCD, np.uint16))
    dec = P.Decoder()
    errs = dec.decode_device_batch(P.DecodeJobs(jobs), P.DecodeTuning(1, 1, 64, 0, graph))
    check_batch(errs[:len(packets)], want, bufs)
    for p, e in zip(short, errs[len(packets):]):
        w = O.decode_result(p)
        assert w[0] != "Ok" and e is not None and (e.kind, e.msg) == (w[0], w[1]), (len(p), e, w)


def test_jobs_without_outputs_fail_alone(batch):
    """A job with no device packet, no planes, or no alpha plane for a frame
    with alpha fails with a configuration error and decodes nothing; the
    other jobs of the batch decode."""
    packets, want, bufs, _ = batch
    print("case: jobs without outputs", flush=True)
    dp, dd, da, n, h = bufs[3]  # the frame with 8-bit alpha
    assert h.alpha
    good = bufs[2]
    good[1].upload(np.zeros(good[3], np.uint16))
    jobs = [(packets[3], dp.ptr, dd.ptr, 0), (len(packets[3]), 0, dd.ptr, da.ptr),
            (packets[3], dp.ptr, 0, da.ptr), (len(packets[2]), good[0].ptr, good[1].ptr, 0)]
    errs = P.Decoder().decode_device_batch(jobs)
    assert [type(e) for e in errs[:3]] == [P.ConfigError] * 3, errs
    assert errs[3] is None
    np.testing.assert_array_equal(good[1].download(good[3] * 2, dtype=np.uint16), want[2][1].data)


@pytest.mark.parametrize("persistent", (0, 1))
def test_batch_past_the_ordering_limit(persistent):
    """22 copies of a 1920x1080 Proxy frame with 1-MB slices: 179,520 slices,
    8,415 entropy items, more than the entropy ordering sorts (the items then
    go in their natural order). Every copy must decode to the oracle's
    samples."""
    print("case: batch past the ordering limit, persistent", persistent, flush=True)
    f = O.test_frame(1920, 1080, 2, 10, 12, 3)
    packet = O.encode(f, O.config(0, log2_slice_mbs=0))
    want = O.decode(packet)
    n = 22
    h = P.FrameHeader.parse(packet)
    planes = P.layout(h.width, h.height, h.chroma)
    total = planes[2][0] + planes[2][1] * planes[2][2]
    dp = P.DeviceBuffer(len(packet) + 128)
    dp.upload(np.frombuffer(packet, np.uint8))
    outs = [P.DeviceBuffer(total * 2 + 64) for _ in range(n)]
    for o in outs:
        o.upload(np.full(total, 0xABCD, np.uint16))
    jobs = P.DecodeJobs([(packet, dp.ptr, o.ptr, 0) for o in outs])
    errs = P.Decoder().decode_device_batch(jobs, P.DecodeTuning(persistent, 4 if persistent else 1, 96, 0, 0))
    assert errs == [None] * n, errs
    for i, o in enumerate(outs):
        np.testing.assert_array_equal(o.download(total * 2, dtype=np.uint16), want.data, err_msg=f"copy {i}")


_FRESH_PLAN = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import prores_amd as P
packet = open(sys.argv[2], "rb").read()
h = P.FrameHeader.parse(packet)
planes = P.layout(h.width, h.height, h.chroma)
n = planes[2][0] + planes[2][1] * planes[2][2]
dp = P.DeviceBuffer(len(packet) + 128)
dp.upload(np.frombuffer(packet, np.uint8))
outs = [P.DeviceBuffer(n * 2 + 64) for _ in range(3)]
dec = P.Decoder()
for k in range(3):  # the first batch outgrows the first capacities; then 3 frames; then 1 again
    jobs = [(len(packet), dp.ptr, o.ptr, 0) for o in outs[:(1, 3, 1)[k]]]
    errs = dec.decode_device_batch(jobs)
    assert errs == [None] * len(jobs), errs
np.save(sys.argv[3], np.stack([o.download(n * 2, dtype=np.uint16) for o in outs]))
"""


def test_device_plan_grows_its_workspace_in_a_fresh_process(tmp_path):
    """The first device-planned batch of a process is planned into first
    capacities; a 1920x1080 Proxy frame of one-macroblock slices (8,160
    slices) outgrows them, so the batch does no work, the workspace grows to
    the batch's totals and it runs again. Every frame equals the oracle's."""
    import os
    import subprocess
    import sys

    f = O.test_frame(1920, 1080, 2, 10, 12, 3)
    packet = O.encode(f, O.config(0, log2_slice_mbs=0))
    want = O.decode(packet).data
    path, out = tmp_path / "p.bin", tmp_path / "out.npy"
    path.write_bytes(packet)
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rivet-prores_amd")
    print("case: fresh-process device plan over the first capacities", flush=True)
    r = subprocess.run([sys.executable, "-c", _FRESH_PLAN, pkg, str(path), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = np.load(out)
    for k in range(3):
        np.testing.assert_array_equal(got[k], want)
