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


def test_autotune_then_batch(batch):
    packets, want, bufs, jobs = batch
    dec = P.Decoder()
    print("case: autotune", flush=True)
    best, us = dec.autotune(jobs, reps=1)
    assert us > 0
    assert P.get_decode_tuning() == best
    check_batch(dec.decode_device_batch(jobs), want, bufs)
    P.set_decode_tuning(P.DecodeTuning())
