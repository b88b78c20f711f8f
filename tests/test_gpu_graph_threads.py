"""Graph replay beside other threads' device work: with the device's batch
tuning set to capture and replay a graph, device batches on one thread run
while other threads copy through the library (prores_memcpy) and a serving
loop decodes and encodes with clients polling and reading results. Every
batch and every request must give the oracle's samples and bytes, and the
batches must actually replay a captured graph.
"""

import threading

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu


def _packets():
    out = []
    for i, (w, h, ch, bd, prof, log2, il) in enumerate([(720, 486, 2, 10, 3, 1, 1), (640, 360, 3, 12, 5, 3, 0),
                                                         (1280, 720, 2, 10, 0, 0, 0)]):
        f = O.test_frame(w, h, ch, bd, 6 + 4 * i, 700 + i)
        f.interlace = il
        out.append(O.encode(f, O.config(prof, log2_slice_mbs=log2)))
    return out


def test_graph_batches_beside_copies_and_serving():
    packets = _packets()
    want = [O.decode(p) for p in packets]
    bufs, jobs, keep = [], [], []
    for p in packets:
        hd = P.FrameHeader.parse(p)
        planes = P.layout(hd.width, hd.height, hd.chroma)
        n = planes[2][0] + planes[2][1] * planes[2][2]
        dp = P.DeviceBuffer(len(p) + 128)
        dp.upload(np.frombuffer(p, np.uint8))
        dd = P.DeviceBuffer(n * 2 + 64)
        bufs.append((dd, n))
        keep.append(dp)  # the packets must outlive every batch that reads them
        jobs.append((p, dp.ptr, dd.ptr, 0))
    jobs = P.DecodeJobs(jobs)
    enc_frame = O.test_frame(512, 288, 2, 10, 9, 77)
    enc_want = O.encode(enc_frame, O.config(3))
    pf = P.Frame(512, 288, 2, 10, enc_frame.data.copy())

    before = P.graph_counters()
    saved = P.get_decode_tuning()
    P.set_decode_tuning(P.DecodeTuning(persistent=1, chunk=2, idct_blocks=64, entropy_wg=256, graph=1))
    failures = []
    stop = threading.Event()

    def copier():
        scratch = P.DeviceBuffer(1 << 20)
        rng = np.random.default_rng(3)
        while not stop.is_set():
            a = rng.integers(0, 256, 1 << 20, dtype=np.uint8)
            scratch.upload(a)
            if not np.array_equal(scratch.download(), a):
                failures.append("copy round trip")
                return

    def server():
        with P.Server(hbm_budget=8 << 30, max_batch_frames=8) as srv:
            for rep in range(6):
                ids = {srv.submit_decode(packets[rep % 3]): ("d", rep % 3), srv.submit_encode(P.Config(3), pf): ("e", 0)}
                while ids:
                    for r in srv.poll(wait_ms=100):
                        kind, k = ids.pop(r.id)
                        if r.error is not None:
                            failures.append(f"serve {kind}: {r.error}")
                        elif kind == "d" and not np.array_equal(r.samples()[:want[k].data.size], want[k].data):
                            failures.append("serve decode samples")
                        elif kind == "e" and r.packet() != enc_want:
                            failures.append("serve encode bytes")
                        srv.release(r.id)

    threads = [threading.Thread(target=copier), threading.Thread(target=server)]
    try:
        for t in threads:
            t.start()
        dec = P.Decoder()
        for rep in range(12):
            print("case: graph batch beside copies and serving, rep", rep, flush=True)
            for dd, n in bufs:
                dd.upload(np.zeros(n, np.uint16))
            errs = dec.decode_device_batch(jobs)
            assert errs == [None] * 3, errs
            for (dd, n), w in zip(bufs, want):
                np.testing.assert_array_equal(dd.download(n * 2, dtype=np.uint16), w.data)
    finally:
        stop.set()
        for t in threads:
            t.join()
        P.set_decode_tuning(saved)
    assert not failures, failures
    after = P.graph_counters()
    print("graph counters (captures, failures, replays):", before, "->", after, flush=True)
    assert after[2] - before[2] >= 1, "no batch replayed a captured graph"
