"""The allocation-free entry point (prores_decode_device_batch_noalloc)
against prores_decode_device_batch, through the C ABI: the same jobs give the
same samples and the same errors, on two calls in a row, the second with a
smaller batch; both equal the oracle's decode. A batch whose workspace does
not fit the scratch array fails with the documented configuration error.
"""

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu


def _packets():
    out = []
    specs = [(1920, 1080, 2, 10, 3, 3, 0, 0), (720, 486, 2, 10, 3, 1, 1, 0), (512, 288, 3, 12, 4, 2, 0, 1),
             (333, 75, 3, 12, 5, 3, 2, 2), (1280, 720, 2, 10, 0, 0, 0, 0)]
    for i, (w, h, ch, bd, prof, log2, il, atype) in enumerate(specs):
        f = O.test_frame(w, h, ch, bd, 10 + 3 * i, 900 + i)
        f.interlace = il
        if atype:
            mx = (1 << bd) - 1
            yy, xx = np.mgrid[0:h, 0:w]
            d = np.hypot((xx - w * 0.5) / (0.3 * w), (yy - h * 0.5) / (0.3 * h))
            f.alpha = (np.clip((1.0 - d) * 4.0, 0, 1) * mx).round().astype(np.uint16).ravel()
        out.append(O.encode(f, O.config(prof, log2_slice_mbs=log2, alpha=atype or 2)))
    # A damaged copy: a slice-size table entry made huge.
    bad = bytearray(out[4])
    pic = 8 + int.from_bytes(bad[8:10], "big")
    phs = bad[pic] >> 3
    bad[pic + phs + 6], bad[pic + phs + 7] = 0xFF, 0xF0
    out.append(bytes(bad))
    return out


def _buffers(packets):
    bufs = []
    for p in packets:
        h = P.FrameHeader.parse(p)
        planes = P.layout(h.width, h.height, h.chroma)
        n = planes[2][0] + planes[2][1] * planes[2][2]
        dp = P.DeviceBuffer(len(p) + 128)
        dp.upload(np.frombuffer(p, np.uint8))
        bufs.append((dp, n, h))
    return bufs


def _decode(dec, packets, bufs, noalloc):
    outs, jobs = [], []
    for p, (dp, n, h) in zip(packets, bufs):
        dd = P.DeviceBuffer(n * 2 + 64)
        dd.upload(np.full(n, 0xABCD, np.uint16))
        da = P.DeviceBuffer(h.width * h.height * 2 + 64) if h.alpha else None
        outs.append((dd, da))
        jobs.append((len(p), dp.ptr, dd.ptr, da.ptr if da else 0))
    errs = dec.decode_device_batch_noalloc(jobs) if noalloc else dec.decode_device_batch(jobs)
    got = []
    for (dd, da), (dp, n, h), e in zip(outs, bufs, errs):
        if e is not None:
            got.append((e.kind, e.msg))
            continue
        a = da.download(h.width * h.height * 2, dtype=np.uint16) if da else None
        got.append(("Ok", dd.download(n * 2, dtype=np.uint16), a))
    return got


def test_noalloc_equals_the_allocating_entry_point():
    assert P.noalloc_scratch_bytes() == 256 << 20
    P.noalloc_init()
    packets = _packets()
    bufs = _buffers(packets)
    dec = P.Decoder()
    want = [O.decode_result(p) for p in packets]
    # Two calls in a row, the second with a smaller batch.
    for sel in (list(range(len(packets))), [1, 3, 5]):
        ps, bs = [packets[i] for i in sel], [bufs[i] for i in sel]
        print("case: noalloc batch of", len(sel), "frames", flush=True)
        got = _decode(dec, ps, bs, True)
        ref = _decode(dec, ps, bs, False)
        for k, (g, r) in enumerate(zip(got, ref)):
            w = want[sel[k]]
            assert g[0] == r[0] == w[0], (sel[k], g[:2], r[:2], w[:2])
            if w[0] != "Ok":
                assert g[1] == r[1] == w[1]
                continue
            np.testing.assert_array_equal(g[1], r[1])
            np.testing.assert_array_equal(g[1], w[1].data)
            if w[1].alpha is not None:
                np.testing.assert_array_equal(g[2], r[2])
                np.testing.assert_array_equal(g[2], w[1].alpha)


def test_noalloc_refuses_a_batch_over_the_scratch():
    P.noalloc_init()
    f = O.test_frame(3840, 2160, 3, 12, 10, 5)
    packet = O.encode(f, O.config(5))
    bufs = _buffers([packet] * 2)
    dec = P.Decoder()
    print("case: noalloc batch over the scratch", flush=True)
    got = _decode(dec, [packet] * 2, bufs, True)
    assert got == [("Config", "the batch needs more than the 268435456 bytes of its scratch memory")] * 2
    # The same device afterwards decodes a batch that fits.
    small = _packets()[:2]
    sb = _buffers(small)
    for g, p in zip(_decode(dec, small, sb, True), small):
        assert g[0] == "Ok"
        np.testing.assert_array_equal(g[1], O.decode(p).data)
