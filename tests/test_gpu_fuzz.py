"""Malformed input on the GPU decoder: the property groups of tests/fuzz.rs
(arbitrary bytes, arbitrary bytes behind a frame identifier, valid frames
with bytes changed, bits flipped, cut and spliced) restated with seeded
generators. The reference demands an Ok or an Err, never a panic or a hang;
the drop-in demands more: the GPU's Result equals the oracle's, message
included, and an Ok frame has the oracle's samples. The same cases decoded
as one device batch must give each frame its own Result. Random frames round
trip (tests/fuzz.rs:101-130) with the oracle's bytes and samples.
"""

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu


def seeds():
    """The three valid frames of tests/fuzz.rs:12-38."""
    out = []
    f = O.test_frame(40, 24, 2, 10, 20, 1)
    out.append(O.encode(f, O.config(2)))
    f = O.test_frame(33, 21, 3, 12, 20, 2)
    f.interlace = 1
    f.alpha = (np.arange(33 * 21) * 7 % 4096).astype(np.uint16)
    out.append(O.encode(f, O.config(5, luma=[5] * 64, chroma=[9] * 64)))
    f.interlace = 0
    out.append(O.encode(f, O.config(4, alpha=1, log2_slice_mbs=0)))
    return out


def cases(seed=31):
    rng = np.random.default_rng(seed)
    S = seeds()
    out = []
    for _ in range(300):  # arbitrary_bytes
        out.append(rng.integers(0, 256, int(rng.integers(0, 400)), dtype=np.uint8).tobytes())
    for _ in range(300):  # arbitrary_bytes_after_a_frame_identifier
        d = bytearray(rng.integers(0, 256, int(rng.integers(28, 2000)), dtype=np.uint8).tobytes())
        d[0:4] = len(d).to_bytes(4, "big")
        d[4:8] = b"icpf"
        out.append(bytes(d))
    for _ in range(600):  # valid_frames_with_bytes_changed
        d = bytearray(S[int(rng.integers(0, 3))])
        for _ in range(int(rng.integers(1, 12))):
            d[int(rng.integers(0, len(d)))] = int(rng.integers(0, 256))
        out.append(bytes(d))
    for _ in range(600):  # valid_frames_with_bits_flipped
        d = bytearray(S[int(rng.integers(0, 3))])
        for _ in range(int(rng.integers(1, 8))):
            bit = int(rng.integers(0, len(d) * 8))
            d[bit // 8] ^= 0x80 >> (bit % 8)
        out.append(bytes(d))
    for _ in range(200):  # valid_frames_cut_and_spliced
        d = S[int(rng.integers(0, 3))]
        at = int(rng.integers(0, len(d)))
        out.append(d[:at])
        junk = rng.integers(0, 256, int(rng.integers(0, 64)), dtype=np.uint8).tobytes()
        out.append(d[:at] + junk + d[at:])
    return out


def gpu_result(dec, data):
    try:
        return ("Ok", dec.decode(data))
    except P.Error as e:
        return (e.kind, e.msg)


def same(got, want):
    if got[0] != want[0]:
        return False
    if want[0] != "Ok":
        return got[1] == want[1]
    g, w = got[1], want[1]
    return (np.array_equal(g.data, w.data) and
            ((w.alpha is None and g.alpha is None) or
             (w.alpha is not None and g.alpha is not None and np.array_equal(g.alpha, w.alpha))))


def test_malformed_input_gives_the_oracles_result():
    dec = P.Decoder()
    kinds = {}
    for i, data in enumerate(cases()):
        want = O.decode_result(data)
        got = gpu_result(dec, data)
        assert same(got, want), (i, len(data), got[:2] if got[0] != "Ok" else "Ok",
                                 want[:2] if want[0] != "Ok" else "Ok")
        kinds[want[0]] = kinds.get(want[0], 0) + 1
    print("results by kind:", kinds, flush=True)
    assert kinds.get("Ok", 0) > 0 and kinds.get("Invalid", 0) > 0


def test_malformed_frames_in_one_batch():
    """Frames that need device work (a header that parses), decoded together:
    each gets its own Result, the oracle's."""
    dec = P.Decoder()
    picked = []
    for data in cases(seed=47):
        try:
            h = P.FrameHeader.parse(data)
        except P.Error:
            continue
        if h.width * h.height > 1 << 20:  # output buffers are sized from the header
            continue
        picked.append((data, h))
        if len(picked) == 96:
            break
    bufs, jobs = [], []
    for data, h in picked:
        n = sum(w * hh for _, w, hh in P.layout(h.width, h.height, h.chroma))
        dp = P.DeviceBuffer(len(data) + 128)
        dp.upload(np.frombuffer(data, np.uint8))
        dd = P.DeviceBuffer(n * 2 + 64)
        da = P.DeviceBuffer(h.width * h.height * 2 + 64) if h.alpha else None
        bufs.append((dp, dd, da, n, h))
        jobs.append((data, dp.ptr, dd.ptr, da.ptr if da else 0))
    for tuning in (P.DecodeTuning(), P.DecodeTuning(1, 2, 32, 64, 1)):
        print("case: malformed batch", tuning, flush=True)
        errs = dec.decode_device_batch(jobs, tuning)
        for i, ((data, h), e, (dp, dd, da, n, _)) in enumerate(zip(picked, errs, bufs)):
            want = O.decode_result(data)
            if want[0] != "Ok":
                assert e is not None and (e.kind, e.msg) == (want[0], want[1]), (i, e, want[:2])
                continue
            assert e is None, (i, e)
            np.testing.assert_array_equal(dd.download(n * 2, dtype=np.uint16), want[1].data)
            if want[1].alpha is not None:
                np.testing.assert_array_equal(da.download(h.width * h.height * 2, dtype=np.uint16), want[1].alpha)


def test_random_frames_round_trip():
    """tests/fuzz.rs:101-130 at 48 cases: any size, format, field order and
    slice size; the GPU's packet and samples are the oracle's."""
    rng = np.random.default_rng(5)
    dec = P.Decoder()
    for case in range(48):
        w, h = int(rng.integers(1, 90)), int(rng.integers(1, 70))
        is444 = bool(rng.integers(0, 2))
        il, log2 = int(rng.integers(0, 3)), int(rng.integers(0, 4))
        grain, seed = int(rng.integers(0, 40)), int(rng.integers(0, 1 << 31))
        chroma, profile, depth = (3, 4, 12) if is444 else (2, 3, 10)
        f = O.test_frame(w, h, chroma, depth, grain, seed)
        f.interlace = il
        print("case: round trip", case, w, h, is444, il, log2, flush=True)
        want = O.encode(f, O.config(profile, log2_slice_mbs=log2, target=1 << 30))
        pf = P.Frame(f.width, f.height, f.chroma, f.bit_depth, f.data.copy(), None, f.interlace,
                     P.Metadata(*f.metadata))
        got = P.Encoder(P.Config(profile, log2_slice_mbs=log2, target_frame_bytes=1 << 30)).encode(pf)
        assert got == want
        d = dec.decode(got)
        np.testing.assert_array_equal(d.data, O.decode(want).data)


def test_malformed_packets_planned_on_the_device():
    """Every fuzz case as a job whose packet is only in device memory: the
    frame and picture headers are parsed and the batch planned on the device
    (pd_plan), 128 jobs a batch. Each job's Result must be the oracle's,
    message included; a job whose header the host cannot parse gets no output
    planes, so a header the device wrongly accepted would fail as a missing
    plane, not write anywhere."""
    dec = P.Decoder()
    items = []
    for data in cases(seed=31) + cases(seed=47)[:600]:
        try:
            h = P.FrameHeader.parse(data)
        except P.Error:
            h = None
        if h is not None and h.width * h.height > 1 << 20:
            continue
        items.append((data, h))
    kinds = {}
    for b0 in range(0, len(items), 128):
        part = items[b0:b0 + 128]
        bufs, jobs = [], []
        for data, h in part:
            dp = P.DeviceBuffer(len(data) + 128)
            if data:
                dp.upload(np.frombuffer(data, np.uint8))
            dd = da = None
            n = 0
            if h is not None:
                n = sum(w * hh for _, w, hh in P.layout(h.width, h.height, h.chroma))
                dd = P.DeviceBuffer(n * 2 + 64)
                da = P.DeviceBuffer(h.width * h.height * 2 + 64) if h.alpha else None
            bufs.append((dp, dd, da, n, h))
            jobs.append((len(data), dp.ptr, dd.ptr if dd else 0, da.ptr if da else 0))
        print("case: device-planned fuzz batch from", b0, flush=True)
        errs = dec.decode_device_batch(jobs)
        for i, ((data, h), e, (dp, dd, da, n, _)) in enumerate(zip(part, errs, bufs)):
            want = O.decode_result(data)
            kinds[want[0]] = kinds.get(want[0], 0) + 1
            if want[0] != "Ok":
                assert e is not None and (e.kind, e.msg) == (want[0], want[1]), (b0 + i, len(data), e, want[:2])
                continue
            assert e is None, (b0 + i, e)
            np.testing.assert_array_equal(dd.download(n * 2, dtype=np.uint16), want[1].data)
            if want[1].alpha is not None:
                np.testing.assert_array_equal(da.download(h.width * h.height * 2, dtype=np.uint16), want[1].alpha)
    print("results by kind:", kinds, flush=True)
    assert kinds.get("Ok", 0) > 0 and kinds.get("Invalid", 0) > 0 and kinds.get("Unsupported", 0) > 0
