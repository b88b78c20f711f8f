"""GPU parity: the HIP decoder and encoder against the CPU oracle, bit for bit.

Decoded samples must equal the oracle's on the same bitstream; encoded frames
must be byte-identical to the oracle's on the same input and Config. The cases
follow the reference's round-trip tests (tests/roundtrip.rs) at sizes the
oracle finishes in seconds.
"""

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu


def to_p(frame):
    """oracle_lib.Frame -> prores_amd.Frame (same layout)."""
    return P.Frame(frame.width, frame.height, frame.chroma, frame.bit_depth, frame.data.copy(),
                   None if frame.alpha is None else frame.alpha.copy(), frame.interlace,
                   P.Metadata(*frame.metadata))


def cfg_pair(profile, **kw):
    oc = O.config(profile, log2_slice_mbs=kw.get("log2", 3), alpha=kw.get("alpha", 2),
                  luma=kw.get("luma"), chroma=kw.get("chroma"), target=kw.get("target", 0))
    pc = P.Config(profile, log2_slice_mbs=kw.get("log2", 3), alpha=kw.get("alpha", 2),
                  luma_matrix=kw.get("luma"), chroma_matrix=kw.get("chroma"),
                  target_frame_bytes=kw.get("target") or None)
    return oc, pc


def assert_decode_parity(packet, bit_depth=0):
    want = O.decode_result(packet, bit_depth)
    try:
        got = ("Ok", P.Decoder(bit_depth).decode(packet))
    except P.Error as e:
        got = (e.kind, e.msg)
    assert got[0] == want[0], (got[:2] if got[0] != "Ok" else "Ok", want if want[0] != "Ok" else "Ok")
    if want[0] != "Ok":
        assert got == want
        return
    g, w = got[1], want[1]
    assert (g.width, g.height, g.bit_depth, int(g.chroma), int(g.interlace)) == \
        (w.width, w.height, w.bit_depth, w.chroma, w.interlace)
    assert g.metadata.astuple() == tuple(w.metadata)
    np.testing.assert_array_equal(g.data, w.data)
    if w.alpha is None:
        assert g.alpha is None
    else:
        np.testing.assert_array_equal(g.alpha, w.alpha)


def assert_encode_parity(frame, profile, **kw):
    oc, pc = cfg_pair(profile, **kw)
    want = O.encode(frame, oc)
    got = P.Encoder(pc).encode(to_p(frame))
    if got != want:
        n = min(len(got), len(want))
        diff = next((i for i in range(n) if got[i] != want[i]), n)
        raise AssertionError(f"encoded bytes differ: len {len(got)} vs {len(want)}, first difference at {diff}")
    return want


def print_case(*a):
    print("case:", *a, flush=True)


@pytest.mark.parametrize("profile", range(6))
def test_decode_1080p_every_profile(profile):
    chroma, bd = (3, 12) if profile >= 4 else (2, 10)
    print_case("decode 1080p", profile)
    f = O.test_frame(1920, 1080, chroma, bd, 24, 7)
    pkt = O.encode(f, O.config(profile))
    assert_decode_parity(pkt)


@pytest.mark.parametrize("profile", range(6))
def test_encode_1080p_every_profile(profile):
    chroma, bd = (3, 12) if profile >= 4 else (2, 10)
    print_case("encode 1080p", profile)
    f = O.test_frame(1920, 1080, chroma, bd, 24, 7)
    pkt = assert_encode_parity(f, profile)
    # README pins (README.md:173-180) hold for the GPU bytes too.
    assert len(pkt) == [187677, 425420, 613107, 917564, 1376337, 2085373][profile]


SIZES = [(1, 1), (2, 2), (17, 9), (31, 33), (48, 16), (100, 37), (129, 75), (250, 19)]


@pytest.mark.parametrize("w,h", SIZES)
def test_odd_sizes_slices_fields(w, h):
    """tests/roundtrip.rs:139-187 at unconstrained targets, and the same with
    the profile's target, both directions."""
    for chroma, profile in ((2, 3), (3, 4)):
        bd = 12 if chroma == 3 else 10
        for interlace in (0, 1, 2):
            for log2 in range(4):
                print_case("odd", w, h, chroma, interlace, log2)
                f = O.test_frame(w, h, chroma, bd, 6, w * 131 + h)
                f.interlace = interlace
                for target in (1 << 30, 0):
                    pkt = assert_encode_parity(f, profile, log2=log2, target=target)
                    assert_decode_parity(pkt)


def test_alpha_lossless_and_depths():
    """tests/roundtrip.rs:228-276 alpha cases, decoded at several depths."""
    rng = np.random.default_rng(99)
    for depth, atype in ((12, 2), (16, 2), (8, 1), (10, 2)):
        for interlace in (0, 1):
            print_case("alpha", depth, atype, interlace)
            f = O.test_frame(97, 41, 3, depth, 4, 5)
            f.interlace = interlace
            mx = (1 << depth) - 1
            a = np.zeros((41, 97), np.int64)
            xs = np.arange(97)
            a[:, :20] = 0
            a[:, 20:40] = mx
            a[:, 40:60] = (xs[40:60] * mx // 97)
            a[:, 60:80] = rng.integers(0, mx + 1, size=(41, 20))
            a[:, 80:] = mx // 2
            f.alpha = a.astype(np.uint16).ravel()
            pkt = assert_encode_parity(f, 4, alpha=atype)
            for bd in (0, 8, depth, 16):
                assert_decode_parity(pkt, bd)


def test_custom_matrices():
    """tests/roundtrip.rs:379-410."""
    luma = [4 + 2 * (u + v) for v in range(8) for u in range(8)]
    chroma = [6 + 3 * (u + v) for v in range(8) for u in range(8)]
    f = O.test_frame(320, 180, 2, 10, 8, 21)
    for l, c in ((luma, chroma), (luma, None), (None, chroma)):
        print_case("matrices", l is not None, c is not None)
        pkt = assert_encode_parity(f, 3, luma=l, chroma=c)
        assert_decode_parity(pkt)


def test_flat_and_extremes():
    """tests/roundtrip.rs:519-545."""
    for profile in range(6):
        chroma, bd = (3, 12) if profile >= 4 else (2, 10)
        f = O.Frame(640, 360, chroma, bd)
        pkt = assert_encode_parity(f, profile)
        assert_decode_parity(pkt)
    f = O.Frame(64, 32, 2, 10)
    for i in range(3):
        p = f.plane(i)
        p[:] = np.where(((np.arange(p.shape[1]) // 3) % 2) == 0, 0, 1023)[None, :]
    pkt = assert_encode_parity(f, 0)
    assert_decode_parity(pkt)


def test_damaged_frames_give_the_same_error():
    """tests/roundtrip.rs:584-610: the first error in picture order."""
    f = O.test_frame(333, 75, 2, 10, 20, 3)
    pkt = bytearray(O.encode(f, O.config(3)))
    n = len(pkt)
    for i in range(n // 2, n // 2 + 64):
        pkt[i] ^= 0x5A
    assert_decode_parity(bytes(pkt))
    rng = np.random.default_rng(5)
    base = O.encode(O.test_frame(40, 24, 2, 10, 20, 1), O.config(2))
    for t in range(200):
        b = bytearray(base)
        for _ in range(rng.integers(1, 8)):
            b[rng.integers(0, len(b))] = rng.integers(0, 256)
        assert_decode_parity(bytes(b))


def _component_bytes(scanned):
    import ctypes as C
    arr = np.ascontiguousarray(scanned, dtype=np.int32)
    cap = 64 * arr.size + 64
    buf = (C.c_uint8 * cap)()
    n = O.lib().oracle_encode_coefficients(arr.ctypes.data, arr.size // 64, buf, cap)
    assert 0 < n <= cap
    return bytes(buf[:n])


def test_coefficients_beyond_16_bits():
    """Coefficients outside +-32767 (a DC of 40000, levels up to 2^29) decode
    like the oracle: the engine keeps them beside its 16-bit coefficient
    store. Streams are coded by the oracle's coefficient coder
    (src/encode.rs:413-437) into one-slice 32x16 4:2:2 frames."""
    from test_handbuilt import frame, frame_header, picture, slice_

    rng = np.random.default_rng(8)
    specials = np.array([32767, -32767, 32768, -32768, 40000, -40000, 1 << 20, -(1 << 20), (1 << 29) + 5,
                         -(1 << 29)], dtype=np.int64)
    for case in range(12):
        comps = []
        for nb in (8, 4, 4):  # two macroblocks: Y' 8 blocks, Cb and Cr 4 each
            c = rng.integers(-40, 41, size=nb * 64).astype(np.int64)
            c[rng.random(nb * 64) < 0.5] = 0
            k = rng.integers(1, 6)
            c[rng.integers(0, nb * 64, size=k)] = rng.choice(specials, size=k)
            if case == 0:
                c[:nb] = 40000  # every DC
            comps.append(_component_bytes(c))
        q = int(rng.integers(1, 225))
        data = frame(frame_header(32, 16, 2, 0, 0, 0, 0), [picture(1, [slice_(q, *comps, 0)])], 0, 0)
        print_case("wide", case, q)
        assert_decode_parity(data)
        assert_decode_parity(data, 16)


def test_many_long_alpha_streams():
    """Random alpha coded at 16 bits, 1-macroblock slices: 4,608 of the
    5,120 slices carry long alpha streams (over 256 bytes, decoded first in
    the alpha work order) beside short ones; samples and alpha equal the
    oracle's (src/decode.rs:365-389, 570-595)."""
    rng = np.random.default_rng(17)
    f = O.test_frame(2048, 640, 3, 12, 20, 5)
    f.alpha = rng.integers(0, 4096, size=2048 * 640).astype(np.uint16)
    f.alpha[: 2048 * 64] = np.repeat(np.arange(0, 2048 * 64 // 16) % 4096, 16).astype(np.uint16)
    print_case("long alpha streams", 2048, 640)
    pkt = assert_encode_parity(f, 5, log2=0)
    assert_decode_parity(pkt)
    assert_decode_parity(pkt, 10)


def test_host_batch_equals_the_oracle():
    """prores_decode_host_batch: many host packets (profiles, sizes, fields,
    alpha, damaged ones) decoded in overlapped chunks give each packet's
    oracle Result, in order; frames own their planes."""
    cases = []
    specs = [(1920, 1080, 2, 10, 3, 3, 0, 0), (333, 75, 3, 12, 5, 1, 1, 2), (720, 486, 2, 10, 0, 2, 1, 0),
             (512, 288, 3, 12, 4, 0, 0, 1), (3840, 2160, 3, 12, 5, 3, 0, 2)]
    for i, (w, h, ch, bd, prof, log2, il, atype) in enumerate(specs):
        f = O.test_frame(w, h, ch, bd, 10 + i, 50 + i)
        f.interlace = il
        if atype:
            f.alpha = ((np.arange(w * h) * (7 + i)) % (1 << bd)).astype(np.uint16)
        cases.append(O.encode(f, O.config(prof, log2_slice_mbs=log2, alpha=atype or 2)))
    bad = bytearray(cases[0])
    bad[len(bad) // 2:len(bad) // 2 + 40] = b"\x5a" * 40
    packets = cases + [bytes(bad), b"\x00\x00\x00\x10icpf" + b"\x00" * 8] + cases[:2]
    print_case("host batch", len(packets))
    got = P.Decoder().decode_batch(packets)
    assert len(got) == len(packets)
    for k, (pkt, g) in enumerate(zip(packets, got)):
        want = O.decode_result(pkt)
        if want[0] != "Ok":
            assert isinstance(g, P.Error) and (g.kind, g.msg) == tuple(want), (k, g, want)
            continue
        w = want[1]
        assert not isinstance(g, Exception), (k, g)
        np.testing.assert_array_equal(g.data, w.data, err_msg=f"packet {k}")
        if w.alpha is None:
            assert g.alpha is None
        else:
            np.testing.assert_array_equal(g.alpha, w.alpha, err_msg=f"packet {k} alpha")


def test_decode_file_hash_equals_the_oracles():
    """decode_file (examples/prores_bench.rs:104-122): every frame of a
    MOV-like file decoded on the GPU; the count and the FNV-1a of the
    pictures equal the oracle's."""
    rng = np.random.default_rng(9)
    buf, pk = bytearray(b"\x00\x00\x00\x20ftypqt  " + b"\x00" * 20), []
    for i, (w, h, ch, bd, prof) in enumerate([(720, 486, 2, 10, 3), (1280, 720, 2, 10, 1), (256, 144, 3, 12, 4)]):
        p = O.encode(O.test_frame(w, h, ch, bd, 12, 70 + i), O.config(prof))
        buf += rng.integers(0, 256, 24, dtype=np.uint8).tobytes() + p
        pk.append(p)
    h = 0xCBF29CE484222325
    for p in pk:
        h = P.fnv1a64(h, O.decode(p).data.astype("<u2"))
    print_case("decode file", len(pk))
    assert P.decode_file(bytes(buf)) == (len(pk), h)
