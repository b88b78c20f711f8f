"""The C-ABI library on the CPU: it loads, exports every entry point
include/prores_hip.h declares, and its host-side logic (header syntax,
profile arithmetic, configuration and frame checks — nothing that launches a
kernel) agrees with the oracle.
"""

import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "prores_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prores_[a-z_0-9]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = P.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(P.EXPORTS) == names
    assert "gfx950" in P.version()


def test_profiles_and_targets():
    for p in P.Profile.ALL:
        assert P.Profile.from_fourcc(p.fourcc()) == p
        for w, h in [(1920, 1080), (16, 16), (1, 1), (3840, 2160), (720, 486), (4096, 2160), (65535, 3)]:
            assert p.target_frame_bytes(w, h) == O.target_frame_bytes(int(p), w, h)
    assert P.Profile.Hq.target_frame_bytes(1920, 1080) == 917_583
    assert P.Profile.from_fourcc(b"xxxx") is None
    assert [p.fourcc() for p in P.Profile.ALL] == [b"apco", b"apcs", b"apcn", b"apch", b"ap4h", b"ap4x"]


def test_frame_header_parse_matches_the_oracle():
    f = O.test_frame(64, 48, 3, 12, 4, 1)
    f.interlace = 1
    f.metadata = (2, 4, 6, 1, 6)
    luma = [2 + k % 62 for k in range(64)]
    pkt = O.encode(f, O.config(5, luma=luma))
    h = P.FrameHeader.parse(pkt)
    assert (h.frame_size, h.width, h.height, int(h.chroma), int(h.interlace)) == (len(pkt), 64, 48, 3, 1)
    assert h.metadata.astuple() == (2, 4, 6, 1, 6)
    assert h.luma_matrix == bytes(luma) and h.chroma_matrix is None
    assert h.encoder_identifier == b"rivt" and h.bitstream_version == 1
    assert (h.picture_height(0), h.picture_height(1)) == (24, 24)


def test_header_errors_are_the_references_without_a_device():
    """Header-level errors are found before any device work, with the
    reference's kinds and messages (src/header.rs:82-170, src/decode.rs:113-128)."""
    pkt = O.encode(O.test_frame(40, 24, 2, 10, 20, 1), O.config(2))
    cases = [pkt[:20], pkt[:-1]]
    for at, val in [(4, ord("x")), (8 + 12, 1 << 6), (8 + 12, (2 << 6) | (3 << 2)), (8 + 17, 3), (8 + 3, 2),
                    (8, 0), (9, 10), (8 + 8, 0), (8 + 19, 2), (0, 0)]:
        b = bytearray(pkt)
        b[at] = val
        cases.append(bytes(b))
    big = bytearray(pkt)
    big[8 + 8:8 + 12] = (65535).to_bytes(2, "big") * 2  # plausibility: far too small for 65535x65535
    cases.append(bytes(big))
    for data in cases:
        want = O.decode_result(data)
        if want[0] == "Ok":
            continue
        try:
            P.FrameHeader.parse(data)
            header_ok = True
        except P.Error as e:
            header_ok = False
            assert (e.kind, e.msg) == want
        if header_ok:
            # A plausibility error: the decoder reports it before touching a device.
            with pytest.raises(P.Error) as ei:
                P.Decoder().decode(data)
            assert (ei.value.kind, ei.value.msg) == want


def test_decoder_bit_depth_checks():
    for bd in (7, 17, 100):
        with pytest.raises(P.ConfigError) as ei:
            P.Decoder.with_bit_depth(bd)
        assert ei.value.msg == f"output bit depth {bd} is outside 8–16"
    for bd in (8, 12, 16):
        P.Decoder.with_bit_depth(bd)


def _oracle_config_error(frame, cfg):
    try:
        O.encode(frame, cfg)
    except O.OracleError as e:
        return e.as_tuple()
    return None


def test_encoder_refuses_what_it_cannot_code():
    """tests/roundtrip.rs:467-515, checked on the host before any device use."""
    f422 = O.Frame(32, 32, 2, 10)
    f444 = O.Frame(32, 32, 3, 12)
    with_alpha = O.Frame(32, 32, 2, 10, alpha=np.zeros(32 * 32, np.uint16))
    short_alpha = O.Frame(32, 32, 3, 12, alpha=np.zeros(10, np.uint16))
    truncated = O.Frame(32, 32, 2, 10)
    truncated.data = truncated.data[:100]
    cases = [(f444, dict(profile=3)), (f422, dict(profile=4)), (with_alpha, dict(profile=3)),
             (short_alpha, dict(profile=4)), (f422, dict(profile=3, luma=[1] * 64)),
             (f422, dict(profile=3, chroma=[64] * 64)), (f422, dict(profile=3, log2=4)), (truncated, dict(profile=3))]
    for frame, kw in cases:
        oc = O.config(kw["profile"], log2_slice_mbs=kw.get("log2", 3), luma=kw.get("luma"), chroma=kw.get("chroma"))
        want = _oracle_config_error(frame, oc)
        assert want is not None and want[0] == "Config"
        pf = P.Frame(frame.width, frame.height, frame.chroma, frame.bit_depth,
                     None, frame.alpha, frame.interlace)
        pf.data = frame.data  # keep a truncated buffer truncated
        pc = P.Config(kw["profile"], log2_slice_mbs=kw.get("log2", 3), luma_matrix=kw.get("luma"),
                      chroma_matrix=kw.get("chroma"))
        with pytest.raises(P.ConfigError) as ei:
            P.Encoder(pc).encode(pf)
        assert (ei.value.kind, ei.value.msg) == want


def test_frame_layout():  # src/frame.rs:361-368
    f = P.Frame.new(15, 3, P.ChromaFormat.Yuv422, 10)
    assert f.planes[1][1] == 8 and f.planes[2][0] == 45 + 24 and f.data.size == 45 + 48
    planes = (P._Plane * 3)()
    assert P.lib().prores_frame_layout(15, 3, 2, planes) == 93
    assert [(p.offset, p.width, p.height) for p in planes] == [(0, 15, 3), (45, 8, 3), (69, 8, 3)]
    with pytest.raises(P.ConfigError):
        P.Frame.new(0, 1, P.ChromaFormat.Yuv422, 10)
    with pytest.raises(P.ConfigError):
        P.Frame.new(1, 1, P.ChromaFormat.Yuv422, 7)


def test_metadata_frame_rates():  # src/frame.rs:370-382
    for c in range(1, 12):
        n, d = P.Metadata(frame_rate_code=c).frame_rate()
        assert P.Metadata.frame_rate_code_for(n, d) == c
    assert P.Metadata.frame_rate_code_for(48, 2) == 2
    assert P.Metadata.frame_rate_code_for(48, 1) == 0


def test_server_config_is_checked_before_the_device():
    """A serving loop with a budget under 1 MiB or no batch room is refused
    with a Config error before any device work."""
    with pytest.raises(P.ConfigError):
        P.Server(hbm_budget=1000)
    with pytest.raises(P.ConfigError):
        P.Server(max_batch_frames=0)


def test_find_frames_in_a_mov_like_file():
    """prores_find_frames (tests/sample.rs:26-41): packets embedded in
    container bytes are found at their offsets, whole; 'icpf' with a
    frame_size under 28 or past the data is passed over; a truncated last
    frame is left out."""
    rng = np.random.default_rng(3)
    pk = [O.encode(O.test_frame(w, h, 2, 10, 8, s), O.config(p, log2_slice_mbs=1))
          for w, h, s, p in ((48, 32, 1, 3), (64, 16, 2, 0), (32, 48, 3, 4 - 2))]
    buf, want = bytearray(), []
    buf += rng.integers(0, 256, 37, dtype=np.uint8).tobytes()
    buf += b"\x00\x00\x00\x10icpf" + b"\x01" * 8          # frame_size 16 < 28
    buf += b"\x7f\xff\xff\xffgarbage icpf"                  # 'icpf' without a plausible size
    for p in pk:
        buf += rng.integers(0, 256, int(rng.integers(0, 50)), dtype=np.uint8).tobytes()
        want.append((len(buf), len(p)))
        buf += p
    buf += b"\x00\x10\x00\x00icpf" + b"\x00" * 40           # runs past the end
    got = P.find_frames(bytes(buf))
    assert got == want
    for (o, n), p in zip(got, pk):
        assert bytes(buf[o:o + n]) == p
    assert P.find_frames(bytes(buf)[:want[-1][0] + want[-1][1] - 1]) == want[:-1]
    assert P.find_frames(b"") == [] and P.find_frames(b"icpf") == []


def test_fnv1a_known_values():
    # FNV-1a 64 published test vectors (the empty string, "a", "foobar").
    base = 0xCBF29CE484222325
    assert P.fnv1a64(base, b"") == 0xCBF29CE484222325
    assert P.fnv1a64(base, b"a") == 0xAF63DC4C8601EC8C
    assert P.fnv1a64(base, b"foobar") == 0x85944171F73967E8
