"""Frames assembled bit by bit from RDD 36's syntax and codebooks, with the
samples §7 says they decode to worked out by hand — the vectors of the
reference's tests/handbuilt.rs:91-317. The CPU oracle and the GPU decoder are
each checked against the hand-computed samples (not against each other).
"""

import math

import numpy as np
import pytest

import oracle_lib as O


def bits(s):
    """Bytes from '0'/'1' characters (spaces ignored), zero-padded."""
    s = s.replace(" ", "")
    s += "0" * (-len(s) % 8)
    return bytes(int(s[i:i + 8], 2) for i in range(0, len(s), 8))


def frame_header(w, h, chroma, interlace, alpha, version, extra):
    v = bytearray((20 + extra).to_bytes(2, "big"))
    v += bytes([0, version]) + b"hand" + w.to_bytes(2, "big") + h.to_bytes(2, "big")
    v += bytes([(chroma << 6) | (interlace << 2), (1 << 4) | 3, 1, 1, 1, alpha, 0, 0])
    v += bytes([0xAA] * extra)
    return bytes(v)


def slice_(q, y, cb, cr, extra):
    v = bytearray([((6 + extra) << 3), q]) + len(y).to_bytes(2, "big") + len(cb).to_bytes(2, "big")
    v += bytes([0x55] * extra) + y + cb + cr
    return bytes(v)


def picture(log2, slices):
    size = 8 + 2 * len(slices) + sum(len(s) for s in slices)
    v = bytearray([8 << 3]) + size.to_bytes(4, "big") + len(slices).to_bytes(2, "big") + bytes([log2 << 4])
    for s in slices:
        v += len(s).to_bytes(2, "big")
    for s in slices:
        v += s
    return bytes(v)


def frame(header, pictures, stuffing, trailing):
    size = 8 + len(header) + sum(len(p) for p in pictures) + stuffing
    return size.to_bytes(4, "big") + b"icpf" + header + b"".join(pictures) + bytes(stuffing) + b"\xff" * trailing


def one_macroblock():
    """Y' DCs 0, 64, -64, 128 and one AC +100 at scanned index 5; Cb flat;
    Cr DCs -32, -32 (tests/handbuilt.rs:91-116)."""
    y = bits("100000 000010001000 00000100000111 00000110000111 010 0000000 1100011 0")
    cb = bits("100000 1000")
    cr = bits("01011111 1000") + b"\x00"
    return y, cb, cr


def sample(v):
    return int(math.floor(2.0 * v + 512.0 + 0.5))


def hand_macroblock_y():
    want = np.zeros((16, 16), np.int64)
    for row in range(16):
        for col in range(16):
            q = (col // 8, row // 8)
            if q == (0, 0):
                want[row, col] = sample(0.0)
            elif q == (1, 0):
                x = col - 8
                ac = 50.0 / (4.0 * math.sqrt(2)) * math.cos((2 * x + 1) * math.pi / 16)
                want[row, col] = sample(4.0 + ac)
            elif q == (0, 1):
                want[row, col] = sample(-4.0)
            else:
                want[row, col] = sample(8.0)
    return want


def decoders():
    """(name, decode(data, bit_depth) -> (kind, frame-or-msg))."""
    out = [("oracle", lambda d, bd=0: O.decode_result(d, bd))]
    return out


def gpu_decode(d, bd=0):
    import prores_amd as P
    try:
        f = P.Decoder(bd).decode(d)
        return ("Ok", f)
    except P.Error as e:
        return (e.kind, e.msg)


def check_macroblock(dec):
    y, cb, cr = one_macroblock()
    data = frame(frame_header(16, 16, 2, 0, 0, 0, 4), [picture(0, [slice_(1, y, cb, cr, 1)])], 5, 7)
    kind, f = dec(data)
    assert kind == "Ok"
    assert f.bit_depth == 10
    np.testing.assert_array_equal(np.asarray(f.data[:256]).reshape(16, 16), hand_macroblock_y())
    assert set(np.asarray(f.data[256:384]).tolist()) == {512}
    assert set(np.asarray(f.data[384:512]).tolist()) == {sample(-2.0)}


def check_fields(dec):
    y_hi = bits("0010100000 1000 1 1")
    y_lo = bits("0010011111 1000 1 1")
    flat = bits("100000 1000")
    for mode, first_row in ((1, 0), (2, 1)):
        first = picture(0, [slice_(1, y_hi, flat, flat, 0)])
        second = picture(0, [slice_(1, y_lo, flat, flat, 0)])
        data = frame(frame_header(16, 2, 2, mode, 0, 0, 0), [first, second], 0, 0)
        kind, f = dec(data)
        assert kind == "Ok"
        assert int(f.interlace) == mode
        y = np.asarray(f.data[:32]).reshape(2, 16)
        assert set(y[first_row].tolist()) == {sample(4.0)}
        assert set(y[1 - first_row].tolist()) == {sample(-4.0)}


def check_alpha(dec):
    flat4 = bits("100000 1000 1 1")
    alpha = bits("1 00000000 01111  1 00000001 00000 00011101110  1 00001010 1")
    s = bytearray([8 << 3, 1]) + len(flat4).to_bytes(2, "big") * 3 + flat4 * 3 + alpha
    data = frame(frame_header(16, 16, 3, 0, 1, 1, 0), [picture(3, [bytes(s)])], 0, 0)
    kind, f = dec(data, 8)
    assert kind == "Ok"
    a = np.asarray(f.alpha)
    assert set(a[:16].tolist()) == {255}
    assert set(a[16:255].tolist()) == {0}
    assert a[255] == 10
    assert set(np.asarray(f.data).tolist()) == {128}


def check_malformed(dec):
    y, cb, cr = one_macroblock()
    good = frame(frame_header(16, 16, 2, 0, 0, 0, 0), [picture(0, [slice_(1, y, cb, cr, 0)])], 0, 0)
    assert dec(good)[0] == "Ok"

    def kind_of(d):
        return dec(bytes(d))[0]
    assert kind_of(good[:20]) == "Invalid"
    assert kind_of(good[:-1]) == "Invalid"
    for at, val, want in [(4, ord("x"), "Invalid"), (8 + 12, 1 << 6, "Invalid"),
                          (8 + 12, (2 << 6) | (3 << 2), "Invalid"), (8 + 17, 3, "Invalid"),
                          (8 + 3, 2, "Unsupported"), (8 + 20 + 8 + 2 + 1, 0, "Invalid"),
                          (8 + 20 + 8, 0xFF, "Invalid"), (8 + 20 + 4, 0xFF, "Invalid")]:
        b = bytearray(good)
        b[at] = val
        assert kind_of(b) == want, (at, val)
    y2 = bits("100000 1000 1 1 00000000 100101101 1 0")
    b = frame(frame_header(16, 16, 2, 0, 0, 0, 0), [picture(0, [slice_(1, y2, cb, cr, 0)])], 0, 0)
    assert dec(b) == ("Invalid", "a run passes the end of the coefficient array")
    y3 = bits("0000000000")
    b = frame(frame_header(16, 16, 2, 0, 0, 0, 0), [picture(0, [slice_(1, y3, cb, cr, 0)])], 0, 0)
    assert dec(b)[0] == "Invalid"


CHECKS = [check_macroblock, check_fields, check_alpha, check_malformed]


@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_oracle_handbuilt(check):
    check(lambda d, bd=0: O.decode_result(d, bd))


@pytest.mark.gpu
@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_gpu_handbuilt(check):
    check(gpu_decode)
