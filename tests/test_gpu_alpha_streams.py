"""Alpha streams written pair by pair (RDD 36 alpha syntax, as the reference
reads it: src/vlc.rs:263-299, src/decode.rs:570-595), spliced into frames the
oracle encoded, and decoded by the oracle and the GPU: the Result and every
sample must be equal. The cases reach both of the GPU's alpha parses (short
streams one pair at a time, long ones by the whole wave) and the reference's
ends of a stream: the last value at `count`, trailing bytes after it, a run
past the end, a stream cut short, the end-of-data rule of a partial
macroblock row, and the longest stream 2048 values can take.
"""

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu


def pair_bits(pairs, nbits):
    """'0'/'1' text of (difference, run) pairs: a difference of magnitude
    1..2^mag as magnitude-1 and sign, any other as nbits raw bits; runs of
    1, 2..16 (4 bits) or 1..2048 (11 bits, `long`)."""
    mag = 6 if nbits == 16 else 3
    out = []
    for d, run, raw, long in pairs:
        if not raw and d != 0 and abs(d) <= (1 << mag):
            out.append("0" + format(abs(d) - 1, f"0{mag}b") + ("1" if d < 0 else "0"))
        else:
            out.append("1" + format(d % (1 << nbits), f"0{nbits}b"))
        if run == 1 and not long:
            out.append("1")
        elif 2 <= run <= 16 and not long:
            out.append("0" + format(run - 1, "04b"))
        else:
            out.append("00000" + format(run - 1, "011b"))
    return "".join(out)


def to_bytes(s):
    s += "0" * (-len(s) % 8)
    return bytes(int(s[i:i + 8], 2) for i in range(0, len(s), 8))


def random_pairs(rng, total, nbits, long_runs=True):
    pairs, n = [], 0
    while n < total:
        r = rng.random()
        run = 1 if r < 0.7 else (int(rng.integers(2, 17)) if r < 0.95 or not long_runs else int(rng.integers(17, 300)))
        run = min(run, total - n)
        small = rng.random() < 0.4
        d = int(rng.integers(-(1 << (6 if nbits == 16 else 3)), (1 << (6 if nbits == 16 else 3)) + 1)) if small \
            else int(rng.integers(0, 1 << nbits))
        pairs.append((d, run, not small, rng.random() < 0.05))
        n += run
    return pairs


def base_packet(width, height, nbits):
    f = O.Frame(width, height, 3, 12)
    f.alpha = np.zeros(width * height, np.uint16)
    return O.encode(f, O.config(5, log2_slice_mbs=3, alpha=1 if nbits == 8 else 2))


def with_alpha(packet, streams):
    """`packet` with slice k's alpha stream replaced by streams[k]."""
    hs = int.from_bytes(packet[8:10], "big")
    pic = 8 + hs
    phs = packet[pic] >> 3
    n = int.from_bytes(packet[pic + 5:pic + 7], "big")
    sizes = [int.from_bytes(packet[pic + phs + 2 * k:pic + phs + 2 * k + 2], "big") for k in range(n)]
    off = pic + phs + 2 * n
    slices = []
    for k, sz in enumerate(sizes):
        s = packet[off:off + sz]
        off += sz
        h = s[0] >> 3
        colour = sum(int.from_bytes(s[2 + 2 * c:4 + 2 * c], "big") for c in range(3))
        slices.append(s[:h + colour] + streams[k % len(streams)])
    body = b"".join(slices)
    assert all(len(s) < 65536 for s in slices)
    psize = phs + 2 * n + len(body)
    ph = bytearray(packet[pic:pic + phs])
    ph[1:5] = psize.to_bytes(4, "big")
    table = b"".join(len(s).to_bytes(2, "big") for s in slices)
    fh = bytearray(packet[:pic])
    fsize = pic + psize
    fh[0:4] = fsize.to_bytes(4, "big")
    return bytes(fh) + bytes(ph) + table + body


def check(packet, nbits):
    want = O.decode_result(packet)
    try:
        got = P.Decoder().decode(packet)
    except P.Error as e:
        assert (e.kind, e.msg) == (want[0], want[1])
        return want[0]
    assert want[0] == "Ok", want
    np.testing.assert_array_equal(np.asarray(got.alpha), want[1].alpha)
    np.testing.assert_array_equal(np.asarray(got.data), want[1].data)
    return "Ok"


def streams_for(rng, nbits):
    """name -> stream bytes, for a slice of 2048 values."""
    out = {}
    pairs = random_pairs(rng, 2048, nbits)
    exact = to_bytes(pair_bits(pairs, nbits))
    out["exact"] = exact
    out["trailing"] = exact + bytes(rng.integers(0, 256, 37, dtype=np.uint8))
    out["beyond 8448 bytes"] = exact + bytes(rng.integers(0, 256, 10000, dtype=np.uint8))
    over = random_pairs(rng, 2040, nbits) + [(5, 20, False, False)]
    out["run past the end"] = to_bytes(pair_bits(over, nbits)) + bytes(rng.integers(0, 256, 600, dtype=np.uint8))
    out["cut short"] = exact[: len(exact) * 2 // 3]
    ramp = [(4097 % (1 << nbits), 1, True, False)] * 2048
    out["ramp"] = to_bytes(pair_bits(ramp, nbits))
    out["longest"] = to_bytes(pair_bits([(int(rng.integers(0, 1 << nbits)), 1, True, True)] * 2048, nbits))
    out["one long run"] = to_bytes(pair_bits([(7, 1, True, False)] * 300 + [(9, 1748, True, True)], nbits))
    for size in (62, 63, 64, 65, 66):
        # a prefix of about `size` bytes, then one run to the end
        ps, bitsum = [], 0
        while True:
            p = (int(rng.integers(0, 1 << nbits)), 1, True, False)
            b = len(pair_bits([p], nbits))
            if bitsum + b + 30 > 8 * size:
                break
            ps.append(p)
            bitsum += b
        ps.append((3, 2048 - len(ps), True, True))
        st = to_bytes(pair_bits(ps, nbits))
        out[f"{len(st)} bytes"] = st
    return out


@pytest.mark.parametrize("nbits", (8, 16))
def test_crafted_alpha_streams(nbits):
    rng = np.random.default_rng(nbits)
    base = base_packet(128, 16, nbits)
    for name, st in streams_for(rng, nbits).items():
        print("case:", nbits, "bit alpha,", name, len(st), "bytes", flush=True)
        check(with_alpha(base, [st]), nbits)


@pytest.mark.parametrize("nbits", (8, 16))
def test_partial_row_end_of_data(nbits):
    """A 128x24 frame: the second row of slices shows 8 of 16 rows, so its
    stream may end (zero bits only) once 1024 values are decoded."""
    rng = np.random.default_rng(100 + nbits)
    base = base_packet(128, 24, nbits)
    full = to_bytes(pair_bits(random_pairs(rng, 2048, nbits), nbits))
    for extra in (0, 1, 3, 40):
        half = to_bytes(pair_bits(random_pairs(rng, 1100, nbits), nbits)) + bytes(extra)
        print("case:", nbits, "bit alpha, partial row, zero bytes", extra, flush=True)
        assert check(with_alpha(base, [full, half]), nbits) == "Ok"
    print("case:", nbits, "bit alpha, partial row, stream too short", flush=True)
    short = to_bytes(pair_bits(random_pairs(rng, 900, nbits), nbits))
    assert check(with_alpha(base, [full, short]), nbits) != "Ok"
