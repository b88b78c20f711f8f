"""GPU encode over frame targets from far below the coarsest quantiser's
size to well above the finest one's, in one batch and one frame at a time:
the binary search over qi (src/encode.rs:685-713) then ends everywhere from
qi 224 (nothing fits) through the upper half of the range (224 fits, 112 does
not) down to qi 1, so the size passes' speculative first pass (224, 112, 56,
28, 14) is cut short at every point of its path. Every packet must be
byte-identical to the oracle's encode of the same frame and Config.
"""

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu


def _frame():
    return O.test_frame(720, 486, 2, 10, 30, 4242)


def _targets(f):
    # Sizes the oracle gives at a few uniform quantisers bound the range; the
    # targets fall between and around them.
    lo = len(O.encode(f, O.config(3, target=1)))  # qi 224 (or over: nothing fits)
    hi = len(O.encode(f, O.config(3, target=1 << 24)))  # qi 1
    return sorted({1, 500, lo - 1, lo, lo + 800, lo + 1600, lo + 2400, lo + 3000,
                   *np.geomspace(lo * 1.05, hi * 1.1, 14).astype(int).tolist()})


def _first_qi(packet):
    """The quantiser of the packet's first slice (frame header, picture
    header, slice table, then the slice header's second byte)."""
    pic = 8 + (packet[8] << 8 | packet[9])
    n = packet[pic + 5] << 8 | packet[pic + 6]
    return packet[pic + 8 + 2 * n + 1]


def test_targets_across_the_quantiser_range():
    f = _frame()
    targets = _targets(f)
    want = [O.encode(f, O.config(3, target=t)) for t in targets]
    pf = P.Frame(f.width, f.height, f.chroma, f.bit_depth, f.data.copy())
    dd = P.DeviceBuffer(f.data.nbytes + 64)
    dd.upload(f.data)
    print("case: encode targets", targets, flush=True)
    items = [(P.Config(P.Profile.Hq, target_frame_bytes=t), pf, dd.ptr, 0) for t in targets]
    out = P.encode_device_batch(items)
    for t, r, w in zip(targets, out, want):
        assert not isinstance(r, Exception), (t, r)
        got = P.download_ptr(*r).tobytes()
        assert got == w, (t, len(got), len(w))
    for t, w in zip(targets[::4], want[::4]):
        print("case: encode target alone", t, flush=True)
        assert P.Encoder(P.Config(P.Profile.Hq, target_frame_bytes=t)).encode(pf) == w
    # The searches ended across the range: at 224 (nothing fits), in the
    # upper half (112 does not fit), and at the finest quantisers.
    qis = [_first_qi(w) for w in want]
    print("first-slice quantisers", qis, flush=True)
    assert qis[0] == 224
    assert any(112 < q < 224 for q in qis)
    assert min(qis) <= 3
