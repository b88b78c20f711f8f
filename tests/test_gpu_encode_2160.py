"""The batched 2160p encode at its real shape: 3840x2160 4:2:2 HQ frames from
10-bit input, 8-macroblock slices (4,050 slices a picture), through one
encode_device_batch call, with the synthetic recipe of the benchmark
(sinusoids plus noise), default and seeded loaded matrices (entries 4-32) and
hard-edged rectangles. Every packet must be byte-identical to the oracle's
encode of the same frame and Config (src/encode.rs:223-287, 657-830) and at
or under the profile's target_frame_bytes (tests/roundtrip.rs:47-73 checks
the same bound at 1080p).
"""

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P
from prores_amd import synth

pytestmark = pytest.mark.gpu

# (seed, loaded matrices, rectangles)
FRAMES = [(9000, False, False), (9017, True, False), (9004, False, True), (9028, True, True)]


def _first_difference(a, b):
    return next((k for k in range(min(len(a), len(b))) if a[k] != b[k]), None)


def test_batched_2160p_hq_encode_equals_the_oracle():
    rng = np.random.default_rng(11)
    items, bufs, want = [], [], []
    for seed, mats, rects in FRAMES:
        f = synth.frame(3840, 2160, P.ChromaFormat.Yuv422, 10, seed=seed, rectangles=rects)
        luma = chroma = None
        if mats:
            luma = [int(v) for v in rng.integers(4, 33, 64)]
            chroma = [int(v) for v in rng.integers(4, 33, 64)]
        cfg = P.Config(P.Profile.Hq, luma_matrix=luma, chroma_matrix=chroma)
        of = O.Frame(f.width, f.height, int(f.chroma), f.bit_depth, f.data, None, 0)
        want.append(O.encode(of, O.config(3, luma=luma, chroma=chroma)))
        dd = P.DeviceBuffer(f.data.nbytes + 64)
        dd.upload(f.data)
        bufs.append(dd)
        items.append((cfg, f, dd.ptr, 0))
    target = P.Profile.Hq.target_frame_bytes(3840, 2160)
    assert target == O.target_frame_bytes(3, 3840, 2160) == 3_643_238
    print("case: 4 x 3840x2160 422 HQ batched encode", flush=True)
    out = P.encode_device_batch(items)
    got = []
    for r in out:
        assert not isinstance(r, Exception), r
        got.append(P.download_ptr(*r).tobytes())
    for i, (g, w) in enumerate(zip(got, want)):
        assert len(w) <= target, (i, len(w))
        assert g == w, (i, FRAMES[i], len(g), len(w), _first_difference(g, w))
    # The same frames one call each give the same bytes.
    for i, it in enumerate(items[:2]):
        print("case: 3840x2160 422 HQ single encode", i, flush=True)
        r = P.encode_device_batch([it])[0]
        assert P.download_ptr(*r).tobytes() == want[i]
