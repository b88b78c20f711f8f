"""GPU batch encode: one launch sequence over device-resident frames of mixed
profile, size, slice width, matrices, alpha and scan. Every packet must be
byte-identical to the oracle's encode of the same frame and Config
(src/encode.rs:223-287), and a frame the Config rejects must report the
reference's error without disturbing the others.
"""

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu

# (width, height, chroma, depth, profile, log2, interlace, alpha type, matrices, rectangles)
CASES = [
    (720, 486, 2, 10, 3, 1, 1, 0, False, False),
    (1280, 720, 2, 10, 1, 2, 0, 0, True, True),
    (640, 360, 2, 10, 0, 0, 0, 0, False, False),
    (512, 270, 3, 12, 4, 3, 0, 1, True, False),
    (333, 75, 3, 12, 5, 3, 2, 2, False, True),
    (1920, 1080, 2, 10, 3, 3, 0, 0, False, True),
]


def make(i, case):
    w, h, chroma, depth, profile, log2, interlace, atype, matrices, rects = case
    rng = np.random.default_rng(200 + i)
    f = O.test_frame(w, h, chroma, depth, int(rng.integers(2, 40)), 70 + i)
    f.interlace = interlace
    if rects:
        data = f.data.astype(np.int64)
        pw = w
        for _ in range(3):
            x0, y0 = int(rng.integers(0, w // 2)), int(rng.integers(0, h // 2))
            rw, rh = int(rng.integers(8, w // 2)), int(rng.integers(8, h // 2))
            step = int(rng.integers(300, 600)) * (1 << (depth - 10))
            plane = data[:w * h].reshape(h, pw)
            plane[y0:y0 + rh, x0:x0 + rw] += step
        f.data = np.clip(data, 0, (1 << depth) - 1).astype(np.uint16)
    if atype:
        mx = (1 << depth) - 1
        yy, xx = np.mgrid[0:h, 0:w]
        d = np.hypot((xx - w * 0.5) / (0.35 * w), (yy - h * 0.5) / (0.35 * h))
        f.alpha = (np.clip((1.0 - d) * min(w, h) * 0.35 / 16.0, 0, 1) * mx).round().astype(np.uint16).ravel()
    luma = chroma_m = None
    if matrices:
        luma = [int(v) for v in rng.integers(4, 33, 64)]
        chroma_m = [int(v) for v in rng.integers(4, 33, 64)]
    oc = O.config(profile, log2_slice_mbs=log2, alpha=atype or 2, luma=luma, chroma=chroma_m)
    pc = P.Config(profile, log2_slice_mbs=log2, alpha=atype or 2, luma_matrix=luma, chroma_matrix=chroma_m)
    return f, oc, pc


def to_p(frame):
    return P.Frame(frame.width, frame.height, frame.chroma, frame.bit_depth, frame.data.copy(),
                   None if frame.alpha is None else frame.alpha.copy(), frame.interlace,
                   P.Metadata(*frame.metadata))


def test_mixed_encode_batch():
    frames, items, bufs, want = [], [], [], []
    for i, c in enumerate(CASES):
        f, oc, pc = make(i, c)
        print("case: encode batch frame", i, c, flush=True)
        want.append(O.encode(f, oc))
        pf = to_p(f)
        dd = P.DeviceBuffer(f.data.nbytes + 64)
        dd.upload(f.data)
        da = None
        if f.alpha is not None:
            da = P.DeviceBuffer(f.alpha.nbytes + 64)
            da.upload(f.alpha)
        bufs.append((dd, da))
        items.append((pc, pf, dd.ptr, da.ptr if da else 0))
    # A frame whose depth the Config cannot take (4:2:2 profile, 4:4:4 frame).
    bad = to_p(make(3, CASES[3])[0])
    dbad = P.DeviceBuffer(bad.data.nbytes + 64)
    dbad.upload(bad.data)
    bad_cfg = P.Config(P.Profile.Hq)
    items.insert(2, (bad_cfg, bad, dbad.ptr, 0))
    try:
        P.Encoder(bad_cfg).encode(bad)
        bad_err = None
    except P.Error as e:
        bad_err = (e.kind, e.msg)
    assert bad_err is not None
    out = P.encode_device_batch(items)
    assert isinstance(out[2], P.Error) and (out[2].kind, out[2].msg) == bad_err
    del out[2]
    for i, (r, w) in enumerate(zip(out, want)):
        assert not isinstance(r, Exception), (i, r)
        ptr, n = r
        got = P.download_ptr(ptr, n).tobytes()
        assert got == w, (i, len(got), len(w), next((k for k in range(min(len(got), len(w))) if got[k] != w[k]), None))


_FRESH = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import prores_amd as P
data = np.fromfile(sys.argv[2], dtype=np.uint16)
f = P.Frame(1920, 1080, 2, 10, data)
items = []
bufs = []
for target in (64, 64, 200):
    dd = P.DeviceBuffer(data.nbytes + 64)
    dd.upload(data)
    bufs.append(dd)
    items.append((P.Config(P.Profile.Hq, target_frame_bytes=target or None), f, dd.ptr, 0))
out = P.encode_device_batch(items)
for r in out:
    if isinstance(r, Exception):
        raise r
    sys.stdout.write(P.download_ptr(*r).tobytes().hex() + "\n")
"""


def test_output_arena_grows_in_a_fresh_process(tmp_path):
    """A first batch whose packets outgrow the output arena's first guess
    (frames whose targets are far below what qi 224 takes) is packed again
    over a larger arena; its packets equal the oracle's. It runs in a fresh
    process, where the arena starts at the batch's guess."""
    import os
    import subprocess
    import sys

    f = O.test_frame(1920, 1080, 2, 10, 12, 31)
    path = tmp_path / "frame.u16"
    f.data.tofile(path)
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rivet-prores_amd")
    print("case: fresh-process encode with targets 64, 64, 200", flush=True)
    r = subprocess.run([sys.executable, "-c", _FRESH, pkg, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = [bytes.fromhex(x) for x in r.stdout.split()]
    want = [O.encode(f, O.config(3, target=t)) for t in (64, 64, 200)]
    # qi 224 does not fit these targets: the packets take more than twice the
    # arena's first guess (each target plus 4 KiB, and a quarter more).
    assert sum(len(w) for w in want) > 2 * 3 * (200 + 4096)
    assert got == want
