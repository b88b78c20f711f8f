"""Seeded synthetic pictures for the benchmark and the batch tests.

The recipe (numpy PCG64 generator, one seed per frame):
- each colour plane is a centre value plus six 2-D sinusoids (period 8-512 px,
  amplitude 20-120 codes at 10 bits, random direction and phase) plus
  Gaussian noise, rounded and clamped: luma 64-940, chroma 64-960 at 10 bits
  (centre 512). At 12 bits the centre and amplitudes scale by 4 and the range
  is 256-3760. Chroma planes use the same recipe at their own width.
- optional hard-edged rectangles (steps of 300-600 codes at 10 bits).
- alpha: soft-edged ellipses, 0 outside and the maximum at the frame's depth
  inside, with 16-px linear ramps.

Sinusoids are evaluated separably (sin(a+b) = sin a cos b + cos a sin b): the
six of a plane are two rank-6 matrix products.
"""

import numpy as np

from . import ChromaFormat, Frame, Interlace


def _plane(rng, h, w, xs, centre, amp, sigma, lo, hi):
    # sum_j a_j sin(kx_j x + ky_j y + phase_j) as two rank-6 products:
    # sin(p + q) = sin p cos q + cos p sin q with p = ky y + phase, q = kx x.
    y = np.arange(h, dtype=np.float64)[:, None]
    x = (np.arange(w, dtype=np.float64) * xs)[:, None]
    params = []
    for _ in range(6):
        period = rng.uniform(8.0, 512.0)
        a = rng.uniform(20.0, 120.0) * amp
        theta = rng.uniform(0.0, 2 * np.pi)
        phase = rng.uniform(0.0, 2 * np.pi)
        params.append((a, 2 * np.pi * np.cos(theta) / period, 2 * np.pi * np.sin(theta) / period, phase))
    a = np.array([p[0] for p in params])[None, :]
    kx = np.array([p[1] for p in params])[None, :]
    ky = np.array([p[2] for p in params])[None, :]
    ph = np.array([p[3] for p in params])[None, :]
    py = ky * y + ph                      # h x 6
    qx = kx * x                           # w x 6
    v = (a * np.sin(py)) @ np.cos(qx).T + (a * np.cos(py)) @ np.sin(qx).T
    v += float(centre)
    if sigma > 0:
        v += sigma * rng.standard_normal(size=(h, w))
    return v


def frame(width, height, chroma=ChromaFormat.Yuv422, bit_depth=10, seed=0, sigma=None, alpha=False,
          rectangles=False, interlace=Interlace.Progressive):
    """One synthetic Frame (see the module docstring)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    chroma = ChromaFormat(int(chroma))
    scale = 1 << (bit_depth - 10) if bit_depth >= 10 else 1.0 / (1 << (10 - bit_depth))
    if sigma is None:
        sigma = 6.0 * scale if bit_depth <= 10 else 24.0
    f = Frame(width, height, chroma, bit_depth, interlace=interlace)
    for i in range(3):
        off, w, h = f.planes[i]
        xs = 1.0 if i == 0 else float(chroma.subsampling()[0])
        lo, hi = (64, 940) if i == 0 else (64, 960)
        if bit_depth == 12:
            lo, hi = 256, 3760
        else:
            lo, hi = lo * scale, hi * scale
        v = _plane(rng, h, w, xs, 512 * scale, scale, sigma, lo, hi)
        if rectangles:
            for _ in range(int(rng.integers(2, 6))):
                rw, rh = int(rng.integers(w // 16 + 1, w // 3 + 2)), int(rng.integers(h // 16 + 1, h // 3 + 2))
                x0, y0 = int(rng.integers(0, max(1, w - rw))), int(rng.integers(0, max(1, h - rh)))
                step = rng.uniform(300, 600) * scale * (1 if rng.random() < 0.5 else -1)
                v[y0:y0 + rh, x0:x0 + rw] += step
        f.data[off:off + w * h] = np.clip(np.rint(v), lo, hi).astype(np.uint16).ravel()
    if alpha:
        f.alpha = ellipses(width, height, bit_depth, rng)
    return f


def ellipses(width, height, bit_depth, rng, count=3, ramp=16.0):
    """Soft-edged ellipses: 0 outside, the maximum inside, linear ramps."""
    mx = (1 << bit_depth) - 1
    y = np.arange(height, dtype=np.float64)[:, None]
    x = np.arange(width, dtype=np.float64)[None, :]
    a = np.zeros((height, width), dtype=np.float64)
    for _ in range(count):
        cx, cy = rng.uniform(0, width), rng.uniform(0, height)
        rx = rng.uniform(0.1, 0.4) * width
        ry = rng.uniform(0.1, 0.4) * height
        d = np.sqrt(((x - cx) / rx) ** 2 + ((y - cy) / ry) ** 2)
        # distance to the edge in pixels, approximately: (1 - d) * min radius
        inside = (1.0 - d) * min(rx, ry)
        a = np.maximum(a, np.clip(inside / ramp, 0.0, 1.0))
    return np.rint(a * mx).astype(np.uint16).ravel()
