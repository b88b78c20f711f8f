"""Writes the golden fixtures in this folder with the CPU oracle.

- readme_1080p.json: for each of the six profiles, test_frame(1920, 1080,
  chroma, depth, 24, 7) encoded with the profile's defaults (the README's
  table, README.md:173-180): the packet's size and SHA-256 and the SHA-256 of
  the decoded samples (u16 little-endian, Y'||Cb||Cr). The sizes are the
  README's own; the hashes pin everything else.
- seed_{0,1,2}.bin and seeds.json: the three small frames of tests/fuzz.rs:12-38
  (4:2:2 Standard 40x24; 4:4:4 XQ 33x21 interlaced with 16-bit alpha and
  flat matrices 5 / 9; 4:4:4 33x21 with 8-bit alpha and 1-macroblock slices),
  with the SHA-256 of their decoded samples and alpha.

Run from the repository root: python tests/golden/make_golden.py
"""

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

PROFILES = [(0, 2, 10), (1, 2, 10), (2, 2, 10), (3, 2, 10), (4, 3, 12), (5, 3, 12)]


def sha(b):
    return hashlib.sha256(bytes(b)).hexdigest()


def seed_packets():
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


def main():
    readme = []
    for prof, chroma, depth in PROFILES:
        f = O.test_frame(1920, 1080, chroma, depth, 24, 7)
        pkt = O.encode(f, O.config(prof))
        d = O.decode(pkt)
        readme.append({"profile": prof, "size": len(pkt), "packet_sha256": sha(pkt),
                       "samples_sha256": sha(d.data.astype("<u2").tobytes())})
    with open(os.path.join(HERE, "readme_1080p.json"), "w") as fh:
        json.dump(readme, fh, indent=1)
    seeds = []
    for i, pkt in enumerate(seed_packets()):
        with open(os.path.join(HERE, f"seed_{i}.bin"), "wb") as fh:
            fh.write(pkt)
        d = O.decode(pkt)
        seeds.append({"file": f"seed_{i}.bin", "size": len(pkt),
                      "samples_sha256": sha(d.data.astype("<u2").tobytes()),
                      "alpha_sha256": None if d.alpha is None else sha(d.alpha.astype("<u2").tobytes())})
    with open(os.path.join(HERE, "seeds.json"), "w") as fh:
        json.dump(seeds, fh, indent=1)


if __name__ == "__main__":
    main()
