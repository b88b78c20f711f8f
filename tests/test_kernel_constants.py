"""Constants compiled into the gfx950 kernels, checked on the CPU.

The IDCT takes the DCT basis as instruction literals from
rivet-prores_amd/csrc/dct_basis.h; every entry must be the oracle's f32 basis
(src/dct.rs:11-25), bit for bit.
"""

import ctypes as C
import os
import re
import struct

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_dct_basis_header_matches_oracle():
    text = open(os.path.join(ROOT, "rivet-prores_amd", "csrc", "dct_basis.h")).read()
    body = text[text.index("BASIS[64] = {"):]
    vals = [float.fromhex(m) for m in re.findall(r"(-?0x[0-9a-f.]+p[-+]?\d+)f", body)]
    assert len(vals) == 64
    ob = (C.c_float * 64)()
    O.lib().oracle_basis(ob)
    for i in range(64):
        assert f32_bits(vals[i]) == f32_bits(ob[i]), (i, vals[i], ob[i])
