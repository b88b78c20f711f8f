"""GPU slice-table walk on a crafted table (src/decode.rs:202-221): a
16384x4160 4:2:2 picture of one-macroblock slices (266,240 entries) whose
sizes before the last 256-slice chunk add up to exactly 2^32 in each wave's
share, so sums kept in 32 bits wrap to a carry of 0 and the table would pass
its bounds check. The reference's walk stops at the first slice that ends
past picture_size; the GPU must report the same error, on the per-call path
and in a device batch beside a good frame, and decode the good frame.
"""

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P
from test_kernel_logic import crafted_wrap_table

pytestmark = pytest.mark.gpu


def crafted_packet():
    w, h = 16384, 4160
    sizes = crafted_wrap_table((w // 16) * (h // 16))
    table = sizes.astype(">u2").tobytes()
    data = b"\x00" * 64
    pic_size = 8 + len(table) + len(data)
    n = len(sizes)
    pic = bytes([8 << 3]) + pic_size.to_bytes(4, "big") + (n & 0xFFFF).to_bytes(2, "big") + bytes([0 << 4])
    hdr = (20).to_bytes(2, "big") + bytes([0, 0]) + b"hand" + w.to_bytes(2, "big") + h.to_bytes(2, "big")
    hdr += bytes([(2 << 6) | (0 << 2), 0, 2, 2, 2, 0, 0, 0])
    body = hdr + pic + table + data
    return (8 + len(body)).to_bytes(4, "big") + b"icpf" + body


def test_wrapping_table_is_a_table_error():
    pkt = crafted_packet()
    want = O.decode_result(pkt)
    assert want == ("Invalid", "a slice runs past picture_size")
    print("case: crafted 266,240-slice table, per-call decode", flush=True)
    with pytest.raises(P.Invalid) as e:
        P.Decoder().decode(pkt)
    assert e.value.msg == want[1]

    print("case: crafted table in a device batch with a good frame", flush=True)
    good = O.encode(O.test_frame(256, 64, 2, 10, 12, 3), O.config(3))
    bufs, jobs = [], []
    for p in (pkt, good):
        hd = P.FrameHeader.parse(p)
        planes = P.layout(hd.width, hd.height, hd.chroma)
        n = planes[2][0] + planes[2][1] * planes[2][2]
        dp = P.DeviceBuffer(len(p) + 128)
        dp.upload(np.frombuffer(p, np.uint8))
        dd = P.DeviceBuffer(n * 2 + 64)
        bufs.append((dp, dd, n))
        jobs.append((p, dp.ptr, dd.ptr, 0))
    errs = P.Decoder().decode_device_batch(jobs)
    assert errs[0] is not None and (errs[0].kind, errs[0].msg) == want
    assert errs[1] is None
    np.testing.assert_array_equal(bufs[1][1].download(bufs[1][2] * 2, dtype=np.uint16), O.decode(good).data)
