"""Pins for the CPU oracle: the reference's own known-answer vectors, and
what the mathematics fixes where the reference has no vector.

Each test names the reference test or text it restates. Together they pin
every oracle function the GPU parity tests rely on:
  - codewords: the hand-worked KATs of src/vlc.rs:459-493, 644-712;
  - the EG == Combo(0, k, k+1) identity (src/vlc.rs:495-503);
  - tables and slice layout (src/tables.rs:143-177, src/header.rs:326-361);
  - the DCT basis against its closed form (50-digit arithmetic);
  - the IDCT against RDD 36 Annex A (src/dct.rs:389-440) and the FDCT
    against an f64 DCT (src/dct.rs:442-467);
  - quantise at ties, signed zeros and saturation;
  - Rescale against the division form for every value (src/decode.rs:603-619);
  - hand-built frames decoded to hand-computed samples (tests/handbuilt.rs);
  - the encoder end to end: the six README packet sizes (README.md:173-180).
"""

import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O

L = O.lib()
EG, CB = 0, 1


def cw(kind, a, b, c, n):
    return O.codeword(kind, a, b, c, n)


# ---- VLC (src/vlc.rs) -------------------------------------------------------

def test_exp_golomb_by_hand():  # src/vlc.rs:459-474
    eg0 = ["1", "010", "011", "00100", "00101", "00110", "00111", "0001000"]
    for n, want in enumerate(eg0):
        assert cw(EG, 0, 0, 0, n) == want
    eg1 = ["10", "11", "0100", "0101", "0110", "0111", "001000"]
    for n, want in enumerate(eg1):
        assert cw(EG, 1, 0, 0, n) == want
    assert cw(EG, 5, 0, 0, 0) == "100000"
    assert cw(EG, 5, 0, 0, 32) == "01000000"


def test_combination_code_by_hand():  # src/vlc.rs:476-493
    table = [((2, 0, 1), 0, "1"), ((2, 0, 1), 1, "01"), ((2, 0, 1), 2, "001"), ((2, 0, 1), 3, "00010"),
             ((2, 0, 1), 4, "00011"), ((2, 0, 1), 5, "0000100"), ((1, 2, 3), 0, "100"),
             ((1, 2, 3), 5, "0101"), ((1, 2, 3), 7, "0111"), ((1, 2, 3), 8, "001000"),
             ((1, 2, 3), 15, "001111"), ((1, 2, 3), 16, "00010000")]
    for (a, b, c), n, want in table:
        assert cw(CB, a, b, c, n) == want


def test_exp_golomb_is_combo_with_last_rice_q_0():  # src/vlc.rs:495-503
    for k in range(6):
        for n in range(0, 2000, 7):
            assert cw(EG, k, 0, 0, n) == cw(CB, 0, k, k + 1, n)


def _eg_closed_form(n, k):
    """RDD 36 §7.1.1.1: q = floor(log2(n + 2^k)) - k zeros, then n + 2^k."""
    v = n + (1 << k)
    q = v.bit_length() - 1 - k
    return "0" * q + format(v, "b")


ALL_BOOKS = ([(EG, 0, 0, 0), (EG, 1, 0, 0), (CB, 1, 2, 3), (EG, 3, 0, 0)]
             + [(CB, 2, 0, 1), (CB, 1, 0, 1), (EG, 0, 0, 0), (CB, 1, 1, 2), (EG, 1, 0, 0), (EG, 2, 0, 0)]
             + [(CB, 2, 0, 2), (CB, 1, 0, 1), (CB, 2, 0, 1), (EG, 5, 0, 0)])


def test_codewords_against_the_closed_forms():
    """Every codebook against §7.1.1.1's definition, and len == bits written."""
    for kind, a, b, c in ALL_BOOKS:
        for n in list(range(0, 600)) + [65535, 1 << 20]:
            if kind == EG:
                want = _eg_closed_form(n, a)
            else:
                th = (a + 1) << b
                if n < th:
                    want = "0" * (n >> b) + "1" + (format(n & ((1 << b) - 1), f"0{b}b") if b else "")
                else:
                    want = "0" * (a + 1) + _eg_closed_form(n - th, c)
            assert cw(kind, a, b, c, n) == want, (kind, a, b, c, n)
            assert L.oracle_code_len(kind, a, b, c, n) == len(want)


def _bytes_of(bits):
    bits = bits + "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def test_every_codebook_reads_back():  # src/vlc.rs:505-531
    msg = C.create_string_buffer(256)
    sym = C.c_uint32()
    for kind, a, b, c in ALL_BOOKS:
        for n in list(range(0, 400, 3)) + [65535, 1 << 20, (1 << 24) + 7, 1 << 30]:
            data = _bytes_of(cw(kind, a, b, c, n))
            assert L.oracle_code_read(kind, a, b, c, data, len(data), C.byref(sym), msg, 256) == 0
            assert sym.value == n


def test_read_errors():
    """src/bits.rs:79-94, src/vlc.rs:224-228: a prefix over 32 zeros, a
    codeword past the end, a value over 2^30."""
    msg = C.create_string_buffer(256)
    sym = C.c_uint32()
    assert L.oracle_code_read(EG, 0, 0, 0, bytes(5), 5, C.byref(sym), msg, 256) == 1
    assert msg.value == b"a variable-length code's prefix is too long"
    assert L.oracle_code_read(EG, 0, 0, 0, b"\x00\x01", 2, C.byref(sym), msg, 256) == 1
    assert msg.value == b"a codeword runs past the end of its data"
    big = _bytes_of(_eg_closed_form((1 << 30) + 1, 0))
    assert L.oracle_code_read(EG, 0, 0, 0, big, len(big), C.byref(sym), msg, 256) == 1
    assert msg.value == b"a variable-length code's value is out of range"


def test_signed_mapping_is_table_8():  # src/vlc.rs:622-632
    for n, s in [(0, 0), (-1, 1), (1, 2), (-2, 3), (2, 4), (-3, 5), (3, 6)]:
        assert L.oracle_signed_to_symbol(n) == s
        assert L.oracle_symbol_to_signed(s) == n
    for n in range(-70000, 70000, 37):
        assert L.oracle_symbol_to_signed(L.oracle_signed_to_symbol(n)) == n


def _alpha_run(run):
    buf = C.create_string_buffer(64)
    n = L.oracle_alpha_run_bits(run, buf, 64)
    return buf.value.decode()[:n]


def _alpha_diff(d, bits):
    buf = C.create_string_buffer(64)
    n = L.oracle_alpha_diff_bits(d, bits, buf, 64)
    return buf.value.decode()[:n]


def test_alpha_runs_are_table_12():  # src/vlc.rs:644-665
    for run, want in [(1, "1"), (2, "00001"), (3, "00010"), (15, "01110"), (16, "01111"),
                      (17, "0000000000010000"), (18, "0000000000010001"), (2047, "0000011111111110"),
                      (2048, "0000011111111111")]:
        assert _alpha_run(run) == want


def test_alpha_differences_are_tables_13_and_14():  # src/vlc.rs:667-712
    for d, want in [(1, "00000"), (-1, "00001"), (2, "00010"), (-2, "00011"), (8, "01110"), (-8, "01111")]:
        assert _alpha_diff(d, 8) == want
    assert _alpha_diff(9, 8) == "100001001"
    assert _alpha_diff(-9, 8) == "111110111"
    assert _alpha_diff(0, 8) == "100000000"
    for d, want in [(1, "00000000"), (-1, "00000001"), (2, "00000010"), (64, "01111110"), (-64, "01111111")]:
        assert _alpha_diff(d, 16) == want
    assert _alpha_diff(65, 16) == "10000000001000001"


# ---- tables, layout (src/tables.rs, src/header.rs) -------------------------

def scans():
    p = (C.c_uint8 * 64)()
    i = (C.c_uint8 * 64)()
    L.oracle_scan(0, p)
    L.oracle_scan(1, i)
    return list(p), list(i)


def test_scans_are_permutations_and_transposes():  # src/tables.rs:143-159
    p, i = scans()
    assert sorted(p) == list(range(64)) and sorted(i) == list(range(64))
    for v in range(8):
        for u in range(8):
            assert p[8 * v + u] == i[8 * u + v]
    assert p[0] == 0 and p[63] == 63
    # Figure 4's first row and column.
    assert p[:8] == [0, 1, 4, 5, 16, 17, 21, 22]
    assert [p[8 * v] for v in range(8)] == [0, 2, 8, 10, 32, 34, 35, 41]


def test_qscale_is_table_15():  # src/tables.rs:161-177
    for i, q in [(1, 1), (2, 2), (126, 126), (127, 127), (128, 128), (129, 132), (130, 136), (223, 508),
                 (224, 512)]:
        assert L.oracle_qscale(i) == q


def slice_sizes(w, log2):
    out = (C.c_uint32 * 4096)()
    n = L.oracle_slice_sizes(w, log2, out, 4096)
    return list(out[:n])


def test_slice_layout_is_section_4():  # src/header.rs:326-334
    assert slice_sizes(45, 3) == [8, 8, 8, 8, 8, 4, 1]
    assert slice_sizes(1, 3) == [1]
    assert slice_sizes(7, 3) == [4, 2, 1]
    assert len(slice_sizes(120, 3)) == 15
    assert slice_sizes(5, 0) == [1] * 5
    for w in range(1, 300):
        for log2 in range(4):
            s = slice_sizes(w, log2)
            assert sum(s) == w and s == sorted(s, reverse=True)
            assert all(x & (x - 1) == 0 and x <= (1 << log2) for x in s)


def test_target_frame_bytes():  # src/encode.rs:988-995, README.md:60-83
    assert O.target_frame_bytes(3, 1920, 1080) == 917_583
    assert O.target_frame_bytes(3, 16, 16) == 155
    readme = [187_687, 425_425, 613_112, 917_583, 1_376_375, 2_085_416]
    for p in range(6):
        assert O.target_frame_bytes(p, 1920, 1080) == readme[p]


# ---- DCT (src/dct.rs) -------------------------------------------------------

def basis():
    b = (C.c_float * 64)()
    L.oracle_basis(b)
    return np.array(b, dtype=np.float32)


def test_basis_is_the_closed_form_rounded():
    import mpmath
    mpmath.mp.dps = 50
    b = basis()
    for x in range(8):
        for u in range(8):
            c = 1 / mpmath.sqrt(2) if u == 0 else mpmath.mpf(1)
            exact = c / 2 * mpmath.cos((2 * x + 1) * u * mpmath.pi / 16)
            assert b[8 * x + u] == np.float32(float(exact)), (x, u)


def _ieee1180(seed):
    state = [seed]

    def nxt(l, h):
        state[0] = (state[0] * 1103515245 + 12345) & 0xFFFFFFFF
        v = state[0] - (1 << 32) if state[0] >= 1 << 31 else state[0]
        i = v & 0x7FFFFFFE
        return int(i / 0x7FFFFFFF * (l + h + 1)) - l
    return nxt


def _cos_matrix():
    c = np.array([1 / math.sqrt(2)] + [1.0] * 7)
    return np.array([[c[u] / 2 * math.cos((2 * x + 1) * u * math.pi / 16) for u in range(8)] for x in range(8)])


def test_idct_annex_a_accuracy():  # src/dct.rs:389-440, 2000 blocks per case
    T = _cos_matrix()
    out = (C.c_float * 64)()
    for l, h in [(256, 255), (5, 5), (300, 300)]:
        for sign in (1.0, -1.0):
            rng = _ieee1180(1)
            n = 2000
            err_sum = np.zeros(64)
            err_sq = np.zeros(64)
            peak = 0.0
            for _ in range(n):
                p = np.array([sign * rng(l, h) for _ in range(64)]).reshape(8, 8)
                f = T.T @ p @ T  # the orthonormal FDCT in f64
                f = np.clip(np.round(f * 4) / 4, -2048, 2047.75)
                ref = (T @ f @ T.T).ravel()
                ff = (C.c_float * 64)(*f.ravel().astype(np.float32))
                L.oracle_idct(ff, out)
                t = np.clip(np.array(out, dtype=np.float64), -256, 256)
                e = t - np.clip(ref, -256, 256)
                err_sum += e
                err_sq += e * e
                peak = max(peak, np.abs(e).max())
            assert peak <= 0.15
            assert (err_sq / n).max() <= 0.002 and np.abs(err_sum / n).max() <= 0.0015
            assert err_sq.sum() / (64 * n) <= 0.001 and abs(err_sum.sum() / (64 * n)) <= 0.00015


def test_fdct_matches_f64_and_inverts():  # src/dct.rs:442-467
    T = _cos_matrix()
    rng = _ieee1180(7)
    out = (C.c_float * 64)()
    back = (C.c_float * 64)()
    for _ in range(300):
        p = np.array([rng(256, 255) for _ in range(64)], dtype=np.float64).reshape(8, 8)
        want = (T.T @ p @ T).ravel()
        L.oracle_fdct((C.c_float * 64)(*p.ravel().astype(np.float32)), out)
        got = np.array(out, dtype=np.float64)
        assert np.abs(want - got).max() < 2e-3
        L.oracle_idct(out, back)
        assert np.abs(np.array(back) - p.ravel()).max() < 2e-3


def test_idct_is_a_separable_product_in_f32_order():
    """The f32 operation order of src/dct.rs:36-66, restated with numpy
    float32 scalars (each product and sum rounded to f32, never fused)."""
    b = basis().reshape(8, 8)
    rng = np.random.default_rng(3)
    out = (C.c_float * 64)()
    for _ in range(200):
        coef = rng.integers(-300, 300, 64).astype(np.float32) * np.float32(4.0)
        coef[rng.random(64) < 0.6] = 0
        tmp = np.zeros((8, 8), np.float32)
        for v in range(8):
            for x in range(8):
                s = np.float32(0)
                for u in range(8):
                    s = np.float32(s + np.float32(coef[8 * v + u] * b[x, u]))
                tmp[v, x] = s
        want = np.zeros((8, 8), np.float32)
        for y in range(8):
            for x in range(8):
                s = np.float32(0)
                for v in range(8):
                    s = np.float32(s + np.float32(b[y, v] * tmp[v, x]))
                want[y, x] = s
        L.oracle_idct((C.c_float * 64)(*coef), out)
        np.testing.assert_array_equal(np.array(out, np.float32), want.ravel())


def test_idct_put_conversion_and_clamp():
    """§7.5.1: s = floor(v 2^b/512 + 2^(b-1) + 0.5) clamped; a DC-only block
    is the constant F00/8 (tests/handbuilt.rs:118-121)."""
    res = (C.c_uint16 * 64)()
    for bd in (8, 10, 12, 16):
        for dc in (-4096, -100, -3, 0, 5, 64, 1000, 8000):
            coef = (C.c_int32 * 64)(dc, *([0] * 63))
            scale = (C.c_float * 64)(*([0.5] * 64))
            L.oracle_idct_put(coef, scale, bd, res)
            v = dc * 0.5 / 8.0
            want = min(max(math.floor(v * (1 << bd) / 512 + (1 << (bd - 1)) + 0.5), 0), (1 << bd) - 1)
            assert set(res) == {want}, (bd, dc)
    # Extremes saturate instead of wrapping (src/dsp/mod.rs:238-240).
    coef = (C.c_int32 * 64)(*([2**31 - 1] * 32 + [-(2**31)] * 32))
    L.oracle_idct_put(coef, (C.c_float * 64)(*([4032.0] * 64)), 10, res)
    assert all(0 <= r <= 1023 for r in res)


def test_quantise_ties_zeros_and_saturation():
    """src/dsp/scalar.rs:19-39: floor(|x| + r) with r 0.5 (DC) and 0.42 (AC),
    the sign from x < 0 (so -0.0 is +), a saturating conversion."""
    n = 1
    inv = np.ones(64, np.float32)
    f = np.zeros(64, np.float32)
    f[0] = 2.5            # DC tie: floor(3.0) = 3
    f[1] = -0.0           # -0 -> 0
    f[2] = 0.58           # floor(0.58 + 0.42f) = 1 (0.58f + 0.42f rounds to 1.0)
    f[3] = 0.5799         # 0
    f[4] = -1.5799        # -floor(2.0) = -1? (1.5799 + 0.42 = 1.9999) -> -1
    f[5] = 3e9            # saturates to i32::MAX
    f[6] = -3e9
    f[7] = float(np.float32(0.58) - np.float32(0.42))
    out = (C.c_int32 * 64)()
    mask = (C.c_uint64 * 1)()
    L.oracle_quantise(f.ctypes.data, n, inv.ctypes.data, C.c_float(0.42), out, mask)
    got = list(out)
    r = np.float32(0.42)
    want = []
    for s in range(64):
        x = np.float32(f[s] * inv[s])
        t = np.floor(np.float32(np.abs(x) + (np.float32(0.5) if s == 0 else r)))
        m = 2**31 - 1 if t >= 2.0**31 else int(t)
        want.append(-m if x < 0 else m)
    assert got == want
    assert got[:7] == [3, 0, 1, 0, -1, 2**31 - 1, -(2**31 - 1)]
    assert mask[0] == sum(1 << i for i, v in enumerate(got) if v)


def test_rescale_is_exact():  # src/decode.rs:603-619
    for bi in range(8, 17):
        a = np.arange(1 << bi, dtype=np.uint16)
        out = np.zeros_like(a)
        for bo in range(8, 17):
            L.oracle_rescale_many(bi, bo, a.ctypes.data, a.size, out.ctypes.data)
            mi, mo = (1 << bi) - 1, (1 << bo) - 1
            want = (2 * mo * a.astype(np.int64) + mi) // (2 * mi)
            np.testing.assert_array_equal(out, want.astype(np.uint16))


def test_end_of_data_needs_fewer_than_32_zero_bits():  # src/bits.rs:228-234
    out = (C.c_uint16 * 64)()
    n = C.c_size_t()
    msg = C.create_string_buffer(256)
    # Alpha decoding with nothing needed stops at endOfData.
    assert L.oracle_decode_alpha(bytes(3), 3, 8, 16, 0, out, C.byref(n), msg, 256) == 0 and n.value == 0
    assert L.oracle_decode_alpha(bytes(4), 4, 8, 16, 0, out, C.byref(n), msg, 256) == 0 and n.value == 1
    # [0, 0, 1] is not endOfData: the reader goes on and runs out of bits.
    assert L.oracle_decode_alpha(b"\x00\x00\x01", 3, 8, 16, 0, out, C.byref(n), msg, 256) == 1
    assert msg.value == b"a codeword runs past the end of its data"


# ---- entropy coder round trips (src/encode.rs:875-986) ----------------------

def test_coefficients_round_trip_and_count():
    seed = [0x12345678]

    def rnd():
        s = seed[0]
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        seed[0] = s
        return s
    for nb in (1, 2, 4, 8, 16, 32):
        for density in (0, 1, 8, 50, 100):
            c = np.zeros(nb * 64, np.int32)
            for i in range(nb * 64):
                if i < nb or rnd() % 100 < density:
                    mag = (rnd() % 3000) >> (rnd() % 10)
                    v = -mag if rnd() & 1 else mag
                    if i >= nb and v == 0:
                        v = 1
                    c[i] = v
            buf = np.zeros(nb * 64 * 16, np.uint8)
            nbytes = L.oracle_encode_coefficients(c.ctypes.data, nb, buf.ctypes.data, buf.size)
            got = np.zeros(nb * 64, np.int32)
            msg = C.create_string_buffer(256)
            assert L.oracle_decode_coefficients(bytes(buf[:nbytes]), nbytes, nb, got.ctypes.data, msg, 256) == 0
            np.testing.assert_array_equal(got, c)
            assert (L.oracle_coefficient_bits(c.ctypes.data, nb) + 7) // 8 == nbytes


def test_alpha_round_trips():
    for bits in (8, 16):
        mx = (1 << bits) - 1
        v = [mx] * 3000 + [0, 1, 2, 200, mx // 2, 3, 3, 3, mx, 0] + [(i * 37) % (mx + 1) for i in range(500)]
        v = np.array(v, np.uint16)
        buf = np.zeros(v.size * 4, np.uint8)
        nbytes = L.oracle_encode_alpha(v.ctypes.data, v.size, bits, buf.ctypes.data, buf.size)
        out = np.zeros(v.size, np.uint16)
        n = C.c_size_t()
        msg = C.create_string_buffer(256)
        assert L.oracle_decode_alpha(bytes(buf[:nbytes]), nbytes, bits, v.size, v.size, out.ctypes.data,
                                     C.byref(n), msg, 256) == 0
        np.testing.assert_array_equal(out[:n.value], v)


# ---- end to end -------------------------------------------------------------

@pytest.mark.parametrize("profile", range(6))
def test_readme_packet_sizes_and_psnr(profile):  # README.md:173-180, tests/roundtrip.rs:47-73
    sizes = [187_677, 425_420, 613_107, 917_564, 1_376_337, 2_085_373]
    psnr = [37.54, 38.82, 40.04, 42.85, 43.03, 49.31]
    chroma, bd = (3, 12) if profile >= 4 else (2, 10)
    f = O.test_frame(1920, 1080, chroma, bd, 24, 7)
    pkt = O.encode(f, O.config(profile))
    assert len(pkt) == sizes[profile]
    d = O.decode(pkt)
    assert round(O.psnr_samples(f.data, d.data, bd), 2) == psnr[profile]


def test_thread_count_changes_nothing():  # tests/roundtrip.rs:551-612
    f = O.test_frame(250, 131, 3, 12, 20, 4)
    f.interlace = 2
    f.alpha = ((np.arange(250 * 131) * 13) % 4096).astype(np.uint16)
    one = O.encode(f, O.config(5, threads=1))
    pic = O.decode(one, threads=1)
    dmg = bytearray(one)
    for i in range(len(dmg) // 2, len(dmg) // 2 + 64):
        dmg[i] ^= 0x5A
    dres = O.decode_result(bytes(dmg), threads=1)
    for t in (0, 2, 3, 16):
        assert O.encode(f, O.config(5, threads=t)) == one
        d = O.decode(one, threads=t)
        assert np.array_equal(d.data, pic.data) and np.array_equal(d.alpha, pic.alpha)
        r = O.decode_result(bytes(dmg), threads=t)
        assert r[0] == dres[0] and (r[0] == "Ok" or r == dres)
