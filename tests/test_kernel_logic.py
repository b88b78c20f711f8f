"""The kernels' closed forms and lane logic, restated on the CPU and checked
against bit-serial readers and the oracle. The GPU tests check the kernels
themselves end to end; these pin the two places where a kernel departs from
the reference's step-by-step structure.

- fast_code (csrc/decode_kernels.hip): a combination code (lastRiceQ, kRice,
  kExp) read in closed form from the top 32 bits of the window, against a
  reader that follows Codebook::read bit by bit (src/vlc.rs:206-230).
- pe_alpha (csrc/encode_kernels.hip): 64 lanes each code the runs that start
  in their 1/64 of a slice's values, a run ending at the next lane's first run
  start (an inclusive suffix minimum shifted by one lane); the bits must equal
  the oracle's encode_alpha (src/encode.rs:457-470).
"""

import ctypes as C
import random

import numpy as np

import oracle_lib as O

BOOKS = [(2, 0, 1), (1, 0, 1), (0, 0, 1), (1, 1, 2), (0, 1, 2), (0, 2, 3), (2, 0, 2), (0, 3, 4), (1, 2, 3),
         (0, 5, 6)]


def reference_read(bits, lq, kr, ke):
    """(symbol, length), 'range' or None (prefix over 32 or out of bits)."""
    q = 0
    while q < len(bits) and bits[q] == "0":
        q += 1
    if q > 32 or q >= len(bits):
        return None
    if q <= lq:
        n = 1 + kr
        if q + n > len(bits):
            return None
        return ((q << kr) | (int(bits[q:q + n], 2) & ((1 << kr) - 1)), q + n)
    n = q - (lq + 1) + ke + 1
    if q + n > len(bits):
        return None
    val = int(bits[q:q + n], 2) - (1 << ke) + ((lq + 1) << kr)
    return "range" if val > (1 << 30) else (val, q + n)


def fast_code(top, lq, kr, ke):
    """The kernel's closed form, in 32-bit arithmetic."""
    q = 32 - top.bit_length() if top else 32
    rice = q <= lq
    L = q + 1 + kr if rice else 2 * q - lq + ke
    p = (top >> ((32 - L) & 31)) & 0xFFFFFFFF
    k = kr if rice else ke
    qq = q if rice else lq + 1
    sym = (p - (1 << k) + (qq << kr)) & 0xFFFFFFFF
    return (L <= 32 and sym <= (1 << 30)), sym, L


def test_fast_code_equals_the_bit_serial_reader():
    rng = random.Random(1)
    n = 0
    for lq, kr, ke in BOOKS:
        tops = list(range(0, 1 << 14)) + [rng.getrandbits(32) >> rng.randrange(32) for _ in range(60000)]
        for t in tops:
            ok, sym, L = fast_code(t, lq, kr, ke)
            r = reference_read(format(t, "032b") + "0" * 40, lq, kr, ke)
            if ok:
                assert r not in (None, "range") and r == (sym, L), (lq, kr, ke, hex(t))
                n += 1
            else:
                assert r is None or r == "range" or r[1] > 32, (lq, kr, ke, hex(t), r)
    assert n > 100000


def lane_alpha(vals, bits):
    """pe_alpha's lane logic: bytes of the coded slice."""
    count = len(vals)
    seg = count // 64
    mag, maxd = (3, 8) if bits == 8 else (6, 64)
    first = [count] * 64
    for L in range(64):
        for i in range(L * seg, L * seg + seg):
            if i == 0 or vals[i] != vals[i - 1]:
                first[L] = i
                break
    incl = first[:]
    d = 1
    while d < 64:
        incl = [min(incl[L], incl[L + d]) if L + d < 64 else incl[L] for L in range(64)]
        d *= 2
    nxt = [incl[L + 1] if L < 63 else count for L in range(64)]
    codes = []
    for L in range(64):
        hi = L * seg + seg
        i = first[L]
        while i < hi:
            j = i + 1
            while j < hi and vals[j] == vals[i]:
                j += 1
            run = (j if j < hi else nxt[L]) - i
            dd = int(vals[i]) - (int(vals[i - 1]) if i else -1)
            if 1 <= abs(dd) <= maxd:
                codes.append((((abs(dd) - 1) << 1) | (dd < 0), 2 + mag))
            else:
                codes.append(((1 << bits) | (dd & ((1 << bits) - 1)), 1 + bits))
            codes.append((1, 1) if run == 1 else ((run - 1, 5) if run <= 16 else (run - 1, 16)))
            i = j
    s = "".join(format(v, f"0{k}b") for v, k in codes)
    s += "0" * (-len(s) % 8)
    return bytes(int(s[k:k + 8], 2) for k in range(0, len(s), 8))


def test_alpha_lane_logic_equals_the_oracle():
    rng = np.random.default_rng(5)
    buf = C.create_string_buffer(8192)
    for trial in range(160):
        bits = 8 if trial % 3 == 0 else 16
        count = 256 * int(rng.choice([1, 2, 4, 8]))
        kind = trial % 4
        if kind == 0:
            v = rng.integers(0, 1 << bits, count)
        elif kind == 1:
            v = np.repeat(rng.integers(0, 1 << bits, count // 7 + 1), rng.integers(1, 40, count // 7 + 1))[:count]
            v = np.concatenate([v, np.full(count - len(v), v[-1])]) if len(v) < count else v
        elif kind == 2:
            v = np.full(count, rng.integers(0, 1 << bits))
        else:
            v = np.clip(np.cumsum(rng.integers(-3, 4, count)) + 100, 0, (1 << bits) - 1)
        v = np.ascontiguousarray(v, dtype=np.uint16)
        n = O.lib().oracle_encode_alpha(v.ctypes.data, len(v), bits, buf, 8192)
        assert lane_alpha(v, bits) == buf.raw[:n], trial


def _basis():
    ob = (C.c_float * 64)()
    O.lib().oracle_basis(ob)
    return np.array(ob[:], dtype=np.float32).reshape(8, 8)  # T[x][u]


def test_idct_shared_products_equal_the_reference_sums():
    """pd_idct's idct8: o[x] = sum_u a[u] T[x][u] from products with |T|,
    each formed once (22 for 8 outputs) and signed inside the sum, must give
    the reference's sums (src/dct.rs:36-66: from 0.0, u ascending, every
    product and sum rounded to f32) bit for bit, up to the sign of a zero."""
    T = _basis()
    # The symmetries the kernel asserts at compile time.
    for x in range(8):
        for u in range(8):
            want = T[x, u] if u % 2 == 0 else -T[x, u]
            assert T[7 - x, u] == want
    mags = np.abs(T[0])
    rng = np.random.default_rng(5)
    n = 200_000
    a = rng.integers(-70000, 70000, size=(n, 8)).astype(np.float32) * rng.choice(
        np.array([0.125, 1.0, 3.75, 1e-30, 1e6], dtype=np.float32), size=(n, 8))
    a[rng.random((n, 8)) < 0.3] = 0.0
    a[: 1000] = rng.standard_normal((1000, 8)).astype(np.float32) * np.float32(1e-39)  # subnormal
    for x in range(8):
        ref = np.zeros(n, dtype=np.float32)
        for u in range(8):
            ref = (ref + a[:, u] * T[x, u]).astype(np.float32)
        # The kernel: sum from the first product, |T| products, signs in the sum.
        xx = x if x < 4 else 7 - x
        s = (a[:, 0] * mags[0]).astype(np.float32)
        for u in range(1, 8):
            c = T[xx, u]
            i = int(np.nonzero(mags == abs(c))[0][0])
            p = (a[:, u] * mags[i]).astype(np.float32)
            neg = (c < 0) != (x >= 4 and u % 2 == 1)
            s = (s - p if neg else s + p).astype(np.float32)
        same = (s == ref) & ((s != 0) | (ref == 0))
        assert same.all(), (x, np.nonzero(~same)[0][:5])
        nz = ref != 0
        assert np.array_equal(s[nz].view(np.uint32), ref[nz].view(np.uint32)), x


# The run and level codebooks of each context (src/vlc.rs:30-68), as
# (lastRiceQ, kRice, kExp).
RUN_CONTEXTS = [(2, 0, 1)] * 2 + [(1, 0, 1)] * 2 + [(0, 0, 1)] + [(1, 1, 2)] * 4 + [(0, 1, 2)] * 6 + [(0, 2, 3)]
LEVEL_CONTEXTS = [(2, 0, 2), (1, 0, 1), (2, 0, 1), (0, 0, 1)] + [(0, 1, 2)] * 4 + [(0, 2, 3)]


def _distinct(books):
    out = []
    for b in books:
        if b not in out:
            out.append(b)
    return out


def test_entropy_tables_equal_the_bit_serial_reader():
    """pd_entropy's LDS tables (prores_entropy_tables): every prefix of the
    next 9 (run) or 10 (level + sign) bits holds the code the bit-serial
    reader (Codebook::read, src/vlc.rs:206-230) decodes from it, with its
    length and the table block of the context the symbol sets, or 0 when the
    code is longer than 9 bits."""
    level = (C.c_uint16 * (6 * 1024))()
    run = (C.c_uint16 * (6 * 512))()
    import prores_amd as P
    P.lib().prores_entropy_tables(level, run)
    runs, levels = _distinct(RUN_CONTEXTS), _distinct(LEVEL_CONTEXTS)
    assert len(runs) == 6 and len(levels) == 6
    run_block = lambda sym: runs.index(RUN_CONTEXTS[min(sym, 15)])
    level_block = lambda sym: levels.index(LEVEL_CONTEXTS[min(sym, 8)])
    covered = 0
    for b, (lq, kr, ke) in enumerate(runs):
        for p in range(512):
            r = reference_read(format(p, "09b"), lq, kr, ke)
            want = 0 if r in (None, "range") else 0x8000 | run_block(r[0]) << 10 | r[0] << 4 | r[1]
            assert run[b * 512 + p] == want, ("run", b, p, r)
            covered += want != 0
    for b, (lq, kr, ke) in enumerate(levels):
        for p in range(1024):
            bits = format(p, "010b")
            r = reference_read(bits[:9], lq, kr, ke)
            if r in (None, "range"):
                want = 0
            else:
                sym, n = r
                lvl = -(sym + 1) if bits[n] == "1" else sym + 1
                want = 0x8000 | level_block(sym) << 11 | (lvl & 127) << 4 | (n + 1)
            assert level[b * 1024 + p] == want, ("level", b, p, r)
            covered += want != 0
    assert covered > 8000


# ---------------------------------------------------------------------------
# pd_slice_table (csrc/decode_kernels.hip): a workgroup per chunk of 256
# slices. Thread t sums table entries t, t + 256, ... before the chunk; the 64
# lanes of each wave add their sums (xor shuffles); the four wave sums are the
# chunk's carry, and the chunk holding the table's last entry checks
# data0 + carry + its own sizes against picture_size (the reference's walk,
# src/decode.rs:202-221, errs at the first slice that ends past it; with
# sizes >= 0 that is "the last end passes picture_size").


def chunk_verdict(sizes, data0, picture_size, bits):
    """The last chunk's bounds check with the sums kept in `bits`-bit words
    (the kernel's per-thread, per-wave and carry words)."""
    mask = (1 << bits) - 1
    n = len(sizes)
    k0 = ((n - 1) // 256) * 256
    sizes = np.asarray(sizes, dtype=np.uint64)
    before = sizes[:k0]
    pad = (-len(before)) % 256
    per_thread = np.concatenate([before, np.zeros(pad, np.uint64)]).reshape(-1, 256).sum(axis=0)
    per_thread = [int(v) & mask for v in per_thread]
    carry = 0
    for w in range(4):
        s = 0
        for t in range(64 * w, 64 * w + 64):
            s = (s + per_thread[t]) & mask
        carry += s  # the four wave sums are added in 64 bits
    total = int(sizes[k0:].sum())
    return data0 + carry + total > picture_size  # True: PIC_TABLE_ERR


def reference_walk(sizes, data0, picture_size):
    """src/decode.rs:211-221: True when some slice runs past picture_size."""
    at = data0
    for s in sizes:
        at += int(s)
        if at > picture_size:
            return True
    return False


def crafted_wrap_table(n=266_240):
    """Sizes whose per-wave sums before the last chunk are exactly 2^32 each
    and whose last chunk is empty: 32-bit sums wrap to a carry of 0."""
    k0 = ((n - 1) // 256) * 256
    sizes = np.zeros(n, np.uint64)
    idx = np.arange(k0)
    for w in range(4):
        sel = idx[(idx % 256) // 64 == w]
        base, rem = divmod(1 << 32, len(sel))
        assert base + 1 <= 65535
        sizes[sel] = base
        sizes[sel[:rem]] += 1
    return sizes


def test_slice_table_carry_does_not_wrap():
    sizes = crafted_wrap_table()
    data0 = 8 + 2 * len(sizes)
    picture_size = data0 + 64
    assert reference_walk(sizes, data0, picture_size)
    # 32-bit sums would pass the check (the bug); the kernel's 64-bit sums fail it.
    assert not chunk_verdict(sizes, data0, picture_size, 32)
    assert chunk_verdict(sizes, data0, picture_size, 64)
    # Ordinary tables: the 64-bit verdict equals the reference's walk.
    rng = np.random.default_rng(5)
    for n in (1, 255, 256, 257, 4050, 70_000):
        s = rng.integers(0, 65536, n).astype(np.uint64)
        total = int(s.sum())
        for ps in (data0, total + data0 - 1, total + data0, total + data0 + 5):
            assert chunk_verdict(s, data0, ps, 64) == reference_walk(s, data0, ps)
