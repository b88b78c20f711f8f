// gfx950 decode kernels for the ProRes slice engine.
//
// Five launches per batch of pictures (and, for a batch with packets the
// caller holds only in device memory, a sixth before them, pd_plan, that
// parses the frame and picture headers and lays out the batch on the device):
// Packet and slice offsets are absolute device addresses.
//   pd_plan         one workgroup: header parse and batch layout (above).
//   pd_slice_table  one workgroup per 256 slices of a picture: the u16
//                   big-endian slice-size table becomes byte offsets (an
//                   exclusive prefix), plus the bounds check of the table walk
//                   (src/decode.rs:202-221).
//   pd_entropy_order one workgroup: the entropy items ordered by cost.
//   pd_entropy      one lane per colour component of a slice: the RDD 36
//                   adaptive Rice/exp-Golomb DC and run/level decode
//                   (src/decode.rs:484-540, src/vlc.rs) into a coefficient
//                   arena in scanned order.
//   pd_alpha        one wavefront per slice with alpha: the run/difference
//                   decode from an LDS copy of the stream, straight to the
//                   output plane (src/decode.rs:570-595, 365-389).
//   pd_idct         one workgroup per slice: dequantisation, the separable f32
//                   IDCT in the reference's operation order (src/dct.rs:36-66,
//                   src/dsp/scalar.rs:8-12) and clamped planar u16 stores
//                   (src/decode.rs:395-438).
//
// Built with -ffp-contract=off: every product and sum is its own IEEE op.
#include <hip/hip_runtime.h>

#include "dct_basis.h"
#include "layout.h"

using namespace prores_hip;

namespace {

constexpr uint32_t MAX_SYMBOL = 1u << 30;
constexpr uint32_t MAX_PREFIX = 32;

// Codebook parameters as (last_q | k_rice << 4 | k_exp << 8); an exp-Golomb
// code of order k is the combination code (0, k, k + 1). Tables 9-11 of
// RDD 36 (src/vlc.rs:30-68): the DC books here, the run and level books in
// run_book_params / level_book_params and the host's tables.
__constant__ uint16_t DC_BOOK[4] = {0x100, 0x210, 0x321, 0x430};
constexpr uint32_t FIRST_DC_BOOK = 0x650;  // EG(5)

__device__ __forceinline__ uint32_t clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
__device__ __forceinline__ uint32_t qscale(uint32_t q) { return q <= 128 ? q : 128 + 4 * (q - 128); }
__device__ __forceinline__ int32_t symbol_to_signed(uint32_t s) {
  return (s & 1) == 0 ? (int32_t)(s >> 1) : -(int32_t)((s >> 1) + 1);
}

// MSB-first reader over one stream of `len` bytes at `base`. The window is
// two big-endian 64-bit words A|B and a bit offset into A (128 - off >= 65
// bits held, so a 64-bit window is always whole). The next two words wait in
// q0/q1, and the two after them are in flight as one raw 16-byte load that is
// byte-swapped only when q0/q1 are refilled, a 16-byte stretch of the stream
// after it was issued: the load's latency hides behind that decode, and no
// register holding a pending load is ever copied.
//
// window() holds whatever bytes follow the stream in memory. A code read
// from it is accepted only when it ends inside the stream, so those bytes
// never reach a result; the general reader uses exact_window(), in which
// bits past the end of the stream read as zero, as the reference's 64-bit
// peek does (src/bits.rs:30-40).
typedef const __attribute__((address_space(1))) uint64_t gword_t;
// Global memory reached through an address held as an integer (a record,
// an LDS list entry, a readlane): a generic (flat) access would also count
// on the LDS counter, and every LDS wait after it would wait for it too.
#define GLOBAL __attribute__((address_space(1)))

struct Reader {
  gword_t *src;
  uint64_t A, B, q0, q1, r0, r1;  // r0/r1: raw words rj, rj + 1
  uint32_t off, half, rj, maxw, pos, lenbits;

  __device__ __forceinline__ static uint64_t fix(uint64_t raw) { return __builtin_bswap64(raw); }
  // Words j and j + 1 (j clamped to the last word holding stream bytes; the
  // caller's buffer has 64 readable bytes past every packet).
  __device__ __forceinline__ void issue(uint32_t j) {
    const uint32_t c = j < maxw ? j : maxw;
    r0 = src[c];
    r1 = src[c + 1];
  }
  __device__ __forceinline__ void init(const GLOBAL uint8_t *base, uint32_t len) {
    const uintptr_t a = (uintptr_t)base;
    const uint32_t m = (uint32_t)(a & 7);
    src = (gword_t *)(a - m);
    maxw = (m + len + 7) / 8;
    maxw = maxw ? maxw - 1 : 0;
    issue(0);
    A = fix(r0);
    B = fix(r1);
    issue(2);
    q0 = fix(r0);
    q1 = fix(r1);
    rj = 4;
    issue(4);
    half = 0;
    off = 8 * m;
    pos = 0;
    lenbits = 8 * len;
  }
  __device__ __forceinline__ uint64_t window() const {
    return off ? (A << off) | (B >> (64 - off)) : A;
  }
  __device__ __forceinline__ uint64_t exact_window() const {
    const uint32_t r = remaining();
    const uint64_t w = window();
    return r >= 64 ? w : (r == 0 ? 0ull : w & (~0ull << (64 - r)));
  }
  __device__ __forceinline__ uint32_t remaining() const { return lenbits - pos; }
  __device__ __forceinline__ void consume(uint32_t n) {  // n <= 64
    pos += n;
    off += n;
    if (off >= 64) {
      off -= 64;
      A = B;
      if (!half) {
        B = q0;
        half = 1;
      } else {
        B = q1;
        q0 = fix(r0);
        q1 = fix(r1);
        rj += 2;
        issue(rj);
        half = 0;
      }
    }
  }
  __device__ __forceinline__ uint32_t take(uint32_t n) {  // n <= 32
    const uint32_t v = n ? (uint32_t)(exact_window() >> (64 - n)) : 0u;
    consume(n);
    return v;
  }
  // RDD 36 endOfData(): 31 or fewer bits remain and all are zero (the top r
  // bits of the window are the stream's last r bits).
  __device__ __forceinline__ bool end_of_data() const {
    const uint32_t r = remaining();
    return r <= 31 && (r == 0 || (window() >> (64 - r)) == 0);
  }
};

// Codebook::read with the reference's checks, in its order
// (src/vlc.rs:206-230 over src/bits.rs:59-94).
__device__ __forceinline__ uint32_t read_code(Reader &r, uint32_t book, uint32_t &sym) {
  const uint32_t last_q = book & 15, k_rice = (book >> 4) & 15, k_exp = book >> 8;
  const uint32_t lz = clz64(r.exact_window());
  if (lz > MAX_PREFIX) return E_PREFIX;
  if (lz >= r.remaining()) return E_PAST_END;
  r.consume(lz);
  if (lz <= last_q) {
    const uint32_t n = 1 + k_rice;
    if (n > r.remaining()) return E_PAST_END;
    const uint32_t v = r.take(n) & ((1u << k_rice) - 1);
    sym = (lz << k_rice) | v;
    return E_NONE;
  }
  const uint32_t n = lz - (last_q + 1) + k_exp + 1;  // <= 39
  if (n > r.remaining()) return E_PAST_END;
  const uint64_t v = r.exact_window() >> (64 - n);
  r.consume(n);
  const uint64_t val = v - (1ull << k_exp) + ((uint64_t)(last_q + 1) << k_rice);
  if (val > MAX_SYMBOL) return E_RANGE;
  sym = (uint32_t)val;
  return E_NONE;
}

// A 16-byte store marked non-temporal: a 64-byte piece of a 128-byte line
// written that way goes past the L2 without the L2 first reading the line's
// other half (which an ordinary store makes it do).
typedef uint32_t nt_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt(GLOBAL uint4 *p, uint4 v) {
  const nt_u32x4 x = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(x, (GLOBAL nt_u32x4 *)p);
}

// Writes one component's scanned coefficients as i16, in increasing index
// order, through a 64-coefficient ring of LDS per lane: half-line h (the 32
// coefficients from 32 h) is kept in half h & 1 of the lane's line, and a
// coefficient is one LDS store at its index mod 64. A lane that moves to a
// later half-line leaves the finished one pending; the wave writes out the
// pending half-lines of all its lanes together: each lane lists its own in
// the wave's LDS list, then four lanes copy each listed half-line as one
// 64-byte store (non-temporal), sixteen a store instruction, clearing it in
// LDS. Zeros between are written too, so the arena needs no clearing. A value outside
// +-32767 is stored as -32768 with the full value beside it in the i32 side
// arena. The *_all calls are made by every lane of the wave together.
struct CoefOut {
  static constexpr uint32_t NONE = 0xFFFFFFFFu;
  GLOBAL int16_t *base;
  GLOBAL int32_t *wide;
  uint32_t *lds;   // the workgroup's lines (LINE_WORDS dwords a lane)
  uint4 *list;     // the wave's 64 list entries
  uint32_t line;   // this lane's line: dword index into lds
  uint32_t cur;    // the half-line being filled
  uint32_t pend;   // the pending half-line, or NONE
  uint32_t *any_wide;  // the batch's flag: set when a value goes to the side arena
  uint32_t spare;      // u16 index in lds of a slot never written out: this lane's list entry
  __device__ __forceinline__ void clear() {
    uint4 *l = (uint4 *)(lds + line);
#pragma unroll
    for (uint32_t i = 0; i < 8; i++) l[i] = make_uint4(0, 0, 0, 0);
  }
  __device__ __forceinline__ void zero_half(uint32_t h) {
    GLOBAL uint4 *o = (GLOBAL uint4 *)(base + 32 * h);
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) store_nt(&o[i], make_uint4(0, 0, 0, 0));
  }
  // Coefficient n, of the half-line being filled or the next one.
  __device__ __forceinline__ void store(uint32_t n, int32_t v) {
    int32_t h = v;
    if (v > 32767 || v < -32767) {
      h = -32768;
      wide[n] = v;
      *any_wide = 1;
    }
    ((uint16_t *)(lds + line))[n & 63] = (uint16_t)(int16_t)h;
  }
  // A value of magnitude below 32768 at ring slot `slot`, or nowhere (the
  // spare slot) when `keep` is false.
  __device__ __forceinline__ void store_slot(bool keep, uint32_t slot, int32_t v) {
    ((uint16_t *)lds)[keep ? 2 * line + slot : spare] = (uint16_t)(int16_t)v;
  }
  // Every lane's pending half-line to the arena; its half of the ring cleared.
  __device__ __forceinline__ void flush_pending_all() {
    const uint32_t L = threadIdx.x & 63;
    const bool has = pend != NONE;
    const uint64_t m = __ballot(has);
    if (!m) return;
    const uint32_t cnt = (uint32_t)__popcll(m);
    if (has) {
      const uint32_t slot = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      const uint64_t dst = (uint64_t)(uintptr_t)(base + 32 * pend);
      list[slot] = make_uint4(line + 16 * (pend & 1), 0u, (uint32_t)dst, (uint32_t)(dst >> 32));
      pend = NONE;
    }
    for (uint32_t r = 0; r < cnt; r += 16) {
      const uint32_t e = r + (L >> 2);
      if (e < cnt) {
        const uint4 ent = list[e];
        uint4 *src = (uint4 *)(lds + ent.x) + (L & 3);
        const uint4 v = *src;
        *src = make_uint4(0, 0, 0, 0);
        store_nt(&((GLOBAL uint4 *)(uintptr_t)(((uint64_t)ent.w << 32) | ent.z))[L & 3], v);
      }
    }
  }
  // One coefficient (`have`) of a lane that may move past the next
  // half-line: the half-lines it leaves go out first, the ones it skips are
  // written as zeros, then it stores the coefficient.
  __device__ __forceinline__ void put_all(bool have, uint32_t n, int32_t v) {
    const uint32_t h = n >> 5;
    const bool moves = have && h != cur;
    if (__ballot(moves)) {
      flush_pending_all();
      if (moves) pend = cur;
      flush_pending_all();
      if (moves) {
        for (uint32_t g = cur + 1; g < h; g++) zero_half(g);
        cur = h;
      }
    }
    if (have) store(n, v);
  }
  // The end of the component for lanes with `done`: the pending half-line,
  // the one being filled, then zeros up to `nhalves`, sixteen half-lines a
  // wave store.
  __device__ __forceinline__ void finish_all(bool done, uint32_t nhalves) {
    const uint32_t L = threadIdx.x & 63;
    flush_pending_all();
    if (done) pend = cur;
    flush_pending_all();
    uint64_t m = __ballot(done && cur + 1 < nhalves);
    while (m) {
      const uint32_t j = (uint32_t)__builtin_ctzll(m);
      m &= m - 1;
      const uint32_t first = lane_u32(cur, j) + 1;
      const uint32_t count = lane_u32(nhalves, j) - first;
      GLOBAL int16_t *const b = lane_base(j);
      for (uint32_t k = L >> 2; k < count; k += 16)
        store_nt(&((GLOBAL uint4 *)(b + 32 * (first + k)))[L & 3], make_uint4(0, 0, 0, 0));
    }
  }
  // Lane j's value of a 32-bit lane variable, and of `base` (j uniform).
  // readlane returns int: zero-extended here, never sign-extended into an
  // address.
  __device__ __forceinline__ static uint32_t lane_u32(uint32_t v, uint32_t j) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j);
  }
  __device__ __forceinline__ GLOBAL int16_t *lane_base(uint32_t j) const {
    const uint64_t lo = lane_u32((uint32_t)(uintptr_t)base, j);
    const uint64_t hi = lane_u32((uint32_t)((uintptr_t)base >> 32), j);
    return (GLOBAL int16_t *)(uintptr_t)((hi << 32) | lo);
  }
};

// The DC codebooks for the closed-form decode, 4 bits per context:
// lastRiceQ in bits 0-1, kRice in bits 2-3; kExp = kRice + 1.
constexpr uint32_t DC_NIB = 0xC940u;  // contexts 0..3

// A combination code (lq, kr, ke) from the top of a 32-bit window `top`
// (RDD 36 §7.1.1.1): q leading zeros; if q <= lq, a Rice code of q + 1 + kr
// bits with value (q << kr) + r; else an exp-Golomb code of 2q - lq + ke bits
// whose trailing 1 + (q - lq - 1) + ke bits v give v - 2^ke + ((lq + 1) << kr).
// Both read as the code's leading bits p minus a base: p - 2^k + (q' << kr).
// True when the code lies in the 32 bits and its value is at most 2^30; the
// general reader takes every other case and reports its error.
__device__ __forceinline__ bool fast_code(uint32_t top, uint32_t lq, uint32_t kr, uint32_t ke,
                                          uint32_t &sym, uint32_t &len) {
  const uint32_t q = top ? (uint32_t)__builtin_clz(top) : 32u;
  const bool rice = q <= lq;
  const uint32_t L = rice ? q + 1 + kr : 2 * q - lq + ke;
  const uint32_t p = top >> ((32 - L) & 31);
  const uint32_t k = rice ? kr : ke;
  const uint32_t qq = rice ? q : lq + 1;
  sym = p - (1u << k) + (qq << kr);
  len = L;
  return L <= 32 && sym <= MAX_SYMBOL;
}

// scanned_coefficients() of one colour component (src/decode.rs:484-540),
// called by every lane of the wave together; lanes with `act` false (no
// stream, or a stream already in error) only help with the stores.
//
// In each step of the AC loop every lane decodes run/level/sign pairs from
// its 64-bit window
// through the LDS tables `ftab` (codes of at most 9 bits; the next 9 bits
// of the stream index the book of the current context), as long as the
// pair lies inside the stream, its coefficient falls in the half-line being
// filled or, once a step, the next one, and the window holds another
// pair's worth. Such a pair is the reference's next pair: every code holds
// a 1 bit, so endOfData (at most 31 bits left, all zero) cannot hold before
// a code that ends inside the stream, and a code of at most 9 bits passes
// every other check. A lane that decodes no pair from the tables (a longer
// code, the end of the stream, a run past a half-line, an error) decodes one
// pair in closed form or with the general reader, check for check, in the
// same step. The step ends with the wave writing out its pending half-lines.
// Table pairs per step at most. A wave steps until its last lane stops, so
// a longer step keeps lanes that stopped early waiting longer; measured on
// the 2160p 4444 XQ batch (entropy alone): 8 pairs 582 us, 6: 590, 10: 587,
// 12: 601, 16: 640.
constexpr uint32_t FAST_PAIRS = 8;
constexpr uint32_t FAST_START = 45;  // a pair needs 9 + 10 bits of the 64-bit window
// The table book of a run or level context, 4 bits per context: run books
// sit 1024 bytes apart after FAST_RUN_BASE, level books 2048 bytes apart.
constexpr uint64_t RUN_CTX_NIB = 0x5444444333321100ull;  // contexts 0..15
constexpr uint64_t LVL_CTX_NIB = 0x544443210ull;         // contexts 0..8
__device__ __forceinline__ uint32_t run_book_off(uint32_t prev) {
  const uint32_t c = prev < 15 ? prev : 15;
  return (uint32_t)((RUN_CTX_NIB >> (4 * c)) & 15) << 10;
}
__device__ __forceinline__ uint32_t level_book_off(uint32_t prev) {
  const uint32_t c = prev < 8 ? prev : 8;
  return (uint32_t)((LVL_CTX_NIB >> (4 * c)) & 15) << 11;
}
// Codebook parameters (lastRiceQ | kRice << 4 | kExp << 8) of a book.
__device__ __forceinline__ uint32_t run_book_params(uint32_t off) {
  const uint32_t b = off >> 10;
  return b == 0 ? 0x102u : b == 1 ? 0x101u : b == 2 ? 0x100u : b == 3 ? 0x211u : b == 4 ? 0x210u : 0x320u;
}
__device__ __forceinline__ uint32_t level_book_params(uint32_t off) {
  const uint32_t b = off >> 11;
  return b == 0 ? 0x202u : b == 1 ? 0x101u : b == 2 ? 0x102u : b == 3 ? 0x100u : b == 4 ? 0x210u : 0x320u;
}

__device__ uint32_t decode_component(Reader &r, uint32_t nb, CoefOut &out, bool act, const uint16_t *ftab) {
  uint32_t e = E_NONE;
  const uint32_t total = nb * 64;
  uint32_t n = nb;
  if (act) {
    // The DC coefficients: nb <= 32, all in half-line 0.
    uint32_t sym = 0;
    e = read_code(r, FIRST_DC_BOOK, sym);
    int32_t prev_dc = symbol_to_signed(sym);
    if (!e) out.store(0, prev_dc);
    int32_t prev_diff = 3;
    for (uint32_t i = 1; i < nb && !e; i++) {
      const uint32_t a = prev_diff < 0 ? (uint32_t)(-prev_diff) : (uint32_t)prev_diff;
      const uint32_t c = a < 3 ? a : 3;
      const uint32_t nib = (DC_NIB >> (4 * c)) & 15;
      const uint32_t top = (uint32_t)(r.window() >> 32);
      uint32_t len;
      if (fast_code(top, nib & 3, nib >> 2, (nib >> 2) + 1, sym, len) && len <= r.remaining()) {
        r.consume(len);
      } else {
        e = read_code(r, DC_BOOK[c], sym);
        if (e) break;
      }
      const int32_t v = symbol_to_signed(sym);
      const int32_t diff = prev_diff < 0 ? -v : v;
      const int64_t sum = (int64_t)prev_dc + diff;
      if (sum > INT32_MAX || sum < INT32_MIN) {
        e = E_DC_OVERFLOW;
        break;
      }
      prev_dc = (int32_t)sum;
      out.store(i, prev_dc);
      prev_diff = diff;
    }
    act = !e;
  }
  const bool started = act;
  // The books of the current run and level contexts (prev_run 4, prev_lvl 1).
  uint32_t rbo = run_book_off(4), lbo = level_book_off(1);
  while (__ballot(act)) {
    // Pairs from the tables. Every lane's half-lines before `cur` are out,
    // so the pairs may run up to the end of the next half-line.
    bool moved = false, blocked = false;
    if (act) {
      // The chain runs on speculative copies (window, contexts, position,
      // bits used), advanced by every pair whatever it finds: an entry past
      // the tables is 0 (book 0, no bits), so a lane that has stopped keeps
      // reading inside the tables. A pair commits only while its lane has
      // not stopped, off the chain.
      uint64_t ws = r.window();
      const uint32_t rem = r.remaining();
      const uint32_t lim0 = 32 * (out.cur + 2), lim = lim0 < total ? lim0 : total;
      uint32_t used = 0, us = 0, rs = rbo, ls = lbo, ns = n;
      bool go = true;
#pragma unroll
      for (uint32_t k = 0; k < FAST_PAIRS; k++) {
        if (k == FAST_PAIRS / 2 && !__ballot(go)) break;
        const uint32_t hi = (uint32_t)(ws >> 32), lo = (uint32_t)ws;
        const uint32_t er =
            *(const uint16_t *)((const char *)ftab + FAST_RUN_BASE + (rs | ((hi >> 22) & 0x3FE)));
        const uint32_t lr = er & 15;
        const uint32_t hi2 = __builtin_amdgcn_alignbit(hi, lo, 32 - lr);  // ws << lr, top word
        const uint32_t el = *(const uint16_t *)((const char *)ftab + (ls | ((hi2 >> 21) & 0x7FE)));
        const uint32_t t = lr + (el & 15), m = ns + ((er >> 4) & 63);
        ws <<= t;
        rs = er & 0x1C00;
        ls = el & 0x3800;
        us += t;
        ns = m + 1;
        const bool ok = go && (er & el & FAST_VALID) != 0 && us <= rem && m < lim;
        blocked = blocked || (go && !ok);
        out.store_slot(ok, m & 63, ((int32_t)(el << 21)) >> 25);
        n = ok ? ns : n;
        used = ok ? us : used;
        rbo = ok ? rs : rbo;
        lbo = ok ? ls : lbo;
        moved = moved || ok;
        go = ok && us <= FAST_START;
      }
      if (moved) {
        r.consume(used);
        const uint32_t c = (n - 1) >> 5;
        if (c != out.cur) {
          out.pend = out.cur;
          out.cur = c;
        }
      }
    }
    // One pair the reference's way for lanes the tables stopped on (a code
    // past the tables, a run past the ring, the stream's last bits), in the
    // same step: a lane with many long codes does not take a step for each.
    const bool slow = act && (!moved || blocked);
    if (__ballot(slow)) {
      bool have = false;
      uint32_t at = 0;
      int32_t val = 0;
      if (slow) {
        const uint64_t w = r.window();
        const uint32_t top = (uint32_t)(w >> 32);
        const uint32_t rem = r.remaining();
        if (top == 0 && rem <= 31) {  // endOfData
          act = false;
        } else {
          const uint32_t rb = run_book_params(rbo), lb = level_book_params(lbo);
          uint32_t run, l1, lsym, l2;
          const bool okr = fast_code(top, rb & 15, (rb >> 4) & 15, rb >> 8, run, l1);
          const uint64_t w2 = w << (l1 & 63);
          const bool okl = fast_code((uint32_t)(w2 >> 32), lb & 15, (lb >> 4) & 15, lb >> 8, lsym, l2);
          const uint32_t T = l1 + l2 + 1;
          if (okr && okl && T <= 64 && T <= rem) {
            n += run;
            if (n >= total) {
              e = E_RUN_PAST;
              act = false;
            } else {
              const bool neg = ((w2 >> (63 - l2)) & 1) != 0;
              const int32_t level = (int32_t)lsym + 1;
              have = true;
              at = n;
              val = neg ? -level : level;
              n++;
              rbo = run_book_off(run);
              lbo = level_book_off(lsym);
              r.consume(T);
            }
          } else if (r.end_of_data()) {
            act = false;
          } else {
            e = read_code(r, rb, run);
            if (!e) {
              rbo = run_book_off(run);
              n += run;
              if (n >= total) e = E_RUN_PAST;
            }
            if (!e) {
              e = read_code(r, lb, lsym);
              if (!e) lbo = level_book_off(lsym);
            }
            if (!e && r.remaining() < 1) e = E_PAST_END;
            if (!e) {
              const uint32_t sgn = r.take(1);
              const int32_t level = (int32_t)lsym + 1;
              have = true;
              at = n;
              val = sgn ? -level : level;
              n++;
            }
            if (e) act = false;
          }
        }
      }
      out.put_all(have, at, val);
    }
    // The half-lines finished this step go out last: vmcnt counts stores
    // as well as loads, so a bitstream refill waits for every store issued
    // before it, and these have the next step's table decode to complete.
    out.flush_pending_all();
  }
  out.finish_all(started && !e, 2 * nb);
  return e;
}

// The slice header (src/decode.rs:301-338): header size, component sizes
// and their checks. Returns the error word, E_NONE when the header is sound.
__device__ __forceinline__ uint32_t slice_header(const GLOBAL uint8_t *s, uint32_t slen, uint32_t alpha_bits,
                                                 uint32_t &hs, uint32_t &ysz, uint32_t &cbsz,
                                                 uint32_t &crsz) {
  const uint32_t min_header = alpha_bits ? 8u : 6u;
  hs = ysz = cbsz = crsz = 0;
  if (slen < min_header) return E_SLICE_SHORT;
  hs = s[0] >> 3;
  const uint32_t q = s[1];
  if (hs < min_header || hs > slen) return E_SLICE_HDR_SIZE | (hs << 8);
  if (q < 1 || q > 224) return E_QINDEX | (q << 8);
  ysz = ((uint32_t)s[2] << 8) | s[3];
  cbsz = ((uint32_t)s[4] << 8) | s[5];
  const uint32_t body = slen - hs;
  if (alpha_bits) {
    crsz = ((uint32_t)s[6] << 8) | s[7];
  } else {
    if (ysz + cbsz > body) return E_COMP_SIZES;
    crsz = body - ysz - cbsz;
  }
  if (ysz + cbsz + crsz > body) return E_COMP_SIZES;
  return E_NONE;
}

__device__ __forceinline__ uint16_t rescale(uint32_t a, uint32_t identity, uint64_t m) {
  return identity ? (uint16_t)a : (uint16_t)(((uint64_t)a * m + (1ull << 39)) >> 40);
}

// Work distribution. In grid form each workgroup (or wave) takes the item
// its index names; in persistent form a grid sized to the device takes items
// from the batch's queue, `chunk` at a time, until the queue is empty.
__device__ __forceinline__ uint32_t wave_pop(uint32_t *head, uint32_t chunk) {
  uint32_t v = 0;
  if ((threadIdx.x & 63) == 0) v = atomicAdd(head, chunk);
  return __shfl(v, 0, 64);
}

// One colour stream: component `comp` of slice `g`.
// One colour stream per lane, every lane of the wave together; `valid`
// lanes hold a slice.
__device__ __forceinline__ void entropy_stream(
    bool valid, uint32_t g, uint32_t comp, uint32_t *lds, uint4 *list, uint32_t *any_wide,
    const uint16_t *ftab, const DecFrame *__restrict__ frames,
    const DecPic *__restrict__ pics, const SliceRec *__restrict__ slices, const uint32_t *__restrict__ pic_status,
    int16_t *__restrict__ coefs, int32_t *__restrict__ coef_wide, uint32_t *__restrict__ slice_err,
    uint32_t *__restrict__ pic_first_err) {
  SliceRec S{};
  uint32_t err = E_NONE, nb = 0, slice_base = 0;
  bool act = false;
  Reader r;
  CoefOut out{(GLOBAL int16_t *)coefs, (GLOBAL int32_t *)coef_wide, lds, list, LINE_WORDS * threadIdx.x, 0,
              CoefOut::NONE, any_wide,
              2 * (uint32_t)((uint32_t *)(list + (threadIdx.x & 63)) - lds)};
  if (valid) {
    S = slices[g];
    if (pic_status[S.pic] == PIC_OK) {
      const DecPic &P = pics[S.pic];
      const DecFrame &F = frames[P.frame];
      slice_base = P.slice_base;
      const GLOBAL uint8_t *s = (const GLOBAL uint8_t *)(uintptr_t)S.off;
      uint32_t hs, ysz, cbsz, crsz;
      err = slice_header(s, S.len, F.alpha_bits, hs, ysz, cbsz, crsz);
      if (!err) {
        const uint32_t cpm = F.chroma == 2 ? 2u : 4u;
        nb = (comp == 0 ? 4u : cpm) * S.mbs;
        const uint32_t off = comp == 0 ? hs : (comp == 1 ? hs + ysz : hs + ysz + cbsz);
        const uint32_t len = comp == 0 ? ysz : (comp == 1 ? cbsz : crsz);
        const uint64_t cofs =
            S.coef + (comp == 0 ? 0ull : (comp == 1 ? 256ull * S.mbs : (256ull + 64ull * cpm) * S.mbs));
        r.init(s + off, len);
        out.base = (GLOBAL int16_t *)(coefs + cofs);
        out.wide = (GLOBAL int32_t *)(coef_wide + cofs);
        out.clear();
        act = true;
      }
    }
  }
  const uint32_t e = decode_component(r, nb, out, act, ftab);
  if (act && e) err = e;
  if (err) {
    slice_err[4 * (size_t)g + comp] = err;
    atomicMin(&pic_first_err[S.pic], g - slice_base);
  }
}

// One slice's alpha (see pd_alpha).
constexpr uint32_t ALPHA_WORDS = (2048 * 33 / 8) / 8 + 6;

// A value the whole wave holds, as a scalar.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Alpha streams of at least this many bytes are parsed by alpha_parallel.
constexpr uint32_t ALPHA_PAR_BYTES = 64;
// No pair of a slice's decode starts at or after this bit: 2048 values take
// at most 2048 pairs of at most 33 bits.
constexpr uint32_t ALPHA_MAX_BITS = 2048 * 33;

// 64 bits of the staged alpha stream from absolute bit q, MSB first; at
// least the first 33 are stream bits. The staged u64 words are big-endian
// values in little-endian LDS, so 32-bit chunk k of the stream is u32 k ^ 1.
__device__ __forceinline__ uint64_t alpha_window(const uint32_t *w32, uint32_t q) {
  const uint32_t c = q >> 5;
  return ((uint64_t)w32[c ^ 1] << 32 | w32[(c + 1) ^ 1]) << (q & 31);
}

// The length of the (difference, run) pair at the head of window x: a
// difference of 1 + bits (raw) or 2 + mag (magnitude and sign), then a run
// of 1, 5 or 16 bits (src/vlc.rs:263-299). Needs the first 22 bits.
__device__ __forceinline__ uint32_t alpha_pair_len(uint64_t x, uint32_t bits, uint32_t mag) {
  const uint32_t hi = (uint32_t)(x >> 32);
  const uint32_t dl = (hi >> 31) ? 1 + bits : 2 + mag;
  const uint32_t r = (hi << dl) >> 27;  // the run code's first 5 bits
  return dl + ((r & 16) ? 1u : ((r & 15) ? 5u : 16u));
}

// The pair at the head of window x (every bit of it inside the window).
__device__ __forceinline__ void alpha_pair(uint64_t x, uint32_t bits, uint32_t mag, int32_t &d, uint32_t &run) {
  const uint32_t used = (x >> 63) ? 1 + bits : 2 + mag;
  const int32_t av = (int32_t)((x >> (63 - mag)) & ((1u << mag) - 1)) + 1;
  const int32_t small = ((x >> (62 - mag)) & 1) ? -av : av;
  d = (x >> 63) ? (int32_t)((x >> (63 - bits)) & ((1u << bits) - 1)) : small;
  const uint64_t x2 = x << used;
  const uint32_t v = (uint32_t)(x2 >> 59) & 15;
  run = (x2 >> 63) ? 1u : (v ? v + 1 : (uint32_t)((x2 >> 48) & 0x7FF) + 1);
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
  const uint32_t L = threadIdx.x & 63;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, 64);
    if (L >= o) v += t;
  }
  return v;
}

// The pairs of a long alpha stream that start at stream bits 0..F (each has
// all of its bits in the stream, so the reference reads them unchecked;
// src/decode.rs:570-595), with the whole wave. For each 64-bit window of the
// stream every lane computes the pair length at one bit position, and a
// scalar walk over those lengths marks the window's pair starts (a few
// scalar steps a pair instead of the whole decode). Every 64 windows the
// lanes decode the marked pairs, one window a lane: the run and difference
// sums, wave prefix sums, then each pair's value at its first output
// position, flagged in `marks` (alpha_fill spreads the runs). Stops where
// the reference's loop does: at the first pair that starts at or past
// `count` values, or at a run past the end (E_ALPHA_RUN). Returns true when
// the decode ended here (`err` says how); otherwise `pos` is the first pair
// start after F, where the reference's checked reads take over, with `n`
// values decoded and `prev` the last value.
__device__ bool alpha_parallel(const uint32_t *w32, uint32_t base, uint32_t F, uint32_t bits, uint32_t mag,
                               uint32_t count, uint32_t ident, uint64_t rm, uint16_t *vals, uint32_t *marks,
                               uint32_t &pos, uint32_t &n, int32_t &prev, uint32_t &err) {
  const uint32_t L = threadIdx.x & 63, mask = (1u << bits) - 1;
  uint32_t w = 0, c = 0;  // the walk: the next pair starts at bit c of window w
  for (;;) {
    const uint32_t w0 = w;
    uint64_t gm = 0;  // lane j: the pair starts of window w0 + j
    uint32_t nw = 0;
    bool tail = false;
    while (nw < 64) {
      const uint32_t pl = alpha_pair_len(alpha_window(w32, base + 64 * w + L), bits, mag);
      const uint32_t room = F - 64 * w + 1;  // 64 * w <= F here
      const uint32_t lim = room < 64 ? room : 64u;
      uint64_t mk = 0;
      while (c < lim) {
        mk |= 1ull << c;
        c += (uint32_t)__builtin_amdgcn_readlane((int)pl, (int)c);
      }
      gm = L == nw ? mk : gm;
      nw++;
      if (c < 64) {  // the next pair starts past F inside this window
        tail = true;
        break;
      }
      c -= 64;
      w++;
      if (64 * w > F) {
        tail = true;
        break;
      }
    }
    // The group's pairs, one window a lane.
    const uint64_t mine = L < nw ? gm : 0ull;
    const uint32_t at = base + 64 * (w0 + L);
    uint32_t rs = 0, ds = 0;
    for (uint64_t t = mine; t; t &= t - 1) {
      int32_t d;
      uint32_t run;
      alpha_pair(alpha_window(w32, at + (uint32_t)__builtin_ctzll(t)), bits, mag, d, run);
      rs += run;
      ds += (uint32_t)d;
    }
    const uint32_t ri = wave_incl_sum(rs), di = wave_incl_sum(ds);
    uint32_t nn = n + ri - rs, acc = (uint32_t)prev + di - ds;
    bool stop = false, bad = false;
    for (uint64_t t = mine; t; t &= t - 1) {
      if (nn >= count) {
        stop = true;
        break;
      }
      int32_t d;
      uint32_t run;
      alpha_pair(alpha_window(w32, at + (uint32_t)__builtin_ctzll(t)), bits, mag, d, run);
      acc += (uint32_t)d;
      if (nn + run > count) {
        bad = true;
        break;
      }
      vals[nn] = rescale(acc & mask, ident, rm);
      atomicOr(&marks[nn >> 5], 1u << (nn & 31));
      nn += run;
    }
    if (__ballot(bad)) {
      err = E_ALPHA_RUN;
      return true;
    }
    if (__ballot(stop)) return true;
    n += uni((uint32_t)__builtin_amdgcn_readlane((int)ri, 63));
    prev = (int32_t)(((uint32_t)prev + uni((uint32_t)__builtin_amdgcn_readlane((int)di, 63))) & mask);
    if (n >= count) return true;
    if (tail) {
      pos = 64 * w + c;
      return false;
    }
  }
}

// Spreads each marked value of `vals` over the positions up to the next
// mark (lane L: positions 32L..32L+31; a mark's value is only read).
__device__ __forceinline__ void alpha_fill(uint16_t *vals, const uint32_t *marks) {
  const uint32_t L = threadIdx.x & 63;
  const uint32_t mw = marks[L];
  int32_t last = mw ? (int32_t)(32 * L + 31 - __builtin_clz(mw)) : -1;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const int32_t t = __shfl_up(last, o, 64);
    if (L >= o) last = t > last ? t : last;
  }
  int32_t before = __shfl_up(last, 1, 64);
  if (L == 0) before = -1;
  uint16_t cur = before >= 0 ? vals[before] : (uint16_t)0;
  for (uint32_t k = 0; k < 32; k++) {
    const uint32_t i = 32 * L + k;
    if ((mw >> k) & 1)
      cur = vals[i];
    else
      vals[i] = cur;
  }
}

__device__ __forceinline__ void alpha_slice(
    const AlphaRec &R, uint64_t *__restrict__ words, uint16_t *__restrict__ vals, uint32_t *__restrict__ marks,
    const DecFrame *__restrict__ frames, const DecPic *__restrict__ pics, const uint32_t *__restrict__ pic_status,
    uint32_t *__restrict__ slice_err, uint32_t *__restrict__ pic_first_err) {
  const uint32_t L = threadIdx.x;
  const uint32_t g = uni(R.g), pic = uni(R.pic);
  // The picture's status is loaded with the stream and looked at once the
  // stream is staged: one round trip to memory instead of two. The record
  // names only a slice whose bytes lie inside its picture, so the staging
  // reads stay inside the packet whatever the status.
  const uint32_t status = pic_status[pic];
  const DecPic &P = pics[pic];
  const DecFrame &F = frames[uni(R.frame)];
  const uint32_t bits = uni(F.alpha_bits);
  const uint32_t mb_x = uni(R.mb) & 0xFFFF, mb_y = uni(R.mb) >> 16, mbs = uni(R.mbs);
  uint32_t err = E_NONE;
  // Stage the stream: aligned words from base - m, bytes outside the stream
  // zeroed, and two zero words after it (the reader holds the word after its
  // current one, and never passes the end of the stream).
  const uint32_t len = uni(R.len);
  const uintptr_t a = (uintptr_t)(((uint64_t)uni((uint32_t)(R.stream >> 32)) << 32) | uni((uint32_t)R.stream));
  // A short stream comes from the record (from its first byte: m = 0).
  const bool from_rec = len <= ALPHA_HEAD_BYTES;
  const uint32_t m = from_rec ? 0u : uni((uint32_t)(a & 7));
  const GLOBAL uint64_t *src = (const GLOBAL uint64_t *)(a - m);
  const uint32_t end_byte = m + len;  // bytes of stream data end here
  const uint32_t have = (end_byte + 7) / 8;
  const uint32_t stage = have + 2 < ALPHA_WORDS ? have + 2 : ALPHA_WORDS;
  if (from_rec) {
    if (L < stage) {
      const uint32_t h0 = uni(R.head[0]), h1 = uni(R.head[1]), h2 = uni(R.head[2]), h3 = uni(R.head[3]);
      uint64_t w = L == 0 ? ((uint64_t)h1 << 32 | h0) : L == 1 ? ((uint64_t)h3 << 32 | h2) : 0ull;
      w = L < have ? __builtin_bswap64(w) : 0ull;
      const uint32_t hi = 8 * L + 8;
      if (L < have && hi > end_byte) w &= ~0ull << (8 * (hi - end_byte));
      words[L] = w;
    }
  } else {
    // Four words a lane in flight at a time.
    for (uint32_t j0 = 0; j0 < stage; j0 += 256) {
      uint64_t raw[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; u++) {
        const uint32_t j = j0 + 64 * u + L;
        raw[u] = j < have && j < stage ? src[j] : 0ull;
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; u++) {
        const uint32_t j = j0 + 64 * u + L;
        uint64_t w = __builtin_bswap64(raw[u]);
        const uint32_t lo = 8 * j, hi = lo + 8;
        if (lo < m) w &= ~0ull >> (8 * (m - lo));
        if (hi > end_byte) w &= hi - end_byte >= 8 ? 0ull : ~0ull << (8 * (hi - end_byte));
        if (j < stage) words[j] = w;
      }
    }
  }
  __syncthreads();
  if (uni(status) != PIC_OK) return;

  // The parse is the same in every lane: its state lives in scalar
  // registers (each staged word is read as a wave-uniform value), and the
  // vector lanes only fill runs.
  const uint32_t mag = bits == 8 ? 3u : 6u, mask = (1u << bits) - 1;
  const uint32_t log2c = 4 + (31 - __builtin_clz(mbs));
  const uint32_t cols = 1u << log2c, count = cols * 16;
  const uint32_t rows =
      uni(mb_y + 1 < P.height_in_mb ? 16u : P.height - 16u * (P.height_in_mb - 1));
  const uint32_t needed = cols * rows;
  const uint32_t ident = uni(F.rescale_identity);
  const uint64_t rm = ((uint64_t)uni((uint32_t)(F.rescale_m >> 32)) << 32) | uni((uint32_t)F.rescale_m);
  auto word = [&](uint32_t j) -> uint64_t {
    const uint64_t w = words[j < ALPHA_WORDS ? j : ALPHA_WORDS - 1];
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)w);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(w >> 32));
    return ((uint64_t)hi << 32) | lo;
  };
  // The reader: bit `pos` of the stream is bit `off` of word `wj`.
  uint32_t wj = 0, off = 8 * m, pos = 0;
  const uint32_t lenbits = 8 * len;
  uint64_t A = word(0), B = word(1);
  auto window = [&]() -> uint64_t { return off ? (A << off) | (B >> (64 - off)) : A; };
  auto consume = [&](uint32_t n) {  // n <= 64
    pos += n;
    off += n;
    if (off >= 64) {
      off -= 64;
      wj++;
      A = B;
      B = word(wj + 1);
    }
  };
  auto take = [&](uint32_t n) -> uint32_t {  // n <= 32
    const uint32_t v = n ? (uint32_t)(window() >> (64 - n)) : 0u;
    consume(n);
    return v;
  };
  uint32_t n = 0;
  int32_t prev = -1;
  const bool par = len >= ALPHA_PAR_BYTES;
  if (par) {
    marks[L] = 0;
    const uint32_t last = lenbits - 33 < ALPHA_MAX_BITS - 1 ? lenbits - 33 : ALPHA_MAX_BITS - 1;
    uint32_t p = 0;
    if (alpha_parallel((const uint32_t *)words, 8 * m, last, bits, mag, count, ident, rm, vals, marks, p, n, prev,
                       err)) {
      n = count;  // decoded to the end (or failed: err)
    } else {
      const uint32_t q = 8 * m + p;
      wj = q >> 6;
      off = q & 63;
      pos = p;
      A = word(wj);
      B = word(wj + 1);
    }
  }
  while (n < count && !err) {
    const uint32_t rem = lenbits - pos;
    int32_t d;
    uint32_t run;
    if (rem >= 33) {
      // Both codes lie in the window and every bit is real.
      const uint64_t w = window();
      uint32_t used_bits;
      if (w >> 63) {
        d = (int32_t)((w >> (63 - bits)) & mask);
        used_bits = 1 + bits;
      } else {
        const int32_t av = (int32_t)((w >> (63 - mag)) & ((1u << mag) - 1)) + 1;
        d = ((w >> (62 - mag)) & 1) ? -av : av;
        used_bits = 2 + mag;
      }
      const uint64_t w2 = w << used_bits;
      if (w2 >> 63) {
        run = 1;
        used_bits += 1;
      } else {
        const uint32_t v = (uint32_t)(w2 >> 59) & 15;
        if (v) {
          run = v + 1;
          used_bits += 5;
        } else {
          run = (uint32_t)((w2 >> 48) & 0x7FF) + 1;
          used_bits += 16;
        }
      }
      consume(used_bits);
    } else {
      // Fewer than 33 bits left: the reference's reads, check for check.
      if (n >= needed && (rem == 0 || (window() >> (64 - rem)) == 0)) break;  // endOfData
      if (rem < 1) { err = E_PAST_END; break; }
      if (take(1)) {
        if (lenbits - pos < bits) { err = E_PAST_END; break; }
        d = (int32_t)take(bits);
      } else {
        if (lenbits - pos < mag) { err = E_PAST_END; break; }
        const int32_t av = (int32_t)take(mag) + 1;
        if (lenbits - pos < 1) { err = E_PAST_END; break; }
        d = take(1) ? -av : av;
      }
      if (lenbits - pos < 1) { err = E_PAST_END; break; }
      if (take(1)) {
        run = 1;
      } else {
        if (lenbits - pos < 4) { err = E_PAST_END; break; }
        const uint32_t v = take(4);
        if (v) {
          run = v + 1;
        } else {
          if (lenbits - pos < 11) { err = E_PAST_END; break; }
          run = take(11) + 1;
        }
      }
    }
    const uint32_t value = (uint32_t)(prev + d) & mask;
    prev = (int32_t)value;
    if (n + run > count) { err = E_ALPHA_RUN; break; }
    const uint16_t out = rescale(value, ident, rm);
    if (par) {
      if (L == 0) {
        vals[n] = out;
        atomicOr(&marks[n >> 5], 1u << (n & 31));
      }
    } else if (run <= 64) {
      if (L < run) vals[n + L] = out;
    } else {
      // A long run: to a 4-value boundary, then 8 bytes per lane.
      const uint32_t head = (4 - (n & 3)) & 3;
      if (L < head) vals[n + L] = out;
      const uint32_t a4 = n + head, quads = (run - head) / 4;
      const uint64_t v4 = (uint64_t)out * 0x0001000100010001ull;
      for (uint32_t k = L; k < quads; k += 64) ((uint64_t *)vals)[a4 / 4 + k] = v4;
      const uint32_t t0 = a4 + 4 * quads;
      if (L < n + run - t0) vals[t0 + L] = out;
    }
    n += run;
  }
  if (err) {
    if (L == 0) {
      slice_err[4 * (size_t)g + 3] = err;
      atomicMin(&pic_first_err[pic], g - P.slice_base);
    }
    return;
  }
  if (par) alpha_fill(vals, marks);
  __syncthreads();
  // Visible rows, clipped to the plane width.
  const uint32_t x0 = 16u * mb_x, width = F.width;
  const uint32_t nvis = cols < width - x0 ? cols : width - x0;
  GLOBAL uint16_t *aout = (GLOBAL uint16_t *)(uintptr_t)F.out_alpha;
  const size_t row0 = P.first_row + (size_t)P.row_step * (16u * mb_y);
  const size_t rstride = (size_t)P.row_step * width;
  if ((width & 7) == 0 && (((uintptr_t)aout) & 15) == 0) {
    // x0 and nvis are multiples of 8: 16-byte stores.
    const uint32_t per_row = nvis / 8;
    for (uint32_t c = L; c < rows * per_row; c += 64) {
      const uint32_t r = c / per_row, k = c - r * per_row;
      *(GLOBAL uint4 *)(aout + (row0 * width + r * rstride) + x0 + 8 * k) =
          *(const uint4 *)(vals + (r << log2c) + 8 * k);
    }
  } else {
    for (uint32_t i = L; i < rows * nvis; i += 64) {
      const uint32_t r = i / nvis, k = i - r * nvis;
      aout[(row0 * width + r * rstride) + x0 + k] = vals[(r << log2c) + k];
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// One workgroup per chunk of ST_CHUNK consecutive slices of a picture (a
// grid-stride loop over the chunks in persistent form), one thread a slice:
// the u16 big-endian slice-size table becomes byte offsets (an exclusive
// prefix: the sizes before the chunk summed, then a scan of the chunk's), with
// the bounds check of the table walk (src/decode.rs:202-221). It also clears
// the batch's queue heads, its slice error words and its first-error slots, so
// the launch sequence needs no memset, and writes the picture's IDCT items.
// Spread over many workgroups, the scattered slice-header reads of a batch
// are in flight on every CU at once.
extern "C" __global__ void __launch_bounds__(ST_CHUNK)
    pd_slice_table(DecBatch *__restrict__ B, const DecFrame *__restrict__ frames,
                   const DecPic *__restrict__ pics, const RowSlot *__restrict__ rows,
                   const uint2 *__restrict__ chunks, SliceRec *__restrict__ slices,
                   uint32_t *__restrict__ pic_status, uint4 *__restrict__ slice_err,
                   uint32_t *__restrict__ pic_first_err, const IdctItem *__restrict__ segs,
                   IdctItem *__restrict__ items, AlphaRec *__restrict__ alpha_order,
                   uint32_t *__restrict__ group_cost) {
  static_assert(ST_CHUNK == 256, "four waves a chunk");
  const uint32_t tid = threadIdx.x;
  if (blockIdx.x == 0 && tid < 4) B->q[tid] = 0;
  const uint32_t nc = B->n_chunks, G = B->n_slices;
  __shared__ uint32_t wsum[4];
  __shared__ uint64_t wpre[4];
  const uint32_t lane = tid & 63, wave = tid >> 6;
  for (uint32_t x = blockIdx.x; x < nc; x += gridDim.x) {
    // The chunk: its picture and its first slice (a multiple of ST_CHUNK).
    const uint2 C = chunks[x];
    const uint32_t p = C.x, k0 = C.y, k = k0 + tid;
    const DecPic P = pics[p];
    const bool has_alpha = frames[P.frame].alpha_bits != 0;
    const GLOBAL uint8_t *tab = (const GLOBAL uint8_t *)(uintptr_t)(P.pic_off + P.header_size);
    const uint64_t data0 = P.header_size + 2ull * P.n_slices;  // relative to the picture
    const bool valid = k < P.n_slices;
    // Table entries before the chunk: thread t sums entries t, t + 256, ...
    // In 64 bits: a picture of more than 65,537 slices can hold more than
    // 2^32 bytes of sizes before a chunk, and a wrapped sum could pass the
    // bounds check below.
    uint64_t pre = 0;
#pragma unroll 4
    for (uint32_t i = tid; i < k0; i += ST_CHUNK) pre += (uint32_t)tab[2 * i] << 8 | tab[2 * i + 1];
    const uint32_t sz = valid ? ((uint32_t)tab[2 * k] << 8 | tab[2 * k + 1]) : 0u;
    RowSlot rs = {0, 0, 0};
    if (valid) rs = rows[P.layout_base + (k - (k / P.row_slices) * P.row_slices)];
    uint32_t incl = sz;  // at most 256 sizes of at most 65535: no wrap
    uint64_t psum = pre;
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint32_t a = __shfl_up(incl, d, 64);
      const uint64_t b = __shfl_xor(psum, d, 64);
      if (lane >= d) incl += a;
      psum += b;
    }
    if (lane == 63) wsum[wave] = incl;
    if (lane == 0) wpre[wave] = psum;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += wsum[w];
    const uint32_t total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    const uint64_t carry = wpre[0] + wpre[1] + wpre[2] + wpre[3];
    const uint64_t at = carry + before + incl - sz;
    if (tid == 0) {
      if (k0 == 0) pic_first_err[p] = 0xFFFFFFFFu;
      // The chunk that holds the table's last entry checks its end.
      if (k0 + ST_CHUNK >= P.n_slices)
        pic_status[p] = (data0 + carry + total > P.picture_size) ? PIC_TABLE_ERR : PIC_OK;
    }
    // The alpha work order: the slices of frames with alpha whose header
    // passes slice_header's checks (the colour decode reports the others),
    // long alpha streams at the front, the rest from the back; each with
    // where its alpha stream is. A slice outside the picture reads nothing.
    bool listed = false, heavy = false;
    uint32_t alen = 0, ahead[4] = {0, 0, 0, 0};
    uint64_t astream = 0;
    // The colour streams' sizes (the entropy items' costs below): the
    // header's size fields, each at most the slice's size; a slice whose
    // first 8 bytes are not inside the picture counts its whole size.
    uint32_t csz[3] = {sz, sz, sz};
    if (valid && sz >= 8 && data0 + at + sz <= P.picture_size) {
      const GLOBAL uint8_t *sl = (const GLOBAL uint8_t *)(uintptr_t)(P.pic_off + data0 + at);
      const uint32_t hs = sl[0] >> 3, q = sl[1];
      const uint32_t y = (uint32_t)sl[2] << 8 | sl[3], u = (uint32_t)sl[4] << 8 | sl[5];
      const uint32_t v = has_alpha ? ((uint32_t)sl[6] << 8 | sl[7]) : (hs + y + u < sz ? sz - hs - y - u : 0u);
      csz[0] = y < sz ? y : sz;
      csz[1] = u < sz ? u : sz;
      csz[2] = v < sz ? v : sz;
      const uint32_t c = y + u + v;
      // slice_header with alpha: 8 <= hs <= size, 1 <= q <= 224, the
      // component sizes inside the body.
      listed = has_alpha && hs >= 8 && hs <= sz && q >= 1 && q <= 224 && c <= sz - hs;
      alen = listed ? sz - hs - c : 0u;
      heavy = listed && alen > ALPHA_HEAVY_BYTES;
      astream = P.pic_off + data0 + at + hs + c;
      if (listed && alen <= ALPHA_HEAD_BYTES) {
        // Inside the packet and its 64 readable bytes past the end.
        const GLOBAL uint8_t *h = (const GLOBAL uint8_t *)(uintptr_t)astream;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++)
          ahead[i] = (uint32_t)h[4 * i] | (uint32_t)h[4 * i + 1] << 8 | (uint32_t)h[4 * i + 2] << 16 |
                     (uint32_t)h[4 * i + 3] << 24;
      }
    }
    if (valid) {
      const uint32_t row = k / P.row_slices;
      SliceRec sr;
      sr.off = P.pic_off + data0 + at;
      sr.coef = P.coef_base + (uint64_t)row * P.row_coefs + rs.coef_prefix;
      sr.len = sz;
      sr.pic = p;
      sr.mb_x = rs.mb_x;
      sr.mb_y = (uint16_t)row;
      sr.mbs = rs.mbs;
      slices[P.slice_base + k] = sr;
      slice_err[P.slice_base + k] = make_uint4(0, 0, 0, 0);
    }
    {
      // The entropy items' costs (pd_entropy_order): for each group of 64
      // consecutive slices of the batch, each colour component's largest
      // stream (item 3 g + c), zero at upload. A wave's slices fall in at
      // most two groups: one atomic each a component.
      const uint32_t g0 = (P.slice_base + k0 + 64 * wave) / 64, gk = (P.slice_base + k) / 64;
      const bool first = valid && gk == g0, second = valid && gk != g0;
      const uint64_t in0 = __ballot(first), in1 = __ballot(second);
#pragma unroll
      for (uint32_t c = 0; c < 3; c++) {
        uint32_t m0 = first ? csz[c] : 0u, m1 = second ? csz[c] : 0u;
        for (uint32_t d = 1; d < 64; d <<= 1) {
          const uint32_t a = __shfl_xor(m0, d, 64), b = __shfl_xor(m1, d, 64);
          m0 = a > m0 ? a : m0;
          m1 = b > m1 ? b : m1;
        }
        if (lane == 0) {
          if (in0) atomicMax(&group_cost[3 * g0 + c], m0);
          if (in1) atomicMax(&group_cost[3 * (g0 + 1) + c], m1);
        }
      }
    }
    if (has_alpha) {
      // Places in the work order: one atomic a wave for each end.
      const uint64_t hm = __ballot(heavy), lm = __ballot(listed && !heavy);
      uint32_t hb = 0, lb = 0;
      if (lane == 0) {
        if (hm) hb = atomicAdd(&B->heavy, (uint32_t)__popcll(hm));
        if (lm) lb = atomicAdd(&B->light, (uint32_t)__popcll(lm));
      }
      hb = __shfl(hb, 0, 64);
      lb = __shfl(lb, 0, 64);
      const uint64_t below = (1ull << lane) - 1;
      if (listed) {
        AlphaRec ar;
        ar.stream = astream;
        ar.g = P.slice_base + k;
        ar.len = alen;
        ar.pic = p;
        ar.frame = P.frame;
        ar.mb = (uint32_t)rs.mb_x | (k / P.row_slices) << 16;
        ar.mbs = rs.mbs;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) ar.head[i] = ahead[i];
        alpha_order[heavy ? hb + (uint32_t)__popcll(hm & below) : G - 1 - lb - (uint32_t)__popcll(lm & below)] = ar;
      }
    }
    // The picture's IDCT items (every row cut into the same segments),
    // shared among the picture's chunks.
    const uint32_t nch = P.n_slices ? (P.n_slices + ST_CHUNK - 1) / ST_CHUNK : 1u;
    for (uint32_t i = k0 + tid; i < P.height_in_mb * P.n_segs; i += nch * ST_CHUNK) {
      const uint32_t row = i / P.n_segs;
      IdctItem it = segs[P.seg_base + (i - row * P.n_segs)];
      it.first += P.slice_base + row * P.row_slices;
      it.pic = p;
      items[P.item_base + i] = it;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// One lane per colour stream; an item is 64 streams of one component (64
// consecutive slices), one wave's work: item i is component i % 3 of slices
// 64 (i / 3) ... 64 (i / 3) + 63. Dynamic LDS: LINE_WORDS dwords per thread.
extern "C" __global__ void __launch_bounds__(ENTROPY_MAX_WG)
    pd_entropy(const DevTables *__restrict__ T, DecBatch *__restrict__ B, const DecFrame *__restrict__ frames,
               const DecPic *__restrict__ pics, const SliceRec *__restrict__ slices,
               const uint32_t *__restrict__ pic_status, int16_t *__restrict__ coefs,
               int32_t *__restrict__ coef_wide, uint32_t *__restrict__ slice_err,
               uint32_t *__restrict__ pic_first_err, const uint32_t *__restrict__ order, uint32_t persistent) {
  // LINE_WORDS dwords a thread, then a 64-entry list a wave.
  extern __shared__ uint32_t lines[];
  uint4 *list = (uint4 *)(lines + LINE_WORDS * blockDim.x) + 64 * (threadIdx.x >> 6);
  __shared__ __attribute__((aligned(16))) uint16_t ftab[6 * 1024 + 6 * 512];
  static_assert(sizeof(ftab) == sizeof(T->fast_level) + sizeof(T->fast_run), "the tables are contiguous");
  for (uint32_t i = threadIdx.x; i < sizeof(ftab) / 16; i += blockDim.x)
    ((uint4 *)ftab)[i] = ((const uint4 *)T->fast_level)[i];
  __syncthreads();
  const uint32_t n_items = B->n_entropy_items, n_slices = B->n_slices;
  const uint32_t lane = threadIdx.x & 63;
  if (!persistent) {
    const uint32_t b = blockIdx.x, w = threadIdx.x >> 6, ew = blockDim.x >> 6, nb = gridDim.x;
    uint32_t item = b * ew + w;
    if (order && B->ordered) {
      // `order` holds the items longest first. Waves w, w + 4, w + 8, ... of
      // a workgroup run on one SIMD (the hardware deals a workgroup's waves
      // to its CU's four SIMDs in turn), and they share its issue slots.
      // With m of them a SIMD and n = 4 nb SIMD slots in all, the waves of
      // SIMD slot s take ranks s, n + s, ..., and the last one rank
      // m n - 1 - s: the longest items share a SIMD with the shortest, so
      // they run alone at the end, and the SIMDs' loads come out even.
      if ((ew & 3) == 0) {
        const uint32_t m = ew >> 2, n = 4 * nb, s = 4 * b + (w & 3), k = w >> 2;
        item = order[(m > 1 && k == m - 1) ? m * n - 1 - s : k * n + s];
      } else {
        item = order[w * nb + ((w & 1) ? nb - 1 - b : b)];
      }
    }
    const uint32_t g = (item / 3) * 64 + lane;
    if (item < n_items)  // the same for the whole wave
      entropy_stream(g < n_slices, g, item % 3, lines, list, &B->wide, ftab, frames, pics, slices, pic_status, coefs,
                     coef_wide, slice_err, pic_first_err);
    return;
  }
  const uint32_t chunk = B->chunk;
  for (;;) {
    const uint32_t first = wave_pop(&B->q[0], chunk);
    if (first >= n_items) break;
    const uint32_t last = first + chunk < n_items ? first + chunk : n_items;
    for (uint32_t k = first; k < last; k++) {
      const uint32_t item = order && B->ordered ? order[k] : k;  // longest first
      const uint32_t g = (item / 3) * 64 + lane;
      entropy_stream(g < n_slices, g, item % 3, lines, list, &B->wide, ftab, frames, pics, slices, pic_status, coefs,
                     coef_wide, slice_err, pic_first_err);
    }
  }
}

// ---------------------------------------------------------------------------
// One workgroup: the entropy items of the batch in order of cost, longest
// first, into order[0, n_slots) (slots past the items hold indices past
// them; n_slots 0: the batch's item count). An item's cost is its longest
// stream in bytes (pd_slice_table's costs): a wave's time follows its
// longest stream. A counting sort on cost / ORDER_BUCKET_BYTES (items of
// one bucket in any order); the host launches this only for
// n_slots <= ORDER_MAX.
constexpr uint32_t ORDER_BUCKETS = 2048, ORDER_BUCKET_BYTES = 32;
extern "C" __global__ void __launch_bounds__(1024)
    pd_entropy_order(DecBatch *__restrict__ B, const uint32_t *__restrict__ item_cost,
                     uint32_t *__restrict__ order, uint32_t n_slots) {
  __shared__ uint32_t cnt[ORDER_BUCKETS];
  __shared__ uint16_t bucket[ORDER_MAX];
  __shared__ uint32_t wsum[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t G = B->n_slices, groups = (G + 63) / 64, n_items = 3 * groups;
  if (n_slots == 0) n_slots = n_items;
  // A batch of more items than the order holds goes unordered (a device
  // plan's count is known only here).
  if (n_slots > ORDER_MAX || n_items > n_slots) return;
  for (uint32_t i = tid; i < ORDER_BUCKETS; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  // Costs, one an item (its longest stream, from pd_slice_table).
  for (uint32_t i = tid; i < n_items; i += blockDim.x) {
    const uint32_t v = item_cost[i];
    // Descending cost: bucket 0 holds the longest items.
    const uint32_t b = ORDER_BUCKETS - 1 - (v / ORDER_BUCKET_BYTES < ORDER_BUCKETS - 1
                                                ? v / ORDER_BUCKET_BYTES : ORDER_BUCKETS - 1);
    bucket[i] = (uint16_t)b;
    atomicAdd(&cnt[b], 1u);
  }
  __syncthreads();
  // Exclusive scan of the counts, two buckets a thread.
  const uint32_t c0 = cnt[2 * tid], c1 = cnt[2 * tid + 1];
  uint32_t incl = c0 + c1;
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; w++) before += wsum[w];
  const uint32_t ex = before + incl - c0 - c1;
  __syncthreads();
  cnt[2 * tid] = ex;
  cnt[2 * tid + 1] = ex + c0;
  __syncthreads();
  for (uint32_t i = tid; i < n_items; i += blockDim.x) order[atomicAdd(&cnt[bucket[i]], 1u)] = i;
  for (uint32_t i = n_items + tid; i < n_slots; i += blockDim.x) order[i] = i;
  if (tid == 0) B->ordered = 1;
}

// ---------------------------------------------------------------------------
// One wavefront per slice with alpha, in the work order pd_slice_table made
// (long streams first): scanned_alpha() (src/decode.rs:570-595) and its
// placement (src/decode.rs:365-389). The wave stages the slice's
// alpha stream in LDS (big-endian 64-bit words, zero past the end), every
// lane parses the same codes from it (a broadcast read; the parse is
// wave-uniform), runs are filled into an LDS copy of the slice's values by
// the whole wave, and the visible rows are stored coalesced at the end. The
// parse never waits on a global store.
// A code is at most 33 bits (a raw difference of 1 + 16, a run of 1 + 4 +
// 11), so 2048 values read at most 8448 bytes of the stream.
extern "C" __global__ void __launch_bounds__(64)
    pd_alpha(DecBatch *__restrict__ B, const DecFrame *__restrict__ frames,
             const DecPic *__restrict__ pics, const SliceRec *__restrict__ slices,
             const uint32_t *__restrict__ pic_status, uint32_t *__restrict__ slice_err,
             uint32_t *__restrict__ pic_first_err, const AlphaRec *__restrict__ order,
             uint32_t persistent) {
  __shared__ uint64_t words[ALPHA_WORDS];
  __shared__ __attribute__((aligned(16))) uint16_t vals[2048];
  __shared__ uint32_t marks[64];
  // The order holds `heavy` records at the front and `light` at the back.
  const uint32_t G = B->n_slices, heavy = B->heavy, n = heavy + B->light;
  if (!persistent) {
    const uint32_t x = blockIdx.x;
    if (x < heavy || (x < G && x >= G - B->light))
      alpha_slice(order[x], words, vals, marks, frames, pics, pic_status, slice_err, pic_first_err);
    return;
  }
  const uint32_t chunk = B->chunk;
  for (;;) {
    const uint32_t first = uni(wave_pop(&B->q[1], chunk));
    if (first >= n) break;
    const uint32_t last = first + chunk < n ? first + chunk : n;
    for (uint32_t i = first; i < last; i++) {
      alpha_slice(order[i < heavy ? i : G - 1 - (i - heavy)], words, vals, marks, frames, pics, pic_status, slice_err,
                  pic_first_err);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// One workgroup per item of consecutive slices of one picture (at most
// IDCT_MAX_BLOCKS blocks): dequantisation, the separable f32 IDCT in the
// reference's operation order (src/dct.rs:36-66, src/dsp/scalar.rs:8-12) and
// clamped planar u16 stores (src/decode.rs:395-438).
//
// Set-up, once per item: the slices' quantisers, weights and scan; the
// dequantisation factors (W q)/8 of every slice and weight matrix in the
// item; and one descriptor per block (its coefficients, its factors, its
// place in the output plane). Then 32 blocks per pass, 8 lanes per block
// (the lanes of a block are in one wave, so the two transposes through LDS
// need no workgroup barrier): a row each in the row pass, a column each in
// the column pass, and a row each again for the 16-byte stores.
//
// Every sum starts from its first product instead of 0.0 + it, which can
// change only the sign of a zero, and no nonzero sum or final sample depends
// on that. Slices with an error are skipped, and so are blocks outside the
// picture.
// Per block in LDS, the transpose buffer: 8 rows of 9 floats; the samples
// (8 rows of 8 u16, 16-byte aligned) reuse the same bytes once the column
// pass has read them.
constexpr uint32_t IDCT_TMP = 72;
constexpr uint32_t IDCT_OUT = 2 * IDCT_TMP;  // the block's stride in u16

// T[7 - x][u] = (-1)^u T[x][u] for every entry of the f32 basis, bit for bit.
constexpr bool basis_mirrored() {
  for (int x = 0; x < 8; x++)
    for (int u = 0; u < 8; u++) {
      const float a = BASIS[8 * x + u], b = BASIS[8 * (7 - x) + u];
      if ((u & 1) ? !(a == -b) : !(a == b)) return false;
    }
  return true;
}
static_assert(basis_mirrored(), "the DCT basis must be mirror-symmetric");
// Every |T[x][u]| is one of the entries of row 0.
constexpr bool basis_magnitudes_in_row0() {
  for (int k = 0; k < 64; k++) {
    const float m = BASIS[k] < 0.0f ? -BASIS[k] : BASIS[k];
    bool found = false;
    for (int i = 0; i < 8; i++) found = found || BASIS[i] == m;
    if (!found) return false;
  }
  return true;
}
static_assert(basis_magnitudes_in_row0(), "the DCT basis must take the magnitudes of its first row");

// o[x] = sum_u a[u] T[x][u] for x = 0..7, u ascending from the u = 0 product,
// every product and sum its own IEEE operation. The products are formed from
// |T[x][u]| and the sign goes into the sum (s + (-p) and s - p are the same
// IEEE operation, and a (-c) = -(a c) exactly). Outputs x and 7 - x then use
// the same products, and equal magnitudes in a column of T (seven distinct
// ones in all) give the same product: 22 multiplications instead of 64, and
// every sum bit for bit the reference's.
// The empty asm keeps each product as it is: without it the compiler turns
// s - a c back into s + a (-c), a product of its own.
// (Not volatile: a product no sum uses is still dropped.)
__device__ __forceinline__ float kept(float p) {
  asm("" : "+v"(p));
  return p;
}

__device__ __forceinline__ void idct8(const float a[8], float o[8]) {
  constexpr float K[8] = {BASIS[0], BASIS[1], BASIS[2], BASIS[3], BASIS[4], BASIS[5], BASIS[6], BASIS[7]};
  // p[u][i]: a[u] times the i-th magnitude of column u of T; |T[x][u]| is
  // one of |T[0][u']|, and the products are made once each.
  float pr[8][8];
#pragma unroll
  for (int u = 0; u < 8; u++)
#pragma unroll
    for (int i = 0; i < 8; i++) pr[u][i] = kept(a[u] * K[i]);
#pragma unroll
  for (int x = 0; x < 4; x++) {
    float s = pr[0][0];  // T[x][0] is the same for every x
    float m = s;         // output 7 - x
#pragma unroll
    for (int u = 1; u < 8; u++) {
      const float c = BASIS[8 * x + u];
      const float mag = c < 0.0f ? -c : c;
      int i = 0;
      while (K[i] != mag) i++;
      const float p = pr[u][i];
      const bool neg = c < 0.0f, neg_m = neg != ((u & 1) != 0);
      s = neg ? s - p : s + p;
      m = neg_m ? m - p : m + p;
    }
    o[x] = s;
    o[7 - x] = m;
  }
}

template <bool WIDE>
__device__ __forceinline__ void idct_pass(uint32_t t, uint32_t nblocks, uint32_t r, uint32_t lb,
                                          const uint2 *s_blk, const int16_t *sc, const float *s_scale,
                                          float *tmp, uint16_t *outb, const uint32_t s2[8],
                                          const int32_t *__restrict__ wide, const uint64_t *s_plane,
                                          const uint32_t *s_rstride, float osc, float obias, float omax) {
  uint2 D = make_uint2(0, 0);
  if (t < nblocks) D = s_blk[t];
  // The packed descriptor (pd_idct): coefficient offset, rows, cols, plane,
  // log2 of the component's blocks, slice within the item.
  const uint32_t bx0 = D.x & 0x1FFF, rows = (D.x >> 13) & 15, cols = (D.x >> 17) & 15, c = (D.x >> 21) & 3;
  const uint32_t lg = (D.x >> 23) & 7, fbase = 128 * (D.x >> 26) + (c ? 64u : 0u);
  const bool active = rows != 0;
  float o[8];
  if (active) {
    // Row pass, lane r: row r of the block, F = coefficient x (W q)/8.
    const float4 *sp = (const float4 *)(s_scale + fbase + 8 * r);
    const float4 f0 = sp[0], f1 = sp[1];
    const float fs[8] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w};
    float a[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) {
      // Byte offset of coefficient nb * scan + block (s2 holds 2 scan).
      const uint32_t at = 2 * bx0 + (s2[u] << lg);
      int32_t v = *(const int16_t *)((const char *)sc + at);
      if (WIDE && v == -32768) v = wide[at / 2];
      a[u] = (float)v * fs[u];
    }
    idct8(a, o);
#pragma unroll
    for (uint32_t x = 0; x < 8; x++) tmp[lb * IDCT_TMP + 9 * r + x] = o[x];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (active) {
    // Column pass, lane r: column r; clamped samples to LDS in rows, over
    // the block's transpose buffer (every sample depends on all eight
    // reads, so every lane of the wave has read before any lane writes).
    float a[8];
#pragma unroll
    for (uint32_t v = 0; v < 8; v++) a[v] = tmp[lb * IDCT_TMP + 9 * v + r];
    idct8(a, o);
#pragma unroll
    for (uint32_t y = 0; y < 8; y++) {
      const float sv = __builtin_amdgcn_fmed3f(floorf(o[y] * osc + obias), 0.0f, omax);
      outb[lb * IDCT_OUT + 8 * y + r] = (uint16_t)(uint32_t)sv;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (active && r < rows) {
    // Row r of the block to its plane (put_block, src/decode.rs:418-438).
    GLOBAL uint16_t *o16 = (GLOBAL uint16_t *)(uintptr_t)s_plane[c] + D.y + (size_t)s_rstride[c] * r;
    const uint4 pk = *(const uint4 *)(outb + lb * IDCT_OUT + 8 * r);
    if (cols == 8 && (((uintptr_t)o16) & 15) == 0) {
      *(GLOBAL uint4 *)o16 = pk;
    } else {
      const uint32_t wv[4] = {pk.x, pk.y, pk.z, pk.w};
      for (uint32_t x = 0; x < cols; x++) o16[x] = (uint16_t)(wv[x >> 1] >> (16 * (x & 1)));
    }
  }
  // The next pass reuses tmp and outb: every lane is done with them.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

extern "C" __global__ void __launch_bounds__(256)
    pd_idct(const DevTables *__restrict__ T, DecBatch *__restrict__ B,
            const DecFrame *__restrict__ frames, const DecPic *__restrict__ pics,
            const SliceRec *__restrict__ slices, const uint32_t *__restrict__ pic_status,
            const uint32_t *__restrict__ slice_err, const int16_t *__restrict__ coefs,
            const int32_t *__restrict__ coef_wide, const IdctItem *__restrict__ items,
            uint32_t persistent) {
  __shared__ __attribute__((aligned(16))) int16_t sc[IDCT_MAX_BLOCKS * 64];
  // 128 factors for each slice an item can hold: the launch sizes this
  // (dynamic LDS) for the batch's largest item, so that one-slice items
  // leave room for more workgroups a CU.
  extern __shared__ __attribute__((aligned(16))) float s_scale[];
  __shared__ __attribute__((aligned(16))) float tmp[32 * IDCT_TMP];
  uint16_t *const outb = (uint16_t *)tmp;
  // Kept small: 23,040 bytes of LDS or less a workgroup leave room for
  // seven workgroups a CU.
  __shared__ __attribute__((aligned(16))) uint2 s_blk[IDCT_MAX_BLOCKS];
  // Per slice: first block; mb_x | mb_y << 16; mbs | ok << 8 | qScale << 16.
  __shared__ uint32_t s_start[IDCT_MAX_SLICES + 1], s_pos[IDCT_MAX_SLICES], s_mq[IDCT_MAX_SLICES];
  __shared__ uint64_t s_plane[3];
  __shared__ uint32_t s_rstride[3];
  __shared__ uint32_t s_item;
  const uint32_t tid = threadIdx.x, n_items = B->n_idct_items;
  const uint32_t chunk = persistent ? B->chunk : 1u;
  // Coefficients outside +-32767 (from the side arena) only when the
  // entropy decode stored any in this batch.
  const bool any_wide = B->wide != 0;
  uint32_t first = blockIdx.x;
  for (;;) {
    if (persistent) {
      if (tid == 0) s_item = atomicAdd(&B->q[2], chunk);
      __syncthreads();
      first = s_item;
      __syncthreads();
    }
    if (first >= n_items) break;
    const uint32_t last = first + chunk < n_items ? first + chunk : n_items;
    for (uint32_t it = first; it < last; it++) {
      const IdctItem I = items[it];
      const SliceRec S0 = slices[I.first];
      const DecPic &P = pics[I.pic];
      const DecFrame &F = frames[P.frame];
      const bool pic_ok = pic_status[I.pic] == PIC_OK;
      const uint32_t chroma = F.chroma, cpm = chroma == 2 ? 2u : 4u;
      // The item's slices: block offsets, qScale, position, error state; the
      // frame's weights, scan and planes.
      if (tid < I.count) {
        const SliceRec S = slices[I.first + tid];
        const uint4 ev = *(const uint4 *)(slice_err + 4 * (size_t)(I.first + tid));
        // A picture whose slice table failed its bounds check has slice
        // offsets that may point anywhere: nothing of it is read. The alpha
        // word is not looked at: alpha is decoded beside this kernel.
        const bool ok = pic_ok && !(ev.x | ev.y | ev.z);
        const uint32_t q = ok ? qscale(((const GLOBAL uint8_t *)(uintptr_t)S.off)[1]) : 0u;
        s_pos[tid] = (uint32_t)S.mb_x | (uint32_t)S.mb_y << 16;
        s_mq[tid] = S.mbs | (ok ? 1u : 0u) << 8 | q << 16;
        s_start[tid + 1] = (uint32_t)((S.coef - S0.coef) / 64) + S.mbs * (4 + 2 * cpm);
        if (tid == 0) s_start[0] = 0;
      }
      if (tid >= 32 && tid < 35) {
        s_plane[tid - 32] = F.out_plane[tid - 32];
        s_rstride[tid - 32] = P.row_step * F.plane_w[tid - 32];
      }
      {
        // The item's coefficients (at most IDCT_MAX_BLOCKS = 96 blocks, three
        // 16-byte pieces a thread), every load issued before the first store.
        static_assert(IDCT_MAX_BLOCKS * 8 <= 3 * 256, "three pieces a thread");
        const uint4 *src = (const uint4 *)(coefs + S0.coef);
        const uint32_t n16 = I.blocks * 8;
        // Indices past the item's end read its last piece again (n16 >= 8).
        const uint4 v0 = src[tid < n16 ? tid : n16 - 1];
        const uint4 v1 = src[tid + 256 < n16 ? tid + 256 : n16 - 1];
        const uint4 v2 = src[tid + 512 < n16 ? tid + 512 : n16 - 1];
        if (tid < n16) ((uint4 *)sc)[tid] = v0;
        if (tid + 256 < n16) ((uint4 *)sc)[tid + 256] = v1;
        if (tid + 512 < n16) ((uint4 *)sc)[tid + 512] = v2;
      }
      __syncthreads();
      // Dequantisation factors, scale[k] = (f32)(W[k] qScale) / 8 for each
      // slice and matrix (src/decode.rs:356).
      for (uint32_t e = tid; e < I.count * 128; e += 256)
        s_scale[e] = (float)((uint32_t)((e & 64) ? F.chroma_w[e & 63] : F.luma_w[e & 63]) * (s_mq[e >> 7] >> 16)) / 8.0f;
      // Block descriptors: x = its first coefficient in sc (the scanned
      // index is nb * scan + block), y = its factors in s_scale, z = its
      // top-left sample in the plane, w = rows | cols << 4 | plane << 8 |
      // log2(nb) << 12; rows is 0 for a block that is skipped.
      for (uint32_t t = tid; t < I.blocks; t += 256) {
        uint32_t j = 0;
        while (j + 1 < I.count && s_start[j + 1] <= t) j++;
        const uint32_t mbs = s_mq[j] & 255, nbY = 4 * mbs, nbC = cpm * mbs, tl = t - s_start[j];
        uint32_t c = 0, b = tl, nb = nbY, cbase = 64 * s_start[j];
        if (tl >= nbY + nbC) {
          c = 2; b = tl - nbY - nbC; nb = nbC; cbase += 64 * (nbY + nbC);
        } else if (tl >= nbY) {
          c = 1; b = tl - nbY; nb = nbC; cbase += 64 * nbY;
        }
        // block_position (src/decode.rs:395-413).
        const uint32_t per_mb = c == 0 ? 4u : cpm;
        const uint32_t mbx = (s_pos[j] & 0xFFFF) + b / per_mb, bb = b % per_mb, mby = s_pos[j] >> 16;
        uint32_t bx, by;
        if (c == 0) {
          bx = 16 * mbx + 8 * (bb & 1); by = 16u * mby + 8 * (bb >> 1);
        } else if (chroma == 2) {
          bx = 8 * mbx; by = 16u * mby + 8 * bb;
        } else {
          bx = 16 * mbx + 8 * (bb >> 1); by = 16u * mby + 8 * (bb & 1);
        }
        const uint32_t pw = F.plane_w[c];
        uint32_t rows = 0, cols = 0, out = 0;
        if (((s_mq[j] >> 8) & 1) && bx < pw && by < P.height) {
          rows = P.height - by < 8 ? P.height - by : 8u;
          cols = pw - bx < 8 ? pw - bx : 8u;
          out = (P.first_row + P.row_step * by) * pw + bx;
        }
        const uint32_t lg = (uint32_t)__builtin_ctz(nb);  // nb is a power of two
        s_blk[t] = make_uint2((cbase + b) | rows << 13 | cols << 17 | c << 21 | lg << 23 | j << 26, out);
      }
      __syncthreads();
      const uint32_t r = tid & 7, lb = tid >> 3;
      uint32_t s2[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) s2[u] = 2u * T->scan[F.interlaced ? 1 : 0][8 * r + u];
      const float osc = F.out_scale, obias = F.out_bias, omax = F.out_max;
      const int32_t *wide = coef_wide + S0.coef;
      for (uint32_t t0 = 0; t0 < I.blocks; t0 += 32) {
        if (any_wide)
          idct_pass<true>(t0 + lb, I.blocks, r, lb, s_blk, sc, s_scale, tmp, outb, s2, wide, s_plane,
                          s_rstride, osc, obias, omax);
        else
          idct_pass<false>(t0 + lb, I.blocks, r, lb, s_blk, sc, s_scale, tmp, outb, s2, wide, s_plane,
                           s_rstride, osc, obias, omax);
      }
      __syncthreads();
    }
    if (!persistent) break;
  }
}

// ---------------------------------------------------------------------------
// pd_plan: a batch of packets held in device memory, planned on the device.
// One workgroup. For each job (one thread a job): FrameHeader::parse
// (src/header.rs:82-170), check_plausible_size (src/decode.rs:107-128), the
// job's device buffers, then each picture's PictureHeader::parse
// (src/header.rs:262-290) and slice-table extent (src/decode.rs:202-210),
// with the reference's checks in its order; then prefix sums over the jobs
// place every good picture, and each job writes its DecFrame, DecPics, row
// slots, IDCT row segments and slice-table chunks, exactly as the host
// planner lays out a batch of host packets. The entropy item costs are
// cleared and the DecBatch written (queues zero). A batch over the
// workspace's capacities writes nothing but the totals, marked overflow, and
// a DecBatch with no work.
namespace {

__device__ __forceinline__ uint32_t rd16(const GLOBAL uint8_t *p) { return (uint32_t)p[0] << 8 | p[1]; }
__device__ __forceinline__ uint32_t rd32(const GLOBAL uint8_t *p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

// Slices of a row of wmb macroblocks at 2^log2 (header::slice_sizes,
// src/header.rs:308-320): wmb >> log2 full slices, then one for each set bit
// of the remainder, largest first. Slice j's first macroblock and size.
__device__ __forceinline__ uint32_t row_slice_count(uint32_t wmb, uint32_t log2) {
  return (wmb >> log2) + __builtin_popcount(wmb & ((1u << log2) - 1));
}
__device__ __forceinline__ void row_slice(uint32_t wmb, uint32_t log2, uint32_t j, uint32_t &mb_x, uint32_t &mbs) {
  const uint32_t full = wmb >> log2;
  if (j < full) {
    mb_x = j << log2;
    mbs = 1u << log2;
    return;
  }
  uint32_t x = full << log2, rem = wmb & ((1u << log2) - 1);
  for (uint32_t k = full; k < j; k++) {  // the tail slices before j (at most 2 of them)
    const uint32_t b = 1u << (31 - __builtin_clz(rem));
    x += b;
    rem -= b;
  }
  mb_x = x;
  mbs = 1u << (31 - __builtin_clz(rem));
}

// The IDCT row segments of a row (the host planner's cut): consecutive
// slices while their blocks fit `cap` and they are at most `max_slices`; a
// larger slice is a segment of its own. Counts them, or writes them at `out`.
__device__ uint32_t row_segments(uint32_t wmb, uint32_t log2, uint32_t cpm, uint32_t cap, uint32_t max_slices,
                                 IdctItem *out, uint32_t &max_count, uint32_t &wanted) {
  const uint32_t n = row_slice_count(wmb, log2);
  IdctItem cur{0, 0, 0, 0};
  uint32_t segs = 0, lcount = 0, lblocks = 0;  // the segment the cut would make without max_slices
  max_count = 0;
  wanted = 0;
  for (uint32_t j = 0; j < n; j++) {
    uint32_t mb_x, mbs;
    row_slice(wmb, log2, j, mb_x, mbs);
    const uint32_t tb = mbs * (4 + 2 * cpm);
    if (cur.count && (cur.blocks + tb > cap || cur.count == max_slices)) {
      if (out) out[segs] = cur;
      max_count = cur.count > max_count ? cur.count : max_count;
      segs++;
      cur = IdctItem{j, 0, 0, 0};
    }
    if (lcount && (lblocks + tb > cap || lcount == IDCT_MAX_SLICES)) lcount = lblocks = 0;
    lcount++;
    lblocks += tb;
    wanted = lcount > wanted ? lcount : wanted;
    cur.count++;
    cur.blocks += tb;
  }
  if (out) out[segs] = cur;
  max_count = cur.count > max_count ? cur.count : max_count;
  return segs + 1;
}

// Parses job j's headers into its record and PlanJob, with the reference's
// checks in its order.
__device__ void plan_parse(const DevJob &J, PlanRec &R, PlanJob &P, uint32_t cap, uint32_t max_slices) {
  R.code = F_NONE;
  R.deferred = F_NONE;
  R.n_pics = 0;
  R.first_pic = 0;
  R.arg[0] = R.arg[1] = R.arg[2] = R.arg[3] = 0;
  R.darg = 0;
  P.n_pics = 0;
  P.slices = P.rows = P.segs = P.items = P.chunks = P.max_count = P.wanted = 0;
  P.coefs = 0;
  P.base_pic = P.base_slice = P.base_row = P.base_seg = P.base_item = P.base_chunk = 0;
  P.base_coef = 0;
  if (!J.packet) {
    R.code = F_NO_PACKET;
    return;
  }
  const GLOBAL uint8_t *d = (const GLOBAL uint8_t *)(uintptr_t)J.packet;
  const uint64_t len = J.len;
  if (len < 28) {
    R.code = F_SHORT;
    R.arg[0] = len;
    return;
  }
  const uint32_t frame_size = rd32(d);
  if (!(d[4] == 'i' && d[5] == 'c' && d[6] == 'p' && d[7] == 'f')) {
    R.code = F_ICPF;
    return;
  }
  if ((uint64_t)frame_size > len) {
    R.code = F_SIZE_EXCEEDS;
    R.arg[0] = frame_size;
    R.arg[1] = len;
    return;
  }
  if (frame_size < 28) {
    R.code = F_SIZE_SMALL;
    return;
  }
  const GLOBAL uint8_t *h = d + 8;
  const uint32_t avail = frame_size - 8, header_size = rd16(h);
  if (header_size < 20 || header_size > avail) {
    R.code = F_HDR_SIZE;
    R.arg[0] = header_size;
    return;
  }
  if (h[3] > 1) {
    R.code = F_VERSION;
    R.arg[0] = h[3];
    return;
  }
  const uint32_t width = rd16(h + 8), height = rd16(h + 10);
  if (width == 0 || height == 0) {
    R.code = F_DIMS;
    R.arg[0] = width;
    R.arg[1] = height;
    return;
  }
  const uint32_t chroma = h[12] >> 6, interlace = (h[12] >> 2) & 3;
  if (chroma != 2 && chroma != 3) {
    R.code = F_CHROMA;
    R.arg[0] = chroma;
    return;
  }
  if (interlace == 3) {
    R.code = F_INTERLACE;
    return;
  }
  const uint32_t alpha = h[17] & 15;
  if (alpha > 2) {
    R.code = F_ALPHA;
    R.arg[0] = alpha;
    return;
  }
  uint32_t mpos = 20;
  if (h[19] & 2) {
    if (mpos + 64 > header_size) {
      R.code = F_MATRIX;
      return;
    }
    mpos += 64;
  }
  if ((h[19] & 1) && mpos + 64 > header_size) {
    R.code = F_MATRIX;
    return;
  }
  // check_plausible_size: 11 bytes a slice at 8-macroblock slices.
  const uint32_t wmb = (width + 15) / 16, n_pics = interlace ? 2u : 1u;
  const uint64_t per_row = row_slice_count(wmb, 3);
  uint64_t need = 8 + (uint64_t)header_size;
  uint32_t heights[2];
  for (uint32_t i = 0; i < n_pics; i++) {
    const bool top = interlace == 0 || ((interlace == 1) == (i == 0));
    heights[i] = interlace == 0 ? height : (top ? (height + 1) / 2 : height / 2);
    need += 8 + 11 * per_row * (((uint64_t)heights[i] + 15) / 16);
  }
  if ((uint64_t)frame_size < need) {
    R.code = F_IMPLAUSIBLE;
    R.arg[0] = frame_size;
    R.arg[1] = width;
    R.arg[2] = height;
    R.arg[3] = need;
    return;
  }
  // The device writes through these: a job without them decodes nothing.
  if (!J.data) {
    R.code = F_NO_PLANES;
    return;
  }
  if (alpha && !J.alpha) {
    R.code = F_NO_ALPHA;
    return;
  }
  P.wmb = wmb;
  P.cpm = chroma == 2 ? 2u : 4u;
  uint64_t pos = 8 + (uint64_t)header_size;
  for (uint32_t i = 0; i < n_pics; i++) {
    if (pos > frame_size) {
      R.deferred = P_START_PAST;
      break;
    }
    const uint64_t plen = frame_size - pos;
    if (plen < 8) {
      R.deferred = P_HDR_SHORT;
      break;
    }
    const GLOBAL uint8_t *ph = d + pos;
    const uint32_t hs = ph[0] >> 3, ps = rd32(ph + 1);
    if (hs < 8 || hs > plen) {
      R.deferred = P_HDR_SIZE;
      R.darg = hs;
      break;
    }
    if (ps < hs || ps > plen) {
      R.deferred = P_SIZE;
      R.darg = ps;
      break;
    }
    const uint32_t log2 = (ph[7] >> 4) & 3, hmb = (heights[i] + 15) / 16;
    const uint32_t rs = row_slice_count(wmb, log2);
    const uint64_t n = (uint64_t)rs * hmb;
    if ((uint64_t)hs + 2 * n > ps) {
      R.deferred = P_TABLE;
      break;
    }
    PlanPic &Q = P.pic[P.n_pics];
    Q.pos = pos;
    Q.header_size = hs;
    Q.picture_size = ps;
    Q.log2 = log2;
    Q.row_slices = rs;
    Q.height = heights[i];
    Q.hmb = hmb;
    uint32_t mc, want;
    Q.n_segs = row_segments(wmb, log2, P.cpm, cap, max_slices, nullptr, mc, want);
    Q.max_count = mc;
    P.max_count = mc > P.max_count ? mc : P.max_count;
    P.wanted = want > P.wanted ? want : P.wanted;
    P.slices += (uint32_t)n;
    P.rows += rs;
    P.segs += Q.n_segs;
    P.items += hmb * Q.n_segs;
    P.chunks += n ? (uint32_t)((n + ST_CHUNK - 1) / ST_CHUNK) : 1u;
    P.coefs += (uint64_t)wmb * (4 + 2 * P.cpm) * 64 * hmb;
    P.n_pics++;
    pos += ps;
  }
  R.n_pics = P.n_pics;
}

}  // namespace

// The plan buffer's parts: frames (one a job), pictures, row slots, IDCT row
// segments, slice-table chunks (uint2), entropy item costs, DecBatch.
extern "C" __global__ void __launch_bounds__(256)
    pd_plan(const DevJob *__restrict__ jobs, uint32_t n_jobs, uint32_t want_depth, uint32_t idct_cap,
            uint32_t chunk, PlanCaps caps, PlanJob *__restrict__ scratch, PlanRec *__restrict__ recs,
            PlanTotals *__restrict__ totals, DecFrame *__restrict__ frames, DecPic *__restrict__ pics,
            RowSlot *__restrict__ rows, IdctItem *__restrict__ segs, uint2 *__restrict__ chunks,
            uint32_t *__restrict__ group_cost, DecBatch *__restrict__ B) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t max_slices = caps.item_slices < IDCT_MAX_SLICES ? caps.item_slices : IDCT_MAX_SLICES;
  for (uint32_t j = tid; j < n_jobs; j += 256) {
    const DevJob J = jobs[j];
    plan_parse(J, recs[j], scratch[j], idct_cap, max_slices);
  }
  __syncthreads();
  // Exclusive prefix sums over the jobs, 256 at a time: pictures, slices,
  // row slots, segments, IDCT items, chunks (u32) and coefficients (u64).
  __shared__ uint32_t wsum[4][6];
  __shared__ uint64_t wcoef[4];
  __shared__ uint32_t carry[6], maxc[2];
  __shared__ uint64_t ccarry;
  if (tid < 6) carry[tid] = 0;
  if (tid < 2) maxc[tid] = 0;
  if (tid == 0) ccarry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < n_jobs; b += 256) {
    const uint32_t j = b + tid;
    uint32_t v[6] = {0, 0, 0, 0, 0, 0};
    uint64_t cv = 0;
    if (j < n_jobs) {
      const PlanJob &P = scratch[j];
      v[0] = P.n_pics;
      v[1] = P.slices;
      v[2] = P.rows;
      v[3] = P.segs;
      v[4] = P.items;
      v[5] = P.chunks;
      cv = P.coefs;
      atomicMax(&maxc[0], P.max_count);
      atomicMax(&maxc[1], P.wanted);
    }
    uint32_t incl[6];
    uint64_t cincl = cv;
    for (int k = 0; k < 6; k++) incl[k] = v[k];
    for (uint32_t d = 1; d < 64; d <<= 1) {
      for (int k = 0; k < 6; k++) {
        const uint32_t o = __shfl_up(incl[k], d, 64);
        if (lane >= d) incl[k] += o;
      }
      const uint64_t co = __shfl_up(cincl, d, 64);
      if (lane >= d) cincl += co;
    }
    if (lane == 63) {
      for (int k = 0; k < 6; k++) wsum[wave][k] = incl[k];
      wcoef[wave] = cincl;
    }
    __syncthreads();
    uint32_t before[6] = {0, 0, 0, 0, 0, 0};
    uint64_t cbefore = 0;
    for (uint32_t w = 0; w < wave; w++) {
      for (int k = 0; k < 6; k++) before[k] += wsum[w][k];
      cbefore += wcoef[w];
    }
    if (j < n_jobs) {
      // The job's bases, kept in its scratch record for the layout pass.
      PlanJob &P = scratch[j];
      P.base_pic = carry[0] + before[0] + incl[0] - v[0];
      P.base_slice = carry[1] + before[1] + incl[1] - v[1];
      P.base_row = carry[2] + before[2] + incl[2] - v[2];
      P.base_seg = carry[3] + before[3] + incl[3] - v[3];
      P.base_item = carry[4] + before[4] + incl[4] - v[4];
      P.base_chunk = carry[5] + before[5] + incl[5] - v[5];
      P.base_coef = ccarry + cbefore + cincl - cv;
    }
    __syncthreads();
    if (tid == 0) {
      for (int k = 0; k < 6; k++) carry[k] += wsum[0][k] + wsum[1][k] + wsum[2][k] + wsum[3][k];
      ccarry += wcoef[0] + wcoef[1] + wcoef[2] + wcoef[3];
    }
    __syncthreads();
  }
  const bool over = carry[0] > caps.pics || carry[1] > caps.slices || carry[2] > caps.rows || carry[3] > caps.segs ||
                    carry[4] > caps.items || carry[5] > caps.chunks || ccarry > caps.coefs;
  if (tid == 0) {
    PlanTotals T;
    T.overflow = over ? 1u : 0u;
    T.pics = carry[0];
    T.slices = carry[1];
    T.rows = carry[2];
    T.segs = carry[3];
    T.items = carry[4];
    T.chunks = carry[5];
    T.max_count = maxc[0];
    T.coefs = ccarry;
    T.wanted_count = maxc[1];
    T.pad = 0;
    *totals = T;
    DecBatch D;
    memset(&D, 0, sizeof D);
    if (!over) {
      D.n_pics = carry[0];
      D.n_slices = carry[1];
      D.n_entropy_items = 3 * ((carry[1] + 63) / 64);
      D.n_idct_items = carry[4];
      D.n_chunks = carry[5];
    }
    D.chunk = chunk;
    *B = D;
  }
  if (over) return;
  // The entropy item costs (three a group of 64 slices) start at zero.
  for (uint32_t i = tid; i < 3 * ((carry[1] + 63) / 64); i += 256) group_cost[i] = 0;
  // Each job's frame, pictures, row slots, segments and chunks.
  for (uint32_t j = tid; j < n_jobs; j += 256) {
    PlanRec &R = recs[j];
    const PlanJob P = scratch[j];
    R.first_pic = P.base_pic;
    if (R.code != F_NONE) continue;
    const DevJob J = jobs[j];
    const GLOBAL uint8_t *d = (const GLOBAL uint8_t *)(uintptr_t)J.packet, *h = d + 8;
    const uint32_t width = rd16(h + 8), height = rd16(h + 10), chroma = h[12] >> 6, interlace = (h[12] >> 2) & 3;
    const uint32_t alpha = h[17] & 15;
    const uint32_t bd = want_depth ? want_depth : (chroma == 2 ? 10u : 12u);
    DecFrame F;
    memset(&F, 0, sizeof F);
    // Frame::layout (src/frame.rs:247-266): chroma width ceil(w/2) for 4:2:2.
    const uint32_t cw = chroma == 2 ? (width + 1) / 2 : width;
    const uint64_t off[3] = {0, (uint64_t)width * height, (uint64_t)width * height + (uint64_t)cw * height};
    for (int i = 0; i < 3; i++) F.out_plane[i] = J.data + 2 * off[i];
    F.plane_w[0] = width;
    F.plane_w[1] = F.plane_w[2] = cw;
    F.out_alpha = J.alpha;
    F.width = width;
    F.height = height;
    F.chroma = chroma;
    F.alpha_bits = alpha == 1 ? 8u : alpha == 2 ? 16u : 0u;
    F.bit_depth = bd;
    F.interlaced = interlace != 0;
    // Output (src/dsp/mod.rs:88-96): scale 2^b/512, bias 2^(b-1) + 0.5, max 2^b - 1.
    F.out_scale = (float)(1u << bd) / 512.0f;
    F.out_bias = (float)(1u << (bd - 1)) + 0.5f;
    F.out_max = (float)((1u << bd) - 1);
    if (F.alpha_bits) {
      // Rescale (src/decode.rs:446-469): (a * ceil(max_out 2^40 / max_in) + 2^39) >> 40.
      const uint64_t max_in = (1ull << F.alpha_bits) - 1, max_out = (1ull << bd) - 1;
      F.rescale_m = ((max_out << 40) + max_in - 1) / max_in;
      F.rescale_identity = F.alpha_bits == bd;
    }
    // Weights (src/header.rs:202-210): the loaded matrices, else all 4s;
    // chroma falls back to the luma weights.
    uint32_t mpos = 20;
    const bool lm = (h[19] & 2) != 0, cm = (h[19] & 1) != 0;
    for (int k = 0; k < 64; k++) F.luma_w[k] = lm ? h[mpos + k] : (uint8_t)4;
    if (lm) mpos += 64;
    for (int k = 0; k < 64; k++) F.chroma_w[k] = cm ? h[mpos + k] : F.luma_w[k];
    frames[j] = F;
    const uint32_t cpm = P.cpm;
    uint32_t pic = P.base_pic, slice = P.base_slice, row = P.base_row, seg = P.base_seg, item = P.base_item,
             ch = P.base_chunk;
    uint64_t coef = P.base_coef;
    for (uint32_t i = 0; i < R.n_pics; i++) {
      const PlanPic &Q = P.pic[i];
      DecPic D;
      memset(&D, 0, sizeof D);
      D.pic_off = J.packet + Q.pos;
      D.coef_base = coef;
      D.header_size = Q.header_size;
      D.picture_size = Q.picture_size;
      D.n_slices = Q.row_slices * Q.hmb;
      D.row_slices = Q.row_slices;
      D.slice_base = slice;
      D.layout_base = row;
      D.frame = j;
      D.height = Q.height;
      D.height_in_mb = Q.hmb;
      const bool top = interlace == 0 || ((interlace == 1) == (i == 0));
      D.first_row = interlace == 0 ? 0u : (top ? 0u : 1u);
      D.row_step = interlace == 0 ? 1u : 2u;
      D.seg_base = seg;
      D.n_segs = Q.n_segs;
      D.item_base = item;
      uint32_t prefix = 0;
      for (uint32_t k = 0; k < Q.row_slices; k++) {
        uint32_t mb_x, mbs;
        row_slice(P.wmb, Q.log2, k, mb_x, mbs);
        rows[row + k] = RowSlot{(uint16_t)mb_x, (uint16_t)mbs, prefix};
        prefix += mbs * (4 + 2 * cpm) * 64;
      }
      D.row_coefs = prefix;
      uint32_t mc, want;
      row_segments(P.wmb, Q.log2, cpm, idct_cap, max_slices, segs + seg, mc, want);
      for (uint32_t k = 0; k == 0 || k < D.n_slices; k += ST_CHUNK) chunks[ch++] = make_uint2(pic, k);
      pics[pic] = D;
      coef += (uint64_t)prefix * Q.hmb;
      slice += D.n_slices;
      row += Q.row_slices;
      seg += Q.n_segs;
      item += Q.hmb * Q.n_segs;
      pic++;
    }
  }
}
