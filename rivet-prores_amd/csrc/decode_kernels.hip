// gfx950 decode kernels for the ProRes slice engine.
//
// Three launches per batch of pictures:
// Packet and slice offsets are absolute device addresses.
//   pd_slice_table  one workgroup per picture: the u16 big-endian slice-size
//                   table becomes byte offsets (an exclusive prefix), plus the
//                   bounds check of the table walk (src/decode.rs:202-221).
//   pd_entropy      one lane per coded stream (a colour component or the alpha
//                   of one slice): the RDD 36 adaptive Rice/exp-Golomb DC and
//                   run/level decode (src/decode.rs:484-540, src/vlc.rs) into a
//                   coefficient arena in scanned order, alpha straight to the
//                   output plane (src/decode.rs:570-595, 365-389).
//   pd_idct         one workgroup per slice: dequantisation, the separable f32
//                   IDCT in the reference's operation order (src/dct.rs:36-66,
//                   src/dsp/scalar.rs:8-12) and clamped planar u16 stores
//                   (src/decode.rs:395-438).
//
// Built with -ffp-contract=off: every product and sum is its own IEEE op.
#include <hip/hip_runtime.h>

#include "layout.h"

using namespace prores_hip;

namespace {

constexpr uint32_t MAX_SYMBOL = 1u << 30;
constexpr uint32_t MAX_PREFIX = 32;

// Codebook parameters as (last_q | k_rice << 4 | k_exp << 8); an exp-Golomb
// code of order k is the combination code (0, k, k + 1). Tables 9-11 of
// RDD 36 (src/vlc.rs:30-68).
__constant__ uint16_t DC_BOOK[4] = {0x100, 0x210, 0x321, 0x430};
__constant__ uint16_t RUN_BOOK[16] = {0x102, 0x102, 0x101, 0x101, 0x100, 0x211, 0x211, 0x211,
                                      0x211, 0x210, 0x210, 0x210, 0x210, 0x210, 0x210, 0x320};
__constant__ uint16_t LEVEL_BOOK[9] = {0x202, 0x101, 0x102, 0x100, 0x210,
                                       0x210, 0x210, 0x210, 0x320};
constexpr uint32_t FIRST_DC_BOOK = 0x650;  // EG(5)

__device__ __forceinline__ uint32_t clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
__device__ __forceinline__ uint32_t qscale(uint32_t q) { return q <= 128 ? q : 128 + 4 * (q - 128); }
__device__ __forceinline__ int32_t symbol_to_signed(uint32_t s) {
  return (s & 1) == 0 ? (int32_t)(s >> 1) : -(int32_t)((s >> 1) + 1);
}

// MSB-first reader over one stream of `len` bytes at `base`. The window
// holds `nbits` loaded bits (at least 33 between operations); bits past the
// end of the stream read as zero, as the reference's 64-bit peek does
// (src/bits.rs:30-40). Words are loaded aligned, one ahead of use.
struct Reader {
  const uint32_t *src;
  uint64_t win;
  uint32_t nbits, wj, maxw, pos, lenbits, nextw;
  int32_t end_byte;

  __device__ __forceinline__ uint32_t load_word(uint32_t j) const {
    return __builtin_bswap32(src[j < maxw ? j : maxw]);
  }
  __device__ __forceinline__ void refill() {
    uint32_t w = nextw;
    int32_t valid = end_byte - 4 * (int32_t)wj;
    if (valid < 4) w = valid <= 0 ? 0u : (w & (0xFFFFFFFFu << (8 * (4 - valid))));
    win |= (uint64_t)w << (32 - nbits);
    nbits += 32;
    wj++;
    nextw = load_word(wj);
  }
  __device__ __forceinline__ void init(const uint8_t *base, uint32_t len) {
    uintptr_t a = (uintptr_t)base;
    uint32_t m = (uint32_t)(a & 3);
    src = (const uint32_t *)(a - m);
    end_byte = (int32_t)(m + len);
    maxw = (m + len + 3) / 4;
    maxw = maxw ? maxw - 1 : 0;
    win = 0;
    nbits = 0;
    wj = 0;
    nextw = load_word(0);
    refill();
    refill();
    win <<= 8 * m;
    nbits -= 8 * m;
    if (nbits <= 32) refill();
    pos = 0;
    lenbits = 8 * len;
  }
  __device__ __forceinline__ uint32_t remaining() const { return lenbits - pos; }
  __device__ __forceinline__ void consume(uint32_t n) {
    win <<= n;
    nbits -= n;
    pos += n;
    if (nbits <= 32) refill();
  }
  __device__ __forceinline__ uint32_t take(uint32_t n) {  // n <= 32
    uint32_t v = n ? (uint32_t)(win >> (64 - n)) : 0u;
    consume(n);
    return v;
  }
  // RDD 36 endOfData(): 31 or fewer bits remain and all are zero.
  __device__ __forceinline__ bool end_of_data() const {
    uint32_t r = remaining();
    return r <= 31 && (r == 0 || (win >> (64 - r)) == 0);
  }
};

// Codebook::read with the reference's checks, in its order
// (src/vlc.rs:206-230 over src/bits.rs:59-94).
__device__ __forceinline__ uint32_t read_code(Reader &r, uint32_t book, uint32_t &sym) {
  uint32_t last_q = book & 15, k_rice = (book >> 4) & 15, k_exp = book >> 8;
  uint32_t lz = clz64(r.win);
  if (lz > MAX_PREFIX) return E_PREFIX;
  if (lz >= r.remaining()) return E_PAST_END;
  r.consume(lz);
  if (lz <= last_q) {
    uint32_t n = 1 + k_rice;
    if (n > r.remaining()) return E_PAST_END;
    uint32_t v = r.take(n) & ((1u << k_rice) - 1);
    sym = (lz << k_rice) | v;
    return E_NONE;
  }
  uint32_t n = lz - (last_q + 1) + k_exp + 1;  // <= 39
  if (n > r.remaining()) return E_PAST_END;
  uint64_t v;
  if (n > 32) {
    uint64_t hi = r.take(n - 32);
    v = (hi << 32) | r.take(32);
  } else {
    v = r.take(n);
  }
  uint64_t val = v - (1ull << k_exp) + ((uint64_t)(last_q + 1) << k_rice);
  if (val > MAX_SYMBOL) return E_RANGE;
  sym = (uint32_t)val;
  return E_NONE;
}

// Writes one component's scanned coefficients as i16, in increasing index
// order, 16 bytes at a time; the zeros between are written too, so the arena
// needs no clearing. A value outside +-32767 is stored as -32768 with the
// full value beside it in the i32 side arena.
struct CoefOut {
  int16_t *base;
  int32_t *wide;
  uint32_t cur, w0, w1, w2, w3;
  __device__ __forceinline__ void store(uint32_t grp, uint32_t a, uint32_t b, uint32_t c,
                                        uint32_t d) {
    *(uint4 *)(base + 8 * grp) = make_uint4(a, b, c, d);
  }
  __device__ __forceinline__ void put(uint32_t n, int32_t v) {
    uint32_t grp = n >> 3;
    if (grp != cur) {
      store(cur, w0, w1, w2, w3);
      for (uint32_t g = cur + 1; g < grp; g++) store(g, 0, 0, 0, 0);
      cur = grp;
      w0 = w1 = w2 = w3 = 0;
    }
    int32_t h = v;
    if (v > 32767 || v < -32767) {
      h = -32768;
      wide[n] = v;
    }
    uint32_t word = (uint32_t)(uint16_t)(int16_t)h << (16 * (n & 1));
    uint32_t wi = (n >> 1) & 3;
    w0 |= wi == 0 ? word : 0u;
    w1 |= wi == 1 ? word : 0u;
    w2 |= wi == 2 ? word : 0u;
    w3 |= wi == 3 ? word : 0u;
  }
  __device__ __forceinline__ void finish(uint32_t total) {
    store(cur, w0, w1, w2, w3);
    for (uint32_t g = cur + 1; g < total / 8; g++) store(g, 0, 0, 0, 0);
  }
};

// scanned_coefficients() of one colour component (src/decode.rs:484-540).
__device__ uint32_t decode_component(Reader &r, uint32_t nb, CoefOut &out,
                                     const uint16_t *__restrict__ lut_run,
                                     const uint16_t *__restrict__ lut_lvl) {
  uint32_t sym;
  uint32_t e = read_code(r, FIRST_DC_BOOK, sym);
  if (e) return e;
  int32_t prev_dc = symbol_to_signed(sym);
  out.put(0, prev_dc);
  int32_t prev_diff = 3;
  for (uint32_t i = 1; i < nb; i++) {
    uint32_t a = prev_diff < 0 ? (uint32_t)(-prev_diff) : (uint32_t)prev_diff;
    e = read_code(r, DC_BOOK[a < 3 ? a : 3], sym);
    if (e) return e;
    int32_t n = symbol_to_signed(sym);
    int32_t diff = prev_diff < 0 ? -n : n;
    int64_t s = (int64_t)prev_dc + diff;
    if (s > INT32_MAX || s < INT32_MIN) return E_DC_OVERFLOW;
    prev_dc = (int32_t)s;
    out.put(i, prev_dc);
    prev_diff = diff;
  }
  uint32_t total = nb * 64, n = nb, prev_run = 4, prev_lvl = 1;
  for (;;) {
    if (r.remaining() >= 32) {
      // Both codewords from the 9-bit tables: at most 19 bits, all real.
      uint32_t e1 = lut_run[(prev_run < 15 ? prev_run : 15) * 512 + (uint32_t)(r.win >> 55)];
      if (e1) {
        uint32_t l1 = e1 & 31;
        uint64_t w2 = r.win << l1;
        uint32_t e2 = lut_lvl[(prev_lvl < 8 ? prev_lvl : 8) * 512 + (uint32_t)(w2 >> 55)];
        if (e2) {
          uint32_t l2 = e2 & 31, run = e1 >> 5, lsym = e2 >> 5;
          bool neg = ((w2 << l2) >> 63) != 0;
          n += run;
          if (n >= total) return E_RUN_PAST;
          r.consume(l1 + l2 + 1);
          int32_t level = (int32_t)lsym + 1;
          out.put(n, neg ? -level : level);
          n++;
          prev_run = run;
          prev_lvl = lsym;
          continue;
        }
      }
    }
    if (r.end_of_data()) break;
    uint32_t run;
    e = read_code(r, RUN_BOOK[prev_run < 15 ? prev_run : 15], run);
    if (e) return e;
    prev_run = run;
    n += run;
    if (n >= total) return E_RUN_PAST;
    uint32_t lsym;
    e = read_code(r, LEVEL_BOOK[prev_lvl < 8 ? prev_lvl : 8], lsym);
    if (e) return e;
    prev_lvl = lsym;
    if (r.remaining() < 1) return E_PAST_END;
    uint32_t sgn = r.take(1);
    int32_t level = (int32_t)lsym + 1;
    out.put(n, sgn ? -level : level);
    n++;
  }
  out.finish(total);
  return E_NONE;
}

__device__ __forceinline__ uint16_t rescale(uint32_t a, uint32_t identity, uint64_t m) {
  return identity ? (uint16_t)a : (uint16_t)(((uint64_t)a * m + (1ull << 39)) >> 40);
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(256)
    pd_slice_table(const DecPic *__restrict__ pics,
                   const RowSlot *__restrict__ rows, SliceRec *__restrict__ slices,
                   uint32_t *__restrict__ pic_status) {
  const uint32_t p = blockIdx.x, tid = threadIdx.x;
  const DecPic P = pics[p];
  const uint8_t *tab = (const uint8_t *)(P.pic_off + P.header_size);
  const uint64_t data0 = P.header_size + 2ull * P.n_slices;  // relative to the picture
  __shared__ uint32_t wsum[4];
  __shared__ uint64_t carry;
  if (tid == 0) carry = 0;
  __syncthreads();
  const uint32_t lane = tid & 63, wave = tid >> 6;
  for (uint32_t base = 0; base < P.n_slices; base += 256) {
    uint32_t k = base + tid;
    uint32_t size = k < P.n_slices ? ((uint32_t)tab[2 * k] << 8) | tab[2 * k + 1] : 0u;
    uint32_t incl = size;
    for (uint32_t d = 1; d < 64; d <<= 1) {
      uint32_t o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += wsum[w];
    uint64_t excl = carry + before + incl - size;
    if (k < P.n_slices) {
      uint32_t row = k / P.row_slices, j = k - row * P.row_slices;
      RowSlot rs = rows[P.layout_base + j];
      SliceRec s;
      s.off = P.pic_off + data0 + excl;
      s.coef = P.coef_base + (uint64_t)row * P.row_coefs + rs.coef_prefix;
      s.len = size;
      s.pic = p;
      s.mb_x = rs.mb_x;
      s.mb_y = (uint16_t)row;
      s.mbs = rs.mbs;
      slices[P.slice_base + k] = s;
    }
    __syncthreads();
    if (tid == 255) carry += before + incl;
    __syncthreads();
  }
  if (tid == 0) pic_status[p] = (data0 + carry > P.picture_size) ? PIC_TABLE_ERR : PIC_OK;
}

// ---------------------------------------------------------------------------
// grid.x covers slices 256 at a time, grid.y is the stream kind: 0 Y', 1 Cb,
// 2 Cr, 3 alpha.
extern "C" __global__ void __launch_bounds__(256)
    pd_entropy(const DevTables *__restrict__ T,
               const DecFrame *__restrict__ frames, const DecPic *__restrict__ pics,
               const SliceRec *__restrict__ slices, const uint32_t *__restrict__ pic_status,
               uint32_t n_slices, int16_t *__restrict__ coefs, int32_t *__restrict__ coef_wide,
               uint32_t *__restrict__ slice_err, uint32_t *__restrict__ pic_first_err) {
  __shared__ uint16_t lut_run[16 * 512];
  __shared__ uint16_t lut_lvl[9 * 512];
  const uint32_t comp = blockIdx.y, tid = threadIdx.x;
  if (comp < 3) {
    const uint4 *sr = (const uint4 *)T->run_lut;
    const uint4 *sl = (const uint4 *)T->level_lut;
    for (uint32_t i = tid; i < 16 * 512 / 8; i += 256) ((uint4 *)lut_run)[i] = sr[i];
    for (uint32_t i = tid; i < 9 * 512 / 8; i += 256) ((uint4 *)lut_lvl)[i] = sl[i];
    __syncthreads();
  }
  const uint32_t g = blockIdx.x * 256 + tid;
  if (g >= n_slices) return;
  const SliceRec S = slices[g];
  if (pic_status[S.pic] != PIC_OK) return;
  const DecPic &P = pics[S.pic];
  const DecFrame &F = frames[P.frame];
  const uint32_t alpha_bits = F.alpha_bits, chroma = F.chroma;
  if (comp == 3 && alpha_bits == 0) return;

  // Slice header (src/decode.rs:301-338).
  const uint8_t *s = (const uint8_t *)S.off;
  const uint32_t slen = S.len;
  const uint32_t min_header = alpha_bits ? 8u : 6u;
  uint32_t err = E_NONE, hs = 0, ysz = 0, cbsz = 0, crsz = 0;
  if (slen < min_header) {
    err = E_SLICE_SHORT;
  } else {
    hs = s[0] >> 3;
    uint32_t q = s[1];
    if (hs < min_header || hs > slen) {
      err = E_SLICE_HDR_SIZE | (hs << 8);
    } else if (q < 1 || q > 224) {
      err = E_QINDEX | (q << 8);
    } else {
      ysz = ((uint32_t)s[2] << 8) | s[3];
      cbsz = ((uint32_t)s[4] << 8) | s[5];
      uint32_t body = slen - hs;
      if (alpha_bits) {
        crsz = ((uint32_t)s[6] << 8) | s[7];
      } else if (ysz + cbsz > body) {
        err = E_COMP_SIZES;
      } else {
        crsz = body - ysz - cbsz;
      }
      if (!err && ysz + cbsz + crsz > body) err = E_COMP_SIZES;
    }
  }
  if (!err) {
    Reader r;
    if (comp < 3) {
      const uint32_t cpm = chroma == 2 ? 2u : 4u;
      const uint32_t nb = (comp == 0 ? 4u : cpm) * S.mbs;
      const uint32_t off = comp == 0 ? hs : (comp == 1 ? hs + ysz : hs + ysz + cbsz);
      const uint32_t len = comp == 0 ? ysz : (comp == 1 ? cbsz : crsz);
      const uint64_t cofs = S.coef + (comp == 0 ? 0ull : (comp == 1 ? 256ull * S.mbs : (256ull + 64ull * cpm) * S.mbs));
      r.init(s + off, len);
      CoefOut out{coefs + cofs, coef_wide + cofs, 0, 0, 0, 0, 0};
      err = decode_component(r, nb, out, lut_run, lut_lvl);
    } else {
      // scanned_alpha() (src/decode.rs:570-595) and its placement (365-389).
      const uint32_t used = hs + ysz + cbsz + crsz;
      r.init(s + used, slen - used);
      const uint32_t bits = alpha_bits, mag = bits == 8 ? 3u : 6u, mask = (1u << bits) - 1;
      const uint32_t log2c = 4 + (31 - __builtin_clz(S.mbs));
      const uint32_t cols = 1u << log2c, count = cols * 16;
      const uint32_t rows =
          (uint32_t)S.mb_y + 1 < P.height_in_mb ? 16u : P.height - 16u * (P.height_in_mb - 1);
      const uint32_t needed = cols * rows;
      const uint32_t x0 = 16u * S.mb_x, width = F.width;
      const uint32_t nvis = cols < width - x0 ? cols : width - x0;
      uint16_t *aout = (uint16_t *)F.out_alpha;
      const uint32_t ident = F.rescale_identity;
      const uint64_t m = F.rescale_m;
      uint32_t len = 0;
      int32_t prev = -1;
      while (len < count) {
        if (len >= needed && r.end_of_data()) break;
        int32_t d;
        if (r.remaining() < 1) { err = E_PAST_END; break; }
        if (r.take(1)) {
          if (r.remaining() < bits) { err = E_PAST_END; break; }
          d = (int32_t)r.take(bits);
        } else {
          if (r.remaining() < mag) { err = E_PAST_END; break; }
          int32_t a = (int32_t)r.take(mag) + 1;
          if (r.remaining() < 1) { err = E_PAST_END; break; }
          d = r.take(1) ? -a : a;
        }
        uint32_t alpha = (uint32_t)(prev + d) & mask;
        prev = (int32_t)alpha;
        uint32_t run;
        if (r.remaining() < 1) { err = E_PAST_END; break; }
        if (r.take(1)) {
          run = 1;
        } else {
          if (r.remaining() < 4) { err = E_PAST_END; break; }
          uint32_t v = r.take(4);
          if (v) {
            run = v + 1;
          } else {
            if (r.remaining() < 11) { err = E_PAST_END; break; }
            run = r.take(11) + 1;
          }
        }
        if (len + run > count) { err = E_ALPHA_RUN; break; }
        const uint16_t val = rescale(alpha, ident, m);
        uint32_t end = len + run;
        if (len < needed) {
          uint32_t stop = end < needed ? end : needed;
          for (uint32_t i = len; i < stop;) {
            uint32_t row = i >> log2c, col = i & (cols - 1);
            uint32_t seg_end = (row + 1) << log2c;
            uint32_t e2 = seg_end < stop ? seg_end : stop;
            uint32_t c_end = col + (e2 - i);
            if (c_end > nvis) c_end = nvis;
            uint16_t *o = aout + (size_t)(P.first_row + P.row_step * (16u * S.mb_y + row)) * width + x0;
            for (uint32_t c = col; c < c_end; c++) o[c] = val;
            i = e2;
          }
        }
        len = end;
      }
    }
  }
  if (err) {
    slice_err[4 * (size_t)g + comp] = err;
    atomicMin(&pic_first_err[S.pic], g - P.slice_base);
  }
}

// ---------------------------------------------------------------------------
// One workgroup per slice; 8 lanes per block (a row each in the row pass, an
// output row each in the column pass), 32 blocks per pass.
extern "C" __global__ void __launch_bounds__(256)
    pd_idct(const DevTables *__restrict__ T,
            const DecFrame *__restrict__ frames, const DecPic *__restrict__ pics,
            const SliceRec *__restrict__ slices, const uint32_t *__restrict__ pic_status,
            const uint32_t *__restrict__ slice_err, const int16_t *__restrict__ coefs,
            const int32_t *__restrict__ coef_wide) {
  __shared__ __attribute__((aligned(16))) int16_t sc[96 * 64];
  __shared__ float tmp[32 * 64];
  const uint32_t g = blockIdx.x, tid = threadIdx.x;
  const SliceRec S = slices[g];
  if (pic_status[S.pic] != PIC_OK) return;
  const uint4 ev = *(const uint4 *)(slice_err + 4 * (size_t)g);
  if (ev.x | ev.y | ev.z | ev.w) return;
  const DecPic &P = pics[S.pic];
  const DecFrame &F = frames[P.frame];
  const uint32_t chroma = F.chroma, cpm = chroma == 2 ? 2u : 4u, mbs = S.mbs;
  const uint32_t nbY = 4 * mbs, nbC = cpm * mbs, TB = nbY + 2 * nbC;
  const uint32_t ncoef = TB * 64;
  {
    const uint4 *src = (const uint4 *)(coefs + S.coef);
    for (uint32_t i = tid; i < ncoef / 8; i += 256) ((uint4 *)sc)[i] = src[i];
  }
  const uint32_t q = qscale(((const uint8_t *)S.off)[1]);
  const uint8_t *scan = T->scan[F.interlaced ? 1 : 0];
  const float *basis = T->basis;
  const float osc = F.out_scale, obias = F.out_bias, omax = F.out_max;
  __syncthreads();
  const uint32_t r = tid & 7, lb = tid >> 3;
  for (uint32_t t0 = 0; t0 < TB; t0 += 32) {
    const uint32_t t = t0 + lb;
    const bool active = t < TB;
    uint32_t c = 0, b = t, nb = nbY, cbase = 0;
    if (t >= nbY + nbC) {
      c = 2; b = t - nbY - nbC; nb = nbC; cbase = 64 * (nbY + nbC);
    } else if (t >= nbY) {
      c = 1; b = t - nbY; nb = nbC; cbase = 64 * nbY;
    }
    if (active) {
      const uint8_t *W = c == 0 ? F.luma_w : F.chroma_w;
      float fu[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) {
        const uint32_t k = 8 * r + u;
        const uint32_t idx = cbase + nb * scan[k] + b;
        int32_t v = sc[idx];
        if (v == -32768) v = coef_wide[S.coef + idx];
        fu[u] = (float)v * ((float)((uint32_t)W[k] * q) / 8.0f);
      }
#pragma unroll
      for (uint32_t x = 0; x < 8; x++) {
        float acc = 0.0f;
#pragma unroll
        for (uint32_t u = 0; u < 8; u++) acc = acc + fu[u] * basis[8 * x + u];
        tmp[lb * 64 + 8 * r + x] = acc;
      }
    }
    __syncthreads();
    if (active) {
      uint16_t outv[8];
#pragma unroll
      for (uint32_t x = 0; x < 8; x++) {
        float acc = 0.0f;
#pragma unroll
        for (uint32_t v = 0; v < 8; v++) acc = acc + basis[8 * r + v] * tmp[lb * 64 + 8 * v + x];
        float sv = floorf(acc * osc + obias);
        sv = sv < 0.0f ? 0.0f : (sv > omax ? omax : sv);
        outv[x] = (uint16_t)(uint32_t)sv;
      }
      // block_position (src/decode.rs:395-413) and put_block (418-438).
      const uint32_t per_mb = c == 0 ? 4u : cpm;
      const uint32_t mbx = S.mb_x + b / per_mb, bb = b % per_mb;
      uint32_t bx, by;
      if (c == 0) {
        bx = 16 * mbx + 8 * (bb & 1); by = 16u * S.mb_y + 8 * (bb >> 1);
      } else if (chroma == 2) {
        bx = 8 * mbx; by = 16u * S.mb_y + 8 * bb;
      } else {
        bx = 16 * mbx + 8 * (bb >> 1); by = 16u * S.mb_y + 8 * (bb & 1);
      }
      const uint32_t pw = F.plane_w[c];
      const uint32_t prow = by + r;
      if (bx < pw && prow < P.height) {
        uint16_t *o = (uint16_t *)F.out_plane[c] + (size_t)(P.first_row + P.row_step * prow) * pw + bx;
        const uint32_t cols = pw - bx < 8 ? pw - bx : 8u;
        if (cols == 8 && (((uintptr_t)o) & 15) == 0) {
          uint4 pk;
          pk.x = outv[0] | ((uint32_t)outv[1] << 16);
          pk.y = outv[2] | ((uint32_t)outv[3] << 16);
          pk.z = outv[4] | ((uint32_t)outv[5] << 16);
          pk.w = outv[6] | ((uint32_t)outv[7] << 16);
          *(uint4 *)o = pk;
        } else {
          for (uint32_t x = 0; x < cols; x++) o[x] = outv[x];
        }
      }
    }
    __syncthreads();
  }
}
