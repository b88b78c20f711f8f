// gfx950 encode kernels for the ProRes slice engine (src/encode.rs).
//
//   pe_slices    one thread per slice: the slice records from each
//                picture's row layout.
//   pe_fdct      one workgroup per slice: edge-replicated 8x8 gathers, the
//                f32 forward DCT in the reference's order (src/dct.rs:70-93,
//                src/dsp/scalar.rs:14-17), scattered to scanned-interleaved
//                order per component (src/encode.rs:518-555).
//   pe_alpha     one wavefront per slice: alpha rescaled and run/difference
//                coded in parallel (src/encode.rs:838-868, 457-470).
//   pe_size      one wavefront per slice: bytes at the picture's pending probe
//                qi (quantise + exact bit count, src/encode.rs:671-681,
//                398-409), kept per slice and summed per picture;
//                pe_size_first the same at the first five probes at once.
//   pe_search    one lane per picture: the steps of the binary search over
//                qi (src/encode.rs:685-713) that the pass settles.
//   pe_window    one wavefront per slice: the share-out's size table, from
//                the probes at uniform and uniform-1 and, for the slices that
//                may step further, the next finer quantisers (down to qi 1
//                while the slice stays a candidate, the deeper ones into the
//                size cache).
//   pe_shareout  one workgroup per picture: the sequential slice-by-slice
//                share-out (src/encode.rs:769-793) solved 64 slices at a time
//                by fixed-point iteration of its triangular recurrence (one
//                wave); a size neither the table nor the cache holds is coded
//                by all the waves, four quantisers a request, one per wave,
//                and cached.
//   pe_offsets   one workgroup per frame: slice and picture offsets.
//   pe_arena     one workgroup: each frame's place in the output arena.
//   pe_pack      one wavefront per slice: quantise at the chosen qi and place
//                every codeword with a wave prefix sum over their lengths.
//
// Built with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "enc_layout.h"
#include "layout.h"

using namespace prores_hip;

// Global memory reached through an address held as an integer (frame
// planes, the output arena): a generic (flat) access would also count on the
// LDS counter, and every LDS wait after it would wait for it too.
#define GLOBAL __attribute__((address_space(1)))

namespace {

__device__ __forceinline__ uint32_t clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
__device__ __forceinline__ uint32_t bitlen64(uint64_t v) { return 64 - clz64(v); }

// Codebooks as combination codes (lastRiceQ, kRice, kExp), packed 4 bits per
// context as lastRiceQ | kRice << 2; kExp is kRice + 1 except for level
// context 0, RICE_EXP_COMBO(2, 0, 2). Tables 9-11 (src/vlc.rs:30-68).
constexpr uint64_t RUN_P = 0x8444444555501122ull;
constexpr uint64_t LVL_P = 0x844440212ull;
constexpr uint32_t DC_P = 0xC940u;

struct Book {
  uint32_t lq, kr, ke;
};
__device__ __forceinline__ Book run_book(uint32_t ctx) {
  uint32_t nib = (uint32_t)(RUN_P >> (4 * ctx)) & 15;
  return Book{nib & 3, nib >> 2, (nib >> 2) + 1};
}
__device__ __forceinline__ Book lvl_book(uint32_t ctx) {
  uint32_t nib = (uint32_t)(LVL_P >> (4 * ctx)) & 15;
  return Book{nib & 3, nib >> 2, (nib >> 2) + 1 + (ctx == 0 ? 1u : 0u)};
}
__device__ __forceinline__ Book dc_book(uint32_t ctx) {
  uint32_t nib = (DC_P >> (4 * ctx)) & 15;
  return Book{nib & 3, nib >> 2, (nib >> 2) + 1};
}
constexpr Book FIRST_DC = {0, 5, 6};  // EG(5)

// Codebook::len (src/vlc.rs:155-172).
__device__ __forceinline__ uint32_t book_len(Book b, uint32_t n) {
  uint32_t th = (b.lq + 1) << b.kr;
  if (n < th) return (n >> b.kr) + 1 + b.kr;
  uint64_t v = (uint64_t)(n - th) + (1ull << b.ke);
  return b.lq + 1 + 2 * bitlen64(v) - b.ke - 1;
}
// Codebook::code (src/vlc.rs:131-151) in 32-bit arithmetic and without branches, for n < 2^30
// (every symbol pe_pack codes is below 2^17): the codeword's value (its low
// `len` bits are the codeword; the bits above the value are zeros).
__device__ __forceinline__ uint32_t book_code32(Book b, uint32_t n, uint32_t &len) {
  const uint32_t th = (b.lq + 1) << b.kr;
  const uint32_t v = n - th + (1u << b.ke);
  const bool rice = n < th;
  len = rice ? (n >> b.kr) + 1 + b.kr : b.lq + 2 * (32 - (uint32_t)__clz((int)v)) - b.ke;
  return rice ? (1u << b.kr) | (n & ((1u << b.kr) - 1)) : v;
}

__device__ __forceinline__ uint32_t signed_to_symbol(int32_t n) {
  return n >= 0 ? 2u * (uint32_t)n : 2u * (0u - (uint32_t)n) - 1u;
}

// The scalar quantiser (src/dsp/scalar.rs:19-39): QF = sign(x) floor(|x| + r),
// x = F inv, with a saturating float-to-int conversion.
__device__ __forceinline__ int32_t quant(float f, float inv, float r) {
  float x = f * inv;
  float t = floorf(fabsf(x) + r);
  int32_t m = t >= 2147483648.0f ? 0x7FFFFFFF : (int32_t)t;
  return x < 0.0f ? -m : m;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = __shfl_xor((uint32_t)v, o, 64), hi = __shfl_xor((uint32_t)(v >> 32), o, 64);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}
// Wave-wide inclusive prefix sum over lanes by DPP row shifts (within rows
// of 16: +1, +2, +3 of the input, then +4 and +8 of the partial sums on the
// upper banks) and row broadcasts (lane 15 into rows 1 and 3, lane 31 into
// rows 2 and 3). Every lane of the wave must be active.
template <int CTRL, int ROWS, int BANKS>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, BANKS, false);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  uint32_t s = v + dpp_move<0x111, 0xf, 0xf>(v);  // row_shr:1
  s += dpp_move<0x112, 0xf, 0xf>(v);              // row_shr:2
  s += dpp_move<0x113, 0xf, 0xf>(v);              // row_shr:3
  s += dpp_move<0x114, 0xf, 0xe>(s);              // row_shr:4, banks 1-3
  s += dpp_move<0x118, 0xf, 0xc>(s);              // row_shr:8, banks 2-3
  s += dpp_move<0x142, 0xa, 0xf>(s);              // row_bcast:15, rows 1 and 3
  s += dpp_move<0x143, 0xc, 0xf>(s);              // row_bcast:31, rows 2 and 3
  return s;
}
// The wave's total (every lane active), a wave-uniform value.
__device__ __forceinline__ uint32_t wave_total(uint32_t v) {
  return __builtin_amdgcn_readlane(wave_incl_scan(v), 63);
}

__device__ __forceinline__ uint32_t excl_scan(uint32_t v, uint32_t lane) {
  uint32_t incl = v;
  for (uint32_t d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  return incl - v;
}
__device__ __forceinline__ int64_t excl_scan64(int64_t v, uint32_t lane) {
  int64_t incl = v;
  for (uint32_t d = 1; d < 64; d <<= 1) {
    uint32_t lo = __shfl_up((uint32_t)(uint64_t)incl, d, 64);
    uint32_t hi = __shfl_up((uint32_t)((uint64_t)incl >> 32), d, 64);
    if (lane >= d) incl += (int64_t)(((uint64_t)hi << 32) | lo);
  }
  return incl - v;
}

__device__ __forceinline__ uint32_t ilog2(uint32_t v) { return 31 - __builtin_clz(v); }
// The wavefront's index in its workgroup, as a wave-uniform value: what
// follows from it (the slice, its records, its chunk walk) is then kept in
// scalar registers and worked on by the scalar unit.
__device__ __forceinline__ uint32_t wave_index() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
// Table 15 of RDD 36 (src/tables.rs:35-38).
__device__ __forceinline__ uint32_t qscale(uint32_t q) { return q <= 128 ? q : 128 + 4 * (q - 128); }

// Component c of a slice: its block count and its offset inside the slice's
// coefficients (Y', then Cb, then Cr; src/encode.rs:557-564).
__device__ __forceinline__ void component(uint32_t c, uint32_t mbs, uint32_t cpm, uint32_t &nb,
                                          uint32_t &off) {
  nb = (c == 0 ? 4u : cpm) * mbs;
  off = c == 0 ? 0u : (c == 1 ? 256u * mbs : (256u + 64u * cpm) * mbs);
}

// Length of the combination code (lq, kr, ke) of n (Codebook::len,
// src/vlc.rs:155-172) without branches: the Rice length below the threshold,
// else the exp-Golomb length, lq + 1 + 2 bitlen(n - th + 2^ke) - ke - 1
// (n <= 2^30, so the sum fits 32 bits; the unused side may wrap).
__device__ __forceinline__ uint32_t code_len(uint32_t lq, uint32_t kr, uint32_t ke, uint32_t n) {
  const uint32_t th = (lq + 1) << kr;
  const uint32_t rice = (n >> kr) + 1 + kr;
  const uint32_t eg = lq + 2 * (32 - (uint32_t)__clz((int)(n - th + (1u << ke)))) - ke;
  return n < th ? rice : eg;
}
__device__ __forceinline__ uint32_t run_code_len(uint32_t ctx, uint32_t n) {
  const uint32_t nib = (uint32_t)(RUN_P >> (4 * ctx)) & 15, kr = nib >> 2;
  return code_len(nib & 3, kr, kr + 1, n);
}
__device__ __forceinline__ uint32_t level_code_len(uint32_t ctx, uint32_t n) {
  const uint32_t nib = (uint32_t)(LVL_P >> (4 * ctx)) & 15, kr = nib >> 2;
  return code_len(nib & 3, kr, kr + 1 + (ctx == 0 ? 1u : 0u), n);
}

// The encoder's code-length tables (DevTables::enc_run_len, enc_level_len)
// into LDS, by the whole workgroup.
__device__ __forceinline__ void load_len_tables(const DevTables *__restrict__ T, uint32_t *lens) {
  const uint32_t *src = (const uint32_t *)&T->enc_run_len[0][0];
  for (uint32_t i = threadIdx.x; i < ENC_LEN_BYTES / 4; i += blockDim.x) lens[i] = src[i];
  __syncthreads();
}

// A chunk of 64 coefficients whose bound (pe_fdct: the largest
// |F| * fl(8 / W) over the chunk) is at most qScale(qi) * ZERO_CHUNK_LIMIT
// quantises to zeros at qi. In f32, fl(|F| fl(8 / (W q))) is at most
// bound / q * (1 + 5 * 2^-24), so below 0.58 - 10^-4, and adding r = 0.42f
// stays under 1: floor gives 0 for every AC coefficient of the chunk. The
// first chunk of a component (its DC coefficients) is always coded.
constexpr float ZERO_CHUNK_LIMIT = 0.5798f;

// Which chunk slice_bytes_q counts: the DC chunk (the first of a component)
// or another.
struct DcChunk {
  static constexpr bool value = true;
};
struct AcChunk {
  static constexpr bool value = false;
};

// Bytes of one slice at each of the quantisers qis[0..nq), one wavefront,
// and whether every size field can hold them (src/encode.rs:671-681): per
// component, the exact bit count of coefficient_bits (src/encode.rs:366-409)
// over the quantised coefficients (src/dsp/scalar.rs:19-39), a 64-coefficient
// chunk of the scanned array a step, one coefficient a lane.
// - Chunks that quantise to zeros (by their bounds `cb`) are skipped: they
//   add no bits and leave the run/level contexts as they are. The chunks are
//   walked once for every quantiser, each chunk loaded once and counted for
//   the quantisers at which it may hold a nonzero.
// - The inverse quantiser of a lane's scan position comes from a register
//   (lane k holds the component's entry k) by a lane permute, not a load.
// - The next coded chunk's coefficients are loaded while this one is counted.
// - Run and level contexts travel as one packed word (min(run, 15) |
//   min(level, 8) << 4): one permute from the previous nonzero lane, and the
//   chunk's last nonzero passes its word to the next chunk by a readlane.
// - Code lengths come from the LDS tables `lens` (runs below 64, levels
//   below 256); a chunk with a larger symbol also takes the closed form.
// No conversion here saturates: a coefficient of fdct_load's output is at
// most 64 * 256 * (1/2)^2 = 4096 in magnitude and an inverse quantiser at
// most 8 / 2, so |x| + r < 2^15.
template <int NQ>
__device__ __forceinline__ void slice_bytes_q(const float *__restrict__ coef, const float *__restrict__ cb,
                                              uint32_t mbs, uint32_t cpm, const float *__restrict__ qt,
                                              const uint32_t *qis, uint32_t nq, uint32_t hb, uint32_t alen,
                                              const uint8_t *lens, uint32_t *size, bool *fits) {
  const uint32_t L = threadIdx.x & 63;
  const uint64_t below = (1ull << L) - 1;
  const uint8_t *run_lens = lens, *level_lens = lens + 16 * 64;
  // Chunks that may hold a nonzero at qis[k]: bit n of live0[k] / live1[k]
  // for chunk n / chunk 64 + n (a slice has at most 96 chunks, and no
  // component crosses chunk 64: slices have 1, 2, 4 or 8 macroblocks).
  const uint32_t TB = (4 + 2 * cpm) * mbs;
  const float b0 = cb[L < TB ? L : 0], b1 = cb[L + 64 < TB ? L + 64 : 0];
  uint64_t live0[NQ], live1[NQ], any0 = 0, any1 = 0;
#pragma unroll
  for (int k = 0; k < NQ; k++) {
    const float qlim = (float)qscale(qis[k]) * ZERO_CHUNK_LIMIT;
    live0[k] = (uint32_t)k < nq ? __ballot(L < TB && b0 > qlim) : 0ull;
    live1[k] = (uint32_t)k < nq ? __ballot(L + 64 < TB && b1 > qlim) : 0ull;
    any0 |= live0[k];
    any1 |= live1[k];
    size[k] = hb + alen;
    fits[k] = true;
  }
  uint32_t k0 = 0;  // the component's first chunk in the slice
#pragma unroll 1
  for (uint32_t c = 0; c < 3; c++) {
    uint32_t nb, off;
    component(c, mbs, cpm, nb, off);
    const uint32_t lg = ilog2(nb);
    const uint64_t in_comp = (1ull << nb) - 1;
    uint64_t todo = ((((k0 >> 6) ? any1 : any0) >> (k0 & 63)) & in_comp) | 1;  // and the DC chunk
    uint64_t lv[NQ];
    float invL[NQ];
    uint32_t bits[NQ], last_idx[NQ], ctx[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      lv[k] = ((((k0 >> 6) ? live1[k] : live0[k]) >> (k0 & 63)) & in_comp) | 1;
      invL[k] = (uint32_t)k < nq ? qt[(qis[k] * 2 + (c > 0 ? 1u : 0u)) * 64 + L] : 0.0f;
      bits[k] = 0;
      last_idx[k] = nb - 1;
      ctx[k] = 4u | 1u << 4;  // no nonzero yet: run context 4, level context 1
    }
    const float *cf = coef + off + L;
    // One chunk, counted for every quantiser it may hold a nonzero at. The
    // DC chunk (j = 0) also codes the DC coefficients, its lanes below nb;
    // in every other chunk each lane's scan position is past 0, so each
    // coefficient rounds with 0.42 and the nonzero test is the quantised
    // value's alone.
    auto count_chunk = [&](auto dc, uint32_t j, float f) {
      constexpr bool DC = decltype(dc)::value;
      const uint32_t base = j * 64, i = base + L, s = i >> lg;
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        if ((uint32_t)k >= nq || !((lv[k] >> j) & 1)) continue;
        const float inv = __shfl(invL[k], (int)s, 64);
        const float x = f * inv;
        const uint32_t m = (uint32_t)(int32_t)floorf(fabsf(x) + (DC && s == 0 ? 0.5f : 0.42f));
        bool nz = m != 0;
        if (DC) {  // the DC coefficients (src/encode.rs:383-395), nb <= 32 of the first chunk
          const int32_t q = x < 0.0f ? -(int32_t)m : (int32_t)m;
          const int32_t d1 = __shfl(q, L ? (int)L - 1 : 0, 64);
          const int32_t d2 = __shfl(q, L > 1 ? (int)L - 2 : 0, 64);
          if (L < nb) {
            if (L == 0) {
              bits[k] += book_len(FIRST_DC, signed_to_symbol(q));
            } else {
              const int32_t diff = (int32_t)((uint32_t)q - (uint32_t)d1);
              const int32_t pd = L == 1 ? 3 : (int32_t)((uint32_t)d1 - (uint32_t)d2);
              const int32_t nn = pd < 0 ? (int32_t)(0u - (uint32_t)diff) : diff;
              const uint32_t a = pd < 0 ? 0u - (uint32_t)pd : (uint32_t)pd;
              bits[k] += book_len(dc_book(a < 3 ? a : 3), signed_to_symbol(nn));
            }
          }
          nz = nz && L >= nb;
        }
        const uint64_t mask = __builtin_amdgcn_ballot_w64(nz);
        if (mask) {
          const uint64_t lower = mask & below;
          const int P = lower ? 63 - (int)__builtin_clzll(lower) : -1;
          const uint32_t prev_idx = P >= 0 ? base + (uint32_t)P : last_idx[k];
          const uint32_t run = i - prev_idx - 1, lvl = m - 1;
          const uint32_t mine = (run < 15 ? run : 15u) | (lvl < 8 ? lvl : 8u) << 4;
          uint32_t pc = __shfl(mine, P >= 0 ? P : (int)L, 64);
          if (P < 0) pc = ctx[k];
          uint32_t pair = run_lens[(pc & 15) << 6 | (run & 63)] + level_lens[(pc >> 4) << 8 | (lvl & 255)] + 1;
          // Symbols past the tables (each ballot of a plain compare is the
          // compare's own mask).
          const uint64_t bigs = (__builtin_amdgcn_ballot_w64(run >= 64) | __builtin_amdgcn_ballot_w64(lvl >= 256)) & mask;
          if (bigs) {
            if ((bigs >> L) & 1) pair = run_code_len(pc & 15, run) + level_code_len(pc >> 4, lvl) + 1;
          }
          if (nz) bits[k] += pair;
          const uint32_t last = 63 - (uint32_t)__builtin_clzll(mask);
          last_idx[k] = base + last;
          ctx[k] = __builtin_amdgcn_readlane(mine, last);
        }
      }
    };
    // The chunks in order, the next coded one loaded while this one is
    // counted (past the last, the last is loaded again).
    float f = cf[0];
    todo &= todo - 1;  // the DC chunk, bit 0
    uint32_t jn = todo ? (uint32_t)__builtin_ctzll(todo) : 0u;
    float fn = cf[jn * 64];
    count_chunk(DcChunk{}, 0u, f);
    while (todo) {
      const uint32_t j = jn;
      f = fn;
      todo &= todo - 1;
      jn = todo ? (uint32_t)__builtin_ctzll(todo) : j;
      fn = cf[jn * 64];
      count_chunk(AcChunk{}, j, f);
    }
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      if ((uint32_t)k >= nq) continue;
      const uint32_t bytes = (wave_total(bits[k]) + 7) / 8;
      size[k] += bytes;
      if (bytes > 65535) fits[k] = false;
    }
    k0 += nb;
  }
#pragma unroll
  for (int k = 0; k < NQ; k++)
    if (size[k] > 65535) fits[k] = false;
}

// Bytes of one slice at quantiser qi (slice_bytes_q for one quantiser).
__device__ void slice_bytes(const float *__restrict__ coef, const float *__restrict__ cb, uint32_t mbs,
                            uint32_t cpm, const float *__restrict__ qt, uint32_t qi, uint32_t hb, uint32_t alen,
                            const uint8_t *lens, uint32_t &size, bool &fits) {
  const uint32_t q[1] = {qi};
  uint32_t sz[1];
  bool ft[1];
  slice_bytes_q<1>(coef, cb, mbs, cpm, qt, q, 1, hb, alen, lens, sz, ft);
  size = sz[0];
  fits = ft[0];
}

__device__ __forceinline__ uint16_t rescale(uint32_t a, uint32_t identity, uint64_t m) {
  return identity ? (uint16_t)a : (uint16_t)(((uint64_t)a * m + (1ull << 39)) >> 40);
}

}  // namespace

// ---------------------------------------------------------------------------
// The slice records, one thread a slice, blockIdx.y the picture: each
// picture's rows of slices laid out as its row slots, coefficients
// contiguous in row-major order.
extern "C" __global__ void __launch_bounds__(256)
    pe_slices(const EncFrame *__restrict__ frames, const EncPic *__restrict__ pics,
              const EncRowSlot *__restrict__ slots, EncSlice *__restrict__ slices) {
  const EncPic &P = pics[blockIdx.y];
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= P.n_slices) return;
  const uint32_t row = k / P.row_slices, j = k - row * P.row_slices;
  const EncRowSlot rs = slots[P.layout_base + j];
  const uint32_t per_mb = 64 * (frames[P.frame].chroma == 2 ? 8u : 12u);
  EncSlice S;
  S.coef = P.coef_base + ((uint64_t)row * P.row_mbs + rs.mbs_before) * per_mb;
  S.pic = blockIdx.y;
  S.mb_x = rs.mb_x;
  S.mb_y = (uint16_t)row;
  S.mbs = rs.mbs;
  S.mbs_from = (P.rows - row - 1) * P.row_mbs + (P.row_mbs - rs.mbs_before);
  slices[P.slice_base + k] = S;
}

// ---------------------------------------------------------------------------
// MAXB: the most blocks a slice of the batch has (64 for 4:2:2 slices of 8
// macroblocks, 96 for 4:4:4): LDS sized for it, so 4:2:2 batches fit six
// workgroups a CU where 96-block LDS fits five (static LDS: the same size
// as dynamic LDS measured slower, 1.16 against 1.00 ms for the 32-frame
// 2160p HQ batch).
template <uint32_t MAXB>
__device__ __forceinline__ void fdct_slice(const DevTables *__restrict__ T, const EncFrame *__restrict__ frames,
                                           const EncPic *__restrict__ pics, const EncSlice *__restrict__ slices,
                                           const float *__restrict__ qtabs, float *__restrict__ coefs,
                                           float *__restrict__ cbound) {
  __shared__ __attribute__((aligned(16))) float outc[MAXB * 64];
  __shared__ float tmp[32 * 64];
  const uint32_t g = blockIdx.x, tid = threadIdx.x;
  const EncSlice S = slices[g];
  const EncPic &P = pics[S.pic];
  const EncFrame &F = frames[P.frame];
  const uint32_t chroma = F.chroma, cpm = chroma == 2 ? 2u : 4u, mbs = S.mbs;
  const uint32_t nbY = 4 * mbs, nbC = cpm * mbs, TB = nbY + 2 * nbC;
  const uint8_t *scan = T->scan[P.interlaced ? 1 : 0];
  const float *basis = T->basis;
  const float scale = F.in_scale;
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
      // block_position (src/decode.rs:395-413), PictureLayout::block (src/encode.rs:488-505).
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
      const uint32_t prow = by + r < P.height - 1 ? by + r : P.height - 1;
      const GLOBAL uint16_t *src =
          (const GLOBAL uint16_t *)(uintptr_t)F.in_plane[c] + (size_t)(P.first_row + P.row_step * prow) * pw;
      float p[8];
      if (bx + 8 <= pw && ((uintptr_t)(src + bx) & 15) == 0) {  // the row's 8 samples in one load
        const uint4 v = *(const GLOBAL uint4 *)(src + bx);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (uint32_t x = 0; x < 8; x++) p[x] = (float)((w[x >> 1] >> (16 * (x & 1))) & 0xFFFFu) * scale - 256.0f;
      } else {  // at the right edge (the last column repeated) or unaligned
#pragma unroll
        for (uint32_t x = 0; x < 8; x++) {
          const uint32_t col = bx + x < pw - 1 ? bx + x : pw - 1;
          p[x] = (float)src[col] * scale - 256.0f;
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) {
        float acc = 0.0f;
#pragma unroll
        for (uint32_t x = 0; x < 8; x++) acc = acc + p[x] * basis[8 * x + u];
        tmp[lb * 64 + 8 * r + u] = acc;
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) {
        float acc = 0.0f;
#pragma unroll
        for (uint32_t y = 0; y < 8; y++) acc = acc + basis[8 * y + r] * tmp[lb * 64 + 8 * y + u];
        outc[cbase + nb * scan[8 * r + u] + b] = acc;
      }
    }
    __syncthreads();
  }
  float *dst = coefs + S.coef;
  for (uint32_t i = tid; i < TB * 16; i += 256) ((float4 *)dst)[i] = ((const float4 *)outc)[i];  // 16 bytes a store
  // Each chunk's bound for the size count: the largest |F| * fl(8 / W) of
  // its 64 coefficients (the quantiser table's row for qi = 1), four
  // threads a chunk.
  const float *qt1 = qtabs + (size_t)P.qtab * 225 * 128 + 2 * 64;
  for (uint32_t k0 = 0; k0 < TB; k0 += 64) {
    const uint32_t k = k0 + (tid >> 2), part = tid & 3;
    float mx = 0.0f;
    if (k < TB) {
      const uint32_t c = k < nbY ? 0u : (k < nbY + nbC ? 1u : 2u);
      const uint32_t kc = c == 0 ? 0u : (c == 1 ? nbY : nbY + nbC), nb = c == 0 ? nbY : nbC;
      const uint32_t lg = 31 - __builtin_clz(nb);
      const float *inv1 = qt1 + (c > 0 ? 64 : 0);
      for (uint32_t e = 16 * part; e < 16 * part + 16; e++) {
        const float v = fabsf(outc[64 * k + e]) * inv1[(64 * (k - kc) + e) >> lg];
        mx = v > mx ? v : mx;
      }
    }
    const float m1 = __shfl_xor(mx, 1, 64);
    mx = m1 > mx ? m1 : mx;
    const float m2 = __shfl_xor(mx, 2, 64);
    mx = m2 > mx ? m2 : mx;
    if (k < TB && part == 0) cbound[S.coef / 64 + k] = mx;
  }
}

extern "C" __global__ void __launch_bounds__(256)
    pe_fdct(const DevTables *__restrict__ T, const EncFrame *__restrict__ frames,
            const EncPic *__restrict__ pics, const EncSlice *__restrict__ slices,
            const float *__restrict__ qtabs, float *__restrict__ coefs, float *__restrict__ cbound) {
  fdct_slice<96>(T, frames, pics, slices, qtabs, coefs, cbound);
}
extern "C" __global__ void __launch_bounds__(256)
    pe_fdct64(const DevTables *__restrict__ T, const EncFrame *__restrict__ frames,
              const EncPic *__restrict__ pics, const EncSlice *__restrict__ slices,
              const float *__restrict__ qtabs, float *__restrict__ coefs, float *__restrict__ cbound) {
  fdct_slice<64>(T, frames, pics, slices, qtabs, coefs, cbound);
}

// ---------------------------------------------------------------------------
// One wavefront per slice. slice_alpha (src/encode.rs:838-868): the slice's
// whole macroblocks, 16 rows of 16 * mbs values, the last column and row
// repeated past the picture, rescaled to the coded depth; then encode_alpha
// (src/encode.rs:457-470) in parallel: every lane codes the runs that start
// in its 1/64 of the values (a run ends at the next start, found across lanes
// with a suffix minimum), a prefix sum places each lane's bits, and the bits
// are OR-ed into an LDS buffer that is copied out word by word.
extern "C" __global__ void __launch_bounds__(64)
    pe_alpha(const EncFrame *__restrict__ frames, const EncPic *__restrict__ pics,
             const EncSlice *__restrict__ slices, uint32_t n_slices, uint8_t *__restrict__ scratch,
             uint32_t *__restrict__ alpha_len, EncState *__restrict__ state) {
  __shared__ uint16_t vals[2048];
  __shared__ uint32_t buf[ALPHA_SLICE_CAP / 4 + 4];
  const uint32_t g = blockIdx.x, L = threadIdx.x;
  if (g >= n_slices) return;
  const EncSlice S = slices[g];
  const EncPic &P = pics[S.pic];
  const EncFrame &F = frames[P.frame];
  if (F.alpha_bits == 0) {
    if (L == 0) alpha_len[g] = 0;
    return;
  }
  const uint32_t bits = F.alpha_bits, mag = bits == 8 ? 3u : 6u, maxd = bits == 8 ? 8u : 64u;
  const uint32_t lg = 4 + ilog2(S.mbs), cols = 1u << lg, count = 16 * cols;
  const uint32_t x0 = 16u * S.mb_x, width = F.width;
  const GLOBAL uint16_t *alpha = (const GLOBAL uint16_t *)(uintptr_t)F.in_alpha;
  const uint32_t ident = F.rescale_identity, max_in = F.max_in;
  const uint64_t m = F.rescale_m;
  for (uint32_t i = L; i < count; i += 64) {
    const uint32_t rr = i >> lg, n = i & (cols - 1);
    const uint32_t pr = 16u * S.mb_y + rr < P.height - 1 ? 16u * S.mb_y + rr : P.height - 1;
    const uint32_t row = P.first_row + P.row_step * pr;
    const uint32_t col = x0 + n < width - 1 ? x0 + n : width - 1;
    uint32_t a = alpha[(size_t)row * width + col];
    a = a < max_in ? a : max_in;
    vals[i] = rescale(a, ident, m);
  }
  for (uint32_t i = L; i < ALPHA_SLICE_CAP / 4 + 4; i += 64) buf[i] = 0;
  __syncthreads();
  // count <= 2048, so a run never reaches the 2048 cap of src/encode.rs:462.
  const uint32_t seg = count / 64, lo = L * seg, hi = lo + seg;
  uint32_t first = count;
  for (uint32_t i = lo; i < hi; i++)
    if (i == 0 || vals[i] != vals[i - 1]) {
      first = i;
      break;
    }
  // The first run start in the lanes after this one: an inclusive suffix
  // minimum over the lanes, shifted down by one lane.
  uint32_t incl = first;
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_down(incl, d, 64);
    if (L + d < 64) incl = o < incl ? o : incl;
  }
  uint32_t next = __shfl_down(incl, 1, 64);
  if (L == 63) next = count;
  // Bits of this lane's runs: a difference code and a run code each
  // (src/vlc.rs:254-287).
  auto diff_len = [&](uint32_t i) -> uint32_t {
    const int32_t prev = i ? (int32_t)vals[i - 1] : -1;
    const int32_t d = (int32_t)vals[i] - prev;
    const uint32_t ad = d < 0 ? (uint32_t)(-d) : (uint32_t)d;
    return (ad >= 1 && ad <= maxd) ? 2 + mag : 1 + bits;
  };
  auto run_len = [](uint32_t run) -> uint32_t { return run == 1 ? 1u : (run <= 16 ? 5u : 16u); };
  uint32_t mybits = 0;
  for (uint32_t i = first; i < hi;) {
    uint32_t j = i + 1;
    while (j < hi && vals[j] == vals[i]) j++;
    const uint32_t end = j < hi ? j : next;
    mybits += diff_len(i) + run_len(end - i);
    i = j;
  }
  const uint32_t off = excl_scan(mybits, L);
  const uint32_t total = __shfl(off + mybits, 63, 64);
  uint32_t o = off;
  auto put = [&](uint32_t v, uint32_t n) {  // n <= 17
    const uint32_t sh = o & 31, w = o >> 5;
    const uint64_t X = (uint64_t)v << (64 - n);
    atomicOr(&buf[w], (uint32_t)(X >> (32 + sh)));
    if (n + sh > 32) atomicOr(&buf[w + 1], (uint32_t)(X >> sh));
    o += n;
  };
  for (uint32_t i = first; i < hi;) {
    uint32_t j = i + 1;
    while (j < hi && vals[j] == vals[i]) j++;
    const uint32_t end = j < hi ? j : next, run = end - i;
    const int32_t prev = i ? (int32_t)vals[i - 1] : -1;
    const int32_t d = (int32_t)vals[i] - prev;
    const uint32_t ad = d < 0 ? (uint32_t)(-d) : (uint32_t)d;
    if (ad >= 1 && ad <= maxd)
      put(((ad - 1) << 1) | (d < 0 ? 1u : 0u), 2 + mag);
    else
      put((1u << bits) | ((uint32_t)d & ((1u << bits) - 1)), 1 + bits);
    if (run == 1)
      put(1, 1);
    else if (run <= 16)
      put(run - 1, 5);
    else
      put(run - 1, 16);
    i = j;
  }
  __syncthreads();
  const uint32_t bytes = (total + 7) / 8;  // zero bits to the byte boundary
  uint32_t *out = (uint32_t *)(scratch + (size_t)g * ALPHA_SLICE_CAP);
  for (uint32_t w = L; w < (bytes + 3) / 4; w += 64) out[w] = __builtin_bswap32(buf[w]);
  if (L == 0) {
    alpha_len[g] = bytes;
    atomicAdd(&state[S.pic].alpha_total, bytes);
  }
}

// ---------------------------------------------------------------------------
// One wavefront per slice, four per workgroup: the slice's bytes at each
// quantiser the picture's pass lists, kept per slice (probe_sizes, one row
// per probe result) and summed per picture.
template <int NQ>
__device__ __forceinline__ void size_pass(const EncFrame *__restrict__ frames, const EncPic *__restrict__ pics,
                                          const EncSlice *__restrict__ slices, uint32_t n_slices,
                                          const float *__restrict__ coefs, const float *__restrict__ qtabs,
                                          const uint32_t *__restrict__ alpha_len, EncState *__restrict__ state,
                                          uint32_t *__restrict__ probe_sizes, const float *__restrict__ cbound,
                                          const uint32_t *lens) {
  const uint32_t g = blockIdx.x * 4 + wave_index();
  if (g >= n_slices) return;
  const EncSlice S = slices[g];
  EncState &st = state[S.pic];
  const uint32_t nq = st.nq < (uint32_t)NQ ? st.nq : (uint32_t)NQ;
  if (nq == 0) return;
  uint32_t qis[NQ];
#pragma unroll
  for (int k = 0; k < NQ; k++) qis[k] = (uint32_t)k < nq ? st.pq[k] : 1u;
  const uint32_t round = st.round;
  const EncPic &P = pics[S.pic];
  const EncFrame &F = frames[P.frame];
  uint32_t size[NQ];
  bool fits[NQ];
  slice_bytes_q<NQ>(coefs + S.coef, cbound + S.coef / 64, S.mbs, F.chroma == 2 ? 2u : 4u,
                    qtabs + (size_t)P.qtab * 225 * 128, qis, nq, P.header_bytes, alpha_len[g], (const uint8_t *)lens,
                    size, fits);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      if ((uint32_t)k >= nq) continue;
      // Every probe's sizes are kept: the share-out starts from the uniform
      // quantiser's and the next finer one's, both of which the search probed.
      if (round + k < ENC_HIST) probe_sizes[(size_t)(round + k) * n_slices + g] = size[k] | (fits[k] ? 0u : 0x80000000u);
      atomicAdd(&st.psum[k], (unsigned long long)size[k]);
      if (!fits[k]) atomicAdd(&st.pnotfit[k], 1u);
    }
  }
}

extern "C" __global__ void __launch_bounds__(256)
    pe_size(const DevTables *__restrict__ T, const EncFrame *__restrict__ frames, const EncPic *__restrict__ pics,
            const EncSlice *__restrict__ slices, uint32_t n_slices, const float *__restrict__ coefs,
            const float *__restrict__ qtabs, const uint32_t *__restrict__ alpha_len,
            EncState *__restrict__ state, uint32_t *__restrict__ probe_sizes, const float *__restrict__ cbound) {
  __shared__ uint32_t lens[ENC_LEN_BYTES / 4];
  load_len_tables(T, lens);
  size_pass<1>(frames, pics, slices, n_slices, coefs, qtabs, alpha_len, state, probe_sizes, cbound, lens);
}

// The first pass: up to ENC_PASS_Q quantisers, each chunk loaded once.
extern "C" __global__ void __launch_bounds__(256)
    pe_size_first(const DevTables *__restrict__ T, const EncFrame *__restrict__ frames,
                  const EncPic *__restrict__ pics, const EncSlice *__restrict__ slices, uint32_t n_slices,
                  const float *__restrict__ coefs, const float *__restrict__ qtabs,
                  const uint32_t *__restrict__ alpha_len, EncState *__restrict__ state,
                  uint32_t *__restrict__ probe_sizes, const float *__restrict__ cbound) {
  __shared__ uint32_t lens[ENC_LEN_BYTES / 4];
  load_len_tables(T, lens);
  size_pass<ENC_PASS_Q>(frames, pics, slices, n_slices, coefs, qtabs, alpha_len, state, probe_sizes, cbound, lens);
}

// One lane per picture: the binary search over qi (src/encode.rs:685-713),
// as many of its steps as the pass just made settles, in the search's own
// order: a step whose probe the pass sized is taken, and the search waits
// at the first one it did not (the next pass sizes that probe alone).
extern "C" __global__ void __launch_bounds__(64)
    pe_search(const EncPic *__restrict__ pics, EncState *__restrict__ state, uint32_t n_pics,
              uint32_t window) {
  const uint32_t p = blockIdx.x * 64 + threadIdx.x;
  if (p >= n_pics) return;
  EncState &st = state[p];
  const uint32_t nq = st.nq;
  if (nq == 0) return;
  if (st.phase == 0) {
    // (budget + alpha) less the picture header and the slice table (src/encode.rs:686-689).
    const uint64_t b = pics[p].budget + st.alpha_total;
    const uint64_t over = 8 + 2ull * pics[p].n_slices;
    st.budget = b > over ? b - over : 0;
  }
  for (uint32_t k = 0; k < nq && st.round < ENC_HIST; k++) {
    st.probe_qi[st.round] = st.pq[k];
    st.probe_sum[st.round] = st.psum[k];
    st.round++;
  }
  for (;;) {
    uint32_t k = 0;
    while (k < nq && st.pq[k] != st.probe) k++;
    if (k == nq) break;
    const bool ok = st.pnotfit[k] == 0 && st.psum[k] <= st.budget;
    if (st.phase == 0) {
      st.fits_at_all = ok;
      if (ok) {
        st.lo = 1;
        st.hi = 224;
        st.phase = 1;
      } else {
        st.uniform = 224;
        st.phase = 2;
      }
    } else if (ok) {
      st.hi = st.probe;
    } else {
      st.lo = st.probe + 1;
    }
    if (st.phase == 1) {
      if (st.lo < st.hi) {
        st.probe = st.lo + (st.hi - st.lo) / 2;
      } else {
        st.uniform = st.hi;
        st.phase = 2;
      }
    }
    if (st.phase == 2) break;
  }
  for (uint32_t k = 0; k < ENC_PASS_Q; k++) {
    st.psum[k] = 0;
    st.pnotfit[k] = 0;
  }
  if (st.phase == 2) {
    st.probe = 0;
    st.nq = 0;
    st.window_k = (st.fits_at_all && st.uniform > 1) ? (st.uniform < window ? st.uniform : window) : 1;
  } else {
    st.nq = 1;
    st.pq[0] = st.probe;
  }
}

// One wavefront per slice: the size table's columns, qi = uniform - t.
// Columns 0 and 1 are the probes at uniform and uniform - 1 (whenever
// uniform > 1 the search probed both: `hi` only ever takes probed values, and
// `lo` reached uniform by lo = mid + 1). Further columns are coded only for
// the slices the reference precomputes them for (src/encode.rs:741-766): a
// slice whose previous column fits and is at most `most`, its size at
// uniform plus the slack shared from it on if no slice before it takes any.
// Past the table's K columns the same candidate chain goes on, down to qi
// 1, into the size cache `deep` (column t of slice g at g * ENC_DEEP + t,
// cleared beforehand): every size the share-out can reach is then found
// there, unless a size finer than uniform came out smaller than the one at
// uniform (slack then grows past `most`), and the share-out codes that one
// itself. A column left 0 is coded by the share-out if it turns out to need
// it.
extern "C" __global__ void __launch_bounds__(256)
    pe_window(const DevTables *__restrict__ T, const EncFrame *__restrict__ frames, const EncPic *__restrict__ pics,
              const EncSlice *__restrict__ slices, uint32_t n_slices, const float *__restrict__ coefs,
              const float *__restrict__ qtabs, const uint32_t *__restrict__ alpha_len,
              const EncState *__restrict__ state, const uint32_t *__restrict__ probe_sizes,
              const float *__restrict__ cbound, uint32_t *__restrict__ size_tab, uint32_t *__restrict__ deep) {
  __shared__ uint32_t lens[ENC_LEN_BYTES / 4];
  load_len_tables(T, lens);
  const uint32_t g = blockIdx.x * 4 + wave_index();
  if (g >= n_slices) return;
  const uint32_t L = threadIdx.x & 63;
  const EncSlice S = slices[g];
  const EncState &st = state[S.pic];
  const uint32_t U = st.uniform, K = st.window_k, R = st.round;
  uint32_t ru = 0, ru1 = 0;
  for (uint32_t r = 0; r < R; r++) {
    if (st.probe_qi[r] == U) ru = r;
    if (st.probe_qi[r] == U - 1) ru1 = r;
  }
  uint32_t col[ENC_MAXK];
#pragma unroll
  for (uint32_t t = 0; t < ENC_MAXK; t++) col[t] = 0;
  col[0] = probe_sizes[(size_t)ru * n_slices + g];
  if (K > 1) {
    col[1] = probe_sizes[(size_t)ru1 * n_slices + g];
    const EncPic &P = pics[S.pic];
    const EncFrame &F = frames[P.frame];
    const float *qt = qtabs + (size_t)P.qtab * 225 * 128;
    const uint32_t cpm = F.chroma == 2 ? 2u : 4u, alen = alpha_len[g];
    // The search found uniform fitting, so its total is within the budget.
    const uint64_t slack = st.budget - st.probe_sum[ru];
    const uint64_t most = (uint64_t)(col[0] & 0x7FFFFFFFu) + slack * S.mbs / S.mbs_from;
    // The slice is a candidate for column t while column t - 1 fits and is
    // at most `most` (src/encode.rs:751-754).
    uint32_t prev = col[1];
    bool candidate = true;
#pragma unroll
    for (uint32_t t = 2; t < ENC_MAXK; t++) {
      if (t >= K) break;
      if ((prev >> 31) || (uint64_t)(prev & 0x7FFFFFFFu) > most) {
        candidate = false;
        break;
      }
      uint32_t sz[1];
      bool ft[1];
      slice_bytes(coefs + S.coef, cbound + S.coef / 64, S.mbs, cpm, qt, U - t, P.header_bytes, alen,
                  (const uint8_t *)lens, sz[0], ft[0]);
      col[t] = sz[0] | (ft[0] ? 0u : 0x80000000u);
      prev = col[t];
    }
    for (uint32_t t = K; candidate && t < U; t++) {
      if ((prev >> 31) || (uint64_t)(prev & 0x7FFFFFFFu) > most) break;
      uint32_t sz;
      bool ft;
      slice_bytes(coefs + S.coef, cbound + S.coef / 64, S.mbs, cpm, qt, U - t, P.header_bytes, alen,
                  (const uint8_t *)lens, sz, ft);
      prev = sz | (ft ? 0u : 0x80000000u);
      if (L == 0) deep[(size_t)g * ENC_DEEP + t] = prev;
    }
  }
  if (L < ENC_MAXK) {
    uint32_t v = 0;
#pragma unroll
    for (uint32_t t = 0; t < ENC_MAXK; t++)
      if (t == L) v = col[t];
    size_tab[(size_t)g * ENC_MAXK + L] = v;
  }
}

// ---------------------------------------------------------------------------
// The share-out (src/encode.rs:769-793). With S the budget left after every
// slice's size at the uniform qi, slice i sees slack_i = S - sum_{j<i} extra_j
// (extra_j = used_j - base_j), a triangular recurrence: 64 slices at a time,
// every lane decides from a guess of its slack, the guesses are replaced by
// the prefix sums of the decisions, and the loop stops at the fixed point,
// which is the sequential answer (lane j is exact after j rounds).
extern "C" __global__ void __launch_bounds__(64 * SHARE_WAVES)
    pe_shareout(const DevTables *__restrict__ T, const EncFrame *__restrict__ frames, const EncPic *__restrict__ pics,
                const EncSlice *__restrict__ slices, const float *__restrict__ coefs,
                const float *__restrict__ qtabs, const uint32_t *__restrict__ alpha_len,
                EncState *__restrict__ state, const uint32_t *__restrict__ size_tab,
                uint32_t *__restrict__ deep, uint32_t *__restrict__ used, uint32_t *__restrict__ qidx,
                const float *__restrict__ cbound) {
  __shared__ uint32_t req_g[64], req_t[64];
  __shared__ uint32_t n_req, done;
  __shared__ uint32_t lens[ENC_LEN_BYTES / 4];
  load_len_tables(T, lens);
  const uint32_t p = blockIdx.x, tid = threadIdx.x, L = tid & 63, wave = wave_index();
  const EncPic &P = pics[p];
  EncState &st = state[p];
  const uint32_t n = P.n_slices, g0 = P.slice_base, U = st.uniform, K = st.window_k;
  if (!(U > 1 && st.fits_at_all)) {
    for (uint32_t i = tid; i < n; i += 64 * SHARE_WAVES) {
      used[g0 + i] = size_tab[(size_t)(g0 + i) * ENC_MAXK] & 0x7FFFFFFFu;
      qidx[g0 + i] = U;
    }
    return;
  }
  const EncFrame &F = frames[P.frame];
  const uint32_t cpm = F.chroma == 2 ? 2u : 4u;
  const float *qt = qtabs + (size_t)P.qtab * 225 * 128;
  // Wave 0 runs the share-out; its state lives across the service rounds.
  int64_t slack = 0, s_i = 0, extra = 0;
  uint64_t total_mbs = 0, mbs_before = 0, M = 0, allowed = 0;
  uint32_t c0 = 0, round = 0, phase = 0;  // 0 new chunk, 1 new round, 2 walking
  uint32_t g = 0, base = 0, m = 0, qi = 0, u = 0, served = 0;
  uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0, w6 = 0, w7 = 0;  // the slice's window columns
  uint4 na = make_uint4(0, 0, 0, 0), nb = na;  // the next chunk's, loaded a chunk ahead
  uint32_t nm_next = 0;
  bool valid = false, walking = false;
  auto prefetch = [&](uint32_t c) {
    const uint32_t i = c + L;
    const uint32_t gg = g0 + (i < n ? i : 0);
    na = ((const uint4 *)(size_tab + (size_t)gg * ENC_MAXK))[0];
    nb = ((const uint4 *)(size_tab + (size_t)gg * ENC_MAXK))[1];
    nm_next = i < n ? slices[gg].mbs : 0u;
  };
  if (wave == 0) {
    uint64_t sum_base = 0;
    for (uint32_t i = L; i < n; i += 64) {
      sum_base += size_tab[(size_t)(g0 + i) * ENC_MAXK] & 0x7FFFFFFFu;
      total_mbs += slices[g0 + i].mbs;
    }
    sum_base = wave_sum64(sum_base);
    total_mbs = wave_sum64(total_mbs);
    slack = (int64_t)(st.budget - sum_base);
    prefetch(0);
  }
  for (;;) {
    if (wave == 0) {
      if (L == 0) {
        n_req = 0;
        done = 0;
      }
      for (;;) {
        if (phase == 0) {  // the next 64 slices
          if (c0 >= n) {
            if (L == 0) done = 1;
            break;
          }
          const uint32_t i = c0 + L;
          valid = i < n;
          g = g0 + (valid ? i : 0);
          w0 = na.x; w1 = na.y; w2 = na.z; w3 = na.w; w4 = nb.x; w5 = nb.y; w6 = nb.z; w7 = nb.w;
          m = nm_next;
          if (c0 + 64 < n) prefetch(c0 + 64);
          base = valid ? (w0 & 0x7FFFFFFFu) : 0u;
          M = total_mbs - mbs_before - (wave_incl_scan(m) - m);
          s_i = slack;
          round = 0;
          phase = 1;
        }
        if (phase == 1) {
          // Slice i's own steps for a given slack (src/encode.rs:773-789): it
          // steps to finer quantisers while the next size fits its share.
          const uint64_t sc = s_i < 0 ? 0ull : (uint64_t)s_i;
          allowed = valid ? base + sc * m / M : 0ull;
          qi = U;
          u = base;
          walking = valid;
          phase = 2;
        }
        bool need = false;
        while (__ballot(walking)) {
          if (walking) {
            const uint32_t t = U - qi + 1;  // the column of qi - 1
            uint32_t v;
            v = 0;
            if (t < K) v = t == 1 ? w1 : t == 2 ? w2 : t == 3 ? w3 : t == 4 ? w4 : t == 5 ? w5 : t == 6 ? w6 : w7;
            // A column the window left out (0), or one past it: the cache.
            if (v == 0) v = deep[(size_t)g * ENC_DEEP + t];
            if (v == 0) {
              need = true;
            } else if ((v >> 31) || (v & 0x7FFFFFFFu) > allowed) {
              walking = false;
            } else {
              qi--;
              u = v & 0x7FFFFFFFu;
              if (qi <= 1) walking = false;
            }
          }
          if (__ballot(need)) break;
        }
        const uint64_t nm = __ballot(need);
        if (nm) {
          // Sizes past the cache: the whole workgroup codes them.
          if (need) {
            const uint32_t slot = (uint32_t)__popcll(nm & ((1ull << L) - 1));
            req_g[slot] = g;
            req_t[slot] = U - qi + 1;
          }
          if (L == 0) n_req = (uint32_t)__popcll(nm);
          break;
        }
        // Round done: replace the guessed slacks by the prefix of the
        // decisions they imply; stop at the fixed point.
        extra = valid ? (int64_t)u - (int64_t)base : 0;
        const int64_t s_new = slack - excl_scan64(extra, L);
        if (__ballot(valid && s_new != s_i) && round < 64) {
          s_i = s_new;
          round++;
          phase = 1;
          continue;
        }
        if (valid) {
          used[g] = u;
          qidx[g] = qi;
        }
        int64_t ext = extra;
        for (int o = 32; o > 0; o >>= 1) {
          const uint32_t lo = __shfl_xor((uint32_t)(uint64_t)ext, o, 64);
          const uint32_t hi = __shfl_xor((uint32_t)((uint64_t)ext >> 32), o, 64);
          ext += (int64_t)(((uint64_t)hi << 32) | lo);
        }
        slack -= ext;
        mbs_before += wave_sum64(m);
        c0 += 64;
        phase = 0;
      }
    }
    __syncthreads();
    if (done) break;
    // Each request: quantisers U - t ... U - t - (DEEP_STEP - 1) (down to 1)
    // of one slice, one quantiser per wave.
    const uint32_t nr = n_req;
    for (uint32_t it = wave; it < nr * DEEP_STEP; it += SHARE_WAVES) {
      const uint32_t r = it / DEEP_STEP, k = it - r * DEEP_STEP;
      const uint32_t gr = req_g[r], t = req_t[r];
      if (k >= U - t) continue;  // quantisers stop at 1
      const uint32_t q = U - t - k;
      uint32_t sz[1];
      bool ft[1];
      slice_bytes(coefs + slices[gr].coef, cbound + slices[gr].coef / 64, slices[gr].mbs, cpm, qt, q,
                  P.header_bytes, alpha_len[gr], (const uint8_t *)lens, sz[0], ft[0]);
      if (L == 0) deep[(size_t)gr * ENC_DEEP + t + k] = sz[0] | (ft[0] ? 0u : 0x80000000u);
    }
    served += nr;
    __syncthreads();
  }
  if (tid == 0) st.fallback = served;
}

// ---------------------------------------------------------------------------
// One workgroup per frame: the picture header, the slice table and the
// slices, back to back, after the frame header (src/encode.rs:820-829,
// 284-286). Offsets are relative to the frame's first byte.
extern "C" __global__ void __launch_bounds__(256)
    pe_offsets(const EncFrame *__restrict__ frames, EncPic *__restrict__ pics,
               const uint32_t *__restrict__ used, uint64_t *__restrict__ slice_out,
               uint64_t *__restrict__ frame_size) {
  const uint32_t f = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const EncFrame &F = frames[f];
  __shared__ uint64_t wsum[4];
  __shared__ uint64_t carry;
  uint64_t pos = F.hdr_len;
  for (uint32_t p = F.first_pic; p < F.first_pic + F.n_pics; p++) {
    const uint32_t n = pics[p].n_slices, g0 = pics[p].slice_base;
    const uint64_t data0 = pos + 8 + 2ull * n;
    if (tid == 0) {
      pics[p].out_off = pos;
      carry = 0;
    }
    __syncthreads();
    for (uint32_t b = 0; b < n; b += 256) {
      const uint32_t k = b + tid;
      const uint32_t v = k < n ? used[g0 + k] : 0u;
      uint32_t incl = v;
      for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
      }
      if (lane == 63) wsum[wave] = incl;
      __syncthreads();
      uint64_t before = 0;
      for (uint32_t w = 0; w < wave; w++) before += wsum[w];
      if (k < n) slice_out[g0 + k] = data0 + carry + before + incl - v;
      __syncthreads();
      if (tid == 255) carry += before + incl;
      __syncthreads();
    }
    pos = data0 + carry;
    __syncthreads();
  }
  if (tid == 0) frame_size[f] = pos;
}

// ---------------------------------------------------------------------------
// One workgroup: each frame's place in the output arena, its packet at a
// 256-byte boundary after the previous frame's, and whether the packets fit
// the arena's capacity (if not, pe_pack writes nothing and the host runs it
// again over a larger arena).
extern "C" __global__ void __launch_bounds__(256)
    pe_arena(const uint64_t *__restrict__ frame_size, uint32_t n_frames, uint64_t base, uint64_t capacity,
             uint64_t *__restrict__ out_addr, uint32_t *__restrict__ overflow) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ uint64_t wsum[4];
  uint64_t carry = 0;
  for (uint32_t b = 0; b < n_frames; b += 256) {
    const uint32_t k = b + tid;
    const uint64_t v = k < n_frames ? (frame_size[k] + 255) & ~255ull : 0ull;
    uint64_t incl = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint64_t o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint64_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += wsum[w];
    if (k < n_frames) out_addr[k] = base + carry + before + incl - v;
    carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
  if (tid == 0) *overflow = carry > capacity ? 1u : 0u;
}

namespace {
constexpr uint32_t RING = 512;  // words of the per-slice bit ring

// ORs the `n` (1..64) low bits of `v` into the ring at bit `o` (MSB first).
__device__ __forceinline__ void ring_put(uint32_t *ring, uint32_t o, uint64_t v, uint32_t n) {
  const uint32_t sh = o & 31, w = o >> 5;
  const uint64_t X = v << (64 - n);
  atomicOr(&ring[w & (RING - 1)], (uint32_t)(X >> (32 + sh)));
  if (n + sh > 32) atomicOr(&ring[(w + 1) & (RING - 1)], (uint32_t)(X >> sh));
  if (n + sh > 64) atomicOr(&ring[(w + 2) & (RING - 1)], (uint32_t)(X << (32 - sh)));
}

// Stores bytes [lo, hi) of big-endian ring word `w` at `A + 4 w`.
__device__ __forceinline__ void store_word(GLOBAL uint8_t *A, uint32_t w, uint32_t val, uint32_t lo,
                                           uint32_t hi) {
  if (lo == 0 && hi == 4) {
    *(GLOBAL uint32_t *)(A + 4 * (size_t)w) = __builtin_bswap32(val);
  } else {
    for (uint32_t b = lo; b < hi; b++) A[4 * (size_t)w + b] = (uint8_t)(val >> (24 - 8 * b));
  }
}
}  // namespace

// Byte b of the frame: frame_size and 'icpf', then FrameHeader::write
// (src/header.rs:174-199): header size, version, encoder identifier, width,
// height, chroma format and interlace mode, aspect ratio and frame rate,
// colour primaries, transfer characteristic, matrix coefficients, alpha
// channel type, the matrix flags, then the loaded luma and chroma matrices.
__device__ __forceinline__ uint8_t frame_header_byte(const EncFrame &F, uint32_t frame_size, uint32_t b) {
  if (b < 4) return (uint8_t)(frame_size >> (24 - 8 * b));
  if (b >= 28) {
    const uint32_t m = b - 28;
    return (F.matrix_flags & 2) ? (m < 64 ? F.luma_matrix[m] : F.chroma_matrix[m - 64]) : F.chroma_matrix[m];
  }
  const uint32_t header_size = F.hdr_len - 8;
  switch (b) {
    case 4: return 'i';
    case 5: return 'c';
    case 6: return 'p';
    case 7: return 'f';
    case 8: return (uint8_t)(header_size >> 8);
    case 9: return (uint8_t)header_size;
    case 11: return F.version;
    case 12: case 13: case 14: case 15: return F.encoder_id[b - 12];
    case 16: return (uint8_t)(F.width >> 8);
    case 17: return (uint8_t)F.width;
    case 18: return (uint8_t)(F.height >> 8);
    case 19: return (uint8_t)F.height;
    case 20: return F.chroma_interlace;
    case 21: return F.aspect_rate;
    case 22: return F.primaries;
    case 23: return F.transfer;
    case 24: return F.matrix;
    case 25: return F.alpha_type;
    case 27: return F.matrix_flags;
    default: return 0;  // bytes 10 and 26 are reserved
  }
}

// One wavefront per slice: slice header, Y', Cb, Cr codewords (src/encode.rs:
// 413-437, 796-819) and the alpha bytes; the slice-table entry; the picture
// header for the first slice of a picture and the frame header for the first
// slice of a frame.
extern "C" __global__ void __launch_bounds__(64)
    pe_pack(const EncFrame *__restrict__ frames, const EncPic *__restrict__ pics,
            const EncSlice *__restrict__ slices, const float *__restrict__ coefs,
            const float *__restrict__ qtabs, const uint32_t *__restrict__ alpha_len,
            const uint8_t *__restrict__ alpha_scratch, const uint32_t *__restrict__ used,
            const uint32_t *__restrict__ qidx, const uint64_t *__restrict__ slice_out,
            const uint64_t *__restrict__ frame_size, const uint64_t *__restrict__ out_addr,
            const uint32_t *__restrict__ overflow, const float *__restrict__ cbound) {
  __shared__ uint32_t ring[RING];
  if (*overflow) return;  // the arena is too small: the host grows it and packs again
  const uint32_t g = blockIdx.x, L = threadIdx.x;
  const EncSlice S = slices[g];
  const EncPic &P = pics[S.pic];
  const EncFrame &F = frames[P.frame];
  GLOBAL uint8_t *out = (GLOBAL uint8_t *)(uintptr_t)out_addr[P.frame];
  const uint32_t k = g - P.slice_base;
  const uint32_t u = used[g], qi = qidx[g], hb = P.header_bytes;
  if (L == 0) {
    GLOBAL uint8_t *te = out + P.out_off + 8 + 2ull * k;
    te[0] = (uint8_t)(u >> 8);
    te[1] = (uint8_t)u;
  }
  if (k == 0 && L == 0) {
    // PictureHeader::write (src/header.rs:292-304) with picture_size patched.
    const uint32_t n = P.n_slices;
    const uint64_t end = (S.pic + 1 < F.first_pic + F.n_pics) ? pics[S.pic + 1].out_off : frame_size[P.frame];
    const uint32_t psz = (uint32_t)(end - P.out_off);
    GLOBAL uint8_t *ph = out + P.out_off;
    ph[0] = 8 << 3;
    ph[1] = (uint8_t)(psz >> 24);
    ph[2] = (uint8_t)(psz >> 16);
    ph[3] = (uint8_t)(psz >> 8);
    ph[4] = (uint8_t)psz;
    const uint32_t dep = n <= 65535 ? n : 0;
    ph[5] = (uint8_t)(dep >> 8);
    ph[6] = (uint8_t)dep;
    ph[7] = (uint8_t)((P.log2_slice_mbs & 3) << 4);
  }
  if (g == F.first_slice) {
    // The frame header (src/header.rs:174-199) with frame_size patched, and
    // the header of any picture without slices (a one-row frame's empty field).
    const uint32_t fs = (uint32_t)frame_size[P.frame];
    for (uint32_t b = L; b < F.hdr_len; b += 64) out[b] = frame_header_byte(F, fs, b);
    if (L == 0)
      for (uint32_t p = F.first_pic; p < F.first_pic + F.n_pics; p++) {
        if (pics[p].n_slices) continue;
        GLOBAL uint8_t *ph = out + pics[p].out_off;
        ph[0] = 8 << 3;
        ph[1] = 0; ph[2] = 0; ph[3] = 0; ph[4] = 8;
        ph[5] = 0; ph[6] = 0;
        ph[7] = (uint8_t)((pics[p].log2_slice_mbs & 3) << 4);
      }
  }
  for (uint32_t i = L; i < RING; i += 64) ring[i] = 0;
  __syncthreads();

  GLOBAL uint8_t *const sa = out + slice_out[g] + hb;  // first byte of the colour data
  const uint32_t lead = (uint32_t)((uintptr_t)sa & 3);
  GLOBAL uint8_t *const A = sa - lead;
  uint32_t pos = 8 * lead, flushed = 0;
  const uint32_t cpm = F.chroma == 2 ? 2u : 4u;
  const float *qt = qtabs + (size_t)P.qtab * 225 * 128;
  const float *cf0 = coefs + S.coef;
  // Chunks that may hold a nonzero at qi (their bounds, as in slice_bytes_q):
  // the others add no codeword and leave the contexts as they are.
  const uint32_t TB = (4 + 2 * cpm) * S.mbs;
  const float *cb = cbound + S.coef / 64;
  const float qlim = (float)qscale(qi) * ZERO_CHUNK_LIMIT;
  const uint64_t live0 = __ballot(L < TB && cb[L < TB ? L : 0] > qlim);
  const uint64_t live1 = __ballot(L + 64 < TB && cb[L + 64 < TB ? L + 64 : 0] > qlim);
  // Stores the ring's complete words [flushed, upto) and clears them.
  auto flush = [&](uint32_t upto) {
    for (uint32_t w = flushed + L; w < upto; w += 64) {
      store_word(A, w, ring[w & (RING - 1)], w == 0 ? lead : 0u, 4);
      ring[w & (RING - 1)] = 0;
    }
    flushed = upto;
  };
  uint32_t comp_bytes[3];
  uint32_t k0 = 0;  // the component's first chunk in the slice
  for (uint32_t c = 0; c < 3; c++) {
    uint32_t nb, off;
    component(c, S.mbs, cpm, nb, off);
    const uint32_t lg = ilog2(nb);
    const float *cf = cf0 + off + L;
    const float invL = qt[(qi * 2 + (c > 0 ? 1 : 0)) * 64 + L];
    uint64_t todo = ((((k0 >> 6) ? live1 : live0) >> (k0 & 63)) & ((1ull << nb) - 1)) | 1;  // and the DC chunk
    const uint32_t start = pos;
    // The last nonzero's index and its packed run/level contexts
    // (min(run, 15) | min(level, 8) << 4); none yet: run context 4, level 1.
    uint32_t last_idx = nb - 1, ctx = 4u | 1u << 4;
    uint32_t j = 0;
    float f = cf[0];
    for (;;) {
      todo &= todo - 1;  // chunk j is the lowest bit
      const uint32_t jn = todo ? (uint32_t)__builtin_ctzll(todo) : j;
      const float fn = cf[jn * 64];  // the next coded chunk, loaded while this one is coded
      const uint32_t base = j * 64, i = base + L;
      const uint32_t s = i >> lg;
      const int32_t q = quant(f, __shfl(invL, (int)s, 64), s == 0 ? 0.5f : 0.42f);
      uint64_t v1 = 0, v2 = 0;
      uint32_t n1 = 0, n2 = 0;
      if (j == 0) {
        const int32_t d1 = __shfl(q, L ? L - 1 : 0, 64);
        const int32_t d2 = __shfl(q, L > 1 ? L - 2 : 0, 64);
        if (L < nb) {
          if (L == 0) {
            v1 = book_code32(FIRST_DC, signed_to_symbol(q), n1);
          } else {
            const int32_t diff = (int32_t)((uint32_t)q - (uint32_t)d1);
            const int32_t pd = L == 1 ? 3 : (int32_t)((uint32_t)d1 - (uint32_t)d2);
            const int32_t nn = pd < 0 ? (int32_t)(0u - (uint32_t)diff) : diff;
            const uint32_t a = pd < 0 ? 0u - (uint32_t)pd : (uint32_t)pd;
            v1 = book_code32(dc_book(a < 3 ? a : 3), signed_to_symbol(nn), n1);
          }
        }
      }
      const bool nz = i >= nb && q != 0;
      const uint64_t mask = __ballot(nz);
      if (mask) {
        const uint64_t lower = mask & ((1ull << L) - 1);
        const int Pn = lower ? 63 - (int)__builtin_clzll(lower) : -1;
        const uint32_t prev_idx = Pn >= 0 ? base + (uint32_t)Pn : last_idx;
        const uint32_t run = i - prev_idx - 1;
        const uint32_t lvl = (q < 0 ? 0u - (uint32_t)q : (uint32_t)q) - 1;
        const uint32_t mine = (run < 15 ? run : 15u) | (lvl < 8 ? lvl : 8u) << 4;
        uint32_t pc = __shfl(mine, Pn >= 0 ? Pn : (int)L, 64);
        if (Pn < 0) pc = ctx;
        if (nz) {
          // Run, then level and sign (src/encode.rs:419-432): one codeword
          // of both when it fits 64 bits.
          v1 = book_code32(run_book(pc & 15), run, n1);
          uint32_t ll;
          const uint64_t lc = book_code32(lvl_book(pc >> 4), lvl, ll);
          v2 = (lc << 1) | (q < 0 ? 1u : 0u);
          n2 = ll + 1;
          if (n1 + n2 <= 64) {
            v1 = (v1 << n2) | v2;
            n1 += n2;
            n2 = 0;
          }
        }
        const uint32_t last = 63 - (uint32_t)__builtin_clzll(mask);
        last_idx = base + last;
        ctx = __builtin_amdgcn_readlane(mine, last);
      }
      const uint32_t T = n1 + n2;
      const uint32_t incl = wave_incl_scan(T);
      const uint32_t o = pos + incl - T;
      const uint32_t chunk = __builtin_amdgcn_readlane(incl, 63);
      if (n1) ring_put(ring, o, v1, n1);
      if (n2) ring_put(ring, o + n1, v2, n2);
      pos += chunk;
      // A chunk adds at most 64 x 128 bits (256 words): with fewer than 256
      // complete words held, the ring's 512 always have room for the next.
      if ((pos >> 5) - flushed >= 256) flush(pos >> 5);
      if (!todo) break;
      j = jn;
      f = fn;
    }
    pos = (pos + 7) & ~7u;  // zero bits to the byte boundary (src/encode.rs:436)
    comp_bytes[c] = (pos - start) / 8;
    k0 += nb;
  }
  // The complete words held, then the last, partial word.
  flush(pos >> 5);
  if ((pos & 31) && L == 0) {
    const uint32_t w = pos >> 5;
    store_word(A, w, ring[w & (RING - 1)], w == 0 ? lead : 0u, (pos & 31) / 8);
  }
  // Slice header (src/encode.rs:800-817).
  uint8_t *sh = out + slice_out[g];
  if (L < hb) {
    uint8_t v = 0;
    if (L == 0) v = (uint8_t)(hb << 3);
    else if (L == 1) v = (uint8_t)qi;
    else if (L < 8) {
      const uint32_t cs = comp_bytes[(L - 2) / 2];
      v = (L & 1) ? (uint8_t)cs : (uint8_t)(cs >> 8);
    }
    sh[L] = v;
  }
  // Alpha bytes after the colour data.
  const uint32_t alen = alpha_len[g];
  uint8_t *ad = sa + comp_bytes[0] + comp_bytes[1] + comp_bytes[2];
  const uint8_t *as = alpha_scratch + (size_t)g * ALPHA_SLICE_CAP;
  for (uint32_t b = L; b < alen; b += 64) ad[b] = as[b];
}
