/*
 * TEST INFRASTRUCTURE ONLY — the CPU oracle.
 *
 * A C++ restatement of the rivet-prores crate (Rust; it cannot be built in
 * this environment) used as the checker for the HIP slice engine. Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg load it; the
 * product library never links it.
 *
 * Every entry point cites the reference file:line it follows. The C API is
 * flat (plain pointers and sizes) so that Python reaches it through ctypes.
 */
#ifndef PRORES_ORACLE_H
#define PRORES_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error kinds, src/error.rs:5-21. 0 = Ok. */
enum { ORACLE_OK = 0, ORACLE_INVALID = 1, ORACLE_UNSUPPORTED = 2, ORACLE_CONFIG = 3 };

/* A frame (src/frame.rs:180-203). data = Y' || Cb || Cr, tightly packed. */
typedef struct {
  uint32_t width, height, bit_depth;
  uint32_t chroma;    /* RDD 36 code: 2 = 4:2:2, 3 = 4:4:4 */
  uint32_t interlace; /* 0 progressive, 1 top field first, 2 bottom field first */
  uint8_t aspect_ratio, frame_rate_code, color_primaries, transfer_characteristic,
      matrix_coefficients;
  uint8_t has_alpha;
  uint16_t *data;
  size_t data_len;
  uint16_t *alpha; /* width*height samples, or NULL */
  size_t alpha_len;
  uint64_t plane_offset[3];
  uint32_t plane_width[3], plane_height[3];
} oracle_frame;

/* Config (src/encode.rs:135-174). */
typedef struct {
  uint32_t profile;        /* 0 Proxy, 1 LT, 2 422, 3 HQ, 4 4444, 5 4444 XQ */
  uint32_t log2_slice_mbs; /* 0..3 */
  uint32_t alpha;          /* 0 none, 1 8-bit, 2 16-bit */
  uint8_t has_luma_matrix, has_chroma_matrix;
  uint8_t luma_matrix[64], chroma_matrix[64];
  uint64_t target_frame_bytes; /* 0: the profile's */
  uint8_t encoder_identifier[4];
  uint32_t threads;
} oracle_config;

/* Decoder::decode, src/decode.rs:61-104. bit_depth 0 = natural. The frame's
 * buffers are malloc'd; release with oracle_frame_free. */
int oracle_decode(const uint8_t *data, size_t len, uint32_t bit_depth, uint32_t threads,
                  oracle_frame *out, char *msg, size_t msg_cap);
void oracle_frame_free(oracle_frame *f);

/* Encoder::encode, src/encode.rs:223-287. *out is malloc'd; oracle_free. */
int oracle_encode(const oracle_config *cfg, const oracle_frame *frame, uint8_t **out,
                  size_t *out_len, char *msg, size_t msg_cap);
void oracle_free(void *p);

/* Config::new defaults, src/encode.rs:163-174. */
void oracle_config_default(uint32_t profile, oracle_config *cfg);
/* Profile::target_frame_bytes, src/encode.rs:112-117. */
uint64_t oracle_target_frame_bytes(uint32_t profile, uint32_t width, uint32_t height);

/* ---- kernels and pieces, for known-answer tests ---- */
/* dsp::idct_put scalar, src/dsp/scalar.rs:8-12 with Output::new, src/dsp/mod.rs:88-96 */
void oracle_idct_put(const int32_t *coef, const float *scale, uint32_t bit_depth, uint16_t *out);
/* dct::idct, src/dct.rs:36-66 */
void oracle_idct(const float *coef, float *out);
/* dct::fdct, src/dct.rs:70-93 */
void oracle_fdct(const float *pix, float *out);
/* dsp::fdct_load scalar, src/dsp/scalar.rs:14-17 */
void oracle_fdct_load(const uint16_t *pix, float scale, float *out);
/* dsp::quantise scalar, src/dsp/scalar.rs:19-39 */
void oracle_quantise(const float *f, size_t n, const float *inv, float ac_rounding, int32_t *out,
                     uint64_t *mask);
/* dct::BASIS, src/dct.rs:11-25 */
void oracle_basis(float *out64);
/* header::slice_sizes, src/header.rs:308-320; returns count, writes up to cap */
size_t oracle_slice_sizes(uint32_t width_in_mb, uint32_t log2, uint32_t *out, size_t cap);
/* tables::qscale, src/tables.rs:33-36 */
uint32_t oracle_qscale(uint32_t qi);
/* the scans, src/tables.rs:6-26: which 0 progressive, 1 interlaced */
void oracle_scan(uint32_t which, uint8_t *out64);
/* A codebook's codeword for symbol n as '0'/'1' characters (src/vlc.rs:106-126).
 * kind 0 = ExpGolomb(a), 1 = Combo(a, b, c). Returns the length. */
uint32_t oracle_codeword(uint32_t kind, uint32_t a, uint32_t b, uint32_t c, uint32_t n,
                         char *bits, size_t cap);
/* Codebook::len (src/vlc.rs:155-172) */
uint32_t oracle_code_len(uint32_t kind, uint32_t a, uint32_t b, uint32_t c, uint32_t n);
/* Codebook::read from a byte string (src/vlc.rs:207-230); returns 0 and the
 * symbol, or the error kind with msg. */
int oracle_code_read(uint32_t kind, uint32_t a, uint32_t b, uint32_t c, const uint8_t *data,
                     size_t len, uint32_t *sym, char *msg, size_t msg_cap);
/* signed mapping, src/vlc.rs:233-251 */
uint32_t oracle_signed_to_symbol(int32_t n);
int32_t oracle_symbol_to_signed(uint32_t s);
/* alpha codes, src/vlc.rs:254-299, as '0'/'1' strings */
uint32_t oracle_alpha_run_bits(uint32_t run, char *bits, size_t cap);
uint32_t oracle_alpha_diff_bits(int32_t diff, uint32_t nbits, char *bits, size_t cap);
/* Rescale, src/decode.rs:446-469 */
uint16_t oracle_rescale(uint32_t bits_in, uint32_t bits_out, uint16_t a);
void oracle_rescale_many(uint32_t bits_in, uint32_t bits_out, const uint16_t *a, size_t n, uint16_t *out);
/* encode_coefficients over a scanned array (src/encode.rs:441-454); returns bytes */
size_t oracle_encode_coefficients(const int32_t *scanned, size_t n_blocks, uint8_t *out,
                                  size_t cap);
/* coefficient_bits (src/encode.rs:398-409) */
size_t oracle_coefficient_bits(const int32_t *scanned, size_t n_blocks);
/* decode_coefficients (src/decode.rs:545-547); scanned order */
int oracle_decode_coefficients(const uint8_t *data, size_t len, size_t n_blocks, int32_t *out,
                               char *msg, size_t msg_cap);
/* encode_alpha / decode_alpha (src/encode.rs:457-470, src/decode.rs:570-595) */
size_t oracle_encode_alpha(const uint16_t *values, size_t n, uint32_t bits, uint8_t *out,
                           size_t cap);
int oracle_decode_alpha(const uint8_t *data, size_t len, uint32_t bits, size_t count,
                        size_t needed, uint16_t *out, size_t *out_n, char *msg, size_t msg_cap);

/* ---- test pictures ---- */
/* tests/common/mod.rs:26-80 test_frame: fills data (Y'||Cb||Cr). */
void oracle_test_frame(uint32_t width, uint32_t height, uint32_t chroma, uint32_t bit_depth,
                       int32_t grain, uint32_t seed, uint16_t *data);

/* Threads the pool would use for `threads` (src/pool.rs:20-30 resolve). */
uint32_t oracle_resolve_threads(uint32_t threads);

#ifdef __cplusplus
}
#endif
#endif
