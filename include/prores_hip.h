/*
 * prores_hip.h — C ABI of the MI355X ProRes slice engine.
 *
 * The drop-in boundary for the rivet-prores crate's API (Rust). Each entry
 * point names the reference item it replaces (file:line in the reference
 * tree). Plain pointers and sizes only; no HIP or torch types cross it
 * except the opaque device pointers of the device-resident entry points,
 * which are plain `void *` / `uint8_t *` values.
 *
 * Errors: every call that can fail returns a prores_status and fills a
 * prores_error whose kind and message text are the reference's
 * (src/error.rs:5-21; the message strings of src/header.rs, src/decode.rs,
 * src/bits.rs, src/vlc.rs, src/frame.rs, src/encode.rs). PRORES_ERR_DEVICE
 * is this library's own: a HIP runtime failure (no device, out of memory).
 */
#ifndef PRORES_HIP_H
#define PRORES_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error kinds: Invalid, Unsupported, Config (src/error.rs:5-21). */
typedef enum {
  PRORES_OK = 0,
  PRORES_ERR_INVALID = 1,
  PRORES_ERR_UNSUPPORTED = 2,
  PRORES_ERR_CONFIG = 3,
  PRORES_ERR_DEVICE = 4
} prores_status;

typedef struct {
  int32_t kind;     /* a prores_status */
  char msg[256];    /* the reference's message (UTF-8), without the "invalid ProRes data: " prefix */
} prores_error;

/* ChromaFormat codes (src/frame.rs:7-41): the header's chroma_format. */
enum { PRORES_CHROMA_422 = 2, PRORES_CHROMA_444 = 3 };
/* Interlace (src/frame.rs:46-77). */
enum { PRORES_PROGRESSIVE = 0, PRORES_TOP_FIELD_FIRST = 1, PRORES_BOTTOM_FIELD_FIRST = 2 };
/* AlphaType (src/header.rs:11-40). */
enum { PRORES_ALPHA_NONE = 0, PRORES_ALPHA_8 = 1, PRORES_ALPHA_16 = 2 };
/* Profile (src/encode.rs:30-44), in Profile::ALL order. */
enum {
  PRORES_PROFILE_PROXY = 0,
  PRORES_PROFILE_LT = 1,
  PRORES_PROFILE_STANDARD = 2,
  PRORES_PROFILE_HQ = 3,
  PRORES_PROFILE_4444 = 4,
  PRORES_PROFILE_4444XQ = 5
};

/* Metadata (src/frame.rs:83-110). */
typedef struct {
  uint8_t aspect_ratio, frame_rate_code, color_primaries, transfer_characteristic,
      matrix_coefficients;
} prores_metadata;

/* Plane (src/frame.rs:150-158): offset in samples into `data`. */
typedef struct {
  uint64_t offset;
  uint32_t width, height;
} prores_plane;

/* Frame (src/frame.rs:180-203). data = Y' || Cb || Cr, u16 native-endian,
 * each plane tightly packed; alpha = width*height samples or NULL. When the
 * library fills a frame (prores_decode) it owns the buffers until
 * prores_frame_free. */
typedef struct {
  uint32_t width, height, bit_depth;
  uint32_t chroma, interlace;
  prores_metadata metadata;
  uint16_t *data;
  size_t data_len;
  prores_plane planes[3];
  uint16_t *alpha;
  size_t alpha_len;
} prores_frame;

/* FrameHeader (src/header.rs:44-72). */
typedef struct {
  uint32_t frame_size;
  uint16_t header_size;
  uint8_t bitstream_version;
  uint8_t encoder_identifier[4];
  uint16_t width, height;
  uint32_t chroma, interlace;
  prores_metadata metadata;
  uint32_t alpha; /* PRORES_ALPHA_* */
  uint8_t has_luma_matrix, has_chroma_matrix;
  uint8_t luma_matrix[64], chroma_matrix[64];
} prores_frame_header;

/* Config (src/encode.rs:135-159). target_frame_bytes 0 = the profile's. */
typedef struct {
  uint32_t profile;
  uint32_t log2_slice_mbs;
  uint32_t alpha;
  uint8_t has_luma_matrix, has_chroma_matrix;
  uint8_t luma_matrix[64], chroma_matrix[64];
  uint64_t target_frame_bytes;
  uint8_t encoder_identifier[4];
  uint32_t threads;
} prores_config;

typedef struct prores_decoder prores_decoder;

/* ---- Decoder (src/decode.rs:19-104) ---- */

/* Decoder::new / Decoder::with_bit_depth / with_threads
 * (src/decode.rs:28-57). bit_depth 0 = natural (10 for 4:2:2, 12 for
 * 4:4:4); otherwise 8..16 or a Config error. threads is accepted for
 * drop-in parity and does not change the GPU path. */
prores_status prores_decoder_create(uint32_t bit_depth, uint32_t threads, prores_decoder **out,
                                    prores_error *err);
void prores_decoder_destroy(prores_decoder *d);

/* Decoder::decode (src/decode.rs:61-104): one frame of a MOV sample, bytes
 * after frame_size ignored. The library allocates out->data / out->alpha. */
prores_status prores_decode(const prores_decoder *d, const uint8_t *data, size_t len,
                            prores_frame *out, prores_error *err);

/* The same into caller buffers: out->data must hold the frame's planes
 * (prores_frame_layout) and out->alpha width*height samples when the frame
 * has alpha (prores_frame_header_parse says). No counterpart in the
 * reference beyond Decoder::decode; for callers that reuse buffers. */
prores_status prores_decode_into(const prores_decoder *d, const uint8_t *data, size_t len,
                                 prores_frame *out, prores_error *err);

/* Many frames from host memory (Decoder::decode on each of n packets): out[i]
 * and errs[i] as prores_decode gives them; a frame in error leaves out[i]
 * empty. The frames go through the GPU in chunks on two streams, so the
 * uploads and downloads of one chunk overlap the kernels of the next; the
 * library's frames live in page-locked memory, so the downloads run at the
 * link's speed. Returns PRORES_OK, or the device error that stopped the
 * call (then frames not decoded carry it). No counterpart in the reference
 * beyond Decoder::decode. */
prores_status prores_decode_host_batch(const prores_decoder *d, size_t n, const uint8_t *const *data,
                                       const size_t *len, prores_frame *out, prores_error *errs);

/* Frees what prores_decode / prores_decode_host_batch allocated (their
 * planes sit in a pool of page-locked buffers kept for reuse). */
void prores_frame_free(prores_frame *f);

/* FrameHeader::parse (src/header.rs:82-170). */
prores_status prores_frame_header_parse(const uint8_t *data, size_t len, prores_frame_header *out,
                                        prores_error *err);

/* Frame::layout (src/frame.rs:247-266): fills planes[3]; returns the total
 * sample count of data. */
size_t prores_frame_layout(uint32_t width, uint32_t height, uint32_t chroma, prores_plane *planes);

/* ---- Encoder (src/encode.rs:195-287) ---- */

/* Config::new (src/encode.rs:163-174). */
void prores_config_default(uint32_t profile, prores_config *cfg);

/* Encoder::encode (src/encode.rs:223-287). *out is malloc'd by the library;
 * release with prores_buffer_free. */
prores_status prores_encode(const prores_config *cfg, const prores_frame *frame, uint8_t **out,
                            size_t *out_len, prores_error *err);
void prores_buffer_free(uint8_t *p);

/* Profile helpers (src/encode.rs:46-118). */
uint64_t prores_profile_target_frame_bytes(uint32_t profile, uint32_t width, uint32_t height);
void prores_profile_fourcc(uint32_t profile, char out[4]);
int32_t prores_profile_from_fourcc(const char code[4]); /* -1 if none */

/* ---- Device-resident batch entry points ----
 * No counterpart in the reference (it runs on CPU threads). For callers that
 * keep bitstreams and planes in GPU memory (the batched path, the serving
 * loop, the benchmark): packets and outputs are device pointers. The host
 * copy of a packet, where the caller has one, is read for its headers only.
 * A batch with a job without one (host_packet NULL) is planned on the
 * device: the frame and picture headers (src/header.rs:82-290) are parsed
 * and the batch laid out by a kernel, so the launch sequence starts with no
 * host round trip, and header errors come back with the results, the
 * reference's messages. The device plan is laid out for the largest batch
 * the library has seen; a larger one is run again over a grown workspace.
 * A job without dev_packet or dev_data, or without dev_alpha for a frame
 * that has alpha, fails with PRORES_ERR_CONFIG and decodes nothing.
 */
typedef struct {
  const uint8_t *host_packet; /* host bytes (at least the frame and picture headers), or NULL */
  size_t len;                 /* bytes of the packet (>= frame_size) */
  const uint8_t *dev_packet;  /* device copy of the packet; 64 readable bytes past len */
  uint16_t *dev_data;         /* device output planes, Frame::layout at the decoder's depth */
  uint16_t *dev_alpha;        /* device output alpha (width*height) or NULL if none */
} prores_dev_decode_job;

/* Decodes n frames with one launch sequence on one of the library's two
 * per-call streams; waits for completion. Calls from two threads run at
 * once, each on its own stream and workspace (a third waits for the first
 * stream), so one batch's planning and launches overlap the other's
 * kernels. errs[i] receives frame i's result (kind 0 = Ok). Returns
 * PRORES_ERR_DEVICE only for runtime failures, else PRORES_OK. Uses the
 * device's batch tuning (the defaults until prores_set_decode_tuning or
 * prores_autotune_decode sets it). No counterpart in the reference: the
 * batched form of Decoder::decode (src/decode.rs:61-104). */
prores_status prores_decode_device_batch(const prores_decoder *d, size_t n,
                                         const prores_dev_decode_job *jobs, prores_error *errs);

/* ---- Allocation-free decode ----
 * No counterpart in the reference. For callers that cannot allocate device
 * memory while they run: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.
 *
 * prores_noalloc_init: set-up, once per device, with that device current:
 * loads the library's kernels, constant tables and stream there and looks up
 * that device's copy of the library's scratch array
 * (prores_noalloc_scratch_bytes(): 256 MiB, held by every GPU the library
 * runs on until the process ends).
 *
 * prores_decode_device_batch_noalloc: prores_decode_device_batch with the
 * batch planned on the device (each job's packet is read from dev_packet;
 * host_packet is not read) and its whole workspace in the scratch array: the
 * call allocates no device or page-locked memory and captures no graph.
 * Samples and errors are those of prores_decode_device_batch. A batch whose
 * workspace does not fit the scratch (for instance more than one 3840x2160
 * 4:4:4 frame) fails, every job with PRORES_ERR_CONFIG "the batch needs more
 * than the 268435456 bytes of its scratch memory"; so does one of more than
 * 2048 jobs, and a call on a device where prores_noalloc_init has not run.
 * Calls on one device share its scratch and run one at a time; the scratch
 * keeps the last call's contents. */
prores_status prores_noalloc_init(prores_error *err);
size_t prores_noalloc_scratch_bytes(void);
prores_status prores_decode_device_batch_noalloc(const prores_decoder *d, size_t n,
                                                 const prores_dev_decode_job *jobs, prores_error *errs);

/* How a decode batch is launched. Every setting gives the same samples and
 * the same errors; only the speed differs. No counterpart in the reference.
 *   persistent   0: one workgroup (entropy: one wave) per work item;
 *                1: a grid the size of the device takes items from a device
 *                   work queue until it is empty.
 *   chunk        items taken per queue pop (persistent form), 1..64.
 *   idct_blocks  8x8 blocks per IDCT workgroup item: consecutive slices of
 *                one picture are packed up to this many blocks, 8..96.
 *   entropy_wg   threads per entropy workgroup: 64, 128, 256, 512 or 768,
 *                or 0 to size workgroups from the batch (one a CU when the
 *                batch has at least 12 streams' waves a CU).
 *   graph        1: the launch sequence is captured once as a hipGraph and
 *                replayed while the workspace and launch shapes stay the
 *                same (with persistent = 1 the shapes never change). The
 *                library's own copies never use the legacy (null) stream;
 *                a capture that fails all the same (another thread's work
 *                invalidated it) is not retried for that shape, and its
 *                batches run as stream launches with the same results.
 *                prores_decode_host_batch and the serving loop always use
 *                stream launches. */
typedef struct {
  uint32_t persistent, chunk, idct_blocks, entropy_wg, graph;
} prores_decode_tuning;

void prores_decode_tuning_default(prores_decode_tuning *t);
/* prores_decode_device_batch with this tuning instead of the device's (the
 * same per-call streams, so two threads' calls run at once). */
prores_status prores_decode_device_batch_tuned(const prores_decoder *d, size_t n,
                                               const prores_dev_decode_job *jobs,
                                               const prores_decode_tuning *t, prores_error *errs);
prores_status prores_set_decode_tuning(const prores_decode_tuning *t, prores_error *err);
void prores_get_decode_tuning(prores_decode_tuning *t);

/* Times every candidate tuning on this batch (`reps` timed runs each after
 * one untimed run) and makes the fastest the device's batch tuning. best /
 * best_us (median microseconds of the whole launch sequence) may be NULL.
 * The candidates: both forms, chunk 1 and 4, idct_blocks 32, 64 and 96,
 * entropy_wg 0 (sized from the batch) and 256, with and without graph
 * replay. */
prores_status prores_autotune_decode(const prores_decoder *d, size_t n, const prores_dev_decode_job *jobs,
                                     uint32_t reps, prores_decode_tuning *best, double *best_us,
                                     prores_error *err);

/* One frame to encode from device memory: frame.data / frame.alpha are
 * DEVICE pointers laid out as frame.planes say; the rest of `frame` is host
 * metadata. On return dev_out / out_len locate the packet inside the
 * library's output arena (valid until the next encode batch on the device). */
typedef struct {
  const prores_config *cfg;
  prores_frame frame;
  const uint8_t *dev_out;
  size_t out_len;
} prores_dev_encode_job;

/* Encodes n frames with one launch sequence (Encoder::encode semantics per
 * frame, src/encode.rs:223-287); errs[i] receives frame i's result. */
prores_status prores_encode_device_batch(size_t n, prores_dev_encode_job *jobs, prores_error *errs);

/* Graph capture counters of the decode batches on the current device, all
 * lanes: [0] captures made, [1] captures that failed (their batches ran as
 * stream launches), [2] batches launched as a captured graph. No
 * counterpart in the reference. */
void prores_graph_counters(uint64_t *out3);

/* Microseconds of the last decode batch's kernels by stage, measured with
 * HIP events on its stream (of the per-call batches, the one that ended
 * last; with two in flight each also shares the GPU with the other while
 * it runs): [0] slice table, [1] colour entropy
 * decode, [2] alpha decode, [3] idct, [4] all kernels. No counterpart in the
 * reference. */
void prores_last_batch_timing(double *us5);

/* Microseconds of the last encode batch by stage: [0] fdct + alpha, [1] rate
 * control (probes, window, share-out, offsets), [2] pack, [3] whole batch
 * (host sync for the output sizes included). */
void prores_last_encode_timing(double *us4);

/* Microseconds of the last encode batch on the per-call lane, kernel by
 * kernel from HIP events on its stream: [0] pe_fdct, [1] pe_alpha, [2] the
 * nine probe rounds (pe_size + pe_search), [3] pe_window, [4] pe_shareout,
 * [5] pe_offsets, [6] pe_pack, [7] the whole batch (the host's wait for the
 * packet sizes included). The reference's bench_api times the same steps
 * per block on the CPU (src/bench_api.rs:95-178). */
void prores_last_encode_stages(double *us8);

/* Counters of the last encode batch on the per-call lane: [0] sizes the
 * share-out computed past its precomputed window, [1] the most for one
 * picture, [2] pictures, [3] mean uniform quantiser index. No counterpart in
 * the reference (rate-control diagnostics). */
void prores_last_encode_counters(uint64_t *out4);

/* ---- Serving loop (no counterpart in the reference) ----
 *
 * Decode and encode requests are submitted at any time from any thread. A
 * request is admitted while its device-memory footprint (its buffers and its
 * share of the batch's working arenas) fits hbm_budget, first in first out;
 * each worker (one for decode, one for encode, on separate streams so their
 * batches overlap on the device) takes every admissible request of its kind
 * each time it comes round, up to max_batch_frames, so requests that arrive
 * while a batch runs ride the next one. Inputs are copied at submit. Results
 * stay resident in device memory (dev_out / dev_alpha) until released. A
 * request's result is the same alone and in any mix: each frame's bytes and
 * samples are those of prores_decode / prores_encode. Times are microseconds
 * on the server's clock. */
typedef struct {
  uint64_t hbm_budget;        /* bytes admitted requests may hold; >= 1 MiB */
  uint32_t max_batch_frames;  /* frames per device batch; >= 1 */
  uint32_t decode_bit_depth;  /* 0 = natural, else 8..16 */
} prores_server_config;

typedef struct prores_server prores_server;

typedef struct {
  uint64_t id;
  uint32_t kind;        /* 1 decode, 2 encode */
  prores_error error;   /* kind 0 = Ok */
  const void *dev_out;  /* decode: u16 planes in Frame layout; encode: the packet */
  size_t out_bytes;
  const void *dev_alpha; /* decode: the alpha plane, or NULL */
  size_t alpha_bytes;
  prores_frame_header header; /* decode: the packet's frame header */
  double submit_us, start_us, done_us;
  uint32_t batch_frames; /* frames in the device batch that carried it */
} prores_server_result;

typedef struct {
  uint64_t submitted, completed, failed;
  uint64_t decode_batches, encode_batches, max_batch_frames;
  uint64_t held_bytes, peak_held_bytes, hbm_budget, pending;
  uint64_t staged_ahead; /* requests uploaded while their kind's previous batch was on the GPU */
} prores_server_stats;

void prores_server_config_default(prores_server_config *cfg);
prores_status prores_server_create(const prores_server_config *cfg, prores_server **out, prores_error *err);
/* Waits for the batches in flight, then frees everything. */
void prores_server_destroy(prores_server *s);
/* Returns the request id. A request that can never run (a bad header or
 * Config, or a footprint over the budget) completes at once with its error. */
uint64_t prores_server_submit_decode(prores_server *s, const uint8_t *packet, size_t len);
uint64_t prores_server_submit_encode(prores_server *s, const prores_config *cfg, const prores_frame *frame);
/* Moves up to max completed results to out; waits up to wait_ms (-1:
 * until one completes) when none is ready. Returns how many. */
size_t prores_server_poll(prores_server *s, prores_server_result *out, size_t max, int32_t wait_ms);
/* Frees a polled result's device memory; PRORES_ERR_CONFIG for an unknown or
 * unpolled id. */
prores_status prores_server_release(prores_server *s, uint64_t id);
void prores_server_get_stats(prores_server *s, prores_server_stats *out);

/* Device memory on the current device, for callers without an allocator of
 * their own (the benchmark, the serving loop, the tests). No counterpart in
 * the reference. kind: 1 host->device, 2 device->host, 3 device->device. */
void *prores_device_alloc(size_t bytes);
void prores_device_free(void *p);
prores_status prores_memcpy(void *dst, const void *src, size_t bytes, int32_t kind);
/* The whole frames of a MOV file (or a prefix of one), found by their
 * 'icpf' identifier and frame_size without parsing the container, as the
 * reference's sample test and bench do (tests/sample.rs:26-41,
 * examples/prores_bench.rs:104-122): returns how many there are and stores
 * the first `max` byte offsets and sizes. */
size_t prores_find_frames(const uint8_t *data, size_t len, uint64_t *offsets, uint64_t *sizes, size_t max);

/* FNV-1a 64 of `len` bytes from state h (start: 0xcbf29ce484222325), the
 * hash the reference's bench prints for packets and pictures
 * (examples/prores_bench.rs:96-101). */
uint64_t prores_fnv1a(uint64_t h, const void *data, size_t len);

/* The HIP error of this thread's last failed prores_memcpy. */
const char *prores_memcpy_error(void);

/* Library build string. */
const char *prores_version(void);

/* The run/level code tables the entropy decode reads (for inspection and
 * tests; no device work): level[6 * 1024] entries for the next 10 bits of a
 * stream (a level code and its sign), run[6 * 512] for the next 9 bits, one
 * block per distinct codebook, in the order the codebooks first appear
 * among the contexts of src/vlc.rs:30-68. An entry is
 *   level: 0x8000 | next << 11 | (signed level & 127) << 4 | (code bits + 1)
 *   run:   0x8000 | next << 10 | run << 4 | code bits
 * where next is the block of the context the symbol sets, or 0 when the
 * code is longer than 9 bits. No counterpart in the reference. */
void prores_entropy_tables(uint16_t *level, uint16_t *run);

#ifdef __cplusplus
}
#endif
#endif
