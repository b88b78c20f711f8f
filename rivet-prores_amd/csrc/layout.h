// Plain-old-data records shared by the host runtime and the gfx950 kernels.
// Both sides include this file; nothing here depends on HIP or on the host
// compiler. Sizes are fixed-width so the layouts agree.
#pragma once
#include <stdint.h>

namespace prores_hip {

// Slice-level error codes the kernels record (one 32-bit word per slice
// component: code | arg << 8). The host turns them into the reference's
// messages (src/decode.rs:301-338, src/bits.rs:59-94, src/vlc.rs:206-230,
// src/decode.rs:484-595).
enum SliceError : uint32_t {
  E_NONE = 0,
  E_SLICE_SHORT = 1,     // "a slice is shorter than its header"
  E_SLICE_HDR_SIZE = 2,  // "slice_header_size {arg} is out of range"
  E_QINDEX = 3,          // "quantization_index {arg} is reserved"
  E_COMP_SIZES = 4,      // "slice component sizes exceed the slice"
  E_PAST_END = 5,        // "a codeword runs past the end of its data"
  E_PREFIX = 6,          // "a variable-length code's prefix is too long"
  E_RANGE = 7,           // "a variable-length code's value is out of range"
  E_DC_OVERFLOW = 8,     // "a DC coefficient overflows"
  E_RUN_PAST = 9,        // "a run passes the end of the coefficient array"
  E_ALPHA_RUN = 10,      // "an alpha run passes the end of the slice"
};

// Process-wide constant tables, uploaded once per device.
struct DevTables {
  float basis[64];              // T[x][u] at 8x+u, C(u)/2 cos((2x+1)u pi/16) rounded to f32
  uint8_t scan[2][64];          // [progressive, interlaced]: scanned index of raster position
  // Run and level codes of at most 9 bits, which the entropy decode reads
  // from LDS; 0 marks a longer code. Each table is indexed by the next bits
  // of the stream and `next` is the book of the context the symbol sets.
  // fast_level[book][next 10 bits]: a level code and its sign bit,
  //   FAST_VALID | next << 11 | (signed level & 127) << 4 | (length + 1);
  // fast_run[book][next 9 bits]:
  //   FAST_VALID | next << 10 | run << 4 | length.
  // Books are the six distinct level and six distinct run codebooks.
  uint16_t fast_level[6][1024];
  uint16_t fast_run[6][512];
  // The encoder's code lengths (Codebook::len, src/vlc.rs:155-172) by
  // context, for the symbols the size count meets most: runs below 64 and
  // levels (|q| - 1) below 256; larger symbols take the closed form.
  uint8_t enc_run_len[16][64];
  uint8_t enc_level_len[9][256];
};
constexpr uint32_t ENC_LEN_BYTES = 16 * 64 + 9 * 256;
constexpr uint32_t FAST_VALID = 0x8000;
constexpr uint32_t FAST_RUN_BASE = 6 * 1024 * 2;  // byte offset of fast_run after fast_level

// A frame to decode into (device pointers and per-frame constants).
struct DecFrame {
  uint64_t out_plane[3];  // u16* Y', Cb, Cr (each tightly packed, stride = width)
  uint64_t out_alpha;     // u16* or 0
  uint32_t plane_w[3];
  uint32_t width, height;
  uint32_t chroma;        // 2 = 4:2:2, 3 = 4:4:4
  uint32_t alpha_bits;    // 0, 8, 16
  uint32_t bit_depth;     // output depth
  uint32_t interlaced;    // selects the scan
  float out_scale, out_bias, out_max;
  uint32_t rescale_identity;
  uint64_t rescale_m;
  uint8_t luma_w[64], chroma_w[64];
};

// One picture (a frame, or one field of an interlaced frame).
struct DecPic {
  uint64_t pic_off;        // device address of the picture header
  uint64_t coef_base;      // element offset of the picture's first slice coefficients
  uint32_t header_size, picture_size;
  uint32_t n_slices, row_slices;
  uint32_t slice_base;     // global index of the first slice
  uint32_t layout_base;    // index of the first RowSlot of this picture
  uint32_t row_coefs;      // coefficients in one row of slices
  uint32_t frame;          // DecFrame index
  uint32_t height, height_in_mb, first_row, row_step;
  uint32_t seg_base, n_segs;  // IDCT segments of one row (the same for every row)
  uint32_t item_base;         // index of the picture's first IDCT item
  uint32_t pad;
};

// Where slice j of a row sits: its first macroblock, its size and the
// offset of its coefficients inside the row.
struct RowSlot {
  uint16_t mb_x, mbs;
  uint32_t coef_prefix;
};

// One slice, located by the device slice-table walk.
struct SliceRec {
  uint64_t off;    // device address of the slice
  uint64_t coef;   // element offset of its coefficients
  uint32_t len;    // bytes
  uint32_t pic;    // DecPic index
  uint16_t mb_x, mb_y;
  uint32_t mbs;
};

// One slice of the alpha work order (pd_slice_table, read by pd_alpha): its
// alpha stream and what it belongs to, so that the alpha decode starts from
// one load. Only slices of frames with alpha whose header passed its checks
// are listed; the colour decode reports the others' header errors.
struct AlphaRec {
  uint64_t stream;   // device address of the alpha stream
  uint32_t g;        // the slice (its SliceRec and error words)
  uint32_t len;      // stream bytes
  uint32_t pic;      // DecPic index
  uint32_t frame;    // DecFrame index
  uint32_t mb;       // mb_x | mb_y << 16
  uint32_t mbs;      // macroblocks
  uint32_t head[4];  // the stream's first ALPHA_HEAD_BYTES bytes, in memory order
};
static_assert(sizeof(AlphaRec) == 48, "AlphaRec layout");
// A stream this short is decoded from its record alone (most alpha slices:
// a constant slice is one difference and one run, 5 bytes).
constexpr uint32_t ALPHA_HEAD_BYTES = 16;

// Counts of a batch and the heads of its work queues, in device memory so
// that the launch arguments stay the same from one batch to the next (a
// captured graph replays unchanged). The host uploads it before every
// launch sequence, so every counter starts at zero.
struct DecBatch {
  uint32_t n_pics, n_slices;
  uint32_t n_entropy_items;  // 64 streams of one component each: 3 * ceil(n_slices / 64)
  uint32_t n_idct_items;
  uint32_t chunk;            // items taken per queue pop by a persistent workgroup
  uint32_t heavy, light;     // alpha work order fill counts (zero at upload)
  uint32_t wide;             // nonzero once the entropy decode stores a value outside +-32767
  uint32_t q[4];             // queue heads: entropy, alpha, idct
  uint32_t n_chunks;         // pd_slice_table's chunks of ST_CHUNK slices
  uint32_t ordered;          // nonzero once pd_entropy_order has written the order
  uint32_t pad[2];
};
// Slices a pd_slice_table workgroup (one thread each); a picture's chunks
// start at multiples of it, and a picture without slices has one chunk.
constexpr uint32_t ST_CHUNK = 256;

// A run of consecutive slices of one picture transformed by one workgroup
// (their coefficients are contiguous in the arena). As a row segment,
// `first` counts from the start of the row; as an item, from the batch's
// first slice. pd_slice_table expands each picture's row segments into its
// items (rows x segments), and gives each item its picture (`pic`: a row
// segment leaves it 0), so pd_idct reads the picture beside the slices.
struct IdctItem {
  uint32_t first, count, blocks, pic;
};
// ---------------------------------------------------------------------------
// Device-planned batches (pd_plan): the frame and picture headers of packets
// held in device memory are parsed on the device (FrameHeader::parse,
// check_plausible_size, PictureHeader::parse: src/header.rs:82-290,
// src/decode.rs:113-128) and the batch laid out there, so the launch sequence
// needs no host round trip first.

// One job: device addresses of the packet, the output planes and alpha, and
// the packet's length (64 readable bytes past it).
struct DevJob {
  uint64_t packet, len, data, alpha;
};

// Header errors, in the order the host's checks make them. A code | the
// arguments the message needs (PlanRec::arg). F_* stop the frame before any
// picture; P_* stop it after `n_pics` good pictures, and count only when
// those decode cleanly (the first error in picture order).
enum PlanError : uint32_t {
  F_NONE = 0,
  F_SHORT = 1,         // "{len} bytes is too short for a ProRes frame header"
  F_ICPF = 2,          // "no 'icpf' frame identifier"
  F_SIZE_EXCEEDS = 3,  // "frame_size {a0} exceeds the {a1} bytes given"
  F_SIZE_SMALL = 4,    // "frame_size leaves no room for the frame header"
  F_HDR_SIZE = 5,      // "frame_header_size {a0} is out of range"
  F_VERSION = 6,       // Unsupported "bitstream_version {a0}"
  F_DIMS = 7,          // "a {a0}x{a1} frame"
  F_CHROMA = 8,        // "chroma_format {a0} is reserved"
  F_INTERLACE = 9,     // "interlace_mode 3 is reserved"
  F_ALPHA = 10,        // "alpha_channel_type {a0} is reserved"
  F_MATRIX = 11,       // "a quantisation matrix runs past frame_header_size"
  F_IMPLAUSIBLE = 12,  // "frame_size {a0} is too small for a {a1}x{a2} frame (at least {a3} bytes)"
  F_NO_PACKET = 13,    // Config: "a decode job has no device packet"
  F_NO_PLANES = 14,    // Config: "a decode job has no device planes"
  F_NO_ALPHA = 15,     // Config: "a decode job has no alpha plane for a frame with alpha"
  P_START_PAST = 16,   // "a picture starts past the end of the frame"
  P_HDR_SHORT = 17,    // "a picture header runs past the end of the frame"
  P_HDR_SIZE = 18,     // "picture_header_size {d} is out of range"
  P_SIZE = 19,         // "picture_size {d} is out of range"
  P_TABLE = 20,        // "the slice table runs past picture_size"
};

// pd_plan's record of one job.
struct PlanRec {
  uint32_t code;      // F_* or F_NONE
  uint32_t deferred;  // P_* after the good pictures, or F_NONE
  uint32_t first_pic, n_pics;  // the good pictures (DecPic indices)
  uint64_t arg[4];    // the arguments of `code`
  uint32_t darg;      // the argument of `deferred`
  uint32_t pad;
};
static_assert(sizeof(PlanRec) == 56, "PlanRec layout");

// pd_plan's scratch: what the parse of one job leaves for the layout pass
// (its good pictures and their counts), and the job's bases in the batch.
struct PlanPic {
  uint64_t pos;  // byte offset of the picture in the packet
  uint32_t header_size, picture_size, log2, row_slices, height, hmb, n_segs, max_count;
};
struct PlanJob {
  uint32_t n_pics, wmb, cpm, pad;
  uint32_t slices, rows, segs, items, chunks, max_count, wanted;
  uint64_t coefs;
  uint32_t base_pic, base_slice, base_row, base_seg, base_item, base_chunk;
  uint64_t base_coef;
  PlanPic pic[2];
};

// The capacities pd_plan lays a batch into (elements), and the batch's own
// counts it found: when any count is over its capacity the batch is marked
// (overflow) and does no work, and the host grows its workspace and runs it
// again.
struct PlanCaps {
  uint32_t pics, rows, segs, chunks, slices, items, item_slices, pad;
  uint64_t coefs;
};
struct PlanTotals {
  uint32_t overflow, pics, rows, segs, chunks, slices, items, max_count;
  uint64_t coefs;
  uint32_t wanted_count;  // the most slices an IDCT item would take without the capacity's limit
  uint32_t pad;
};
static_assert(sizeof(PlanTotals) == 48, "PlanTotals layout");

// Entropy work placement: batches of at most ORDER_MAX grid slots have
// their entropy items ordered by cost (the longest of the item's 64 streams,
// in bytes), and the waves that share a SIMD get a long, a middle and a
// short item.
constexpr uint32_t ORDER_MAX = 8192;

constexpr uint32_t IDCT_MAX_BLOCKS = 96;  // blocks one workgroup holds in LDS
constexpr uint32_t IDCT_MAX_SLICES = 12;  // slices of one item (a slice is at least 8 blocks)
// Entropy decode LDS: two 32-coefficient half-lines a lane (16-byte aligned),
// and a 64-entry list of 16-byte entries a wave. Kept to 126 KB for a
// 12-wave workgroup with its tables, so two alpha workgroups fit beside it
// on a CU.
constexpr uint32_t LINE_WORDS = 32;
constexpr uint32_t ENTROPY_LDS_PER_WAVE = 64 * LINE_WORDS * 4 + 64 * 16;
constexpr uint32_t ENTROPY_MAX_WG = 768;  // 12 waves: 108 KB of lines and lists plus 18 KB of tables

// Slices whose alpha stream is longer than this many bytes go first in the
// alpha work order: their decode is one long serial chain.
constexpr uint32_t ALPHA_HEAVY_BYTES = 256;

// Picture status words written by the kernels.
enum PicStatus : uint32_t { PIC_OK = 0, PIC_TABLE_ERR = 1 };

}  // namespace prores_hip
