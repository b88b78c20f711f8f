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
  uint32_t pad[3];
};
// Slices a pd_slice_table workgroup (one thread each); a picture's chunks
// start at multiples of it, and a picture without slices has one chunk.
constexpr uint32_t ST_CHUNK = 256;

// A run of consecutive slices of one picture transformed by one workgroup
// (their coefficients are contiguous in the arena). As a row segment,
// `first` counts from the start of the row; as an item, from the batch's
// first slice. pd_slice_table expands each picture's row segments into its
// items (rows x segments).
struct IdctItem {
  uint32_t first, count, blocks, pad;
};
// Header bytes pd_headers gathers from a packet held only in device memory,
// one HDR_BLOB record a packet: the packet's first HDR_FRAME_BYTES bytes (a
// frame header with both quantisation matrices ends at byte 156), the first
// HDR_PIC_BYTES bytes of each of the two pictures, and the two pictures'
// byte offsets (u64; ~0 where the packet is too short to locate one).
constexpr uint32_t HDR_FRAME_BYTES = 160, HDR_PIC_BYTES = 8, HDR_BLOB = 192;
constexpr uint32_t HDR_PIC_AT = HDR_FRAME_BYTES, HDR_POS_AT = HDR_FRAME_BYTES + 2 * HDR_PIC_BYTES;
static_assert(HDR_POS_AT + 16 == HDR_BLOB, "header record layout");

// Entropy work placement: batches of at most ORDER_MAX grid slots have
// their entropy items ordered by cost (the largest slice of the item's 64,
// in bytes) so that every workgroup gets a mix of long and short streams.
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
