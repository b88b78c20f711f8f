// Encoder records shared by the host runtime and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace prores_hip {

constexpr uint32_t ENC_MAXK = 8;         // window columns of the size table
// The search over qi probes 224, then bisects [1, 224] in at most 8 steps
// (224 < 2^8). The first pass sizes every slice at the first five probes of
// the path on which each one fits (224, 112, 56, 28, 14); each later pass at
// the one probe pending. The first pass settles at least two probes (224 and
// 112, or 224 alone when it does not fit), so 8 passes settle every search,
// and at most 5 + 7 probe results are kept.
constexpr uint32_t ENC_PASS_Q = 5;       // quantisers of the first pass
constexpr uint32_t ENC_PASSES = 8;       // size passes
constexpr uint32_t ENC_HIST = 12;        // probe results kept per picture
constexpr uint32_t ENC_DEEP = 224;       // size cache columns per slice (column t: qi = uniform - t)
constexpr uint32_t SHARE_WAVES = 8;      // waves of a pe_shareout workgroup
constexpr uint32_t DEEP_STEP = 4;        // quantisers coded per cache fill
constexpr uint32_t ALPHA_SLICE_CAP = 4608;  // bytes: 2048 values x 18 bits, the worst case

struct EncFrame {
  uint64_t in_plane[3];   // u16* Y', Cb, Cr (tightly packed)
  uint64_t in_alpha;      // u16* or 0 (alpha coded only when alpha_bits != 0)
  uint32_t plane_w[3];
  uint32_t width, height, chroma, bit_depth, interlaced;
  uint32_t alpha_bits;    // 0, 8, 16: as coded
  uint32_t rescale_identity, max_in;
  uint64_t rescale_m;     // frame depth -> alpha bits
  uint32_t first_pic, n_pics;
  uint32_t hdr_len;       // bytes of the frame header (with the frame_size and 'icpf' prefix)
  // FrameHeader fields pe_pack writes (src/header.rs:174-199).
  uint8_t version, chroma_interlace, aspect_rate, alpha_type;
  uint8_t encoder_id[4];
  uint8_t primaries, transfer, matrix, matrix_flags;
  uint8_t luma_matrix[64], chroma_matrix[64];
  float in_scale;         // 512 / 2^bit_depth
  uint32_t first_slice;   // global index of the frame's first slice
};

struct EncPic {
  uint64_t budget;        // per-picture budget before alpha and overhead
  uint64_t out_off;       // picture offset inside the frame (pe_offsets)
  uint64_t coef_base;     // element offset of the picture's first slice coefficients
  uint32_t frame, n_slices, slice_base, height;
  uint32_t first_row, row_step, header_bytes, log2_slice_mbs;
  uint32_t qtab;          // index of the 225 x 2 x 64 inverse-quantiser table
  uint32_t interlaced;
  // Slice layout: `rows` rows of `row_slices` slices, laid out as the
  // EncRowSlot entries from `layout_base`; `row_mbs` macroblocks a row.
  uint32_t rows, row_slices, layout_base, row_mbs;
};

// Slice j of a picture's row: its first macroblock, its macroblocks and the
// macroblocks of the row before it (pe_slices expands the slice records).
struct EncRowSlot {
  uint16_t mb_x, mbs;
  uint32_t mbs_before;
};

struct EncSlice {
  uint64_t coef;          // element offset of the slice's f32 coefficients
  uint32_t pic;
  uint16_t mb_x, mb_y;
  uint32_t mbs;
  uint32_t mbs_from;      // macroblocks from this slice to the end of its picture
};

// The picture-level rate-control state (src/encode.rs:685-794).
struct EncState {
  unsigned long long psum[ENC_PASS_Q];  // pass accumulators: total bytes at each listed qi
  uint32_t pnotfit[ENC_PASS_Q];         // and the slices that do not fit at it
  uint32_t nq;                          // quantisers listed for the next pass (0: search done)
  uint32_t pq[ENC_PASS_Q];              // the list, coarsest first
  uint32_t alpha_total;    // alpha bytes of the picture
  uint64_t budget;         // budget after alpha and overhead
  uint32_t phase;          // 0: probing 224, 1: bisecting, 2: done
  uint32_t probe;          // qi of the pending probe, 0 when none
  uint32_t lo, hi;
  uint32_t fits_at_all, uniform;
  uint32_t window_k;       // columns of the size table: uniform, uniform-1, ...
  uint32_t fallback;       // sizes the share-out computed itself (diagnostic)
  uint32_t round;          // probe results kept
  uint32_t probe_qi[ENC_HIST];            // the qi of each
  unsigned long long probe_sum[ENC_HIST];  // and the picture's total bytes at it
};

}  // namespace prores_hip
