// Host-side internals shared by the decode and encode runtimes: the error
// type, header syntax, slice layout, constant tables and the device context.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/prores_hip.h"
#include "layout.h"

namespace prores_hip {

// The crate's error (src/error.rs:5-21): a kind and the message text.
struct Err {
  int kind = PRORES_OK;
  std::string msg;
  bool ok() const { return kind == PRORES_OK; }
};
Err invalid(const std::string &m);
Err unsupported(const std::string &m);
Err config_error(const std::string &m);
Err device_error(const std::string &where, hipError_t e);
std::string fmtstr(const char *f, ...) __attribute__((format(printf, 1, 2)));
void fill_error(prores_error *out, const Err &e);

extern const char *const TIMES;    // U+00D7 in the reference's size messages
extern const char *const EN_DASH;  // U+2013 in its ranges

// header::slice_sizes (src/header.rs:308-320).
std::vector<uint32_t> slice_sizes(uint32_t width_in_mb, uint32_t log2);
inline uint32_t qscale(uint32_t q) { return q <= 128 ? q : 128 + 4 * (q - 128); }

// FrameHeader (src/header.rs:44-255).
struct Header {
  prores_frame_header h;
  uint32_t alpha_bits() const { return h.alpha == 1 ? 8 : h.alpha == 2 ? 16 : 0; }
  uint32_t picture_count() const { return h.interlace == 0 ? 1 : 2; }
  bool top_field(uint32_t i) const { return (h.interlace == 1) == (i == 0); }
  uint32_t picture_height(uint32_t i) const {
    if (h.interlace == 0) return h.height;
    return top_field(i) ? (h.height + 1u) / 2 : h.height / 2u;
  }
  void picture_rows(uint32_t i, uint32_t &first, uint32_t &step) const {
    if (h.interlace == 0) {
      first = 0;
      step = 1;
    } else {
      first = top_field(i) ? 0 : 1;
      step = 2;
    }
  }
  void luma_weights(uint8_t *w) const;
  void chroma_weights(uint8_t *w) const;
};
Err parse_frame_header(const uint8_t *data, size_t len, Header &out);
Err check_plausible_size(const Header &h);

// PictureHeader (src/header.rs:262-290).
struct PicHeader {
  uint32_t header_size, picture_size, log2_slice_mbs;
};
Err parse_picture_header(const uint8_t *d, size_t len, PicHeader &out);

// One line of work on a device: a stream, its timing events, the decode and
// encode workspaces, and the lock that makes its batches one at a time.
// Lane 0 serves the per-call API; the serving loop's workers take other
// lanes, so their batches overlap on the device. Streams keep the default
// priority.
struct Lane {
  int index = 0;
  hipStream_t stream = nullptr;
  // A second stream for work that runs beside the main one (alpha decode),
  // forked from and joined back into `stream` with the two untimed events.
  hipStream_t aux = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // The encoder's stage markers: before the transform, after the transform,
  // the alpha coding, the probes, the window, the share-out, the offsets;
  // before and after the pack.
  static constexpr int ENC_EVENTS = 9;
  hipEvent_t eev[ENC_EVENTS] = {};
  std::mutex mu;
  void *dec = nullptr;  // decode workspace (decode_host.cpp)
  void *enc = nullptr;  // encode workspace (encode_host.cpp)
  double dec_us[5] = {0, 0, 0, 0, 0};  // last decode batch by stage
  uint64_t dec_seq = 0;                // when it ended, in the device's batch count
  double enc_us[4] = {0, 0, 0, 0};     // last encode batch by stage
  double enc_stage_us[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // see prores_last_encode_stages
  uint64_t enc_counters[4] = {0, 0, 0, 0};  // last encode batch: see prores_last_encode_counters
};
constexpr int MAX_LANES = 8;
// The lanes of the per-call device batch API: a call takes the first that
// is free, so two callers' batches are in flight at once.
constexpr int CALL_LANES[2] = {0, 6};

// The per-device context: modules, kernels, constant tables, lanes.
struct Device {
  int id = -1;
  hipModule_t dec_mod = nullptr, enc_mod = nullptr;
  hipFunction_t k_slice_table = nullptr, k_entropy = nullptr, k_alpha = nullptr, k_idct = nullptr,
                k_plan = nullptr, k_order = nullptr;
  hipFunction_t k_fdct = nullptr, k_alpha_enc = nullptr, k_size = nullptr, k_search = nullptr,
                k_shareout = nullptr, k_offsets = nullptr, k_pack = nullptr,
                k_window = nullptr, k_arena = nullptr, k_slices = nullptr, k_size_first = nullptr,
                k_fdct64 = nullptr;  // pe_fdct for batches whose slices have at most 64 blocks
  DevTables *tables = nullptr;
  std::mutex mu;  // lane creation, encoder loading, the batch tuning
  Lane *lanes[MAX_LANES] = {nullptr};
  prores_decode_tuning tuning;
  bool tuned = false;
  std::atomic<uint64_t> batches{0};  // decode batches ended (Lane::dec_seq)
};
// The context of the current device, created on first use.
Err device(Device *&out);
// Lane `index` of `d`, created on first use.
Err lane(Device &d, int index, Lane *&out);
// Loads the encoder's kernels into `d` (once; takes d.mu).
Err ensure_encoder(Device &d);

// A grow-only device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  Err reserve(size_t bytes);
  template <class T>
  T *as() const {
    return (T *)p;
  }
};

// Page-locked host memory, grow-only: a lane's staging for a batch plan
// going up and its results coming back, so that each is one asynchronous
// copy. The caller holds the lane's lock and has waited for the last copy
// out of it before it grows or refills.
struct HostBuf {
  void *p = nullptr;
  size_t cap = 0;
  Err reserve(size_t bytes);
};

// Host memory for staging copies: page-locked from the library's reuse pool
// (so a copy to the device runs at the link's speed), else malloc'd;
// returned with host_staging_put either way.
void *host_staging_get(size_t bytes);
void host_staging_put(void *p);

// `lds`: bytes of dynamic LDS per workgroup.
Err launch(Lane &L, hipFunction_t f, unsigned gx, unsigned gy, unsigned bx, void **args, unsigned lds = 0);
Err launch_on(hipStream_t s, hipFunction_t f, unsigned gx, unsigned gy, unsigned bx, void **args,
              unsigned lds = 0);
Err check(hipError_t e, const char *where);
// Copies that wait for their completion without touching the legacy stream:
// on a library copy stream of the current device, or on a lane's stream
// (the caller holds L.mu).
Err copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
Err copy_on(Lane &L, void *dst, const void *src, size_t bytes, hipMemcpyKind kind);

void build_tables(DevTables &t);

// Batches on a given lane (the caller holds L.mu): the bodies of
// prores_decode_device_batch and prores_encode_device_batch.
Err decode_batch_on(Device &d, Lane &L, uint32_t bit_depth, size_t n, const prores_dev_decode_job *jobs,
                    const prores_decode_tuning *t, prores_error *errs);
Err encode_batch_on(Device &d, Lane &L, size_t n, prores_dev_encode_job *jobs, prores_error *errs);
// A device-planned decode batch whose workspace is carved from `region`
// (device memory and host staging the caller owns) instead of the lane's
// buffers: it allocates no device or page-locked memory, and replays no
// graph. The caller holds L.mu.
// Memory a device-planned batch lays its workspace out in instead of the
// lane's own buffers (the allocation-free entry point): device bytes from
// `dev`, host staging from `host`, carved afresh for each attempt.
struct FixedRegion {
  uint8_t *dev = nullptr;
  size_t dev_bytes = 0, dev_used = 0;
  uint8_t *host = nullptr;
  size_t host_bytes = 0;
  // Points `b` at the next `bytes` of the region (256-byte aligned); false
  // when they do not fit.
  bool take(DevBuf &b, size_t bytes) {
    const size_t at = (dev_used + 255) & ~(size_t)255;
    if (at > dev_bytes || bytes > dev_bytes - at) return false;
    b.p = dev + at;
    b.cap = bytes;
    dev_used = at + bytes;
    return true;
  }
};
Err decode_batch_fixed(Device &d, Lane &L, const prores_decoder *dec, size_t n, const prores_dev_decode_job *jobs,
                       FixedRegion &region, prores_error *errs);
// The lane the allocation-free entry point runs on, and its decode
// workspace record (made on first use).
constexpr int NOALLOC_LANE = 5;
void *decode_workspace(Lane &L);

// The Config and Frame checks of Encoder::encode, without device work.
Err encode_check(const prores_config &cfg, const prores_frame &frame);

}  // namespace prores_hip
