// Decoder host runtime: Decoder::decode (src/decode.rs:61-104) over the
// gfx950 kernels. The host parses the frame and picture headers (a few bytes
// each) and lays out the batch; the slice-table walk, entropy decoding, IDCT
// and placement run on the device. Errors come back per slice and are
// reported as the reference does: the first in picture order.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <thread>

#include "host_internal.h"

struct prores_decoder {
  uint32_t bit_depth;  // 0: natural
  uint32_t threads;
};

namespace prores_hip {
namespace {

const char *slice_message(uint32_t code) {
  switch (code) {
    case E_SLICE_SHORT: return "a slice is shorter than its header";
    case E_COMP_SIZES: return "slice component sizes exceed the slice";
    case E_PAST_END: return "a codeword runs past the end of its data";
    case E_PREFIX: return "a variable-length code's prefix is too long";
    case E_RANGE: return "a variable-length code's value is out of range";
    case E_DC_OVERFLOW: return "a DC coefficient overflows";
    case E_RUN_PAST: return "a run passes the end of the coefficient array";
    case E_ALPHA_RUN: return "an alpha run passes the end of the slice";
    default: return "unknown slice error";
  }
}

Err slice_error(uint32_t word) {
  uint32_t code = word & 0xFF, arg = word >> 8;
  if (code == E_SLICE_HDR_SIZE) return invalid(fmtstr("slice_header_size %u is out of range", arg));
  if (code == E_QINDEX) return invalid(fmtstr("quantization_index %u is reserved", arg));
  return invalid(slice_message(code));
}

// One frame of a batch: where its pictures sit in the batch arrays, and the
// error the host found in its headers, which counts only when every picture
// before it decodes cleanly.
struct FramePlan {
  Err early;            // header errors before any picture: no device work
  Err deferred;         // a picture header error after `n_pics` good pictures
  uint32_t first_pic = 0, n_pics = 0;
};

struct Batch {
  std::vector<DecFrame> frames;
  std::vector<DecPic> pics;
  std::vector<RowSlot> rows;
  std::vector<FramePlan> plans;
  uint64_t n_slices = 0, n_coefs = 0;
  bool any_alpha = false;
};

// A captured launch sequence and what it was captured against: the kernel
// arguments are workspace addresses and launch shapes, so it replays while
// those stay the same.
struct GraphKey {
  void *ptrs[22];
  unsigned grids[4];
  unsigned ewg, idct_lds, n_jobs, idct_cap;
  prores_decode_tuning t;
  PlanCaps caps;
  bool alpha, device_plan;
  bool operator==(const GraphKey &o) const { return memcmp(this, &o, sizeof *this) == 0; }
};

struct Workspace {
  DevBuf slices, slice_err, coefs, wide, packet, planes, items, alpha_order, order;
  // The batch plan (frames, pictures, row slots, IDCT segments, slice-table
  // chunks, entropy item costs, header) in one device buffer, each part at a
  // fixed offset while it fits its capacity; the per-picture status and
  // first-error words in another.
  DevBuf plan, picres;
  HostBuf hplan, hres;
  size_t cap_frames = 0, cap_pics = 0, cap_rows = 0, cap_segs = 0, cap_chunks = 0, cap_cost = 0, cap_np = 0;
  // Device-planned batches: the jobs, pd_plan's scratch, and its results
  // (job records, totals, picture status and first-error words) in one
  // buffer that comes back in one copy; the capacities it plans into (the
  // largest batch seen, grow-only).
  DevBuf jobs, scratch, dres;
  PlanCaps dcap{};
  uint32_t idct_cap_slices = IDCT_MAX_SLICES;
  hipGraphExec_t graph = nullptr;
  GraphKey key;
  bool graph_failed = false;    // the capture for `key` failed
  // Captures made, captures that failed, graph launches (this lane).
  uint64_t graph_captures = 0, graph_failures = 0, graph_replays = 0;
};
Workspace &workspace(Lane &L) {
  if (!L.dec) L.dec = new Workspace();
  return *(Workspace *)L.dec;
}

// Lays out one frame's pictures from its packet in host memory (`host`, or
// the header `parsed` already). `dev_packet` is the device address of the
// packet; `dev_data` / `dev_alpha` the device output planes.
void plan_frame(const uint8_t *host, size_t len, uint64_t dev_packet, uint32_t want_depth,
                uint16_t *dev_data, uint16_t *dev_alpha, Batch &b, const Header *parsed) {
  FramePlan fp;
  Header hd;
  if (parsed) {
    hd = *parsed;
  } else {
    fp.early = parse_frame_header(host, len, hd);
    if (fp.early.ok()) fp.early = check_plausible_size(hd);
  }
  // The device writes through these: a job without them decodes nothing.
  if (fp.early.ok() && (!dev_packet || !dev_data || (hd.alpha_bits() && !dev_alpha)))
    fp.early = config_error(!dev_packet ? "a decode job has no device packet"
                            : !dev_data ? "a decode job has no device planes"
                                        : "a decode job has no alpha plane for a frame with alpha");
  if (!fp.early.ok()) {
    b.plans.push_back(fp);
    return;
  }
  const prores_frame_header &h = hd.h;
  DecFrame F;
  memset(&F, 0, sizeof F);
  const uint32_t bd = want_depth ? want_depth : (h.chroma == 2 ? 10 : 12);
  prores_plane planes[3];
  prores_frame_layout(h.width, h.height, h.chroma, planes);
  for (int i = 0; i < 3; i++) {
    F.out_plane[i] = (uint64_t)(dev_data + planes[i].offset);
    F.plane_w[i] = planes[i].width;
  }
  F.out_alpha = (uint64_t)dev_alpha;
  F.width = h.width;
  F.height = h.height;
  F.chroma = h.chroma;
  F.alpha_bits = hd.alpha_bits();
  F.bit_depth = bd;
  F.interlaced = h.interlace != 0;
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
  hd.luma_weights(F.luma_w);
  hd.chroma_weights(F.chroma_w);
  const uint32_t frame_index = (uint32_t)b.frames.size();
  b.frames.push_back(F);
  if (F.alpha_bits) b.any_alpha = true;

  fp.first_pic = (uint32_t)b.pics.size();
  const uint32_t wmb = (h.width + 15u) / 16u, cpm = h.chroma == 2 ? 2u : 4u;
  size_t pos = 8 + (size_t)h.header_size;
  for (uint32_t index = 0; index < hd.picture_count(); index++) {
    if (pos > h.frame_size) {
      fp.deferred = invalid("a picture starts past the end of the frame");
      break;
    }
    PicHeader ph;
    Err e = parse_picture_header(host + pos, h.frame_size - pos, ph);
    if (!e.ok()) {
      fp.deferred = e;
      break;
    }
    const uint32_t height = hd.picture_height(index), hmb = (height + 15u) / 16u;
    const std::vector<uint32_t> sizes = slice_sizes(wmb, ph.log2_slice_mbs);
    const uint64_t n = (uint64_t)sizes.size() * hmb;
    if ((uint64_t)ph.header_size + 2 * n > ph.picture_size) {
      fp.deferred = invalid("the slice table runs past picture_size");
      break;
    }
    DecPic P;
    memset(&P, 0, sizeof P);
    P.pic_off = dev_packet + pos;
    P.header_size = ph.header_size;
    P.picture_size = ph.picture_size;
    P.n_slices = (uint32_t)n;
    P.row_slices = (uint32_t)sizes.size();
    P.slice_base = (uint32_t)b.n_slices;
    P.layout_base = (uint32_t)b.rows.size();
    P.frame = frame_index;
    P.height = height;
    P.height_in_mb = hmb;
    hd.picture_rows(index, P.first_row, P.row_step);
    uint32_t mb_x = 0, prefix = 0;
    for (uint32_t m : sizes) {
      b.rows.push_back(RowSlot{(uint16_t)mb_x, (uint16_t)m, prefix});
      mb_x += m;
      prefix += m * (4 + 2 * cpm) * 64;
    }
    P.row_coefs = prefix;
    P.coef_base = b.n_coefs;
    b.n_coefs += (uint64_t)prefix * hmb;
    b.n_slices += n;
    b.pics.push_back(P);
    fp.n_pics++;
    pos += ph.picture_size;
  }
  b.plans.push_back(fp);
}

// Workgroups of kernel `f` that fit on the device at once (persistent grids).
unsigned resident_grid(Device &d, hipFunction_t f, unsigned block, size_t lds = 0) {
  int per_cu = 0, cus = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, (int)block, lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d.id) != hipSuccess || cus < 1) cus = 1;
  return (unsigned)(per_cu * cus);
}

// Threads per entropy workgroup. A set value is used as it is; 0 picks the
// size from the batch: enough waves a workgroup that one workgroup per CU
// holds the batch's waves (up to ENTROPY_MAX_WG threads). Every workgroup
// copies the decode tables into its LDS once, so fewer, larger workgroups
// leave LDS for more waves: at 12 waves a CU one workgroup holds them all,
// where 256-thread workgroups fit only three to a CU.
unsigned entropy_wg(Device &d, const prores_decode_tuning &t, uint32_t items) {
  if (t.entropy_wg) return t.entropy_wg;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d.id) != hipSuccess || cus < 1) cus = 1;
  const uint32_t per_cu = (items + (uint32_t)cus - 1) / (uint32_t)cus;
  for (unsigned wg : {64u, 128u, 256u, 512u})
    if (per_cu <= wg / 64) return wg;
  return ENTROPY_MAX_WG;
}

// IDCT row segments: runs of consecutive slices of a row holding at most
// `cap` blocks (a larger slice is a segment of its own) and 16 slices. Every
// row of a picture is cut the same way; returns the batch's item count.
uint64_t plan_segments(Batch &b, uint32_t cap, std::vector<IdctItem> &segs) {
  segs.clear();
  uint64_t items = 0;
  for (DecPic &P : b.pics) {
    const uint32_t cpm = b.frames[P.frame].chroma == 2 ? 2u : 4u;
    P.seg_base = (uint32_t)segs.size();
    P.item_base = (uint32_t)items;
    IdctItem cur{0, 0, 0, 0};
    for (uint32_t j = 0; j < P.row_slices; j++) {
      const uint32_t tb = b.rows[P.layout_base + j].mbs * (4 + 2 * cpm);
      if (cur.count && (cur.blocks + tb > cap || cur.count == IDCT_MAX_SLICES)) {
        segs.push_back(cur);
        cur = IdctItem{j, 0, 0, 0};
      }
      cur.count++;
      cur.blocks += tb;
    }
    segs.push_back(cur);
    P.n_segs = (uint32_t)segs.size() - P.seg_base;
    items += (uint64_t)P.height_in_mb * P.n_segs;
  }
  return items;
}

Err validate_tuning(const prores_decode_tuning &t) {
  if (t.persistent > 1 || t.graph > 1 || t.chunk < 1 || t.chunk > 64 || t.idct_blocks < 8 ||
      t.idct_blocks > IDCT_MAX_BLOCKS ||
      (t.entropy_wg != 0 && t.entropy_wg != 64 && t.entropy_wg != 128 && t.entropy_wg != 256 &&
       t.entropy_wg != 512 && t.entropy_wg != ENTROPY_MAX_WG))
    return config_error("decode tuning out of range");
  return Err{};
}

// The device addresses and launch shapes of one batch's launch sequence:
// pd_plan first for a device-planned batch, then the slice table, the entropy
// order, the colour entropy decode, the alpha decode on the second stream
// and the IDCT.
struct Launch {
  void *B = nullptr, *frames = nullptr, *pics = nullptr, *rows = nullptr, *chunks = nullptr, *slices = nullptr;
  void *status = nullptr, *serr = nullptr, *ferr = nullptr, *segs = nullptr, *items = nullptr, *aord = nullptr;
  void *cost = nullptr, *tables = nullptr, *coefs = nullptr, *wide = nullptr, *order = nullptr;
  unsigned grids[4] = {0, 0, 0, 0};
  unsigned ewg = 64, idct_lds = 0;
  uint32_t pers = 0, n_slots = 0;
  bool alpha = false;
  // pd_plan's arguments (device_plan only).
  bool device_plan = false;
  void *jobs = nullptr, *scratch = nullptr, *recs = nullptr, *totals = nullptr;
  uint32_t n_jobs = 0, want_depth = 0, idct_cap = 0, chunk = 1;
  PlanCaps caps{};
};

Err enqueue_sequence(Device &d, Lane &L, Launch &A, bool timed) {
  Err x;
  const unsigned ewaves = A.ewg / 64;
  if (timed) (void)hipEventRecord(L.ev[0], L.stream);
  if (A.device_plan) {
    void *args[] = {&A.jobs, &A.n_jobs, &A.want_depth, &A.idct_cap, &A.chunk, &A.caps, &A.scratch, &A.recs,
                    &A.totals, &A.frames, &A.pics, &A.rows, &A.segs, &A.chunks, &A.cost, &A.B};
    if (!(x = launch(L, d.k_plan, 1, 1, 256, args)).ok()) return x;
  }
  {
    void *args[] = {&A.B, &A.frames, &A.pics, &A.rows, &A.chunks, &A.slices, &A.status, &A.serr, &A.ferr, &A.segs,
                    &A.items, &A.aord, &A.cost};
    if (!(x = launch(L, d.k_slice_table, A.grids[0], 1, ST_CHUNK, args)).ok()) return x;
  }
  if (A.order) {
    void *args[] = {&A.B, &A.cost, &A.order, &A.n_slots};
    if (!(x = launch(L, d.k_order, 1, 1, 1024, args)).ok()) return x;
  }
  if (timed && !(x = check(hipEventRecord(L.ev[1], L.stream), "hipEventRecord")).ok()) return x;
  // Alpha needs only the slice table: the second stream waits for the work
  // so far and decodes alpha beside the colour entropy decode and the IDCT,
  // which leave most of the machine's wave slots free. The entropy decode is
  // launched first and its workgroups (most of a CU's LDS each) are placed
  // before the alpha workgroups.
  bool forked = false, joined = false;
  if (A.alpha) {
    // A timed batch forks at its stage event: one marker less between the
    // order and entropy launches.
    hipEvent_t fork = timed ? L.ev[1] : L.fork;
    if (!timed && !(x = check(hipEventRecord(fork, L.stream), "hipEventRecord")).ok()) return x;
    if (!(x = check(hipStreamWaitEvent(L.aux, fork, 0), "hipStreamWaitEvent")).ok()) return x;
    forked = true;
  }
  {
    void *args[] = {&A.tables, &A.B, &A.frames, &A.pics, &A.slices, &A.status, &A.coefs, &A.wide, &A.serr, &A.ferr,
                    &A.order, &A.pers};
    x = launch(L, d.k_entropy, A.grids[1], 1, A.ewg, args, ewaves * ENTROPY_LDS_PER_WAVE);
  }
  if (x.ok() && forked) {
    void *args[] = {&A.B, &A.frames, &A.pics, &A.slices, &A.status, &A.serr, &A.ferr, &A.aord, &A.pers};
    const Err a = launch_on(L.aux, d.k_alpha, A.grids[2], 1, 64, args);
    if (timed) (void)hipEventRecord(L.ev[5], L.aux);
    Err j = check(hipEventRecord(L.join, L.aux), "hipEventRecord");
    if (!j.ok()) return j;  // the stream joins nothing: fail before more work
    joined = true;
    x = a;
  }
  if (x.ok()) {
    if (timed) (void)hipEventRecord(L.ev[2], L.stream);
    void *args[] = {&A.tables, &A.B, &A.frames, &A.pics, &A.slices, &A.status, &A.serr, &A.coefs, &A.wide, &A.items,
                    &A.pers};
    x = launch(L, d.k_idct, A.grids[3], 1, 256, args, A.idct_lds);
    if (timed) (void)hipEventRecord(L.ev[3], L.stream);
  }
  // Whatever happened, the main stream waits for the alpha work before
  // anything else on it can reuse the workspace.
  if (joined) {
    Err j = check(hipStreamWaitEvent(L.stream, L.join, 0), "hipStreamWaitEvent");
    if (x.ok()) x = j;
  }
  return x;
}

// Launches the sequence (replaying a captured graph when the tuning asks
// for one), copies `res_bytes` of results from `res` to `host`, waits, and
// records the stage times.
Err run_sequence(Device &d, Lane &L, Launch &A, const prores_decode_tuning &t, const void *res, size_t res_bytes,
                 void *host) {
  Workspace &ws = workspace(L);
  Err e;
  // A capture that fails (another thread's work invalidated it, or the
  // runtime refused a node) leaves the stream out of capture mode with
  // nothing launched: the batch then runs as stream launches.
  bool replay = t.graph != 0;
  if (replay) {
    GraphKey key;
    memset(&key, 0, sizeof key);
    void *ptrs[22] = {A.B,    A.pics,  A.rows,  A.slices, A.status, A.frames, A.tables, A.coefs,
                      A.wide, A.serr,  A.ferr,  A.items,  A.segs,   A.aord,   A.cost,   A.order,
                      A.chunks, A.jobs, A.scratch, A.recs, A.totals, nullptr};
    memcpy(key.ptrs, ptrs, sizeof ptrs);
    memcpy(key.grids, A.grids, sizeof A.grids);
    key.ewg = A.ewg;
    key.idct_lds = A.idct_lds;
    key.n_jobs = A.n_jobs;
    key.idct_cap = A.idct_cap;
    key.t = t;
    key.caps = A.caps;
    key.alpha = A.alpha;
    key.device_plan = A.device_plan;
    // Captured once per key; a key whose capture failed runs as stream
    // launches without trying again.
    if (!(key == ws.key) || (!ws.graph && !ws.graph_failed)) {
      if (ws.graph) {
        (void)hipGraphExecDestroy(ws.graph);
        ws.graph = nullptr;
      }
      hipGraph_t graph = nullptr;
      bool captured = false;
      if (hipStreamBeginCapture(L.stream, hipStreamCaptureModeRelaxed) == hipSuccess) {
        const Err ce = enqueue_sequence(d, L, A, false);
        const hipError_t ee = hipStreamEndCapture(L.stream, &graph);
        captured = ce.ok() && ee == hipSuccess && graph;
      }
      if (captured && hipGraphInstantiate(&ws.graph, graph, nullptr, nullptr, 0) != hipSuccess) {
        ws.graph = nullptr;
        captured = false;
      }
      if (graph) (void)hipGraphDestroy(graph);
      if (!captured) (void)hipGetLastError();  // the failed capture's error is not this batch's
      ws.key = key;
      ws.graph_failed = !captured;
      ws.graph_captures += captured;
      ws.graph_failures += !captured;
    }
    replay = ws.graph != nullptr;
  }
  if (replay) {
    (void)hipEventRecord(L.ev[0], L.stream);
    if (!(e = check(hipGraphLaunch(ws.graph, L.stream), "hipGraphLaunch")).ok()) return e;
    ws.graph_replays++;
    (void)hipEventRecord(L.ev[4], L.stream);
  } else {
    if (!(e = enqueue_sequence(d, L, A, true)).ok()) return e;
    (void)hipEventRecord(L.ev[4], L.stream);
  }
  if (res_bytes && !(e = check(hipMemcpyAsync(host, res, res_bytes, hipMemcpyDeviceToHost, L.stream),
                               "hipMemcpyAsync(results)")).ok())
    return e;
  if (!(e = check(hipStreamSynchronize(L.stream), "hipStreamSynchronize")).ok()) return e;
  // Stages: slice table, colour entropy, alpha (from the end of the slice
  // table to its own end, beside the other two), IDCT, whole batch.
  float ms[4] = {0, 0, 0, 0}, all = 0;
  if (!replay) {
    (void)hipEventElapsedTime(&ms[0], L.ev[0], L.ev[1]);
    (void)hipEventElapsedTime(&ms[1], L.ev[1], L.ev[2]);
    if (A.alpha) (void)hipEventElapsedTime(&ms[2], L.ev[1], L.ev[5]);
    (void)hipEventElapsedTime(&ms[3], L.ev[2], L.ev[3]);
  }
  (void)hipEventElapsedTime(&all, L.ev[0], L.ev[4]);
  for (int i = 0; i < 4; i++) L.dec_us[i] = ms[i] * 1000.0;
  L.dec_us[4] = all * 1000.0;
  L.dec_seq = ++d.batches;
  return Err{};
}

// Grids, workgroup size and LDS of a launch for batches of up to `slices`
// slices, `chunks` slice-table chunks and `idct_items` IDCT items.
void launch_shape(Device &d, const prores_decode_tuning &t, uint64_t slices, uint64_t chunks, uint64_t idct_items,
                  uint32_t max_count, Launch &A) {
  const uint32_t entropy_items = (uint32_t)(3 * ((slices + 63) / 64));
  A.idct_lds = max_count * 128 * (unsigned)sizeof(float);
  A.ewg = entropy_wg(d, t, entropy_items);
  const unsigned ewaves = A.ewg / 64;
  A.pers = t.persistent;
  if (t.persistent) {
    A.grids[0] = resident_grid(d, d.k_slice_table, ST_CHUNK);
    A.grids[1] = resident_grid(d, d.k_entropy, A.ewg, ewaves * ENTROPY_LDS_PER_WAVE);
    A.grids[2] = resident_grid(d, d.k_alpha, 64);
    A.grids[3] = resident_grid(d, d.k_idct, 256, A.idct_lds);
  } else {
    A.grids[0] = (unsigned)chunks;
    A.grids[1] = (unsigned)((entropy_items + ewaves - 1) / ewaves);
    // Workgroups of several waves (one a CU) fill every CU: the slots past
    // the items are empty, and the order deals them to the SIMDs that hold
    // the longest items.
    int cus = 0;
    if (ewaves >= 4 && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d.id) == hipSuccess &&
        A.grids[1] < (unsigned)cus && (uint64_t)cus * ewaves <= ORDER_MAX)
      A.grids[1] = (unsigned)cus;
    A.grids[2] = (unsigned)slices;
    A.grids[3] = (unsigned)idct_items;
  }
  // Entropy items longest first (pd_entropy_order) when the sort fits: the
  // grid's slots, or the queue's items (0: the kernel reads the count from
  // the batch, so a captured graph replays for any batch of its shape).
  A.n_slots = t.persistent ? 0u : A.grids[1] * ewaves;
  A.order = nullptr;
}

// The picture results of a batch: the first error of each frame's good
// pictures in picture order (a slice table past picture_size, else the
// failing slice's first error word), then the frame's deferred error.
Err picture_result(Lane &L, Workspace &ws, const uint32_t *st, const uint32_t *fe, uint32_t first_pic,
                   uint32_t n_pics, const DecPic *host_pics, const void *dev_pics, Err deferred, Err &out) {
  Err r, e;
  for (uint32_t p = first_pic; p < first_pic + n_pics && r.ok(); p++) {
    if (st[p] == PIC_TABLE_ERR) {
      r = invalid("a slice runs past picture_size");
    } else if (fe[p] != 0xFFFFFFFFu) {
      DecPic P;
      if (host_pics) {
        P = host_pics[p];
      } else if (!(e = copy_on(L, &P, (const DecPic *)dev_pics + p, sizeof P, hipMemcpyDeviceToHost)).ok()) {
        return e;
      }
      uint32_t words[4];
      const uint64_t g = P.slice_base + (uint64_t)fe[p];
      if (!(e = copy_on(L, words, (uint8_t *)ws.slice_err.p + 16 * g, 16, hipMemcpyDeviceToHost)).ok()) return e;
      for (int c = 0; c < 4; c++)
        if (words[c]) {
          r = slice_error(words[c]);
          break;
        }
    }
  }
  out = r.ok() ? deferred : r;
  return Err{};
}

// Runs the kernels over a batch the host planned and turns the device's
// error records into one Result per frame (`results`, in frame order).
Err run_batch(Device &d, Lane &L, Batch &b, const prores_decode_tuning &t, std::vector<Err> &results) {
  Workspace &ws = workspace(L);
  Err e;
  if (!(e = validate_tuning(t)).ok()) return e;
  const uint64_t G = b.n_slices;
  const size_t np = b.pics.size();
  if (G > 0xFFFFFFF0ull) return invalid("a batch of more than 2^32 slices");
  std::vector<IdctItem> segs;
  const uint64_t n_items = plan_segments(b, t.idct_blocks, segs);
  // pd_idct's dynamic LDS: 128 dequantisation factors a slice of its
  // largest item.
  uint32_t max_count = 1;
  for (const IdctItem &it : segs) max_count = it.count > max_count ? it.count : max_count;
  DecBatch hdr;
  memset(&hdr, 0, sizeof hdr);
  hdr.n_pics = (uint32_t)np;
  hdr.n_slices = (uint32_t)G;
  hdr.n_entropy_items = (uint32_t)(3 * ((G + 63) / 64));
  hdr.n_idct_items = (uint32_t)n_items;
  hdr.chunk = t.persistent ? t.chunk : 1;
  // pd_slice_table's chunks: (picture, first slice), ST_CHUNK slices each,
  // at least one a picture.
  std::vector<uint32_t> chunks;
  for (size_t p = 0; p < np; p++) {
    const uint32_t n = b.pics[p].n_slices;
    for (uint32_t k = 0; k == 0 || k < n; k += ST_CHUNK) {
      chunks.push_back((uint32_t)p);
      chunks.push_back(k);
    }
  }
  if (chunks.size() / 2 > 0xFFFFFFFFull) return invalid("a batch of more than 2^32 slice table chunks");
  hdr.n_chunks = (uint32_t)(chunks.size() / 2);
  // The plan goes up in one copy from page-locked memory.
  auto fit = [](size_t &cap, size_t need) {
    if (need > cap) cap = (need + need / 4 + 4095) & ~(size_t)4095;
  };
  fit(ws.cap_frames, b.frames.size() * sizeof(DecFrame));
  fit(ws.cap_pics, b.pics.size() * sizeof(DecPic));
  fit(ws.cap_rows, b.rows.size() * sizeof(RowSlot));
  fit(ws.cap_segs, segs.size() * sizeof(IdctItem));
  fit(ws.cap_chunks, chunks.size() * sizeof(uint32_t));
  // Entropy item costs, one word an item (three a group of 64 slices),
  // zero at upload.
  const size_t cost_bytes = (size_t)((G + 63) / 64) * 12;
  fit(ws.cap_cost, cost_bytes);
  const size_t o_pics = ws.cap_frames, o_rows = o_pics + ws.cap_pics, o_segs = o_rows + ws.cap_rows,
               o_chunks = o_segs + ws.cap_segs, o_cost = o_chunks + ws.cap_chunks, o_hdr = o_cost + ws.cap_cost,
               plan_bytes = o_hdr + sizeof(DecBatch);
  if (!(e = ws.plan.reserve(plan_bytes)).ok()) return e;
  if (!(e = ws.hplan.reserve(plan_bytes)).ok()) return e;
  {
    uint8_t *h = (uint8_t *)ws.hplan.p;
    if (!b.frames.empty()) memcpy(h, b.frames.data(), b.frames.size() * sizeof(DecFrame));
    if (!b.pics.empty()) memcpy(h + o_pics, b.pics.data(), b.pics.size() * sizeof(DecPic));
    if (!b.rows.empty()) memcpy(h + o_rows, b.rows.data(), b.rows.size() * sizeof(RowSlot));
    if (!segs.empty()) memcpy(h + o_segs, segs.data(), segs.size() * sizeof(IdctItem));
    if (!chunks.empty()) memcpy(h + o_chunks, chunks.data(), chunks.size() * sizeof(uint32_t));
    memset(h + o_cost, 0, cost_bytes);
    memcpy(h + o_hdr, &hdr, sizeof hdr);
    if (!(e = check(hipMemcpyAsync(ws.plan.p, ws.hplan.p, plan_bytes, hipMemcpyHostToDevice, L.stream),
                    "hipMemcpyAsync(plan)")).ok())
      return e;
  }
  fit(ws.cap_np, np * 4);
  if (!(e = ws.picres.reserve(2 * ws.cap_np)).ok()) return e;
  if (!(e = ws.items.reserve(n_items * sizeof(IdctItem) + 64)).ok()) return e;
  if (!(e = ws.slices.reserve(G * sizeof(SliceRec) + 64)).ok()) return e;
  if (!(e = ws.slice_err.reserve(G * 16 + 64)).ok()) return e;
  if (!(e = ws.coefs.reserve(b.n_coefs * 2 + 64)).ok()) return e;
  if (!(e = ws.wide.reserve(b.n_coefs * 4 + 64)).ok()) return e;
  if (!(e = ws.alpha_order.reserve(G * sizeof(AlphaRec) + 64)).ok()) return e;

  Launch A;
  uint8_t *const pl = (uint8_t *)ws.plan.p;
  A.B = pl + o_hdr;
  A.frames = pl;
  A.pics = pl + o_pics;
  A.rows = pl + o_rows;
  A.segs = pl + o_segs;
  A.chunks = pl + o_chunks;
  A.cost = pl + o_cost;
  A.slices = ws.slices.p;
  A.status = ws.picres.p;
  A.ferr = (uint8_t *)ws.picres.p + ws.cap_np;
  A.serr = ws.slice_err.p;
  A.items = ws.items.p;
  A.aord = ws.alpha_order.p;
  A.tables = d.tables;
  A.coefs = ws.coefs.p;
  A.wide = ws.wide.p;
  A.alpha = b.any_alpha;
  launch_shape(d, t, G, hdr.n_chunks, n_items, max_count, A);
  const uint32_t slots = t.persistent ? hdr.n_entropy_items : A.n_slots;
  if (slots <= ORDER_MAX) {
    if (!(e = ws.order.reserve((size_t)ORDER_MAX * 4 + 64)).ok()) return e;
    A.order = ws.order.p;
  }
  if (np == 0) {
    L.dec_us[0] = L.dec_us[1] = L.dec_us[2] = L.dec_us[3] = L.dec_us[4] = 0;
    L.dec_seq = ++d.batches;
  } else if (!(e = ws.hres.reserve(2 * ws.cap_np + 64)).ok() ||
             !(e = run_sequence(d, L, A, t, ws.picres.p, 2 * ws.cap_np, ws.hres.p)).ok()) {
    return e;
  }
  const uint32_t *st = (const uint32_t *)ws.hres.p, *fe = st + ws.cap_np / 4;
  results.assign(b.plans.size(), Err{});
  for (size_t f = 0; f < b.plans.size(); f++) {
    const FramePlan &fp = b.plans[f];
    if (!fp.early.ok()) {
      results[f] = fp.early;
      continue;
    }
    if (!(e = picture_result(L, ws, st, fe, fp.first_pic, fp.n_pics, b.pics.data(), nullptr, fp.deferred,
                             results[f])).ok())
      return e;
  }
  return Err{};
}

// The reference's message for a pd_plan error (the host checks' text,
// host_common.cpp parse_frame_header / check_plausible_size /
// parse_picture_header, and plan_frame).
Err plan_error(uint32_t code, const uint64_t *a, uint32_t darg) {
  switch (code) {
    case F_SHORT: return invalid(fmtstr("%llu bytes is too short for a ProRes frame header", (unsigned long long)a[0]));
    case F_ICPF: return invalid("no 'icpf' frame identifier");
    case F_SIZE_EXCEEDS:
      return invalid(fmtstr("frame_size %llu exceeds the %llu bytes given", (unsigned long long)a[0],
                            (unsigned long long)a[1]));
    case F_SIZE_SMALL: return invalid("frame_size leaves no room for the frame header");
    case F_HDR_SIZE: return invalid(fmtstr("frame_header_size %llu is out of range", (unsigned long long)a[0]));
    case F_VERSION: return unsupported(fmtstr("bitstream_version %llu", (unsigned long long)a[0]));
    case F_DIMS:
      return invalid(fmtstr("a %llu%s%llu frame", (unsigned long long)a[0], TIMES, (unsigned long long)a[1]));
    case F_CHROMA: return invalid(fmtstr("chroma_format %llu is reserved", (unsigned long long)a[0]));
    case F_INTERLACE: return invalid("interlace_mode 3 is reserved");
    case F_ALPHA: return invalid(fmtstr("alpha_channel_type %llu is reserved", (unsigned long long)a[0]));
    case F_MATRIX: return invalid("a quantisation matrix runs past frame_header_size");
    case F_IMPLAUSIBLE:
      return invalid(fmtstr("frame_size %llu is too small for a %llu%s%llu frame (at least %llu bytes)",
                            (unsigned long long)a[0], (unsigned long long)a[1], TIMES, (unsigned long long)a[2],
                            (unsigned long long)a[3]));
    case F_NO_PACKET: return config_error("a decode job has no device packet");
    case F_NO_PLANES: return config_error("a decode job has no device planes");
    case F_NO_ALPHA: return config_error("a decode job has no alpha plane for a frame with alpha");
    case P_START_PAST: return invalid("a picture starts past the end of the frame");
    case P_HDR_SHORT: return invalid("a picture header runs past the end of the frame");
    case P_HDR_SIZE: return invalid(fmtstr("picture_header_size %u is out of range", darg));
    case P_SIZE: return invalid(fmtstr("picture_size %u is out of range", darg));
    case P_TABLE: return invalid("the slice table runs past picture_size");
    default: return Err{};
  }
}

// A batch of jobs planned on the device (pd_plan): the jobs go up, the
// launch sequence runs with no host round trip before it, and the results
// (each job's record, the batch totals, the picture status words) come back
// in one copy. The workspace is laid out for the capacities of the largest
// batch seen on the lane; a batch over them does no work and runs again over
// a workspace grown to its totals. With `fixed`, the workspace is carved from
// that region instead: nothing is allocated, the first capacities split the
// region, and a batch whose totals do not fit it fails.
Err run_device_batch(Device &d, Lane &L, uint32_t bit_depth, size_t n, const prores_dev_decode_job *jobs,
                     const prores_decode_tuning &t, std::vector<Err> &results, FixedRegion *fixed = nullptr) {
  Workspace &ws = workspace(L);
  Err e;
  if (!(e = validate_tuning(t)).ok()) return e;
  if (n > (1u << 24)) return invalid("a batch of more than 2^24 frames");
  results.assign(n, Err{});
  if (!n) return Err{};
  // Device buffers: the lane's (grown as needed) or carved from the region.
  auto need = [&](DevBuf &b, size_t bytes) -> Err {
    if (!fixed) return b.reserve(bytes);
    if (fixed->take(b, bytes)) return Err{};
    return Err{PRORES_ERR_CONFIG, fmtstr("the batch needs more than the %zu bytes of its scratch memory",
                                         fixed->dev_bytes)};
  };
  const size_t job_bytes = n * sizeof(DevJob);
  if (fixed) {
    fixed->dev_used = 0;
    if (job_bytes > fixed->host_bytes) return config_error("too many jobs for the scratch memory's host staging");
  } else if (!(e = ws.hplan.reserve(job_bytes + 64)).ok()) {
    return e;
  }
  if (!(e = need(ws.jobs, job_bytes + 64)).ok()) return e;
  const size_t jobs_used = fixed ? fixed->dev_used : 0;
  // Jobs up, from the page-locked plan staging (or the region's host part).
  {
    DevJob *hj = (DevJob *)(fixed ? (void *)fixed->host : ws.hplan.p);
    for (size_t i = 0; i < n; i++)
      hj[i] = DevJob{(uint64_t)jobs[i].dev_packet, jobs[i].len, (uint64_t)jobs[i].dev_data, (uint64_t)jobs[i].dev_alpha};
    if (!(e = check(hipMemcpyAsync(ws.jobs.p, hj, job_bytes, hipMemcpyHostToDevice, L.stream),
                    "hipMemcpyAsync(jobs)")).ok())
      return e;
  }
  // First capacities: a 4:4:4 frame of 2160p's slices at 8 macroblocks a
  // job; the totals of a batch over them set the next. In a fixed region
  // the first capacities fill it: a slice's records and its share of
  // coefficients (1,024 a slice: one 4:2:2 macroblock) from what is left
  // after the per-job parts.
  PlanCaps &c = ws.dcap;
  const uint32_t np = (uint32_t)(2 * n);
  if (fixed) {
    const size_t per_job = 2 * sizeof(DecPic) + sizeof(DecFrame) + sizeof(PlanJob) + sizeof(PlanRec) + 64 * 8 +
                           4 * 256 + 16;
    const size_t fixed_part = n * per_job + (size_t)ORDER_MAX * 4 + 16 * 4096;
    const size_t per_slice = sizeof(SliceRec) + 16 + sizeof(AlphaRec) + 2 * sizeof(IdctItem) + sizeof(RowSlot) + 8 + 1 +
                             1024 * 6;
    const size_t room = fixed->dev_bytes > fixed_part + jobs_used ? fixed->dev_bytes - fixed_part - jobs_used : 0;
    c = PlanCaps{};
    c.slices = (uint32_t)std::min<size_t>(room / per_slice, 0xFFFFFF00u);
    c.coefs = (uint64_t)c.slices * 1024;
    c.rows = c.segs = c.items = c.slices;
    c.chunks = c.slices / ST_CHUNK + np;
  } else if (!c.slices) {
    c.slices = (uint32_t)std::min<uint64_t>(4096ull * n, 0xFFFFFF00u);
    c.coefs = 32400ull * 768 * n;
  }
  PlanTotals T;
  for (int attempt = 0;; attempt++) {
    c.pics = np;
    if (fixed) {
      fixed->dev_used = jobs_used;
    } else {
      c.rows = std::max<uint32_t>(c.rows, 64 * np);
      c.segs = std::max<uint32_t>(c.segs, c.rows);
      c.chunks = std::max<uint32_t>(c.chunks, c.slices / ST_CHUNK + np);
      c.items = std::max<uint32_t>(c.items, c.slices);
    }
    c.item_slices = ws.idct_cap_slices;
    // The plan buffer: frames (one a job), pictures, row slots, segments,
    // chunks, item costs, header.
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_pics = al(n * sizeof(DecFrame)), o_rows = o_pics + al((size_t)c.pics * sizeof(DecPic)),
                 o_segs = o_rows + al((size_t)c.rows * sizeof(RowSlot)),
                 o_chunks = o_segs + al((size_t)c.segs * sizeof(IdctItem)),
                 o_cost = o_chunks + al((size_t)c.chunks * 8), o_hdr = o_cost + al(((size_t)c.slices + 63) / 64 * 12),
                 plan_bytes = o_hdr + sizeof(DecBatch);
    // Results: job records, totals, picture status, first errors.
    const size_t r_tot = al(n * sizeof(PlanRec)), r_st = r_tot + al(sizeof(PlanTotals)), r_fe = r_st + al(np * 4),
                 res_bytes = r_fe + np * 4;
    if (fixed && res_bytes > fixed->host_bytes) return config_error("too many jobs for the scratch memory's host staging");
    if (!(e = need(ws.plan, plan_bytes)).ok() || !(e = need(ws.dres, res_bytes + 64)).ok() ||
        !(e = need(ws.scratch, n * sizeof(PlanJob) + 64)).ok() ||
        !(e = need(ws.items, (size_t)c.items * sizeof(IdctItem) + 64)).ok() ||
        !(e = need(ws.slices, (size_t)c.slices * sizeof(SliceRec) + 64)).ok() ||
        !(e = need(ws.slice_err, (size_t)c.slices * 16 + 64)).ok() || !(e = need(ws.coefs, c.coefs * 2 + 64)).ok() ||
        !(e = need(ws.wide, c.coefs * 4 + 64)).ok() ||
        !(e = need(ws.alpha_order, (size_t)c.slices * sizeof(AlphaRec) + 64)).ok())
      return e;
    if (!fixed && !(e = ws.hres.reserve(res_bytes + 64)).ok()) return e;
    void *const hres = fixed ? (void *)fixed->host : ws.hres.p;
    Launch A;
    uint8_t *const pl = (uint8_t *)ws.plan.p, *rs = (uint8_t *)ws.dres.p;
    A.device_plan = true;
    A.jobs = ws.jobs.p;
    A.n_jobs = (uint32_t)n;
    A.want_depth = bit_depth;
    A.idct_cap = t.idct_blocks;
    A.chunk = t.persistent ? t.chunk : 1;
    A.caps = c;
    A.scratch = ws.scratch.p;
    A.recs = rs;
    A.totals = rs + r_tot;
    A.B = pl + o_hdr;
    A.frames = pl;
    A.pics = pl + o_pics;
    A.rows = pl + o_rows;
    A.segs = pl + o_segs;
    A.chunks = pl + o_chunks;
    A.cost = pl + o_cost;
    A.slices = ws.slices.p;
    A.status = rs + r_st;
    A.ferr = rs + r_fe;
    A.serr = ws.slice_err.p;
    A.items = ws.items.p;
    A.aord = ws.alpha_order.p;
    A.tables = d.tables;
    A.coefs = ws.coefs.p;
    A.wide = ws.wide.p;
    A.alpha = true;  // whether any frame has alpha is known only on the device
    launch_shape(d, t, c.slices, c.chunks, c.items, c.item_slices, A);
    const uint32_t slots = t.persistent ? 3 * ((c.slices + 63) / 64) : A.n_slots;
    if (slots <= ORDER_MAX) {
      if (!(e = need(ws.order, (size_t)ORDER_MAX * 4 + 64)).ok()) return e;
      A.order = ws.order.p;
    }
    if (!(e = run_sequence(d, L, A, t, ws.dres.p, res_bytes, hres)).ok()) return e;
    const uint8_t *h = (const uint8_t *)hres;
    memcpy(&T, h + r_tot, sizeof T);
    // The next batches' IDCT items may hold as many slices as this one's
    // wanted (the dynamic LDS is sized for it).
    ws.idct_cap_slices = std::max<uint32_t>(std::min<uint32_t>(T.wanted_count, IDCT_MAX_SLICES), 1);
    if (T.overflow) {
      if (attempt) return Err{PRORES_ERR_DEVICE, "a device-planned batch outgrew the workspace sized for it"};
      if (fixed) {
        // The batch's own totals, if they fit the region.
        c.slices = T.slices;
        c.rows = T.rows;
        c.segs = T.segs;
        c.chunks = T.chunks;
        c.items = T.items;
        c.coefs = T.coefs;
        continue;
      }
      if (T.slices > 0xFFFFFFF0u || T.chunks > 0xFFFFFFF0u) return invalid("a batch of more than 2^32 slices");
      auto grow = [](uint32_t v) { return (uint32_t)std::min<uint64_t>((uint64_t)v + v / 4 + 64, 0xFFFFFFF0u); };
      c.slices = std::max(c.slices, grow(T.slices));
      c.rows = std::max(c.rows, grow(T.rows));
      c.segs = std::max(c.segs, grow(T.segs));
      c.chunks = std::max(c.chunks, grow(T.chunks));
      c.items = std::max(c.items, grow(T.items));
      c.coefs = std::max<uint64_t>(c.coefs, T.coefs + T.coefs / 4 + 4096);
      continue;
    }
    const PlanRec *R = (const PlanRec *)h;
    const uint32_t *st = (const uint32_t *)(h + r_st), *fe = (const uint32_t *)(h + r_fe);
    for (size_t i = 0; i < n; i++) {
      if (R[i].code != F_NONE) {
        results[i] = plan_error(R[i].code, R[i].arg, 0);
        continue;
      }
      if (!(e = picture_result(L, ws, st, fe, R[i].first_pic, R[i].n_pics, nullptr, A.pics,
                               plan_error(R[i].deferred, R[i].arg, R[i].darg), results[i])).ok())
        return e;
    }
    return Err{};
  }
}

// Host memory of the frames the library allocates: page-locked
// (hipHostMalloc), so device-to-host copies run at the link's speed with no
// staging, and kept for reuse after prores_frame_free (a fresh allocation of
// a frame's size costs page faults, or pinning, every time). Up to KEEP
// bytes wait for reuse; beyond that buffers go back to the system. Never
// destroyed: frames may be freed while the process exits.
class PinnedPool {
  std::mutex mu;
  std::multimap<size_t, void *> idle;  // capacity -> buffer
  std::map<void *, size_t> cap;        // every pinned buffer handed out or idle
  size_t idle_bytes = 0;

 public:
  static constexpr size_t KEEP = 2ull << 30;
  // A buffer of at least n bytes, or nullptr when none can be pinned.
  void *get(size_t n) {
    const size_t c = (n + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = idle.lower_bound(c);
      if (it != idle.end() && it->first <= c + c / 4) {
        void *p = it->second;
        idle_bytes -= it->first;
        idle.erase(it);
        return p;
      }
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    cap[p] = c;
    return p;
  }
  // Takes back a buffer of this pool (true), or leaves p alone (false).
  bool put(void *p) {
    if (!p) return true;
    std::vector<void *> release;
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = cap.find(p);
      if (it == cap.end()) return false;
      idle.emplace(it->second, p);
      idle_bytes += it->second;
      while (idle_bytes > KEEP) {
        auto big = std::prev(idle.end());
        idle_bytes -= big->first;
        cap.erase(big->second);
        release.push_back(big->second);
        idle.erase(big);
      }
    }
    for (void *q : release) (void)hipHostFree(q);
    return true;
  }
};
PinnedPool &pinned() {
  static PinnedPool *p = new PinnedPool();
  return *p;
}

}  // namespace

void *host_staging_get(size_t bytes) {
  void *p = pinned().get(bytes);
  return p ? p : malloc(bytes ? bytes : 1);
}

void host_staging_put(void *p) {
  if (!pinned().put(p)) free(p);
}

namespace {

// A decoded frame's fields from its header, and its planes allocated from
// the pinned pool (malloc when none can be pinned).
Err frame_outputs(const prores_decoder *d, const Header &hd, prores_frame *out, bool allocate, size_t &total,
                  size_t &alpha_n) {
  const prores_frame_header &h = hd.h;
  const uint32_t bd = d->bit_depth ? d->bit_depth : (h.chroma == 2 ? 10 : 12);
  prores_plane planes[3];
  total = prores_frame_layout(h.width, h.height, h.chroma, planes);
  alpha_n = h.alpha ? (size_t)h.width * h.height : 0;
  if (allocate) {
    memset(out, 0, sizeof *out);
    out->data = (uint16_t *)pinned().get(total * 2 + 2);
    if (!out->data) out->data = (uint16_t *)malloc(total * 2 + 2);
    if (alpha_n) {
      out->alpha = (uint16_t *)pinned().get(alpha_n * 2 + 2);
      if (!out->alpha) out->alpha = (uint16_t *)malloc(alpha_n * 2 + 2);
    }
    if (!out->data || (alpha_n && !out->alpha)) {
      prores_frame_free(out);
      return Err{PRORES_ERR_DEVICE, "host allocation failed"};
    }
  } else if (!out->data || out->data_len < total || (alpha_n && (!out->alpha || out->alpha_len < alpha_n))) {
    return config_error("the output buffers are smaller than the frame");
  }
  out->width = h.width;
  out->height = h.height;
  out->bit_depth = bd;
  out->chroma = h.chroma;
  out->interlace = h.interlace;
  out->metadata = h.metadata;
  out->data_len = total;
  out->alpha_len = alpha_n;
  for (int i = 0; i < 3; i++) out->planes[i] = planes[i];
  return Err{};
}

}  // namespace
}  // namespace prores_hip

using namespace prores_hip;

extern "C" {

size_t prores_frame_layout(uint32_t width, uint32_t height, uint32_t chroma, prores_plane *p) {
  // Frame::layout (src/frame.rs:247-266): chroma width ceil(w/2) for 4:2:2.
  const uint32_t cw = chroma == PRORES_CHROMA_422 ? (width + 1) / 2 : width;
  p[0] = prores_plane{0, width, height};
  p[1] = prores_plane{(uint64_t)width * height, cw, height};
  p[2] = prores_plane{p[1].offset + (uint64_t)cw * height, cw, height};
  return (size_t)(p[2].offset + (uint64_t)cw * height);
}

prores_status prores_decoder_create(uint32_t bit_depth, uint32_t threads, prores_decoder **out,
                                    prores_error *err) {
  *out = nullptr;
  if (bit_depth != 0 && (bit_depth < 8 || bit_depth > 16)) {
    // Decoder::with_bit_depth (src/decode.rs:37-47)
    fill_error(err, config_error(fmtstr("output bit depth %u is outside 8%s16", bit_depth, EN_DASH)));
    return PRORES_ERR_CONFIG;
  }
  *out = new prores_decoder{bit_depth, threads};
  fill_error(err, Err{});
  return PRORES_OK;
}

void prores_decoder_destroy(prores_decoder *d) { delete d; }

prores_status prores_frame_header_parse(const uint8_t *data, size_t len, prores_frame_header *out,
                                        prores_error *err) {
  Header h;
  Err e = parse_frame_header(data, len, h);
  if (e.ok()) *out = h.h;
  fill_error(err, e);
  return (prores_status)e.kind;
}

void prores_frame_free(prores_frame *f) {
  if (!f) return;
  if (!pinned().put(f->data)) free(f->data);
  if (!pinned().put(f->alpha)) free(f->alpha);
  f->data = nullptr;
  f->alpha = nullptr;
}

static prores_status decode_common(const prores_decoder *d, const uint8_t *data, size_t len,
                                   prores_frame *out, prores_error *err, bool allocate) {
  Header hd;
  Err e = parse_frame_header(data, len, hd);
  if (e.ok()) e = check_plausible_size(hd);
  size_t total = 0, alpha_n = 0;
  if (e.ok()) e = frame_outputs(d, hd, out, allocate, total, alpha_n);
  if (!e.ok()) {
    fill_error(err, e);
    return (prores_status)e.kind;
  }
  Device *dev = nullptr;
  Lane *lane0 = nullptr;
  if (!(e = device(dev)).ok() || !(e = lane(*dev, 0, lane0)).ok()) {
    if (allocate) prores_frame_free(out);
    fill_error(err, e);
    return (prores_status)e.kind;
  }
  std::lock_guard<std::mutex> g(lane0->mu);
  Workspace &ws = workspace(*lane0);
  const size_t fsz = hd.h.frame_size;
  const size_t plane_bytes = (total * 2 + 255) & ~(size_t)255;
  if ((e = ws.packet.reserve(fsz + 128)).ok() && (e = ws.planes.reserve(plane_bytes + alpha_n * 2 + 64)).ok() &&
      (e = check(hipMemcpyAsync(ws.packet.p, data, fsz, hipMemcpyHostToDevice, lane0->stream), "H2D packet")).ok()) {
    Batch b;
    uint16_t *dd = ws.planes.as<uint16_t>();
    uint16_t *da = alpha_n ? (uint16_t *)((uint8_t *)ws.planes.p + plane_bytes) : nullptr;
    plan_frame(data, len, (uint64_t)ws.packet.p, d->bit_depth, dd, da, b, &hd);
    std::vector<Err> results;
    prores_decode_tuning t;
    prores_decode_tuning_default(&t);
    e = run_batch(*dev, *lane0, b, t, results);
    if (e.ok()) e = results[0];
    if (e.ok()) {
      // Into pinned memory (prores_decode) this runs at the link's speed.
      e = check(hipMemcpyAsync(out->data, dd, total * 2, hipMemcpyDeviceToHost, lane0->stream), "D2H planes");
      if (e.ok() && alpha_n)
        e = check(hipMemcpyAsync(out->alpha, da, alpha_n * 2, hipMemcpyDeviceToHost, lane0->stream), "D2H alpha");
      Err s2 = check(hipStreamSynchronize(lane0->stream), "hipStreamSynchronize");
      if (e.ok()) e = s2;
    }
  }
  if (!e.ok() && allocate) prores_frame_free(out);
  fill_error(err, e);
  return (prores_status)e.kind;
}

prores_status prores_decode(const prores_decoder *d, const uint8_t *data, size_t len, prores_frame *out,
                            prores_error *err) {
  return decode_common(d, data, len, out, err, true);
}

prores_status prores_decode_into(const prores_decoder *d, const uint8_t *data, size_t len,
                                 prores_frame *out, prores_error *err) {
  return decode_common(d, data, len, out, err, false);
}

}  // extern "C"

namespace prores_hip {

// A batch is planned on the device when any job's packet is only in device
// memory (host_packet NULL); else the host plans it from the host copies.
bool device_planned(size_t n, const prores_dev_decode_job *jobs) {
  for (size_t i = 0; i < n; i++)
    if (!jobs[i].host_packet) return true;
  return false;
}

// One batch of device jobs either way, into one Result per job.
Err run_jobs(Device &d, Lane &L, uint32_t bit_depth, size_t n, const prores_dev_decode_job *jobs,
             const prores_decode_tuning &t, std::vector<Err> &results) {
  if (device_planned(n, jobs)) return run_device_batch(d, L, bit_depth, n, jobs, t, results);
  Batch b;
  b.frames.reserve(n);
  b.pics.reserve(2 * n);
  b.plans.reserve(n);
  b.rows.reserve(2 * n * 64);
  for (size_t i = 0; i < n; i++) {
    const prores_dev_decode_job &j = jobs[i];
    plan_frame(j.host_packet, j.len, (uint64_t)j.dev_packet, bit_depth, j.dev_data, j.dev_alpha, b, nullptr);
  }
  return run_batch(d, L, b, t, results);
}

Err decode_batch_on(Device &d, Lane &L, uint32_t bit_depth, size_t n, const prores_dev_decode_job *jobs,
                    const prores_decode_tuning *tuning, prores_error *errs) {
  prores_decode_tuning t;
  if (tuning) {
    t = *tuning;
  } else {
    std::lock_guard<std::mutex> g(d.mu);
    t = d.tuning;
  }
  std::vector<Err> results;
  Err e = run_jobs(d, L, bit_depth, n, jobs, t, results);
  if (!e.ok()) {
    for (size_t i = 0; i < n; i++) fill_error(&errs[i], e);
    return e;
  }
  for (size_t i = 0; i < n; i++) fill_error(&errs[i], results[i]);
  return Err{};
}

void *decode_workspace(Lane &L) { return &workspace(L); }

Err decode_batch_fixed(Device &d, Lane &L, const prores_decoder *dec, size_t n, const prores_dev_decode_job *jobs,
                       FixedRegion &region, prores_error *errs) {
  prores_decode_tuning t;
  {
    std::lock_guard<std::mutex> g(d.mu);
    t = d.tuning;
  }
  t.graph = 0;  // instantiating a graph allocates
  std::vector<Err> results;
  Err e = run_device_batch(d, L, dec->bit_depth, n, jobs, t, results, &region);
  for (size_t i = 0; i < n; i++) fill_error(&errs[i], e.ok() ? results[i] : e);
  return e;
}

}  // namespace prores_hip

extern "C" {

static prores_status batch_common(const prores_decoder *d, size_t n, const prores_dev_decode_job *jobs,
                                  const prores_decode_tuning *tuning, prores_error *errs) {
  Device *dev = nullptr;
  Lane *ls[2] = {nullptr, nullptr};
  Err e = device(dev);
  for (int i = 0; i < 2 && e.ok(); i++) e = lane(*dev, CALL_LANES[i], ls[i]);
  if (!e.ok()) {
    for (size_t i = 0; i < n; i++) fill_error(&errs[i], e);
    return (prores_status)e.kind;
  }
  // The first free call lane; when both are busy, wait for the first.
  Lane *L = ls[0];
  std::unique_lock<std::mutex> g(ls[0]->mu, std::try_to_lock);
  if (!g.owns_lock()) {
    g = std::unique_lock<std::mutex>(ls[1]->mu, std::try_to_lock);
    if (g.owns_lock()) {
      L = ls[1];
    } else {
      g = std::unique_lock<std::mutex>(ls[0]->mu);
    }
  }
  e = decode_batch_on(*dev, *L, d->bit_depth, n, jobs, tuning, errs);
  return e.ok() ? PRORES_OK : (prores_status)e.kind;
}

// Host packets in, host frames out, many at a time: the frames are cut into
// chunks (up to HOST_CHUNK_BYTES of planes or HOST_CHUNK_FRAMES frames); two
// lanes, each driven by its own thread, take alternate chunks through
// upload, one batched launch sequence and download, so one chunk's copies
// overlap the other's kernels (and uploads overlap downloads: the link
// carries both directions at once).
static constexpr size_t HOST_CHUNK_BYTES = 256ull << 20, HOST_CHUNK_FRAMES = 16;
static constexpr int HOST_LANES[2] = {3, 4};

prores_status prores_decode_host_batch(const prores_decoder *d, size_t n, const uint8_t *const *data,
                                       const size_t *len, prores_frame *out, prores_error *errs) {
  struct Item {
    Header hd;
    size_t total = 0, alpha_n = 0;
    Err e;
  };
  std::vector<Item> items(n);
  for (size_t i = 0; i < n; i++) {
    Item &it = items[i];
    it.e = parse_frame_header(data[i], len[i], it.hd);
    if (it.e.ok()) it.e = check_plausible_size(it.hd);
    if (it.e.ok()) it.e = frame_outputs(d, it.hd, &out[i], true, it.total, it.alpha_n);
    if (!it.e.ok()) memset(&out[i], 0, sizeof out[i]);
  }
  std::vector<std::pair<size_t, size_t>> chunks;
  for (size_t i = 0; i < n;) {
    size_t j = i, bytes = 0;
    while (j < n && j - i < HOST_CHUNK_FRAMES && (j == i || bytes < HOST_CHUNK_BYTES)) {
      bytes += 2 * (items[j].total + items[j].alpha_n);
      j++;
    }
    chunks.push_back({i, j});
    i = j;
  }
  Device *dev = nullptr;
  Err de = device(dev);
  prores_decode_tuning t;
  prores_decode_tuning_default(&t);
  if (de.ok()) {
    std::lock_guard<std::mutex> g(dev->mu);
    t = dev->tuning;
  }
  // Stream launches: the two lanes' threads copy while the other launches,
  // and a replayed graph saves nothing next to the link's transfer time.
  t.graph = 0;
  auto work = [&](int which, Err &fail) {
    Lane *L = nullptr;
    if (!(fail = lane(*dev, HOST_LANES[which], L)).ok()) return;
    std::lock_guard<std::mutex> g(L->mu);
    Workspace &ws = workspace(*L);
    for (size_t c = (size_t)which; c < chunks.size(); c += 2) {
      const size_t a = chunks[c].first, b = chunks[c].second;
      // Arena offsets: each packet with 64 readable bytes past it, planes
      // then alpha, 256-byte aligned.
      std::vector<size_t> poff(b - a), qoff(b - a), aoff(b - a);
      size_t pbytes = 0, qbytes = 0;
      for (size_t i = a; i < b; i++) {
        if (!items[i].e.ok()) continue;
        poff[i - a] = pbytes;
        pbytes += (items[i].hd.h.frame_size + 128 + 255) & ~(size_t)255;
        qoff[i - a] = qbytes;
        qbytes += (items[i].total * 2 + 255) & ~(size_t)255;
        aoff[i - a] = qbytes;
        qbytes += (items[i].alpha_n * 2 + 255) & ~(size_t)255;
      }
      if (!pbytes) continue;
      Err e;
      if (!(e = ws.packet.reserve(pbytes + 64)).ok() || !(e = ws.planes.reserve(qbytes + 64)).ok()) {
        fail = e;
        return;
      }
      Batch bt;
      std::vector<size_t> idx;
      for (size_t i = a; i < b && e.ok(); i++) {
        if (!items[i].e.ok()) continue;
        uint8_t *dp = (uint8_t *)ws.packet.p + poff[i - a];
        uint16_t *dd = (uint16_t *)((uint8_t *)ws.planes.p + qoff[i - a]);
        uint16_t *da = items[i].alpha_n ? (uint16_t *)((uint8_t *)ws.planes.p + aoff[i - a]) : nullptr;
        e = check(hipMemcpyAsync(dp, data[i], items[i].hd.h.frame_size, hipMemcpyHostToDevice, L->stream),
                  "H2D packet");
        plan_frame(data[i], len[i], (uint64_t)dp, d->bit_depth, dd, da, bt, &items[i].hd);
        idx.push_back(i);
      }
      std::vector<Err> results;
      if (e.ok()) e = run_batch(*dev, *L, bt, t, results);
      if (!e.ok()) {
        fail = e;
        return;
      }
      for (size_t k = 0; k < idx.size() && e.ok(); k++) {
        const size_t i = idx[k];
        items[i].e = results[k];
        if (!results[k].ok()) continue;
        e = check(hipMemcpyAsync(out[i].data, (uint8_t *)ws.planes.p + qoff[i - a], items[i].total * 2,
                                 hipMemcpyDeviceToHost, L->stream),
                  "D2H planes");
        if (e.ok() && items[i].alpha_n)
          e = check(hipMemcpyAsync(out[i].alpha, (uint8_t *)ws.planes.p + aoff[i - a], items[i].alpha_n * 2,
                                   hipMemcpyDeviceToHost, L->stream),
                    "D2H alpha");
      }
      Err s2 = check(hipStreamSynchronize(L->stream), "hipStreamSynchronize");
      if (e.ok()) e = s2;
      if (!e.ok()) {
        fail = e;
        return;
      }
    }
  };
  Err fails[2];
  if (de.ok() && !chunks.empty()) {
    std::thread other;
    if (chunks.size() > 1) other = std::thread(work, 1, std::ref(fails[1]));
    work(0, fails[0]);
    if (other.joinable()) other.join();
  }
  Err call = !de.ok() ? de : !fails[0].ok() ? fails[0] : fails[1];
  for (size_t i = 0; i < n; i++) {
    Err e = call.ok() ? items[i].e : (items[i].e.ok() ? call : items[i].e);
    if (!e.ok() && (out[i].data || out[i].alpha)) {
      prores_frame_free(&out[i]);
      memset(&out[i], 0, sizeof out[i]);
    }
    fill_error(&errs[i], e);
  }
  return (prores_status)call.kind;
}

prores_status prores_decode_device_batch(const prores_decoder *d, size_t n, const prores_dev_decode_job *jobs,
                                         prores_error *errs) {
  return batch_common(d, n, jobs, nullptr, errs);
}

prores_status prores_decode_device_batch_tuned(const prores_decoder *d, size_t n,
                                               const prores_dev_decode_job *jobs,
                                               const prores_decode_tuning *t, prores_error *errs) {
  return batch_common(d, n, jobs, t, errs);
}

void prores_decode_tuning_default(prores_decode_tuning *t) {
  t->persistent = 0;
  t->chunk = 1;
  t->idct_blocks = IDCT_MAX_BLOCKS;
  t->entropy_wg = 0;
  t->graph = 0;
}

prores_status prores_set_decode_tuning(const prores_decode_tuning *t, prores_error *err) {
  Err e = validate_tuning(*t);
  Device *dev = nullptr;
  if (e.ok()) e = device(dev);
  if (e.ok()) {
    std::lock_guard<std::mutex> g(dev->mu);
    dev->tuning = *t;
    dev->tuned = true;
  }
  fill_error(err, e);
  return (prores_status)e.kind;
}

void prores_get_decode_tuning(prores_decode_tuning *t) {
  Device *dev = nullptr;
  prores_decode_tuning_default(t);
  if (!device(dev).ok()) return;
  std::lock_guard<std::mutex> g(dev->mu);
  *t = dev->tuning;
}

prores_status prores_autotune_decode(const prores_decoder *d, size_t n, const prores_dev_decode_job *jobs,
                                     uint32_t reps, prores_decode_tuning *best, double *best_us,
                                     prores_error *err) {
  Device *dev = nullptr;
  Lane *lane0 = nullptr;
  Err e = device(dev);
  if (e.ok()) e = lane(*dev, 0, lane0);
  if (!e.ok()) {
    fill_error(err, e);
    return (prores_status)e.kind;
  }
  std::lock_guard<std::mutex> g(lane0->mu);
  if (reps < 1) reps = 1;
  std::vector<prores_decode_tuning> cands;
  for (uint32_t pers = 0; pers < 2; pers++)
    for (uint32_t chunk : {1u, 4u})
      for (uint32_t blocks : {32u, 64u, 96u})
        for (uint32_t wg : {0u, 256u})
          for (uint32_t graph = 0; graph < 2; graph++) {
            if (!pers && chunk != 1) continue;
            cands.push_back(prores_decode_tuning{pers, chunk, blocks, wg, graph});
          }
  double best_t = 1e300;
  prores_decode_tuning best_c = cands[0];
  std::vector<Err> results;
  std::vector<double> times;
  for (const prores_decode_tuning &c : cands) {
    times.clear();
    for (uint32_t r = 0; r <= reps; r++) {
      if (!(e = run_jobs(*dev, *lane0, d->bit_depth, n, jobs, c, results)).ok()) {
        fill_error(err, e);
        return (prores_status)e.kind;
      }
      if (r) times.push_back(lane0->dec_us[4]);
    }
    std::sort(times.begin(), times.end());
    const double med = times[times.size() / 2];
    if (med < best_t) {
      best_t = med;
      best_c = c;
    }
  }
  {
    std::lock_guard<std::mutex> g2(dev->mu);
    dev->tuning = best_c;
    dev->tuned = true;
  }
  if (best) *best = best_c;
  if (best_us) *best_us = best_t;
  fill_error(err, Err{});
  return PRORES_OK;
}

void prores_last_batch_timing(double *us5) {
  Device *dev = nullptr;
  for (int i = 0; i < 5; i++) us5[i] = 0;
  if (!device(dev).ok()) return;
  // The call lane whose batch ended last.
  uint64_t last = 0;
  bool any = false;
  for (int k : CALL_LANES) {
    Lane *L = nullptr;
    {
      std::lock_guard<std::mutex> g(dev->mu);
      L = dev->lanes[k];
    }
    if (!L) continue;
    std::lock_guard<std::mutex> g(L->mu);
    if (any && L->dec_seq <= last) continue;
    any = true;
    last = L->dec_seq;
    for (int i = 0; i < 5; i++) us5[i] = L->dec_us[i];
  }
}

void prores_graph_counters(uint64_t *out3) {
  out3[0] = out3[1] = out3[2] = 0;
  Device *dev = nullptr;
  if (!device(dev).ok()) return;
  for (int i = 0; i < MAX_LANES; i++) {
    Lane *L = nullptr;
    {
      std::lock_guard<std::mutex> g(dev->mu);
      L = dev->lanes[i];
    }
    if (!L) continue;
    std::lock_guard<std::mutex> g(L->mu);
    if (!L->dec) continue;
    const Workspace &ws = *(const Workspace *)L->dec;
    out3[0] += ws.graph_captures;
    out3[1] += ws.graph_failures;
    out3[2] += ws.graph_replays;
  }
}

const char *prores_version(void) { return "prores-hip 0.2 gfx950"; }

void prores_entropy_tables(uint16_t *level, uint16_t *run) {
  static DevTables t;  // built once; the tables are constant
  static std::once_flag once;
  std::call_once(once, [] { build_tables(t); });
  memcpy(level, t.fast_level, sizeof t.fast_level);
  memcpy(run, t.fast_run, sizeof t.fast_run);
}

}  // extern "C"
