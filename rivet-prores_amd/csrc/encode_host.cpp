// Encoder host runtime: Encoder::encode (src/encode.rs:223-287) over the
// gfx950 kernels. The host checks the configuration and the frame, writes
// the frame header and lays out pictures and slices; the transform, the rate
// control (probe passes, binary search, share-out) and the entropy coding
// run on the device. One launch sequence covers a whole batch of frames.
#include <stdlib.h>

#include <algorithm>
#include <map>

#include "enc_layout.h"
#include "host_internal.h"

namespace prores_hip {
namespace {

const char *PROFILE_NAMES[6] = {"Proxy", "Lt", "Standard", "Hq", "P4444", "P4444Xq"};
const uint64_t BITRATES[6] = {45000000, 102000000, 147000000, 220000000, 330000000, 500000000};
const char *FOURCC[6] = {"apco", "apcs", "apcn", "apch", "ap4h", "ap4x"};
const uint8_t SCAN_P[64] = {0,  1,  4,  5,  16, 17, 21, 22, 2,  3,  6,  7,  18, 20, 23, 28,
                            8,  9,  12, 13, 19, 24, 27, 29, 10, 11, 14, 15, 25, 26, 30, 31,
                            32, 33, 37, 38, 45, 46, 53, 54, 34, 36, 39, 44, 47, 52, 55, 60,
                            35, 40, 43, 48, 51, 56, 59, 61, 41, 42, 49, 50, 57, 58, 62, 63};
constexpr uint32_t WINDOW = 6;  // size-table columns below the uniform qi

uint32_t profile_chroma(uint32_t p) { return p >= PRORES_PROFILE_4444 ? 3u : 2u; }
const char *chroma_name(uint32_t c) { return c == 2 ? "Yuv422" : "Yuv444"; }

uint64_t default_headers(uint32_t w, uint32_t h) {
  // Frame header, picture header and per slice a table entry and a 6-byte
  // slice header, at 8-macroblock slices (src/encode.rs:128-131).
  const uint64_t slices = slice_sizes((w + 15) / 16, 3).size() * (uint64_t)((h + 15) / 16);
  return 8 + 20 + 8 + 8 * slices;
}

uint64_t target_frame_bytes(uint32_t p, uint32_t w, uint32_t h) {
  // The reference rate's bytes per macroblock at 1920x1080 times the frame's
  // macroblocks, plus its headers; f64 as in src/encode.rs:112-117.
  const double per_frame = (double)BITRATES[p] * 1001.0 / 30000.0 / 8.0;
  const double per_mb = (per_frame - (double)default_headers(1920, 1080)) / (120.0 * 68.0);
  const double t = per_mb * (double)((uint64_t)((w + 15) / 16) * ((h + 15) / 16));
  return (uint64_t)(t < 0 ? 0.0 : t) + default_headers(w, h);
}

Err validate_config(const prores_config &c) {  // src/encode.rs:176-192
  if (c.profile > 5) return config_error(fmtstr("profile %u is not a ProRes profile", c.profile));
  if (c.log2_slice_mbs > 3)
    return config_error(fmtstr("log2_slice_mbs %u is outside 0%s3", c.log2_slice_mbs, EN_DASH));
  for (int m = 0; m < 2; m++) {
    if (!(m == 0 ? c.has_luma_matrix : c.has_chroma_matrix)) continue;
    const uint8_t *w = m == 0 ? c.luma_matrix : c.chroma_matrix;
    for (int k = 0; k < 64; k++)
      if (w[k] < 2 || w[k] > 63) return config_error(fmtstr("quantisation weights must be 2%s63", EN_DASH));
  }
  return Err{};
}

Err validate_frame(const prores_frame &f) {  // Frame::validate, src/frame.rs:299-340
  if (f.width == 0 || f.height == 0 || f.width > 65535 || f.height > 65535)
    return config_error(fmtstr("a frame of %u%s%u cannot be coded", f.width, TIMES, f.height));
  if (f.bit_depth < 8 || f.bit_depth > 16)
    return config_error(fmtstr("bit depth %u is outside 8%s16", f.bit_depth, EN_DASH));
  if (f.chroma != 2 && f.chroma != 3) return config_error(fmtstr("chroma format %u is not 4:2:2 or 4:4:4", f.chroma));
  prores_plane want[3];
  prores_frame_layout(f.width, f.height, f.chroma, want);
  for (int i = 0; i < 3; i++) {
    const prores_plane &p = f.planes[i];
    if (p.width != want[i].width || p.height != want[i].height)
      return config_error(fmtstr("plane %d is %u%s%u; a %s frame of %u%s%u needs %u%s%u", i, p.width, TIMES,
                                 p.height, chroma_name(f.chroma), f.width, TIMES, f.height, want[i].width,
                                 TIMES, want[i].height));
    const uint64_t end = p.offset + (uint64_t)p.width * p.height;
    if (end < p.offset || end > f.data_len)
      return config_error(fmtstr("plane %d runs past the end of the frame's data", i));
  }
  if (f.alpha && f.alpha_len != (size_t)f.width * f.height)
    return config_error(fmtstr("the alpha plane has %zu samples, not %u%s%u", f.alpha_len, f.width, TIMES, f.height));
  return Err{};
}

struct EncBatch {
  std::vector<EncFrame> frames;
  std::vector<EncPic> pics;
  std::vector<EncRowSlot> slots;  // the pictures' row layouts
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> slot_index;  // (width in MBs, log2) -> first slot
  uint64_t n_slices = 0;
  uint32_t max_pic_slices = 0;
  uint32_t max_slice_blocks = 0;  // the largest slice's blocks: pe_fdct's LDS is sized for it
  std::vector<EncState> states;
  // The distinct quantiser-table keys of the batch (luma and chroma
  // weights, interlaced); a picture's qtab names one until run_encode maps
  // it to the table's slot in the lane's device table cache.
  std::vector<std::vector<uint8_t>> qkeys;
  std::map<std::vector<uint8_t>, uint32_t> qtab_index;
  std::vector<int> job_frame;  // batch frame index per job, -1 when the job failed its checks
  // Each frame's output arena guess: its target_frame_bytes plus a margin,
  // or the most its slices could take if that is less.
  std::vector<uint64_t> targets;
  uint64_t n_coefs = 0;
};

constexpr size_t QTAB_FLOATS = 225 * 128;

// 8 / (W qScale) at scanned positions for luma and chroma and every
// quantization_index (index 0 as 1), IEEE single division (src/encode.rs:
// 303-324), for a key (luma weights, chroma weights, interlaced), into
// `out` (QTAB_FLOATS floats). Tables computed before are kept, process-wide:
// the same matrices recur from frame to frame.
void quantiser_values(const std::vector<uint8_t> &key, float *out) {
  static std::mutex mu;
  static std::map<std::vector<uint8_t>, std::vector<float>> *cache =
      new std::map<std::vector<uint8_t>, std::vector<float>>();  // never destroyed: encodes may run at exit
  std::lock_guard<std::mutex> g(mu);
  auto c = cache->find(key);
  if (c == cache->end()) {
    if (cache->size() >= 4096) cache->clear();
    std::vector<float> t(QTAB_FLOATS);
    for (uint32_t qi = 0; qi <= 224; qi++) {
      const uint32_t q = qscale(qi ? qi : 1);
      for (int k = 0; k < 64; k++) {
        const uint32_t s = key[128] ? SCAN_P[8 * (k % 8) + k / 8] : SCAN_P[k];
        t[qi * 128 + s] = 8.0f / (float)((uint32_t)key[k] * q);
        t[qi * 128 + 64 + s] = 8.0f / (float)((uint32_t)key[64 + k] * q);
      }
    }
    c = cache->emplace(key, std::move(t)).first;
  }
  memcpy(out, c->second.data(), QTAB_FLOATS * sizeof(float));
}

// The batch-local index of a frame's quantiser table key.
uint32_t quantiser_table(EncBatch &b, const Header &hd) {
  std::vector<uint8_t> key(129);
  hd.luma_weights(key.data());
  hd.chroma_weights(key.data() + 64);
  key[128] = hd.h.interlace != 0;
  auto it = b.qtab_index.find(key);
  if (it != b.qtab_index.end()) return it->second;
  const uint32_t index = (uint32_t)b.qkeys.size();
  b.qtab_index[key] = index;
  b.qkeys.push_back(key);
  return index;
}

// Checks one job and lays out its pictures and slices.
Err plan_encode(const prores_config &cfg, const prores_frame &f, uint64_t dev_data, uint64_t dev_alpha,
                EncBatch &b) {
  Err e = validate_config(cfg);
  if (e.ok()) e = validate_frame(f);
  if (!e.ok()) return e;
  const uint32_t pc = profile_chroma(cfg.profile);
  if (f.chroma != pc)
    return config_error(fmtstr("%s codes %s, and the frame is %s", PROFILE_NAMES[cfg.profile], chroma_name(pc),
                               chroma_name(f.chroma)));
  uint32_t alpha = PRORES_ALPHA_NONE;
  if (f.alpha) {
    if (pc == 2) return config_error(fmtstr("%s carries no alpha channel", PROFILE_NAMES[cfg.profile]));
    alpha = cfg.alpha;
  }
  Header hd;
  memset(&hd.h, 0, sizeof hd.h);
  prores_frame_header &h = hd.h;
  h.bitstream_version = (f.chroma == 2 && alpha == 0) ? 0 : 1;
  memcpy(h.encoder_identifier, cfg.encoder_identifier, 4);
  h.width = (uint16_t)f.width;
  h.height = (uint16_t)f.height;
  h.chroma = f.chroma;
  h.interlace = f.interlace;
  h.metadata = f.metadata;
  h.alpha = alpha;
  h.has_luma_matrix = cfg.has_luma_matrix;
  h.has_chroma_matrix = cfg.has_chroma_matrix;
  memcpy(h.luma_matrix, cfg.luma_matrix, 64);
  memcpy(h.chroma_matrix, cfg.chroma_matrix, 64);

  EncFrame F;
  memset(&F, 0, sizeof F);
  for (int i = 0; i < 3; i++) {
    F.in_plane[i] = dev_data + 2 * f.planes[i].offset;
    F.plane_w[i] = f.planes[i].width;
  }
  F.in_alpha = alpha ? dev_alpha : 0;
  F.width = f.width;
  F.height = f.height;
  F.chroma = f.chroma;
  F.bit_depth = f.bit_depth;
  F.interlaced = f.interlace != 0;
  F.alpha_bits = hd.alpha_bits();
  if (F.alpha_bits) {
    const uint64_t max_in = (1ull << f.bit_depth) - 1, max_out = (1ull << F.alpha_bits) - 1;
    F.rescale_m = ((max_out << 40) + max_in - 1) / max_in;
    F.rescale_identity = f.bit_depth == F.alpha_bits;
  }
  F.max_in = 0xFFFFu >> (16 - f.bit_depth);
  F.in_scale = 512.0f / (float)(1u << f.bit_depth);
  // The frame header's fields; pe_pack writes its bytes.
  F.hdr_len = 28 + 64 * ((uint32_t)h.has_luma_matrix + h.has_chroma_matrix);
  F.version = h.bitstream_version;
  memcpy(F.encoder_id, h.encoder_identifier, 4);
  F.chroma_interlace = (uint8_t)((h.chroma << 6) | (h.interlace << 2));
  F.aspect_rate = (uint8_t)((h.metadata.aspect_ratio << 4) | (h.metadata.frame_rate_code & 15));
  F.primaries = h.metadata.color_primaries;
  F.transfer = h.metadata.transfer_characteristic;
  F.matrix = h.metadata.matrix_coefficients;
  F.alpha_type = (uint8_t)h.alpha;
  F.matrix_flags = (uint8_t)((h.has_luma_matrix << 1) | h.has_chroma_matrix);
  memcpy(F.luma_matrix, h.luma_matrix, 64);
  memcpy(F.chroma_matrix, h.chroma_matrix, 64);
  const uint64_t target = cfg.target_frame_bytes ? cfg.target_frame_bytes : target_frame_bytes(cfg.profile, f.width, f.height);
  const uint32_t npics = hd.picture_count();
  const uint64_t per_picture = (target > F.hdr_len ? target - F.hdr_len : 0) / npics;
  const uint32_t qtab = quantiser_table(b, hd);
  F.first_pic = (uint32_t)b.pics.size();
  F.n_pics = npics;
  F.first_slice = (uint32_t)b.n_slices;
  const uint32_t fi = (uint32_t)b.frames.size();
  const uint32_t wmb = (f.width + 15) / 16, cpm = f.chroma == 2 ? 2u : 4u;
  const std::vector<uint32_t> sizes = slice_sizes(wmb, cfg.log2_slice_mbs);
  uint32_t layout_base;
  {
    const auto key = std::make_pair(wmb, cfg.log2_slice_mbs);
    auto it = b.slot_index.find(key);
    if (it != b.slot_index.end()) {
      layout_base = it->second;
    } else {
      layout_base = (uint32_t)b.slots.size();
      b.slot_index[key] = layout_base;
      uint32_t mx = 0;
      for (uint32_t m : sizes) {
        b.slots.push_back(EncRowSlot{(uint16_t)mx, (uint16_t)m, mx});
        mx += m;
      }
    }
  }
  for (uint32_t index = 0; index < npics; index++) {
    EncPic P;
    memset(&P, 0, sizeof P);
    P.budget = per_picture;
    P.frame = fi;
    P.height = hd.picture_height(index);
    hd.picture_rows(index, P.first_row, P.row_step);
    P.header_bytes = alpha ? 8 : 6;
    P.log2_slice_mbs = cfg.log2_slice_mbs;
    P.qtab = qtab;
    P.interlaced = F.interlaced;
    P.slice_base = (uint32_t)b.n_slices;
    P.rows = (P.height + 15) / 16;
    P.row_slices = (uint32_t)sizes.size();
    P.layout_base = layout_base;
    P.row_mbs = wmb;
    P.n_slices = P.rows * P.row_slices;
    P.coef_base = b.n_coefs;
    b.n_coefs += 64ull * (4 + 2 * cpm) * wmb * P.rows;
    b.n_slices += P.n_slices;
    b.max_pic_slices = std::max(b.max_pic_slices, P.n_slices);
    for (uint32_t m : sizes) b.max_slice_blocks = std::max(b.max_slice_blocks, (4 + 2 * cpm) * m);
    b.pics.push_back(P);
    EncState st;
    memset(&st, 0, sizeof st);
    st.probe = 224;
    // The probes of the search while every one fits.
    const uint32_t path[5] = {224, 112, 56, 28, 14};
    static_assert(ENC_PASS_Q <= 5, "first pass list");
    st.nq = ENC_PASS_Q;
    memcpy(st.pq, path, ENC_PASS_Q * sizeof(uint32_t));
    b.states.push_back(st);
  }
  b.frames.push_back(F);
  const uint64_t most = F.hdr_len + 8ull * npics + (b.n_slices - F.first_slice) * (2 + 65535);
  b.targets.push_back(std::min<uint64_t>(target + 4096, most));
  return Err{};
}

struct EncWorkspace {
  DevBuf plan, slices, coefs, alpha_scratch, alpha_len, size_tab, deep, used, qidx, slice_out, frame_size, out, planes,
      alpha_in, probe_sizes, chunk_bounds, out_addr;
  // The plan (frames, pictures, slices, rate-control states, quantiser
  // tables, frame headers) goes up in one copy from page-locked staging; the
  // frame sizes and the arena's overflow flag come back in one.
  HostBuf hplan, hres;
  // Quantiser tables resident on the device, one QTAB_FLOATS slot a key,
  // uploaded the first time a batch on this lane needs them.
  DevBuf qarena;
  std::map<std::vector<uint8_t>, uint32_t> qslot;
};
EncWorkspace &enc_workspace(Lane &L) {
  if (!L.enc) L.enc = new EncWorkspace();
  return *(EncWorkspace *)L.enc;
}

// Runs the kernels over a planned batch; on return the packets are in the
// output arena at `offsets[i]` with `sizes[i]` bytes (batch frame order).
Err run_encode(Device &d, Lane &L, EncBatch &b, std::vector<uint64_t> &offsets, std::vector<uint64_t> &sizes) {
  EncWorkspace &ws = enc_workspace(L);
  Err e;
  const uint64_t G = b.n_slices;
  if (G > 0xFFFFFFF0ull) return invalid("a batch of more than 2^32 slices");
  const uint32_t NP = (uint32_t)b.pics.size(), NF = (uint32_t)b.frames.size();
  // The plan's parts at 256-byte offsets of one buffer.
  size_t at = 0;
  auto part = [&](size_t bytes) {
    const size_t o = at;
    at += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_frames = part(b.frames.size() * sizeof(EncFrame)), o_pics = part(b.pics.size() * sizeof(EncPic)),
               o_slots = part(b.slots.size() * sizeof(EncRowSlot)),
               o_states = part(b.states.size() * sizeof(EncState)),
               plan_bytes = at;
  // The batch's quantiser tables in the lane's device cache: keys not there
  // get the next slots (the arena grows by clearing the cache: every table
  // of the batch is then uploaded again) and go up with the plan.
  std::vector<uint32_t> slot_of(b.qkeys.size());
  std::vector<size_t> fresh;
  for (int attempt = 0; attempt < 2; attempt++) {
    fresh.clear();
    uint32_t next = (uint32_t)ws.qslot.size();
    for (size_t k = 0; k < b.qkeys.size(); k++) {
      auto it = ws.qslot.find(b.qkeys[k]);
      if (it != ws.qslot.end()) {
        slot_of[k] = it->second;
      } else {
        slot_of[k] = next++;
        fresh.push_back(k);
      }
    }
    if ((size_t)next * QTAB_FLOATS * 4 <= ws.qarena.cap) break;
    ws.qslot.clear();
    if (!(e = ws.qarena.reserve((size_t)std::max<uint32_t>(next, 32) * QTAB_FLOATS * 4)).ok()) return e;
  }
  const uint32_t first_fresh = (uint32_t)ws.qslot.size();
  for (size_t k : fresh) ws.qslot[b.qkeys[k]] = slot_of[k];
  for (EncPic &P : b.pics) P.qtab = slot_of[P.qtab];
  const size_t o_fresh = part(fresh.size() * QTAB_FLOATS * 4), staged = at;
  if (!(e = ws.plan.reserve(plan_bytes + 64)).ok() || !(e = ws.hplan.reserve(staged + 64)).ok()) return e;
  {
    uint8_t *h = (uint8_t *)ws.hplan.p;
    auto put = [&](size_t o, const void *src, size_t n) {
      if (n) memcpy(h + o, src, n);
    };
    put(o_frames, b.frames.data(), b.frames.size() * sizeof(EncFrame));
    put(o_pics, b.pics.data(), b.pics.size() * sizeof(EncPic));
    put(o_slots, b.slots.data(), b.slots.size() * sizeof(EncRowSlot));
    put(o_states, b.states.data(), b.states.size() * sizeof(EncState));
    for (size_t i = 0; i < fresh.size(); i++) quantiser_values(b.qkeys[fresh[i]], (float *)(h + o_fresh) + i * QTAB_FLOATS);
    if (!(e = check(hipMemcpyAsync(ws.plan.p, ws.hplan.p, plan_bytes, hipMemcpyHostToDevice, L.stream),
                    "hipMemcpyAsync(encode plan)")).ok())
      return e;
    // New slots are consecutive from first_fresh.
    if (!fresh.empty() &&
        !(e = check(hipMemcpyAsync((uint8_t *)ws.qarena.p + (size_t)first_fresh * QTAB_FLOATS * 4, h + o_fresh,
                                   fresh.size() * QTAB_FLOATS * 4, hipMemcpyHostToDevice, L.stream),
                    "hipMemcpyAsync(quantiser tables)")).ok())
      return e;
  }
  if (!(e = ws.slices.reserve(G * sizeof(EncSlice) + 64)).ok()) return e;
  if (!(e = ws.coefs.reserve(b.n_coefs * 4 + 512)).ok()) return e;
  if (!(e = ws.alpha_scratch.reserve(G * ALPHA_SLICE_CAP + 64)).ok()) return e;
  if (!(e = ws.alpha_len.reserve(G * 4 + 64)).ok()) return e;
  if (!(e = ws.size_tab.reserve(G * ENC_MAXK * 4 + 64)).ok()) return e;
  if (!(e = ws.probe_sizes.reserve(G * ENC_HIST * 4 + 64)).ok()) return e;
  if (!(e = ws.chunk_bounds.reserve(b.n_coefs / 64 * 4 + 512)).ok()) return e;
  if (!(e = ws.deep.reserve(G * ENC_DEEP * 4 + 64)).ok()) return e;
  if (!(e = ws.used.reserve(G * 4 + 64)).ok()) return e;
  if (!(e = ws.qidx.reserve(G * 4 + 64)).ok()) return e;
  if (!(e = ws.slice_out.reserve(G * 8 + 64)).ok()) return e;
  if (!(e = ws.frame_size.reserve(NF * 8 + 64)).ok()) return e;
  if (!(e = ws.out_addr.reserve(NF * 8 + 64)).ok()) return e;
  uint8_t *const pl = (uint8_t *)ws.plan.p;
  void *frames = pl + o_frames, *pics = pl + o_pics, *slots = pl + o_slots, *states = pl + o_states;
  void *slices = ws.slices.p;
  void *qtabs = ws.qarena.p, *coefs = ws.coefs.p, *ascr = ws.alpha_scratch.p;
  void *alen = ws.alpha_len.p, *stab = ws.size_tab.p, *used = ws.used.p, *qidx = ws.qidx.p;
  void *sout = ws.slice_out.p, *fsize = ws.frame_size.p, *tables = d.tables, *deep = ws.deep.p;
  void *psz = ws.probe_sizes.p, *cbnd = ws.chunk_bounds.p;
  uint32_t g32 = (uint32_t)G, np32 = NP, window = WINDOW;
  hipEvent_t *ev = L.eev;
  (void)hipEventRecord(ev[0], L.stream);
  {
    void *args[] = {&frames, &pics, &slots, &slices};
    if (!(e = launch(L, d.k_slices, (b.max_pic_slices + 255) / 256, np32, 256, args)).ok()) return e;
  }
  {
    // LDS for the batch's largest slice: a batch of slices of at most 64
    // blocks (4:2:2 up to 8 macroblocks) fits six workgroups a CU.
    void *args[] = {&tables, &frames, &pics, &slices, &qtabs, &coefs, &cbnd};
    if (!(e = launch(L, b.max_slice_blocks <= 64 ? d.k_fdct64 : d.k_fdct, g32, 1, 256, args)).ok()) return e;
  }
  (void)hipEventRecord(ev[1], L.stream);
  {
    void *args[] = {&frames, &pics, &slices, &g32, &ascr, &alen, &states};
    if (!(e = launch(L, d.k_alpha_enc, g32, 1, 64, args)).ok()) return e;
  }
  (void)hipEventRecord(ev[2], L.stream);
  // The search's size passes: the first at five quantisers, then one each.
  for (uint32_t pass = 0; pass < ENC_PASSES; pass++) {
    void *a1[] = {&tables, &frames, &pics, &slices, &g32, &coefs, &qtabs, &alen, &states, &psz, &cbnd};
    if (!(e = launch(L, pass ? d.k_size : d.k_size_first, (g32 + 3) / 4, 1, 256, a1)).ok()) return e;
    void *a2[] = {&pics, &states, &np32, &window};
    if (!(e = launch(L, d.k_search, (np32 + 63) / 64, 1, 64, a2)).ok()) return e;
  }
  (void)hipEventRecord(ev[3], L.stream);
  {
    // The size cache starts empty: the window fills its deep columns.
    if (G && !(e = check(hipMemsetAsync(ws.deep.p, 0, G * ENC_DEEP * 4, L.stream), "memset")).ok()) return e;
    void *args[] = {&tables, &frames, &pics, &slices, &g32, &coefs, &qtabs, &alen, &states, &psz, &cbnd, &stab, &deep};
    if (!(e = launch(L, d.k_window, (g32 + 3) / 4, 1, 256, args)).ok()) return e;
  }
  (void)hipEventRecord(ev[4], L.stream);
  {
    void *args[] = {&tables, &frames, &pics, &slices, &coefs, &qtabs, &alen, &states, &stab, &deep, &used, &qidx,
                    &cbnd};
    if (!(e = launch(L, d.k_shareout, np32, 1, 64 * SHARE_WAVES, args)).ok()) return e;
  }
  (void)hipEventRecord(ev[5], L.stream);
  {
    void *args[] = {&frames, &pics, &used, &sout, &fsize};
    if (!(e = launch(L, d.k_offsets, NF, 1, 256, args)).ok()) return e;
  }
  (void)hipEventRecord(ev[6], L.stream);
  // The packets' places in the output arena are found on the device; the
  // arena holds every frame at its target (a frame over it is one that does
  // not fit at qi 224) or the largest batch seen, else the pack runs again
  // over a larger one.
  uint64_t want = 0;
  for (uint32_t f = 0; f < NF; f++) want += (b.targets[f] + 255) & ~255ull;
  if (!(e = ws.out.reserve(want + 256)).ok()) return e;  // grow-only: keeps the largest batch's arena
  void *oaddr = ws.out_addr.p, *oflow = (uint8_t *)ws.out_addr.p + NF * 8;
  if (!(e = ws.hres.reserve(NF * 8 + 64)).ok()) return e;
  (void)hipEventRecord(ev[7], L.stream);
  auto pack = [&](uint64_t capacity) -> Err {
    Err x;
    uint64_t base = (uint64_t)ws.out.p;
    uint32_t nf = NF;
    void *aargs[] = {&fsize, &nf, &base, &capacity, &oaddr, &oflow};
    if (!(x = launch(L, d.k_arena, 1, 1, 256, aargs)).ok()) return x;
    void *args[] = {&frames, &pics,  &slices, &coefs, &qtabs, &alen,  &ascr,
                    &used,   &qidx,  &sout,   &fsize, &oaddr, &oflow, &cbnd};
    if (!(x = launch(L, d.k_pack, g32, 1, 64, args)).ok()) return x;
    return check(hipMemcpyAsync(ws.hres.p, fsize, NF * 8, hipMemcpyDeviceToHost, L.stream), "frame sizes");
  };
  if (!(e = pack(ws.out.cap - 256)).ok()) return e;
  uint8_t *const hres = (uint8_t *)ws.hres.p;
  if (!(e = check(hipMemcpyAsync(hres + NF * 8, oflow, 4, hipMemcpyDeviceToHost, L.stream), "arena flag")).ok())
    return e;
  (void)hipEventRecord(ev[8], L.stream);
  if (!(e = check(hipStreamSynchronize(L.stream), "hipStreamSynchronize")).ok()) return e;
  uint32_t overflow = 0;
  memcpy(&overflow, hres + NF * 8, 4);
  sizes.assign((const uint64_t *)hres, (const uint64_t *)hres + NF);
  offsets.assign(NF, 0);
  uint64_t total = 0;
  for (uint32_t f = 0; f < NF; f++) {
    if (sizes[f] > 0xFFFFFFFFull) return invalid("an encoded frame exceeds 4 GiB");
    offsets[f] = total;
    total += (sizes[f] + 255) & ~255ull;
  }
  if (overflow) {
    if (!(e = ws.out.reserve(total + 256)).ok()) return e;
    if (!(e = pack(ws.out.cap - 256)).ok()) return e;
    (void)hipEventRecord(ev[8], L.stream);
    if (!(e = check(hipStreamSynchronize(L.stream), "hipStreamSynchronize")).ok()) return e;
  }
  {
    // Diagnostics: sizes the share-out computed past the window.
    std::vector<EncState> st(NP);
    if (NP && copy_on(L, st.data(), states, NP * sizeof(EncState), hipMemcpyDeviceToHost).ok()) {
      uint64_t total = 0, most = 0, uni = 0;
      for (const EncState &x : st) {
        total += x.fallback;
        most = x.fallback > most ? x.fallback : most;
        uni += x.uniform;
      }
      L.enc_counters[0] = total;
      L.enc_counters[1] = most;
      L.enc_counters[2] = NP;
      L.enc_counters[3] = NP ? uni / NP : 0;
    }
  }
  // Stages: transform, alpha, probes, window, share-out, offsets, pack, and
  // the whole batch (the host's wait for the packet sizes included).
  const int from[8] = {0, 1, 2, 3, 4, 5, 7, 0}, to[8] = {1, 2, 3, 4, 5, 6, 8, 8};
  for (int i = 0; i < 8; i++) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[from[i]], ev[to[i]]);
    L.enc_stage_us[i] = ms * 1000.0;
  }
  const double *st = L.enc_stage_us;
  L.enc_us[0] = st[0] + st[1];
  L.enc_us[1] = st[2] + st[3] + st[4] + st[5];
  L.enc_us[2] = st[6];
  L.enc_us[3] = st[7];
  return Err{};
}

}  // namespace
}  // namespace prores_hip

namespace prores_hip {

Err encode_check(const prores_config &cfg, const prores_frame &frame) {
  EncBatch probe;
  return plan_encode(cfg, frame, 0, 0, probe);
}

Err encode_batch_on(Device &d, Lane &L, size_t n, prores_dev_encode_job *jobs, prores_error *errs) {
  Err e = ensure_encoder(d);
  if (!e.ok()) {
    for (size_t i = 0; i < n; i++) fill_error(&errs[i], e);
    return e;
  }
  EncBatch b;
  std::vector<int> frame_of(n, -1);
  for (size_t i = 0; i < n; i++) {
    jobs[i].dev_out = nullptr;
    jobs[i].out_len = 0;
    const uint32_t before = (uint32_t)b.frames.size();
    Err je = plan_encode(*jobs[i].cfg, jobs[i].frame, (uint64_t)jobs[i].frame.data, (uint64_t)jobs[i].frame.alpha, b);
    fill_error(&errs[i], je);
    if (je.ok()) frame_of[i] = (int)before;
  }
  std::vector<uint64_t> offsets, sizes;
  if (!b.frames.empty()) e = run_encode(d, L, b, offsets, sizes);
  if (!e.ok()) {
    for (size_t i = 0; i < n; i++) fill_error(&errs[i], e);
    return e;
  }
  EncWorkspace &ws = enc_workspace(L);
  for (size_t i = 0; i < n; i++)
    if (frame_of[i] >= 0) {
      jobs[i].dev_out = (const uint8_t *)ws.out.p + offsets[frame_of[i]];
      jobs[i].out_len = sizes[frame_of[i]];
    }
  return Err{};
}

}  // namespace prores_hip

using namespace prores_hip;

extern "C" {

void prores_config_default(uint32_t profile, prores_config *c) {  // Config::new, src/encode.rs:163-174
  memset(c, 0, sizeof *c);
  c->profile = profile;
  c->log2_slice_mbs = 3;
  c->alpha = PRORES_ALPHA_16;
  memcpy(c->encoder_identifier, "rivt", 4);
}

uint64_t prores_profile_target_frame_bytes(uint32_t profile, uint32_t width, uint32_t height) {
  return profile > 5 ? 0 : target_frame_bytes(profile, width, height);
}

void prores_profile_fourcc(uint32_t profile, char out[4]) {
  memcpy(out, FOURCC[profile > 5 ? 0 : profile], 4);
}

int32_t prores_profile_from_fourcc(const char code[4]) {
  for (int p = 0; p < 6; p++)
    if (memcmp(code, FOURCC[p], 4) == 0) return p;
  return -1;
}

void prores_buffer_free(uint8_t *p) { free(p); }

prores_status prores_encode(const prores_config *cfg, const prores_frame *frame, uint8_t **out, size_t *out_len,
                            prores_error *err) {
  *out = nullptr;
  *out_len = 0;
  // Checks first: a refused frame or configuration needs no device.
  EncBatch probe;
  Err e = plan_encode(*cfg, *frame, 0, 0, probe);
  if (!e.ok()) {
    fill_error(err, e);
    return (prores_status)e.kind;
  }
  Device *dev = nullptr;
  Lane *lane0 = nullptr;
  if (!(e = device(dev)).ok() || !(e = lane(*dev, 0, lane0)).ok()) {
    fill_error(err, e);
    return (prores_status)e.kind;
  }
  std::lock_guard<std::mutex> g(lane0->mu);
  if (!(e = ensure_encoder(*dev)).ok()) {
    fill_error(err, e);
    return (prores_status)e.kind;
  }
  EncWorkspace &ws = enc_workspace(*lane0);
  const size_t nsamp = frame->data_len, nalpha = frame->alpha ? frame->alpha_len : 0;
  if ((e = ws.planes.reserve(nsamp * 2 + 64)).ok() && (e = ws.alpha_in.reserve(nalpha * 2 + 64)).ok() &&
      (e = check(hipMemcpyAsync(ws.planes.p, frame->data, nsamp * 2, hipMemcpyHostToDevice, lane0->stream), "H2D frame")).ok() &&
      (!nalpha || (e = check(hipMemcpyAsync(ws.alpha_in.p, frame->alpha, nalpha * 2, hipMemcpyHostToDevice, lane0->stream),
                             "H2D alpha")).ok())) {
    EncBatch b;
    e = plan_encode(*cfg, *frame, (uint64_t)ws.planes.p, (uint64_t)ws.alpha_in.p, b);
    std::vector<uint64_t> offsets, sizes;
    if (e.ok()) e = run_encode(*dev, *lane0, b, offsets, sizes);
    if (e.ok()) {
      *out = (uint8_t *)malloc(sizes[0] ? sizes[0] : 1);
      if (!*out) e = Err{PRORES_ERR_DEVICE, "host allocation failed"};
    }
    if (e.ok()) e = copy_on(*lane0, *out, (uint8_t *)ws.out.p + offsets[0], sizes[0], hipMemcpyDeviceToHost);
    if (e.ok()) {
      *out_len = sizes[0];
    } else {
      free(*out);
      *out = nullptr;
    }
  }
  fill_error(err, e);
  return (prores_status)e.kind;
}

prores_status prores_encode_device_batch(size_t n, prores_dev_encode_job *jobs, prores_error *errs) {
  Device *dev = nullptr;
  Lane *lane0 = nullptr;
  Err e = device(dev);
  if (e.ok()) e = lane(*dev, 0, lane0);
  if (e.ok()) {
    std::lock_guard<std::mutex> g(lane0->mu);
    e = encode_batch_on(*dev, *lane0, n, jobs, errs);
    return e.ok() ? PRORES_OK : (prores_status)e.kind;
  }
  for (size_t i = 0; i < n; i++) fill_error(&errs[i], e);
  return (prores_status)e.kind;
}

void prores_last_encode_counters(uint64_t *out4) {
  Device *dev = nullptr;
  Lane *lane0 = nullptr;
  for (int i = 0; i < 4; i++) out4[i] = 0;
  if (!device(dev).ok() || !lane(*dev, 0, lane0).ok()) return;
  std::lock_guard<std::mutex> g(lane0->mu);
  for (int i = 0; i < 4; i++) out4[i] = lane0->enc_counters[i];
}

void prores_last_encode_stages(double *us8) {
  Device *dev = nullptr;
  Lane *lane0 = nullptr;
  for (int i = 0; i < 8; i++) us8[i] = 0;
  if (!device(dev).ok() || !lane(*dev, 0, lane0).ok()) return;
  std::lock_guard<std::mutex> g(lane0->mu);
  for (int i = 0; i < 8; i++) us8[i] = lane0->enc_stage_us[i];
}

void prores_last_encode_timing(double *us4) {
  Device *dev = nullptr;
  Lane *lane0 = nullptr;
  for (int i = 0; i < 4; i++) us4[i] = 0;
  if (!device(dev).ok() || !lane(*dev, 0, lane0).ok()) return;
  std::lock_guard<std::mutex> g(lane0->mu);
  for (int i = 0; i < 4; i++) us4[i] = lane0->enc_us[i];
}

}  // extern "C"
