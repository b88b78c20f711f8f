// Host runtime pieces shared by decode and encode: errors, frame and picture
// header syntax (src/header.rs), the constant tables the kernels read, and
// the per-device context (code objects loaded with hipModuleLoadData).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include "host_internal.h"

extern "C" const unsigned char prores_decode_co[];
extern "C" const unsigned char prores_encode_co[];

namespace prores_hip {

const char *const TIMES = "\xC3\x97";
const char *const EN_DASH = "\xE2\x80\x93";

std::string fmtstr(const char *f, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

Err invalid(const std::string &m) { return Err{PRORES_ERR_INVALID, m}; }
Err unsupported(const std::string &m) { return Err{PRORES_ERR_UNSUPPORTED, m}; }
Err config_error(const std::string &m) { return Err{PRORES_ERR_CONFIG, m}; }
Err device_error(const std::string &where, hipError_t e) {
  return Err{PRORES_ERR_DEVICE, where + ": " + hipGetErrorString(e)};
}
Err check(hipError_t e, const char *where) { return e == hipSuccess ? Err{} : device_error(where, e); }

void fill_error(prores_error *out, const Err &e) {
  if (!out) return;
  out->kind = e.kind;
  size_t n = e.msg.size() < sizeof(out->msg) - 1 ? e.msg.size() : sizeof(out->msg) - 1;
  memcpy(out->msg, e.msg.data(), n);
  out->msg[n] = 0;
}

std::vector<uint32_t> slice_sizes(uint32_t width_in_mb, uint32_t log2) {
  // Slices of 2^log2 macroblocks from the left, then halving (RDD 36 §6.2).
  std::vector<uint32_t> v;
  uint32_t size = 1u << log2, left = width_in_mb;
  while (left) {
    while (left >= size) {
      v.push_back(size);
      left -= size;
    }
    size >>= 1;
  }
  return v;
}

void Header::luma_weights(uint8_t *w) const {
  if (h.has_luma_matrix)
    memcpy(w, h.luma_matrix, 64);
  else
    memset(w, 4, 64);
}
void Header::chroma_weights(uint8_t *w) const {
  if (h.has_chroma_matrix)
    memcpy(w, h.chroma_matrix, 64);
  else
    luma_weights(w);
}

static uint32_t be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static uint16_t be16(const uint8_t *p) { return (uint16_t)((p[0] << 8) | p[1]); }

Err parse_frame_header(const uint8_t *data, size_t len, Header &out) {
  // The checks and their order follow FrameHeader::parse (src/header.rs:82-170).
  prores_frame_header &h = out.h;
  memset(&h, 0, sizeof h);
  if (len < 28) return invalid(fmtstr("%zu bytes is too short for a ProRes frame header", len));
  h.frame_size = be32(data);
  if (memcmp(data + 4, "icpf", 4) != 0) return invalid("no 'icpf' frame identifier");
  if ((size_t)h.frame_size > len)
    return invalid(fmtstr("frame_size %u exceeds the %zu bytes given", h.frame_size, len));
  if (h.frame_size < 28) return invalid("frame_size leaves no room for the frame header");
  const uint8_t *p = data + 8;
  const size_t avail = h.frame_size - 8;
  h.header_size = be16(p);
  if (h.header_size < 20 || h.header_size > avail)
    return invalid(fmtstr("frame_header_size %u is out of range", h.header_size));
  h.bitstream_version = p[3];
  if (h.bitstream_version > 1) return unsupported(fmtstr("bitstream_version %u", h.bitstream_version));
  memcpy(h.encoder_identifier, p + 4, 4);
  h.width = be16(p + 8);
  h.height = be16(p + 10);
  if (h.width == 0 || h.height == 0) return invalid(fmtstr("a %u%s%u frame", h.width, TIMES, h.height));
  const uint32_t chroma = p[12] >> 6;
  if (chroma != 2 && chroma != 3) return invalid(fmtstr("chroma_format %u is reserved", chroma));
  h.chroma = chroma;
  const uint32_t il = (p[12] >> 2) & 3;
  if (il == 3) return invalid("interlace_mode 3 is reserved");
  h.interlace = il;
  h.metadata.aspect_ratio = p[13] >> 4;
  h.metadata.frame_rate_code = p[13] & 15;
  h.metadata.color_primaries = p[14];
  h.metadata.transfer_characteristic = p[15];
  h.metadata.matrix_coefficients = p[16];
  const uint32_t alpha = p[17] & 15;
  if (alpha > 2) return invalid(fmtstr("alpha_channel_type %u is reserved", alpha));
  h.alpha = alpha;
  size_t pos = 20;
  if (p[19] & 2) {
    if (pos + 64 > h.header_size) return invalid("a quantisation matrix runs past frame_header_size");
    memcpy(h.luma_matrix, p + pos, 64);
    h.has_luma_matrix = 1;
    pos += 64;
  }
  if (p[19] & 1) {
    if (pos + 64 > h.header_size) return invalid("a quantisation matrix runs past frame_header_size");
    memcpy(h.chroma_matrix, p + pos, 64);
    h.has_chroma_matrix = 1;
  }
  return Err{};
}

Err check_plausible_size(const Header &hd) {
  // Every slice needs at least 11 bytes and a picture has at least one slice
  // per 8 macroblocks (src/decode.rs:107-128).
  const uint64_t wmb = ((uint64_t)hd.h.width + 15) / 16;
  const uint64_t per_row = slice_sizes((uint32_t)wmb, 3).size();
  uint64_t need = 8 + (uint64_t)hd.h.header_size;
  for (uint32_t i = 0; i < hd.picture_count(); i++)
    need += 8 + 11 * per_row * (((uint64_t)hd.picture_height(i) + 15) / 16);
  if ((uint64_t)hd.h.frame_size < need)
    return invalid(fmtstr("frame_size %u is too small for a %u%s%u frame (at least %llu bytes)",
                          hd.h.frame_size, hd.h.width, TIMES, hd.h.height,
                          (unsigned long long)need));
  return Err{};
}

Err parse_picture_header(const uint8_t *d, size_t len, PicHeader &out) {
  // src/header.rs:269-290
  if (len < 8) return invalid("a picture header runs past the end of the frame");
  out.header_size = d[0] >> 3;
  out.picture_size = be32(d + 1);
  if (out.header_size < 8 || out.header_size > len)
    return invalid(fmtstr("picture_header_size %u is out of range", out.header_size));
  if (out.picture_size < out.header_size || out.picture_size > len)
    return invalid(fmtstr("picture_size %u is out of range", out.picture_size));
  out.log2_slice_mbs = (d[7] >> 4) & 3;
  return Err{};
}

// ---------------------------------------------------------------------------
// Constant tables. Codebooks as combination codes (lastRiceQ, kRice, kExp);
// an exp-Golomb code of order k is (0, k, k+1) (RDD 36 §7.1.1.1).
namespace {
struct Book {
  uint32_t lq, kr, ke;
};
constexpr Book RUN_BOOKS[16] = {{2, 0, 1}, {2, 0, 1}, {1, 0, 1}, {1, 0, 1}, {0, 0, 1}, {1, 1, 2},
                            {1, 1, 2}, {1, 1, 2}, {1, 1, 2}, {0, 1, 2}, {0, 1, 2}, {0, 1, 2},
                            {0, 1, 2}, {0, 1, 2}, {0, 1, 2}, {0, 2, 3}};
constexpr Book LEVEL_BOOKS[9] = {{2, 0, 2}, {1, 0, 1}, {2, 0, 1}, {0, 0, 1}, {0, 1, 2},
                             {0, 1, 2}, {0, 1, 2}, {0, 1, 2}, {0, 2, 3}};
uint32_t bitlen(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }
// (bits, length) of symbol n.
void code_of(Book b, uint32_t n, uint64_t &bits, uint32_t &len) {
  uint64_t th = (uint64_t)(b.lq + 1) << b.kr;
  if (n < th) {
    bits = (1ull << b.kr) | (n & ((1ull << b.kr) - 1));
    len = (uint32_t)(n >> b.kr) + 1 + b.kr;
  } else {
    uint64_t v = n - th + (1ull << b.ke);
    bits = v;
    len = b.lq + 1 + 2 * bitlen(v) - b.ke - 1;
  }
}
// The distinct codebooks of the run and level contexts (RUN_BOOKS and
// LEVEL_BOOKS hold each context's), for the entropy decode's tables.
constexpr Book FAST_RUN_BOOKS[6] = {{2, 0, 1}, {1, 0, 1}, {0, 0, 1}, {1, 1, 2}, {0, 1, 2}, {0, 2, 3}};
constexpr Book FAST_LEVEL_BOOKS[6] = {{2, 0, 2}, {1, 0, 1}, {2, 0, 1}, {0, 0, 1}, {0, 1, 2}, {0, 2, 3}};
constexpr uint8_t RUN_CTX_BOOK[16] = {0, 0, 1, 1, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 5};
constexpr uint8_t LEVEL_CTX_BOOK[9] = {0, 1, 2, 3, 4, 4, 4, 4, 5};
constexpr bool same_book(Book a, Book b) { return a.lq == b.lq && a.kr == b.kr && a.ke == b.ke; }
constexpr bool fast_books_match() {
  for (int c = 0; c < 16; c++)
    if (RUN_CTX_BOOK[c] > 5 || !same_book(RUN_BOOKS[c], FAST_RUN_BOOKS[RUN_CTX_BOOK[c]])) return false;
  for (int c = 0; c < 9; c++)
    if (LEVEL_CTX_BOOK[c] > 5 || !same_book(LEVEL_BOOKS[c], FAST_LEVEL_BOOKS[LEVEL_CTX_BOOK[c]])) return false;
  return true;
}
static_assert(fast_books_match(), "each context's codebook must be its distinct book");
// Every prefix that starts a code of at most 9 bits (for levels, with its
// sign bit) holds that code's entry (layout.h).
void fill_fast(DevTables &t) {
  memset(t.fast_level, 0, sizeof t.fast_level);
  memset(t.fast_run, 0, sizeof t.fast_run);
  for (int b = 0; b < 6; b++)
    for (uint32_t n = 0; n < 4096; n++) {
      uint64_t bits;
      uint32_t len;
      code_of(FAST_RUN_BOOKS[b], n, bits, len);
      if (len <= 9) {
        if (n >= 64) abort();  // 6 bits of run; no code this short has more
        const uint32_t first = (uint32_t)bits << (9 - len);
        for (uint32_t p = 0; p < (1u << (9 - len)); p++)
          t.fast_run[b][first + p] = (uint16_t)(FAST_VALID | RUN_CTX_BOOK[n < 15 ? n : 15] << 10 | n << 4 | len);
      }
      code_of(FAST_LEVEL_BOOKS[b], n, bits, len);
      if (len <= 9) {
        if (n + 1 > 63) abort();  // a 7-bit signed level; no code this short has more
        for (uint32_t sgn = 0; sgn < 2; sgn++) {
          const int32_t level = sgn ? -(int32_t)(n + 1) : (int32_t)(n + 1);
          const uint32_t first = (uint32_t)(bits << 1 | sgn) << (9 - len);
          for (uint32_t p = 0; p < (1u << (9 - len)); p++)
            t.fast_level[b][first + p] =
                (uint16_t)(FAST_VALID | LEVEL_CTX_BOOK[n < 8 ? n : 8] << 11 | ((uint32_t)level & 127) << 4 | (len + 1));
        }
      }
    }
}
const uint8_t SCAN_P[64] = {0,  1,  4,  5,  16, 17, 21, 22, 2,  3,  6,  7,  18, 20, 23, 28,
                            8,  9,  12, 13, 19, 24, 27, 29, 10, 11, 14, 15, 25, 26, 30, 31,
                            32, 33, 37, 38, 45, 46, 53, 54, 34, 36, 39, 44, 47, 52, 55, 60,
                            35, 40, 43, 48, 51, 56, 59, 61, 41, 42, 49, 50, 57, 58, 62, 63};
}  // namespace

void build_tables(DevTables &t) {
  // The DCT basis T[x][u] = C(u)/2 cos((2x+1)u pi/16) in double, rounded to
  // f32 (src/dct.rs:11-25). Every value is far from an f32 rounding tie.
  const double PI = 3.14159265358979323846;
  for (int x = 0; x < 8; x++)
    for (int u = 0; u < 8; u++) {
      double c = u == 0 ? 0.70710678118654752440 : 1.0;
      t.basis[8 * x + u] = (float)(c / 2.0 * cos((double)((2 * x + 1) * u) * PI / 16.0));
    }
  // Figures 4 and 5 of RDD 36: the interlaced scan is the progressive one
  // transposed (src/tables.rs:6-26).
  for (int v = 0; v < 8; v++)
    for (int u = 0; u < 8; u++) {
      t.scan[0][8 * v + u] = SCAN_P[8 * v + u];
      t.scan[1][8 * u + v] = SCAN_P[8 * v + u];
    }
  fill_fast(t);
  for (int c = 0; c < 16; c++)
    for (uint32_t n = 0; n < 64; n++) {
      uint64_t bits;
      uint32_t len;
      code_of(RUN_BOOKS[c], n, bits, len);
      t.enc_run_len[c][n] = (uint8_t)len;
    }
  for (int c = 0; c < 9; c++)
    for (uint32_t n = 0; n < 256; n++) {
      uint64_t bits;
      uint32_t len;
      code_of(LEVEL_BOOKS[c], n, bits, len);
      t.enc_level_len[c][n] = (uint8_t)len;
    }
}

// ---------------------------------------------------------------------------
Err DevBuf::reserve(size_t bytes) {
  if (bytes <= cap) return Err{};
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  size_t want = bytes + bytes / 4 + 256;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) return device_error("hipMalloc", e);
  cap = want;
  return Err{};
}

// Streams for copies outside a lane: non-blocking, default priority, kept in
// an idle list per device, so copies from many threads run at once and none
// goes through the legacy stream (whose implicit synchronisation with other
// streams can invalidate a graph capture running on a lane).
namespace {
struct CopyStreams {
  std::mutex mu;
  std::vector<hipStream_t> idle[64];
};
CopyStreams &copy_streams() {
  static CopyStreams *c = new CopyStreams();  // never destroyed: copies may run at exit
  return *c;
}
}  // namespace

Err copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  if (!bytes) return Err{};
  int id = 0;
  hipError_t he = hipGetDevice(&id);
  if (he != hipSuccess) return device_error("hipGetDevice", he);
  if (id < 0 || id >= 64) return Err{PRORES_ERR_DEVICE, "device index out of range"};
  CopyStreams &cs = copy_streams();
  hipStream_t s = nullptr;
  {
    std::lock_guard<std::mutex> g(cs.mu);
    if (!cs.idle[id].empty()) {
      s = cs.idle[id].back();
      cs.idle[id].pop_back();
    }
  }
  if (!s && (he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess)
    return device_error("hipStreamCreate(copy)", he);
  he = hipMemcpyAsync(dst, src, bytes, kind, s);
  const hipError_t se = hipStreamSynchronize(s);
  {
    std::lock_guard<std::mutex> g(cs.mu);
    cs.idle[id].push_back(s);
  }
  if (he != hipSuccess) return device_error("hipMemcpyAsync", he);
  if (se != hipSuccess) return device_error("hipStreamSynchronize(copy)", se);
  return Err{};
}

Err copy_on(Lane &L, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  if (!bytes) return Err{};
  Err e = check(hipMemcpyAsync(dst, src, bytes, kind, L.stream), "hipMemcpyAsync");
  Err s = check(hipStreamSynchronize(L.stream), "hipStreamSynchronize");
  return e.ok() ? s : e;
}

Err HostBuf::reserve(size_t bytes) {
  if (bytes <= cap) return Err{};
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  const size_t want = bytes + bytes / 4 + 4096;
  const hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  if (e != hipSuccess) return device_error("hipHostMalloc", e);
  cap = want;
  return Err{};
}

Err launch_on(hipStream_t s, hipFunction_t f, unsigned gx, unsigned gy, unsigned bx, void **args,
              unsigned lds) {
  if (gx == 0 || gy == 0) return Err{};
  return check(hipModuleLaunchKernel(f, gx, gy, 1, bx, 1, 1, lds, s, args, nullptr), "hipModuleLaunchKernel");
}

Err launch(Lane &L, hipFunction_t f, unsigned gx, unsigned gy, unsigned bx, void **args, unsigned lds) {
  return launch_on(L.stream, f, gx, gy, bx, args, lds);
}

static Err load_functions(hipModule_t m, const char *const *names, hipFunction_t *const *out,
                          int n) {
  for (int i = 0; i < n; i++) {
    Err e = check(hipModuleGetFunction(out[i], m, names[i]), names[i]);
    if (!e.ok()) return e;
  }
  return Err{};
}

static Err init_device(Device &d, int id) {
  Err e;
  d.id = id;
  if (!(e = check(hipModuleLoadData(&d.dec_mod, prores_decode_co), "hipModuleLoadData(decode)")).ok())
    return e;
  const char *const names[] = {"pd_slice_table", "pd_entropy", "pd_alpha", "pd_idct", "pd_plan",
                               "pd_entropy_order"};
  hipFunction_t *const fns[] = {&d.k_slice_table, &d.k_entropy, &d.k_alpha, &d.k_idct, &d.k_plan, &d.k_order};
  if (!(e = load_functions(d.dec_mod, names, fns, 6)).ok()) return e;
  DevTables host;
  build_tables(host);
  if (!(e = check(hipMalloc(&d.tables, sizeof(DevTables)), "hipMalloc(tables)")).ok()) return e;
  if (!(e = copy_sync(d.tables, &host, sizeof host, hipMemcpyHostToDevice)).ok()) return e;
  prores_decode_tuning_default(&d.tuning);
  return Err{};
}

Err lane(Device &d, int index, Lane *&out) {
  if (index < 0 || index >= MAX_LANES) return Err{PRORES_ERR_DEVICE, "lane index out of range"};
  std::lock_guard<std::mutex> g(d.mu);
  if (!d.lanes[index]) {
    Lane *L = new Lane();
    L->index = index;
    Err e = check(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (e.ok()) e = check(hipStreamCreateWithFlags(&L->aux, hipStreamNonBlocking), "hipStreamCreate");
    for (auto &ev : L->ev)
      if (e.ok()) e = check(hipEventCreate(&ev), "hipEventCreate");
    for (auto &ev : L->eev)
      if (e.ok()) e = check(hipEventCreate(&ev), "hipEventCreate");
    if (e.ok()) e = check(hipEventCreateWithFlags(&L->fork, hipEventDisableTiming), "hipEventCreate");
    if (e.ok()) e = check(hipEventCreateWithFlags(&L->join, hipEventDisableTiming), "hipEventCreate");
    if (!e.ok()) {
      for (auto &ev : L->ev)
        if (ev) (void)hipEventDestroy(ev);
      for (auto &ev : L->eev)
        if (ev) (void)hipEventDestroy(ev);
      if (L->fork) (void)hipEventDestroy(L->fork);
      if (L->join) (void)hipEventDestroy(L->join);
      if (L->aux) (void)hipStreamDestroy(L->aux);
      if (L->stream) (void)hipStreamDestroy(L->stream);
      delete L;
      return e;
    }
    d.lanes[index] = L;
  }
  out = d.lanes[index];
  return Err{};
}

Err ensure_encoder(Device &d) {
  // The encoder's code object is loaded on the first encode.
  std::lock_guard<std::mutex> g(d.mu);
  if (d.enc_mod) return Err{};
  Err e = check(hipModuleLoadData(&d.enc_mod, prores_encode_co), "hipModuleLoadData(encode)");
  if (!e.ok()) return e;
  const char *const names[] = {"pe_fdct",     "pe_alpha",   "pe_size", "pe_search", "pe_window",   "pe_shareout",
                               "pe_offsets", "pe_pack",    "pe_arena", "pe_slices", "pe_size_first", "pe_fdct64"};
  hipFunction_t *const fns[] = {&d.k_fdct,    &d.k_alpha_enc, &d.k_size,  &d.k_search, &d.k_window,
                                &d.k_shareout, &d.k_offsets,  &d.k_pack,  &d.k_arena,  &d.k_slices,
                                &d.k_size_first, &d.k_fdct64};
  e = load_functions(d.enc_mod, names, fns, 12);
  if (!e.ok()) {
    (void)hipModuleUnload(d.enc_mod);
    d.enc_mod = nullptr;
  }
  return e;
}

Err device(Device *&out) {
  static std::mutex mu;
  static Device *devices[64] = {nullptr};
  int id = 0;
  hipError_t he = hipGetDevice(&id);
  if (he != hipSuccess) return device_error("hipGetDevice", he);
  if (id < 0 || id >= 64) return Err{PRORES_ERR_DEVICE, "device index out of range"};
  std::lock_guard<std::mutex> g(mu);
  if (!devices[id]) {
    Device *d = new Device();
    Err e = init_device(*d, id);
    if (!e.ok()) {
      delete d;
      return e;
    }
    devices[id] = d;
  }
  out = devices[id];
  return Err{};
}

}  // namespace prores_hip

using namespace prores_hip;

extern "C" {

void *prores_device_alloc(size_t bytes) {
  void *p = nullptr;
  return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? p : nullptr;
}

void prores_device_free(void *p) {
  if (p) (void)hipFree(p);
}

size_t prores_find_frames(const uint8_t *data, size_t len, uint64_t *offsets, uint64_t *sizes, size_t max) {
  // tests/sample.rs:26-41: a frame is its BE u32 frame_size, 'icpf', at
  // least 28 bytes and wholly inside the data; the scan goes on after it.
  size_t count = 0, i = 4;
  while (i + 4 <= len) {
    if (data[i] == 'i' && data[i + 1] == 'c' && data[i + 2] == 'p' && data[i + 3] == 'f') {
      const size_t size = ((size_t)data[i - 4] << 24) | ((size_t)data[i - 3] << 16) | ((size_t)data[i - 2] << 8) |
                          data[i - 1];
      if (size >= 28 && size <= len - (i - 4)) {
        if (count < max) {
          offsets[count] = i - 4;
          sizes[count] = size;
        }
        count++;
        i += size - 4;
        continue;
      }
    }
    i++;
  }
  return count;
}

uint64_t prores_fnv1a(uint64_t h, const void *data, size_t len) {
  const uint8_t *p = (const uint8_t *)data;
  for (size_t i = 0; i < len; i++) h = (h ^ p[i]) * 0x100000001b3ull;
  return h;
}

static thread_local char last_memcpy_error[160];

prores_status prores_memcpy(void *dst, const void *src, size_t bytes, int32_t kind) {
  const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                          : kind == 2 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
  const Err e = copy_sync(dst, src, bytes, k);
  if (e.ok()) return PRORES_OK;
  snprintf(last_memcpy_error, sizeof last_memcpy_error, "copy of %zu bytes: %s", bytes, e.msg.c_str());
  return PRORES_ERR_DEVICE;
}

const char *prores_memcpy_error(void) { return last_memcpy_error; }

}  // extern "C"
