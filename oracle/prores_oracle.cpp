// TEST INFRASTRUCTURE ONLY — the CPU oracle for the HIP ProRes slice engine.
//
// A C++ restatement of the rivet-prores crate (a Rust crate; no Rust toolchain
// exists here, so the reference itself cannot be compiled or run). Each part
// cites the reference file:line it follows. It is the checker: the product
// library never links it, and only tests, the smoke check and the benchmark's
// CPU-baseline leg load it.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off: the reference's f32
// arithmetic is a multiply then an add, never fused, src/dsp/mod.rs:11-17).

#include "prores_oracle.h"

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <thread>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------------------
// Errors (src/error.rs:5-42). The decoder reports malformed input as a value,
// never a crash; here an exception carries it to the C boundary.
struct Error {
  int kind;
  std::string msg;
};
[[noreturn]] static void invalid(const std::string &m) { throw Error{ORACLE_INVALID, m}; }
[[noreturn]] static void unsupported(const std::string &m) { throw Error{ORACLE_UNSUPPORTED, m}; }
[[noreturn]] static void config_err(const std::string &m) { throw Error{ORACLE_CONFIG, m}; }

static std::string fmt(const char *f, ...) __attribute__((format(printf, 1, 2)));
static std::string fmt(const char *f, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

static const char *TIMES = "\xC3\x97";  // U+00D7, as the reference's messages write sizes
static const char *EN_DASH = "\xE2\x80\x93";

static inline uint64_t be64(const uint8_t *p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v = (v << 8) | p[i];
  return v;
}
static inline uint32_t clz64(uint64_t v) { return v ? (uint32_t)__builtin_clzll(v) : 64; }

// ---------------------------------------------------------------------------
// Thread pool (src/pool.rs). Process-wide workers, `parallelism − 1` of them
// plus the caller; tasks claimed `chunk` at a time from an atomic counter:
// 1 for decode rows (TaskSize::Large), clamp(tasks/(4·threads), 1, 8) for
// encode slices (TaskSize::Small), src/pool.rs:179-183.
static size_t parallelism() {
  static size_t n = std::max<size_t>(1, std::thread::hardware_concurrency());
  return n;
}
static size_t resolve(size_t threads) { return threads == 0 ? parallelism() : threads; }

class Pool {
 public:
  struct Job {
    const std::function<void(size_t)> *f;
    size_t tasks, chunk;
    alignas(128) std::atomic<size_t> next{0};
    size_t max_helpers;
    size_t joined = 0, left = 0;
    std::exception_ptr error;
    void work() {
      for (;;) {
        size_t first = next.fetch_add(chunk, std::memory_order_relaxed);
        if (first >= tasks) return;
        size_t end = std::min(first + chunk, tasks);
        for (size_t i = first; i < end; i++) (*f)(i);
      }
    }
  };
  static Pool &get() {
    static Pool *p = new Pool();
    return *p;
  }
  void run(size_t threads, size_t tasks, bool small, const std::function<void(size_t)> &f) {
    size_t helpers = std::min(resolve(threads), tasks);
    helpers = helpers > 0 ? helpers - 1 : 0;
    if (helpers == 0) {
      for (size_t i = 0; i < tasks; i++) f(i);
      return;
    }
    auto job = std::make_shared<Job>();
    job->f = &f;
    job->tasks = tasks;
    job->chunk = small ? std::clamp<size_t>(tasks / (4 * (helpers + 1)), 1, 8) : 1;
    job->max_helpers = helpers;
    {
      std::lock_guard<std::mutex> g(mu_);
      queue_.push_back(job);
    }
    cv_.notify_all();
    std::exception_ptr mine;
    try {
      job->work();
    } catch (...) {
      mine = std::current_exception();
      job->next.store(tasks);
    }
    std::unique_lock<std::mutex> g(mu_);
    queue_.erase(std::remove(queue_.begin(), queue_.end(), job), queue_.end());
    done_.wait(g, [&] { return job->left >= job->joined; });
    if (mine) std::rethrow_exception(mine);
    if (job->error) std::rethrow_exception(job->error);
  }

 private:
  Pool() {
    for (size_t i = 0; i + 1 < parallelism(); i++) std::thread([this] { worker(); }).detach();
  }
  void worker() {
    std::unique_lock<std::mutex> g(mu_);
    for (;;) {
      std::shared_ptr<Job> job;
      for (auto &j : queue_)
        if (j->joined < j->max_helpers && j->next.load(std::memory_order_relaxed) < j->tasks) {
          job = j;
          break;
        }
      if (!job) {
        cv_.wait(g);
        continue;
      }
      job->joined++;
      g.unlock();
      std::exception_ptr err;
      try {
        job->work();
      } catch (...) {
        err = std::current_exception();
        job->next.store(job->tasks);
      }
      g.lock();
      if (err && !job->error) job->error = err;
      job->left++;
      done_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, done_;
  std::vector<std::shared_ptr<Job>> queue_;
};

// ---------------------------------------------------------------------------
// Bit reading and writing, MSB first (src/bits.rs).
struct BitReader {  // src/bits.rs:10-102
  const uint8_t *d;
  size_t len;
  size_t pos = 0;
  BitReader(const uint8_t *data, size_t n) : d(data), len(n) {}
  size_t remaining() const { return len * 8 - pos; }
  uint64_t peek64() const {
    size_t byte = pos >> 3;
    uint8_t buf[8] = {0};
    if (byte + 8 <= len)
      memcpy(buf, d + byte, 8);
    else if (byte < len)
      memcpy(buf, d + byte, len - byte);
    return be64(buf) << (pos & 7);
  }
  bool peek_fast(uint64_t &w) const {
    size_t byte = pos >> 3;
    if (byte + 8 > len) return false;
    w = be64(d + byte) << (pos & 7);
    return true;
  }
  void skip(uint32_t n) { pos += n; }
  uint64_t read(uint32_t n) {
    if (n == 0) return 0;
    if (n > remaining()) invalid("a codeword runs past the end of its data");
    uint64_t v = peek64() >> (64 - n);
    pos += n;
    return v;
  }
  uint32_t read_bit() { return (uint32_t)read(1); }
  uint32_t leading_zeros(uint32_t max) {
    uint32_t lz = clz64(peek64());
    if (lz > max) invalid("a variable-length code's prefix is too long");
    if (lz >= remaining()) invalid("a codeword runs past the end of its data");
    pos += lz;
    return lz;
  }
  bool end_of_data() const {
    size_t r = remaining();
    return r <= 31 && (r == 0 || (peek64() >> (64 - r)) == 0);
  }
};

struct BitWriter {  // src/bits.rs:138-202 (a plain accumulator; same bits)
  std::vector<uint8_t> *out;
  size_t start;
  uint64_t acc = 0;
  uint32_t nacc = 0;
  explicit BitWriter(std::vector<uint8_t> *o) : out(o), start(o->size()) {}
  void put(uint64_t v, uint32_t n) {
    if (n == 0) return;
    v &= (n == 64) ? ~0ull : ((1ull << n) - 1);
    uint32_t free_ = 64 - nacc;
    if (n < free_) {
      acc = (acc << n) | v;
      nacc += n;
    } else {
      uint32_t rest = n - free_;
      uint64_t full = (free_ == 64 ? 0 : (acc << free_)) | (v >> rest);
      for (int i = 7; i >= 0; i--) out->push_back((uint8_t)(full >> (8 * i)));
      acc = rest ? (v & ((1ull << rest) - 1)) : 0;
      nacc = rest;
    }
  }
  size_t bits() const { return (out->size() - start) * 8 + nacc; }
  void align() {
    size_t r = bits() % 8;
    if (r) put(0, 8 - (uint32_t)r);
  }
  size_t finish() {
    align();
    if (nacc) {
      uint64_t a = acc << (64 - nacc);
      for (uint32_t i = 0; i < nacc / 8; i++) out->push_back((uint8_t)(a >> (56 - 8 * i)));
      acc = 0;
      nacc = 0;
    }
    return out->size() - start;
  }
};

struct BitCounter {
  size_t n = 0;
  void put(uint64_t, uint32_t k) { n += k; }
  size_t bits() const { return n; }
};

// ---------------------------------------------------------------------------
// Variable-length codes (src/vlc.rs).
static const uint32_t MAX_PREFIX = 32;       // src/vlc.rs:23
static const uint32_t MAX_SYMBOL = 1u << 30;  // src/vlc.rs:26

struct Codebook {  // src/vlc.rs:11-15: ExpGolomb(k) or Combo(lastQ, kRice, kExp)
  bool eg;
  uint32_t a, b, c;
  // The combination-code parameters (an EG(k) is Combo(0, k, k+1), src/vlc.rs:181-186)
  uint32_t last_q() const { return eg ? 0 : a; }
  uint32_t k_rice() const { return eg ? a : b; }
  uint32_t k_exp() const { return eg ? a + 1 : c; }
};
static constexpr Codebook EG(uint32_t k) { return Codebook{true, k, 0, 0}; }
static constexpr Codebook CB(uint32_t q, uint32_t r, uint32_t e) { return Codebook{false, q, r, e}; }

static const Codebook DC_CODEBOOKS[4] = {EG(0), EG(1), CB(1, 2, 3), EG(3)};  // src/vlc.rs:30-31
static const Codebook RUN_CODEBOOKS[16] = {                                  // src/vlc.rs:34-51
    CB(2, 0, 1), CB(2, 0, 1), CB(1, 0, 1), CB(1, 0, 1), EG(0), CB(1, 1, 2),
    CB(1, 1, 2), CB(1, 1, 2), CB(1, 1, 2), EG(1),       EG(1), EG(1),
    EG(1),       EG(1),       EG(1),       EG(2)};
static const Codebook LEVEL_CODEBOOKS[9] = {  // src/vlc.rs:55-65
    CB(2, 0, 2), CB(1, 0, 1), CB(2, 0, 1), EG(0), EG(1), EG(1), EG(1), EG(1), EG(2)};
static const Codebook FIRST_DC_CODEBOOK = EG(5);  // src/vlc.rs:68

static inline Codebook dc_codebook(int32_t prev_diff) {  // src/vlc.rs:70-73
  uint32_t a = prev_diff < 0 ? (uint32_t)(-(int64_t)prev_diff) : (uint32_t)prev_diff;
  return DC_CODEBOOKS[std::min<uint32_t>(a, 3)];
}
static inline Codebook run_codebook(uint32_t prev) { return RUN_CODEBOOKS[std::min<uint32_t>(prev, 15)]; }
static inline Codebook level_codebook(uint32_t prev) { return LEVEL_CODEBOOKS[std::min<uint32_t>(prev, 8)]; }

static inline uint32_t bit_len(uint64_t v) { return 64 - clz64(v); }

template <class S>
static void put_exp_golomb(S &s, uint64_t n, uint32_t k) {  // src/vlc.rs:93-104
  uint64_t v = n + (1ull << k);
  uint32_t len = bit_len(v);
  uint32_t q = len - k - 1;
  if (len + q <= 57) {
    s.put(v, len + q);
  } else {
    s.put(0, q);
    s.put(v, len);
  }
}

template <class S>
static void cb_put(Codebook cb, S &s, uint32_t n32) {  // src/vlc.rs:108-126
  uint64_t n = n32;
  if (cb.eg) return put_exp_golomb(s, n, cb.a);
  uint64_t threshold = (uint64_t)(cb.a + 1) << cb.b;
  if (n < threshold) {
    uint32_t q = (uint32_t)(n >> cb.b);
    uint64_t r = n & ((1ull << cb.b) - 1);
    s.put((1ull << cb.b) | r, q + 1 + cb.b);
  } else {
    s.put(0, cb.a + 1);
    put_exp_golomb(s, n - threshold, cb.c);
  }
}

// Codebook::code (src/vlc.rs:131-151): (bits, length) when length <= 57.
static bool cb_code(Codebook cb, uint32_t n32, uint64_t &bits, uint32_t &len) {
  uint32_t last_q = cb.last_q(), k_rice = cb.k_rice(), k_exp = cb.k_exp();
  uint64_t n = n32;
  uint64_t threshold = (uint64_t)(last_q + 1) << k_rice;
  if (n < threshold) {
    uint32_t q = (uint32_t)(n >> k_rice);
    bits = (1ull << k_rice) | (n & ((1ull << k_rice) - 1));
    len = q + 1 + k_rice;
    return true;
  }
  uint64_t v = n - threshold + (1ull << k_exp);
  uint32_t b = bit_len(v);
  len = last_q + 1 + 2 * b - k_exp - 1;
  bits = v;
  return len <= 57;
}

static uint32_t cb_len(Codebook cb, uint32_t n32) {  // src/vlc.rs:155-172
  auto eg = [](uint64_t n, uint32_t k) -> uint32_t {
    uint32_t len = bit_len(n + (1ull << k));
    return 2 * len - k - 1;
  };
  uint64_t n = n32;
  if (cb.eg) return eg(n, cb.a);
  uint64_t threshold = (uint64_t)(cb.a + 1) << cb.b;
  if (n < threshold) return (uint32_t)(n >> cb.b) + 1 + cb.b;
  return cb.a + 1 + eg(n - threshold, cb.c);
}

// Codebook::peek (src/vlc.rs:179-204): decode from a 64-bit window whose
// first 57 bits are real; false when longer than 57 bits or invalid.
static inline bool cb_peek(Codebook cb, uint64_t w, uint32_t &sym, uint32_t &len) {
  uint32_t last_q = cb.last_q(), k_rice = cb.k_rice(), k_exp = cb.k_exp();
  uint32_t q = clz64(w);
  if (q <= last_q) {
    len = q + 1 + k_rice;
    uint64_t v = (w >> (64 - len)) & ((1ull << k_rice) - 1);
    sym = (uint32_t)(((uint64_t)q << k_rice) | v);
    return true;
  }
  if (q > MAX_PREFIX) return false;
  uint32_t tail = q - (last_q + 1) + k_exp + 1;
  len = q + tail;
  if (len > 57) return false;
  uint64_t v = (w << q) >> (64 - tail);
  uint64_t n = v - (1ull << k_exp) + ((uint64_t)(last_q + 1) << k_rice);
  if (n > MAX_SYMBOL) return false;
  sym = (uint32_t)n;
  return true;
}

static uint32_t cb_read(Codebook cb, BitReader &r) {  // src/vlc.rs:206-230
  uint32_t q = r.leading_zeros(MAX_PREFIX);
  uint32_t q_exp, k_exp;
  uint64_t offset;
  if (cb.eg) {
    q_exp = q;
    k_exp = cb.a;
    offset = 0;
  } else {
    if (q <= cb.a) {
      uint64_t v = r.read(1 + cb.b) & ((1ull << cb.b) - 1);
      return (uint32_t)(((uint64_t)q << cb.b) | v);
    }
    q_exp = q - (cb.a + 1);
    k_exp = cb.c;
    offset = (uint64_t)(cb.a + 1) << cb.b;
  }
  uint64_t v = r.read(q_exp + k_exp + 1);
  uint64_t n = v - (1ull << k_exp) + offset;
  if (n > MAX_SYMBOL) invalid("a variable-length code's value is out of range");
  return (uint32_t)n;
}

static inline uint32_t signed_to_symbol(int32_t n) {  // src/vlc.rs:233-241
  return n >= 0 ? 2u * (uint32_t)n : 2u * (uint32_t)(-(int64_t)n) - 1;
}
static inline int32_t symbol_to_signed(uint32_t s) {  // src/vlc.rs:243-251
  return (s & 1) == 0 ? (int32_t)(s >> 1) : -(int32_t)((s + 1) / 2);
}

template <class S>
static void put_alpha_run(S &s, uint32_t run) {  // src/vlc.rs:254-261
  if (run == 1)
    s.put(1, 1);
  else if (run <= 16)
    s.put(run - 1, 5);
  else
    s.put(run - 1, 16);
}
static uint32_t read_alpha_run(BitReader &r) {  // src/vlc.rs:263-272
  if (r.read_bit() == 1) return 1;
  uint32_t v = (uint32_t)r.read(4);
  if (v != 0) return v + 1;
  return (uint32_t)r.read(11) + 1;
}
template <class S>
static void put_alpha_difference(S &s, int32_t diff, uint32_t bits) {  // src/vlc.rs:277-287
  uint32_t mag_bits = bits == 8 ? 3 : 6, max = bits == 8 ? 8 : 64;
  uint32_t a = diff < 0 ? (uint32_t)(-(int64_t)diff) : (uint32_t)diff;
  if (a >= 1 && a <= max) {
    uint64_t code = ((uint64_t)(a - 1) << 1) | (diff < 0 ? 1u : 0u);
    s.put(code, 1 + mag_bits + 1);
  } else {
    uint32_t m = (uint32_t)diff & ((1u << bits) - 1);
    s.put((1ull << bits) | m, 1 + bits);
  }
}
static int32_t read_alpha_difference(BitReader &r, uint32_t bits) {  // src/vlc.rs:291-299
  uint32_t mag_bits = bits == 8 ? 3 : 6;
  if (r.read_bit() == 1) return (int32_t)r.read(bits);
  int32_t a = (int32_t)r.read(mag_bits) + 1;
  bool neg = r.read_bit() == 1;
  return neg ? -a : a;
}

// Decode tables on 9-bit prefixes (src/vlc.rs:301-366).
static const uint32_t DECODE_BITS = 9;
struct DecodeTables {
  uint16_t run[16][512];
  uint16_t level[9][512];
  uint32_t run_codes[16][32];
  uint32_t level_codes[9][32];
  DecodeTables() {
    auto fill = [](const Codebook *books, int nb, uint16_t (*t)[512]) {
      for (int c = 0; c < nb; c++) {
        memset(t[c], 0, sizeof(uint16_t) * 512);
        for (uint32_t n = 0;; n++) {
          uint32_t len = cb_len(books[c], n);
          if (len <= DECODE_BITS) {
            uint64_t bits;
            uint32_t l2;
            cb_code(books[c], n, bits, l2);
            uint32_t first = (uint32_t)bits << (DECODE_BITS - len);
            for (uint32_t p = 0; p < (1u << (DECODE_BITS - len)); p++)
              t[c][first + p] = (uint16_t)((n << 5) | len);
          } else if (n > (1u << DECODE_BITS)) {
            break;
          }
        }
      }
    };
    fill(RUN_CODEBOOKS, 16, run);
    fill(LEVEL_CODEBOOKS, 9, level);
    auto codes = [](const Codebook *books, int nb, uint32_t (*t)[32]) {  // src/vlc.rs:374-390
      for (int c = 0; c < nb; c++)
        for (uint32_t n = 0; n < 32; n++) {
          uint64_t bits;
          uint32_t len;
          cb_code(books[c], n, bits, len);
          t[c][n] = ((uint32_t)bits << 5) | len;
        }
    };
    codes(RUN_CODEBOOKS, 16, run_codes);
    codes(LEVEL_CODEBOOKS, 9, level_codes);
  }
};
static const DecodeTables &tables() {
  static DecodeTables t;
  return t;
}

static inline bool peek_run(uint32_t prev_run, uint64_t w, uint32_t &sym, uint32_t &len) {
  uint32_t c = std::min<uint32_t>(prev_run, 15);
  uint16_t e = tables().run[c][w >> (64 - DECODE_BITS)];
  if (e) {
    sym = e >> 5;
    len = e & 31;
    return true;
  }
  return cb_peek(RUN_CODEBOOKS[c], w, sym, len);
}
static inline bool peek_level(uint32_t prev, uint64_t w, uint32_t &sym, uint32_t &len) {
  uint32_t c = std::min<uint32_t>(prev, 8);
  uint16_t e = tables().level[c][w >> (64 - DECODE_BITS)];
  if (e) {
    sym = e >> 5;
    len = e & 31;
    return true;
  }
  return cb_peek(LEVEL_CODEBOOKS[c], w, sym, len);
}
static inline uint32_t run_len(uint32_t prev_run, uint32_t run) {  // src/vlc.rs:396-404
  uint32_t c = std::min<uint32_t>(prev_run, 15);
  if (run < 32) return tables().run_codes[c][run] & 31;
  return cb_len(RUN_CODEBOOKS[c], run);
}
static inline uint32_t level_len(uint32_t prev, uint32_t sym) {  // src/vlc.rs:408-415
  uint32_t c = std::min<uint32_t>(prev, 8);
  if (sym < 32) return tables().level_codes[c][sym] & 31;
  return cb_len(LEVEL_CODEBOOKS[c], sym);
}
static inline bool run_code(uint32_t prev_run, uint32_t run, uint64_t &bits, uint32_t &len) {
  uint32_t c = std::min<uint32_t>(prev_run, 15);
  if (run < 32) {
    uint32_t e = tables().run_codes[c][run];
    bits = e >> 5;
    len = e & 31;
    return true;
  }
  return cb_code(RUN_CODEBOOKS[c], run, bits, len);
}
static inline bool level_code(uint32_t prev, uint32_t sym, uint64_t &bits, uint32_t &len) {
  uint32_t c = std::min<uint32_t>(prev, 8);
  if (sym < 32) {
    uint32_t e = tables().level_codes[c][sym];
    bits = e >> 5;
    len = e & 31;
    return true;
  }
  return cb_code(LEVEL_CODEBOOKS[c], sym, bits, len);
}

// ---------------------------------------------------------------------------
// Scans and quantiser scale (src/tables.rs).
static const uint8_t PROGRESSIVE_SCAN[64] = {
    0,  1,  4,  5,  16, 17, 21, 22, 2,  3,  6,  7,  18, 20, 23, 28, 8,  9,  12, 13, 19, 24,
    27, 29, 10, 11, 14, 15, 25, 26, 30, 31, 32, 33, 37, 38, 45, 46, 53, 54, 34, 36, 39, 44,
    47, 52, 55, 60, 35, 40, 43, 48, 51, 56, 59, 61, 41, 42, 49, 50, 57, 58, 62, 63};
static const uint8_t INTERLACED_SCAN[64] = {
    0,  2,  8,  10, 32, 34, 35, 41, 1,  3,  9,  11, 33, 36, 40, 42, 4,  6,  12, 14, 37, 39,
    43, 49, 5,  7,  13, 15, 38, 44, 48, 50, 16, 18, 19, 25, 45, 47, 51, 57, 17, 20, 24, 26,
    46, 52, 56, 58, 21, 23, 27, 30, 53, 55, 59, 62, 22, 28, 29, 31, 54, 60, 61, 63};
static inline uint32_t qscale(uint32_t q) { return q <= 128 ? q : 128 + 4 * (q - 128); }

// ---------------------------------------------------------------------------
// DCT (src/dct.rs) and the block kernels (src/dsp/scalar.rs).
struct Basis {
  float t[64];
  Basis() {  // src/dct.rs:11-25: C(u)/2 · cos((2x+1)uπ/16), in f64, then f32
    const double PI = 3.14159265358979323846;
    for (int x = 0; x < 8; x++)
      for (int u = 0; u < 8; u++) {
        double c = u == 0 ? 0.70710678118654752440 : 1.0;
        double a = (double)((2 * x + 1) * u) * PI / 16.0;
        t[8 * x + u] = (float)(c / 2.0 * std::cos(a));
      }
  }
};
static const float *basis() {
  static Basis b;
  return b.t;
}

static void idct(const float *coef, float *out) {  // src/dct.rs:36-66
  const float *t = basis();
  float tmp[64] = {0};
  for (int v = 0; v < 8; v++) {
    const float *row = coef + 8 * v;
    bool zero = true;
    for (int u = 0; u < 8; u++) zero = zero && row[u] == 0.0f;
    if (zero) continue;
    for (int x = 0; x < 8; x++) {
      const float *b = t + 8 * x;
      float s = 0.0f;
      for (int u = 0; u < 8; u++) s += row[u] * b[u];
      tmp[8 * v + x] = s;
    }
  }
  for (int k = 0; k < 64; k++) out[k] = 0.0f;
  for (int y = 0; y < 8; y++) {
    const float *b = t + 8 * y;
    for (int v = 0; v < 8; v++) {
      float w = b[v];
      for (int x = 0; x < 8; x++) out[8 * y + x] += w * tmp[8 * v + x];
    }
  }
}

static void fdct(const float *pix, float *out) {  // src/dct.rs:70-93
  const float *t = basis();
  float tmp[64] = {0};
  for (int y = 0; y < 8; y++)
    for (int x = 0; x < 8; x++) {
      float p = pix[8 * y + x];
      for (int u = 0; u < 8; u++) tmp[8 * y + u] += p * t[8 * x + u];
    }
  for (int k = 0; k < 64; k++) out[k] = 0.0f;
  for (int y = 0; y < 8; y++)
    for (int v = 0; v < 8; v++) {
      float w = t[8 * y + v];
      for (int u = 0; u < 8; u++) out[8 * v + u] += w * tmp[8 * y + u];
    }
}

struct Output {  // src/dsp/mod.rs:82-96
  float scale, bias, max;
  explicit Output(uint32_t bd)
      : scale((float)(1u << bd) / 512.0f),
        bias((float)(1u << (bd - 1)) + 0.5f),
        max((float)((1u << bd) - 1)) {}
};

static void idct_put(const int32_t *coef, const float *scale, Output o, uint16_t *res) {
  // src/dsp/scalar.rs:8-12
  float f[64], v[64];
  for (int k = 0; k < 64; k++) f[k] = (float)coef[k] * scale[k];
  idct(f, v);
  for (int k = 0; k < 64; k++) {
    float x = std::floor(v[k] * o.scale + o.bias);
    x = x < 0.0f ? 0.0f : (x > o.max ? o.max : x);
    res[k] = (uint16_t)x;
  }
}

static void fdct_load(const uint16_t *pix, float scale, float *out) {  // src/dsp/scalar.rs:14-17
  float p[64];
  for (int k = 0; k < 64; k++) p[k] = (float)pix[k] * scale - 256.0f;
  fdct(p, out);
}

static inline int32_t f32_to_i32_sat(float x) {  // Rust `as i32`: saturating, NaN → 0
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x < -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}

static void quantise(const float *f, size_t n, const float *inv, float ac_rounding, int32_t *out,
                     uint64_t *mask) {  // src/dsp/scalar.rs:19-39
  memset(mask, 0, n * sizeof(uint64_t));
  for (int s = 0; s < 64; s++) {
    float r = s == 0 ? 0.5f : ac_rounding;
    float iv = inv[s];
    for (size_t b = 0; b < n; b++) {
      size_t i = n * s + b;
      float x = f[i] * iv;
      int32_t m = f32_to_i32_sat(std::floor(std::fabs(x) + r));
      int32_t q = x < 0.0f ? -m : m;
      out[i] = q;
      mask[i / 64] |= (uint64_t)(q != 0) << (i % 64);
    }
  }
}

// ---------------------------------------------------------------------------
// Frame, headers and slice layout (src/frame.rs, src/header.rs).
struct Plane {
  size_t offset;
  uint32_t width, height;
  size_t len() const { return (size_t)width * height; }
};
struct Metadata {
  uint8_t aspect_ratio = 0, frame_rate_code = 0, color_primaries = 2, transfer_characteristic = 2,
          matrix_coefficients = 2;
};
struct Frame {
  uint32_t width = 0, height = 0, bit_depth = 0, chroma = 2, interlace = 0;
  Metadata metadata;
  std::vector<uint16_t> data;
  Plane planes[3];
  bool has_alpha = false;
  std::vector<uint16_t> alpha;
};

static void layout(uint32_t width, uint32_t height, uint32_t chroma, Plane p[3]) {  // src/frame.rs:247-266
  uint32_t sw = chroma == 2 ? 2 : 1;
  uint32_t cw = (width + sw - 1) / sw;
  p[0] = {0, width, height};
  p[1] = {p[0].len(), cw, height};
  p[2] = {p[1].offset + p[1].len(), cw, height};
}

static Frame frame_zeroed(uint32_t w, uint32_t h, uint32_t chroma, uint32_t bd) {  // src/frame.rs:218-245
  if (w == 0 || h == 0 || w > 65535 || h > 65535)
    config_err(fmt("a frame of %u%s%u cannot be coded (1%s65535 each way)", w, TIMES, h, EN_DASH));
  if (bd < 8 || bd > 16) config_err(fmt("bit depth %u is outside 8%s16", bd, EN_DASH));
  Frame f;
  f.width = w;
  f.height = h;
  f.bit_depth = bd;
  f.chroma = chroma;
  layout(w, h, chroma, f.planes);
  f.data.assign(f.planes[2].offset + f.planes[2].len(), 0);
  return f;
}

static const char *chroma_name(uint32_t c) { return c == 2 ? "Yuv422" : "Yuv444"; }

static void frame_validate(const Frame &f) {  // src/frame.rs:299-340
  if (f.width == 0 || f.height == 0 || f.width > 65535 || f.height > 65535)
    config_err(fmt("a frame of %u%s%u cannot be coded", f.width, TIMES, f.height));
  if (f.bit_depth < 8 || f.bit_depth > 16)
    config_err(fmt("bit depth %u is outside 8%s16", f.bit_depth, EN_DASH));
  Plane want[3];
  layout(f.width, f.height, f.chroma, want);
  for (int i = 0; i < 3; i++) {
    const Plane &p = f.planes[i];
    if (p.width != want[i].width || p.height != want[i].height)
      config_err(fmt("plane %d is %u%s%u; a %s frame of %u%s%u needs %u%s%u", i, p.width, TIMES,
                     p.height, chroma_name(f.chroma), f.width, TIMES, f.height, want[i].width,
                     TIMES, want[i].height));
    if (p.offset + p.len() < p.offset || p.offset + p.len() > f.data.size())
      config_err(fmt("plane %d runs past the end of the frame's data", i));
  }
  if (f.has_alpha && f.alpha.size() != (size_t)f.width * f.height)
    config_err(fmt("the alpha plane has %zu samples, not %u%s%u", f.alpha.size(), f.width, TIMES,
                   f.height));
}

static std::vector<uint32_t> slice_sizes(uint32_t width_in_mb, uint32_t log2) {  // src/header.rs:308-320
  std::vector<uint32_t> sizes;
  uint32_t size = 1u << log2, remaining = width_in_mb;
  while (remaining > 0) {
    while (remaining >= size) {
      sizes.push_back(size);
      remaining -= size;
    }
    size /= 2;
  }
  return sizes;
}

static const size_t FRAME_HEADER_FIXED = 20;   // src/header.rs:75
static const size_t PICTURE_HEADER_SIZE = 8;   // src/header.rs:258
static const uint8_t DEFAULT_MATRIX_W = 4;     // src/tables.rs:29

struct FrameHeader {  // src/header.rs:44-72
  uint32_t frame_size = 0;
  uint16_t header_size = 0;
  uint8_t bitstream_version = 0;
  uint8_t encoder_identifier[4] = {0, 0, 0, 0};
  uint16_t width = 0, height = 0;
  uint32_t chroma = 2, interlace = 0;
  Metadata metadata;
  uint32_t alpha = 0;  // 0 none, 1 8-bit, 2 16-bit
  bool has_luma = false, has_chroma = false;
  uint8_t luma_matrix[64], chroma_matrix[64];

  uint32_t alpha_bits() const { return alpha == 1 ? 8 : alpha == 2 ? 16 : 0; }
  void luma_weights(uint8_t *w) const {  // src/header.rs:202-204
    if (has_luma)
      memcpy(w, luma_matrix, 64);
    else
      memset(w, DEFAULT_MATRIX_W, 64);
  }
  void chroma_weights(uint8_t *w) const {  // src/header.rs:208-210
    if (has_chroma)
      memcpy(w, chroma_matrix, 64);
    else
      luma_weights(w);
  }
  size_t picture_count() const { return interlace == 0 ? 1 : 2; }
  bool is_top_field(size_t index) const { return (interlace == 1) == (index == 0); }
  uint32_t picture_height(size_t index) const {  // src/header.rs:222-234
    uint32_t v = height;
    if (interlace == 0) return v;
    return is_top_field(index) ? (v + 1) / 2 : v / 2;
  }
  void picture_rows(size_t index, size_t &first, size_t &step) const {  // src/header.rs:243-254
    if (interlace == 0) {
      first = 0;
      step = 1;
    } else if (is_top_field(index)) {
      first = 0;
      step = 2;
    } else {
      first = 1;
      step = 2;
    }
  }

  static FrameHeader parse(const uint8_t *data, size_t len) {  // src/header.rs:82-170
    if (len < 8 + FRAME_HEADER_FIXED)
      invalid(fmt("%zu bytes is too short for a ProRes frame header", len));
    FrameHeader h;
    h.frame_size = ((uint32_t)data[0] << 24) | ((uint32_t)data[1] << 16) |
                   ((uint32_t)data[2] << 8) | data[3];
    if (memcmp(data + 4, "icpf", 4) != 0) invalid("no 'icpf' frame identifier");
    if ((size_t)h.frame_size > len)
      invalid(fmt("frame_size %u exceeds the %zu bytes given", h.frame_size, len));
    if ((size_t)h.frame_size < 8 + FRAME_HEADER_FIXED)
      invalid("frame_size leaves no room for the frame header");
    const uint8_t *p = data + 8;
    size_t plen = h.frame_size - 8;
    h.header_size = (uint16_t)((p[0] << 8) | p[1]);
    if (h.header_size < FRAME_HEADER_FIXED || h.header_size > plen)
      invalid(fmt("frame_header_size %u is out of range", h.header_size));
    h.bitstream_version = p[3];
    if (h.bitstream_version > 1) unsupported(fmt("bitstream_version %u", h.bitstream_version));
    memcpy(h.encoder_identifier, p + 4, 4);
    h.width = (uint16_t)((p[8] << 8) | p[9]);
    h.height = (uint16_t)((p[10] << 8) | p[11]);
    if (h.width == 0 || h.height == 0) invalid(fmt("a %u%s%u frame", h.width, TIMES, h.height));
    uint32_t c = p[12] >> 6;
    if (c != 2 && c != 3) invalid(fmt("chroma_format %u is reserved", c));
    h.chroma = c;
    uint32_t il = (p[12] >> 2) & 3;
    if (il == 3) invalid("interlace_mode 3 is reserved");
    h.interlace = il;
    h.metadata.aspect_ratio = p[13] >> 4;
    h.metadata.frame_rate_code = p[13] & 15;
    h.metadata.color_primaries = p[14];
    h.metadata.transfer_characteristic = p[15];
    h.metadata.matrix_coefficients = p[16];
    uint32_t a = p[17] & 15;
    if (a > 2) invalid(fmt("alpha_channel_type %u is reserved", a));
    h.alpha = a;
    bool load_luma = (p[19] & 2) != 0, load_chroma = (p[19] & 1) != 0;
    size_t pos = FRAME_HEADER_FIXED;
    if (load_luma) {
      if (pos + 64 > h.header_size) invalid("a quantisation matrix runs past frame_header_size");
      memcpy(h.luma_matrix, p + pos, 64);
      h.has_luma = true;
      pos += 64;
    }
    if (load_chroma) {
      if (pos + 64 > h.header_size) invalid("a quantisation matrix runs past frame_header_size");
      memcpy(h.chroma_matrix, p + pos, 64);
      h.has_chroma = true;
      pos += 64;
    }
    return h;
  }

  void write(std::vector<uint8_t> &out) const {  // src/header.rs:174-199
    size_t size = FRAME_HEADER_FIXED + 64 * ((size_t)has_luma + (size_t)has_chroma);
    for (int i = 0; i < 4; i++) out.push_back(0);
    out.insert(out.end(), {'i', 'c', 'p', 'f'});
    out.push_back((uint8_t)(size >> 8));
    out.push_back((uint8_t)size);
    out.push_back(0);
    out.push_back(bitstream_version);
    out.insert(out.end(), encoder_identifier, encoder_identifier + 4);
    out.push_back((uint8_t)(width >> 8));
    out.push_back((uint8_t)width);
    out.push_back((uint8_t)(height >> 8));
    out.push_back((uint8_t)height);
    out.push_back((uint8_t)((chroma << 6) | (interlace << 2)));
    out.push_back((uint8_t)((metadata.aspect_ratio << 4) | (metadata.frame_rate_code & 15)));
    out.push_back(metadata.color_primaries);
    out.push_back(metadata.transfer_characteristic);
    out.push_back(metadata.matrix_coefficients);
    out.push_back((uint8_t)alpha);
    out.push_back(0);
    out.push_back((uint8_t)(((uint8_t)has_luma << 1) | (uint8_t)has_chroma));
    if (has_luma) out.insert(out.end(), luma_matrix, luma_matrix + 64);
    if (has_chroma) out.insert(out.end(), chroma_matrix, chroma_matrix + 64);
  }
};

struct PictureHeader {  // src/header.rs:262-304
  size_t header_size, picture_size;
  uint32_t log2_slice_mbs;
  static PictureHeader parse(const uint8_t *d, size_t len) {
    if (len < PICTURE_HEADER_SIZE) invalid("a picture header runs past the end of the frame");
    PictureHeader h;
    h.header_size = d[0] >> 3;
    h.picture_size = ((size_t)d[1] << 24) | ((size_t)d[2] << 16) | ((size_t)d[3] << 8) | d[4];
    if (h.header_size < PICTURE_HEADER_SIZE || h.header_size > len)
      invalid(fmt("picture_header_size %zu is out of range", h.header_size));
    if (h.picture_size < h.header_size || h.picture_size > len)
      invalid(fmt("picture_size %zu is out of range", h.picture_size));
    h.log2_slice_mbs = (d[7] >> 4) & 3;
    return h;
  }
  static void write(std::vector<uint8_t> &out, uint32_t picture_size, size_t slice_count,
                    uint32_t log2) {
    out.push_back((uint8_t)(PICTURE_HEADER_SIZE << 3));
    for (int i = 3; i >= 0; i--) out.push_back((uint8_t)(picture_size >> (8 * i)));
    uint16_t dep = slice_count <= 65535 ? (uint16_t)slice_count : 0;
    out.push_back((uint8_t)(dep >> 8));
    out.push_back((uint8_t)dep);
    out.push_back((uint8_t)((log2 & 3) << 4));
  }
};

// ---------------------------------------------------------------------------
// Decoder (src/decode.rs).
struct Rescale {  // src/decode.rs:446-469
  uint64_t m;
  bool identity;
  Rescale(uint32_t in, uint32_t out) {
    uint64_t max_in = (1ull << in) - 1, max_out = (1ull << out) - 1;
    m = ((max_out << 40) + max_in - 1) / max_in;
    identity = in == out;
  }
  uint16_t apply(uint16_t a) const {
    return identity ? a : (uint16_t)(((uint64_t)a * m + (1ull << 39)) >> 40);
  }
};

static void check_plausible_size(const FrameHeader &h) {  // src/decode.rs:113-128
  uint64_t width_in_mb = ((uint64_t)h.width + 15) / 16;
  uint64_t need = 8 + (uint64_t)h.header_size;
  for (size_t i = 0; i < h.picture_count(); i++) {
    uint64_t slices = (uint64_t)slice_sizes((uint32_t)width_in_mb, 3).size() *
                      (((uint64_t)h.picture_height(i) + 15) / 16);
    need += PICTURE_HEADER_SIZE + 11 * slices;
  }
  if ((uint64_t)h.frame_size < need)
    invalid(fmt("frame_size %u is too small for a %u%s%u frame (at least %llu bytes)",
                h.frame_size, h.width, TIMES, h.height, (unsigned long long)need));
}

// scanned_coefficients() (src/decode.rs:484-540), calling put(n, value).
template <class Put>
static void decode_scanned(const uint8_t *data, size_t len, size_t n_blocks, Put put) {
  BitReader r(data, len);
  int32_t first = symbol_to_signed(cb_read(FIRST_DC_CODEBOOK, r));
  put(0, first);
  int32_t prev_dc = first, prev_diff = 3;
  for (size_t i = 1; i < n_blocks; i++) {
    int32_t n = symbol_to_signed(cb_read(dc_codebook(prev_diff), r));
    int32_t diff = prev_diff < 0 ? -n : n;
    int64_t s = (int64_t)prev_dc + diff;
    if (s > INT32_MAX || s < INT32_MIN) invalid("a DC coefficient overflows");
    prev_dc = (int32_t)s;
    put(i, prev_dc);
    prev_diff = diff;
  }
  size_t total = n_blocks * 64, n = n_blocks;
  uint32_t prev_run = 4, prev_level = 1;
  for (;;) {
    uint64_t w;
    uint32_t run, l1, lsym, l2;
    if (r.peek_fast(w) && peek_run(prev_run, w, run, l1) && peek_level(prev_level, w << l1, lsym, l2) &&
        l1 + l2 < 57) {
      n += run;
      if (n >= total) invalid("a run passes the end of the coefficient array");
      int32_t level = (int32_t)lsym + 1;
      bool negative = ((w << (l1 + l2)) >> 63) == 1;
      r.skip(l1 + l2 + 1);
      put(n, negative ? -level : level);
      n += 1;
      prev_run = run;
      prev_level = lsym;
      continue;
    }
    if (r.end_of_data()) break;
    run = cb_read(run_codebook(prev_run), r);
    prev_run = run;
    n += run;
    if (n >= total) invalid("a run passes the end of the coefficient array");
    lsym = cb_read(level_codebook(prev_level), r);
    prev_level = lsym;
    int32_t level = (int32_t)lsym + 1;
    put(n, r.read_bit() == 1 ? -level : level);
    n += 1;
  }
}

// scanned_alpha() (src/decode.rs:570-595).
static void decode_alpha(const uint8_t *data, size_t len, uint32_t bits, size_t count,
                         size_t needed, std::vector<uint16_t> &out) {
  BitReader r(data, len);
  int32_t mask = (int32_t)((1u << bits) - 1);
  out.clear();
  int32_t prev = -1;
  while (out.size() < count) {
    if (out.size() >= needed && r.end_of_data()) break;
    int32_t diff = read_alpha_difference(r, bits);
    int32_t alpha = (int32_t)((uint32_t)prev + (uint32_t)diff) & mask;
    prev = alpha;
    size_t run = read_alpha_run(r);
    if (out.size() + run > count) invalid("an alpha run passes the end of the slice");
    out.resize(out.size() + run, (uint16_t)alpha);
  }
}

static void inverse_scan(const uint8_t *scan, uint8_t *out) {  // src/decode.rs:472-478
  for (int k = 0; k < 64; k++) out[scan[k]] = (uint8_t)k;
}

static void block_position(int plane, uint32_t chroma, uint32_t mb_x, size_t mb_y, size_t m,
                           size_t b, size_t &bx, size_t &by) {  // src/decode.rs:395-413
  size_t mx = mb_x + m, y0 = 16 * mb_y;
  if (plane == 0) {
    bx = 16 * mx + 8 * (b & 1);
    by = y0 + 8 * (b >> 1);
  } else if (chroma == 2) {
    bx = 8 * mx;
    by = y0 + 8 * b;
  } else {
    bx = 16 * mx + 8 * (b >> 1);
    by = y0 + 8 * (b & 1);
  }
}

struct DecodeContext {  // src/decode.rs:131-158
  const FrameHeader *hdr;
  uint8_t luma_w[64], chroma_w[64], unscan[64];
  uint32_t bit_depth;
  Output out;
  DecodeContext(const FrameHeader &h, uint32_t bd) : hdr(&h), bit_depth(bd), out(bd) {
    h.luma_weights(luma_w);
    h.chroma_weights(chroma_w);
    inverse_scan(h.interlace == 0 ? PROGRESSIVE_SCAN : INTERLACED_SCAN, unscan);
  }
};

// decode_slice (src/decode.rs:292-391), writing into the frame directly.
// `place_height` / `first_row` / `row_step` locate the picture (src/decode.rs:160-167).
static void decode_slice(const DecodeContext &ctx, uint32_t mb_x, uint32_t mbs, const uint8_t *s,
                         size_t slen, size_t mb_y, size_t height_in_mb, size_t place_height,
                         size_t first_row, size_t row_step, Frame &frame,
                         std::vector<int32_t> &coeffs, std::vector<uint16_t> &alpha_vals) {
  const FrameHeader &hdr = *ctx.hdr;
  bool has_alpha = hdr.alpha != 0;
  size_t min_header = has_alpha ? 8 : 6;
  if (slen < min_header) invalid("a slice is shorter than its header");
  size_t header_size = s[0] >> 3;
  if (header_size < min_header || header_size > slen)
    invalid(fmt("slice_header_size %zu is out of range", header_size));
  uint32_t qindex = s[1];
  if (qindex < 1 || qindex > 224) invalid(fmt("quantization_index %u is reserved", qindex));
  uint32_t q = qscale(qindex);
  size_t y_size = ((size_t)s[2] << 8) | s[3];
  size_t cb_size = ((size_t)s[4] << 8) | s[5];
  size_t body = slen - header_size;
  size_t cr_size;
  if (has_alpha) {
    cr_size = ((size_t)s[6] << 8) | s[7];
  } else {
    if (y_size + cb_size > body) invalid("slice component sizes exceed the slice");
    cr_size = body - (y_size + cb_size);
  }
  if (y_size + cb_size + cr_size > body) invalid("slice component sizes exceed the slice");
  const uint8_t *comp_data[3] = {s + header_size, s + header_size + y_size,
                                 s + header_size + y_size + cb_size};
  size_t comp_len[3] = {y_size, cb_size, cr_size};
  const uint8_t *alpha_data = s + header_size + y_size + cb_size + cr_size;
  size_t alpha_len = slen - (header_size + y_size + cb_size + cr_size);

  size_t cpm = hdr.chroma == 2 ? 2 : 4;
  size_t per_mb[3] = {4, cpm, cpm};
  const uint8_t *weights[3] = {ctx.luma_w, ctx.chroma_w, ctx.chroma_w};
  size_t band_y = 16 * mb_y;
  for (int plane = 0; plane < 3; plane++) {
    size_t nb = per_mb[plane] * mbs;
    coeffs.assign(nb * 64, 0);
    size_t shift = (size_t)__builtin_ctzll(nb), mask = nb - 1;
    int32_t *out = coeffs.data();
    const uint8_t *unscan = ctx.unscan;
    decode_scanned(comp_data[plane], comp_len[plane], nb,
                   [&](size_t n, int32_t v) { out[((n & mask) << 6) | unscan[n >> shift]] = v; });
    float scale[64];
    for (int k = 0; k < 64; k++) scale[k] = (float)((uint32_t)weights[plane][k] * q) / 8.0f;
    const Plane &pl = frame.planes[plane];
    size_t pw = pl.width;
    uint16_t *pdata = frame.data.data() + pl.offset;
    for (size_t blk = 0; blk < nb; blk++) {
      uint16_t v[64];
      idct_put(out + 64 * blk, scale, ctx.out, v);
      size_t bx, by;
      block_position(plane, hdr.chroma, mb_x, mb_y, blk / per_mb[plane], blk % per_mb[plane], bx,
                     by);
      // put_block (src/decode.rs:418-438)
      if (bx >= pw) continue;
      size_t cols = std::min<size_t>(pw - bx, 8);
      size_t rows = place_height > by ? std::min<size_t>(place_height - by, 8) : 0;
      for (size_t y = 0; y < rows; y++) {
        size_t frow = first_row + row_step * (by + y);
        memcpy(pdata + frow * pw + bx, v + 8 * y, cols * sizeof(uint16_t));
      }
    }
  }
  (void)band_y;
  if (has_alpha) {
    size_t rows = mb_y + 1 < height_in_mb ? 16 : place_height - 16 * (height_in_mb - 1);
    size_t cols = 16 * (size_t)mbs;
    decode_alpha(alpha_data, alpha_len, hdr.alpha_bits(), cols * 16, cols * rows, alpha_vals);
    Rescale convert(hdr.alpha_bits(), ctx.bit_depth);
    size_t width = frame.width;
    size_t x0 = 16 * (size_t)mb_x;
    size_t n = std::min(cols, width > x0 ? width - x0 : 0);
    for (size_t r = 0; r < rows; r++) {
      size_t frow = first_row + row_step * (16 * mb_y + r);
      uint16_t *o = frame.alpha.data() + frow * width + x0;
      for (size_t i = 0; i < n; i++) o[i] = convert.apply(alpha_vals[cols * r + i]);
    }
  }
}

static void decode_picture(const DecodeContext &ctx, size_t index, const PictureHeader &ph,
                           const uint8_t *pic, Frame &frame, size_t threads) {  // src/decode.rs:188-290
  const FrameHeader &hdr = *ctx.hdr;
  uint32_t width_in_mb = (hdr.width + 15u) / 16u;
  uint32_t height = hdr.picture_height(index);
  size_t height_in_mb = (height + 15u) / 16u;
  std::vector<uint32_t> sizes = slice_sizes(width_in_mb, ph.log2_slice_mbs);
  size_t n_slices = sizes.size() * height_in_mb;
  if (ph.header_size + 2 * n_slices > ph.picture_size)
    invalid("the slice table runs past picture_size");
  const uint8_t *table = pic + ph.header_size;
  size_t first_row, row_step;
  hdr.picture_rows(index, first_row, row_step);
  std::vector<size_t> offs(n_slices), lens(n_slices);
  size_t pos = ph.header_size + 2 * n_slices;
  for (size_t k = 0; k < n_slices; k++) {
    size_t size = ((size_t)table[2 * k] << 8) | table[2 * k + 1];
    if (pos + size > ph.picture_size) invalid("a slice runs past picture_size");
    offs[k] = pos;
    lens[k] = size;
    pos += size;
  }
  // The band split (src/decode.rs:223-266) never runs out of rows for a
  // header that parsed, since picture heights come from the frame height.
  size_t row_slices = sizes.size();
  std::vector<std::optional<Error>> results(height_in_mb);
  Pool::get().run(threads, height_in_mb, false, [&](size_t i) {
    std::vector<int32_t> coeffs;
    std::vector<uint16_t> alpha_vals;
    uint32_t mb_x = 0;
    try {
      for (size_t j = 0; j < row_slices; j++) {
        size_t k = i * row_slices + j;
        decode_slice(ctx, mb_x, sizes[j], pic + offs[k], lens[k], i, height_in_mb, height,
                     first_row, row_step, frame, coeffs, alpha_vals);
        mb_x += sizes[j];
      }
    } catch (const Error &e) {
      results[i] = e;
    }
  });
  for (auto &r : results)  // the first error in picture order (src/decode.rs:288-289)
    if (r) throw *r;
}

static Frame decode(const uint8_t *data, size_t len, uint32_t want_depth, size_t threads) {
  // Decoder::decode, src/decode.rs:61-104
  FrameHeader hdr = FrameHeader::parse(data, len);
  check_plausible_size(hdr);
  uint32_t bd = want_depth ? want_depth : (hdr.chroma == 2 ? 10 : 12);
  Frame frame = frame_zeroed(hdr.width, hdr.height, hdr.chroma, bd);
  frame.interlace = hdr.interlace;
  frame.metadata = hdr.metadata;
  if (hdr.alpha != 0) {
    frame.has_alpha = true;
    frame.alpha.assign((size_t)frame.width * frame.height, 0);
  }
  size_t flen = hdr.frame_size;
  size_t pos = 8 + hdr.header_size;
  DecodeContext ctx(hdr, bd);
  for (size_t index = 0; index < hdr.picture_count(); index++) {
    if (pos > flen) invalid("a picture starts past the end of the frame");
    PictureHeader ph = PictureHeader::parse(data + pos, flen - pos);
    decode_picture(ctx, index, ph, data + pos, frame, threads);
    pos += ph.picture_size;
  }
  return frame;
}

// ---------------------------------------------------------------------------
// Encoder (src/encode.rs).
enum Profile { PROXY, LT, STANDARD, HQ, P4444, P4444XQ };
static const char *PROFILE_NAMES[6] = {"Proxy", "Lt", "Standard", "Hq", "P4444", "P4444Xq"};
static uint32_t profile_chroma(uint32_t p) { return p >= P4444 ? 3 : 2; }
static uint64_t reference_bitrate(uint32_t p) {  // src/encode.rs:85-94
  static const uint64_t r[6] = {45000000, 102000000, 147000000, 220000000, 330000000, 500000000};
  return r[p];
}
static uint64_t macroblocks(uint32_t w, uint32_t h) {  // src/encode.rs:121-123
  return (uint64_t)((w + 15) / 16) * ((h + 15) / 16);
}
static size_t default_headers(uint32_t w, uint32_t h) {  // src/encode.rs:128-131
  size_t slices = slice_sizes((w + 15) / 16, 3).size() * ((h + 15) / 16);
  return 8 + FRAME_HEADER_FIXED + PICTURE_HEADER_SIZE + (2 + 6) * slices;
}
static size_t target_frame_bytes(uint32_t p, uint32_t w, uint32_t h) {  // src/encode.rs:112-117
  double per_frame = (double)reference_bitrate(p) * 1001.0 / 30000.0 / 8.0;
  double per_mb = (per_frame - (double)default_headers(1920, 1080)) / (double)macroblocks(1920, 1080);
  double t = per_mb * (double)macroblocks(w, h);
  return (size_t)(t < 0 ? 0 : t) + default_headers(w, h);
}

struct Config {
  uint32_t profile = HQ, log2_slice_mbs = 3, alpha = 2;
  bool has_luma = false, has_chroma = false;
  uint8_t luma[64], chroma[64];
  size_t target = 0;  // 0: the profile's
  uint8_t encoder_identifier[4] = {'r', 'i', 'v', 't'};
  size_t threads = 0;
  void validate() const {  // src/encode.rs:176-192
    if (log2_slice_mbs > 3)
      config_err(fmt("log2_slice_mbs %u is outside 0%s3", log2_slice_mbs, EN_DASH));
    for (int m = 0; m < 2; m++) {
      bool has = m == 0 ? has_luma : has_chroma;
      const uint8_t *w = m == 0 ? luma : chroma;
      if (!has) continue;
      for (int k = 0; k < 64; k++)
        if (w[k] < 2 || w[k] > 63)
          config_err(fmt("quantisation weights must be 2%s63", EN_DASH));
    }
  }
};

static const float AC_ROUNDING = 0.42f;  // src/encode.rs:204

struct SliceBlocks {  // src/encode.rs:294-298
  uint32_t mbs;
  const float *comps[3];
  size_t comp_len[3];
  std::vector<uint8_t> alpha;
};

struct Quantised {  // src/encode.rs:331-353
  std::vector<int32_t> values;
  std::vector<uint64_t> mask;
  void quantise(const float *f, size_t len, const float *inv) {
    size_t n = len / 64;
    values.resize(64 * n);
    mask.resize(n);
    oracle::quantise(f, n, inv, AC_ROUNDING, values.data(), mask.data());
  }
};

template <class F>
static void for_each_set_bit(const uint64_t *mask, size_t words, size_t from, F f) {  // src/encode.rs:366-378
  for (size_t w = from / 64; w < words; w++) {
    uint64_t bits = mask[w];
    if (w == from / 64) bits &= ~0ull << (from % 64);
    while (bits) {
      f(64 * w + (size_t)__builtin_ctzll(bits));
      bits &= bits - 1;
    }
  }
}

template <class F>
static void dc_symbols(const int32_t *scanned, size_t n_blocks, F f) {  // src/encode.rs:383-395
  int32_t first = scanned[0];
  f(FIRST_DC_CODEBOOK, signed_to_symbol(first));
  int32_t prev_dc = first, prev_diff = 3;
  for (size_t i = 1; i < n_blocks; i++) {
    int32_t dc = scanned[i];
    int32_t diff = (int32_t)((uint32_t)dc - (uint32_t)prev_dc);
    int32_t n = prev_diff < 0 ? (int32_t)(0u - (uint32_t)diff) : diff;
    f(dc_codebook(prev_diff), signed_to_symbol(n));
    prev_diff = diff;
    prev_dc = dc;
  }
}

static size_t coefficient_bits(const int32_t *values, const uint64_t *mask, size_t n_blocks) {
  // src/encode.rs:398-409
  uint32_t bits = 0;
  dc_symbols(values, n_blocks, [&](Codebook cb, uint32_t n) { bits += cb_len(cb, n); });
  uint32_t prev_run = 4, prev_level = 1;
  size_t pos = n_blocks;
  for_each_set_bit(mask, n_blocks, n_blocks, [&](size_t i) {
    uint32_t run = (uint32_t)(i - pos);
    int32_t v = values[i];
    uint32_t level = (v < 0 ? (uint32_t)(-(int64_t)v) : (uint32_t)v) - 1;
    bits += run_len(prev_run, run) + level_len(prev_level, level) + 1;
    prev_run = run;
    prev_level = level;
    pos = i + 1;
  });
  return bits;
}

template <class S>
static void encode_coefficients_masked(S &s, const int32_t *values, const uint64_t *mask,
                                       size_t n_blocks) {  // src/encode.rs:413-437
  dc_symbols(values, n_blocks, [&](Codebook cb, uint32_t n) { cb_put(cb, s, n); });
  uint32_t prev_run = 4, prev_level = 1;
  size_t pos = n_blocks;
  for_each_set_bit(mask, n_blocks, n_blocks, [&](size_t i) {
    uint32_t run = (uint32_t)(i - pos);
    int32_t c = values[i];
    uint32_t level = (c < 0 ? (uint32_t)(-(int64_t)c) : (uint32_t)c) - 1;
    uint64_t r, l;
    uint32_t rl, ll;
    bool okr = run_code(prev_run, run, r, rl), okl = level_code(prev_level, level, l, ll);
    if (okr && okl && rl + ll < 57) {
      s.put((((r << ll) | l) << 1) | (uint64_t)(c < 0), rl + ll + 1);
    } else {
      cb_put(run_codebook(prev_run), s, run);
      cb_put(level_codebook(prev_level), s, level);
      s.put((uint64_t)(c < 0), 1);
    }
    prev_run = run;
    prev_level = level;
    pos = i + 1;
  });
  s.align();
}

template <class S>
static void encode_alpha(S &s, const uint16_t *values, size_t n, uint32_t bits) {  // src/encode.rs:457-470
  int32_t prev = -1;
  size_t i = 0;
  while (i < n) {
    uint16_t a = values[i];
    size_t end = std::min(i + 2048, n);
    size_t run = 1;
    while (i + run < end && values[i + run] == a) run++;
    put_alpha_difference(s, (int32_t)a - prev, bits);
    put_alpha_run(s, (uint32_t)run);
    prev = a;
    i += run;
  }
  s.align();
}

struct PictureLayout {  // src/encode.rs:473-565
  const Frame *frame;
  const FrameHeader *hdr;
  size_t height, first_row, step;
  const uint8_t *scan;

  void block(int plane, size_t bx, size_t by, uint16_t *pix) const {  // src/encode.rs:488-505
    const Plane &p = frame->planes[plane];
    const uint16_t *src = frame->data.data() + p.offset;
    size_t w = p.width;
    for (size_t y = 0; y < 8; y++) {
      size_t row = first_row + step * std::min(by + y, height - 1);
      const uint16_t *line = src + row * w;
      if (bx + 8 <= w) {
        memcpy(pix + 8 * y, line + bx, 16);
      } else {
        for (size_t x = 0; x < 8; x++) pix[8 * y + x] = line[std::min(bx + x, w - 1)];
      }
    }
  }
  size_t slice_len(uint32_t mbs) const {
    size_t cpm = frame->chroma == 2 ? 2 : 4;
    return 64 * (4 + 2 * cpm) * mbs;
  }
  std::vector<uint8_t> transform(uint32_t mb_x, uint32_t mbs, size_t mb_y, float *out) const {
    size_t cpm = frame->chroma == 2 ? 2 : 4;
    float scale = 512.0f / (float)(1u << frame->bit_depth);
    float *rest = out;
    for (int plane = 0; plane < 3; plane++) {
      size_t per_mb = plane == 0 ? 4 : cpm;
      size_t n = per_mb * mbs;
      for (size_t blk = 0; blk < n; blk++) {
        size_t bx, by;
        block_position(plane, frame->chroma, mb_x, mb_y, blk / per_mb, blk % per_mb, bx, by);
        uint16_t pix[64];
        block(plane, bx, by, pix);
        float f[64];
        fdct_load(pix, scale, f);
        for (int k = 0; k < 64; k++) rest[n * scan[k] + blk] = f[k];
      }
      rest += 64 * n;
    }
    if (frame->has_alpha && hdr->alpha != 0) return slice_alpha(mb_x, mbs, mb_y);
    return {};
  }
  std::vector<uint8_t> slice_alpha(uint32_t mb_x, uint32_t mbs, size_t mb_y) const {
    // src/encode.rs:838-868
    uint32_t bits = hdr->alpha_bits();
    size_t width = frame->width;
    uint16_t max_in = (uint16_t)(0xFFFFu >> (16 - frame->bit_depth));
    Rescale convert(frame->bit_depth, bits);
    size_t cols = 16 * (size_t)mbs, x0 = 16 * (size_t)mb_x;
    std::vector<uint16_t> values;
    values.reserve(cols * 16);
    for (size_t r = 0; r < 16; r++) {
      size_t row = first_row + step * std::min(16 * mb_y + r, height - 1);
      for (size_t n = 0; n < cols; n++) {
        uint16_t a = std::min(frame->alpha[row * width + std::min(x0 + n, width - 1)], max_in);
        values.push_back(convert.apply(a));
      }
    }
    std::vector<uint8_t> out;
    BitWriter w(&out);
    encode_alpha(w, values.data(), values.size(), bits);
    w.finish();
    return out;
  }
};

struct SizeFit {
  size_t n;
  bool fits;
};

static void code_picture(const FrameHeader &hdr, const std::vector<SliceBlocks> &slices,
                         const uint8_t *scan, uint32_t log2, size_t budget, size_t threads,
                         std::vector<uint8_t> &out) {  // src/encode.rs:657-830
  size_t header_bytes = hdr.alpha != 0 ? 8 : 6;
  // Quantisers::new (src/encode.rs:303-324): 8 / (W · qScale) at scanned positions.
  uint8_t lw[64], cw[64];
  hdr.luma_weights(lw);
  hdr.chroma_weights(cw);
  std::vector<float> quant(225 * 2 * 64);
  for (uint32_t qi = 0; qi <= 224; qi++) {
    uint32_t q = qscale(std::max<uint32_t>(qi, 1));
    for (int k = 0; k < 64; k++) {
      quant[(qi * 2 + 0) * 64 + scan[k]] = 8.0f / (float)((uint32_t)lw[k] * q);
      quant[(qi * 2 + 1) * 64 + scan[k]] = 8.0f / (float)((uint32_t)cw[k] * q);
    }
  }
  auto inv_of = [&](uint32_t qi, int plane) { return &quant[(qi * 2 + (plane > 0)) * 64]; };
  auto slice_size = [&](const SliceBlocks &s, uint32_t qi) -> SizeFit {
    thread_local Quantised q;
    size_t comp[3];
    for (int plane = 0; plane < 3; plane++) {
      q.quantise(s.comps[plane], s.comp_len[plane], inv_of(qi, plane));
      size_t nb = s.comp_len[plane] / 64;
      comp[plane] = (coefficient_bits(q.values.data(), q.mask.data(), nb) + 7) / 8;
    }
    size_t total = header_bytes + comp[0] + comp[1] + comp[2] + s.alpha.size();
    return {total, total <= 65535 && comp[0] <= 65535 && comp[1] <= 65535 && comp[2] <= 65535};
  };
  size_t ns = slices.size();
  auto sizes_at = [&](uint32_t qi) {
    std::vector<SizeFit> v(ns);
    Pool::get().run(threads, ns, true, [&](size_t i) { v[i] = slice_size(slices[i], qi); });
    return v;
  };
  size_t overhead = PICTURE_HEADER_SIZE + 2 * ns;
  size_t alpha_bytes = 0;
  for (auto &s : slices) alpha_bytes += s.alpha.size();
  size_t b = budget + alpha_bytes;
  budget = b > overhead ? b - overhead : 0;
  std::vector<std::optional<std::vector<SizeFit>>> probed(225);
  auto totals = [&](uint32_t qi) {
    auto v = sizes_at(qi);
    bool ok = true;
    size_t sum = 0;
    for (auto &x : v) {
      ok = ok && x.fits;
      sum += x.n;
    }
    ok = ok && sum <= budget;
    probed[qi] = std::move(v);
    return ok;
  };
  uint32_t lo = 1, hi = 224;
  bool fits_at_all = totals(224);
  if (fits_at_all) {
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (totals(mid))
        hi = mid;
      else
        lo = mid + 1;
    }
  }
  uint32_t uniform = hi;
  std::vector<size_t> used(ns);
  for (size_t i = 0; i < ns; i++) used[i] = (*probed[uniform])[i].n;

  const size_t WINDOW = 3;
  std::vector<uint32_t> qindices(ns, uniform);
  if (uniform > 1 && fits_at_all) {
    size_t used_sum = 0;
    for (auto u : used) used_sum += u;
    uint64_t slack = budget - used_sum;
    std::vector<uint64_t> mbs_from(ns + 1, 0);
    for (size_t i = ns; i-- > 0;) mbs_from[i] = mbs_from[i + 1] + slices[i].mbs;
    auto most = [&](size_t i) { return used[i] + (size_t)(slack * slices[i].mbs / mbs_from[i]); };
    std::vector<std::vector<std::optional<SizeFit>>> finer;
    for (size_t j = 0; j < std::min<size_t>(uniform - 1, WINDOW); j++) {
      uint32_t qi = uniform - 1 - (uint32_t)j;
      if (j == 0 && probed[qi]) {
        std::vector<std::optional<SizeFit>> row(ns);
        for (size_t i = 0; i < ns; i++) row[i] = (*probed[qi])[i];
        finer.push_back(std::move(row));
        continue;
      }
      std::vector<size_t> cand;
      for (size_t i = 0; i < ns; i++) {
        bool wanted = j == 0 || (finer[j - 1][i] && finer[j - 1][i]->fits &&
                                 finer[j - 1][i]->n <= most(i));
        if (wanted) cand.push_back(i);
      }
      if (cand.empty()) break;
      std::vector<SizeFit> found(cand.size());
      Pool::get().run(threads, cand.size(), true,
                      [&](size_t c) { found[c] = slice_size(slices[cand[c]], qi); });
      std::vector<std::optional<SizeFit>> row(ns);
      for (size_t c = 0; c < cand.size(); c++) row[cand[c]] = found[c];
      finer.push_back(std::move(row));
    }
    size_t remaining_budget = budget, remaining_base = used_sum;
    uint64_t remaining_mbs = mbs_from[0];
    for (size_t i = 0; i < ns; i++) {
      size_t base = used[i];
      uint64_t sl = remaining_budget - remaining_base;
      size_t allowed = base + (size_t)(sl * slices[i].mbs / remaining_mbs);
      uint32_t qi = uniform;
      while (qi > 1) {
        size_t step = uniform - qi;
        SizeFit sz;
        if (step < finer.size() && finer[step][i])
          sz = *finer[step][i];
        else
          sz = slice_size(slices[i], qi - 1);
        if (!(sz.fits && sz.n <= allowed)) break;
        qi -= 1;
        used[i] = sz.n;
      }
      qindices[i] = qi;
      remaining_budget -= used[i];
      remaining_base -= base;
      remaining_mbs -= slices[i].mbs;
    }
  }

  std::vector<std::vector<uint8_t>> coded(ns);
  Pool::get().run(threads, ns, true, [&](size_t i) {
    const SliceBlocks &s = slices[i];
    uint32_t qi = qindices[i];
    std::vector<uint8_t> &bytes = coded[i];
    bytes.reserve(used[i]);
    bytes.push_back((uint8_t)(header_bytes << 3));
    bytes.push_back((uint8_t)qi);
    bytes.resize(header_bytes, 0);
    size_t comp_sizes[3];
    thread_local Quantised q;
    for (int plane = 0; plane < 3; plane++) {
      q.quantise(s.comps[plane], s.comp_len[plane], inv_of(qi, plane));
      BitWriter bw(&bytes);
      encode_coefficients_masked(bw, q.values.data(), q.mask.data(), s.comp_len[plane] / 64);
      comp_sizes[plane] = bw.finish();
    }
    bytes[2] = (uint8_t)(comp_sizes[0] >> 8);
    bytes[3] = (uint8_t)comp_sizes[0];
    bytes[4] = (uint8_t)(comp_sizes[1] >> 8);
    bytes[5] = (uint8_t)comp_sizes[1];
    if (header_bytes == 8) {
      bytes[6] = (uint8_t)(comp_sizes[2] >> 8);
      bytes[7] = (uint8_t)comp_sizes[2];
    }
    bytes.insert(bytes.end(), s.alpha.begin(), s.alpha.end());
  });
  size_t pic_start = out.size();
  PictureHeader::write(out, 0, ns, log2);
  for (auto &s : coded) {
    out.push_back((uint8_t)(s.size() >> 8));
    out.push_back((uint8_t)s.size());
  }
  for (auto &s : coded) out.insert(out.end(), s.begin(), s.end());
  uint32_t picture_size = (uint32_t)(out.size() - pic_start);
  for (int i = 0; i < 4; i++) out[pic_start + 1 + i] = (uint8_t)(picture_size >> (24 - 8 * i));
}

static void encode_picture(const FrameHeader &hdr, size_t index, const Frame &frame, uint32_t log2,
                           size_t budget, size_t threads, std::vector<uint8_t> &out) {
  // src/encode.rs:567-653
  uint32_t width_in_mb = (frame.width + 15) / 16;
  size_t height = hdr.picture_height(index);
  size_t height_in_mb = (height + 15) / 16;
  size_t first_row, step;
  hdr.picture_rows(index, first_row, step);
  std::vector<uint32_t> sizes = slice_sizes(width_in_mb, log2);
  const uint8_t *scan = hdr.interlace == 0 ? PROGRESSIVE_SCAN : INTERLACED_SCAN;
  PictureLayout lay{&frame, &hdr, height, first_row, step, scan};
  struct Place {
    uint32_t mb_x, mbs;
    size_t mb_y;
  };
  std::vector<Place> places;
  for (size_t mb_y = 0; mb_y < height_in_mb; mb_y++) {
    uint32_t mb_x = 0;
    for (uint32_t m : sizes) {
      places.push_back({mb_x, m, mb_y});
      mb_x += m;
    }
  }
  std::vector<size_t> starts(places.size() + 1, 0);
  for (size_t i = 0; i < places.size(); i++) starts[i + 1] = starts[i] + lay.slice_len(places[i].mbs);
  thread_local std::vector<float> coefficients;
  coefficients.resize(starts.back());
  std::vector<SliceBlocks> slices(places.size());
  // The buffer is this thread's; the workers get its address, not their own.
  float *coef_base = coefficients.data();
  Pool::get().run(threads, places.size(), true, [&](size_t i) {
    const Place &pl = places[i];
    slices[i].alpha = lay.transform(pl.mb_x, pl.mbs, pl.mb_y, coef_base + starts[i]);
  });
  size_t cpm = frame.chroma == 2 ? 2 : 4;
  for (size_t i = 0; i < places.size(); i++) {
    SliceBlocks &s = slices[i];
    s.mbs = places[i].mbs;
    size_t luma = 64 * 4 * (size_t)s.mbs, chroma = 64 * cpm * (size_t)s.mbs;
    const float *c = coef_base + starts[i];
    s.comps[0] = c;
    s.comp_len[0] = luma;
    s.comps[1] = c + luma;
    s.comp_len[1] = chroma;
    s.comps[2] = c + luma + chroma;
    s.comp_len[2] = chroma;
  }
  code_picture(hdr, slices, scan, log2, budget, threads, out);
}

static std::vector<uint8_t> encode(const Config &cfg, const Frame &frame) {  // src/encode.rs:223-287
  cfg.validate();
  frame_validate(frame);
  uint32_t pc = profile_chroma(cfg.profile);
  if (frame.chroma != pc)
    config_err(fmt("%s codes %s, and the frame is %s", PROFILE_NAMES[cfg.profile], chroma_name(pc),
                   chroma_name(frame.chroma)));
  uint32_t alpha = 0;
  if (frame.has_alpha) {
    if (pc == 2) config_err(fmt("%s carries no alpha channel", PROFILE_NAMES[cfg.profile]));
    alpha = cfg.alpha;
  }
  FrameHeader hdr;
  hdr.bitstream_version = (frame.chroma == 2 && alpha == 0) ? 0 : 1;
  memcpy(hdr.encoder_identifier, cfg.encoder_identifier, 4);
  hdr.width = (uint16_t)frame.width;
  hdr.height = (uint16_t)frame.height;
  hdr.chroma = frame.chroma;
  hdr.interlace = frame.interlace;
  hdr.metadata = frame.metadata;
  hdr.alpha = alpha;
  hdr.has_luma = cfg.has_luma;
  hdr.has_chroma = cfg.has_chroma;
  if (cfg.has_luma) memcpy(hdr.luma_matrix, cfg.luma, 64);
  if (cfg.has_chroma) memcpy(hdr.chroma_matrix, cfg.chroma, 64);
  std::vector<uint8_t> out;
  hdr.write(out);
  size_t target = cfg.target ? cfg.target : target_frame_bytes(cfg.profile, frame.width, frame.height);
  out.reserve(std::max(target, out.size()));
  size_t pictures = hdr.picture_count();
  size_t per_picture = (target > out.size() ? target - out.size() : 0) / pictures;
  for (size_t index = 0; index < pictures; index++)
    encode_picture(hdr, index, frame, cfg.log2_slice_mbs, per_picture, cfg.threads, out);
  uint32_t size = (uint32_t)out.size();
  for (int i = 0; i < 4; i++) out[i] = (uint8_t)(size >> (24 - 8 * i));
  return out;
}

// ---------------------------------------------------------------------------
// Test pictures: tests/common/mod.rs:7-80.
static void test_frame(uint32_t width, uint32_t height, uint32_t chroma, uint32_t bit_depth,
                       int32_t grain, uint32_t seed, uint16_t *data) {
  Plane planes[3];
  layout(width, height, chroma, planes);
  uint32_t rng = seed | 1;
  auto next = [&]() {
    rng ^= rng << 13;
    rng ^= rng >> 17;
    rng ^= rng << 5;
    return rng;
  };
  auto noise = [&](int32_t a) { return (int32_t)(next() % (2 * (uint32_t)a + 1)) - a; };
  int shift = (int)bit_depth - 10;
  auto scale = [&](double v) -> uint16_t {
    v = v < 64.0 ? 64.0 : (v > 940.0 ? 940.0 : v);
    double s = shift >= 0 ? v * (double)(1 << shift) : v / (double)(1 << -shift);
    return (uint16_t)std::round(s);
  };
  for (int plane = 0; plane < 3; plane++) {
    const Plane &p = planes[plane];
    double xs = plane == 0 ? 1.0 : (chroma == 2 ? 2.0 : 1.0);
    uint16_t *d = data + p.offset;
    for (size_t y = 0; y < p.height; y++)
      for (size_t x = 0; x < p.width; x++) {
        double fx = (double)x * xs, fy = (double)y;
        double w = width, h = height;
        double v;
        if (plane == 0)
          v = 200.0 + 500.0 * fx / w + 120.0 * std::sin(fx / 37.0) * std::cos(fy / 23.0);
        else if (plane == 1)
          v = 512.0 + 150.0 * (fx / w - 0.5) + 60.0 * std::sin(fy / 50.0);
        else
          v = 512.0 - 120.0 * (fy / h - 0.5) + 50.0 * std::cos(fx / 70.0);
        if (fx > w * 0.15 && fx < w * 0.35 && fy > h * 0.2 && fy < h * 0.5)
          v = plane == 0 ? 900.0 : v + 150.0;
        if (fx > w * 0.6 && fx < w * 0.8 && fy > h * 0.55 && fy < h * 0.85)
          v = plane == 0 ? 90.0 : v - 150.0;
        if (fx < w * 0.25 && fy > h * 0.7) v += 150.0 * std::sin(fx * 1.3);
        if (grain > 0) v += (double)noise(grain);
        d[y * p.width + x] = scale(v);
      }
  }
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C boundary.
using namespace oracle;

static int report(const Error &e, char *msg, size_t cap) {
  if (msg && cap) {
    size_t n = std::min(cap - 1, e.msg.size());
    memcpy(msg, e.msg.data(), n);
    msg[n] = 0;
  }
  return e.kind;
}

static Frame frame_from_c(const oracle_frame *f) {
  Frame fr;
  fr.width = f->width;
  fr.height = f->height;
  fr.bit_depth = f->bit_depth;
  fr.chroma = f->chroma;
  fr.interlace = f->interlace;
  fr.metadata.aspect_ratio = f->aspect_ratio;
  fr.metadata.frame_rate_code = f->frame_rate_code;
  fr.metadata.color_primaries = f->color_primaries;
  fr.metadata.transfer_characteristic = f->transfer_characteristic;
  fr.metadata.matrix_coefficients = f->matrix_coefficients;
  fr.data.assign(f->data, f->data + f->data_len);
  for (int i = 0; i < 3; i++) fr.planes[i] = {(size_t)f->plane_offset[i], f->plane_width[i], f->plane_height[i]};
  fr.has_alpha = f->has_alpha != 0;
  if (fr.has_alpha) fr.alpha.assign(f->alpha, f->alpha + f->alpha_len);
  return fr;
}

extern "C" {

int oracle_decode(const uint8_t *data, size_t len, uint32_t bit_depth, uint32_t threads,
                  oracle_frame *out, char *msg, size_t cap) {
  memset(out, 0, sizeof *out);
  try {
    if (bit_depth != 0 && (bit_depth < 8 || bit_depth > 16))  // Decoder::with_bit_depth
      config_err(fmt("output bit depth %u is outside 8%s16", bit_depth, EN_DASH));
    Frame f = oracle::decode(data, len, bit_depth, threads);
    out->width = f.width;
    out->height = f.height;
    out->bit_depth = f.bit_depth;
    out->chroma = f.chroma;
    out->interlace = f.interlace;
    out->aspect_ratio = f.metadata.aspect_ratio;
    out->frame_rate_code = f.metadata.frame_rate_code;
    out->color_primaries = f.metadata.color_primaries;
    out->transfer_characteristic = f.metadata.transfer_characteristic;
    out->matrix_coefficients = f.metadata.matrix_coefficients;
    out->data_len = f.data.size();
    out->data = (uint16_t *)malloc(f.data.size() * 2 + 2);
    memcpy(out->data, f.data.data(), f.data.size() * 2);
    for (int i = 0; i < 3; i++) {
      out->plane_offset[i] = f.planes[i].offset;
      out->plane_width[i] = f.planes[i].width;
      out->plane_height[i] = f.planes[i].height;
    }
    out->has_alpha = f.has_alpha;
    if (f.has_alpha) {
      out->alpha = (uint16_t *)malloc(f.alpha.size() * 2 + 2);
      memcpy(out->alpha, f.alpha.data(), f.alpha.size() * 2);
      out->alpha_len = f.alpha.size();
    }
    return ORACLE_OK;
  } catch (const Error &e) {
    return report(e, msg, cap);
  }
}

void oracle_frame_free(oracle_frame *f) {
  free(f->data);
  free(f->alpha);
  f->data = nullptr;
  f->alpha = nullptr;
}

int oracle_encode(const oracle_config *c, const oracle_frame *frame, uint8_t **out, size_t *out_len,
                  char *msg, size_t cap) {
  *out = nullptr;
  *out_len = 0;
  try {
    Config cfg;
    cfg.profile = c->profile;
    cfg.log2_slice_mbs = c->log2_slice_mbs;
    cfg.alpha = c->alpha;
    cfg.has_luma = c->has_luma_matrix;
    cfg.has_chroma = c->has_chroma_matrix;
    memcpy(cfg.luma, c->luma_matrix, 64);
    memcpy(cfg.chroma, c->chroma_matrix, 64);
    cfg.target = (size_t)c->target_frame_bytes;
    memcpy(cfg.encoder_identifier, c->encoder_identifier, 4);
    cfg.threads = c->threads;
    Frame f = frame_from_c(frame);
    std::vector<uint8_t> bytes = oracle::encode(cfg, f);
    *out = (uint8_t *)malloc(bytes.size());
    memcpy(*out, bytes.data(), bytes.size());
    *out_len = bytes.size();
    return ORACLE_OK;
  } catch (const Error &e) {
    return report(e, msg, cap);
  }
}

void oracle_free(void *p) { free(p); }

void oracle_config_default(uint32_t profile, oracle_config *c) {
  memset(c, 0, sizeof *c);
  c->profile = profile;
  c->log2_slice_mbs = 3;
  c->alpha = 2;
  memcpy(c->encoder_identifier, "rivt", 4);
}

uint64_t oracle_target_frame_bytes(uint32_t profile, uint32_t w, uint32_t h) {
  return target_frame_bytes(profile, w, h);
}

void oracle_idct_put(const int32_t *coef, const float *scale, uint32_t bd, uint16_t *out) {
  idct_put(coef, scale, Output(bd), out);
}
void oracle_idct(const float *coef, float *out) { idct(coef, out); }
void oracle_fdct(const float *pix, float *out) { fdct(pix, out); }
void oracle_fdct_load(const uint16_t *pix, float scale, float *out) { fdct_load(pix, scale, out); }
void oracle_quantise(const float *f, size_t n, const float *inv, float r, int32_t *out, uint64_t *mask) {
  quantise(f, n, inv, r, out, mask);
}
void oracle_basis(float *out) { memcpy(out, basis(), 64 * sizeof(float)); }
size_t oracle_slice_sizes(uint32_t w, uint32_t log2, uint32_t *out, size_t cap) {
  auto v = slice_sizes(w, log2);
  for (size_t i = 0; i < v.size() && i < cap; i++) out[i] = v[i];
  return v.size();
}
uint32_t oracle_qscale(uint32_t qi) { return qscale(qi); }
void oracle_scan(uint32_t which, uint8_t *out) {
  memcpy(out, which ? INTERLACED_SCAN : PROGRESSIVE_SCAN, 64);
}

static Codebook book(uint32_t kind, uint32_t a, uint32_t b, uint32_t c) {
  return kind == 0 ? EG(a) : CB(a, b, c);
}
static uint32_t to_string(const std::vector<uint8_t> &v, size_t bits, char *out, size_t cap) {
  size_t n = std::min(bits, cap ? cap - 1 : 0);
  for (size_t i = 0; i < n; i++) out[i] = ((v[i / 8] >> (7 - i % 8)) & 1) ? '1' : '0';
  if (cap) out[n] = 0;
  return (uint32_t)bits;
}
uint32_t oracle_codeword(uint32_t kind, uint32_t a, uint32_t b, uint32_t c, uint32_t n, char *bits,
                         size_t cap) {
  std::vector<uint8_t> v;
  BitWriter w(&v);
  cb_put(book(kind, a, b, c), w, n);
  size_t nb = w.bits();
  w.finish();
  return to_string(v, nb, bits, cap);
}
uint32_t oracle_code_len(uint32_t kind, uint32_t a, uint32_t b, uint32_t c, uint32_t n) {
  return cb_len(book(kind, a, b, c), n);
}
int oracle_code_read(uint32_t kind, uint32_t a, uint32_t b, uint32_t c, const uint8_t *data,
                     size_t len, uint32_t *sym, char *msg, size_t cap) {
  try {
    BitReader r(data, len);
    *sym = cb_read(book(kind, a, b, c), r);
    return 0;
  } catch (const Error &e) {
    return report(e, msg, cap);
  }
}
uint32_t oracle_signed_to_symbol(int32_t n) { return signed_to_symbol(n); }
int32_t oracle_symbol_to_signed(uint32_t s) { return symbol_to_signed(s); }
uint32_t oracle_alpha_run_bits(uint32_t run, char *bits, size_t cap) {
  std::vector<uint8_t> v;
  BitWriter w(&v);
  put_alpha_run(w, run);
  size_t nb = w.bits();
  w.finish();
  return to_string(v, nb, bits, cap);
}
uint32_t oracle_alpha_diff_bits(int32_t d, uint32_t nbits, char *bits, size_t cap) {
  std::vector<uint8_t> v;
  BitWriter w(&v);
  put_alpha_difference(w, d, nbits);
  size_t nb = w.bits();
  w.finish();
  return to_string(v, nb, bits, cap);
}
uint16_t oracle_rescale(uint32_t in, uint32_t out, uint16_t a) { return Rescale(in, out).apply(a); }
void oracle_rescale_many(uint32_t in, uint32_t out, const uint16_t *a, size_t n, uint16_t *o) {
  Rescale r(in, out);
  for (size_t i = 0; i < n; i++) o[i] = r.apply(a[i]);
}

size_t oracle_encode_coefficients(const int32_t *scanned, size_t nb, uint8_t *out, size_t cap) {
  size_t n = nb * 64;
  std::vector<uint64_t> mask((n + 63) / 64, 0);
  for (size_t i = 0; i < n; i++) mask[i / 64] |= (uint64_t)(scanned[i] != 0) << (i % 64);
  std::vector<uint8_t> v;
  BitWriter w(&v);
  encode_coefficients_masked(w, scanned, mask.data(), nb);
  w.finish();
  memcpy(out, v.data(), std::min(cap, v.size()));
  return v.size();
}
size_t oracle_coefficient_bits(const int32_t *scanned, size_t nb) {
  size_t n = nb * 64;
  std::vector<uint64_t> mask((n + 63) / 64, 0);
  for (size_t i = 0; i < n; i++) mask[i / 64] |= (uint64_t)(scanned[i] != 0) << (i % 64);
  return coefficient_bits(scanned, mask.data(), nb);
}
int oracle_decode_coefficients(const uint8_t *data, size_t len, size_t nb, int32_t *out, char *msg,
                               size_t cap) {
  try {
    for (size_t i = 0; i < nb * 64; i++) out[i] = 0;
    decode_scanned(data, len, nb, [&](size_t n, int32_t v) { out[n] = v; });
    return 0;
  } catch (const Error &e) {
    return report(e, msg, cap);
  }
}
size_t oracle_encode_alpha(const uint16_t *values, size_t n, uint32_t bits, uint8_t *out, size_t cap) {
  std::vector<uint8_t> v;
  BitWriter w(&v);
  encode_alpha(w, values, n, bits);
  w.finish();
  memcpy(out, v.data(), std::min(cap, v.size()));
  return v.size();
}
int oracle_decode_alpha(const uint8_t *data, size_t len, uint32_t bits, size_t count, size_t needed,
                        uint16_t *out, size_t *out_n, char *msg, size_t cap) {
  try {
    std::vector<uint16_t> v;
    decode_alpha(data, len, bits, count, needed, v);
    memcpy(out, v.data(), v.size() * 2);
    *out_n = v.size();
    return 0;
  } catch (const Error &e) {
    return report(e, msg, cap);
  }
}

void oracle_test_frame(uint32_t w, uint32_t h, uint32_t chroma, uint32_t bd, int32_t grain,
                       uint32_t seed, uint16_t *data) {
  test_frame(w, h, chroma, bd, grain, seed, data);
}

uint32_t oracle_resolve_threads(uint32_t threads) { return (uint32_t)resolve(threads); }

}  // extern "C"
