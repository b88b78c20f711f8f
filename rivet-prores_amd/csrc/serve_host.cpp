// The serving loop: decode and encode requests arrive at any time, are
// admitted while their device-memory footprint fits the server's budget, and
// are batched as they arrive: a worker takes every admissible request each
// time it comes round, so requests that arrive while a batch is on the GPU
// ride the next one. Admission is first in, first out. While a worker's
// batch is on the GPU, a stager thread of its kind admits the requests that
// arrive and uploads their inputs on a stream of its own, so the next
// batch's uploads overlap the kernels of the one running; an idle worker
// admits and uploads requests itself. Decode
// batches and encode batches run on lanes of their own (separate streams),
// so their work overlaps on the device. Results stay resident in device
// memory until the client releases them.
//
// Device buffers come from a caching pool: freed buffers wait in size-class
// lists for reuse and are returned to HIP only when the server is destroyed
// (hipFree synchronises the device, which would stall the other lane).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <set>
#include <thread>

#include "host_internal.h"

using namespace prores_hip;

namespace {

double now_us(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

// Size classes: powers of two up to 1 MiB, then multiples of 1 MiB.
size_t size_class(size_t n) {
  if (n <= 256) return 256;
  if (n <= (1u << 20)) {
    size_t c = 256;
    while (c < n) c <<= 1;
    return c;
  }
  return (n + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
}

struct Pool {
  std::mutex mu;
  std::map<size_t, std::vector<void *>> free_lists;
  std::vector<void *> all;
  Err get(size_t n, void *&out) {
    const size_t c = size_class(n);
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = free_lists.find(c);
      if (it != free_lists.end() && !it->second.empty()) {
        out = it->second.back();
        it->second.pop_back();
        return Err{};
      }
    }
    Err e = check(hipMalloc(&out, c), "hipMalloc(serve)");
    if (e.ok()) {
      std::lock_guard<std::mutex> g(mu);
      all.push_back(out);
    }
    return e;
  }
  void put(void *p, size_t n) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu);
    free_lists[size_class(n)].push_back(p);
  }
  ~Pool() {
    for (void *p : all) (void)hipFree(p);
  }
};

enum Kind : uint32_t { DECODE = 1, ENCODE = 2 };

// A host copy of a request's input, taken at submit into page-locked
// memory where the library can pin it, so the worker's upload is one DMA at
// the link's speed.
struct HostCopy {
  void *p = nullptr;
  size_t bytes = 0;
  void assign(const void *src, size_t n) {
    release();
    p = host_staging_get(n);
    bytes = n;
    if (n) memcpy(p, src, n);
  }
  void release() {
    if (p) host_staging_put(p);
    p = nullptr;
    bytes = 0;
  }
  ~HostCopy() { release(); }
};

struct Request {
  uint64_t id = 0;
  Kind kind = DECODE;
  // Inputs (host copies taken at submit).
  HostCopy packet;
  prores_frame_header hdr;
  prores_config cfg;
  prores_frame frame;  // metadata; data / alpha point into the copies below
  HostCopy data, alpha;
  // Device memory.
  size_t hold = 0;       // bytes counted against the budget while admitted
  size_t keep = 0;       // of which the outputs, held until release
  void *dev_in = nullptr, *dev_in_alpha = nullptr, *dev_out = nullptr, *dev_out_alpha = nullptr;
  size_t in_bytes = 0, in_alpha_bytes = 0, out_bytes = 0, out_alpha_bytes = 0;
  Err err;
  hipEvent_t uploaded = nullptr;  // recorded after the input's upload on its stager's stream
  double t_submit = 0, t_start = 0, t_done = 0;
  uint32_t batch_frames = 0;
  ~Request() {
    if (uploaded) (void)hipEventDestroy(uploaded);
  }
};

// Bytes a decode holds while in flight: the packet, the planes, the alpha
// plane and its share of the batch's coefficient arenas (i16 + i32 escape
// slots per coefficient).
size_t decode_footprint(const prores_frame_header &h, size_t packet_len, size_t &keep) {
  prores_plane pl[3];
  const size_t samples = prores_frame_layout(h.width, h.height, h.chroma, pl);
  const size_t alpha = h.alpha ? (size_t)h.width * h.height * 2 : 0;
  keep = size_class(samples * 2 + 64) + (alpha ? size_class(alpha + 64) : 0);
  const size_t wmb = (h.width + 15) / 16, hmb = (h.height + 15) / 16 + (h.interlace ? 1 : 0);
  const size_t coefs = wmb * hmb * (h.chroma == 2 ? 8 : 12) * 64;
  return keep + size_class(packet_len + 128) + coefs * 6;
}

// Bytes an encode holds while in flight: the input planes, the f32
// coefficient arena, the alpha scratch and size tables, and the packet
// (counted at twice the target, the rate control's bound in practice).
size_t encode_footprint(const prores_config &c, const prores_frame &f, size_t &keep) {
  const size_t samples = f.data_len, alpha = f.alpha ? f.alpha_len * 2 : 0;
  const uint64_t target = c.target_frame_bytes ? c.target_frame_bytes
                                               : prores_profile_target_frame_bytes(c.profile, f.width, f.height);
  keep = size_class(2 * target + 4096);
  const size_t wmb = (f.width + 15) / 16, hmb = (f.height + 15) / 16 + (f.interlace ? 1 : 0);
  const size_t slices = wmb * hmb;  // at most one per macroblock
  return keep + size_class(samples * 2 + 64) + (alpha ? size_class(alpha + 64) : 0) + samples * 4 +
         slices * (4608 + 64);
}

}  // namespace

struct prores_server {
  prores_server_config cfg;
  Device *dev = nullptr;
  Lane *dec_lane = nullptr, *enc_lane = nullptr;
  prores_decoder *decoder = nullptr;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  Pool pool;
  std::mutex mu;
  std::condition_variable cv_work, cv_ready, cv_done;
  bool stop = false;
  uint64_t next_id = 1;
  std::deque<Request *> pend[2];   // decode, encode: waiting for admission
  std::deque<Request *> ready[2];  // admitted, inputs uploading or uploaded
  hipStream_t upload[2] = {nullptr, nullptr};
  bool busy[2] = {false, false};   // the worker has a batch on the GPU
  std::deque<Request *> done;
  std::map<uint64_t, Request *> live;  // admitted and not yet released
  uint64_t held = 0;
  prores_server_stats stats;
  std::thread workers[2], stagers[2];

  bool fits(const Request *r) const { return held + r->hold <= cfg.hbm_budget; }
  Request *admit(int which);
  void stage(Request *r, hipStream_t st, bool event);
  void stager(int which);
  void worker(int which);
  void finish(std::vector<Request *> &batch);
};

// The front request of a kind's queue, admitted (mu held; it fits).
Request *prores_server::admit(int which) {
  Request *r = pend[which].front();
  pend[which].pop_front();
  held += r->hold;
  stats.held_bytes = held;
  stats.peak_held_bytes = std::max(stats.peak_held_bytes, held);
  live[r->id] = r;
  return r;
}

// An admitted request's device buffers and its upload on stream `st`; with
// `event`, an event recorded after it for the batch to wait on (an upload on
// the batch's own stream needs none).
void prores_server::stage(Request *r, hipStream_t st, bool event) {
  Err e;
  if (r->kind == DECODE) {
    prores_plane pl[3];
    r->out_bytes = prores_frame_layout(r->hdr.width, r->hdr.height, r->hdr.chroma, pl) * 2;
    r->out_alpha_bytes = r->hdr.alpha ? (size_t)r->hdr.width * r->hdr.height * 2 : 0;
    r->in_bytes = r->packet.bytes + 128;  // the decoder reads up to 64 bytes past a packet
    e = pool.get(r->in_bytes + 64, r->dev_in);
    if (e.ok()) e = pool.get(r->out_bytes + 64, r->dev_out);
    if (e.ok() && r->out_alpha_bytes) e = pool.get(r->out_alpha_bytes + 64, r->dev_out_alpha);
    if (e.ok())
      e = check(hipMemcpyAsync(r->dev_in, r->packet.p, r->packet.bytes, hipMemcpyHostToDevice, st), "H2D packet");
  } else {
    r->in_bytes = r->data.bytes;
    r->in_alpha_bytes = r->alpha.bytes;
    e = pool.get(r->in_bytes + 64, r->dev_in);
    if (e.ok() && r->in_alpha_bytes) e = pool.get(r->in_alpha_bytes + 64, r->dev_in_alpha);
    if (e.ok()) e = check(hipMemcpyAsync(r->dev_in, r->data.p, r->in_bytes, hipMemcpyHostToDevice, st), "H2D frame");
    if (e.ok() && r->in_alpha_bytes)
      e = check(hipMemcpyAsync(r->dev_in_alpha, r->alpha.p, r->in_alpha_bytes, hipMemcpyHostToDevice, st),
                "H2D alpha");
  }
  if (e.ok() && event) e = check(hipEventCreateWithFlags(&r->uploaded, hipEventDisableTiming), "hipEventCreate");
  if (e.ok() && event) e = check(hipEventRecord(r->uploaded, st), "hipEventRecord");
  // A request that failed here goes back to the pool without a batch: no
  // copy of it may still be in flight then.
  if (!e.ok()) (void)hipStreamSynchronize(st);
  r->err = e;
}

// While the kind's worker has a batch on the GPU: admission and upload of
// the requests that arrive, one at a time, for the worker's next batch.
void prores_server::stager(int which) {
  for (;;) {
    Request *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || (busy[which] && !pend[which].empty() && fits(pend[which].front())); });
      if (stop) return;
      r = admit(which);
    }
    stage(r, upload[which], true);
    {
      std::lock_guard<std::mutex> g(mu);
      ready[which].push_back(r);
      stats.staged_ahead++;
    }
    cv_ready.notify_all();
  }
}

// One batch of one kind: the requests the stager uploaded, then any the
// worker admits and uploads itself, the batch call, results.
void prores_server::worker(int which) {
  Lane &L = which == 0 ? *dec_lane : *enc_lane;
  for (;;) {
    std::vector<Request *> batch;
    size_t staged = 0;
    {
      std::unique_lock<std::mutex> lk(mu);
      busy[which] = false;
      cv_ready.wait(lk, [&] {
        return stop || !ready[which].empty() || (!pend[which].empty() && fits(pend[which].front()));
      });
      if (stop) return;
      while (!ready[which].empty() && batch.size() < cfg.max_batch_frames) {
        batch.push_back(ready[which].front());
        ready[which].pop_front();
      }
      staged = batch.size();
      while (!pend[which].empty() && batch.size() < cfg.max_batch_frames && fits(pend[which].front()))
        batch.push_back(admit(which));
      busy[which] = true;
    }
    cv_work.notify_all();  // the stager takes the arrivals from here on
    std::lock_guard<std::mutex> lane_lock(L.mu);
    for (size_t i = staged; i < batch.size(); i++) stage(batch[i], L.stream, false);
    const double start = now_us(t0);
    for (Request *r : batch) {
      r->t_start = start;
      r->batch_frames = (uint32_t)batch.size();
    }
    // The batch's kernels run after the stager's uploads have arrived.
    for (Request *r : batch)
      if (r->err.ok() && r->uploaded)
        r->err = check(hipStreamWaitEvent(L.stream, r->uploaded, 0), "hipStreamWaitEvent");
    std::vector<Request *> ok;
    for (Request *r : batch)
      if (r->err.ok()) ok.push_back(r);
    if (!ok.empty()) {
      std::vector<prores_error> errs(ok.size());
      if (which == 0) {
        std::vector<prores_dev_decode_job> jobs(ok.size());
        for (size_t i = 0; i < ok.size(); i++)
          jobs[i] = prores_dev_decode_job{(const uint8_t *)ok[i]->packet.p, ok[i]->packet.bytes,
                                          (const uint8_t *)ok[i]->dev_in, (uint16_t *)ok[i]->dev_out,
                                          (uint16_t *)ok[i]->dev_out_alpha};
        // The device's batch tuning, launched directly: a capture on this
        // worker's streams can be invalidated by other threads of the
        // process (clients reading their results through the legacy
        // stream), and the batches of a serving loop change shape too often
        // for a captured graph to be replayed.
        prores_decode_tuning t;
        {
          std::lock_guard<std::mutex> g(dev->mu);
          t = dev->tuning;
        }
        t.graph = 0;
        (void)decode_batch_on(*dev, L, cfg.decode_bit_depth, jobs.size(), jobs.data(), &t, errs.data());
        // A batch that failed part-way may still have copies or kernels on
        // the stream: none may touch the inputs once they go back to the pool.
        (void)hipStreamSynchronize(L.stream);
        (void)hipStreamSynchronize(L.aux);
      } else {
        std::vector<prores_dev_encode_job> jobs(ok.size());
        for (size_t i = 0; i < ok.size(); i++) {
          prores_frame f = ok[i]->frame;
          f.data = (uint16_t *)ok[i]->dev_in;
          f.alpha = (uint16_t *)ok[i]->dev_in_alpha;
          jobs[i] = prores_dev_encode_job{&ok[i]->cfg, f, nullptr, 0};
        }
        (void)encode_batch_on(*dev, L, jobs.size(), jobs.data(), errs.data());
        // The packets sit in the lane's output arena, which the next batch
        // reuses: each moves to a buffer of its own.
        for (size_t i = 0; i < ok.size(); i++) {
          if (errs[i].kind != 0) continue;
          Request *r = ok[i];
          r->out_bytes = jobs[i].out_len;
          Err e = pool.get(r->out_bytes + 64, r->dev_out);
          if (e.ok())
            e = check(hipMemcpyAsync(r->dev_out, jobs[i].dev_out, r->out_bytes, hipMemcpyDeviceToDevice, L.stream),
                      "D2D packet");
          if (!e.ok()) fill_error(&errs[i], e);
        }
        (void)hipStreamSynchronize(L.stream);
      }
      for (size_t i = 0; i < ok.size(); i++)
        ok[i]->err = Err{(prores_status)errs[i].kind, errs[i].msg};
    }
    // Inputs go back to the pool now; outputs stay until released.
    for (Request *r : batch) {
      pool.put(r->dev_in, r->in_bytes + 64);
      pool.put(r->dev_in_alpha, r->in_alpha_bytes + 64);
      r->dev_in = r->dev_in_alpha = nullptr;
      r->packet.release();
      r->data.release();
      r->alpha.release();
    }
    finish(batch);
  }
}

void prores_server::finish(std::vector<Request *> &batch) {
  std::lock_guard<std::mutex> g(mu);
  const double t = now_us(t0);
  for (Request *r : batch) {
    r->t_done = t;
    // The transient part of the footprint is free again.
    held -= r->hold - r->keep;
    r->hold = r->keep;
    stats.completed++;
    if (!r->err.ok()) stats.failed++;
    done.push_back(r);
  }
  stats.held_bytes = held;
  if (!batch.empty()) {
    if (batch[0]->kind == DECODE)
      stats.decode_batches++;
    else
      stats.encode_batches++;
    stats.max_batch_frames = std::max<uint64_t>(stats.max_batch_frames, batch.size());
  }
  cv_done.notify_all();
  cv_work.notify_all();
  cv_ready.notify_all();
}

extern "C" {

void prores_server_config_default(prores_server_config *c) {
  c->hbm_budget = 64ull << 30;
  c->max_batch_frames = 64;
  c->decode_bit_depth = 0;
}

prores_status prores_server_create(const prores_server_config *cfg, prores_server **out, prores_error *err) {
  *out = nullptr;
  Err e;
  if (cfg->max_batch_frames < 1 || cfg->hbm_budget < (1u << 20))
    e = config_error("a server needs a budget of at least 1 MiB and batches of at least one frame");
  prores_decoder *decoder = nullptr;
  if (e.ok()) {
    prores_error de;
    if (prores_decoder_create(cfg->decode_bit_depth, 0, &decoder, &de) != PRORES_OK)
      e = Err{(prores_status)de.kind, de.msg};
  }
  prores_server *s = nullptr;
  if (e.ok()) {
    s = new prores_server();
    s->cfg = *cfg;
    s->decoder = decoder;
    memset(&s->stats, 0, sizeof s->stats);
    s->stats.hbm_budget = cfg->hbm_budget;
    e = device(s->dev);
    if (e.ok()) e = lane(*s->dev, 1, s->dec_lane);
    if (e.ok()) e = lane(*s->dev, 2, s->enc_lane);
    if (e.ok()) e = ensure_encoder(*s->dev);
  }
  if (!e.ok()) {
    if (s) {
      prores_decoder_destroy(s->decoder);
      delete s;
    } else if (decoder) {
      prores_decoder_destroy(decoder);
    }
    fill_error(err, e);
    return (prores_status)e.kind;
  }
  for (int k = 0; k < 2 && e.ok(); k++)
    e = check(hipStreamCreateWithFlags(&s->upload[k], hipStreamNonBlocking), "hipStreamCreate(serve upload)");
  if (!e.ok()) {
    for (hipStream_t st : s->upload)
      if (st) (void)hipStreamDestroy(st);
    prores_decoder_destroy(s->decoder);
    delete s;
    fill_error(err, e);
    return (prores_status)e.kind;
  }
  for (int k = 0; k < 2; k++) {
    s->stagers[k] = std::thread([s, k] { s->stager(k); });
    s->workers[k] = std::thread([s, k] { s->worker(k); });
  }
  *out = s;
  fill_error(err, Err{});
  return PRORES_OK;
}

void prores_server_destroy(prores_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_ready.notify_all();
  s->cv_done.notify_all();
  for (auto &w : s->workers)
    if (w.joinable()) w.join();
  for (auto &w : s->stagers)
    if (w.joinable()) w.join();
  // Uploads still in flight end before their buffers go.
  for (hipStream_t st : s->upload)
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  // Every request once: the queued ones, the admitted ones (live, polled or
  // not) and the ones completed at submit without admission (done only).
  std::set<Request *> all;
  for (auto &q : s->pend) all.insert(q.begin(), q.end());
  for (auto &q : s->ready) all.insert(q.begin(), q.end());
  for (auto &kv : s->live) all.insert(kv.second);
  all.insert(s->done.begin(), s->done.end());
  for (Request *r : all) delete r;
  prores_decoder_destroy(s->decoder);
  delete s;
}

// Queues a request, or completes it at once with an error when it can never
// be admitted.
static uint64_t enqueue(prores_server *s, Request *r, int which) {
  std::lock_guard<std::mutex> g(s->mu);
  r->id = s->next_id++;
  r->t_submit = now_us(s->t0);
  s->stats.submitted++;
  if (!r->err.ok() || r->hold > s->cfg.hbm_budget) {
    if (r->err.ok())
      r->err = config_error(fmtstr("the request needs %zu bytes of device memory, over the server's budget of %llu",
                                   r->hold, (unsigned long long)s->cfg.hbm_budget));
    r->hold = r->keep = 0;
    r->t_start = r->t_done = r->t_submit;
    s->stats.completed++;
    s->stats.failed++;
    s->live[r->id] = r;
    s->done.push_back(r);
    s->cv_done.notify_all();
    return r->id;
  }
  s->pend[which].push_back(r);
  s->cv_work.notify_all();
  s->cv_ready.notify_all();
  return r->id;
}

uint64_t prores_server_submit_decode(prores_server *s, const uint8_t *packet, size_t len) {
  Request *r = new Request();
  r->kind = DECODE;
  memset(&r->hdr, 0, sizeof r->hdr);
  Header h;
  r->err = parse_frame_header(packet, len, h);
  if (r->err.ok()) r->err = check_plausible_size(h);
  if (r->err.ok()) {
    r->hdr = h.h;
    const size_t n = std::min<size_t>(len, h.h.frame_size);
    r->packet.assign(packet, n);
    r->hold = decode_footprint(r->hdr, r->packet.bytes, r->keep);
  }
  return enqueue(s, r, 0);
}

uint64_t prores_server_submit_encode(prores_server *s, const prores_config *cfg, const prores_frame *frame) {
  Request *r = new Request();
  r->kind = ENCODE;
  r->cfg = *cfg;
  r->frame = *frame;
  // The checks of Encoder::encode, so a refused request fails at once.
  r->err = encode_check(*cfg, *frame);
  if (r->err.ok()) {
    r->data.assign(frame->data, frame->data_len * 2);
    if (frame->alpha) r->alpha.assign(frame->alpha, frame->alpha_len * 2);
    r->frame.data = (uint16_t *)r->data.p;
    r->frame.alpha = frame->alpha ? (uint16_t *)r->alpha.p : nullptr;
    r->hold = encode_footprint(r->cfg, r->frame, r->keep);
  }
  return enqueue(s, r, 1);
}

size_t prores_server_poll(prores_server *s, prores_server_result *out, size_t max, int32_t wait_ms) {
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && wait_ms != 0) {
    if (wait_ms < 0)
      s->cv_done.wait(lk, [&] { return s->stop || !s->done.empty(); });
    else
      s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return s->stop || !s->done.empty(); });
  }
  size_t n = 0;
  while (n < max && !s->done.empty()) {
    Request *r = s->done.front();
    s->done.pop_front();
    prores_server_result &o = out[n++];
    memset(&o, 0, sizeof o);
    o.id = r->id;
    o.kind = r->kind;
    fill_error(&o.error, r->err);
    if (r->err.ok()) {
      o.dev_out = r->dev_out;
      o.out_bytes = r->out_bytes;
      o.dev_alpha = r->dev_out_alpha;
      o.alpha_bytes = r->out_alpha_bytes;
    }
    o.header = r->hdr;
    o.submit_us = r->t_submit;
    o.start_us = r->t_start;
    o.done_us = r->t_done;
    o.batch_frames = r->batch_frames;
  }
  return n;
}

prores_status prores_server_release(prores_server *s, uint64_t id) {
  Request *r = nullptr;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end()) return PRORES_ERR_CONFIG;
    r = it->second;
    // Only a result the client has polled can be released.
    if (std::find(s->done.begin(), s->done.end(), r) != s->done.end()) return PRORES_ERR_CONFIG;
    s->live.erase(it);
    s->held -= r->hold;
    s->stats.held_bytes = s->held;
  }
  s->pool.put(r->dev_out, r->out_bytes + 64);
  s->pool.put(r->dev_out_alpha, r->out_alpha_bytes + 64);
  delete r;
  s->cv_work.notify_all();
  s->cv_ready.notify_all();
  return PRORES_OK;
}

void prores_server_get_stats(prores_server *s, prores_server_stats *out) {
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->stats;
  out->pending = s->pend[0].size() + s->pend[1].size();
}

}  // extern "C"
