// tszp_serve.cpp -- request loop over the batch entry points (tszp_plain_*_batch):
// compress / decompress requests of any size arrive from any host thread, are admitted
// in arrival order while their streams fit an HBM budget, and every request waiting
// when the previous batch finishes goes into the next one (continuous batching).
// Streams and side-cars live in one device arena owned by the server until their
// ticket is released.  No counterpart in the reference (its callers run one field per
// call, SZp/src/szp_cuda_compress.cu:718-723).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/toposzp_hip.h"

namespace {

thread_local std::string g_serve_err;

int serve_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_serve_err = buf;
  return code;
}

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr uint64_t kAlign = 256;
uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation; free ranges kept coalesced.
struct Arena {
  uint8_t* base = nullptr;
  uint64_t size = 0, used = 0, peak = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length

  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const uint64_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used += bytes;
      if (used > peak) peak = used;
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = align_up(bytes);
    used -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
};

enum State { kQueued, kRunning, kDone, kFailed };

struct Request {
  bool compress = true;
  float* d_field = nullptr;  // compress: input; decompress: output
  uint64_t n = 0;
  float eb = 0;
  int bs = 0;
  uint64_t src = 0;          // decompress: the compress ticket whose stream it reads
  uint64_t off = 0, mem = 0; // compress: arena range of stream + side-car
  uint64_t cap = 0, bytes = 0;
  int users = 0;             // queued / running decompressions of this stream
  bool released = false;
  State state = kQueued;
  std::string err;
  double t_submit = 0, t_start = 0, t_done = 0;
};

}  // namespace

struct tszp_server {
  int device = 0;
  hipStream_t st = nullptr;
  uint32_t max_batch = 0;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Request> req;
  uint64_t next_id = 1;
  uint64_t batches = 0, served = 0;
  bool stop = false;
  std::thread worker;

  uint8_t* stream_of(const Request& r) { return arena.base + r.off; }
  uint32_t* blk_of(const Request& r) {
    return reinterpret_cast<uint32_t*>(arena.base + r.off + align_up(r.cap));
  }
  void free_stream(Request& r) {
    if (r.mem) arena.release(r.off, r.mem);
    r.mem = 0;
  }
  void run();
};

void tszp_server::run() {
  (void)hipSetDevice(device);
  std::vector<uint64_t> ids, comp, dec;
  std::vector<tszp_batch_field> fc, fd;
  for (;;) {
    ids.clear();
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !queue.empty(); });
      if (queue.empty()) return;  // stop requested and nothing left
      // admit in arrival order: a compression while its stream fits the arena, a
      // decompression once its source stream is written
      while (!queue.empty() && ids.size() < max_batch) {
        const uint64_t id = queue.front();
        Request& r = req[id];
        if (r.compress) {
          const uint64_t need = align_up(r.cap) + align_up(4 * (tszp_blocks(r.n, r.bs ? r.bs : 16) + 1));
          if (!arena.alloc(need, &r.off)) {
            if (need > arena.size) {  // can never fit
              r.state = kFailed;
              r.err = "request larger than the server's HBM budget";
              r.t_done = now_us();
              queue.pop_front();
              cv_done.notify_all();
              continue;
            }
            break;  // wait for releases
          }
          r.mem = need;
        } else {
          const Request& s = req[r.src];
          if (s.state != kDone && s.state != kFailed) break;  // its stream is in this or a later batch
        }
        r.state = kRunning;
        r.t_start = now_us();
        ids.push_back(id);
        queue.pop_front();
      }
      if (ids.empty()) {  // the head waits for memory
        if (stop) {  // no release will come: fail what is left
          const double t = now_us();
          for (uint64_t id : queue) {
            Request& r = req[id];
            r.state = kFailed;
            r.err = "server stopped before the request fit its HBM budget";
            r.t_done = t;
          }
          queue.clear();
          cv_done.notify_all();
          return;
        }
        cv_work.wait(lk);  // a release or a stop
        continue;
      }
      comp.clear();
      dec.clear();
      fc.clear();
      fd.clear();
      for (uint64_t id : ids) {
        Request& r = req[id];
        if (r.compress) {
          tszp_batch_field f{};
          f.d_field = r.d_field;
          f.n = r.n;
          f.eb = r.eb;
          f.block_size = r.bs;
          f.d_stream = stream_of(r);
          f.stream_cap = r.cap;
          f.d_blk = blk_of(r);
          fc.push_back(f);
          comp.push_back(id);
        } else {
          Request& s = req[r.src];
          if (s.state == kFailed) {
            r.state = kFailed;
            r.err = "source compression failed";
            continue;
          }
          tszp_batch_field f{};
          f.d_field = r.d_field;
          f.n = s.n;
          f.eb = s.eb;
          f.block_size = s.bs;
          f.d_stream = stream_of(s);
          f.stream_cap = s.cap;
          f.d_blk = blk_of(s);
          fd.push_back(f);
          dec.push_back(id);
        }
      }
    }
    // the GPU work of the batch, outside the lock: new requests queue meanwhile
    std::string ec, ed;
    bool tune = false;
    for (const tszp_batch_field& f : fc) tune = tune || f.block_size == 0;
    if (tune && tszp_batch_tune(fc.data(), (uint32_t)fc.size(), st) != TSZP_OK) ec = tszp_last_error();
    if (!fc.empty() && ec.empty() && tszp_plain_compress_batch(fc.data(), (uint32_t)fc.size(), st) != TSZP_OK)
      ec = tszp_last_error();
    if (!fd.empty() && tszp_plain_decompress_batch(fd.data(), (uint32_t)fd.size(), st) != TSZP_OK)
      ed = tszp_last_error();
    const double t = now_us();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (size_t k = 0; k < comp.size(); k++) {
        Request& r = req[comp[k]];
        r.t_done = t;
        if (ec.empty()) {
          r.state = kDone;
          r.bytes = fc[k].stream_bytes;
          r.bs = fc[k].block_size;  // the tuner's choice for block_size 0
        } else {
          r.state = kFailed;
          r.err = ec;
          free_stream(r);
        }
      }
      for (size_t k = 0; k < dec.size(); k++) {
        Request& r = req[dec[k]];
        r.t_done = t;
        r.state = ed.empty() ? kDone : kFailed;
        if (!ed.empty()) r.err = ed;
      }
      for (uint64_t id : ids) {
        Request& r = req[id];
        if (r.compress) continue;
        if (r.state == kFailed && r.t_done == 0) r.t_done = t;
        Request& s = req[r.src];
        if (--s.users == 0 && s.released) {
          free_stream(s);
          req.erase(r.src);
        }
      }
      batches++;
      served += ids.size();
    }
    cv_done.notify_all();
  }
}

extern "C" {

const char* tszp_server_last_error(void) { return g_serve_err.c_str(); }

int tszp_server_create(uint64_t hbm_budget, uint32_t max_batch, tszp_server** out) {
  g_serve_err.clear();
  if (!out || hbm_budget < kAlign || max_batch == 0)
    return serve_fail(TSZP_EINVAL, "server_create: bad arguments");
  tszp_server* s = new tszp_server();
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return serve_fail(TSZP_EHIP, "server_create: no device");
  }
  s->max_batch = max_batch;
  s->arena.size = hbm_budget & ~(kAlign - 1);
  hipError_t e = hipMalloc((void**)&s->arena.base, s->arena.size);
  if (e != hipSuccess) {
    delete s;
    return serve_fail(TSZP_EHIP, "server_create: hipMalloc(%llu): %s", (unsigned long long)hbm_budget,
                      hipGetErrorString(e));
  }
  s->arena.free_[0] = s->arena.size;
  e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)hipFree(s->arena.base);
    delete s;
    return serve_fail(TSZP_EHIP, "server_create: hipStreamCreate: %s", hipGetErrorString(e));
  }
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return TSZP_OK;
}

int tszp_server_destroy(tszp_server* s) {
  g_serve_err.clear();
  if (!s) return serve_fail(TSZP_EINVAL, "server_destroy: null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->worker.join();  // the queue drains first; a head that never fits fails the wait
  (void)hipStreamDestroy(s->st);
  (void)hipFree(s->arena.base);
  delete s;
  return TSZP_OK;
}

int tszp_server_compress(tszp_server* s, const float* d_field, uint64_t n, float eb, int block_size,
                         uint64_t* ticket) {
  g_serve_err.clear();
  if (!s || !d_field || !ticket) return serve_fail(TSZP_EINVAL, "server_compress: bad arguments");
  if (n == 0 || n >= 0x7fffffffull) return serve_fail(TSZP_EINVAL, "server_compress: element count out of range");
  if (!(eb > 0) || eb != eb || eb > 3.4e38f) return serve_fail(TSZP_EINVAL, "server_compress: bad error bound");
  if (block_size != 0 && (block_size < 2 || block_size > 1025))
    return serve_fail(TSZP_EINVAL, "server_compress: blockSize %d out of range [2, 1025] (0: tuned)", block_size);
  Request r;
  r.compress = true;
  r.d_field = const_cast<float*>(d_field);
  r.n = n;
  r.eb = eb;
  r.bs = block_size;
  // block size 0 is chosen by the tuner among 16, 32, 64 and 128: room for the largest
  // of their streams (and, at admission, for the side-car of block size 16)
  r.cap = block_size ? tszp_plain_bound(n, block_size, 1) : 0;
  for (int c : {16, 32, 64, 128})
    if (!block_size && tszp_plain_bound(n, c, 1) > r.cap) r.cap = tszp_plain_bound(n, c, 1);
  r.t_submit = now_us();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return serve_fail(TSZP_EINVAL, "server_compress: server stopping");
    *ticket = s->next_id++;
    s->req[*ticket] = r;
    s->queue.push_back(*ticket);
  }
  s->cv_work.notify_one();
  return TSZP_OK;
}

int tszp_server_decompress(tszp_server* s, uint64_t src, float* d_out, uint64_t* ticket) {
  g_serve_err.clear();
  if (!s || !d_out || !ticket) return serve_fail(TSZP_EINVAL, "server_decompress: bad arguments");
  Request r;
  r.compress = false;
  r.d_field = d_out;
  r.src = src;
  r.t_submit = now_us();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->req.find(src);
    if (it == s->req.end() || !it->second.compress || it->second.released)
      return serve_fail(TSZP_EINVAL, "server_decompress: %llu is not a live compress ticket",
                        (unsigned long long)src);
    if (s->stop) return serve_fail(TSZP_EINVAL, "server_decompress: server stopping");
    it->second.users++;
    r.n = it->second.n;
    *ticket = s->next_id++;
    s->req[*ticket] = r;
    s->queue.push_back(*ticket);
  }
  s->cv_work.notify_one();
  return TSZP_OK;
}

int tszp_server_wait(tszp_server* s, uint64_t ticket, tszp_ticket_info* info) {
  g_serve_err.clear();
  if (!s || !info) return serve_fail(TSZP_EINVAL, "server_wait: bad arguments");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->req.find(ticket);
  if (it == s->req.end()) return serve_fail(TSZP_EINVAL, "server_wait: unknown ticket %llu", (unsigned long long)ticket);
  s->cv_done.wait(lk, [&] {
    auto f = s->req.find(ticket);
    return f == s->req.end() || f->second.state == kDone || f->second.state == kFailed;
  });
  it = s->req.find(ticket);
  if (it == s->req.end()) return serve_fail(TSZP_EINVAL, "server_wait: ticket %llu was released", (unsigned long long)ticket);
  const Request& r = it->second;
  info->stream_bytes = r.bytes;
  info->queue_us = r.t_start > 0 ? r.t_start - r.t_submit : r.t_done - r.t_submit;
  info->latency_us = r.t_done - r.t_submit;
  if (r.state == kFailed) return serve_fail(TSZP_EDEVICE, "ticket %llu: %s", (unsigned long long)ticket, r.err.c_str());
  return TSZP_OK;
}

int tszp_server_stream(tszp_server* s, uint64_t ticket, const uint8_t** d_stream, const uint32_t** d_blk,
                       uint64_t* bytes, int* block_size) {
  g_serve_err.clear();
  if (!s) return serve_fail(TSZP_EINVAL, "server_stream: null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->req.find(ticket);
  if (it == s->req.end() || !it->second.compress || it->second.state != kDone || it->second.released)
    return serve_fail(TSZP_EINVAL, "server_stream: %llu is not a finished compress ticket",
                      (unsigned long long)ticket);
  const Request& r = it->second;
  if (d_stream) *d_stream = s->stream_of(r);
  if (d_blk) *d_blk = s->blk_of(r);
  if (bytes) *bytes = r.bytes;
  if (block_size) *block_size = r.bs;
  return TSZP_OK;
}

int tszp_server_copy_stream(tszp_server* s, uint64_t ticket, uint8_t* d_dst, uint64_t cap, void* stream) {
  g_serve_err.clear();
  if (!s || !d_dst) return serve_fail(TSZP_EINVAL, "server_copy_stream: bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);  // the ticket cannot be released meanwhile
  auto it = s->req.find(ticket);
  if (it == s->req.end() || !it->second.compress || it->second.state != kDone || it->second.released)
    return serve_fail(TSZP_EINVAL, "server_copy_stream: %llu is not a finished compress ticket",
                      (unsigned long long)ticket);
  const Request& r = it->second;
  if (cap < r.bytes) return serve_fail(TSZP_ECAP, "server_copy_stream: %llu bytes do not fit %llu",
                                       (unsigned long long)r.bytes, (unsigned long long)cap);
  hipError_t e = hipMemcpyAsync(d_dst, s->stream_of(r), r.bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return serve_fail(TSZP_EHIP, "server_copy_stream: %s", hipGetErrorString(e));
  return TSZP_OK;
}

int tszp_server_release(tszp_server* s, uint64_t ticket) {
  g_serve_err.clear();
  if (!s) return serve_fail(TSZP_EINVAL, "server_release: null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->req.find(ticket);
    if (it == s->req.end() || it->second.released)
      return serve_fail(TSZP_EINVAL, "server_release: unknown ticket %llu", (unsigned long long)ticket);
    Request& r = it->second;
    if (r.state != kDone && r.state != kFailed)
      return serve_fail(TSZP_EINVAL, "server_release: ticket %llu is not finished", (unsigned long long)ticket);
    r.released = true;
    if (r.users == 0) {
      s->free_stream(r);
      s->req.erase(it);
    }
  }
  s->cv_work.notify_one();  // a request waiting for memory may fit now
  return TSZP_OK;
}

int tszp_server_stats(tszp_server* s, tszp_server_stats_t* out) {
  g_serve_err.clear();
  if (!s || !out) return serve_fail(TSZP_EINVAL, "server_stats: bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  out->batches = s->batches;
  out->requests = s->served;
  out->queued = s->queue.size();
  out->hbm_in_use = s->arena.used;
  out->hbm_peak = s->arena.peak;
  out->hbm_budget = s->arena.size;
  return TSZP_OK;
}

}  // extern "C"
