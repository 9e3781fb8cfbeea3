// tszp_topo_serve.cpp -- request loop for the topology pipeline over the batch entry
// points (tszp_topo_batch_*): compress requests (2D / 3D fields of any size) and
// decompress requests (decode, then the topology restoration) arrive from any host
// thread and are admitted in arrival order while their streams fit an HBM budget.
// Several lanes -- each a worker thread with its own HIP stream, batch memory and
// restoration workspace -- take turns at the queue: while one lane's batch runs on the
// device, the next lane forms a batch of whatever arrived meanwhile and starts it, so
// consecutive batches overlap on the device.  No counterpart in the reference, whose
// callers run the per-field chain once per field (example/test_cuda_topology.cu:374-530,
// SZp/src/szp_cuda_topology.cu:77-772).
#include <hip/hip_runtime.h>

#include <algorithm>
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

thread_local std::string g_err;

int tfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr uint64_t kAlign = 256;
uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }
// restoration workspace per element of a lane's batch (mismatch lists, bitmaps, the
// mismatch map, relation masks, the smoothed field of the plateau search)
constexpr uint64_t kWsBytesPerElement = 32;

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
  const float* d_field = nullptr;  // compress: input
  float* d_out = nullptr;          // decompress: output values
  uint8_t* d_types = nullptr;      // decompress: output types (may be null)
  bool restore = true;             // decompress: restored (or decoded) values
  int ndim = 0;
  uint32_t dims[3] = {1, 1, 1};
  uint64_t n = 0;
  float eb = 0;
  int bs = 0;
  uint64_t src = 0;                // decompress: the compress ticket whose streams it reads
  // compress: arena range [off, off + mem): topology stream, its side-car, sort-position
  // stream, its side-car
  uint64_t off = 0, mem = 0, topo_cap = 0, sp_cap = 0, blk_words = 0;
  tszp_topo_info info{};
  tszp_restore_info rinfo{};
  int users = 0;                   // queued / running decompressions of these streams
  bool released = false;
  State state = kQueued;
  std::string err;
  double t_submit = 0, t_start = 0, t_done = 0;
};

struct Lane {
  hipStream_t st = nullptr;
  void* mem = nullptr;
  uint64_t mem_bytes = 0;
  tszp_workspace* ws = nullptr;
  std::thread th;
};

}  // namespace

struct tszp_topo_server {
  int device = 0;
  uint32_t max_batch = 0;
  uint64_t max_elems = 0;
  Arena arena;
  std::vector<Lane> lanes;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Request> req;
  uint64_t next_id = 1;
  uint64_t batches = 0, served = 0, reserved = 0;
  bool stop = false;

  uint8_t* topo_of(const Request& r) { return arena.base + r.off; }
  uint32_t* topo_blk_of(const Request& r) {
    return reinterpret_cast<uint32_t*>(arena.base + r.off + align_up(r.topo_cap));
  }
  uint8_t* sp_of(const Request& r) { return arena.base + r.off + align_up(r.topo_cap) + align_up(4 * r.blk_words); }
  uint32_t* sp_blk_of(const Request& r) {
    return reinterpret_cast<uint32_t*>(sp_of(r) + align_up(r.sp_cap));
  }
  void free_streams(Request& r) {
    if (r.mem) arena.release(r.off, r.mem);
    r.mem = 0;
  }
  void run(Lane& lane);
  // runs one lane's batch on the device (outside the lock); fills the requests' results
  void execute(Lane& lane, std::vector<uint64_t>& comp, std::vector<uint64_t>& dec);
};

namespace {

tszp_topo_batch_field field_of(const Request& r) {
  tszp_topo_batch_field f{};
  f.ndim = r.ndim;
  memcpy(f.dims, r.dims, sizeof f.dims);
  f.eb = r.eb;
  f.block_size = r.bs;
  return f;
}

}  // namespace

void tszp_topo_server::execute(Lane& lane, std::vector<uint64_t>& comp, std::vector<uint64_t>& dec) {
  // the tables are built under the lock (the request map may rehash meanwhile)
  std::vector<tszp_topo_batch_field> fc, fd;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (uint64_t id : comp) {
      Request& r = req[id];
      tszp_topo_batch_field f = field_of(r);
      f.d_field = r.d_field;
      f.d_topo = topo_of(r);
      f.topo_cap = r.topo_cap;
      f.d_topo_blk = topo_blk_of(r);
      f.d_sortpos = sp_of(r);
      f.sortpos_cap = r.sp_cap;
      f.d_sortpos_blk = sp_blk_of(r);
      fc.push_back(f);
    }
    for (uint64_t id : dec) {
      const Request& s = req[req[id].src];
      tszp_topo_batch_field f = field_of(s);
      f.d_topo = topo_of(s);
      f.topo_cap = s.topo_cap;
      f.d_topo_blk = topo_blk_of(s);
      f.d_sortpos = sp_of(s);
      f.sortpos_cap = s.sp_cap;
      f.d_sortpos_blk = sp_blk_of(s);
      fd.push_back(f);
    }
  }
  std::string ec, ed;
  if (!fc.empty()) {  // compression batch
    tszp_topo_batch* B = nullptr;
    int r = tszp_topo_batch_create_mem(fc.data(), (uint32_t)fc.size(), lane.mem, lane.mem_bytes, &B, lane.st);
    if (!r) r = tszp_topo_batch_compress(B, lane.st);
    if (!r) r = tszp_topo_batch_results(B, fc.data(), lane.st);
    if (r) ec = tszp_last_error();
    if (B) tszp_topo_batch_destroy(B);
  }
  std::vector<tszp_restore_info> ri(fd.size());
  if (!fd.empty()) {  // decompression batch: decode, restore where asked, copy out
    bool any_restore = false;
    {
      std::lock_guard<std::mutex> lk(mu);
      for (uint64_t id : dec) any_restore = any_restore || req[id].restore;
    }
    std::vector<Request> copy(dec.size());
    {
      std::lock_guard<std::mutex> lk(mu);
      for (size_t k = 0; k < dec.size(); k++) copy[k] = req[dec[k]];
    }
    tszp_topo_batch* B = nullptr;
    int r = tszp_topo_batch_create_mem(fd.data(), (uint32_t)fd.size(), lane.mem, lane.mem_bytes, &B, lane.st);
    if (!r) r = tszp_topo_batch_decompress(B, lane.st);
    if (!r && any_restore) r = tszp_topo_batch_restore_ws(B, 8, lane.ws, lane.st);
    if (!r) r = tszp_topo_batch_results(B, fd.data(), lane.st);
    for (size_t k = 0; !r && k < copy.size(); k++) {
      const Request& q = copy[k];
      r = tszp_topo_batch_copy(B, (uint32_t)k, q.restore ? 1 : 0, q.d_out, q.n, lane.st);
      if (!r && q.d_types) r = tszp_topo_batch_copy(B, (uint32_t)k, 2, q.d_types, q.n, lane.st);
      ri[k] = fd[k].restore;
    }
    if (!r && hipStreamSynchronize(lane.st) != hipSuccess) r = tfail(TSZP_EHIP, "topo server: stream error");
    if (r) ed = tszp_last_error();
    if (B) tszp_topo_batch_destroy(B);
  }
  const double t = now_us();
  std::lock_guard<std::mutex> lk(mu);
  for (size_t k = 0; k < comp.size(); k++) {
    Request& r = req[comp[k]];
    r.t_done = t;
    if (ec.empty()) {
      r.state = kDone;
      r.info = fc[k].info;
      r.bs = fc[k].block_size;  // the tuner's choice for block size 0
    } else {
      r.state = kFailed;
      r.err = ec;
      free_streams(r);
    }
  }
  for (size_t k = 0; k < dec.size(); k++) {
    Request& r = req[dec[k]];
    r.t_done = t;
    r.state = ed.empty() ? kDone : kFailed;
    if (ed.empty()) r.rinfo = ri[k];
    else r.err = ed;
    Request& s = req[r.src];
    if (--s.users == 0 && s.released) {
      free_streams(s);
      req.erase(r.src);
    }
  }
  batches++;
  served += comp.size() + dec.size();
}

void tszp_topo_server::run(Lane& lane) {
  (void)hipSetDevice(device);
  std::vector<uint64_t> comp, dec;
  for (;;) {
    comp.clear();
    dec.clear();
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        if (queue.empty()) {
          if (stop) return;
          cv_work.wait(lk);
          continue;
        }
        // admit in arrival order while the batch has room: a compression while its
        // streams fit the arena, a decompression once its source streams are written
        uint64_t elems = 0;
        while (!queue.empty() && comp.size() + dec.size() < max_batch) {
          const uint64_t id = queue.front();
          Request& r = req[id];
          const uint64_t n = r.compress ? r.n : req[r.src].n;
          if (n > max_elems) {
            r.state = kFailed;
            r.err = "request larger than the server's batch";
            r.t_done = now_us();
            if (!r.compress) req[r.src].users--;
            queue.pop_front();
            cv_done.notify_all();
            continue;
          }
          if (elems + n > max_elems) break;  // the batch is full
          if (r.compress) {
            if (!arena.alloc(r.mem, &r.off)) {
              if (align_up(r.mem) > arena.size) {  // can never fit
                r.state = kFailed;
                r.err = "request larger than the server's HBM budget";
                r.mem = 0;
                r.t_done = now_us();
                queue.pop_front();
                cv_done.notify_all();
                continue;
              }
              break;  // waits for releases
            }
          } else {
            const Request& s = req[r.src];
            if (s.state == kFailed) {
              r.state = kFailed;
              r.err = "source compression failed";
              r.t_done = now_us();
              req[r.src].users--;
              queue.pop_front();
              cv_done.notify_all();
              continue;
            }
            if (s.state != kDone) break;  // its streams are still being written
          }
          r.state = kRunning;
          r.t_start = now_us();
          (r.compress ? comp : dec).push_back(id);
          elems += n;
          queue.pop_front();
        }
        if (!comp.empty() || !dec.empty()) break;
        if (stop && !queue.empty() && req[queue.front()].compress) {
          // nothing runs that could release memory: what is left cannot be served
          bool busy = false;
          for (auto& kv : req) busy = busy || kv.second.state == kRunning;
          if (!busy) {
            const double t = now_us();
            for (uint64_t id : queue) {
              Request& r = req[id];
              r.state = kFailed;
              r.err = "server stopped before the request fit its HBM budget";
              r.mem = 0;
              r.t_done = t;
              if (!r.compress) req[r.src].users--;
            }
            queue.clear();
            cv_done.notify_all();
            cv_work.notify_all();
            return;
          }
        }
        cv_work.wait(lk);  // a release, a finished batch or a stop
      }
    }
    execute(lane, comp, dec);
    cv_done.notify_all();
    cv_work.notify_all();  // finished sources and freed streams may unblock the queue head
  }
}

extern "C" {

const char* tszp_topo_server_last_error(void) { return g_err.c_str(); }

int tszp_topo_server_create(uint64_t hbm_budget, uint32_t max_batch, uint64_t max_batch_elements, uint32_t lanes,
                            tszp_topo_server** out) {
  g_err.clear();
  if (!out || max_batch == 0 || max_batch > 65536 || max_batch_elements == 0 || max_batch_elements >= (1ull << 31) ||
      lanes == 0 || lanes > 8)
    return tfail(TSZP_EINVAL, "topo_server_create: bad arguments");
  *out = nullptr;
  auto* s = new tszp_topo_server();
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return tfail(TSZP_EHIP, "topo_server_create: no device");
  }
  s->max_batch = max_batch;
  s->max_elems = max_batch_elements;
  const uint64_t lane_mem = align_up(tszp_topo_batch_mem_bound(max_batch_elements, max_batch));
  s->reserved = (uint64_t)lanes * (lane_mem + kWsBytesPerElement * max_batch_elements);
  if (hbm_budget <= s->reserved + kAlign) {
    delete s;
    return tfail(TSZP_EINVAL, "topo_server_create: budget %llu below the lanes' working memory %llu",
                 (unsigned long long)hbm_budget, (unsigned long long)s->reserved);
  }
  s->arena.size = (hbm_budget - s->reserved) & ~(kAlign - 1);
  hipError_t e = hipMalloc((void**)&s->arena.base, s->arena.size);
  s->arena.free_[0] = s->arena.size;
  s->lanes.resize(lanes);
  for (Lane& l : s->lanes) {
    if (e != hipSuccess) break;
    l.mem_bytes = lane_mem;
    e = hipMalloc(&l.mem, lane_mem);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking);
    if (e == hipSuccess && tszp_workspace_create(&l.ws) != TSZP_OK) e = hipErrorOutOfMemory;
  }
  if (e != hipSuccess) {
    for (Lane& l : s->lanes) {
      if (l.ws) tszp_workspace_destroy(l.ws);
      if (l.st) (void)hipStreamDestroy(l.st);
      if (l.mem) (void)hipFree(l.mem);
    }
    if (s->arena.base) (void)hipFree(s->arena.base);
    delete s;
    return tfail(TSZP_EHIP, "topo_server_create: %s", hipGetErrorString(e));
  }
  for (Lane& l : s->lanes) l.th = std::thread([s, &l] { s->run(l); });
  *out = s;
  return TSZP_OK;
}

int tszp_topo_server_destroy(tszp_topo_server* s) {
  g_err.clear();
  if (!s) return tfail(TSZP_EINVAL, "topo_server_destroy: null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (Lane& l : s->lanes) l.th.join();  // the queue drains first
  for (Lane& l : s->lanes) {
    tszp_workspace_destroy(l.ws);
    (void)hipStreamDestroy(l.st);
    (void)hipFree(l.mem);
  }
  (void)hipFree(s->arena.base);
  delete s;
  return TSZP_OK;
}

int tszp_topo_server_compress(tszp_topo_server* s, const float* d_field, int ndim, const uint32_t* dims, float eb,
                              int block_size, uint64_t* ticket) {
  g_err.clear();
  if (!s || !d_field || !dims || !ticket) return tfail(TSZP_EINVAL, "topo_server_compress: bad arguments");
  if (ndim != 2 && ndim != 3) return tfail(TSZP_EINVAL, "topo_server_compress: ndim must be 2 or 3");
  Request r;
  r.compress = true;
  r.d_field = d_field;
  r.ndim = ndim;
  r.n = 1;
  for (int a = 0; a < ndim; a++) {
    if (dims[a] == 0) return tfail(TSZP_EINVAL, "topo_server_compress: zero dimension");
    r.dims[a] = dims[a];
    r.n *= dims[a];
  }
  if (r.n >= 0x7fffffffull) return tfail(TSZP_EINVAL, "topo_server_compress: more than 2^31 elements");
  if (!(eb > 0) || eb != eb || eb > 3.4e38f) return tfail(TSZP_EINVAL, "topo_server_compress: bad error bound");
  if (block_size < 0 || block_size > 257)
    return tfail(TSZP_EINVAL, "topo_server_compress: blockSize %d out of range [1, 257] (0: tuned)", block_size);
  r.eb = eb;
  r.bs = block_size;
  // block size 0 is chosen by the tuner among 16, 32, 64 and 128: room for the largest
  // of their streams, and the side-cars of block size 16
  const int cands[4] = {16, 32, 64, 128};
  for (int k = 0; k < 4; k++) {
    const int b = block_size ? block_size : cands[k];
    r.topo_cap = std::max<uint64_t>(r.topo_cap, tszp_topo_bound(r.n, b, 1));
    r.sp_cap = std::max<uint64_t>(r.sp_cap, tszp_sortpos_bound(r.n, b, 1));
    r.blk_words = std::max<uint64_t>(r.blk_words, tszp_blocks(r.n, b) + 1);
  }
  r.mem = align_up(r.topo_cap) + align_up(4 * r.blk_words) + align_up(r.sp_cap) + align_up(4 * r.blk_words);
  r.t_submit = now_us();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return tfail(TSZP_EINVAL, "topo_server_compress: server stopping");
    *ticket = s->next_id++;
    s->req[*ticket] = r;
    s->queue.push_back(*ticket);
  }
  s->cv_work.notify_one();
  return TSZP_OK;
}

int tszp_topo_server_decompress(tszp_topo_server* s, uint64_t src, float* d_out, uint8_t* d_types, int restore,
                                uint64_t* ticket) {
  g_err.clear();
  if (!s || !d_out || !ticket) return tfail(TSZP_EINVAL, "topo_server_decompress: bad arguments");
  Request r;
  r.compress = false;
  r.d_out = d_out;
  r.d_types = d_types;
  r.restore = restore != 0;
  r.src = src;
  r.t_submit = now_us();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->req.find(src);
    if (it == s->req.end() || !it->second.compress || it->second.released)
      return tfail(TSZP_EINVAL, "topo_server_decompress: %llu is not a live compress ticket",
                   (unsigned long long)src);
    if (s->stop) return tfail(TSZP_EINVAL, "topo_server_decompress: server stopping");
    it->second.users++;
    r.n = it->second.n;
    *ticket = s->next_id++;
    s->req[*ticket] = r;
    s->queue.push_back(*ticket);
  }
  s->cv_work.notify_one();
  return TSZP_OK;
}

int tszp_topo_server_wait(tszp_topo_server* s, uint64_t ticket, tszp_ticket_info* info) {
  g_err.clear();
  if (!s || !info) return tfail(TSZP_EINVAL, "topo_server_wait: bad arguments");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->req.find(ticket) == s->req.end())
    return tfail(TSZP_EINVAL, "topo_server_wait: unknown ticket %llu", (unsigned long long)ticket);
  s->cv_done.wait(lk, [&] {
    auto f = s->req.find(ticket);
    return f == s->req.end() || f->second.state == kDone || f->second.state == kFailed;
  });
  auto it = s->req.find(ticket);
  if (it == s->req.end())
    return tfail(TSZP_EINVAL, "topo_server_wait: ticket %llu was released", (unsigned long long)ticket);
  const Request& r = it->second;
  info->stream_bytes = r.compress ? r.info.topo_bytes + r.info.sortpos_bytes : 0;
  info->queue_us = r.t_start > 0 ? r.t_start - r.t_submit : r.t_done - r.t_submit;
  info->latency_us = r.t_done - r.t_submit;
  if (r.state == kFailed) return tfail(TSZP_EDEVICE, "ticket %llu: %s", (unsigned long long)ticket, r.err.c_str());
  return TSZP_OK;
}

int tszp_topo_server_result(tszp_topo_server* s, uint64_t ticket, tszp_topo_info* info, tszp_restore_info* rinfo,
                            int* block_size) {
  g_err.clear();
  if (!s) return tfail(TSZP_EINVAL, "topo_server_result: null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->req.find(ticket);
  if (it == s->req.end() || it->second.state != kDone)
    return tfail(TSZP_EINVAL, "topo_server_result: %llu is not a finished ticket", (unsigned long long)ticket);
  const Request& r = it->second;
  if (info) *info = r.compress ? r.info : s->req[r.src].info;
  if (rinfo) *rinfo = r.rinfo;
  if (block_size) *block_size = r.compress ? r.bs : s->req[r.src].bs;
  return TSZP_OK;
}

int tszp_topo_server_streams(tszp_topo_server* s, uint64_t ticket, const uint8_t** d_topo, const uint32_t** d_topo_blk,
                             const uint8_t** d_sortpos, const uint32_t** d_sortpos_blk) {
  g_err.clear();
  if (!s) return tfail(TSZP_EINVAL, "topo_server_streams: null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->req.find(ticket);
  if (it == s->req.end() || !it->second.compress || it->second.state != kDone || it->second.released)
    return tfail(TSZP_EINVAL, "topo_server_streams: %llu is not a finished compress ticket",
                 (unsigned long long)ticket);
  const Request& r = it->second;
  if (d_topo) *d_topo = s->topo_of(r);
  if (d_topo_blk) *d_topo_blk = s->topo_blk_of(r);
  if (d_sortpos) *d_sortpos = s->sp_of(r);
  if (d_sortpos_blk) *d_sortpos_blk = s->sp_blk_of(r);
  return TSZP_OK;
}

int tszp_topo_server_release(tszp_topo_server* s, uint64_t ticket) {
  g_err.clear();
  if (!s) return tfail(TSZP_EINVAL, "topo_server_release: null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->req.find(ticket);
    if (it == s->req.end() || it->second.released)
      return tfail(TSZP_EINVAL, "topo_server_release: unknown ticket %llu", (unsigned long long)ticket);
    Request& r = it->second;
    if (r.state != kDone && r.state != kFailed)
      return tfail(TSZP_EINVAL, "topo_server_release: ticket %llu is not finished", (unsigned long long)ticket);
    r.released = true;
    if (r.users == 0) {
      s->free_streams(r);
      s->req.erase(it);
    }
  }
  s->cv_work.notify_all();  // a request waiting for memory may fit now
  return TSZP_OK;
}

int tszp_topo_server_stats(tszp_topo_server* s, tszp_server_stats_t* out) {
  g_err.clear();
  if (!s || !out) return tfail(TSZP_EINVAL, "topo_server_stats: bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  out->batches = s->batches;
  out->requests = s->served;
  out->queued = s->queue.size();
  out->hbm_in_use = s->arena.used + s->reserved;
  out->hbm_peak = s->arena.peak + s->reserved;
  out->hbm_budget = s->arena.size + s->reserved;
  return TSZP_OK;
}

}  // extern "C"
