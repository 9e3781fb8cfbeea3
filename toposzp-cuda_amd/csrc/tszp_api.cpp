// tszp_api.cpp -- host side of libtoposzp_hip.so: the native device-pointer API
// (include/toposzp_hip.h) and the reference-compatible host-pointer drop-in
// (include/szp_cuda_*.h), both driving the HIP kernels of this directory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/szp_cuda_compress.h"
#include "../../include/szp_cuda_decompress.h"
#include "../../include/szp_cuda_topology.h"
#include "../../include/szp_cuda_topology_3d.h"
#include "../../include/toposzp_hip.h"
#include "tszp_kernels.h"

using namespace tszp;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(TSZP_EHIP, "%s: %s", #x, hipGetErrorString(e_));   \
  } while (0)

// ---- per-device context: a grow-only workspace, a stream for drop-in calls ------------
enum Slot {
  kWsStatus, kWsResult, kWsCpKey, kWsCpPay, kWsKeyAlt, kWsPayAlt, kWsHiKey, kWsHiAlt, kWsHist, kWsSortSt,
  kWsSp, kWsSpBlk, kWsIn, kWsOut, kWsOut2, kWsBlk, kWsTypes, kWsStream, kWsStream2, kWsCps, kWsMinMax,
  kWsPpD, kWsPpBad, kWsPpBad2, kWsPpAct, kWsPpBits, kWsPpTBits, kWsPpTiles, kWsPpMisc, kWsPpSt,
  kWsBinMeta, kWsBinKey, kWsBinPay, kWsBinBits, kWsPpPart, kWsPpBadmap, kWsQuality,
  kWsQ, kWsMoves, kWsSaddleList, kWsSaddleMeta, kWsPpDone, kWsBatchTab, kWsBatchMap, kWsBatchRes, kWsBatchSt,
  kWsPpFields, kWsPpSm0, kWsPpSm1, kWsPpEsc, kWsPpStage, kWsPpCtl, kWsElemEnds, kWsPpOrd, kWsPpFflags, kWsCount
};

constexpr size_t kHostSmall = 512;  // bytes of a context's pinned read buffer

struct Ctx {
  int device = -1;
  hipStream_t stream = nullptr;
  std::mutex mu;
  void* ptr[kWsCount] = {};
  size_t cap[kWsCount] = {};
  CompressResult* host_res = nullptr;  // pinned
  uint32_t* host_err = nullptr;        // pinned: a device error word read at a later synchronisation
  uint32_t* host_small = nullptr;      // pinned: small device reads (counters, error words)
  void* host_batch = nullptr;          // pinned: batch tables and results
  size_t host_batch_cap = 0;

  int host_batch_get(size_t bytes, void** out) {
    if (host_batch_cap < bytes) {
      if (host_batch) {
        hipError_t e = hipHostFree(host_batch);
        host_batch = nullptr;
        host_batch_cap = 0;
        if (e != hipSuccess) return fail(TSZP_EHIP, "hipHostFree: %s", hipGetErrorString(e));
      }
      const size_t want = bytes + bytes / 4 + 4096;
      hipError_t e = hipHostMalloc(&host_batch, want, hipHostMallocDefault);
      if (e != hipSuccess) return fail(TSZP_EHIP, "hipHostMalloc(%zu): %s", want, hipGetErrorString(e));
      host_batch_cap = want;
    }
    *out = host_batch;
    return TSZP_OK;
  }

  int get(int slot, size_t bytes, void** out) {
    if (bytes == 0) bytes = 4;
    if (cap[slot] < bytes) {
      if (ptr[slot]) {
        hipError_t e = hipFree(ptr[slot]);
        ptr[slot] = nullptr;
        cap[slot] = 0;
        if (e != hipSuccess) return fail(TSZP_EHIP, "hipFree: %s", hipGetErrorString(e));
      }
      size_t want = bytes + bytes / 4 + 256;
      hipError_t e = hipMalloc(&ptr[slot], want);
      if (e != hipSuccess) return fail(TSZP_EHIP, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
      cap[slot] = want;
    }
    *out = ptr[slot];
    return TSZP_OK;
  }
};

std::mutex g_ctx_mu;
std::vector<Ctx*> g_ctx;

// Optional per-kernel timing: HIP events recorded on the launch stream around the
// kernels of each stage, accumulated until tszp_prof_reset().
enum ProfKind { kProfCodecTopo, kProfCodecInt, kProfCodecPlain, kProfDecode, kProfSort, kProfRank, kProfRestore,
                kProfScatter, kProfSolve, kProfQuery, kProfElem, kProfElemDecode, kProfSaddle, kProfBatch, kProfBatchDecode,
                kProfCount = kProfTopoBatchSort + 1 };
int g_prof = 0;
double g_prof_ms[kProfCount];
uint64_t g_prof_n[kProfCount];

struct ProfScope {
  hipEvent_t a = nullptr, b = nullptr;
  int kind;
  hipStream_t st;
  ProfScope(int k, hipStream_t s) : kind(k), st(s) {
    if (!g_prof) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(a, st);
  }
  void stop() {
    if (a) (void)hipEventRecord(b, st);
  }
  // call after the stream has been synchronized
  void collect() {
    if (!a) return;
    float ms = 0.f;
    if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) {
      g_prof_ms[kind] += ms;
      g_prof_n[kind]++;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    a = b = nullptr;
  }
  ~ProfScope() { collect(); }
};

// Times individual launches of one kind: event pairs are recorded around each launch
// and read when the timer goes out of scope (after the caller has synchronized).
struct EventTimer : LaunchTimer {
  int kind;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  explicit EventTimer(int k) : kind(k) {}
  void begin(hipStream_t st) override {
    if (!g_prof) return;
    hipEvent_t a = nullptr;
    if (hipEventCreate(&a) != hipSuccess) return;
    (void)hipEventRecord(a, st);
    ev.emplace_back(a, nullptr);
  }
  void end(hipStream_t st) override {
    if (!g_prof || ev.empty() || ev.back().second) return;
    hipEvent_t b = nullptr;
    if (hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(b, st);
    ev.back().second = b;
  }
  ~EventTimer() override {
    for (auto& p : ev) {
      float ms = 0.f;
      if (p.second && hipEventSynchronize(p.second) == hipSuccess &&
          hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) {
        g_prof_ms[kind] += ms;
        g_prof_n[kind]++;
      }
      (void)hipEventDestroy(p.first);
      if (p.second) (void)hipEventDestroy(p.second);
    }
  }
};

int get_ctx(Ctx** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(TSZP_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if ((int)g_ctx.size() <= dev) g_ctx.resize(dev + 1, nullptr);
  if (!g_ctx[dev]) {
    Ctx* c = new Ctx();
    c->device = dev;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(TSZP_EHIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    e = hipHostMalloc((void**)&c->host_res, sizeof(CompressResult), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_err, 64, hipHostMallocDefault);
    if (e != hipSuccess) { delete c; return fail(TSZP_EHIP, "hipHostMalloc: %s", hipGetErrorString(e)); }
    g_ctx[dev] = c;
  }
  *out = g_ctx[dev];
  return TSZP_OK;
}

// A small device read (counters, error words) through the context's pinned buffer: a
// pageable target makes the copy stage through a bounce buffer (measured: the
// restoration's dozen counter reads per call took 0.1 ms longer that way).
int read_small(Ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes > kHostSmall) return fail(TSZP_EINVAL, "read_small: %zu bytes", bytes);
  if (!c->host_small) {
    const hipError_t e = hipHostMalloc((void**)&c->host_small, kHostSmall, hipHostMallocDefault);
    if (e != hipSuccess) {
      c->host_small = nullptr;
      return fail(TSZP_EHIP, "hipHostMalloc: %s", hipGetErrorString(e));
    }
  }
  HIPCHK(hipMemcpyAsync(c->host_small, src, bytes, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  memcpy(dst, c->host_small, bytes);
  return TSZP_OK;
}

#define WS(slot, bytes, ptrvar)                                     \
  do {                                                              \
    void* p_;                                                       \
    int r_ = c->get((slot), (bytes), &p_);                          \
    if (r_) return r_;                                              \
    ptrvar = reinterpret_cast<decltype(ptrvar)>(p_);                \
  } while (0)

FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  const uint64_t num = 1ull << (31 + l);
  FastDiv f;
  f.mul = (uint32_t)((num + d - 1) / d);
  f.shift = 31 + l;
  return f;
}

Grid make_grid(int ndim, const uint32_t* dims) {
  Grid g{};
  g.ndim = ndim;
  g.d0 = dims[0];
  g.d1 = ndim >= 2 ? dims[1] : 1;
  g.d2 = ndim == 3 ? dims[2] : 1;
  g.plane = g.d1 * g.d2;
  g.div_row = make_fastdiv(ndim == 2 ? g.d1 : g.d2);
  g.div_plane = make_fastdiv(g.plane ? g.plane : 1);
  return g;
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

size_t block_worst(int bs, bool topo) {
  const size_t n = (size_t)bs - 1;
  return 5 + (n + 7) / 8 + 4 * n + (topo ? (2 * (size_t)bs + 7) / 8 : 0);
}

int check_bs(int bs, bool topo) {
  if (topo ? (bs < 1 || bs > 257) : (bs < 2 || bs > 1025))
    return fail(TSZP_EINVAL, "blockSize %d out of range [%d, %d]", bs, topo ? 1 : 2, topo ? 257 : 1025);
  return TSZP_OK;
}

// zero the look-back words + tile counter + result of one codec launch
int codec_prepare(Ctx* c, uint32_t ntiles, uint64_t** st_bytes, uint64_t** st_cnt, uint32_t** counter,
                  CompressResult** res, hipStream_t st) {
  uint64_t* s;
  const size_t words = 2 * (size_t)ntiles + 2;
  WS(kWsStatus, words * 8, s);
  CompressResult* r;
  WS(kWsResult, sizeof(CompressResult), r);
  FillRanges fr{};
  fr.add(reinterpret_cast<uint32_t*>(s), 0u, words * 2);
  fr.add(reinterpret_cast<uint32_t*>(r), 0u, sizeof(CompressResult) / 4);
  HIPCHK(launch_fill_multi(fr, st));
  *st_bytes = s;
  *st_cnt = s + ntiles;
  *counter = reinterpret_cast<uint32_t*>(s + 2 * ntiles);
  *res = r;
  return TSZP_OK;
}

int read_result(Ctx* c, CompressResult* d_res, hipStream_t st, CompressResult* out) {
  HIPCHK(hipMemcpyAsync(c->host_res, d_res, sizeof(CompressResult), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *out = *c->host_res;
  if (out->error)
    return fail(TSZP_EDEVICE, "device error bits 0x%x (%s%s%s)", out->error,
                (out->error & TSZP_ERR_LOOKBACK) ? "look-back timeout " : "",
                (out->error & TSZP_ERR_STREAM) ? "corrupt stream " : "",
                (out->error & TSZP_ERR_OVERFLOW) ? "capacity " : "");
  return TSZP_OK;
}

// One codec pass (compress) over `n` values; writes blocks at d_stream + hdr and
// the block side-car.  *payload_bytes receives the payload size.
int run_codec(Ctx* c, int mode, const void* d_in, uint32_t n, int bs, float eb, double inv, const Grid* grid,
              const uint8_t* d_tmap, uint8_t* d_stream, size_t cap, size_t hdr, uint32_t* d_blk,
              uint32_t* cp_key, uint32_t* cp_pay, uint32_t cp_cap, hipStream_t st, CompressResult* out) {
  int S, EPL;
  uint32_t bpt;
  block_geometry((uint32_t)bs, &S, &EPL, &bpt);
  const uint32_t nb = cdiv(n, bs);
  const uint32_t ntiles = cdiv(nb, bpt);
  if (((uintptr_t)(d_stream + hdr)) & 3u) return fail(TSZP_EINVAL, "stream payload must be 4-byte aligned");
  if (cap < hdr + 4) return fail(TSZP_ECAP, "stream capacity %zu too small", cap);
  CompressArgs a{};
  uint64_t *stb, *stc;
  uint32_t* counter;
  CompressResult* res;
  int r = codec_prepare(c, ntiles, &stb, &stc, &counter, &res, st);
  if (r) return r;
  a.in = d_in;
  a.n = n;
  a.bs = (uint32_t)bs;
  a.nb = nb;
  a.ntiles = ntiles;
  a.eb = eb;
  a.inv = inv;
  if (grid) a.grid = *grid;
  a.lds_rows = getenv("TSZP_CODEC_LDS_ROWS") ? 1u : 0u;  // the same bytes either way
  a.tmap = d_tmap;
  a.payload = d_stream + hdr;
  uint64_t pc = cap - hdr - 4;  // keep 4 readable bytes past the end
  a.cap = pc > 0xffffffffull ? 0xffffffffull : pc;
  a.blk = d_blk;
  a.st_bytes = stb;
  a.st_cnt = stc;
  a.tile_counter = dyn_counter(counter);
  a.cp_key = cp_key;
  a.cp_pay = cp_pay;
  a.cp_cap = cp_cap;
  a.res = res;
  ProfScope prof(mode == kModeInt ? kProfCodecInt : mode == kModePlain ? kProfCodecPlain : kProfCodecTopo, st);
  HIPCHK(launch_compress(mode, a, st));
  prof.stop();
  r = read_result(c, res, st, out);
  prof.collect();
  return r;
}

SortWork sort_work(Ctx* c, uint32_t count, bool hi, int* rc) {
  SortWork w{};
  *rc = TSZP_OK;
  auto get = [&](int slot, size_t bytes) -> void* {
    void* p = nullptr;
    if (*rc == TSZP_OK) *rc = c->get(slot, bytes, &p);
    return p;
  };
  w.keys_alt = (uint32_t*)get(kWsKeyAlt, (size_t)count * 4);
  w.pay_alt = (uint32_t*)get(kWsPayAlt, (size_t)count * 4);
  w.hi_alt = hi ? (uint32_t*)get(kWsHiAlt, (size_t)count * 4) : nullptr;
  w.hist = (uint32_t*)get(kWsHist, sort_hist_words(count) * 4);
  const size_t stw = sort_status_words(count) + 16 / 2 + 1;
  uint64_t* s = (uint64_t*)get(kWsSortSt, stw * 8 + 64);
  if (*rc) return w;
  w.st = s;
  w.counters = reinterpret_cast<uint32_t*>(s + sort_status_words(count));
  w.err = w.counters + 15;
  return w;
}

// sort (keys[, hi], pay) and write dense in-bin ranks to out[pay]
int sort_and_rank(Ctx* c, uint32_t* keys, uint32_t* hi, uint32_t* pay, uint32_t count, float eb, double inv,
                  int32_t* out, uint32_t out_cap, hipStream_t st) {
  if (!count) return TSZP_OK;
  int rc;
  SortWork w = sort_work(c, count, hi != nullptr, &rc);
  if (rc) return rc;
  const size_t stw = sort_status_words(count) + 8;
  HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(w.st), 0u, stw * 2, st));
  EventTimer scatter_timer(kProfScatter);
  ProfScope ps(kProfSort, st);
  HIPCHK(radix_sort_pairs(keys, hi, pay, count, w, st, &scatter_timer));
  ps.stop();
  ProfScope pr(kProfRank, st);
  HIPCHK(launch_rank(keys, hi, pay, count, eb, inv, out, out_cap, w, st));
  pr.stop();
  uint32_t herr = 0;
  if (int rr = read_small(c, &herr, w.err, 4, st)) return rr;
  if (herr) return fail(TSZP_EDEVICE, "sort: device error bits 0x%x", herr);
  return TSZP_OK;
}

// TSZP_SORTPOS=radix selects the full radix sort for sort positions by value (the
// two paths give identical ranks; the switch exists to compare them).
bool force_radix_sortpos() {
  const char* e = getenv("TSZP_SORTPOS");
  return e && strcmp(e, "radix") == 0;
}

// Sort positions by value through per-bin bitmaps (tszp_binrank.hip): writes
// out[pay] = dense in-bin rank, like sort_and_rank with by-value bins.  Sets
// *done = false, having written nothing, when the bins do not fit (more than kBinMax
// bins, or bins that are not monotone in the value); the caller then uses the radix
// path.  The few bins whose key span exceeds kBinWideBits are ranked by the radix
// sort on their buckets.
// With defer_err the radix pass's error word goes to c->host_err without a
// synchronisation: the caller checks it after its next one.
int bin_sort_positions(Ctx* c, const uint32_t* keys, const uint32_t* pay, uint32_t count, float eb, double inv,
                       int32_t* out, uint32_t out_cap, hipStream_t st, bool* done, bool defer_err = false) {
  *done = false;
  if (!count || force_radix_sortpos()) return TSZP_OK;
  BinWork w{};
  uint32_t* meta;
  WS(kWsBinMeta, bin_meta_words() * 4 + 64, meta);
  bin_work_carve(meta, &w);
  uint32_t h[16];
  ProfScope ps(kProfSort, st);
  HIPCHK(bin_range(keys, count, eb, inv, w, st));
  if (int rr = read_small(c, h, w.mm, 16, st)) return rr;
  const int64_t bmin = (int64_t)(int32_t)(h[2] ^ 0x80000000u), bmax = (int64_t)(int32_t)(~h[3] ^ 0x80000000u);
  if (bmax < bmin || bmax - bmin + 1 > (int64_t)kBinMax) return TSZP_OK;
  const uint32_t B = (uint32_t)(bmax - bmin + 1);
  uint32_t nslice, slice;
  bin_slices(count, &nslice, &slice);
  HIPCHK(bin_count(keys, count, eb, inv, (int)bmin, B, w, st));
  HIPCHK(bin_plan(B, nslice, w, st));
  if (int rr = read_small(c, h, w.info, 64, st)) return rr;
  if (h[8]) return fail(TSZP_EDEVICE, "sort positions: device error bits 0x%x", h[8]);
  if (h[6] != count) return TSZP_OK;  // some key outside [bmin, bmax]: bins not monotone
  *done = true;
  const uint32_t nwords = h[0], nchunks = h[1], wide_s = h[4], wide_e = h[5];
  if (getenv("TSZP_SORTPOS_STATS"))
    fprintf(stderr, "[sortpos] keys %u bins %u slices %u bitmap words %u chunks %u wide bins %d..%d keys %u\n",
            count, B, nslice, nwords, nchunks, h[2] == 0xffffffffu ? 0 : (int)(h[2] + bmin),
            h[2] == 0xffffffffu ? -1 : (int)(h[3] + bmin), wide_e - wide_s);
  WS(kWsBinKey, (size_t)count * 4, w.bkey);
  WS(kWsBinPay, (size_t)count * 4, w.bpay);
  uint2* gw;
  WS(kWsBinBits, (size_t)nwords * 8, gw);
  {
    EventTimer t(kProfScatter);
    t.begin(st);
    HIPCHK(bin_scatter(keys, pay, count, eb, inv, (int)bmin, B, w, st));
    t.end(st);
  }
  HIPCHK(bin_bitmap(nchunks, w, gw, st));
  HIPCHK(bin_chunk_scan(B, w, st));
  ps.stop();
  ProfScope pr(kProfRank, st);
  {
    EventTimer t(kProfQuery);
    t.begin(st);
    HIPCHK(bin_query(keys, pay, count, eb, inv, (int)bmin, B, w, gw, out, out_cap, st));
    t.end(st);
  }
  if (wide_e > wide_s) {
    const uint32_t wn = wide_e - wide_s;
    int rc;
    SortWork sw = sort_work(c, wn, false, &rc);
    if (rc) return rc;
    const size_t stw = sort_status_words(wn) + 8;
    HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(sw.st), 0u, stw * 2, st));
    HIPCHK(radix_sort_pairs(w.bkey + wide_s, nullptr, w.bpay + wide_s, wn, sw, st));
    HIPCHK(launch_rank(w.bkey + wide_s, nullptr, w.bpay + wide_s, wn, eb, inv, out, out_cap, sw, st));
    if (defer_err) {
      HIPCHK(hipMemcpyAsync(c->host_err, sw.err, 4, hipMemcpyDeviceToHost, st));
    } else {
      uint32_t herr = 0;
      HIPCHK(hipMemcpyAsync(&herr, sw.err, 4, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      if (herr) return fail(TSZP_EDEVICE, "sort: device error bits 0x%x", herr);
    }
  }
  pr.stop();
  return TSZP_OK;
}

// ---- host helpers for streams that arrive without a side-car -------------------------
inline uint64_t rd64(const unsigned char* p) {
  uint64_t v;
  memcpy(&v, p, 8);
  return v;
}

inline size_t fixed_len(size_t n, unsigned bc) {
  const unsigned rem = bc & 7u;
  return (size_t)(bc >> 3) * n + (rem ? (rem * n + 7) / 8 : 0);
}

// Walk `nb` consecutive blocks from `p` (chunks are concatenated in block order, so
// the walk needs no partition rule).  topo: blocks carry types and always a bit
// count byte; otherwise a one-element block has no bit-count byte for the reader.
// Returns false when a bit count is impossible or the walk passes `lim` bytes.
bool walk_blocks(const unsigned char* p, size_t n, int bs, bool topo, size_t lim, uint32_t* blk, size_t* total) {
  const size_t nb = (n + bs - 1) / bs;
  size_t pos = 0;
  for (size_t b = 0; b < nb; b++) {
    const size_t i = b * (size_t)bs;
    const size_t cur = (i + bs > n) ? n - i : (size_t)bs;
    if (blk) blk[b] = (uint32_t)pos;
    if (pos + 4 > lim) return false;
    pos += 4;
    if (topo || cur > 1) {
      if (pos >= lim) return false;
      const unsigned bc = p[pos++];
      if (bc > 32) return false;
      if (bc) pos += (cur - 1 + 7) / 8 + fixed_len(cur - 1, bc);
    }
    if (topo) pos += (2 * cur + 7) / 8;
    if (pos > lim) return false;
  }
  if (blk) blk[nb] = (uint32_t)pos;
  *total = pos;
  return true;
}

// Chunk count of a stream whose offset table starts at cmp: the first T in
// 1..128 whose offsets are non-decreasing and bounded and whose blocks walk within
// `lim` bytes (SZp/include/szp_detect_threads.h:161-309; the offsets of T > 1 must
// moreover fall on block boundaries).  blk / total receive the accepted walk (block
// offsets from the payload start, payload bytes), or stay empty when none walks.
unsigned detect_chunks(const unsigned char* cmp, size_t n, int bs, bool topo, size_t lim, std::vector<uint32_t>& blk,
                       size_t* total_out) {
  const size_t nb = (n + bs - 1) / bs;
  blk.assign(nb + 1, 0u);
  *total_out = 0;
  if (rd64(cmp) != 0) {
    if (!walk_blocks(cmp + 8, n, bs, topo, lim, blk.data(), total_out)) blk.clear();
    return 1;
  }
  for (unsigned t = 1; t <= 128; t++) {
    bool ok = true;
    for (unsigned k = 1; k < t && ok; k++) {
      const uint64_t a = rd64(cmp + 8 * k), b = rd64(cmp + 8 * (k - 1));
      if (a > lim || a < b) ok = false;
    }
    if (!ok) break;
    size_t total;
    if (!walk_blocks(cmp + 8 * (size_t)t, n, bs, topo, lim, blk.data(), &total)) continue;
    for (unsigned k = 1; k < t && ok; k++) {
      const uint64_t o = rd64(cmp + 8 * k);
      if (o > total) { ok = false; break; }
      // the offset must coincide with a block start or the end
      size_t lo = 0, hi = nb;
      while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (blk[mid] < o) lo = mid + 1; else hi = mid;
      }
      if (blk[lo] != o) ok = false;
    }
    if (ok) {
      *total_out = total;
      return t;
    }
  }
  if (!walk_blocks(cmp + 8, n, bs, topo, lim, blk.data(), total_out)) blk.clear();
  return 1;
}

// Upload a host stream (payload after a T-chunk table) and its block offsets.
int upload_stream(Ctx* c, const unsigned char* cmp, size_t n, int bs, bool topo, size_t lim, unsigned* T_out,
                  uint8_t** d_payload, uint32_t** d_blk, hipStream_t st) {
  // one walk: the detection's accepted walk gives the block offsets and the length
  std::vector<uint32_t> blk;
  size_t total = 0;
  const unsigned T = detect_chunks(cmp, n, bs, topo, lim, blk, &total);
  const size_t nb = (n + bs - 1) / bs;
  if (blk.empty()) return fail(TSZP_EINVAL, "corrupt stream (block walk failed)");
  uint8_t* dp;
  uint32_t* db;
  WS(kWsStream, total + 8, dp);
  WS(kWsBlk, (nb + 1) * 4, db);
  HIPCHK(hipMemcpyAsync(dp, cmp + 8 * (size_t)T, total, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(db, blk.data(), (nb + 1) * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the host vector dies at return
  *T_out = T;
  *d_payload = dp;
  *d_blk = db;
  return TSZP_OK;
}

int decode(Ctx* c, int mode, const uint8_t* payload, const uint32_t* blk, size_t n, int bs, float eb, void* out,
           uint8_t* t8, int32_t* t32, hipStream_t st) {
  uint32_t* err;
  WS(kWsMinMax, 16, err);
  HIPCHK(launch_fill_u32(err, 0u, 1, st));
  DecodeArgs a{};
  a.payload = payload;
  a.blk = blk;
  a.n = (uint32_t)n;
  a.bs = (uint32_t)bs;
  a.nb = cdiv(n, bs);
  a.eb = eb;
  a.out = out;
  a.types8 = t8;
  a.types32 = t32;
  a.err = err;
  ProfScope prof(kProfDecode, st);
  HIPCHK(launch_decode(mode, a, st));
  prof.stop();
  uint32_t herr = 0;
  if (int rr = read_small(c, &herr, err, 4, st)) return rr;
  if (herr) return fail(TSZP_EDEVICE, "decode: device error bits 0x%x", herr);
  return TSZP_OK;
}

int check_n(size_t n) {
  if (n == 0 || n >= 0x7fffffffull) return fail(TSZP_EINVAL, "element count %zu out of range [1, 2^31)", n);
  return TSZP_OK;
}

}  // namespace

// error reporting and timers for entry points defined in other source files
int tszp::set_error(int code, const char* msg) { return fail(code, "%s", msg); }
void tszp::clear_error() { g_err.clear(); }
Grid tszp::grid_of(int ndim, const uint32_t* dims) { return make_grid(ndim, dims); }
bool tszp::prof_on() { return g_prof != 0; }
void tszp::prof_add(int kind, double ms) {
  if (kind < 0 || kind >= kProfCount) return;
  g_prof_ms[kind] += ms;
  g_prof_n[kind]++;
}

// =========================================================================================
// native API
// =========================================================================================
extern "C" {

const char* tszp_last_error(void) { return g_err.c_str(); }

void tszp_prof_enable(int on) { g_prof = on; }
void tszp_prof_reset(void) {
  for (int k = 0; k < kProfCount; k++) g_prof_ms[k] = 0, g_prof_n[k] = 0;
}
int tszp_prof_read(int kind, double* ms, uint64_t* launches) {
  if (kind < 0 || kind >= kProfCount) return TSZP_EINVAL;
  *ms = g_prof_ms[kind];
  *launches = g_prof_n[kind];
  return TSZP_OK;
}
int tszp_version(void) {
  g_err.clear(); return 1; }

size_t tszp_blocks(size_t n, int bs) {
  g_err.clear(); return bs > 0 ? (n + bs - 1) / bs : 0; }
size_t tszp_topo_bound(size_t n, int bs, unsigned T) {
  g_err.clear();
  return bs > 0 ? 8 * (size_t)T + tszp_blocks(n, bs) * block_worst(bs, true) + 8 : 0;
}
size_t tszp_sortpos_bound(size_t e, int bs, unsigned T) {
  g_err.clear();
  return bs > 0 ? 8 + 8 * (size_t)T + tszp_blocks(e, bs) * block_worst(bs, false) + 8 : 0;
}
size_t tszp_plain_bound(size_t n, int bs, unsigned T) {
  g_err.clear();
  return bs > 0 ? 4 + 8 * (size_t)T + tszp_blocks(n, bs) * block_worst(bs, false) + 8 : 0;
}

int tszp_rel_to_abs(const float* d_field, size_t n, float rel, float* eb, float* range, void* stream) {
  g_err.clear();
  if (!d_field || n == 0 || !(rel > 0) || !eb) return fail(TSZP_EINVAL, "rel_to_abs: bad arguments");
  Ctx* c;
  int r = get_ctx(&c);
  if (r) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  uint32_t* mm;
  WS(kWsMinMax, 16, mm);
  HIPCHK(launch_minmax(d_field, n, mm, st));
  uint32_t h[2];
  HIPCHK(hipMemcpyAsync(h, mm, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  auto val = [](uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    memcpy(&f, &u, 4);
    return f;
  };
  const float lo = val(h[0]), hi = val(h[1]);
  float rg = hi - lo;
  if (rg <= 0) rg = FLT_EPSILON;
  *eb = rel * rg;
  if (range) *range = rg;
  return TSZP_OK;
}

int tszp_topo_compress(const float* d_field, int ndim, const uint32_t* dims, float eb, int bs, unsigned T,
                       uint8_t* d_topo, size_t topo_cap, uint32_t* d_topo_blk, uint8_t* d_sp, size_t sp_cap,
                       uint32_t* d_sp_blk, int32_t* d_sp_values, tszp_topo_info* info, void* stream) {
  g_err.clear();
  if (!d_field || !dims || !d_topo || !d_topo_blk || !info || (ndim != 2 && ndim != 3))
    return fail(TSZP_EINVAL, "topo_compress: bad arguments");
  if (!(eb > 0)) return fail(TSZP_EINVAL, "topo_compress: error bound must be positive");
  if (T < 1 || T > 4096) return fail(TSZP_EINVAL, "topo_compress: nchunks %u out of range", T);
  int r = check_bs(bs, true);
  if (r) return r;
  size_t n = 1;
  for (int k = 0; k < ndim; k++) {
    if (dims[k] == 0) return fail(TSZP_EINVAL, "topo_compress: zero dimension");
    n *= dims[k];
  }
  if ((r = check_n(n))) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  const Grid g = make_grid(ndim, dims);
  const double inv = 1.0 / (double)eb;

  uint32_t *cpk, *cpp;
  WS(kWsCpKey, n * 4, cpk);
  WS(kWsCpPay, n * 4, cpp);
  CompressResult res;
  r = run_codec(c, kModeTopoClassify, d_field, (uint32_t)n, bs, eb, inv, &g, nullptr, d_topo, topo_cap, 8 * (size_t)T,
                d_topo_blk, cpk, cpp, (uint32_t)n, st, &res);
  if (r) return r;
  const uint32_t nb = cdiv(n, bs);
  HIPCHK(launch_write_header(d_topo, 0, T, ndim == 2, nb, d_topo_blk, 0, eb, st));
  info->topo_bytes = 8 * (uint64_t)T + res.payload_bytes;
  info->cp_count = res.cp_count;
  info->ext_count = res.ext_count;
  info->sortpos_bytes = 0;

  const uint32_t E = (uint32_t)res.ext_count;
  if (E == 0) return TSZP_OK;
  if (!d_sp || !d_sp_blk) return fail(TSZP_EINVAL, "topo_compress: sort-position buffers missing");
  int32_t* sp = d_sp_values;
  if (!sp) WS(kWsSp, (size_t)E * 4, sp);
  bool by_bins;
  c->host_err[0] = 0;
  r = bin_sort_positions(c, cpk, cpp, (uint32_t)res.cp_count, eb, inv, sp, E, st, &by_bins, true);
  if (!r && !by_bins) r = sort_and_rank(c, cpk, nullptr, cpp, (uint32_t)res.cp_count, eb, inv, sp, E, st);
  if (r) return r;
  CompressResult r2;
  r = run_codec(c, kModeInt, sp, E, bs, 0.f, 0.0, nullptr, nullptr, d_sp, sp_cap, 8 + 8 * (size_t)T, d_sp_blk,
                nullptr, nullptr, 0, st, &r2);  // synchronises: the deferred sort error is in
  if (!r && c->host_err[0]) r = fail(TSZP_EDEVICE, "sort: device error bits 0x%x", c->host_err[0]);
  if (r) return r;
  // the header kernel is ordered on the stream before anything the caller does next on it
  HIPCHK(launch_write_header(d_sp, 1, T, ndim == 2, cdiv(E, bs), d_sp_blk, E, 0.f, st));
  info->sortpos_bytes = 8 + 8 * (uint64_t)T + r2.payload_bytes;
  return TSZP_OK;
}

int tszp_topo_decompress(const uint8_t* d_topo, const uint32_t* d_blk, unsigned T, size_t n, float eb, int bs,
                         float* d_out, uint8_t* d_types, void* stream) {
  g_err.clear();
  if (!d_topo || !d_blk || !d_out || T < 1) return fail(TSZP_EINVAL, "topo_decompress: bad arguments");
  int r = check_bs(bs, true);
  if (r || (r = check_n(n))) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  return decode(c, kModeTopoMap, d_topo + 8 * (size_t)T, d_blk, n, bs, eb, d_out, d_types, nullptr,
                (hipStream_t)stream);
}

int tszp_sortpos_decompress(const uint8_t* d_sp, const uint32_t* d_blk, unsigned T, size_t e, int bs,
                            int32_t* d_out, void* stream) {
  g_err.clear();
  if (!d_sp || !d_blk || !d_out || T < 1) return fail(TSZP_EINVAL, "sortpos_decompress: bad arguments");
  int r = check_bs(bs, true);
  if (r || (r = check_n(e))) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  return decode(c, kModeInt, d_sp + 8 + 8 * (size_t)T, d_blk, e, bs, 0.f, d_out, nullptr, nullptr,
                (hipStream_t)stream);
}

int tszp_plain_compress(const float* d_in, size_t n, float eb, int bs, unsigned T, uint8_t* d_out, size_t cap,
                        uint32_t* d_blk, uint64_t* out_bytes, void* stream) {
  g_err.clear();
  if (!d_in || !d_out || !d_blk || !out_bytes || T < 1 || T > 4096)
    return fail(TSZP_EINVAL, "plain_compress: bad arguments");
  if (!(eb > 0)) return fail(TSZP_EINVAL, "plain_compress: error bound must be positive");
  int r = check_bs(bs, false);
  if (r || (r = check_n(n))) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  // SZp/src/szp_float.cc:2125: the reciprocal is an int / float division, then widened
  volatile float finv = 1 / eb;
  CompressResult res;
  r = run_codec(c, kModePlain, d_in, (uint32_t)n, bs, eb, (double)finv, nullptr, nullptr, d_out, cap,
                4 + 8 * (size_t)T, d_blk, nullptr, nullptr, 0, st, &res);
  if (r) return r;
  HIPCHK(launch_write_header(d_out, 2, T, 1, cdiv(n, bs), d_blk, 0, eb, st));
  HIPCHK(hipStreamSynchronize(st));
  *out_bytes = 4 + 8 * (uint64_t)T + res.payload_bytes;
  return TSZP_OK;
}

int tszp_plain_decompress(const uint8_t* d_stream, const uint32_t* d_blk, unsigned T, size_t n, float eb, int bs,
                          float* d_out, void* stream) {
  g_err.clear();
  if (!d_stream || !d_blk || !d_out || T < 1) return fail(TSZP_EINVAL, "plain_decompress: bad arguments");
  int r = check_bs(bs, false);
  if (r || (r = check_n(n))) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  return decode(c, kModePlain, d_stream + 4 + 8 * (size_t)T, d_blk, n, bs, eb, d_out, nullptr, nullptr,
                (hipStream_t)stream);
}

// ---- batches of plain float fields -----------------------------------------------------
namespace {

struct BatchClass {
  uint32_t bs;      // a block size of the class (selects the kernel geometry)
  bool v4;          // compression by the 16-byte path
  uint32_t base;    // first tile in the batch's numbering
  uint32_t ntiles;
};

// Compression of the field by the 16-byte path (k_compress_v4): block size 16, 32, 64, 128
// or 256 and a 16-byte aligned input, as for one field; its tiles hold the same blocks.
bool batch_v4(const tszp_batch_field& f, bool compress) {
  return compress &&
         (f.block_size == 16 || f.block_size == 32 || f.block_size == 64 || f.block_size == 128 ||
          f.block_size == 256) &&
         !((uintptr_t)f.d_field & 15u) &&
         !getenv("TSZP_CODEC_GENERIC");
}

// Validates the fields, orders them by block-size class and numbers their tiles: the
// table goes to tab (count entries), order[i] is the caller's index of table entry i.
int batch_layout(const tszp_batch_field* fl, uint32_t count, bool compress, const char* what, BatchField* tab,
                 std::vector<BatchClass>* classes, std::vector<uint32_t>* order, uint32_t* total_tiles) {
  if (!fl || count == 0) return fail(TSZP_EINVAL, "%s: no fields", what);
  if (count > (1u << 20)) return fail(TSZP_EINVAL, "%s: %u fields, at most 2^20 per batch", what, count);
  std::vector<std::pair<uint64_t, uint32_t>> key(count);  // (geometry class, caller index)
  for (uint32_t i = 0; i < count; i++) {
    const tszp_batch_field& f = fl[i];
    if (!f.d_field || !f.d_stream || !f.d_blk) return fail(TSZP_EINVAL, "%s: field %u: null pointer", what, i);
    if (f.n == 0 || f.n >= 0x7fffffffull)
      return fail(TSZP_EINVAL, "%s: field %u: element count %llu out of range [1, 2^31)", what, i,
                  (unsigned long long)f.n);
    if (!(f.eb > 0) || !std::isfinite(f.eb))
      return fail(TSZP_EINVAL, "%s: field %u: error bound must be positive and finite", what, i);
    if (f.block_size < 2 || f.block_size > 1025)
      return fail(TSZP_EINVAL, "%s: field %u: blockSize %d out of range [2, 1025] (0: run tszp_batch_tune first)",
                  what, i, f.block_size);
    if ((uintptr_t)f.d_stream & 3u) return fail(TSZP_EINVAL, "%s: field %u: stream must be 4-byte aligned", what, i);
    if (compress && f.stream_cap < kBatchHdr + 4)
      return fail(TSZP_ECAP, "%s: field %u: stream capacity %llu too small", what, i,
                  (unsigned long long)f.stream_cap);
    int S, EPL;
    uint32_t bpt;
    block_geometry((uint32_t)f.block_size, &S, &EPL, &bpt);
    key[i] = {((uint64_t)batch_v4(f, compress) << 40) | ((uint64_t)S << 16) | (uint32_t)EPL, i};
  }
  std::stable_sort(key.begin(), key.end(),
                   [](const std::pair<uint64_t, uint32_t>& a, const std::pair<uint64_t, uint32_t>& b) {
                     return a.first < b.first;
                   });
  order->resize(count);
  classes->clear();
  uint64_t tiles = 0;
  for (uint32_t k = 0; k < count; k++) {
    const uint32_t i = key[k].second;
    const tszp_batch_field& f = fl[i];
    int S, EPL;
    uint32_t bpt;
    block_geometry((uint32_t)f.block_size, &S, &EPL, &bpt);
    if (k == 0 || key[k].first != key[k - 1].first)
      classes->push_back({(uint32_t)f.block_size, batch_v4(f, compress), (uint32_t)tiles, 0});
    BatchField b{};
    b.n = (uint32_t)f.n;
    b.bs = (uint32_t)f.block_size;
    b.nb = cdiv(f.n, f.block_size);
    b.ntiles = cdiv(b.nb, bpt);
    b.tile0 = (uint32_t)tiles;
    b.eb = f.eb;
    b.blk = f.d_blk;
    b.hdr = kBatchHdr;
    if (compress) {
      b.in = f.d_field;
      b.out = f.d_stream;
      const uint64_t pc = f.stream_cap - kBatchHdr - 4;  // keep 4 readable bytes past the end
      b.cap = pc > 0xffffffffull ? 0xffffffffull : pc;
      volatile float finv = 1 / f.eb;  // the plain codec's reciprocal (SZp/src/szp_float.cc:2125)
      b.inv = (double)finv;
    } else {
      b.in = f.d_stream;
      b.out = f.d_field;
    }
    tab[k] = b;
    (*order)[k] = i;
    tiles += b.ntiles;
    classes->back().ntiles += b.ntiles;
    if (tiles >= 0x7fffffffull) return fail(TSZP_EINVAL, "%s: more than 2^31 tiles in one batch", what);
  }
  *total_tiles = (uint32_t)tiles;
  return TSZP_OK;
}

// The layout in the pinned buffer, uploaded with its tile -> field map.
int batch_plan(Ctx* c, const tszp_batch_field* fl, uint32_t count, bool compress, const char* what,
               hipStream_t st, BatchField** d_tab, uint32_t** d_tmap, std::vector<BatchClass>* classes,
               std::vector<uint32_t>* order, uint32_t* total_tiles) {
  if (!fl || count == 0) return fail(TSZP_EINVAL, "%s: no fields", what);
  if (count > (1u << 20)) return fail(TSZP_EINVAL, "%s: %u fields, at most 2^20 per batch", what, count);
  void* hp;
  // room for the results too, so that reading them back never reallocates this buffer
  int r = c->host_batch_get((size_t)count * (sizeof(BatchField) + sizeof(CompressResult)), &hp);
  if (r) return r;
  BatchField* tab = static_cast<BatchField*>(hp);
  if ((r = batch_layout(fl, count, compress, what, tab, classes, order, total_tiles))) return r;
  const uint32_t tiles = *total_tiles;
  WS(kWsBatchTab, (size_t)count * sizeof(BatchField), *d_tab);
  WS(kWsBatchMap, (size_t)tiles * 4, *d_tmap);
  HIPCHK(hipMemcpyAsync(*d_tab, tab, (size_t)count * sizeof(BatchField), hipMemcpyHostToDevice, st));
  HIPCHK(launch_batch_tilemap(*d_tab, count, *d_tmap, st));
  return TSZP_OK;
}

// Results of a batch launch back to the host (the pinned table is free again after it)
int batch_results(Ctx* c, CompressResult* d_res, uint32_t count, hipStream_t st, CompressResult** out) {
  void* hp;
  int r = c->host_batch_get((size_t)count * sizeof(BatchField) + (size_t)count * sizeof(CompressResult), &hp);
  if (r) return r;
  CompressResult* h = reinterpret_cast<CompressResult*>(static_cast<char*>(hp) + (size_t)count * sizeof(BatchField));
  HIPCHK(hipMemcpyAsync(h, d_res, (size_t)count * sizeof(CompressResult), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *out = h;
  return TSZP_OK;
}

int batch_error(const char* what, uint32_t field, uint32_t bits) {
  return fail(TSZP_EDEVICE, "%s: field %u: device error bits 0x%x (%s%s%s)", what, field, bits,
              (bits & TSZP_ERR_LOOKBACK) ? "look-back timeout " : "", (bits & TSZP_ERR_STREAM) ? "corrupt stream " : "",
              (bits & TSZP_ERR_OVERFLOW) ? "capacity " : "");
}

}  // namespace

int tszp_plain_compress_batch(tszp_batch_field* fields, uint32_t count, void* stream) {
  g_err.clear();
  Ctx* c;
  int r = get_ctx(&c);
  if (r) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  BatchField* tab;
  uint32_t *tmap, tiles;
  std::vector<BatchClass> classes;
  std::vector<uint32_t> order;
  r = batch_plan(c, fields, count, true, "plain_compress_batch", st, &tab, &tmap, &classes, &order, &tiles);
  if (r) {
    (void)hipStreamSynchronize(st);  // the pinned table may still be in a copy
    return r;
  }
  uint64_t* status;
  CompressResult* res;
  // look-back words of every tile, then one tile counter per class
  WS(kWsBatchSt, (size_t)tiles * 8 + 4 * classes.size() + 8, status);
  WS(kWsBatchRes, (size_t)count * sizeof(CompressResult), res);
  uint32_t* counters = reinterpret_cast<uint32_t*>(status + tiles);
  FillRanges fr{};
  fr.add(reinterpret_cast<uint32_t*>(status), 0u, (size_t)tiles * 2 + classes.size());
  fr.add(reinterpret_cast<uint32_t*>(res), 0u, (size_t)count * sizeof(CompressResult) / 4);
  HIPCHK(launch_fill_multi(fr, st));
  ProfScope prof(kProfBatch, st);
  for (size_t k = 0; k < classes.size(); k++) {
    const BatchClass& cl = classes[k];
    HIPCHK(launch_compress_batch(kModePlain, cl.bs, cl.v4, 2, tab, tmap, cl.base, cl.ntiles, status, nullptr, res,
                                 dyn_counter(counters + k), st));
  }
  prof.stop();
  CompressResult* h;
  if ((r = batch_results(c, res, count, st, &h))) return r;
  prof.collect();
  for (uint32_t k = 0; k < count; k++) {
    const uint32_t i = order[k];
    if (h[k].error) return batch_error("plain_compress_batch", i, h[k].error);
    fields[i].stream_bytes = kBatchHdr + h[k].payload_bytes;
  }
  return TSZP_OK;
}

int tszp_plain_decompress_batch(const tszp_batch_field* fields, uint32_t count, void* stream) {
  g_err.clear();
  Ctx* c;
  int r = get_ctx(&c);
  if (r) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  BatchField* tab;
  uint32_t *tmap, tiles;
  std::vector<BatchClass> classes;
  std::vector<uint32_t> order;
  r = batch_plan(c, fields, count, false, "plain_decompress_batch", st, &tab, &tmap, &classes, &order, &tiles);
  if (r) {
    (void)hipStreamSynchronize(st);
    return r;
  }
  CompressResult* res;
  WS(kWsBatchRes, (size_t)count * sizeof(CompressResult), res);
  HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(res), 0u, (size_t)count * sizeof(CompressResult) / 4, st));
  ProfScope prof(kProfBatchDecode, st);
  for (const BatchClass& k : classes)
    HIPCHK(launch_decode_batch(kModePlain, k.bs, tab, tmap, k.base, k.ntiles, res, st));
  prof.stop();
  CompressResult* h;
  if ((r = batch_results(c, res, count, st, &h))) return r;
  prof.collect();
  for (uint32_t k = 0; k < count; k++)
    if (h[k].error) return batch_error("plain_decompress_batch", order[k], h[k].error);
  return TSZP_OK;
}

int tszp_batch_tune(tszp_batch_field* fields, uint32_t count, void* stream) {
  g_err.clear();
  if (!fields || count == 0) return fail(TSZP_EINVAL, "batch_tune: no fields");
  if (count > (1u << 20)) return fail(TSZP_EINVAL, "batch_tune: %u fields, at most 2^20 per batch", count);
  Ctx* c;
  int r = get_ctx(&c);
  if (r) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  std::vector<uint32_t> pick;  // the fields to tune: block_size 0
  for (uint32_t i = 0; i < count; i++) {
    const tszp_batch_field& f = fields[i];
    if (f.block_size != 0) continue;
    if (!f.d_field || f.n == 0 || f.n >= 0x7fffffffull || !(f.eb > 0) || !std::isfinite(f.eb))
      return fail(TSZP_EINVAL, "batch_tune: field %u: bad field, element count or error bound", i);
    pick.push_back(i);
  }
  if (pick.empty()) return TSZP_OK;
  const uint32_t m = (uint32_t)pick.size();
  void* hp;
  if ((r = c->host_batch_get((size_t)m * sizeof(BatchField) + (size_t)m * kTuneCands * 4, &hp))) return r;
  BatchField* tab = static_cast<BatchField*>(hp);
  uint32_t* hest = reinterpret_cast<uint32_t*>(tab + m);
  for (uint32_t k = 0; k < m; k++) {
    const tszp_batch_field& f = fields[pick[k]];
    BatchField b{};
    b.in = f.d_field;
    b.n = (uint32_t)f.n;
    b.eb = f.eb;
    volatile float finv = 1 / f.eb;
    b.inv = (double)finv;
    tab[k] = b;
  }
  BatchField* dtab;
  uint32_t* dest;
  WS(kWsBatchTab, (size_t)m * sizeof(BatchField), dtab);
  WS(kWsBatchRes, (size_t)m * kTuneCands * 4, dest);
  HIPCHK(hipMemcpyAsync(dtab, tab, (size_t)m * sizeof(BatchField), hipMemcpyHostToDevice, st));
  HIPCHK(launch_batch_tune(dtab, m, dest, st));
  HIPCHK(hipMemcpyAsync(hest, dest, (size_t)m * kTuneCands * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (uint32_t k = 0; k < m; k++) {
    // the smallest estimated stream; on a tie 64, then the larger block size (fewer
    // blocks to walk); 64 when the field is too short to sample (every estimate 0)
    int best = 2;
    for (int j = kTuneCands - 1; j >= 0; j--)
      if (hest[(size_t)k * kTuneCands + j] < hest[(size_t)k * kTuneCands + best]) best = j;
    fields[pick[k]].block_size = (int)kTuneSizes[best];
  }
  return TSZP_OK;
}

// ---- batch plans: the layout kept on the device, launches without host work -----------
struct tszp_batch_plan {
  int device = 0;
  uint32_t count = 0, tiles = 0;
  std::vector<BatchClass> classes;
  std::vector<uint32_t> order;
  void* mem = nullptr;  // one device allocation for everything below
  BatchField *ctab = nullptr, *dtab = nullptr;
  uint32_t* tmap = nullptr;
  uint64_t* status = nullptr;
  CompressResult *cres = nullptr, *dres = nullptr;
  CompressResult* host = nullptr;  // pinned: 2 * count results
};

int tszp_batch_plan_create(const tszp_batch_field* fields, uint32_t count, float* const* d_outs,
                           tszp_batch_plan** out, void* stream) {
  g_err.clear();
  if (!out) return fail(TSZP_EINVAL, "batch_plan_create: null output");
  if (!fields || count == 0) return fail(TSZP_EINVAL, "batch_plan_create: no fields");
  if (count > (1u << 20)) return fail(TSZP_EINVAL, "batch_plan_create: %u fields, at most 2^20", count);
  hipStream_t st = (hipStream_t)stream;
  auto* p = new tszp_batch_plan();
  std::vector<BatchField> tab(count);
  int r = batch_layout(fields, count, true, "batch_plan_create", tab.data(), &p->classes, &p->order, &p->tiles);
  if (r) {
    delete p;
    return r;
  }
  for (uint32_t i = 0; d_outs && i < count; i++)
    if (!d_outs[i]) {
      delete p;
      return fail(TSZP_EINVAL, "batch_plan_create: field %u: null output", i);
    }
  p->count = count;
  (void)hipGetDevice(&p->device);
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t tb = up((size_t)count * sizeof(BatchField)), mb = up((size_t)p->tiles * 4),
               sb = up((size_t)p->tiles * 8 + 4 * p->classes.size() + 8), rb = up((size_t)count * sizeof(CompressResult));
  hipError_t e = hipMalloc(&p->mem, 2 * tb + mb + sb + 2 * rb);
  if (e == hipSuccess) e = hipHostMalloc((void**)&p->host, 2 * (size_t)count * sizeof(CompressResult), hipHostMallocDefault);
  if (e != hipSuccess) {
    if (p->mem) (void)hipFree(p->mem);
    delete p;
    return fail(TSZP_EHIP, "batch_plan_create: %s", hipGetErrorString(e));
  }
  char* m = static_cast<char*>(p->mem);
  p->ctab = reinterpret_cast<BatchField*>(m);
  p->dtab = reinterpret_cast<BatchField*>(m + tb);
  p->tmap = reinterpret_cast<uint32_t*>(m + 2 * tb);
  p->status = reinterpret_cast<uint64_t*>(m + 2 * tb + mb);
  p->cres = reinterpret_cast<CompressResult*>(m + 2 * tb + mb + sb);
  p->dres = reinterpret_cast<CompressResult*>(m + 2 * tb + mb + sb + rb);
  std::vector<BatchField> dtab(tab);
  for (uint32_t k = 0; k < count; k++) {  // decode: stream in, field (or its own output) out
    BatchField& b = dtab[k];
    b.in = tab[k].out;
    b.out = d_outs ? static_cast<void*>(d_outs[p->order[k]]) : const_cast<void*>(tab[k].in);
  }
  auto fin = [&](int code) {
    (void)hipStreamSynchronize(st);
    (void)hipHostFree(p->host);
    (void)hipFree(p->mem);
    delete p;
    return code;
  };
  if ((e = hipMemcpyAsync(p->ctab, tab.data(), (size_t)count * sizeof(BatchField), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(p->dtab, dtab.data(), (size_t)count * sizeof(BatchField), hipMemcpyHostToDevice, st)) ||
      (e = launch_batch_tilemap(p->ctab, count, p->tmap, st)))
    return fin(fail(TSZP_EHIP, "batch_plan_create: %s", hipGetErrorString(e)));
  FillRanges fr{};
  fr.add(reinterpret_cast<uint32_t*>(p->cres), 0u, (size_t)count * sizeof(CompressResult) / 4);
  fr.add(reinterpret_cast<uint32_t*>(p->dres), 0u, (size_t)count * sizeof(CompressResult) / 4);
  if ((e = launch_fill_multi(fr, st)) || (e = hipStreamSynchronize(st)))  // the host vectors die at return
    return fin(fail(TSZP_EHIP, "batch_plan_create: %s", hipGetErrorString(e)));
  *out = p;
  return TSZP_OK;
}

int tszp_batch_plan_destroy(tszp_batch_plan* plan) {
  g_err.clear();
  if (!plan) return fail(TSZP_EINVAL, "batch_plan_destroy: null plan");
  (void)hipHostFree(plan->host);
  hipError_t e = hipFree(plan->mem);
  delete plan;
  return e == hipSuccess ? TSZP_OK : fail(TSZP_EHIP, "batch_plan_destroy: %s", hipGetErrorString(e));
}

int tszp_batch_plan_compress(tszp_batch_plan* plan, void* stream) {
  g_err.clear();
  if (!plan) return fail(TSZP_EINVAL, "batch_plan_compress: null plan");
  hipStream_t st = (hipStream_t)stream;
  FillRanges fr{};
  // look-back words, then one tile counter per class
  fr.add(reinterpret_cast<uint32_t*>(plan->status), 0u, (size_t)plan->tiles * 2 + plan->classes.size());
  fr.add(reinterpret_cast<uint32_t*>(plan->cres), 0u, (size_t)plan->count * sizeof(CompressResult) / 4);
  HIPCHK(launch_fill_multi(fr, st));  // launches only: safe under stream capture
  uint32_t* counters = reinterpret_cast<uint32_t*>(plan->status + plan->tiles);
  for (size_t k = 0; k < plan->classes.size(); k++) {
    const BatchClass& cl = plan->classes[k];
    HIPCHK(launch_compress_batch(kModePlain, cl.bs, cl.v4, 2, plan->ctab, plan->tmap, cl.base, cl.ntiles, plan->status,
                                 nullptr, plan->cres, dyn_counter(counters + k), st));
  }
  return TSZP_OK;
}

int tszp_batch_plan_decompress(tszp_batch_plan* plan, void* stream) {
  g_err.clear();
  if (!plan) return fail(TSZP_EINVAL, "batch_plan_decompress: null plan");
  hipStream_t st = (hipStream_t)stream;
  FillRanges fr{};
  fr.add(reinterpret_cast<uint32_t*>(plan->dres), 0u, (size_t)plan->count * sizeof(CompressResult) / 4);
  HIPCHK(launch_fill_multi(fr, st));
  for (const BatchClass& k : plan->classes)
    HIPCHK(launch_decode_batch(kModePlain, k.bs, plan->dtab, plan->tmap, k.base, k.ntiles, plan->dres, st));
  return TSZP_OK;
}

int tszp_batch_plan_results(tszp_batch_plan* plan, uint64_t* stream_bytes, void* stream) {
  g_err.clear();
  if (!plan) return fail(TSZP_EINVAL, "batch_plan_results: null plan");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(plan->host, plan->cres, (size_t)plan->count * sizeof(CompressResult), hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipMemcpyAsync(plan->host + plan->count, plan->dres, (size_t)plan->count * sizeof(CompressResult),
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const CompressResult* hc = plan->host;
  const CompressResult* hd = plan->host + plan->count;
  for (uint32_t k = 0; k < plan->count; k++) {
    if (hc[k].error) return batch_error("batch_plan (compression)", plan->order[k], hc[k].error);
    if (hd[k].error) return batch_error("batch_plan (decompression)", plan->order[k], hd[k].error);
    if (stream_bytes) stream_bytes[plan->order[k]] = kBatchHdr + hc[k].payload_bytes;
  }
  return TSZP_OK;
}

// ---- element-partitioned streams: threadblock codec (float, double), double RA ---------
namespace {

int check_elem(size_t n, int bs, unsigned T, int dsize, const char* what) {
  if (dsize != 4 && dsize != 8) return fail(TSZP_EINVAL, "%s: element size must be 4 or 8", what);
  if (T < 1 || T > 4096) return fail(TSZP_EINVAL, "%s: chunk count %u out of range [1, 4096]", what, T);
  if (bs < 1 || bs > 1025) return fail(TSZP_EINVAL, "%s: blockSize %d out of range [1, 1025]", what, bs);
  return check_n(n);
}

ElemArgs elem_geometry(size_t n, int bs, unsigned T, bool ra) {
  ElemArgs a{};
  a.n = (uint32_t)n;
  a.bs = (uint32_t)bs;
  a.T = T;
  a.ce = (uint32_t)(n / T);
  a.nu = elem_units(n, (uint32_t)bs, T, ra, &a.ukc);
  a.div_ukc = make_fastdiv(a.ukc ? a.ukc : 1u);
  a.ntiles = elem_tiles(a.nu, (uint32_t)bs);
  return a;
}

// the caller holds c->mu; arguments are checked
int elem_compress_impl(Ctx* c, const void* d_in, int dsize, size_t n, double eb, int bs, unsigned T, bool ra,
                       uint8_t* d_out, size_t cap, uint32_t* d_blk, uint64_t* out_bytes, hipStream_t st) {
  const size_t hdr = (size_t)dsize + 8 * (size_t)T;
  ElemArgs a = elem_geometry(n, bs, T, ra);
  uint64_t *stb, *stc;
  uint32_t* counter;
  CompressResult* res;
  int r = codec_prepare(c, a.ntiles, &stb, &stc, &counter, &res, st);
  if (r) return r;
  a.in = d_in;
  // float data: the reciprocal is an int / float division, then widened
  // (SZp/src/szp_float.cc:1489); double data: 1 / eb (SZp/src/szp_double.cc:188, :787)
  if (dsize == 4) {
    volatile float finv = 1 / (float)eb;
    a.inv = (double)finv;
  } else {
    a.inv = 1 / eb;
  }
  a.payload = d_out + hdr;
  const uint64_t pc = cap - hdr - 4;  // keep 4 readable bytes past the end
  a.cap = pc > 0xffffffffull ? 0xffffffffull : pc;
  a.blk = d_blk;
  a.st_bytes = stb;
  a.tile_counter = dyn_counter(counter);
  a.res = res;
  ProfScope prof(kProfElem, st);
  HIPCHK(launch_elem_compress(dsize, ra, a, st));
  HIPCHK(launch_elem_header(d_out, dsize, T, a.ukc, d_blk, eb, st));
  prof.stop();
  CompressResult out;
  r = read_result(c, res, st, &out);
  prof.collect();
  if (r) return r;
  *out_bytes = hdr + out.payload_bytes;
  return TSZP_OK;
}

int elem_compress(const void* d_in, int dsize, size_t n, double eb, int bs, unsigned T, bool ra, uint8_t* d_out,
                  size_t cap, uint32_t* d_blk, uint64_t* out_bytes, hipStream_t st, const char* what) {
  g_err.clear();
  if (!d_in || !d_out || !d_blk || !out_bytes) return fail(TSZP_EINVAL, "%s: bad arguments", what);
  if (!(eb > 0) || (dsize == 4 && !((float)eb > 0))) return fail(TSZP_EINVAL, "%s: error bound must be positive", what);
  int r = check_elem(n, bs, T, dsize, what);
  if (r) return r;
  const size_t hdr = (size_t)dsize + 8 * (size_t)T;
  if (((uintptr_t)(d_out + hdr)) & 3u) return fail(TSZP_EINVAL, "%s: stream must be 4-byte aligned", what);
  if (cap < hdr + 4) return fail(TSZP_ECAP, "%s: stream capacity %zu too small", what, cap);
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  return elem_compress_impl(c, d_in, dsize, n, eb, bs, T, ra, d_out, cap, d_blk, out_bytes, st);
}

// the caller holds c->mu; arguments are checked
int elem_decompress_impl(Ctx* c, const uint8_t* d_payload, const uint32_t* d_blk, unsigned T, size_t n, double eb,
                         int bs, int dsize, bool ra, void* d_out, hipStream_t st) {
  ElemDecodeArgs a{};
  a.geo = elem_geometry(n, bs, T, ra);
  uint64_t *stb, *stc;
  uint32_t* counter;
  CompressResult* res;
  int r = codec_prepare(c, a.geo.ntiles, &stb, &stc, &counter, &res, st);
  if (r) return r;
  a.payload = d_payload;
  a.blk = d_blk;
  a.eb = dsize == 4 ? (double)(float)eb : eb;
  a.out = d_out;
  a.st = stb;
  a.tile_counter = dyn_counter(counter);
  a.err = &res->error;
  ProfScope prof(kProfElemDecode, st);
  HIPCHK(launch_elem_decode(dsize, ra, a, st));
  prof.stop();
  CompressResult out;
  r = read_result(c, res, st, &out);
  prof.collect();
  return r;
}

int elem_decompress(const uint8_t* d_stream, const uint32_t* d_blk, unsigned T, size_t n, double eb, int bs,
                    int dsize, bool ra, void* d_out, hipStream_t st, const char* what) {
  g_err.clear();
  if (!d_stream || !d_blk || !d_out) return fail(TSZP_EINVAL, "%s: bad arguments", what);
  int r = check_elem(n, bs, T, dsize, what);
  if (r) return r;
  const size_t hdr = (size_t)dsize + 8 * (size_t)T;
  if (((uintptr_t)(d_stream + hdr)) & 3u) return fail(TSZP_EINVAL, "%s: stream must be 4-byte aligned", what);
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  return elem_decompress_impl(c, d_stream + hdr, d_blk, T, n, eb, bs, dsize, ra, d_out, st);
}

// ---- element-partitioned streams in host memory (the drop-in entry points) ----
// Walk the units of every chunk of a T-chunk stream (cmp at its offset table): unit
// offsets relative to the payload into blk, chunks required to be contiguous (chunk t
// starts where chunk t - 1 ended), every bit count possible, nothing past `lim`
// payload bytes.  The layout rules are those of the writers
// (SZp/src/szp_float.cc:1505-1570, szp_double.cc:200-278, :810-880).
bool walk_elem(const unsigned char* cmp, size_t n, int bs, unsigned T, bool ra, size_t lim, uint32_t* blk,
               size_t* total) {
  const unsigned char* p = cmp + 8 * (size_t)T;
  const size_t ce = n / T;
  size_t pos = 0, u = 0;
  auto run = [&](size_t m, bool bc_always) -> bool {
    if (m == 0 && !bc_always) return true;
    if (pos >= lim) return false;
    const unsigned bc = p[pos++];
    if (bc > 32) return false;
    if (bc) pos += (m + 7) / 8 + fixed_len((unsigned)m, bc);
    return pos <= lim;
  };
  for (unsigned t = 0; t < T; t++) {
    if (rd64(cmp + 8 * (size_t)t) != pos) return false;
    const size_t lo = (size_t)t * ce, hi = t == T - 1 ? n : lo + ce;
    if (lo >= hi) continue;
    if (ra) {
      for (size_t i = lo; i < hi; i += (size_t)bs) {
        const size_t cur = i + bs > hi ? hi - i : (size_t)bs;
        blk[u++] = (uint32_t)pos;
        pos += 4;
        if (pos > lim || !run(cur - 1, true)) return false;
      }
    } else {
      blk[u++] = (uint32_t)pos;
      pos += 4;
      if (pos > lim) return false;
      for (size_t i = lo + 1; i < hi; i += (size_t)bs) {
        const size_t cur = i + bs > hi ? hi - i : (size_t)bs;
        if (i > lo + 1) blk[u++] = (uint32_t)pos;
        if (!run(cur, false)) return false;
      }
    }
  }
  blk[u] = (uint32_t)pos;
  *total = pos;
  return true;
}

// Chunk count of a host stream (SZp/include/szp_detect_threads.h:196-222 lists the
// candidates): none when off[0] != 0; else, among the T in 2..128 whose offsets are
// non-decreasing and bounded, the largest whose chunks walk, each ending exactly where
// the next one starts; else 1 if one chunk walks.  (The reference takes the smallest
// such T; small chunks make that ambiguous, since an offset table of small numbers
// between zero bytes reads as a run of all-zero blocks under a smaller T.)
unsigned detect_elem_chunks(const unsigned char* cmp, size_t n, int bs, bool ra, size_t lim,
                            std::vector<uint32_t>& blk, size_t* total) {
  if (rd64(cmp) != 0) return 0;
  unsigned tmax = 1;
  for (unsigned t = 2; t <= 128; t++) {
    const uint64_t o = rd64(cmp + 8 * (size_t)(t - 1));
    if (o < rd64(cmp + 8 * (size_t)(t - 2)) || o > lim) break;
    tmax = t;
  }
  for (unsigned t = tmax; t >= 2; t--) {
    blk.assign(elem_units(n, (uint32_t)bs, t, ra, nullptr) + 1, 0u);
    if (walk_elem(cmp, n, bs, t, ra, lim, blk.data(), total)) return t;
  }
  blk.assign(elem_units(n, (uint32_t)bs, 1, ra, nullptr) + 1, 0u);
  return walk_elem(cmp, n, bs, 1, ra, lim, blk.data(), total) ? 1u : 0u;
}

// host stream (at its offset table) -> device values in the context's workspace
int elem_decode_host(Ctx* c, const unsigned char* cmp, size_t n, double eb, int bs, int dsize, bool ra, void** d_out,
                     hipStream_t st) {
  std::vector<uint32_t> blk;
  size_t total = 0;
  const size_t lim = tszp_elem_bound(n, bs, 128, dsize, ra);
  const unsigned T = detect_elem_chunks(cmp, n, bs, ra, lim, blk, &total);
  if (!T) return fail(TSZP_EINVAL, "corrupt stream (no chunk count walks)");
  const size_t hdr = 8 * (size_t)T;
  uint8_t* dp;
  uint32_t* db;
  void* dout;
  WS(kWsStream, total + 8, dp);
  WS(kWsBlk, blk.size() * 4, db);
  WS(kWsOut2, n * (size_t)dsize, dout);
  HIPCHK(hipMemcpyAsync(dp, cmp + hdr, total, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(db, blk.data(), blk.size() * 4, hipMemcpyHostToDevice, st));
  int r = elem_decompress_impl(c, dp, db, T, n, eb, bs, dsize, ra, dout, st);  // synchronises
  if (r) return r;
  *d_out = dout;
  return TSZP_OK;
}

// host values -> host stream [eb big-endian][u64 0][payload] (one chunk, as the CUDA
// writers); out must hold tszp_elem_bound(n, bs, 1, dsize, ra) bytes
int elem_compress_host(const void* data, int dsize, size_t n, double eb, int bs, bool ra, unsigned char* out,
                       size_t* outSize) {
  *outSize = 0;
  if (!data || !out) return fail(TSZP_EINVAL, "bad arguments");
  if (!(eb > 0) || (dsize == 4 && !((float)eb > 0))) return fail(TSZP_EINVAL, "error bound must be positive");
  int r = check_elem(n, bs, 1, dsize, "compress");
  if (r) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  const size_t cap = tszp_elem_bound(n, bs, 1, dsize, ra);
  void* din;
  uint8_t* dout;
  uint32_t* dblk;
  WS(kWsIn, n * (size_t)dsize, din);
  WS(kWsOut, cap, dout);
  WS(kWsBlk, (tszp_elem_units(n, bs, 1, ra) + 1) * 4, dblk);
  HIPCHK(hipMemcpyAsync(din, data, n * (size_t)dsize, hipMemcpyHostToDevice, st));
  uint64_t nbytes = 0;
  if ((r = elem_compress_impl(c, din, dsize, n, eb, bs, 1, ra, dout, cap, dblk, &nbytes, st))) return r;
  HIPCHK(hipMemcpyAsync(out, dout, nbytes, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *outSize = nbytes;
  return TSZP_OK;
}

}  // namespace

size_t tszp_elem_units(size_t n, int block_size, unsigned nchunks, int ra) {
  if (n == 0 || block_size < 1 || nchunks < 1) return 0;
  return elem_units(n, (uint32_t)block_size, nchunks, ra != 0, nullptr);
}

size_t tszp_elem_bound(size_t n, int block_size, unsigned nchunks, int elem_bytes, int ra) {
  const size_t nu = tszp_elem_units(n, block_size, nchunks, ra);
  if (!nu) return 0;
  // anchor, bc byte, signs and 32-bit magnitudes of every unit, 4 readable bytes past the end
  return (size_t)elem_bytes + 8 * (size_t)nchunks + nu * (5 + ((size_t)block_size + 7) / 8 + 4 * (size_t)block_size) +
         4;
}

int tszp_threadblock_compress(const void* d_in, int elem_bytes, size_t n, double eb, int bs, unsigned T,
                              uint8_t* d_out, size_t cap, uint32_t* d_blk, uint64_t* out_bytes, void* stream) {
  return elem_compress(d_in, elem_bytes, n, eb, bs, T, false, d_out, cap, d_blk, out_bytes, (hipStream_t)stream,
                       "threadblock_compress");
}

int tszp_threadblock_decompress(const uint8_t* d_stream, const uint32_t* d_blk, unsigned T, size_t n, double eb,
                                int bs, int elem_bytes, void* d_out, void* stream) {
  return elem_decompress(d_stream, d_blk, T, n, eb, bs, elem_bytes, false, d_out, (hipStream_t)stream,
                         "threadblock_decompress");
}

int tszp_plain_compress_f64(const double* d_in, size_t n, double eb, int bs, unsigned T, uint8_t* d_out, size_t cap,
                            uint32_t* d_blk, uint64_t* out_bytes, void* stream) {
  return elem_compress(d_in, 8, n, eb, bs, T, true, d_out, cap, d_blk, out_bytes, (hipStream_t)stream,
                       "plain_compress_f64");
}

int tszp_plain_decompress_f64(const uint8_t* d_stream, const uint32_t* d_blk, unsigned T, size_t n, double eb, int bs,
                              double* d_out, void* stream) {
  return elem_decompress(d_stream, d_blk, T, n, eb, bs, 8, true, d_out, (hipStream_t)stream, "plain_decompress_f64");
}

namespace {

int check_topo64(const uint32_t* dims, double eb, int bs, uint64_t cp_count, size_t* n_out) {
  if (!(eb > 0)) return fail(TSZP_EINVAL, "topo_compress_f64_3d: error bound must be positive");
  int r = check_bs(bs, true);
  if (r) return r;
  size_t n = 1;
  for (int k = 0; k < 3; k++) {
    if (dims[k] < 3) return fail(TSZP_EINVAL, "topo_compress_f64_3d: every dimension must be at least 3");
    n *= dims[k];
  }
  if ((r = check_n(n))) return r;
  if (cp_count > 0x7fffffffull) return fail(TSZP_EINVAL, "topo_compress_f64_3d: too many critical points");
  *n_out = n;
  return TSZP_OK;
}

// the caller holds c->mu; arguments are checked
int topo64_impl(Ctx* c, const double* d_field, const uint32_t* dims, size_t n, double eb, int bs,
                const int32_t* d_cps, uint64_t cp_count, uint8_t* d_out, size_t cap, uint32_t* d_blk,
                uint64_t* out_bytes, int32_t* d_q, uint64_t* moved, hipStream_t st) {
  int r;
  const Grid g = make_grid(3, dims);
  uint8_t* types;
  int32_t* q;
  int8_t* cd;
  uint32_t *list, *meta;
  const size_t nr = (n + 255) & ~(size_t)255;
  const uint32_t sb = saddle_blocks((uint32_t)n);
  const size_t scan_words = scan_status_words(sb + 1);
  // meta: [moved][scan err][scan counter][pad][scan status words (u64)][counts sb + 1]
  const size_t meta_words = 4 + 2 * scan_words + sb + 1;
  WS(kWsTypes, nr, types);
  WS(kWsQ, n * 4 + 16, q);
  WS(kWsMoves, 2 * nr, cd);
  WS(kWsSaddleList, n * 4 + 16, list);
  WS(kWsSaddleMeta, meta_words * 4, meta);
  int8_t* cd2 = cd + nr;
  uint32_t* cnt = meta + 4 + 2 * scan_words;
  {
    ProfScope prof(kProfSaddle, st);
    HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(types), 0u, nr / 4, st));
    HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(cd), 0u, 2 * nr / 4, st));
    HIPCHK(launch_fill_u32(meta, 0u, meta_words, st));
    HIPCHK(launch_mark_cps3(d_cps, (uint32_t)cp_count, g, types, st));
    HIPCHK(launch_quant_f64(d_field, (uint32_t)n, 1.0 / eb, q, st));
    HIPCHK(launch_saddle_list(g, types, (uint32_t)n, cnt, reinterpret_cast<uint64_t*>(meta + 4), meta + 2, meta + 1,
                              list, st));
    const uint32_t* total = cnt + sb;
    // three passes, as the reference: moves of pass k are read by pass k + 1
    int8_t* upd = cd;  // the first map is free again for the last pass
    HIPCHK(launch_saddle_pass(g, q, types, list, total, nullptr, cd, nullptr, st));
    HIPCHK(launch_saddle_pass(g, q, types, list, total, cd, cd2, nullptr, st));
    HIPCHK(launch_saddle_pass(g, q, types, list, total, cd2, nullptr, upd, st));
    HIPCHK(launch_saddle_apply(q, upd, list, total, meta, st));
    prof.stop();
    prof.collect();
  }
  if (d_q) HIPCHK(hipMemcpyAsync(d_q, q, n * 4, hipMemcpyDeviceToDevice, st));
  CompressResult res;
  r = run_codec(c, kModeTopoInt, q, (uint32_t)n, bs, (float)eb, 1.0 / eb, nullptr, types, d_out, cap, 8, d_blk,
                nullptr, nullptr, 0, st, &res);
  if (r) return r;
  HIPCHK(launch_write_header(d_out, 0, 1, 0, cdiv(n, bs), d_blk, 0, 0.f, st));
  uint32_t hm[2] = {0u, 0u};
  if (int rr = read_small(c, hm, meta, 8, st)) return rr;
  if (hm[1]) return fail(TSZP_EDEVICE, "topo_compress_f64_3d: saddle list scan failed");
  if (moved) *moved = hm[0];
  *out_bytes = 8 + res.payload_bytes;
  return TSZP_OK;
}

}  // namespace

// szp_cuda_double_compress_topology_3d (SZp/src/szp_cuda_topology_3d.cu:1065-1188)
int tszp_topo_compress_f64_3d(const double* d_field, const uint32_t* dims, double eb, int bs, const int32_t* d_cps,
                              uint64_t cp_count, uint8_t* d_out, size_t cap, uint32_t* d_blk, uint64_t* out_bytes,
                              int32_t* d_q, uint64_t* moved, void* stream) {
  g_err.clear();
  if (!d_field || !dims || !d_out || !d_blk || !out_bytes || (cp_count && !d_cps))
    return fail(TSZP_EINVAL, "topo_compress_f64_3d: bad arguments");
  size_t n;
  int r = check_topo64(dims, eb, bs, cp_count, &n);
  if (r) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  return topo64_impl(c, d_field, dims, n, eb, bs, d_cps, cp_count, d_out, cap, d_blk, out_bytes, d_q, moved,
                     (hipStream_t)stream);
}

// q[i] - q[i - 1] (q[-1] = 0) with q = C cast of x * (1 / (double)eb)
// (SZp/src/szp_cuda_compress.cu:360-377, :664-710)
int tszp_predict_quantization(const float* d_in, size_t n, float eb, int32_t* d_out, void* stream) {
  g_err.clear();
  if (!d_in || !d_out) return fail(TSZP_EINVAL, "predict_quantization: bad arguments");
  if (!(eb > 0)) return fail(TSZP_EINVAL, "predict_quantization: error bound must be positive");
  int r = check_n(n);
  if (r) return r;
  HIPCHK(launch_predict_quant(d_in, (uint32_t)n, 1.0 / (double)eb, d_out, (hipStream_t)stream));
  return TSZP_OK;
}

int tszp_classify(const float* d_field, int ndim, const uint32_t* dims, uint8_t* d_types, uint64_t* cp_count,
                  void* stream) {
  g_err.clear();
  if (!d_field || !dims || (ndim != 2 && ndim != 3)) return fail(TSZP_EINVAL, "classify: bad arguments");
  size_t n = 1;
  for (int k = 0; k < ndim; k++) n *= dims[k];
  int r = check_n(n);
  if (r) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  const uint32_t nt = classify_tiles((uint32_t)n);
  uint64_t *stb, *stc;
  uint32_t* counter;
  CompressResult* res;
  if ((r = codec_prepare(c, nt, &stb, &stc, &counter, &res, st))) return r;
  ClassifyArgs a{};
  a.f = d_field;
  a.n = (uint32_t)n;
  a.grid = make_grid(ndim, dims);
  a.eb = 1.f;
  a.inv = 1.0;
  a.ntiles = nt;
  a.st_cnt = stb;
  a.tile_counter = dyn_counter(counter);
  a.types = d_types;
  a.res = res;
  HIPCHK(launch_classify(a, st));
  CompressResult out;
  if ((r = read_result(c, res, st, &out))) return r;
  if (cp_count) *cp_count = out.cp_count;
  return TSZP_OK;
}

}  // extern "C"

namespace {

// Counts of one restoration run
struct RestoreTotals {
  uint64_t verify[4];  // missing, spurious, changed, max |y - d| bits (float)
  uint64_t moves, windows, freed;
  uint32_t rounds;
  uint32_t solve_tiles, solve_chain;  // last window solve: tiles, longest dependent chain
};

// Longest chain of dependent tiles of a one-launch window solve over one field: tiles in
// phase order, each waiting for its listed neighbours (3^ndim block) of lower phases.
// ids: the listed tiles, phase by phase.
uint32_t solve_chain_length(const std::vector<uint32_t>& ids, int ndim, uint32_t tx, uint32_t ty, uint32_t tz) {
  const size_t nt = (size_t)tx * ty * (ndim == 3 ? tz : 1);
  std::vector<uint16_t> len(nt, 0);
  uint32_t best = 0;
  for (uint32_t id : ids) {
    int tc[3] = {0, 0, 0};
    uint32_t r = id;
    if (ndim == 3) {
      tc[2] = (int)(r % tz);
      r /= tz;
    }
    tc[1] = (int)(r % ty);
    tc[0] = (int)(r / ty);
    int ph = 0;
    for (int a = 0; a < ndim; a++) ph |= (tc[a] & 1) << a;
    uint32_t m = 0;
    const int kz = ndim == 3 ? 1 : 0;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -kz; dz <= kz; dz++) {
          const int q[3] = {tc[0] + dx, tc[1] + dy, tc[2] + dz};
          if (q[0] < 0 || q[1] < 0 || q[2] < 0 || (uint32_t)q[0] >= tx || (uint32_t)q[1] >= ty ||
              (ndim == 3 && (uint32_t)q[2] >= tz))
            continue;
          int qph = 0;
          for (int a = 0; a < ndim; a++) qph |= (q[a] & 1) << a;
          if (qph >= ph) continue;
          const size_t qid = ndim == 3 ? ((size_t)q[0] * ty + q[1]) * tz + q[2] : (size_t)q[0] * ty + q[1];
          m = std::max<uint32_t>(m, len[qid]);
        }
    len[id] = (uint16_t)(m + 1);
    best = std::max(best, m + 1);
  }
  return best;
}

// The restoration's rounds over one field or a batch of fields in flat arrays (geo):
// start values, small plateaus, local-search rounds re-checked around the previous
// mismatches, exact window solves, a final verifying pass.  d: decoded values (not y),
// y: output; n_ext_one: extrema of one field (checked against the types).
int restore_run(Ctx* c, const PpGeo& geo, const float* d, const uint8_t* d_types, const int32_t* d_sp,
                uint64_t n_ext_one, bool use_sp, float* d_field, int max_rounds, RestoreTotals* out, hipStream_t st) {
  const uint32_t N = geo.n;
  const size_t n = N;
  const int ndim = geo.ndim;
  uint32_t *bad, *bad2, *act, *bits, *misc;
  uint64_t* status;
  // mismatch lists: a quarter of the points (more are listed on the next full pass)
  const uint32_t cap = N / 4 + 4096;
  WS(kWsPpBad, (size_t)cap * 4, bad);
  WS(kWsPpBad2, (size_t)cap * 4, bad2);
  const size_t bit_words = n / 32 + 4;
  WS(kWsPpBits, bit_words * 4, bits);
  WS(kWsPpMisc, kPpMiscWords * 4, misc);
  WS(kWsPpSt, pp_init_work_words(N) * 4 + 64, status);
  uint32_t h[kPpMiscWords];
  static_assert(sizeof(h) <= kHostSmall, "the counters fit the pinned read buffer");
  auto read_misc = [&]() -> int { return read_small(c, h, misc, sizeof(h), st); };
  HIPCHK(launch_fill_u32(misc, 0u, kPpMiscWords, st));
  if (use_sp) HIPCHK(pp_spmax(geo, d_sp, (uint32_t)n_ext_one, misc, st));
  HIPCHK(pp_init(geo, d, d_types, use_sp ? d_sp : nullptr, (uint32_t)n_ext_one, d_field,
                 reinterpret_cast<uint32_t*>(status), misc, st));
  // full marking pass: fresh list of mismatches and the verification counts
  // The list `bad` is exact when it holds every current mismatch; the verification
  // counts in misc are current when the last full pass saw the present field.
  bool exact = false, verified = false;
  uint32_t* part;
  WS(kWsPpPart, ((size_t)4 * pp_mark_parts() + pp_tiny_parts(cap) + 64) * 4, part);
  const bool vec_mark = pp_mark_maps(geo);
  uint32_t* badmap = nullptr;  // mismatch bitmap of the first full pass
  uint64_t* ordst = nullptr;
  if (vec_mark && !geo.fields) WS(kWsPpOrd, (size_t)pp_compact_tiles(N) * 8 + 64, ordst);
  auto full_mark = [&](uint32_t* nb_out, bool final_pass, uint32_t* bmap = nullptr) -> int {
    FillRanges fr{};
    fr.add(misc + kPpBad, 0u, 1);
    fr.add(misc + kPpVerify, 0u, 8);
    HIPCHK(launch_fill_multi(fr, st));
    // one field with a mismatch map: the list is compacted from the map in index order
    // (the plateau passes then visit neighbourhoods in the order of the field; a batch's
    // map leaves the padding between fields unwritten)
    const bool ordered = bmap && ordst;
    HIPCHK(pp_mark(geo, d_field, d_types, final_pass ? d : nullptr, ordered ? nullptr : bad, cap, part, bmap,
                   final_pass, misc, st));
    if (ordered) HIPCHK(pp_compact_map(bmap, N, ordst, bad, cap, misc, st));
    int rr = read_misc();
    if (rr) return rr;
    *nb_out = h[kPpBad];
    exact = *nb_out <= cap;
    verified = final_pass;
    return TSZP_OK;
  };
  uint32_t nb = 0;
  int r;
  WS(kWsPpBadmap, (n / 32 + 64) * 4, badmap);  // (a batch pass writes whole 1024-point chunks)
  if (vec_mark) {
    if ((r = full_mark(&nb, false, badmap))) return r;
  } else {
    if ((r = full_mark(&nb, false))) return r;
  }
  const bool stats = getenv("TSZP_RESTORE_STATS") != nullptr;
  if (stats) fprintf(stderr, "[restore] after init: %u mismatches\n", nb);
  if (!geo.fields && use_sp && h[kPpExt] != n_ext_one)
    return fail(TSZP_EINVAL, "topo_restore: %u extrema in the types, %llu sort positions given", h[kPpExt],
                (unsigned long long)n_ext_one);
  if (h[kPpErr]) return fail(TSZP_EHIP, "topo_restore: kernel error bits 0x%x", h[kPpErr]);
  uint32_t tiny = 0;
  // plateaus of two or three points: exact in one pass; what stays listed is re-checked
  if (nb && nb <= cap) {
    if (!vec_mark) {  // the mismatch map from the list
      HIPCHK(launch_fill_u32(badmap, 0u, (n + 31) / 32, st));
      HIPCHK(pp_badmap(bad, nb, badmap, st));
    }
    // a solved plateau's new values change no other point's type (its other neighbours
    // lie in other cells), so the listed points its solve left unmatched, and those of
    // the plateaus not solved, are the exact list (compacted from the sparse one)
    FillRanges fr{};
    fr.add(misc + kPpTiny, 0u, 1);
    fr.add(misc + kPpBad, 0u, 1);
    HIPCHK(launch_fill_multi(fr, st));
    HIPCHK(pp_tiny(geo, bad, nb, d, badmap, d_types, d_field, use_sp, bad2, part, misc, st));
    HIPCHK(pp_relist(bad2, nb, bad, cap, misc, st));
    if ((r = read_misc())) return r;
    if (stats) fprintf(stderr, "[restore] small plateaus solved %u: %u -> %u mismatches\n", h[kPpTiny], nb, h[kPpBad]);
    nb = h[kPpBad];
    tiny = h[kPpTiny];
    exact = nb <= cap;  // only the plateau points changed, all of them listed or fixed
    verified = false;
  }
  // whole plateaus by min-conflicts (tszp_platmc.hip): a short pass over every plateau
  // left after the small ones, then the local rounds and window solves below for what it
  // leaves, then a long pass, guided by the smoothed field, for what they leave
  uint64_t mc_solved = 0;
  const bool mc_on = !getenv("TSZP_RESTORE_NO_MC");
  auto env_u32 = [](const char* name, uint32_t dflt) {
    const char* v = getenv(name);
    return v ? (uint32_t)atoi(v) : dflt;
  };
  const float* blur = nullptr;
  // plateaus a first search (first steps) does not finish are searched again by replicas
  // with the whole budget; their staging memory bounds the entries per launch
  const size_t stage_floats = (size_t)env_u32("TSZP_MC_STAGE_MB", 64u) << 18;
  auto mc_pass = [&](uint32_t budget, uint32_t first, uint32_t reps, bool guided, const char* tag) -> int {
    if (!exact && (r = full_mark(&nb, false))) return r;
    const uint32_t nl = nb < cap ? nb : cap;
    if (guided) {  // the smoothed decoded field as the start order (a batch: the listed points' fields)
      float *s0, *s1;
      WS(kWsPpSm0, n * 4 + 16, s0);
      WS(kWsPpSm1, n * 4 + 16, s1);
      uint32_t* fflags = nullptr;
      if (geo.fields) {
        WS(kWsPpFflags, (size_t)geo.nfields * 4 + 16, fflags);
        HIPCHK(pp_field_flags(geo, bad, nl, fflags, st));
      }
      const float* src = d;
      for (int a = 0; a < ndim; a++) {
        float* dst = a % 2 ? s1 : s0;
        HIPCHK(pp_blur(geo, src, dst, a, fflags, st));
        src = dst;
      }
      blur = src;
    }
    const float* guide = guided ? blur : nullptr;
    FillRanges fr{};
    fr.add(bits, 0u, bit_words);
    fr.add(misc + kPpMcSolved, 0u, 5);
    fr.add(misc + kPpMcTicks, 0u, 5);
    fr.add(misc + kPpMcEsc, 0u, 1);
    fr.add(misc + kPpBad, 0u, 1);
    HIPCHK(launch_fill_multi(fr, st));
    // the variants by capacity, each over the same list while plateaus exceeded the last;
    // the claim bits stay, so a plateau is searched by one variant only; the searches
    // list what they leave mismatched into bad2 (exact: the list they read was)
    int rr = 0;
    uint32_t solved_small = 0, escalated = 0;
    const bool use_reps = reps > 1 && first < budget;
    uint32_t* esc = nullptr;
    if (use_reps) WS(kWsPpEsc, (size_t)nl * 4 + 16, esc);
    for (int variant = 0; variant < 3; variant++) {
      FillRanges fb{};
      fb.add(misc + kPpMcBig, 0u, 1);
      fb.add(misc + kPpMcEsc, 0u, 1);
      HIPCHK(launch_fill_multi(fb, st));
      HIPCHK(pp_platmc(geo, bad, nl, d, d_types, d_field, guide, bits, variant, use_reps ? first : budget, bad2, cap,
                       stats, misc, esc, st));
      if ((rr = read_misc())) return rr;
      if (variant == 0) solved_small = h[kPpMcSolved];
      const uint32_t big = h[kPpMcBig], nesc = use_reps ? h[kPpMcEsc] : 0u;
      if (nesc) {  // the replicas, in launches that fit the staging memory
        const size_t per = (size_t)reps * pp_platmc_capacity(variant);
        const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(nesc, stage_floats / per));
        float* stage;
        uint32_t* ctl;
        WS(kWsPpStage, (size_t)chunk * per * 4, stage);
        WS(kWsPpCtl, (size_t)chunk * 8 + 16, ctl);
        for (uint32_t c0 = 0; c0 < nesc; c0 += chunk) {
          const uint32_t cnt = std::min(chunk, nesc - c0);
          FillRanges fc{};
          fc.add(ctl, ~0u, cnt);
          fc.add(ctl + cnt, 0u, cnt);
          HIPCHK(launch_fill_multi(fc, st));
          HIPCHK(pp_platmc_replicas(geo, esc + c0, cnt, d, d_types, d_field, guide, variant, budget, reps, stage, ctl,
                                    bad2, cap, stats, misc, st));
        }
        escalated += nesc;
        if ((rr = read_misc())) return rr;
      }
      if (!big) break;
    }
    const uint32_t before = nb;
    std::swap(bad, bad2);
    nb = h[kPpBad];
    exact = nb <= cap;
    verified = false;
    if (getenv("TSZP_MC_CHECK")) {  // diagnostics: the relisted count against a full pass
      const uint32_t listed = nb;
      uint32_t sv[5];
      for (int k = 0; k < 5; k++) sv[k] = h[kPpMcSolved + k];
      if ((rr = full_mark(&nb, true))) return rr;
      fprintf(stderr, "[restore] check: relisted %u, full pass %u (missing %llu spurious %llu changed %llu)\n", listed,
              nb, (unsigned long long)reinterpret_cast<const uint64_t*>(h + kPpVerify)[0],
              (unsigned long long)reinterpret_cast<const uint64_t*>(h + kPpVerify)[1],
              (unsigned long long)reinterpret_cast<const uint64_t*>(h + kPpVerify)[2]);
      for (int k = 0; k < 5; k++) h[kPpMcSolved + k] = sv[k];
    }
    mc_solved += h[kPpMcSolved];
    if (stats)
      fprintf(stderr, "[restore] min-conflicts (%s, %u steps) plateaus solved %u (%u of up to 64 points) failed %u "
              "beyond the largest variant %u: %u -> %u mismatches; max steps %u, max us %.1f; to the replicas %u\n", tag,
              budget, h[kPpMcSolved], solved_small, h[kPpMcFailed], h[kPpMcBig], before, nb, h[kPpMcMaxSteps],
              h[kPpMcMaxTicks] * 0.01, escalated);
    if (stats && h[kPpMcSteps])
      fprintf(stderr, "[restore]   %u steps, us per step: draw + records %.2f, candidates %.2f, move %.2f, renorm %.2f\n",
              h[kPpMcSteps], h[kPpMcTicks] * 0.01 / h[kPpMcSteps], h[kPpMcTicks + 1] * 0.01 / h[kPpMcSteps],
              h[kPpMcTicks + 2] * 0.01 / h[kPpMcSteps], h[kPpMcTicks + 3] * 0.01 / h[kPpMcSteps]);
    return TSZP_OK;
  };
  if (nb && mc_on &&
      (r = mc_pass(env_u32("TSZP_MC_BUDGET1", 192u), env_u32("TSZP_MC_FIRST1", 16u), env_u32("TSZP_MC_REPS1", 8u),
                   false, "short")))
    return r;
  // window solves: anchors (bit 0) and, behind the plateau solves, only windows that hold
  // their whole plateau (bit 1)
  const int colours = ndim == 2 ? 5 : 7;
  uint64_t moves = 0, solved = 0, freed = 0;
  uint32_t rounds = 0;
  const bool stop_mc = getenv("TSZP_RESTORE_STOP_MC") != nullptr;  // diagnostics: nothing after the plateau solves
  auto window_rounds = [&](int win_mode) -> int {
  for (int outer = 0; outer < 4 && nb && !stop_mc; outer++) {
    // local search, re-checking only around the previous mismatches; a handful of
    // mismatches goes straight to the window solves (a local round costs about as much
    // as solving them; what the windows leave gets the local rounds of the next pass)
    bool complete = nb <= cap;
    const int local_rounds = outer == 0 && exact && nb <= kPpLocalFew ? 0 : max_rounds;
    for (int round = 0; round < local_rounds && nb; round++) {
      const uint32_t nl = nb < cap ? nb : cap;
      const uint32_t nb_before = nb;
      // a closed neighbourhood holds at most one point of each colour class, so every
      // colour list has room for one entry per listed mismatch
      WS(kWsPpAct, (size_t)nl * colours * 4 + 16, act);
      FillRanges fr{};
      fr.add(bits, 0u, bit_words);
      fr.add(misc + kPpMoves, 0u, 1);
      fr.add(misc + kPpActC, 0u, colours);
      HIPCHK(launch_fill_multi(fr, st));
      HIPCHK(pp_active(geo, bad, nl, d, d_types, bits, act, nl, use_sp, misc, st));
      for (int col = 0; col < colours; col++) HIPCHK(pp_move(geo, act, nl, nl, col, d, d_types, d_field, misc, st));
      rounds++;
      verified = false;
      // re-check around the listed mismatches, or the whole field when that is cheaper
      const uint64_t ring = ndim == 2 ? 13 : 25;
      if (complete && (uint64_t)nl * ring < n / 4) {
        FillRanges fr{};
        fr.add(bits, 0u, bit_words);
        fr.add(misc + kPpBad, 0u, 1);
        HIPCHK(launch_fill_multi(fr, st));
        HIPCHK(pp_remark(geo, bad, nl, d_field, d_types, bits, bad2, cap, 2, misc, st));
        if ((r = read_misc())) return r;
        std::swap(bad, bad2);
        nb = h[kPpBad];
        exact = nb <= cap;
      } else {
        if ((r = full_mark(&nb, false))) return r;
      }
      complete = nb <= cap;
      moves += h[kPpMoves];
      if (stats) fprintf(stderr, "[restore] round %u: listed %u moved %u -> %u mismatches\n", rounds, nl, h[kPpMoves], nb);
      // stop once a round removes less than a quarter of the mismatches (what is left is
      // mostly plateaus that only the window solver can order), or once few are left:
      // the one-launch solve takes about as long for a few thousand windows as for one
      // (its time is the longest chain of dependent tiles), less than another round
      if (!h[kPpMoves] || (uint64_t)nb * 4 > (uint64_t)nb_before * 3 || nb <= kPpSolveFew) break;
    }
    if (!nb) break;
    if (!exact) {
      if ((r = full_mark(&nb, false))) return r;
      if (!nb) break;
    }
    // exact window solve around what is left
    const size_t ntiles = geo.ntiles;
    const uint32_t nl = nb < cap ? nb : cap;
    const uint32_t tcap = (uint32_t)(nl < ntiles ? nl : ntiles);
    uint32_t *tbits, *tiles;
    const int phases = 1 << ndim;
    WS(kWsPpTBits, (ntiles / 32 + 4) * 4, tbits);
    WS(kWsPpTiles, (size_t)tcap * phases * 4 + 16, tiles);
    {
      FillRanges fr{};
      fr.add(bits, 0u, bit_words);
      fr.add(tbits, 0u, ntiles / 32 + 4);
      fr.add(misc + kPpSolved, 0u, 2);
      fr.add(misc + kPpFreed, 0u, 1);
      fr.add(misc + kPpTilesP, 0u, 8);
      fr.add(misc + kPpStatNodes, 0u, 12);
      fr.add(misc + kPpStatSumBuild, 0u, 3);
      HIPCHK(launch_fill_multi(fr, st));
    }
    HIPCHK(pp_tiles(geo, bad, nl, bits, tbits, tiles, tcap, misc, st));
    if ((r = read_misc())) return r;
    uint32_t total = 0;
    for (int ph = 0; ph < phases; ph++) total += h[kPpTilesP + ph] < tcap ? h[kPpTilesP + ph] : tcap;
    out->solve_tiles = total;
    if ((prof_on() || stats) && !geo.fields && total) {  // the solve's dependent-tile chain, for the profile
      std::vector<uint32_t> ids(total);
      size_t off = 0;
      for (int ph = 0; ph < phases; ph++) {
        const uint32_t cnt = h[kPpTilesP + ph] < tcap ? h[kPpTilesP + ph] : tcap;
        if (cnt) HIPCHK(hipMemcpyAsync(ids.data() + off, tiles + (size_t)ph * tcap, cnt * 4, hipMemcpyDeviceToHost, st));
        off += cnt;
      }
      HIPCHK(hipStreamSynchronize(st));
      out->solve_chain = solve_chain_length(ids, ndim, geo.tx, geo.ty, geo.tz);
      if (stats) fprintf(stderr, "[restore] solve tiles %u, longest dependent chain %u tiles\n", total, out->solve_chain);
    }
    EventTimer solve_timer(kProfSolve);
    if (getenv("TSZP_SOLVE_PHASES")) {  // one launch per parity phase
      for (int ph = 0; ph < phases; ph++) {
        const uint32_t cnt = h[kPpTilesP + ph] < tcap ? h[kPpTilesP + ph] : tcap;
        if (!cnt) continue;
        solve_timer.begin(st);
        HIPCHK(pp_solve(geo, tiles, cnt, tcap, ph, bits, d, d_types, d_field, win_mode, misc, st));
        solve_timer.end(st);
      }
    } else if (total) {  // every phase in one launch, ordered per neighbouring tiles
      uint32_t* done;
      WS(kWsPpDone, ntiles * 4 + 16, done);
      FillRanges fr{};
      fr.add(done, 0u, ntiles);
      fr.add(misc + kPpFlowCounter, 0u, 1);
      HIPCHK(launch_fill_multi(fr, st));
      solve_timer.begin(st);
      HIPCHK(pp_solve_flow(geo, tiles, total, tcap, tbits, done, bits, d, d_types, d_field, win_mode, misc, st));
      solve_timer.end(st);
    }
    if ((r = full_mark(&nb, true))) return r;
    if (stats)
      fprintf(stderr, "[restore] outer %d: listed %u tiles %u solved %u failed %u freed %u nodes %u max nodes/tile %u "
              "max windows/tile %u max tile us %.1f leaves %u verify fails %u max build us %.1f max attempt us %.1f max search us %.1f left %u\n",
              outer, nl, total, h[kPpSolved], h[kPpFailed], h[kPpFreed], h[kPpStatNodes], h[kPpStatMaxNodes],
              h[kPpStatMaxWin], h[kPpStatMaxTicks] * 0.01, h[kPpStatLeaves],
              h[kPpStatLeaves + 1], h[kPpStatLeaves + 2] * 0.01, h[kPpStatLeaves + 3] * 0.01,
              h[kPpStatLeaves + 4] * 0.01, nb);
    if (stats) {
      const double nwin = (double)(h[kPpSolved] + h[kPpFailed]);
      fprintf(stderr, "[restore]   mean per window: build us %.2f attempts us %.2f; mean tile us %.2f\n",
              nwin ? h[kPpStatSumBuild] * 0.01 / nwin : 0.0, nwin ? h[kPpStatSumAttempt] * 0.01 / nwin : 0.0,
              total ? h[kPpStatSumTile] * 0.01 / total : 0.0);
    }
    solved += h[kPpSolved];
    freed += h[kPpFreed];
    // another outer round only while many mismatches are left and the windows still
    // remove a quarter of what was listed: the rest lies on plateaus, for the
    // whole-plateau stage below
    if (nb <= kPpSolveFew || (uint64_t)nb * 4 > (uint64_t)nl * 3) break;
  }
  return TSZP_OK;
  };
  // behind the plateau solves, windows only order plateaus they hold whole (bit 1 of the
  // window mode; bit 0: anchored extrema)
  if ((r = window_rounds((use_sp ? 1 : 0) | (mc_on ? 2 : 0)))) return r;
  if (!verified && (r = full_mark(&nb, true))) return r;
  if (nb && mc_on && !stop_mc &&
      (r = mc_pass(env_u32("TSZP_MC_BUDGET2", 30000u), env_u32("TSZP_MC_FIRST2", 128u), env_u32("TSZP_MC_REPS2", 8u),
                   true, "long")))
    return r;
  // plateaus beyond the plateau solves' capacity: windows over their parts
  if (nb && mc_on && !stop_mc) {
    if ((r = window_rounds(use_sp ? 1 : 0))) return r;
    if (!verified && (r = full_mark(&nb, true))) return r;
  }
  // what is still left: whole plateaus by insertion searches and annealing
  if (nb && !stop_mc) {
    if (!exact && (r = full_mark(&nb, true))) return r;
    const uint32_t nl = nb < cap ? nb : cap;
    float *s0, *s1;
    WS(kWsPpSm0, n * 4 + 16, s0);
    WS(kWsPpSm1, n * 4 + 16, s1);
    uint32_t* fflags = nullptr;
    if (geo.fields) {  // a batch smooths only the fields of the listed points
      WS(kWsPpFflags, (size_t)geo.nfields * 4 + 16, fflags);
      HIPCHK(pp_field_flags(geo, bad, nl, fflags, st));
    }
    const float* src = d;
    for (int a = 0; a < ndim; a++) {
      float* dst = a % 2 ? s1 : s0;
      HIPCHK(pp_blur(geo, src, dst, a, fflags, st));
      src = dst;
    }
    FillRanges fr{};
    fr.add(bits, 0u, bit_words);
    fr.add(misc + kPpPlatSolved, 0u, 3);
    fr.add(misc + kPpFreed, 0u, 2);  // freed, annealed
    fr.add(misc + kPpPlatMaxTicks, 0u, 3);
    HIPCHK(launch_fill_multi(fr, st));
    HIPCHK(pp_plateau(geo, bad, nl, d, d_types, d_field, src, bits, false, misc, st));
    const uint32_t before = nb;
    if ((r = read_misc())) return r;
    if (h[kPpPlatBig]) {  // plateaus beyond the small variant: the large one over the same list
      FillRanges fb{};
      fb.add(bits, 0u, bit_words);
      HIPCHK(launch_fill_multi(fb, st));
      HIPCHK(pp_plateau(geo, bad, nl, d, d_types, d_field, src, bits, true, misc, st));
    }
    if ((r = full_mark(&nb, true))) return r;
    freed += h[kPpFreed];
    if (stats)
      fprintf(stderr, "[restore] plateaus solved %u (annealed %u) failed %u too large %u: %u -> %u mismatches; "
              "max plateau us %.1f, max anneal us %.1f, max sweeps %u\n", h[kPpPlatSolved], h[kPpPlatAnneal],
              h[kPpPlatFailed], h[kPpPlatBig], before, nb, h[kPpPlatMaxTicks] * 0.01, h[kPpPlatMaxAnneal] * 0.01,
              h[kPpPlatMaxSweeps]);
    solved += h[kPpPlatSolved];
  }
  // the verification counts come from a full pass over the final values
  if (!verified && (r = full_mark(&nb, true))) return r;
  if (h[kPpErr]) return fail(TSZP_EHIP, "topo_restore: kernel error bits 0x%x", h[kPpErr]);
  const uint64_t* hv = reinterpret_cast<const uint64_t*>(h + kPpVerify);
  for (int k = 0; k < 4; k++) out->verify[k] = hv[k];
  out->moves = moves;
  out->windows = solved + tiny + mc_solved;
  out->freed = freed;
  out->rounds = rounds;
  return TSZP_OK;
}

}  // namespace

// A workspace of its own: device buffers that grow on demand and stay until destroyed,
// for callers that restore on several streams at once (the device's own workspace
// serves one call at a time).
struct tszp_workspace {
  Ctx c;
};

int tszp::restore_batch(const RestoreBatchField* fields, uint32_t count, uint64_t flat, const float* d,
                        const uint8_t* t, const int32_t* sp, float* y, int max_rounds, tszp_restore_info* info,
                        tszp_workspace* ws, hipStream_t st) {
  if (!fields || count == 0 || !d || !t || !sp || !y || !info) return fail(TSZP_EINVAL, "restore_batch: bad arguments");
  if (flat == 0 || flat % kFlatAlign || flat >= 0x80000000ull)
    return fail(TSZP_EINVAL, "restore_batch: flat size must be a positive multiple of %u below 2^31", kFlatAlign);
  const int nd = fields[0].ndim;
  Ctx* c;
  int r = 0;
  if (ws) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev != ws->c.device)
      return fail(TSZP_EINVAL, "restore_batch: workspace of device %d used on device %d", ws->c.device, dev);
    c = &ws->c;
  } else if ((r = get_ctx(&c))) {
    return r;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  // the fields of each dimensionality form one restoration (its kernels are per dimension)
  std::vector<PpField> pf;
  std::vector<uint32_t> fmap(flat >> kFlatShift, 0u), tmap;
  for (int dim = 2; dim <= 3; dim++) {
    pf.clear();
    std::vector<uint32_t> which;
    uint32_t tiles = 0;
    for (uint32_t i = 0; i < count; i++) {
      const RestoreBatchField& f = fields[i];
      if (f.ndim != dim) continue;
      if (f.base % kFlatAlign) return fail(TSZP_EINVAL, "restore_batch: field %u: offset not aligned", i);
      PpField p{};
      p.g = make_grid(f.ndim, f.dims);
      p.base = (uint32_t)f.base;
      uint64_t n = 1;
      for (int a = 0; a < f.ndim; a++) n *= f.dims[a];
      p.n = (uint32_t)n;
      if (f.base + n > flat) return fail(TSZP_EINVAL, "restore_batch: field %u beyond the flat arrays", i);
      p.eb = f.eb;
      pp_tile_counts(p.g, &p.tx, &p.ty, &p.tz);
      p.tile0 = tiles;
      tiles += p.tx * p.ty * p.tz;
      p.n_ext = f.n_ext_dev;
      which.push_back(i);
      pf.push_back(p);
    }
    if (pf.empty()) continue;
    const uint32_t m = (uint32_t)pf.size();
    // entry m: an empty field for the chunks of the other dimensionality (every pass
    // skips points at or past base + n of their field)
    pf.push_back(PpField{});
    tmap.assign(tiles, 0u);
    std::fill(fmap.begin(), fmap.end(), m);
    for (uint32_t k = 0; k < m; k++) {
      const uint32_t c0 = pf[k].base >> kFlatShift, c1 = (uint32_t)((pf[k].base + (uint64_t)pf[k].n + kFlatAlign - 1) >> kFlatShift);
      for (uint32_t q = c0; q < c1; q++) fmap[q] = k;
      for (uint32_t q = 0; q < pf[k].tx * pf[k].ty * pf[k].tz; q++) tmap[pf[k].tile0 + q] = k;
    }
    // device tables: fields, flat map, tile map, per-field sort maxima and counts
    const size_t tb = ((m + 1) * sizeof(PpField) + 255) & ~(size_t)255, fb = (fmap.size() * 4 + 255) & ~(size_t)255,
                 mb = (tmap.size() * 4 + 255) & ~(size_t)255, sb = ((m + 1) * 4 + 255) & ~(size_t)255,
                 cb = ((m + 1) * 32 + 255) & ~(size_t)255;
    char* tabs;
    WS(kWsPpFields, tb + fb + mb + sb + cb + 256, tabs);
    // the empty field's extremum count: a zero word after the per-field counts
    uint64_t* zero_word = reinterpret_cast<uint64_t*>(tabs + tb + fb + mb + sb + cb);
    pf[m].n_ext = zero_word;
    pf[m].g = pf[0].g;
    PpGeo geo{};
    geo.ndim = dim;
    geo.n = (uint32_t)flat;
    geo.fields = reinterpret_cast<const PpField*>(tabs);
    geo.fmap = reinterpret_cast<const uint32_t*>(tabs + tb);
    geo.tmap = reinterpret_cast<const uint32_t*>(tabs + tb + fb);
    geo.spmax = reinterpret_cast<uint32_t*>(tabs + tb + fb + mb);
    geo.fcount = reinterpret_cast<unsigned long long*>(tabs + tb + fb + mb + sb);
    geo.nfields = m + 1;
    geo.ntiles = tiles;
    HIPCHK(launch_fill_u32(reinterpret_cast<uint32_t*>(zero_word), 0u, 2, st));
    HIPCHK(hipMemcpyAsync(tabs, pf.data(), (m + 1) * sizeof(PpField), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(tabs + tb, fmap.data(), fmap.size() * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(tabs + tb + fb, tmap.data(), tmap.size() * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // the host tables are reused for the next dimensionality
    RestoreTotals tot{};
    if ((r = restore_run(c, geo, d, t, sp, 0, true, y, max_rounds, &tot, st))) return r;
    std::vector<unsigned long long> fc((size_t)m * 4);
    HIPCHK(hipMemcpyAsync(fc.data(), geo.fcount, fc.size() * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (uint32_t k = 0; k < m; k++) {
      tszp_restore_info& ri = info[which[k]];
      ri.missing = fc[4 * k];
      ri.spurious = fc[4 * k + 1];
      ri.changed = fc[4 * k + 2];
      const uint32_t dv = (uint32_t)fc[4 * k + 3];
      float dev;
      memcpy(&dev, &dv, 4);
      ri.max_dev_over_eb = dev / pf[k].eb;
      ri.moves = tot.moves;
      ri.windows = tot.windows;
      ri.windows_reordered = tot.freed;
      ri.rounds = tot.rounds;
    }
  }
  (void)nd;
  return TSZP_OK;
}

extern "C" {

int tszp_workspace_create(tszp_workspace** out) {
  if (!out) return fail(TSZP_EINVAL, "workspace_create: null output");
  *out = nullptr;
  auto* w = new tszp_workspace();
  if (hipGetDevice(&w->c.device) != hipSuccess) {
    delete w;
    return fail(TSZP_EDEVICE, "workspace_create: no device");
  }
  *out = w;
  return TSZP_OK;
}

int tszp_workspace_destroy(tszp_workspace* ws) {
  if (!ws) return fail(TSZP_EINVAL, "workspace_destroy: null workspace");
  int r = TSZP_OK;
  for (int k = 0; k < kWsCount; k++)
    if (ws->c.ptr[k] && hipFree(ws->c.ptr[k]) != hipSuccess) r = fail(TSZP_EHIP, "workspace_destroy: hipFree failed");
  if (ws->c.host_small && hipHostFree(ws->c.host_small) != hipSuccess)
    r = fail(TSZP_EHIP, "workspace_destroy: hipHostFree failed");
  delete ws;
  return r;
}

int tszp_topo_restore(const float* d_decoded, float* d_field, const uint8_t* d_types, const int32_t* d_sp,
                      uint64_t n_ext, int ndim, const uint32_t* dims, float eb, int max_rounds,
                      tszp_restore_info* info, void* stream) {
  g_err.clear();
  if (!d_decoded || !d_field || !d_types || !dims || !info || (ndim != 2 && ndim != 3) || !(eb > 0) ||
      max_rounds < 0)
    return fail(TSZP_EINVAL, "topo_restore: bad arguments");
  if (reinterpret_cast<uintptr_t>(d_types) & 3u) return fail(TSZP_EINVAL, "topo_restore: types must be 4-byte aligned");
  size_t n = 1;
  for (int k = 0; k < ndim; k++) n *= dims[k];
  int r = check_n(n);
  if (r) return r;
  if (n_ext > n) return fail(TSZP_EINVAL, "topo_restore: %llu extrema for %zu points", (unsigned long long)n_ext, n);
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  PpGeo geo{};
  geo.ndim = ndim;
  geo.n = (uint32_t)n;
  geo.g = make_grid(ndim, dims);
  geo.eb = eb;
  pp_tile_counts(geo.g, &geo.tx, &geo.ty, &geo.tz);
  geo.ntiles = geo.tx * geo.ty * geo.tz;
  const float* d = d_decoded;  // the decoded values define every point's interval
  ProfScope prof(kProfRestore, st);
  if (d_decoded == d_field) {  // in place: keep a copy
    float* dcopy;
    WS(kWsPpD, n * 4, dcopy);
    HIPCHK(hipMemcpyAsync(dcopy, d_decoded, n * 4, hipMemcpyDeviceToDevice, st));
    d = dcopy;
  }
  const bool use_sp = d_sp != nullptr && n_ext > 0;
  RestoreTotals tot{};
  if ((r = restore_run(c, geo, d, d_types, d_sp, n_ext, use_sp, d_field, max_rounds, &tot, st))) return r;
  prof.stop();
  info->missing = tot.verify[0];
  info->spurious = tot.verify[1];
  info->changed = tot.verify[2];
  info->moves = tot.moves;
  info->windows = tot.windows;
  info->windows_reordered = tot.freed;
  info->rounds = tot.rounds;
  info->solve_tiles = tot.solve_tiles;
  info->solve_chain = tot.solve_chain;
  const uint32_t dv = (uint32_t)tot.verify[3];
  float dev;
  memcpy(&dev, &dv, 4);
  info->max_dev_over_eb = dev / eb;
  return TSZP_OK;
}

int tszp_topo_verify(const float* d_field, const uint8_t* d_types, int ndim, const uint32_t* dims, uint64_t* counts,
                     void* stream) {
  g_err.clear();
  if (!d_field || !d_types || !dims || !counts || (ndim != 2 && ndim != 3))
    return fail(TSZP_EINVAL, "topo_verify: bad arguments");
  size_t n = 1;
  for (int k = 0; k < ndim; k++) n *= dims[k];
  int r = check_n(n);
  if (r) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  uint32_t* misc;
  WS(kWsPpMisc, kPpMiscWords * 4, misc);
  HIPCHK(launch_fill_u32(misc, 0u, kPpMiscWords, st));
  uint32_t* part;
  WS(kWsPpPart, ((size_t)4 * pp_mark_parts() + 64) * 4, part);
  PpGeo geo{};
  geo.ndim = ndim;
  geo.n = (uint32_t)n;
  geo.g = make_grid(ndim, dims);
  HIPCHK(pp_mark(geo, d_field, d_types, nullptr, nullptr, 0, part, nullptr, false, misc, st));
  uint64_t hv[4];
  HIPCHK(hipMemcpyAsync(hv, misc + kPpVerify, 32, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int k = 0; k < 3; k++) counts[k] = hv[k];
  return TSZP_OK;
}

int tszp_quality_metrics(const float* d_x, const float* d_y, int ndim, const uint32_t* dims, float eb,
                         tszp_quality* q, void* stream) {
  g_err.clear();
  static_assert(sizeof(tszp_quality) == (kQualCounts + 4) * 8, "tszp_quality layout");
  if (!d_x || !d_y || !dims || !q || (ndim != 2 && ndim != 3) || !(eb > 0))
    return fail(TSZP_EINVAL, "quality_metrics: bad arguments");
  size_t n = 1;
  for (int k = 0; k < ndim; k++) n *= dims[k];
  int r = check_n(n);
  if (r) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  const uint32_t nblocks = 1024;
  uint64_t* part;
  WS(kWsQuality, ((size_t)nblocks + 1) * kQualWords * 8, part);
  uint64_t* out = part + (size_t)nblocks * kQualWords;
  HIPCHK(launch_quality(make_grid(ndim, dims), d_x, d_y, (uint32_t)n, eb, part, nblocks, out, st));
  HIPCHK(hipMemcpyAsync(q, out, sizeof(tszp_quality), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return TSZP_OK;
}

// Block offsets of a device-resident stream without a side-car.  kind: 0 topology
// ([u64 off[T]][blocks]), 1 sort positions ([u64 E][u64 off[T]][blocks]), 2 plain
// ([f32 eb][u64 off[T]][blocks]).  The chunk count is chosen by the rule of
// detect_chunks (first T whose offsets are non-decreasing and whose blocks walk), with
// the walk on the device; the host only reads the head of the stream to list the
// candidates.  d_blk receives tszp_blocks(n, bs) + 1 offsets relative to the payload.
int tszp_parse_blocks(const uint8_t* d_stream, size_t len, size_t n, int bs, int kind, uint32_t* d_blk,
                      unsigned* T_out, void* stream) {
  g_err.clear();
  if (!d_stream || !d_blk || !T_out || kind < 0 || kind > 2 || bs < 1 || n == 0)
    return fail(TSZP_EINVAL, "parse_blocks: bad arguments");
  int r = check_n(n);
  if (r) return r;
  const bool topo = kind == 0;
  const size_t nb = (n + bs - 1) / bs;
  const size_t skip = kind == 1 ? 8 : kind == 2 ? 4 : 0;  // count / eb before the chunk table
  if (len <= skip + 8) return fail(TSZP_EINVAL, "parse_blocks: stream too short");
  const uint8_t* d_tab = d_stream + skip;
  const size_t avail = len - skip;
  // sizes of full blocks by bit count
  ParseArgs a{};
  a.topo = topo;
  a.has_bc = topo || bs > 1;
  a.bc_off = 4;  // SZp blocks: anchor, bit count, cur - 1 deltas
  a.anchor = 1;
  a.dm1 = 1;
  a.bc_always = topo;
  uint32_t maxb = 0, minb = 0xffffffffu;
  for (unsigned bc = 0; bc <= 32; bc++) {
    const size_t m = (size_t)bs - 1;
    size_t z = 4 + (a.has_bc ? 1 : 0) + (bc ? (m + 7) / 8 + fixed_len(m, bc) : 0) + (topo ? (2 * (size_t)bs + 7) / 8 : 0);
    if (z > kParseMaxBlock) return fail(TSZP_EINVAL, "parse_blocks: block size %d too large for the device parse", bs);
    a.sz[bc] = (uint16_t)z;
    maxb = z > maxb ? (uint32_t)z : maxb;
    minb = z < minb ? (uint32_t)z : minb;
  }
  a.sz[33] = (uint16_t)minb;  // an impossible bit count: keep walking, the final walk rejects it
  a.maxb = maxb;
  a.rounds = 1;
  while (((uint64_t)1 << a.rounds) < kParseSeg / minb + 2) a.rounds++;
  // candidates: the head of the chunk table, and enough of the payload for a first
  // look at the blocks of each candidate
  const size_t hb = std::min<size_t>(avail, 8 * 129 + 32 * (size_t)kParseMaxBlock);
  std::vector<unsigned char> head(hb);
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(head.data(), d_tab, hb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const uint64_t lim = avail;
  std::vector<unsigned> cand;
  if (rd64(head.data()) != 0) {
    cand.push_back(1);
  } else {
    for (unsigned t = 1; t <= 128 && 8 * (size_t)t + 8 <= hb; t++) {
      bool ok = true;
      for (unsigned k = 1; k < t && ok; k++) {
        const uint64_t x = rd64(head.data() + 8 * k), y = rd64(head.data() + 8 * (k - 1));
        if (x > lim || x < y) ok = false;
      }
      if (!ok) break;
      cand.push_back(t);
    }
    if (cand.empty()) cand.push_back(1);
  }
  uint32_t* err;
  WS(kWsMinMax, 16, err);
  for (unsigned t : cand) {
    if (avail <= 8 * (size_t)t) break;
    {  // the first blocks must have possible bit counts (cheap rejection of wrong T)
      size_t pos = 8 * (size_t)t;
      bool bad = false;
      for (size_t b = 0; b < std::min<size_t>(nb, 32) && !bad; b++) {
        const size_t cur = std::min<size_t>((size_t)bs, n - b * (size_t)bs);
        if (!(topo || cur > 1)) break;
        if (pos + 4 >= hb) break;
        const unsigned bc = head[pos + 4];
        if (bc > 32) bad = true;
        pos += 5 + (bc ? (cur - 1 + 7) / 8 + fixed_len(cur - 1, bc) : 0) + (topo ? (2 * cur + 7) / 8 : 0);
      }
      if (bad) continue;
    }
    a.p = d_tab + 8 * (size_t)t;
    a.len = avail - 8 * (size_t)t;
    if (a.len > 0xffffffffull) return fail(TSZP_EINVAL, "parse_blocks: payload beyond 4 GiB");
    void* work;
    WS(kWsStream2, parse_work_bytes(a.len, a.maxb), work);
    HIPCHK(launch_fill_u32(err, 0u, 2, st));
    HIPCHK(parse_blocks(a, n, (uint32_t)bs, (uint32_t)nb, d_blk, work, err, st));
    HIPCHK(parse_check_offsets(d_tab, t, d_blk, (uint32_t)nb, err, st));
    uint32_t h[2];
    HIPCHK(hipMemcpyAsync(h, err, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!h[0] && h[1]) {
      *T_out = t;
      return TSZP_OK;
    }
  }
  return fail(TSZP_EINVAL, "parse_blocks: no chunk count gives a consistent block sequence (corrupt stream?)");
}

// Unit offsets of an element-partitioned stream in device memory that arrives without
// its side-car (the host walk of the drop-in decoders, walk_elem, on the device): the
// chunk count is the host rule's (detect_elem_chunks: the largest T in 2..128 whose
// offsets are non-decreasing and whose chunks walk, each ending where the next one
// starts; else 1), and every chunk is cut into units by the segment-table parse of
// tszp_parse.hip with the units' layout (threadblock: a bit-count byte and up to bs
// deltas after the chunk's anchor; double random access: anchor, bit-count byte and
// cur - 1 deltas per unit).
int tszp_elem_parse(const uint8_t* d_full, size_t full_len, size_t n, int bs, int elem_bytes, int ra,
                    uint32_t* d_blk, unsigned* T_out, void* stream) {
  g_err.clear();
  if (!d_full || !d_blk || !T_out || n == 0 || bs < 1 || bs > 1025 || (elem_bytes != 4 && elem_bytes != 8) ||
      (ra && elem_bytes != 8))
    return fail(TSZP_EINVAL, "elem_parse: bad arguments");
  int r = check_n(n);
  if (r) return r;
  if (full_len <= (size_t)elem_bytes + 8) return fail(TSZP_EINVAL, "elem_parse: stream too short");
  const uint8_t* d_stream = d_full + elem_bytes;  // the chunk table, after the error bound
  const size_t len = full_len - elem_bytes;
  // unit layout and the sizes of full units by bit count
  ParseArgs a{};
  a.topo = 0;
  a.has_bc = 1;
  a.bc_off = ra ? 4u : 0u;
  a.anchor = ra ? 1 : 0;
  a.dm1 = ra ? 1 : 0;
  a.bc_always = ra ? 1 : 0;
  const size_t m = ra ? (size_t)bs - 1 : (size_t)bs;  // deltas of a full unit
  uint32_t maxb = 0, minb = 0xffffffffu;
  for (unsigned bc = 0; bc <= 32; bc++) {
    const size_t z = (ra ? 4 : 0) + 1 + (bc ? (m + 7) / 8 + fixed_len((unsigned)m, bc) : 0);
    if (z > kParseMaxBlock) return fail(TSZP_EINVAL, "elem_parse: block size %d too large for the device parse", bs);
    a.sz[bc] = (uint16_t)z;
    maxb = z > maxb ? (uint32_t)z : maxb;
    minb = z < minb ? (uint32_t)z : minb;
  }
  a.sz[33] = (uint16_t)minb;
  a.maxb = maxb;
  a.rounds = 1;
  while (((uint64_t)1 << a.rounds) < kParseSeg / minb + 2) a.rounds++;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  // the chunk table's head: the candidates
  const size_t hb = std::min<size_t>(len, 8 * 129);
  std::vector<unsigned char> head(hb);
  HIPCHK(hipMemcpyAsync(head.data(), d_stream, hb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (rd64(head.data()) != 0) return fail(TSZP_EINVAL, "elem_parse: corrupt stream (first chunk offset not 0)");
  std::vector<unsigned> cand;
  {
    unsigned tmax = 1;
    for (unsigned t = 2; t <= 128 && 8 * (size_t)t <= hb; t++) {
      const uint64_t o = rd64(head.data() + 8 * (size_t)(t - 1));
      if (o < rd64(head.data() + 8 * (size_t)(t - 2)) || 8 * (size_t)t + o > len) break;
      tmax = t;
    }
    for (unsigned t = tmax; t >= 2; t--) cand.push_back(t);
    cand.push_back(1);
  }
  uint32_t *err, *ends;
  WS(kWsMinMax, 16, err);
  WS(kWsElemEnds, 128 * 4 + 16, ends);
  void* work;
  WS(kWsStream2, parse_work_bytes(len, a.maxb), work);
  for (unsigned T : cand) {
    if (len <= 8 * (size_t)T) continue;
    const uint8_t* payload = d_stream + 8 * (size_t)T;
    const uint64_t avail = len - 8 * (size_t)T;
    if (avail > 0xffffffffull) return fail(TSZP_EINVAL, "elem_parse: payload beyond 4 GiB");
    uint32_t ukc = 0;
    elem_units(n, (uint32_t)bs, T, ra != 0, &ukc);
    const uint64_t ce = n / T;
    HIPCHK(launch_fill_u32(err, 0u, 2, st));
    HIPCHK(launch_fill_u32(ends, 0xffffffffu, T, st));
    bool bad = false;
    for (unsigned t = 0; t < T && !bad; t++) {
      const uint64_t lo = (uint64_t)t * ce, hi = t == T - 1 ? n : lo + ce, cl = hi - lo;
      const uint64_t o = rd64(head.data() + 8 * (size_t)t);
      const uint64_t e = t + 1 < T ? rd64(head.data() + 8 * (size_t)(t + 1)) : avail;
      if (o > avail || e < o) {
        bad = true;
        break;
      }
      const uint64_t skip = ra ? 0 : 4;                 // the threadblock chunk's anchor
      const uint64_t nd = ra ? cl : (cl >= 2 ? cl - 1 : 0);  // elements (RA) or deltas cut into units
      if (!nd) continue;                                // (chunk starts and ends: parse_elem_finish)
      if (o + skip >= e) {
        bad = true;
        break;
      }
      ParseArgs ac = a;
      ac.p = payload + o + skip;
      ac.len = e - o - skip;
      ac.out_base = (uint32_t)(o + skip);
      ac.end_out = ends + t;
      const uint32_t nu = (uint32_t)((nd + bs - 1) / bs);
      HIPCHK(parse_blocks(ac, nd, (uint32_t)bs, nu, d_blk + (size_t)t * ukc, work, err, st));
    }
    if (bad) continue;
    HIPCHK(parse_elem_finish(d_stream, T, n, ukc, ra != 0, avail, d_blk, ends, err, st));
    uint32_t h[2];
    HIPCHK(hipMemcpyAsync(h, err, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!h[0]) {
      *T_out = T;
      return TSZP_OK;
    }
  }
  return fail(TSZP_EINVAL, "elem_parse: no chunk count gives a consistent unit sequence (corrupt stream?)");
}

// .SZp container of the reference's CPU drivers: [topology stream][sort-position
// stream][u64 topology bytes][u64 sort-position bytes]
// (example/testfloat_compress_fastmode1_topology.cc:105-139).  Pointers may be host or
// device memory (unified addressing); the call returns when the copies are done.
int tszp_szp_pack(const uint8_t* topo, size_t topo_bytes, const uint8_t* sp, size_t sp_bytes, uint8_t* out,
                  size_t cap, size_t* out_bytes, void* stream) {
  g_err.clear();
  if (!topo || !out || !out_bytes || (sp_bytes && !sp)) return fail(TSZP_EINVAL, "szp_pack: bad arguments");
  const size_t total = topo_bytes + sp_bytes + 2 * sizeof(uint64_t);
  if (total < topo_bytes) return fail(TSZP_EINVAL, "szp_pack: size overflow");
  if (cap < total) return fail(TSZP_ECAP, "szp_pack: %zu bytes needed, capacity %zu", total, cap);
  hipStream_t st = (hipStream_t)stream;
  const uint64_t footer[2] = {(uint64_t)topo_bytes, (uint64_t)sp_bytes};
  HIPCHK(hipMemcpyAsync(out, topo, topo_bytes, hipMemcpyDefault, st));
  if (sp_bytes) HIPCHK(hipMemcpyAsync(out + topo_bytes, sp, sp_bytes, hipMemcpyDefault, st));
  HIPCHK(hipMemcpyAsync(out + topo_bytes + sp_bytes, footer, sizeof(footer), hipMemcpyDefault, st));
  HIPCHK(hipStreamSynchronize(st));  // the footer source lives on this stack
  *out_bytes = total;
  return TSZP_OK;
}

// Reads the footer as the reference's decompress driver does
// (example/testfloat_decompress_fastmode1_topology.cc:722-746): sizes that do not fit
// the buffer mean an old single-stream file, all topology.  *n_extrema is the u64 count
// at the head of the sort-position stream (0 without one).
int tszp_szp_unpack(const uint8_t* buf, size_t len, size_t* topo_bytes, size_t* sp_offset, size_t* sp_bytes,
                    uint64_t* n_extrema) {
  g_err.clear();
  if (!buf || !topo_bytes || !sp_offset || !sp_bytes || !n_extrema) return fail(TSZP_EINVAL, "szp_unpack: bad arguments");
  uint64_t f[2] = {0, 0};
  size_t tb = len, sb = 0;
  if (len >= sizeof(f)) {
    HIPCHK(hipMemcpy(f, buf + len - sizeof(f), sizeof(f), hipMemcpyDefault));
    // the reference's test is f[0] + f[1] + 16 <= len + 16 with f[0] in (0, len); written
    // as f[1] <= len - f[0] it cannot wrap around, so sp_offset + sp_bytes <= len
    if (f[0] > 0 && f[0] < len && f[1] <= (uint64_t)len - f[0]) {
      tb = (size_t)f[0];
      sb = (size_t)f[1];
    }
  }
  uint64_t e = 0;
  if (sb >= sizeof(uint64_t) && tb + sizeof(uint64_t) <= len) HIPCHK(hipMemcpy(&e, buf + tb, sizeof(e), hipMemcpyDefault));
  *topo_bytes = tb;
  *sp_offset = tb;
  *sp_bytes = sb;
  *n_extrema = e;
  return TSZP_OK;
}

int tszp_gen_blobs(float* d, int ndim, const uint32_t* dims, uint64_t seed, int nblobs, float smin, float smax,
                   float noise, int waves, void* stream) {
  g_err.clear();
  if (!d || !dims || (ndim != 2 && ndim != 3) || nblobs < 0) return fail(TSZP_EINVAL, "gen_blobs: bad arguments");
  HIPCHK(launch_gen_blobs(d, ndim, dims, seed, nblobs, smin, smax, noise, waves, (hipStream_t)stream));
  return TSZP_OK;
}

int tszp_gen_fbm(float* d, const uint32_t* dims, uint64_t seed, int octaves, float noise, void* stream) {
  g_err.clear();
  if (!d || !dims) return fail(TSZP_EINVAL, "gen_fbm: bad arguments");
  HIPCHK(launch_gen_fbm(d, dims, seed, octaves, noise, (hipStream_t)stream));
  return TSZP_OK;
}

}  // extern "C"

// =========================================================================================
// drop-in host-pointer API (reference names; NULL / zero size on error, message in
// tszp_last_error() -- the reference exits the process instead)
// =========================================================================================
namespace {

template <typename Rec>
int find_cps_impl(const float* data, int ndim, const uint32_t* dims, float eb, Rec** out, size_t* count) {
  *out = nullptr;
  *count = 0;
  size_t n = 1;
  for (int k = 0; k < ndim; k++) n *= dims[k];
  int r = check_n(n);
  if (r) return r;
  Ctx* c;
  if ((r = get_ctx(&c))) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  float* df;
  WS(kWsIn, n * 4, df);
  HIPCHK(hipMemcpyAsync(df, data, n * 4, hipMemcpyHostToDevice, st));
  size_t interior = 1;
  for (int k = 0; k < ndim; k++) interior *= (dims[k] > 2 ? dims[k] - 2 : 0);
  const int rec = (int)(sizeof(Rec) / 4);
  int32_t* dc;
  WS(kWsCps, interior * sizeof(Rec) + 4, dc);
  const uint32_t nt = classify_tiles((uint32_t)n);
  uint64_t *stb, *stc;
  uint32_t* counter;
  CompressResult* res;
  if ((r = codec_prepare(c, nt, &stb, &stc, &counter, &res, st))) return r;
  ClassifyArgs a{};
  a.f = df;
  a.n = (uint32_t)n;
  a.grid = make_grid(ndim, dims);
  a.eb = eb;
  a.inv = 1.0 / (double)eb;
  a.ntiles = nt;
  a.st_cnt = stb;
  a.tile_counter = dyn_counter(counter);
  a.cps = dc;
  a.rec = rec;
  a.cp_cap = (uint32_t)interior;
  a.res = res;
  HIPCHK(launch_classify(a, st));
  CompressResult o;
  if ((r = read_result(c, res, st, &o))) return r;
  if (o.cp_count == 0) return TSZP_OK;
  Rec* h = (Rec*)malloc(o.cp_count * sizeof(Rec));
  if (!h) return fail(TSZP_EINVAL, "malloc failed");
  hipError_t e = hipMemcpy(h, dc, o.cp_count * sizeof(Rec), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    free(h);
    return fail(TSZP_EHIP, "hipMemcpy: %s", hipGetErrorString(e));
  }
  *out = h;
  *count = o.cp_count;
  return TSZP_OK;
}

// sort positions of a caller's record array, grouped by the stored bins
template <typename Rec>
int sort_cps_impl(Rec* cps, size_t count, const float* data, uint32_t row, uint64_t plane) {
  if (count == 0 || count >= 0x7fffffffull) return count ? fail(TSZP_EINVAL, "too many points") : TSZP_OK;
  const int rec = (int)(sizeof(Rec) / 4);
  // the largest flat index tells how much of the field has to travel
  // (SZp/src/szp_cuda_topology.cu:202-206)
  uint64_t max_flat = 0;
  for (size_t i = 0; i < count; i++) {
    const int32_t* p = reinterpret_cast<const int32_t*>(&cps[i]);
    const int64_t f = rec == 5 ? (int64_t)p[0] * row + p[1] : (int64_t)p[0] * (int64_t)plane + (int64_t)p[1] * row + p[2];
    if (f < 0) return fail(TSZP_EINVAL, "negative coordinate in critical point %zu", i);
    if ((uint64_t)f > max_flat) max_flat = (uint64_t)f;
  }
  Ctx* c;
  int r = get_ctx(&c);
  if (r) return r;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  float* df;
  int32_t* dc;
  uint32_t *k, *hi, *pay, *err;
  int32_t* rank;
  WS(kWsIn, (max_flat + 1) * 4, df);
  WS(kWsCps, count * sizeof(Rec), dc);
  WS(kWsCpKey, count * 4, k);
  WS(kWsHiKey, count * 4, hi);
  WS(kWsCpPay, count * 4, pay);
  WS(kWsSp, count * 4, rank);
  WS(kWsMinMax, 16, err);
  HIPCHK(hipMemcpyAsync(df, data, (max_flat + 1) * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dc, cps, count * sizeof(Rec), hipMemcpyHostToDevice, st));
  HIPCHK(launch_fill_u32(err, 0u, 1, st));
  HIPCHK(launch_cp_keys(dc, (uint32_t)count, rec, df, max_flat + 1, row, plane, k, hi, pay, err, st));
  if ((r = sort_and_rank(c, k, hi, pay, (uint32_t)count, 0.f, 0.0, rank, (uint32_t)count, st))) return r;
  std::vector<int32_t> h(count);
  HIPCHK(hipMemcpyAsync(h.data(), rank, count * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (size_t i = 0; i < count; i++) cps[i].sort_position = h[i];
  return TSZP_OK;
}

// int codec of the extrema sort positions in record order -> [u64 0][blocks]
template <typename Rec>
unsigned char* compress_sp_impl(const Rec* cps, size_t count, size_t* outSize, int bs) {
  *outSize = 0;
  if (!cps || count == 0 || check_bs(bs, true)) return nullptr;
  std::vector<int32_t> sp;
  sp.reserve(count);
  for (size_t i = 0; i < count; i++)
    if (cps[i].type == 1 || cps[i].type == 2) sp.push_back(cps[i].sort_position);
  if (sp.empty() || check_n(sp.size())) return nullptr;
  Ctx* c;
  if (get_ctx(&c)) return nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  const size_t e = sp.size();
  int32_t* din;
  uint8_t* dout;
  uint32_t* dblk;
  const size_t cap = tszp_sortpos_bound(e, bs, 1);
  if (c->get(kWsSp, e * 4, (void**)&din) || c->get(kWsOut, cap, (void**)&dout) ||
      c->get(kWsBlk, (tszp_blocks(e, bs) + 1) * 4, (void**)&dblk))
    return nullptr;
  if (hipMemcpyAsync(din, sp.data(), e * 4, hipMemcpyHostToDevice, st) != hipSuccess) return nullptr;
  CompressResult res;
  // header is the single zero chunk offset; the payload follows at byte 8
  if (run_codec(c, kModeInt, din, (uint32_t)e, bs, 0.f, 0.0, nullptr, nullptr, dout, cap, 8, dblk, nullptr,
                nullptr, 0, st, &res))
    return nullptr;
  const size_t total = 8 + res.payload_bytes;
  unsigned char* h = (unsigned char*)malloc(total);
  if (!h) return nullptr;
  memset(h, 0, 8);
  if (hipMemcpy(h + 8, dout + 8, res.payload_bytes, hipMemcpyDeviceToHost) != hipSuccess) {
    free(h);
    return nullptr;
  }
  *outSize = total;
  return h;
}

template <typename Rec>
unsigned char* compress_topo_impl(const float* data, size_t* outSize, float eb, size_t n, int bs, const Rec* cps,
                                  int count, int ndim, const uint32_t* dims) {
  *outSize = 0;
  if (!data || n == 0 || !(eb > 0) || check_bs(bs, true) || check_n(n)) {
    if (!(eb > 0)) fail(TSZP_EINVAL, "error bound must be positive");
    return nullptr;
  }
  // type map from the caller's list (SZp/src/szp_float.cc:1040-1064)
  std::vector<uint8_t> types(n, 0);
  for (int i = 0; cps && i < count; i++) {
    const int32_t* p = reinterpret_cast<const int32_t*>(&cps[i]);
    const int t = ndim == 2 ? p[2] : p[3];
    int64_t f;
    if (ndim == 2) {
      if (p[0] < 0 || p[0] >= (int)dims[0] || p[1] < 0 || p[1] >= (int)dims[1]) continue;
      f = (int64_t)p[0] * dims[1] + p[1];
    } else {
      if (p[0] < 0 || p[0] >= (int)dims[0] || p[1] < 0 || p[1] >= (int)dims[1] || p[2] < 0 ||
          p[2] >= (int)dims[2])
        continue;
      f = ((int64_t)p[0] * dims[1] + p[1]) * dims[2] + p[2];
    }
    if (t >= 1 && t <= 3 && (size_t)f < n) types[f] = (uint8_t)t;
  }
  Ctx* c;
  if (get_ctx(&c)) return nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  float* din;
  uint8_t *dt, *dout;
  uint32_t* dblk;
  const size_t cap = tszp_topo_bound(n, bs, 1);
  if (c->get(kWsIn, n * 4, (void**)&din) || c->get(kWsTypes, n, (void**)&dt) || c->get(kWsOut, cap, (void**)&dout) ||
      c->get(kWsBlk, (tszp_blocks(n, bs) + 1) * 4, (void**)&dblk))
    return nullptr;
  if (hipMemcpyAsync(din, data, n * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dt, types.data(), n, hipMemcpyHostToDevice, st) != hipSuccess)
    return nullptr;
  CompressResult res;
  if (run_codec(c, kModeTopoMap, din, (uint32_t)n, bs, eb, 1.0 / (double)eb, nullptr, dt, dout, cap, 8, dblk,
                nullptr, nullptr, 0, st, &res))
    return nullptr;
  const size_t total = 8 + res.payload_bytes;
  unsigned char* h = (unsigned char*)malloc(total);
  if (!h) return nullptr;
  memset(h, 0, 8);
  if (hipMemcpy(h + 8, dout + 8, res.payload_bytes, hipMemcpyDeviceToHost) != hipSuccess) {
    free(h);
    return nullptr;
  }
  *outSize = total;
  return h;
}

}  // namespace

extern "C" {

CriticalPoint* szp_cuda_find_critical_points(float* data, size_t* outCount, int rows, int cols, float eb) {
  g_err.clear();
  *outCount = 0;
  if (!data || rows <= 2 || cols <= 2) return nullptr;
  const uint32_t dims[2] = {(uint32_t)rows, (uint32_t)cols};
  CriticalPoint* out;
  if (find_cps_impl(data, 2, dims, eb, &out, outCount)) return nullptr;
  return out;
}

CriticalPoint3D* szp_cuda_find_critical_points_3d(float* data, size_t* outCount, int d1, int d2, int d3, float eb) {
  g_err.clear();
  *outCount = 0;
  if (!data || d1 <= 2 || d2 <= 2 || d3 <= 2) return nullptr;
  const uint32_t dims[3] = {(uint32_t)d1, (uint32_t)d2, (uint32_t)d3};
  CriticalPoint3D* out;
  if (find_cps_impl(data, 3, dims, eb, &out, outCount)) return nullptr;
  return out;
}

void szp_cuda_sort_critical_points_by_original_data(CriticalPoint* cps, size_t count, float* data, int cols) {
  g_err.clear();
  if (!cps || count == 0 || !data || cols <= 0) return;
  sort_cps_impl(cps, count, data, (uint32_t)cols, 0);
}

void szp_cuda_sort_critical_points_3d(CriticalPoint3D* cps, size_t count, float* data, int d2, int d3) {
  g_err.clear();
  if (!cps || count == 0 || !data || d2 <= 0 || d3 <= 0) return;
  sort_cps_impl(cps, count, data, (uint32_t)d3, (uint64_t)d2 * d3);
}

unsigned char* szp_cuda_compress_sort_positions(CriticalPoint* cps, size_t count, size_t* outSize, int bs) {
  g_err.clear();
  return compress_sp_impl(cps, count, outSize, bs);
}

unsigned char* szp_cuda_compress_sort_positions_3d(CriticalPoint3D* cps, size_t count, size_t* outSize, int bs) {
  g_err.clear();
  return compress_sp_impl(cps, count, outSize, bs);
}

unsigned char* szp_cuda_float_compress_randomaccess_topology_preserved(float* data, size_t* outSize, float eb,
                                                                       size_t n, int bs, CriticalPoint* cps,
                                                                       int count, int rows, int cols) {
  g_err.clear();
  *outSize = 0;
  if (rows <= 0 || cols <= 0) return nullptr;
  const uint32_t dims[2] = {(uint32_t)rows, (uint32_t)cols};
  return compress_topo_impl(data, outSize, eb, n, bs, cps, count, 2, dims);
}

unsigned char* szp_cuda_float_compress_topology_3d(float* data, size_t* outSize, float eb, size_t n, int bs,
                                                   CriticalPoint3D* cps, int count, int d1, int d2, int d3) {
  g_err.clear();
  *outSize = 0;
  if (d1 <= 0 || d2 <= 0 || d3 <= 0) return nullptr;
  const uint32_t dims[3] = {(uint32_t)d1, (uint32_t)d2, (uint32_t)d3};
  return compress_topo_impl(data, outSize, eb, n, bs, cps, count, 3, dims);
}

size_t szp_cuda_float_compress_randomaccess_bound(size_t n, int bs) {
  g_err.clear(); return tszp_plain_bound(n, bs, 1); }

void szp_cuda_float_compress_randomaccess_arg(unsigned char* output, float* data, size_t* outSize, float eb,
                                              size_t n, int bs) {
  g_err.clear();
  *outSize = 0;
  if (!output || !data || !(eb > 0) || check_bs(bs, false) || check_n(n)) return;
  Ctx* c;
  if (get_ctx(&c)) return;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  float* din;
  uint8_t* dout;
  uint32_t* dblk;
  const size_t cap = tszp_plain_bound(n, bs, 1);
  if (c->get(kWsIn, n * 4, (void**)&din) || c->get(kWsOut, cap, (void**)&dout) ||
      c->get(kWsBlk, (tszp_blocks(n, bs) + 1) * 4, (void**)&dblk))
    return;
  if (hipMemcpyAsync(din, data, n * 4, hipMemcpyHostToDevice, st) != hipSuccess) return;
  volatile float finv = 1 / eb;
  CompressResult res;
  if (run_codec(c, kModePlain, din, (uint32_t)n, bs, eb, (double)finv, nullptr, nullptr, dout, cap, 12, dblk,
                nullptr, nullptr, 0, st, &res))
    return;
  uint32_t u;
  memcpy(&u, &eb, 4);
  output[0] = (unsigned char)(u >> 24);
  output[1] = (unsigned char)(u >> 16);
  output[2] = (unsigned char)(u >> 8);
  output[3] = (unsigned char)u;
  memset(output + 4, 0, 8);
  if (hipMemcpy(output + 12, dout + 12, res.payload_bytes, hipMemcpyDeviceToHost) != hipSuccess) return;
  *outSize = 12 + res.payload_bytes;
}

unsigned char* szp_cuda_float_compress_randomaccess(float* data, size_t* outSize, float eb, size_t n, int bs) {
  g_err.clear();
  *outSize = 0;
  if (!data || check_bs(bs, false) || check_n(n)) return nullptr;
  unsigned char* out = (unsigned char*)malloc(tszp_plain_bound(n, bs, 1));
  if (!out) return nullptr;
  szp_cuda_float_compress_randomaccess_arg(out, data, outSize, eb, n, bs);
  if (*outSize == 0) {
    free(out);
    return nullptr;
  }
  unsigned char* shrunk = (unsigned char*)realloc(out, *outSize);
  return shrunk ? shrunk : out;
}

void szp_cuda_float_decompress_randomaccess_arg(float* newData, size_t n, float eb, int bs, unsigned char* cmp) {
  g_err.clear();
  if (!newData || !cmp || check_bs(bs, false) || check_n(n)) return;
  Ctx* c;
  if (get_ctx(&c)) return;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  unsigned T;
  uint8_t* dp;
  uint32_t* db;
  float* dout;
  if (upload_stream(c, cmp, n, bs, false, 4 * n + 4, &T, &dp, &db, st)) return;
  if (c->get(kWsOut2, n * 4, (void**)&dout)) return;
  if (decode(c, kModePlain, dp, db, n, bs, eb, dout, nullptr, nullptr, st)) return;
  if (hipMemcpy(newData, dout, n * 4, hipMemcpyDeviceToHost) != hipSuccess) fail(TSZP_EHIP, "hipMemcpy");
}

float* szp_cuda_float_decompress_randomaccess(size_t n, float eb, int bs, unsigned char* cmp) {
  g_err.clear();
  if (!cmp || check_n(n)) return nullptr;
  float* out = (float*)malloc(n * 4);
  if (!out) return nullptr;
  g_err.clear();
  szp_cuda_float_decompress_randomaccess_arg(out, n, eb, bs, cmp);
  if (!g_err.empty()) {
    free(out);
    return nullptr;
  }
  return out;
}

// ---- double random access and the threadblock codec from host buffers -------------
// Streams are the CPU writers' with one chunk: [eb big-endian][u64 0][payload]
// (SZp/src/szp_float.cc:1445-1590, szp_double.cc:148-318, :759-920).  Decoders take
// cmp at the offset table (after the eb header), detect the chunk count and walk the
// unit offsets on the host, as SZp/src/szp_cuda_decompress.cu:480-525, 907-970 do.
extern "C++" {
namespace {

template <typename T>
void elem_compress_arg(unsigned char* output, T* data, size_t* outSize, T eb, size_t n, int bs, bool ra) {
  g_err.clear();
  if (!outSize) return;
  elem_compress_host(data, (int)sizeof(T), n, (double)eb, bs, ra, output, outSize);
}

template <typename T>
unsigned char* elem_compress_alloc(T* data, size_t* outSize, T eb, size_t n, int bs, bool ra) {
  g_err.clear();
  if (!outSize) return nullptr;
  *outSize = 0;
  if (!data || check_elem(n, bs, 1, (int)sizeof(T), "compress")) return nullptr;
  unsigned char* out = (unsigned char*)malloc(tszp_elem_bound(n, bs, 1, (int)sizeof(T), ra));
  if (!out) return nullptr;
  if (elem_compress_host(data, (int)sizeof(T), n, (double)eb, bs, ra, out, outSize)) {
    free(out);
    *outSize = 0;
    return nullptr;
  }
  unsigned char* shrunk = (unsigned char*)realloc(out, *outSize);
  return shrunk ? shrunk : out;
}

template <typename T>
void elem_decompress_arg(T* newData, size_t n, T eb, int bs, unsigned char* cmp, bool ra) {
  g_err.clear();
  if (!newData || !cmp) {
    fail(TSZP_EINVAL, "decompress: bad arguments");
    return;
  }
  if (check_elem(n, bs, 1, (int)sizeof(T), "decompress")) return;
  Ctx* c;
  if (get_ctx(&c)) return;
  std::lock_guard<std::mutex> lk(c->mu);
  void* dout;
  if (elem_decode_host(c, cmp, n, (double)eb, bs, (int)sizeof(T), ra, &dout, c->stream)) return;
  if (hipMemcpy(newData, dout, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) fail(TSZP_EHIP, "hipMemcpy");
}

template <typename T>
T* elem_decompress_alloc(size_t n, T eb, int bs, unsigned char* cmp, bool ra) {
  g_err.clear();
  if (!cmp || check_n(n)) return nullptr;
  T* out = (T*)malloc(n * sizeof(T));
  if (!out) return nullptr;
  elem_decompress_arg<T>(out, n, eb, bs, cmp, ra);
  if (!g_err.empty()) {
    free(out);
    return nullptr;
  }
  return out;
}

}  // namespace
}  // extern "C++"

unsigned char* szp_cuda_double_compress_randomaccess(double* data, size_t* outSize, double eb, size_t n, int bs) {
  return elem_compress_alloc<double>(data, outSize, eb, n, bs, true);
}
void szp_cuda_double_compress_randomaccess_arg(unsigned char* output, double* data, size_t* outSize, double eb,
                                               size_t n, int bs) {
  elem_compress_arg<double>(output, data, outSize, eb, n, bs, true);
}
double* szp_cuda_double_decompress_randomaccess(size_t n, double eb, int bs, unsigned char* cmp) {
  return elem_decompress_alloc<double>(n, eb, bs, cmp, true);
}
void szp_cuda_double_decompress_randomaccess_arg(double* newData, size_t n, double eb, int bs, unsigned char* cmp) {
  elem_decompress_arg<double>(newData, n, eb, bs, cmp, true);
}
unsigned char* szp_cuda_float_compress_threadblock(float* data, size_t* outSize, float eb, size_t n, int bs) {
  return elem_compress_alloc<float>(data, outSize, eb, n, bs, false);
}
void szp_cuda_float_compress_threadblock_arg(unsigned char* output, float* data, size_t* outSize, float eb, size_t n,
                                             int bs) {
  elem_compress_arg<float>(output, data, outSize, eb, n, bs, false);
}
unsigned char* szp_cuda_double_compress_threadblock(double* data, size_t* outSize, double eb, size_t n, int bs) {
  return elem_compress_alloc<double>(data, outSize, eb, n, bs, false);
}
void szp_cuda_double_compress_threadblock_arg(unsigned char* output, double* data, size_t* outSize, double eb,
                                              size_t n, int bs) {
  elem_compress_arg<double>(output, data, outSize, eb, n, bs, false);
}
float* szp_cuda_float_decompress_threadblock(size_t n, float eb, int bs, unsigned char* cmp) {
  return elem_decompress_alloc<float>(n, eb, bs, cmp, false);
}
void szp_cuda_float_decompress_threadblock_arg(float* newData, size_t n, float eb, int bs, unsigned char* cmp) {
  elem_decompress_arg<float>(newData, n, eb, bs, cmp, false);
}
double* szp_cuda_double_decompress_threadblock(size_t n, double eb, int bs, unsigned char* cmp) {
  return elem_decompress_alloc<double>(n, eb, bs, cmp, false);
}
void szp_cuda_double_decompress_threadblock_arg(double* newData, size_t n, double eb, int bs, unsigned char* cmp) {
  elem_decompress_arg<double>(newData, n, eb, bs, cmp, false);
}

// SZp/src/szp_cuda_compress.cu:664-710; *outSize is left alone, as there
int* szp_cuda_float_direct_predict_quantization(float* data, size_t* outSize, float eb, size_t n, int bs) {
  (void)outSize;
  (void)bs;
  g_err.clear();
  if (!data || !(eb > 0) || check_n(n)) {
    if (g_err.empty()) fail(TSZP_EINVAL, "predict_quantization: bad arguments");
    return nullptr;
  }
  Ctx* c;
  if (get_ctx(&c)) return nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  float* din;
  int32_t* dout;
  if (c->get(kWsIn, n * 4, (void**)&din) || c->get(kWsOut2, n * 4, (void**)&dout)) return nullptr;
  int* out = (int*)malloc(n * sizeof(int));
  if (!out) return nullptr;
  if (hipMemcpyAsync(din, data, n * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      launch_predict_quant(din, (uint32_t)n, 1.0 / (double)eb, dout, st) != hipSuccess ||
      hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    fail(TSZP_EHIP, "predict_quantization: device call failed");
    free(out);
    return nullptr;
  }
  return out;
}

// SZp/src/szp_cuda_compress.cu:700-709: the same result as the direct variant
int* szp_cuda_float_threadblock_predict_quantization(float* data, size_t* outSize, float eb, size_t n, int bs) {
  return szp_cuda_float_direct_predict_quantization(data, outSize, eb, n, bs);
}

// SZp/src/szp_cuda_topology_3d.cu:1065-1188: [u64 0][blocks], host buffers
unsigned char* szp_cuda_double_compress_topology_3d(double* data, size_t* outSize, double eb, size_t nbEle, int bs,
                                                    CriticalPoint3D* cps, int count, int d1, int d2, int d3) {
  g_err.clear();
  if (!outSize) return nullptr;
  *outSize = 0;
  if (!data || d1 <= 0 || d2 <= 0 || d3 <= 0 || count < 0 || (count > 0 && !cps)) {
    fail(TSZP_EINVAL, "szp_cuda_double_compress_topology_3d: bad arguments");
    return nullptr;
  }
  const uint32_t dims[3] = {(uint32_t)d1, (uint32_t)d2, (uint32_t)d3};
  size_t n;
  if (check_topo64(dims, eb, bs, (uint64_t)count, &n)) return nullptr;
  if (n != nbEle) {
    fail(TSZP_EINVAL, "szp_cuda_double_compress_topology_3d: nbEle != d1 * d2 * d3");
    return nullptr;
  }
  Ctx* c;
  if (get_ctx(&c)) return nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  const size_t cap = tszp_topo_bound(n, bs, 1);
  double* din;
  int32_t* dc = nullptr;
  uint8_t* dout;
  uint32_t* dblk;
  if (c->get(kWsIn, n * 8, (void**)&din) || c->get(kWsOut, cap, (void**)&dout) ||
      c->get(kWsBlk, (tszp_blocks(n, bs) + 1) * 4, (void**)&dblk))
    return nullptr;
  if (count && c->get(kWsCps, (size_t)count * sizeof(CriticalPoint3D), (void**)&dc)) return nullptr;
  if (hipMemcpyAsync(din, data, n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
      (count && hipMemcpyAsync(dc, cps, (size_t)count * sizeof(CriticalPoint3D), hipMemcpyHostToDevice, st) !=
                    hipSuccess)) {
    fail(TSZP_EHIP, "hipMemcpyAsync");
    return nullptr;
  }
  uint64_t nbytes = 0;
  if (topo64_impl(c, din, dims, n, eb, bs, dc, (uint64_t)count, dout, cap, dblk, &nbytes, nullptr, nullptr, st))
    return nullptr;
  unsigned char* out = (unsigned char*)malloc(nbytes);
  if (!out) return nullptr;
  if (hipMemcpy(out, dout, nbytes, hipMemcpyDeviceToHost) != hipSuccess) {
    fail(TSZP_EHIP, "hipMemcpy");
    free(out);
    return nullptr;
  }
  *outSize = nbytes;
  return out;
}

void szp_cuda_float_decompress_randomaccess_topology_preserved(float** newData, size_t n, float eb, int bs,
                                                               unsigned char* cmp, int** FN) {
  g_err.clear();
  *newData = nullptr;
  *FN = nullptr;
  if (!cmp || n == 0 || check_bs(bs, true) || check_n(n)) return;
  Ctx* c;
  if (get_ctx(&c)) return;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  unsigned T;
  uint8_t* dp;
  uint32_t* db;
  float* dout;
  int32_t* dfn;
  if (upload_stream(c, cmp, n, bs, true, 8 * n + 1024, &T, &dp, &db, st)) return;
  if (c->get(kWsOut2, n * 4, (void**)&dout) || c->get(kWsTypes, n * 4, (void**)&dfn)) return;
  if (decode(c, kModeTopoMap, dp, db, n, bs, eb, dout, nullptr, dfn, st)) return;
  float* h = (float*)malloc(n * 4);
  int* f = (int*)malloc(n * 4);
  if (!h || !f || hipMemcpy(h, dout, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(f, dfn, n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
    free(h);
    free(f);
    fail(TSZP_EHIP, "copy back failed");
    return;
  }
  *newData = h;
  *FN = f;
}

int* szp_cuda_decompress_sort_positions(unsigned char* cmp, size_t count, int bs) {
  g_err.clear();
  if (!cmp || count == 0 || check_bs(bs, true) || check_n(count)) return nullptr;
  Ctx* c;
  if (get_ctx(&c)) return nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = c->stream;
  unsigned T;
  uint8_t* dp;
  uint32_t* db;
  int32_t* dout;
  if (upload_stream(c, cmp, count, bs, false, 4 * count + 4, &T, &dp, &db, st)) return nullptr;
  if (c->get(kWsOut2, count * 4, (void**)&dout)) return nullptr;
  if (decode(c, kModeInt, dp, db, count, bs, 0.f, dout, nullptr, nullptr, st)) return nullptr;
  int* h = (int*)malloc(count * 4);
  if (!h) return nullptr;
  if (hipMemcpy(h, dout, count * 4, hipMemcpyDeviceToHost) != hipSuccess) {
    free(h);
    return nullptr;
  }
  return h;
}

}  // extern "C"
