// tszp_batch_api.cpp -- host side of the batched topology pipeline (include/toposzp_hip.h,
// tszp_topo_batch_*): many 2D / 3D fields through one launch per stage and block-size
// class.  The per-field chain it replaces is the reference's
// example/test_cuda_topology.cu:374-530 (SZp/src/szp_cuda_topology.cu:77-772 for 2D,
// SZp/src/szp_cuda_topology_3d.cu:174-1059 for 3D), run once per field.
//
// Stages (every launch covers all fields; no host round trip in compress or decompress):
//   compress    topology codec with on-the-fly classification (one launch per class),
//               per-field offsets of the critical-point keys, their gather into one
//               key set with the field as the high word, one radix sort by
//               (field, value), one segmented rank -> every field's sort positions,
//               sort-position codec (one launch per class, counts read on the device)
//   decompress  topology decode (values + types) and sort-position decode per class
//   restore     the restoration over the batch's flat arrays (tszp_api.cpp)
// The field order inside the batch is the topology codec's class order; results go
// back to the caller's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/toposzp_hip.h"
#include "tszp_kernels.h"

using namespace tszp;

namespace {

int bfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return set_error(code, buf);
}

#define BHIP(x)                                                                            \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return bfail(TSZP_EHIP, "%s: %s", #x, hipGetErrorString(e_));    \
  } while (0)

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

bool pow2_v4(int bs) { return bs == 16 || bs == 32 || bs == 64 || bs == 128 || bs == 256; }

// One stage: the fields in the order of its classes, numbered tiles, the device table
// and tile map.
struct Stage {
  struct Class {
    uint32_t bs;
    bool v4;
    int nd;
    uint32_t base, ntiles;
  };
  std::vector<Class> classes;
  std::vector<uint32_t> order;  // table entry -> batch index
  std::vector<BatchField> tab;
  uint32_t tiles = 0;
  BatchField* dtab = nullptr;
  uint32_t* tmap = nullptr;
};

// HIP event pair timing a stage when the library's timers are on
struct StageTimer {
  hipEvent_t a = nullptr, b = nullptr;
  int kind;
  StageTimer(int k, hipStream_t st) : kind(k) {
    if (!prof_on()) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      a = b = nullptr;
      return;
    }
    (void)hipEventRecord(a, st);
  }
  void stop(hipStream_t st) {
    if (a) (void)hipEventRecord(b, st);
  }
  ~StageTimer() {
    if (!a) return;
    float ms = 0.f;
    if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) prof_add(kind, ms);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
};

}  // namespace

struct tszp_topo_batch {
  int device = 0;
  uint32_t count = 0;
  std::vector<tszp_topo_batch_field> f;  // batch order
  std::vector<uint32_t> caller;          // batch index -> caller's index
  std::vector<uint64_t> base;            // flat offset of each field
  std::vector<uint64_t> n;               // elements of each field
  uint64_t flat = 0, total_n = 0;
  bool can_compress = true, compressed = false, decompressed = false, restored = false;
  Stage tc, ic, td, id;
  int hi_bits = 8;
  uint32_t nchunks = 0;
  void* mem = nullptr;
  bool own_mem = true;  // false: mem is the caller's (tszp_topo_batch_create_mem), host results pageable
  // device arrays (carved from mem)
  TopoBatchDev* dev = nullptr;
  float* feb = nullptr;
  double* finv = nullptr;
  uint32_t *cpk = nullptr, *cpp = nullptr, *chunk_field = nullptr, *cbase = nullptr, *misc = nullptr;
  uint32_t *skey = nullptr, *shi = nullptr, *spay = nullptr;
  SortWork sw{};
  size_t sort_st_words = 0;
  int32_t* spv = nullptr;
  float *dec = nullptr, *rest = nullptr;
  uint8_t* types = nullptr;
  uint64_t *st_tc = nullptr, *st_ic = nullptr;
  uint32_t* counters = nullptr;  // one per class of tc, then of ic
  CompressResult *res_tc = nullptr, *res_ic = nullptr, *res_td = nullptr, *res_id = nullptr;
  CompressResult* host = nullptr;  // pinned: 4 * count results + misc
  // restoration results of the last tszp_topo_batch_restore (batch order)
  std::vector<tszp_restore_info> rinfo;
};

namespace {

// Orders the batch's fields into the classes of a stage (key: kernel geometry) and
// numbers their tiles.  fill(k, i, b) fills table entry b of batch field i.
template <class KeyFn, class FillFn>
int build_stage(tszp_topo_batch* B, Stage& s, KeyFn key, FillFn fill) {
  const uint32_t count = B->count;
  std::vector<std::pair<uint64_t, uint32_t>> k(count);
  for (uint32_t i = 0; i < count; i++) k[i] = {key(i), i};
  std::stable_sort(k.begin(), k.end(), [](const std::pair<uint64_t, uint32_t>& a, const std::pair<uint64_t, uint32_t>& b) {
    return a.first < b.first;
  });
  s.order.resize(count);
  s.tab.resize(count);
  s.classes.clear();
  uint64_t tiles = 0;
  for (uint32_t j = 0; j < count; j++) {
    const uint32_t i = k[j].second;
    const tszp_topo_batch_field& fl = B->f[i];
    int S, EPL;
    uint32_t bpt;
    block_geometry((uint32_t)fl.block_size, &S, &EPL, &bpt);
    BatchField b{};
    b.n = (uint32_t)B->n[i];
    b.bs = (uint32_t)fl.block_size;
    b.nb = cdiv(b.n, b.bs);
    b.ntiles = cdiv(b.nb, bpt);
    b.tile0 = (uint32_t)tiles;
    b.eb = fl.eb;
    fill(i, b);
    const bool v4 = (k[j].first >> 40) & 1u;
    const int nd = (int)((k[j].first >> 36) & 15u);
    if (j == 0 || k[j].first != k[j - 1].first) s.classes.push_back({b.bs, v4, nd, (uint32_t)tiles, 0});
    s.classes.back().ntiles += b.ntiles;
    s.order[j] = i;
    s.tab[j] = b;
    tiles += b.ntiles;
    if (tiles >= 0x7fffffffull) return bfail(TSZP_EINVAL, "topo_batch: more than 2^31 tiles in one batch");
  }
  s.tiles = (uint32_t)tiles;
  return TSZP_OK;
}

// key of a geometry class: the 16-byte path by block size (and dimension for the
// classifying codec), the generic kernels by lane geometry
uint64_t class_key(int bs, bool v4, int nd) {
  int S, EPL;
  uint32_t bpt;
  block_geometry((uint32_t)bs, &S, &EPL, &bpt);
  if (v4) return (1ull << 40) | ((uint64_t)nd << 36) | (uint64_t)bs;
  return ((uint64_t)S << 16) | (uint32_t)EPL;
}

struct Carve {
  size_t off = 0;
  template <class T>
  T* take(size_t count, T* base_dummy = nullptr) {
    (void)base_dummy;
    const size_t at = off;
    off += (count * sizeof(T) + 255) & ~(size_t)255;
    return reinterpret_cast<T*>(at);
  }
};

template <class T>
void rebase(T*& p, void* mem) {
  p = reinterpret_cast<T*>(static_cast<char*>(mem) + reinterpret_cast<size_t>(p));
}

int batch_device_error(const tszp_topo_batch* B, const char* what, uint32_t bi, uint32_t bits) {
  return bfail(TSZP_EDEVICE, "%s: field %u: device error bits 0x%x (%s%s%s)", what, B->caller[bi], bits,
               (bits & TSZP_ERR_LOOKBACK) ? "look-back timeout " : "", (bits & TSZP_ERR_STREAM) ? "corrupt stream " : "",
               (bits & TSZP_ERR_OVERFLOW) ? "capacity " : "");
}

void destroy(tszp_topo_batch* B) {
  if (!B) return;
  if (B->host) {
    if (B->own_mem) (void)hipHostFree(B->host);
    else free(B->host);
  }
  if (B->mem && B->own_mem) (void)hipFree(B->mem);
  delete B;
}

}  // namespace

extern "C" {

namespace {
int topo_batch_create(tszp_topo_batch_field* fields, uint32_t count, void* ext_mem, uint64_t ext_bytes,
                      tszp_topo_batch** out, void* stream);
}

int tszp_topo_batch_create(tszp_topo_batch_field* fields, uint32_t count, tszp_topo_batch** out, void* stream) {
  return topo_batch_create(fields, count, nullptr, 0, out, stream);
}

int tszp_topo_batch_create_mem(tszp_topo_batch_field* fields, uint32_t count, void* d_mem, uint64_t mem_bytes,
                               tszp_topo_batch** out, void* stream) {
  if (!d_mem || ((uintptr_t)d_mem & 255u)) {
    clear_error();
    return bfail(TSZP_EINVAL, "topo_batch_create_mem: memory must be non-null and 256-byte aligned");
  }
  return topo_batch_create(fields, count, d_mem, mem_bytes, out, stream);
}

uint64_t tszp_topo_batch_mem_bound(uint64_t elements, uint32_t count) {
  // per element: critical-point keys and payloads, sort keys and their alternates,
  // sort positions, decoded and restored values, types, tile maps and look-back words;
  // per field: tables, results and the padding of its flat range
  return 80ull * (elements + (uint64_t)count * kFlatAlign) + 4096ull * count + (4ull << 20);
}

}  // extern "C"

namespace {
int topo_batch_create(tszp_topo_batch_field* fields, uint32_t count, void* ext_mem, uint64_t ext_bytes,
                      tszp_topo_batch** out, void* stream) {
  clear_error();
  if (!out) return bfail(TSZP_EINVAL, "topo_batch_create: null output");
  *out = nullptr;
  if (!fields || count == 0) return bfail(TSZP_EINVAL, "topo_batch_create: no fields");
  if (count > (1u << 16)) return bfail(TSZP_EINVAL, "topo_batch_create: %u fields, at most 65536 per batch", count);
  hipStream_t st = (hipStream_t)stream;
  // validation, and block sizes for the entries that ask the tuner
  std::vector<tszp_batch_field> tune;
  std::vector<uint32_t> tune_idx;
  bool compressible = true;
  for (uint32_t i = 0; i < count; i++) {
    tszp_topo_batch_field& f = fields[i];
    if (f.ndim != 2 && f.ndim != 3) return bfail(TSZP_EINVAL, "topo_batch_create: field %u: ndim must be 2 or 3", i);
    uint64_t n = 1;
    for (int a = 0; a < f.ndim; a++) {
      if (f.dims[a] == 0) return bfail(TSZP_EINVAL, "topo_batch_create: field %u: zero dimension", i);
      n *= f.dims[a];
    }
    if (n >= 0x7fffffffull) return bfail(TSZP_EINVAL, "topo_batch_create: field %u: more than 2^31 elements", i);
    if (!(f.eb > 0) || !std::isfinite(f.eb))
      return bfail(TSZP_EINVAL, "topo_batch_create: field %u: error bound must be positive and finite", i);
    if (f.block_size < 0 || f.block_size > 257)
      return bfail(TSZP_EINVAL, "topo_batch_create: field %u: blockSize %d out of range [1, 257] (0: tuned)", i,
                   f.block_size);
    if (!f.d_topo || !f.d_topo_blk || !f.d_sortpos || !f.d_sortpos_blk)
      return bfail(TSZP_EINVAL, "topo_batch_create: field %u: null stream or side-car", i);
    if (((uintptr_t)f.d_topo | (uintptr_t)f.d_sortpos) & 3u)
      return bfail(TSZP_EINVAL, "topo_batch_create: field %u: streams must be 4-byte aligned", i);
    if (f.topo_cap < kBatchTopoHdr + 4 || f.sortpos_cap < kBatchSpHdr + 4)
      return bfail(TSZP_ECAP, "topo_batch_create: field %u: stream capacity too small", i);
    if (!f.d_field) compressible = false;
    if (f.block_size == 0) {
      if (!f.d_field) return bfail(TSZP_EINVAL, "topo_batch_create: field %u: blockSize 0 needs the field", i);
      tszp_batch_field t{};
      t.d_field = const_cast<float*>(f.d_field);
      t.n = n;
      t.eb = f.eb;
      tune.push_back(t);
      tune_idx.push_back(i);
    }
  }
  if (!tune.empty()) {
    int r = tszp_batch_tune(tune.data(), (uint32_t)tune.size(), stream);
    if (r) return r;
    for (size_t k = 0; k < tune.size(); k++) fields[tune_idx[k]].block_size = tune[k].block_size;
  }
  auto* B = new tszp_topo_batch();
  B->count = count;
  B->can_compress = compressible;
  (void)hipGetDevice(&B->device);
  const bool generic = getenv("TSZP_CODEC_GENERIC") != nullptr;
  // batch order: the topology codec's classes
  {
    std::vector<std::pair<uint64_t, uint32_t>> k(count);
    for (uint32_t i = 0; i < count; i++) {
      const tszp_topo_batch_field& f = fields[i];
      const uint32_t fast = f.ndim == 2 ? f.dims[1] : f.dims[2];
      const bool v4 = !generic && pow2_v4(f.block_size) && f.d_field && !((uintptr_t)f.d_field & 15u) && fast % 4 == 0;
      k[i] = {class_key(f.block_size, v4, f.ndim), i};
    }
    std::stable_sort(k.begin(), k.end(), [](const std::pair<uint64_t, uint32_t>& a,
                                            const std::pair<uint64_t, uint32_t>& b) { return a.first < b.first; });
    B->f.resize(count);
    B->caller.resize(count);
    B->base.resize(count);
    B->n.resize(count);
    uint64_t flat = 0, tot = 0;
    uint32_t chunks = 0;
    for (uint32_t j = 0; j < count; j++) {
      const uint32_t i = k[j].second;
      B->f[j] = fields[i];
      B->caller[j] = i;
      uint64_t n = 1;
      for (int a = 0; a < fields[i].ndim; a++) n *= fields[i].dims[a];
      B->n[j] = n;
      B->base[j] = flat;
      flat += (n + kFlatAlign - 1) / kFlatAlign * kFlatAlign;
      tot += n;
      chunks += cdiv(n, kSortTile);
    }
    if (flat >= 0xffffffffull) {
      destroy(B);
      return bfail(TSZP_EINVAL, "topo_batch_create: the fields need more than 2^32 elements of flat arrays");
    }
    B->flat = flat;
    B->total_n = tot;
    B->nchunks = chunks;
    int hb = 8;
    while (hb < 32 && (1ull << hb) < count) hb += 8;
    B->hi_bits = hb;
  }
  // per-stage layouts (device pointers filled after the carve)
  int r = build_stage(B, B->tc,
                      [&](uint32_t i) {
                        const tszp_topo_batch_field& f = B->f[i];
                        const uint32_t fast = f.ndim == 2 ? f.dims[1] : f.dims[2];
                        const bool v4 = !generic && pow2_v4(f.block_size) && f.d_field &&
                                        !((uintptr_t)f.d_field & 15u) && fast % 4 == 0;
                        return class_key(f.block_size, v4, f.ndim);
                      },
                      [&](uint32_t, BatchField&) {});
  if (!r) r = build_stage(B, B->ic, [&](uint32_t i) { return class_key(B->f[i].block_size, !generic && pow2_v4(B->f[i].block_size), 2); },
                          [&](uint32_t, BatchField&) {});
  if (!r) r = build_stage(B, B->td, [&](uint32_t i) { return class_key(B->f[i].block_size, false, 2); },
                          [&](uint32_t, BatchField&) {});
  if (!r) r = build_stage(B, B->id, [&](uint32_t i) { return class_key(B->f[i].block_size, false, 2); },
                          [&](uint32_t, BatchField&) {});
  if (r) {
    destroy(B);
    return r;
  }
  // the tc order is the batch order by construction
  for (uint32_t j = 0; j < count; j++)
    if (B->tc.order[j] != j) {
      destroy(B);
      return bfail(TSZP_EINVAL, "topo_batch_create: internal field order");
    }
  // one device allocation for everything
  const uint64_t flat = B->flat, tn = B->total_n;
  Carve cv;
  B->dev = cv.take<TopoBatchDev>(count);
  B->feb = cv.take<float>(count);
  B->finv = cv.take<double>(count);
  B->tc.dtab = cv.take<BatchField>(count);
  B->ic.dtab = cv.take<BatchField>(count);
  B->td.dtab = cv.take<BatchField>(count);
  B->id.dtab = cv.take<BatchField>(count);
  B->tc.tmap = cv.take<uint32_t>(B->tc.tiles);
  B->ic.tmap = cv.take<uint32_t>(B->ic.tiles);
  B->td.tmap = cv.take<uint32_t>(B->td.tiles);
  B->id.tmap = cv.take<uint32_t>(B->id.tiles);
  B->chunk_field = cv.take<uint32_t>(B->nchunks);
  B->cbase = cv.take<uint32_t>(count);
  B->misc = cv.take<uint32_t>(64);
  const size_t ncls = B->tc.classes.size() + B->ic.classes.size();
  B->st_tc = cv.take<uint64_t>(2 * (size_t)B->tc.tiles);
  B->st_ic = cv.take<uint64_t>(B->ic.tiles);
  B->counters = cv.take<uint32_t>(ncls + 2);
  B->res_tc = cv.take<CompressResult>(4 * (size_t)count);  // the four stages' results, contiguous
  if (compressible) {
    B->cpk = cv.take<uint32_t>(flat);
    B->cpp = cv.take<uint32_t>(flat);
    B->skey = cv.take<uint32_t>(tn);
    B->shi = cv.take<uint32_t>(tn);
    B->spay = cv.take<uint32_t>(tn);
    B->sw.keys_alt = cv.take<uint32_t>(tn);
    B->sw.hi_alt = cv.take<uint32_t>(tn);
    B->sw.pay_alt = cv.take<uint32_t>(tn);
    B->sw.hist = cv.take<uint32_t>(sort_hist_words((uint32_t)tn));
    B->sort_st_words = sort_status_words((uint32_t)tn) + 8;
    B->sw.st = cv.take<uint64_t>(B->sort_st_words + 8);
  }
  B->spv = cv.take<int32_t>(flat + 4);
  B->dec = cv.take<float>(flat);
  B->rest = cv.take<float>(flat);
  B->types = cv.take<uint8_t>(flat + 16);
  hipError_t e = hipSuccess;
  if (ext_mem) {  // the caller's memory; pageable host results
    B->own_mem = false;
    if (cv.off > ext_bytes) {
      destroy(B);
      return bfail(TSZP_ECAP, "topo_batch_create_mem: the batch needs %zu bytes, %llu given", cv.off,
                   (unsigned long long)ext_bytes);
    }
    B->mem = ext_mem;
    B->host = static_cast<CompressResult*>(malloc((4 * (size_t)count + 8) * sizeof(CompressResult)));
    if (!B->host) {
      destroy(B);
      return bfail(TSZP_EHIP, "topo_batch_create_mem: host allocation failed");
    }
  } else {
    e = hipMalloc(&B->mem, cv.off);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&B->host, (4 * (size_t)count + 8) * sizeof(CompressResult), hipHostMallocDefault);
    if (e != hipSuccess) {
      destroy(B);
      return bfail(TSZP_EHIP, "topo_batch_create: allocation of %zu bytes: %s", cv.off, hipGetErrorString(e));
    }
  }
  void* m = B->mem;
  rebase(B->dev, m); rebase(B->feb, m); rebase(B->finv, m);
  rebase(B->tc.dtab, m); rebase(B->ic.dtab, m); rebase(B->td.dtab, m); rebase(B->id.dtab, m);
  rebase(B->tc.tmap, m); rebase(B->ic.tmap, m); rebase(B->td.tmap, m); rebase(B->id.tmap, m);
  rebase(B->chunk_field, m); rebase(B->cbase, m); rebase(B->misc, m);
  rebase(B->st_tc, m); rebase(B->st_ic, m); rebase(B->counters, m);
  rebase(B->res_tc, m);
  B->res_ic = B->res_tc + count;
  B->res_td = B->res_tc + 2 * (size_t)count;
  B->res_id = B->res_tc + 3 * (size_t)count;
  if (compressible) {
    rebase(B->cpk, m); rebase(B->cpp, m); rebase(B->skey, m); rebase(B->shi, m); rebase(B->spay, m);
    rebase(B->sw.keys_alt, m); rebase(B->sw.hi_alt, m); rebase(B->sw.pay_alt, m); rebase(B->sw.hist, m);
    rebase(B->sw.st, m);
    B->sw.counters = reinterpret_cast<uint32_t*>(B->sw.st + sort_status_words((uint32_t)tn));
    B->sw.err = B->sw.counters + 15;
  }
  rebase(B->spv, m); rebase(B->dec, m); rebase(B->rest, m); rebase(B->types, m);
  // tables
  std::vector<TopoBatchDev> dv(count);
  std::vector<float> feb(count);
  std::vector<double> finv(count);
  uint32_t c0 = 0;
  for (uint32_t j = 0; j < count; j++) {
    dv[j].base = B->base[j];
    dv[j].n = (uint32_t)B->n[j];
    dv[j].chunk0 = c0;
    dv[j].nchunks = cdiv(B->n[j], kSortTile);
    dv[j].sp_stream = B->f[j].d_sortpos;
    c0 += dv[j].nchunks;
    feb[j] = B->f[j].eb;
    finv[j] = 1.0 / (double)B->f[j].eb;
  }
  for (uint32_t k = 0; k < count; k++) {  // topology compression
    const uint32_t i = B->tc.order[k];
    const tszp_topo_batch_field& f = B->f[i];
    BatchField& b = B->tc.tab[k];
    b.in = f.d_field;
    b.out = f.d_topo;
    b.blk = f.d_topo_blk;
    const uint64_t pc = f.topo_cap - kBatchTopoHdr - 4;  // keep 4 readable bytes past the end
    b.cap = pc > 0xffffffffull ? 0xffffffffull : pc;
    b.inv = 1.0 / (double)f.eb;  // the topology path's reciprocal (SZp/src/szp_float.cc:1067)
    b.hdr = kBatchTopoHdr;
    b.grid = grid_of(f.ndim, f.dims);
    if (compressible) {
      b.cp_key = B->cpk + B->base[i];
      b.cp_pay = B->cpp + B->base[i];
    }
    b.cp_cap = (uint32_t)B->n[i];
  }
  for (uint32_t k = 0; k < count; k++) {  // sort-position compression: E on the device
    const uint32_t i = B->ic.order[k];
    const tszp_topo_batch_field& f = B->f[i];
    BatchField& b = B->ic.tab[k];
    b.in = B->spv + B->base[i];
    b.out = f.d_sortpos;
    b.blk = f.d_sortpos_blk;
    const uint64_t pc = f.sortpos_cap - kBatchSpHdr - 4;
    b.cap = pc > 0xffffffffull ? 0xffffffffull : pc;
    b.eb = 0.f;
    b.hdr = kBatchSpHdr;
    b.n_dev = &B->res_tc[i].ext_count;
  }
  for (uint32_t k = 0; k < count; k++) {  // topology decode into the flat arrays
    const uint32_t i = B->td.order[k];
    const tszp_topo_batch_field& f = B->f[i];
    BatchField& b = B->td.tab[k];
    b.in = f.d_topo;
    b.out = B->dec + B->base[i];
    b.types8 = B->types + B->base[i];
    b.blk = f.d_topo_blk;
    b.hdr = kBatchTopoHdr;
  }
  for (uint32_t k = 0; k < count; k++) {  // sort-position decode: E from the stream head
    const uint32_t i = B->id.order[k];
    const tszp_topo_batch_field& f = B->f[i];
    BatchField& b = B->id.tab[k];
    b.in = f.d_sortpos;
    b.out = B->spv + B->base[i];
    b.blk = f.d_sortpos_blk;
    b.eb = 0.f;
    b.hdr = kBatchSpHdr;
    b.n_dev = reinterpret_cast<const uint64_t*>(f.d_sortpos);
  }
  auto fin = [&](int code) {
    (void)hipStreamSynchronize(st);
    destroy(B);
    return code;
  };
  if ((e = hipMemcpyAsync(B->dev, dv.data(), count * sizeof(TopoBatchDev), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(B->feb, feb.data(), count * sizeof(float), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(B->finv, finv.data(), count * sizeof(double), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(B->tc.dtab, B->tc.tab.data(), count * sizeof(BatchField), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(B->ic.dtab, B->ic.tab.data(), count * sizeof(BatchField), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(B->td.dtab, B->td.tab.data(), count * sizeof(BatchField), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(B->id.dtab, B->id.tab.data(), count * sizeof(BatchField), hipMemcpyHostToDevice, st)) ||
      (e = launch_batch_tilemap(B->tc.dtab, count, B->tc.tmap, st)) ||
      (e = launch_batch_tilemap(B->ic.dtab, count, B->ic.tmap, st)) ||
      (e = launch_batch_tilemap(B->td.dtab, count, B->td.tmap, st)) ||
      (e = launch_batch_tilemap(B->id.dtab, count, B->id.tmap, st)) ||
      (e = launch_tb_chunkmap(B->dev, count, B->chunk_field, st)) || (e = hipStreamSynchronize(st)))
    return fin(bfail(TSZP_EHIP, "topo_batch_create: %s", hipGetErrorString(e)));
  *out = B;
  return TSZP_OK;
}
}  // namespace

extern "C" {

int tszp_topo_batch_destroy(tszp_topo_batch* B) {
  clear_error();
  if (!B) return bfail(TSZP_EINVAL, "topo_batch_destroy: null batch");
  destroy(B);
  return TSZP_OK;
}

int tszp_topo_batch_compress(tszp_topo_batch* B, void* stream) {
  clear_error();
  if (!B) return bfail(TSZP_EINVAL, "topo_batch_compress: null batch");
  if (!B->can_compress) return bfail(TSZP_EINVAL, "topo_batch_compress: the batch was created without fields");
  hipStream_t st = (hipStream_t)stream;
  const uint32_t count = B->count;
  StageTimer timer(kProfTopoBatch, st);
  {
    FillRanges fr{};
    fr.add(reinterpret_cast<uint32_t*>(B->st_tc), 0u, 4 * (size_t)B->tc.tiles);
    fr.add(reinterpret_cast<uint32_t*>(B->st_ic), 0u, 2 * (size_t)B->ic.tiles);
    fr.add(B->counters, 0u, B->tc.classes.size() + B->ic.classes.size() + 2);
    fr.add(reinterpret_cast<uint32_t*>(B->res_tc), 0u, 2 * (size_t)count * sizeof(CompressResult) / 4);
    fr.add(B->misc, 0u, 64);
    fr.add(reinterpret_cast<uint32_t*>(B->sw.st), 0u, 2 * B->sort_st_words);
    BHIP(launch_fill_multi(fr, st));  // launches only: safe under stream capture
  }
  size_t ci = 0;
  for (const Stage::Class& cl : B->tc.classes)
    BHIP(launch_compress_batch(kModeTopoClassify, cl.bs, cl.v4, cl.nd, B->tc.dtab, B->tc.tmap, cl.base, cl.ntiles,
                               B->st_tc, B->st_tc + B->tc.tiles, B->res_tc, dyn_counter(B->counters + ci++), st));
  {
    StageTimer ts(kProfTopoBatchSort, st);
    BHIP(launch_tb_offsets(B->dev, count, B->res_tc, B->cbase, B->misc, B->misc + 1, st));
    BHIP(launch_tb_gather(B->dev, B->chunk_field, B->nchunks, B->res_tc, B->cbase, B->cpk, B->cpp, B->skey, B->shi,
                          B->spay, st));
    BHIP(radix_sort_pairs(B->skey, B->shi, B->spay, (uint32_t)B->total_n, B->sw, st, nullptr, B->misc, B->hi_bits));
    BHIP(launch_rank(B->skey, B->shi, B->spay, (uint32_t)B->total_n, 0.f, 0.0, B->spv, (uint32_t)B->flat, B->sw, st,
                     B->misc, B->feb, B->finv));
    ts.stop(st);
  }
  for (const Stage::Class& cl : B->ic.classes)
    BHIP(launch_compress_batch(kModeInt, cl.bs, cl.v4, 2, B->ic.dtab, B->ic.tmap, cl.base, cl.ntiles, B->st_ic,
                               nullptr, B->res_ic, dyn_counter(B->counters + ci++), st));
  timer.stop(st);
  B->compressed = true;
  return TSZP_OK;
}

int tszp_topo_batch_decompress(tszp_topo_batch* B, void* stream) {
  clear_error();
  if (!B) return bfail(TSZP_EINVAL, "topo_batch_decompress: null batch");
  hipStream_t st = (hipStream_t)stream;
  const uint32_t count = B->count;
  StageTimer timer(kProfTopoBatchDecode, st);
  BHIP(launch_fill_u32(reinterpret_cast<uint32_t*>(B->res_td), 0u, 2 * (size_t)count * sizeof(CompressResult) / 4, st));
  for (const Stage::Class& cl : B->td.classes)
    BHIP(launch_decode_batch(kModeTopoMap, cl.bs, B->td.dtab, B->td.tmap, cl.base, cl.ntiles, B->res_td, st));
  for (const Stage::Class& cl : B->id.classes)
    BHIP(launch_decode_batch(kModeInt, cl.bs, B->id.dtab, B->id.tmap, cl.base, cl.ntiles, B->res_id, st));
  timer.stop(st);
  B->decompressed = true;
  return TSZP_OK;
}

int tszp_topo_batch_restore(tszp_topo_batch* B, int max_rounds, void* stream) {
  return tszp_topo_batch_restore_ws(B, max_rounds, nullptr, stream);
}

int tszp_topo_batch_restore_ws(tszp_topo_batch* B, int max_rounds, tszp_workspace* ws, void* stream) {
  clear_error();
  if (!B) return bfail(TSZP_EINVAL, "topo_batch_restore: null batch");
  if (max_rounds < 0) return bfail(TSZP_EINVAL, "topo_batch_restore: max_rounds must be >= 0");
  hipStream_t st = (hipStream_t)stream;
  StageTimer timer(kProfTopoBatchRestore, st);
  std::vector<RestoreBatchField> rf(B->count);
  for (uint32_t j = 0; j < B->count; j++) {
    rf[j].ndim = B->f[j].ndim;
    memcpy(rf[j].dims, B->f[j].dims, sizeof rf[j].dims);
    rf[j].eb = B->f[j].eb;
    rf[j].base = B->base[j];
    rf[j].n_ext_dev = reinterpret_cast<const uint64_t*>(B->f[j].d_sortpos);
  }
  B->rinfo.assign(B->count, tszp_restore_info{});
  int r = restore_batch(rf.data(), B->count, B->flat, B->dec, B->types, B->spv, B->rest, max_rounds,
                        B->rinfo.data(), ws, st);
  timer.stop(st);
  if (r) return r;
  B->restored = true;
  return TSZP_OK;
}

int tszp_topo_batch_results(tszp_topo_batch* B, tszp_topo_batch_field* fields, void* stream) {
  clear_error();
  if (!B || !fields) return bfail(TSZP_EINVAL, "topo_batch_results: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const uint32_t count = B->count;
  CompressResult* h = B->host;
  BHIP(hipMemcpyAsync(h, B->res_tc, 4 * (size_t)count * sizeof(CompressResult), hipMemcpyDeviceToHost, st));
  uint32_t* hm = reinterpret_cast<uint32_t*>(h + 4 * (size_t)count);
  BHIP(hipMemcpyAsync(hm, B->misc, 64 * 4, hipMemcpyDeviceToHost, st));
  uint32_t sort_err = 0;
  if (B->compressed) BHIP(hipMemcpyAsync(&sort_err, B->sw.err, 4, hipMemcpyDeviceToHost, st));
  BHIP(hipStreamSynchronize(st));
  const CompressResult *htc = h, *hic = h + count, *htd = h + 2 * (size_t)count, *hid = h + 3 * (size_t)count;
  if (B->compressed) {
    if (hm[1] || sort_err)
      return bfail(TSZP_EDEVICE, "topo_batch_compress: sort positions: device error bits 0x%x", hm[1] | sort_err);
    for (uint32_t j = 0; j < count; j++) {
      if (htc[j].error) return batch_device_error(B, "topo_batch_compress (topology)", j, htc[j].error);
      tszp_topo_info& info = fields[B->caller[j]].info;
      info.topo_bytes = kBatchTopoHdr + htc[j].payload_bytes;
      info.cp_count = htc[j].cp_count;
      info.ext_count = htc[j].ext_count;
      info.sortpos_bytes = 0;
    }
    // the sort-position codec's results, by its own table order
    for (uint32_t k = 0; k < count; k++)
      if (hic[k].error) return batch_device_error(B, "topo_batch_compress (sort positions)", B->ic.order[k], hic[k].error);
  }
  if (B->decompressed) {
    for (uint32_t k = 0; k < count; k++) {
      // decode results are indexed by each decode stage's table order
      if (htd[k].error) return batch_device_error(B, "topo_batch_decompress (topology)", B->td.order[k], htd[k].error);
      if (hid[k].error)
        return batch_device_error(B, "topo_batch_decompress (sort positions)", B->id.order[k], hid[k].error);
    }
  }
  if (B->compressed) {
    for (uint32_t k = 0; k < count; k++) {
      const uint32_t j = B->ic.order[k];
      tszp_topo_info& info = fields[B->caller[j]].info;
      info.sortpos_bytes = info.ext_count ? kBatchSpHdr + hic[k].payload_bytes : 0;
    }
  }
  if (B->restored)
    for (uint32_t j = 0; j < count; j++) fields[B->caller[j]].restore = B->rinfo[j];
  return TSZP_OK;
}

int tszp_topo_batch_output(tszp_topo_batch* B, uint32_t i, float** d_decoded, float** d_restored, uint8_t** d_types,
                           int32_t** d_sortpos) {
  clear_error();
  if (!B || i >= B->count) return bfail(TSZP_EINVAL, "topo_batch_output: bad arguments");
  uint32_t j = 0;
  while (j < B->count && B->caller[j] != i) j++;
  const uint64_t b = B->base[j];
  if (d_decoded) *d_decoded = B->dec + b;
  if (d_restored) *d_restored = B->rest + b;
  if (d_types) *d_types = B->types + b;
  if (d_sortpos) *d_sortpos = B->spv + b;
  return TSZP_OK;
}

int tszp_topo_batch_copy(tszp_topo_batch* B, uint32_t i, int which, void* d_dst, uint64_t count, void* stream) {
  clear_error();
  if (!B || i >= B->count || !d_dst || which < 0 || which > 3) return bfail(TSZP_EINVAL, "topo_batch_copy: bad arguments");
  uint32_t j = 0;
  while (j < B->count && B->caller[j] != i) j++;
  if (count > B->n[j]) return bfail(TSZP_EINVAL, "topo_batch_copy: %llu elements, the field has %llu",
                                    (unsigned long long)count, (unsigned long long)B->n[j]);
  const uint64_t b = B->base[j];
  const void* src = which == 0 ? (const void*)(B->dec + b) : which == 1 ? (const void*)(B->rest + b)
                  : which == 2 ? (const void*)(B->types + b) : (const void*)(B->spv + b);
  const size_t bytes = (size_t)count * (which == 2 ? 1 : 4);
  if (bytes) BHIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return TSZP_OK;
}

}  // extern "C"
