// tszp_binrank.hip -- sort positions without a full sort.
//
// The sort position of a critical point is the number of distinct critical-point
// values of its quantized bin that are smaller than its own value
// (SZp/src/szp_float.cc:137-509, SZp/src/szp_topology_3d.cc:135-247).  With
// order-preserving 32-bit value keys, the distinct values of one bin are the set
// bits of a bitmap over the bin's key span, and a sort position is a prefix
// popcount.  The path:
//   1. range: smallest and largest key and bin;
//   2. count: keys per (bin, slice) -- a slice is a contiguous run of keys owned by one
//      workgroup -- and the lowest key of every bin's span (binary search; the bin is
//      a monotone function of the key);
//   3. an exclusive scan of the bin-major (bin, slice) counts gives every slice its
//      own range inside every bin's bucket;
//   4. plan (one workgroup): bucket offsets, bitmap word and chunk offsets; bins whose
//      span exceeds kBinWideBits are "wide" and go to the radix sort;
//   5. scatter: each slice writes its keys into its ranges of the buckets through LDS
//      cursors, so no global atomics; neighbouring critical points have close values,
//      so the lanes of a wave fall into few bins and their stores come out as runs;
//   6. bitmap: one workgroup per chunk of kBinChunkWords words of a bin's bitmap; sets
//      the bits of the bin's keys in LDS, then stores {word, distinct keys before the
//      word in the chunk} and the chunk's total;
//   7. query: every critical point, in its original row-major order, reads its rank
//      from its bin's bitmap, so ranks are stored in extremum order.
// The order inside a bucket is not deterministic; ranks are.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

constexpr int kChunkShift = 19;  // key offsets per chunk: kBinChunkWords * 32
constexpr int kUnroll = 4;       // keys per thread and loop step (independent loads in flight)
constexpr int kVec = 2;          // uint4 key vectors per thread and loop step
constexpr int kQueryUnroll = 8;  // critical points per thread of the rank query (random reads in flight)

__device__ __forceinline__ uint32_t comp(const uint4& v, int c) {
  return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}
static_assert((1u << kChunkShift) == kBinChunkWords * 32, "chunk geometry");
static_assert(kBinMax % kBlockThreads == 0, "plan layout");
constexpr uint32_t kBinRangeParts = 2048;  // workgroups of the range pass
constexpr uint32_t kScanSt = (kBinMax * kBinSlices + 1 + kSortTile - 1) / kSortTile;  // scan status words

__device__ __forceinline__ int key_bin(uint32_t k, float eb, double inv) { return cp_bin(okey_value(k), eb, inv); }

size_t bin_meta_words() {
  return 8 + 2 * kBinRangeParts           // mm (padded), range partials
         + (kBinMax + 1) + (kBinMax + 1)  // klo, boff
         + 6 + 4 * (kBinMax + 1)          // alignment slack, bt
         + 2 * kBinMax * kBinMaxChunks    // cbin, ctot
         + 16                             // info
         + 2 + 2 * kScanSt + 2            // alignment slack, scan status, tile counter (+ pad)
         + kBinMax * kBinSlices + 1;      // hist
}

void bin_work_carve(uint32_t* meta, BinWork* w) {
  uint32_t* p = meta;
  w->mm = p;
  p += 8;
  w->part = p;
  p += 2 * kBinRangeParts;
  w->klo = p;
  p += kBinMax + 1;
  w->boff = p;
  p += kBinMax + 1;
  p += (4 - ((uintptr_t)p / 4) % 4) % 4;  // 16-byte alignment for uint4
  w->bt = reinterpret_cast<uint4*>(p);
  p += 4 * (kBinMax + 1);
  w->cbin = p;
  p += kBinMax * kBinMaxChunks;
  w->ctot = p;
  p += kBinMax * kBinMaxChunks;
  w->info = p;
  p += 16;
  p += ((uintptr_t)p / 4) % 2;  // 8-byte alignment
  w->scan_st = reinterpret_cast<uint64_t*>(p);
  p += 2 * kScanSt;
  w->scan_ctr = p;
  p += 2;
  w->hist = p;
}

// Slices are whole multiples of 1024 keys, so that they start on uint4 boundaries.
void bin_slices(uint32_t count, uint32_t* nslice, uint32_t* slice) {
  uint32_t ns = (count + 2047) / 2048;
  ns = ns < 1 ? 1 : (ns > kBinSlices ? kBinSlices : ns);
  const uint64_t s = (((uint64_t)count + ns - 1) / ns + 1023) / 1024 * 1024;
  *slice = (uint32_t)(s ? s : 1024);
  *nslice = count ? (uint32_t)(((uint64_t)count + *slice - 1) / *slice) : 1;
}

// ---- 1. smallest and largest key and bin (bins biased by 2^31: unsigned order) -------
// Each workgroup writes its smallest and largest key to part[2 * block]; k_bin_range_fin
// reduces them (same-address atomics from thousands of workgroups would serialize).
// Also clears the scan's status words, tile counter and error word.
__global__ __launch_bounds__(kBlockThreads) void k_bin_range(const uint32_t* __restrict__ key, uint32_t P,
                                                            uint32_t* __restrict__ part,
                                                            uint64_t* __restrict__ scan_st,
                                                            uint32_t* __restrict__ scan_ctr,
                                                            uint32_t* __restrict__ scan_err) {
  __shared__ uint32_t red[2][kWavesPerBlock];
  const uint32_t gid = blockIdx.x * kBlockThreads + threadIdx.x;
  for (uint32_t z = gid; z < kScanSt; z += gridDim.x * kBlockThreads) scan_st[z] = 0;
  if (gid == 0) {
    *scan_ctr = 0;
    *scan_err = 0;
  }
  uint32_t mn = 0xffffffffu, mx = 0u;
  const uint4* k4 = reinterpret_cast<const uint4*>(key);
  const uint32_t nv = (P + 3) / 4;
  for (uint32_t j = gid; j < nv; j += gridDim.x * kBlockThreads) {
    const uint4 v = k4[j];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const uint32_t k = comp(v, c);
      if (4 * j + c < P) {
        mn = k < mn ? k : mn;
        mx = k > mx ? k : mx;
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t a = __shfl_xor(mn, o, kWave), c = __shfl_xor(mx, o, kWave);
    mn = a < mn ? a : mn;
    mx = c > mx ? c : mx;
  }
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[0][w] = mn;
    red[1][w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
      mn = red[0][k] < mn ? red[0][k] : mn;
      mx = red[1][k] > mx ? red[1][k] : mx;
    }
    part[2 * blockIdx.x] = mn;
    part[2 * blockIdx.x + 1] = mx;
  }
}

// mm: min key, ~max key, min bin ^ 2^31, ~(max bin ^ 2^31) (one workgroup)
__global__ __launch_bounds__(kBlockThreads) void k_bin_range_fin(const uint32_t* __restrict__ part, uint32_t nparts,
                                                                float eb, double inv, uint32_t* __restrict__ mm) {
  __shared__ uint32_t red[2][kWavesPerBlock];
  uint32_t mn = 0xffffffffu, mx = 0u;
  for (uint32_t k = threadIdx.x; k < nparts; k += kBlockThreads) {
    const uint32_t a = part[2 * k], c = part[2 * k + 1];
    mn = a < mn ? a : mn;
    mx = c > mx ? c : mx;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t a = __shfl_xor(mn, o, kWave), c = __shfl_xor(mx, o, kWave);
    mn = a < mn ? a : mn;
    mx = c > mx ? c : mx;
  }
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[0][w] = mn;
    red[1][w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
      mn = red[0][k] < mn ? red[0][k] : mn;
      mx = red[1][k] > mx ? red[1][k] : mx;
    }
    // the bin is monotone in the key, so the extreme keys give the extreme bins
    mm[0] = mn;
    mm[1] = ~mx;
    mm[2] = (uint32_t)key_bin(mn, eb, inv) ^ 0x80000000u;
    mm[3] = ~((uint32_t)key_bin(mx, eb, inv) ^ 0x80000000u);
  }
}

// ---- 2. keys per (bin, slice); lowest key of each bin's span -------------------------
__global__ __launch_bounds__(kBlockThreads) void k_bin_count(const uint32_t* __restrict__ key, uint32_t P, float eb,
                                                            double inv, int bmin, uint32_t B, uint32_t slice,
                                                            const uint32_t* __restrict__ mm,
                                                            uint32_t* __restrict__ klo, uint32_t* __restrict__ hist) {
  __shared__ uint32_t lc[kBinMax];
  const uint32_t nslice = gridDim.x;
  for (uint32_t g = blockIdx.x * kBlockThreads + threadIdx.x; g <= B; g += nslice * kBlockThreads) {
    // smallest key in [min key, max key + 1] whose bin is >= bmin + g
    const int target = bmin + (int)g;
    uint64_t lo = mm[0], hi = (uint64_t)(~mm[1]) + 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (key_bin((uint32_t)mid, eb, inv) < target)
        lo = mid + 1;
      else
        hi = mid;
    }
    klo[g] = (uint32_t)lo;
  }
  for (uint32_t b = threadIdx.x; b < B; b += kBlockThreads) lc[b] = 0;
  __syncthreads();
  const uint32_t s = blockIdx.x * slice, e = min(s + slice, P);
  const uint4* k4 = reinterpret_cast<const uint4*>(key);
  for (uint32_t j0 = s / 4 + threadIdx.x; 4ull * j0 < e; j0 += kBlockThreads * kVec) {
    uint4 v[kVec];
#pragma unroll
    for (int u = 0; u < kVec; u++) {
      const uint32_t j = j0 + u * kBlockThreads;
      v[u] = 4ull * j < e ? k4[j] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < kVec; u++) {
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const uint64_t i = 4ull * (j0 + u * kBlockThreads) + c;
        const uint32_t b = (uint32_t)(key_bin(comp(v[u], c), eb, inv) - bmin);
        if (i < e && b < B) atomicAdd(&lc[b], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < B; b += kBlockThreads) hist[(size_t)b * nslice + blockIdx.x] = lc[b];
  if (blockIdx.x == 0 && threadIdx.x == 0) hist[(size_t)B * nslice] = 0;  // becomes the total
}

// ---- 4. plan (one workgroup) -----------------------------------------------------
// info: [0] bitmap words, [1] chunks, [2] first wide bin, [3] last wide bin,
//       [4] bucket start of the wide range, [5] its end, [6] keys counted
__global__ __launch_bounds__(kBlockThreads) void k_bin_plan(uint32_t B, uint32_t nslice,
                                                           const uint32_t* __restrict__ hist,
                                                           const uint32_t* __restrict__ klo,
                                                           uint32_t* __restrict__ boff, uint4* __restrict__ bt,
                                                           uint32_t* __restrict__ cbin, uint32_t* info) {
  __shared__ uint32_t lds4[kWavesPerBlock];
  __shared__ uint32_t sh_wlo, sh_whi;
  constexpr uint32_t PER = kBinMax / kBlockThreads;
  constexpr uint32_t kWideMark = 0xffffffffu;
  if (threadIdx.x == 0) {
    sh_wlo = 0xffffffffu;
    sh_whi = 0;
  }
  __syncthreads();
  const uint32_t b0 = threadIdx.x * PER;
  uint32_t nw[PER], nch[PER];
  uint32_t ws = 0, chs = 0;
#pragma unroll
  for (uint32_t k = 0; k < PER; k++) {
    const uint32_t b = b0 + k;
    nw[k] = nch[k] = 0;
    if (b < B) {
      const uint32_t c = hist[(size_t)(b + 1) * nslice] - hist[(size_t)b * nslice];
      const uint64_t span = (uint64_t)klo[b + 1] - klo[b];  // keys klo[b] .. klo[b+1]-1
      if (c && span > kBinWideBits) {
        atomicMin(&sh_wlo, b);
        atomicMax(&sh_whi, b);
        nch[k] = kWideMark;
      } else if (c) {
        nw[k] = (uint32_t)((span + 31) >> 5);
        nch[k] = (nw[k] + kBinChunkWords - 1) / kBinChunkWords;
      }
    }
    ws += nw[k];
    chs += nch[k] == kWideMark ? 0u : nch[k];
  }
  uint32_t tw, tch;
  uint32_t run_w = block_excl_scan(ws, lds4, &tw);
  uint32_t run_ch = block_excl_scan(chs, lds4, &tch);
  const uint32_t wlo = sh_wlo, whi = sh_whi;  // block_excl_scan ended with a barrier
#pragma unroll
  for (uint32_t k = 0; k < PER; k++) {
    const uint32_t b = b0 + k;
    if (b < B) {
      const bool wide = nch[k] == kWideMark;
      const uint32_t start = hist[(size_t)b * nslice];
      boff[b] = start;
      bt[b] = make_uint4(klo[b], run_w, run_ch, wide ? 1u : 0u);
      if (b == wlo) info[4] = start;
      if (b == whi) info[5] = hist[(size_t)(b + 1) * nslice];
      if (!wide) {
        for (uint32_t j = 0; j < nch[k]; j++) cbin[run_ch + j] = b;
        run_ch += nch[k];
      }
    }
    run_w += nw[k];
  }
  if (threadIdx.x == 0) {
    const uint32_t total = hist[(size_t)B * nslice];
    boff[B] = total;
    bt[B] = make_uint4(klo[B], tw, tch, 0u);
    info[0] = tw;
    info[1] = tch;
    info[2] = wlo;
    info[3] = whi;
    if (wlo == 0xffffffffu) info[4] = info[5] = 0;
    info[6] = total;
  }
}

// ---- 5. scatter keys into buckets ---------------------------------------------------
constexpr uint32_t kScatterChunk = 8192;  // keys per workgroup and step, ordered by bin in LDS
__global__ __launch_bounds__(kBlockThreads) void k_bin_scatter(const uint32_t* __restrict__ key,
                                                              const uint32_t* __restrict__ pay, uint32_t P, float eb,
                                                              double inv, int bmin, uint32_t B, uint32_t slice,
                                                              const uint32_t* __restrict__ hist,
                                                              const uint32_t* __restrict__ info,
                                                              uint32_t* __restrict__ bkey,
                                                              uint32_t* __restrict__ bpay) {
  __shared__ uint32_t cur[kBinMax];      // next slot of this slice in each bin's bucket
  __shared__ uint32_t run[kBinMax + 1];  // per step: keys of each bin, then where its run starts
  __shared__ uint32_t skey[kScatterChunk];
  __shared__ uint32_t lds4[kWavesPerBlock];
  const int tid = threadIdx.x;
  const uint32_t nslice = gridDim.x;
  for (uint32_t b = tid; b < B; b += kBlockThreads) {
    cur[b] = hist[(size_t)b * nslice + blockIdx.x];
    run[b] = 0;
  }
  const uint32_t wlo = info[2], whi = info[3];
  // the bins this thread scans and advances
  const uint32_t PB = (B + kBlockThreads - 1) / kBlockThreads;
  const uint32_t b0 = min(tid * PB, B), b1 = min(b0 + PB, B);
  __syncthreads();
  const uint32_t s = blockIdx.x * slice, e = min(s + slice, P);
  constexpr int U = kScatterChunk / kBlockThreads;
  constexpr uint32_t kSkip = 0xffffffffu;
  for (uint32_t c0 = s; c0 < e; c0 += kScatterChunk) {
    uint32_t kv[U], bv[U], rk[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t i = c0 + u * kBlockThreads + tid;
      kv[u] = i < e ? key[i] : 0u;
    }
    // 1. each key's bin and its rank among the step's keys of that bin; keys of the wide
    //    bins (a few, for the radix path, which needs their payloads) go straight out
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t i = c0 + u * kBlockThreads + tid;
      uint32_t b = (uint32_t)(key_bin(kv[u], eb, inv) - bmin);
      if (i >= e || b >= B) b = kSkip;
      if (b != kSkip && b >= wlo && b <= whi) {
        const uint32_t pos = atomicAdd(&cur[b], 1u);
        if (pos < P) {
          bkey[pos] = kv[u];
          bpay[pos] = pay[i];
        }
        b = kSkip;
      }
      bv[u] = b;
      rk[u] = b != kSkip ? atomicAdd(&run[b], 1u) : 0u;
    }
    __syncthreads();
    // 2. counts -> run starts (exclusive scan over the bins)
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; b++) sum += run[b];
    uint32_t tot;
    uint32_t r = block_excl_scan(sum, lds4, &tot);  // ends with a barrier
    for (uint32_t b = b0; b < b1; b++) {
      const uint32_t c = run[b];
      run[b] = r;
      r += c;
    }
    if (tid == 0) run[B] = tot;
    __syncthreads();
    // 3. the step's keys in bin order
#pragma unroll
    for (int u = 0; u < U; u++)
      if (bv[u] != kSkip) skey[run[bv[u]] + rk[u]] = kv[u];
    __syncthreads();
    // 4. consecutive lanes store consecutive slots of each bin's run
    for (uint32_t j = tid; j < tot; j += kBlockThreads) {
      const uint32_t k = skey[j];
      const uint32_t b = (uint32_t)(key_bin(k, eb, inv) - bmin);
      const uint32_t pos = cur[b] + (j - run[b]);
      if (pos < P) bkey[pos] = k;
    }
    __syncthreads();
    // 5. advance the cursors by the step's counts, then clear the counts
    for (uint32_t b = b0; b < b1; b++) cur[b] += run[b + 1] - run[b];
    __syncthreads();
    for (uint32_t b = b0; b < b1; b++) run[b] = 0;
    __syncthreads();
  }
}

// ---- 6. per-chunk bitmaps -----------------------------------------------------------
constexpr int kBitmapThreads = 1024;  // the largest bins hold a few hundred thousand keys

// exclusive scan of one value per thread over a kBitmapThreads workgroup
__device__ __forceinline__ uint32_t bitmap_excl_scan(uint32_t v, uint32_t* ldsw, uint32_t* total) {
  constexpr int NW = kBitmapThreads / kWave;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += t;
  }
  if (lane == kWave - 1) ldsw[w] = incl;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < NW; k++) {
    const uint32_t t = ldsw[k];
    pre += k < w ? t : 0u;
    tot += t;
  }
  *total = tot;
  return pre + incl - v;
}

__global__ __launch_bounds__(kBitmapThreads) void k_bin_bitmap(const uint32_t* __restrict__ bkey,
                                                             const uint32_t* __restrict__ boff,
                                                             const uint4* __restrict__ bt,
                                                             const uint32_t* __restrict__ cbin,
                                                             uint32_t* __restrict__ ctot, uint2* __restrict__ gw) {
  __shared__ uint32_t bm[kBinChunkWords];
  __shared__ uint32_t ldsw[kBitmapThreads / kWave];
  const uint32_t c = blockIdx.x, b = cbin[c];
  const uint4 m = bt[b];
  const uint32_t ci = c - m.z;
  const uint32_t nwb = bt[b + 1].y - m.y;  // words of the bin
  const uint32_t w0 = ci * kBinChunkWords;
  const uint32_t cw = min(kBinChunkWords, nwb - w0);
  for (uint32_t i = threadIdx.x; i < cw; i += kBitmapThreads) bm[i] = 0;
  __syncthreads();
  const uint32_t s = boff[b], e = boff[b + 1];
  constexpr int U = kUnroll;
  for (uint32_t i0 = s + threadIdx.x; i0 < e; i0 += kBitmapThreads * U) {
    uint32_t off[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t i = i0 + u * kBitmapThreads;
      off[u] = i < e ? bkey[i] - m.x : 0xffffffffu;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if ((off[u] >> kChunkShift) == ci) {
        const uint32_t wi = (off[u] >> 5) & (kBinChunkWords - 1);
        if (wi < cw) atomicOr(&bm[wi], 1u << (off[u] & 31u));
      }
    }
  }
  __syncthreads();
  // exclusive prefix of popcounts over the chunk's words: each thread a contiguous run
  const uint32_t per = (cw + kBitmapThreads - 1) / kBitmapThreads;
  const uint32_t a0 = min(threadIdx.x * per, cw), a1 = min(a0 + per, cw);
  uint32_t n = 0;
  for (uint32_t i = a0; i < a1; i++) n += (uint32_t)__popc(bm[i]);
  uint32_t tot;
  uint32_t run = bitmap_excl_scan(n, ldsw, &tot);
  uint2* out = gw + m.y + w0;
  for (uint32_t i = a0; i < a1; i++) {
    const uint32_t v = bm[i];
    out[i] = make_uint2(v, run);
    run += (uint32_t)__popc(v);
  }
  if (threadIdx.x == 0) ctot[c] = tot;
}

// ---- 7. ranks in the original order ---------------------------------------------------
// chunk totals -> exclusive prefix over each bin's chunks, in place (one thread per
// bin; a bin has at most kBinMaxChunks chunks), so that a rank query adds one word
// instead of walking the chunks before its own
__global__ __launch_bounds__(kBlockThreads) void k_bin_chunk_scan(const uint4* __restrict__ bt, uint32_t B,
                                                                 uint32_t* __restrict__ ctot) {
  const uint32_t b = blockIdx.x * kBlockThreads + threadIdx.x;
  if (b >= B) return;
  const uint4 m = bt[b];
  if (m.w) return;  // wide bins: ranked by the radix path
  const uint32_t nch = (bt[b + 1].y - m.y + kBinChunkWords - 1) / kBinChunkWords;
  uint32_t run = 0;
  for (uint32_t j = 0; j < nch; j++) {
    const uint32_t v = ctot[m.z + j];
    ctot[m.z + j] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kBlockThreads) void k_bin_query(const uint32_t* __restrict__ key,
                                                            const uint32_t* __restrict__ pay, uint32_t P, float eb,
                                                            double inv, int bmin, uint32_t B,
                                                            const uint4* __restrict__ bt,
                                                            const uint32_t* __restrict__ ctot,
                                                            const uint2* __restrict__ gw, int32_t* __restrict__ out,
                                                            uint32_t out_cap) {
  // consecutive lanes take consecutive critical points: their bitmap words are close.
  // The records are read once (non-temporal: the L2 keeps bitmap words instead).
  constexpr int QU = kQueryUnroll;
  const uint32_t base = blockIdx.x * (kBlockThreads * QU) + threadIdx.x;
  uint32_t pv[QU], kv[QU];
#pragma unroll
  for (int u = 0; u < QU; u++) {
    const uint32_t i = base + u * kBlockThreads;
    pv[u] = i < P ? __builtin_nontemporal_load(&pay[i]) : 0xffffffffu;
    kv[u] = i < P ? __builtin_nontemporal_load(&key[i]) : 0u;
  }
  uint4 m[QU];
  uint32_t off[QU];
  bool live[QU];
#pragma unroll
  for (int u = 0; u < QU; u++) {
    const uint32_t b = (uint32_t)(key_bin(kv[u], eb, inv) - bmin);
    live[u] = pv[u] < out_cap && b < B;  // saddles carry 0xffffffff
    m[u] = live[u] ? bt[b] : make_uint4(0, 0, 0, 1);
    live[u] = live[u] && !m[u].w;  // wide bins: ranked by the radix path
    off[u] = kv[u] - m[u].x;
  }
  uint2 v[QU];
#pragma unroll
  for (int u = 0; u < QU; u++) v[u] = live[u] ? gw[m[u].y + (off[u] >> 5)] : make_uint2(0, 0);
#pragma unroll
  for (int u = 0; u < QU; u++) {
    if (!live[u]) continue;
    // ctot holds, after k_bin_chunk_scan, the distinct keys of the bin's earlier chunks
    const uint32_t ci = off[u] >> kChunkShift;
    const uint32_t r = v[u].y + (uint32_t)__popc(v[u].x & ((1u << (off[u] & 31u)) - 1u)) + ctot[m[u].z + ci];
    __builtin_nontemporal_store((int32_t)r, &out[pv[u]]);
  }
}

// ---------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------
hipError_t bin_range(const uint32_t* key, uint32_t P, float eb, double inv, const BinWork& w, hipStream_t st) {
  uint64_t g = ((uint64_t)P / 4 + kBlockThreads - 1) / kBlockThreads;
  g = g < 4 ? 4 : (g > kBinRangeParts ? kBinRangeParts : g);
  hipLaunchKernelGGL(k_bin_range, dim3((unsigned)g), dim3(kBlockThreads), 0, st, key, P, w.part, w.scan_st,
                     w.scan_ctr, w.info + 8);
  hipLaunchKernelGGL(k_bin_range_fin, dim3(1), dim3(kBlockThreads), 0, st, w.part, (uint32_t)g, eb, inv, w.mm);
  return hipGetLastError();
}

hipError_t bin_count(const uint32_t* key, uint32_t P, float eb, double inv, int bmin, uint32_t B, const BinWork& w,
                     hipStream_t st) {
  uint32_t ns, sl;
  bin_slices(P, &ns, &sl);
  hipLaunchKernelGGL(k_bin_count, dim3(ns), dim3(kBlockThreads), 0, st, key, P, eb, inv, bmin, B, sl, w.mm, w.klo,
                     w.hist);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_scan_excl(w.hist, B * ns + 1, w.scan_st, w.scan_ctr, w.info + 8, st);
}

hipError_t bin_plan(uint32_t B, uint32_t nslice, const BinWork& w, hipStream_t st) {
  hipLaunchKernelGGL(k_bin_plan, dim3(1), dim3(kBlockThreads), 0, st, B, nslice, w.hist, w.klo, w.boff, w.bt, w.cbin,
                     w.info);
  return hipGetLastError();
}

hipError_t bin_scatter(const uint32_t* key, const uint32_t* pay, uint32_t P, float eb, double inv, int bmin,
                       uint32_t B, const BinWork& w, hipStream_t st) {
  uint32_t ns, sl;
  bin_slices(P, &ns, &sl);
  hipLaunchKernelGGL(k_bin_scatter, dim3(ns), dim3(kBlockThreads), 0, st, key, pay, P, eb, inv, bmin, B, sl, w.hist,
                     w.info, w.bkey, w.bpay);
  return hipGetLastError();
}

hipError_t bin_bitmap(uint32_t nchunks, const BinWork& w, uint2* gw, hipStream_t st) {
  if (!nchunks) return hipSuccess;
  hipLaunchKernelGGL(k_bin_bitmap, dim3(nchunks), dim3(kBitmapThreads), 0, st, w.bkey, w.boff, w.bt, w.cbin, w.ctot,
                     gw);
  return hipGetLastError();
}

hipError_t bin_chunk_scan(uint32_t B, const BinWork& w, hipStream_t st) {
  if (!B) return hipSuccess;
  hipLaunchKernelGGL(k_bin_chunk_scan, dim3((B + kBlockThreads - 1) / kBlockThreads), dim3(kBlockThreads), 0, st,
                     w.bt, B, w.ctot);
  return hipGetLastError();
}

hipError_t bin_query(const uint32_t* key, const uint32_t* pay, uint32_t P, float eb, double inv, int bmin, uint32_t B,
                     const BinWork& w, const uint2* gw, int32_t* out, uint32_t out_cap, hipStream_t st) {
  if (!P) return hipSuccess;
  const uint32_t per = kBlockThreads * kQueryUnroll;
  hipLaunchKernelGGL(k_bin_query, dim3((P + per - 1) / per), dim3(kBlockThreads), 0, st, key, pay, P, eb, inv, bmin,
                     B, w.bt, w.ctot, gw, out, out_cap);
  return hipGetLastError();
}

}  // namespace tszp
