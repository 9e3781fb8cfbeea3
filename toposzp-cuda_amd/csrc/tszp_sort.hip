// tszp_sort.hip -- critical-point sort positions on the GPU.
//
// The reference sorts the critical points of every quantized bin by their original
// value and gives each the number of distinct smaller values in its bin
// (SZp/src/szp_float.cc:137-509, SZp/src/szp_topology_3d.cc:135-247).  Because the
// bin is a monotone function of the value, sorting all critical points by an
// order-preserving value key orders them by (bin, value) at once.  The sort is a
// stable LSD radix sort with 8-bit digits (histogram -> look-back scan -> stable
// scatter ranked by wave ballots), and the ranks come from one segmented
// look-back scan over the sorted keys.  Everything is deterministic.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

constexpr int kItemsPerThread = kSortTile / kBlockThreads;  // 16
static_assert(kItemsPerThread == 16, "sort tile layout");

static inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

size_t sort_hist_words(uint32_t count) { return (size_t)256 * ceil_div(count, kSortTile); }

size_t sort_status_words(uint32_t count) {
  const uint32_t t = ceil_div(count, kSortTile);
  return (size_t)8 * ceil_div((uint64_t)256 * t, kSortTile) + t;
}

// ---- per-tile digit histogram, written digit-major: hist[d * ntiles + tile] ----------
// dcount (nullable): the item count on the device (<= count); the grid covers `count`
__device__ __forceinline__ uint32_t dev_count(uint32_t count, const uint32_t* dcount) {
  if (!dcount) return count;
  const uint32_t c = *dcount;
  return c < count ? c : count;
}

__global__ __launch_bounds__(kBlockThreads) void k_hist(const uint32_t* __restrict__ words, uint32_t count,
                                                       int shift, uint32_t* __restrict__ hist, uint32_t ntiles,
                                                       const uint32_t* __restrict__ dcount) {
  __shared__ uint32_t h[256];
  const int tid = threadIdx.x;
  count = dev_count(count, dcount);
  h[tid] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
  for (int i = 0; i < kItemsPerThread; i++) {
    const uint32_t idx = base + i * kBlockThreads + tid;
    if (idx < count) atomicAdd(&h[(words[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)tid * ntiles + blockIdx.x] = h[tid];
}

// ---- in-place exclusive scan of a u32 array (single pass, decoupled look-back) -------
__global__ __launch_bounds__(kBlockThreads) void k_scan_excl(uint32_t* a, uint32_t count, uint64_t* st,
                                                            uint32_t* counter, uint32_t* err) {
  __shared__ uint32_t lds4[kWavesPerBlock];
  __shared__ uint32_t sh_tile;
  __shared__ uint64_t sh_pre;
  const int tid = threadIdx.x;
  if (tid == 0) sh_tile = tile_id(counter);
  __syncthreads();
  const uint32_t tile = sh_tile;
  const uint32_t base = tile * kSortTile + tid * kItemsPerThread;
  uint32_t v[kItemsPerThread];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kItemsPerThread; k++) {
    v[k] = base + k < count ? a[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t pre = block_excl_scan(sum, lds4, &total);
  if (tid < kWave) {
    uint64_t ex = lookback(st, tile, total, err);
    if (tid == 0) sh_pre = ex;
  }
  __syncthreads();
  uint32_t run = (uint32_t)sh_pre + pre;
#pragma unroll
  for (int k = 0; k < kItemsPerThread; k++) {
    if (base + k < count) a[base + k] = run;
    run += v[k];
  }
}

// ---- stable scatter: rank = (scanned tile base for the digit) + (earlier items of
//      the tile with the same digit); items are visited in index order (wave, step,
//      lane).  The tile is first reordered by digit in LDS, so that the global writes
//      go out as contiguous runs of one digit instead of single scattered words.
template <bool HI>
__global__ __launch_bounds__(kBlockThreads) void k_scatter(const uint32_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ hin,
                                                          const uint32_t* __restrict__ pin,
                                                          uint32_t* __restrict__ kout, uint32_t* __restrict__ hout,
                                                          uint32_t* __restrict__ pout, uint32_t count, int shift,
                                                          int digit_from_hi, const uint32_t* __restrict__ hist,
                                                          uint32_t ntiles, const uint32_t* __restrict__ dcount) {
  constexpr int NS = kWavesPerBlock * kItemsPerThread;  // 64 (wave, step) groups
  __shared__ uint16_t cnt[NS][256];
  __shared__ uint32_t gbase[256], lstart[256];
  __shared__ uint32_t lds4[kWavesPerBlock];
  __shared__ uint32_t skey[kSortTile], spay[kSortTile];
  __shared__ uint32_t shi[HI ? kSortTile : 1];
  const int tid = threadIdx.x, w = tid / kWave, lane = tid & (kWave - 1);
  count = dev_count(count, dcount);
  if (blockIdx.x * kSortTile >= count) return;  // past the items counted on the device
  uint32_t* cz = reinterpret_cast<uint32_t*>(&cnt[0][0]);
  for (int i = tid; i < NS * 256 / 2; i += kBlockThreads) cz[i] = 0;
  gbase[tid] = hist[(size_t)tid * ntiles + blockIdx.x];
  __syncthreads();

  const uint32_t tile0 = blockIdx.x * kSortTile;
  const uint32_t base = tile0 + w * (kWave * kItemsPerThread);
  const uint64_t lt = lanemask_lt(lane);
  uint32_t kv[kItemsPerThread], pv[kItemsPerThread], hv[kItemsPerThread];
  uint32_t digit[kItemsPerThread];
  uint8_t rnk[kItemsPerThread];
#pragma unroll
  for (int s = 0; s < kItemsPerThread; s++) {
    const uint32_t idx = base + s * kWave + lane;
    const bool valid = idx < count;
    kv[s] = valid ? kin[idx] : 0u;
    pv[s] = valid ? pin[idx] : 0u;
    hv[s] = (HI && valid) ? hin[idx] : 0u;
    const uint32_t d = ((digit_from_hi ? hv[s] : kv[s]) >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t r = (uint32_t)__popcll(peers & lt);
    if (valid && r == 0) cnt[w * kItemsPerThread + s][d] = (uint16_t)__popcll(peers);
    digit[s] = d;
    rnk[s] = (uint8_t)r;
  }
  __syncthreads();
  uint32_t run = 0;
  for (int k = 0; k < NS; k++) {  // exclusive prefix over the (wave, step) groups of digit tid
    const uint32_t c = cnt[k][tid];
    cnt[k][tid] = (uint16_t)run;
    run += c;
  }
  uint32_t total;
  lstart[tid] = block_excl_scan(run, lds4, &total);  // tile-local start of digit tid
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kItemsPerThread; s++) {
    const uint32_t idx = base + s * kWave + lane;
    if (idx >= count) continue;
    const uint32_t d = digit[s];
    const uint32_t loc = lstart[d] + cnt[w * kItemsPerThread + s][d] + rnk[s];
    skey[loc] = kv[s];
    spay[loc] = pv[s];
    if (HI) shi[loc] = hv[s];
  }
  __syncthreads();
  const uint32_t n_tile = count - tile0 < kSortTile ? count - tile0 : kSortTile;
  for (uint32_t i = tid; i < n_tile; i += kBlockThreads) {
    const uint32_t kk = skey[i];
    const uint32_t hh = HI ? shi[i] : 0u;
    const uint32_t d = ((digit_from_hi ? hh : kk) >> shift) & 255u;
    const uint32_t pos = gbase[d] + (i - lstart[d]);
    if (pos >= count) continue;  // cannot happen for a consistent histogram
    kout[pos] = kk;
    pout[pos] = spay[i];
    if (HI) hout[pos] = hh;
  }
}

size_t scan_status_words(uint32_t count) { return ceil_div(count, kSortTile); }

// back from the alternate buffers after an odd number of passes
__global__ __launch_bounds__(kBlockThreads) void k_copy3(const uint32_t* __restrict__ k, const uint32_t* __restrict__ h,
                                                        const uint32_t* __restrict__ p, uint32_t* __restrict__ ko,
                                                        uint32_t* __restrict__ ho, uint32_t* __restrict__ po,
                                                        uint32_t count, const uint32_t* __restrict__ dcount) {
  count = dev_count(count, dcount);
  for (uint32_t i = blockIdx.x * kSortTile + threadIdx.x; i < count && i < (blockIdx.x + 1) * kSortTile;
       i += kBlockThreads) {
    ko[i] = k[i];
    po[i] = p[i];
    if (h) ho[i] = h[i];
  }
}

hipError_t launch_scan_excl(uint32_t* a, uint32_t count, uint64_t* status, uint32_t* counter, uint32_t* err,
                            hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_scan_excl, dim3(ceil_div(count, kSortTile)), dim3(kBlockThreads), 0, st, a, count, status,
                     dyn_counter(counter), err);
  return hipGetLastError();
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* key_hi, uint32_t* pay, uint32_t count,
                            const SortWork& wk, hipStream_t st, LaunchTimer* scatter_timer, const uint32_t* dcount,
                            int hi_bits) {
  if (count <= 1 && !dcount) return hipSuccess;
  if (count == 0) return hipSuccess;
  const uint32_t ntiles = ceil_div(count, kSortTile);
  const uint32_t hwords = 256u * ntiles;
  const uint32_t stiles = ceil_div(hwords, kSortTile);
  uint32_t *k0 = keys, *k1 = wk.keys_alt, *h0 = key_hi, *h1 = key_hi ? wk.hi_alt : nullptr;
  uint32_t *p0 = pay, *p1 = wk.pay_alt;
  const int words = key_hi ? 2 : 1;
  int pass = 0;
  for (int wd = 0; wd < words; wd++) {
    const int bits = wd ? hi_bits : 32;  // passes over the significant bits of the word
    for (int shift = 0; shift < bits; shift += 8, pass++) {
      const uint32_t* dw = wd ? h0 : k0;
      hipLaunchKernelGGL(k_hist, dim3(ntiles), dim3(kBlockThreads), 0, st, dw, count, shift, wk.hist, ntiles,
                         dcount);
      hipLaunchKernelGGL(k_scan_excl, dim3(stiles), dim3(kBlockThreads), 0, st, wk.hist, hwords,
                         wk.st + (size_t)pass * stiles, dyn_counter(wk.counters + pass), wk.err);
      if (scatter_timer) scatter_timer->begin(st);
      if (h0)
        hipLaunchKernelGGL(k_scatter<true>, dim3(ntiles), dim3(kBlockThreads), 0, st, k0, h0, p0, k1, h1, p1, count,
                           shift, wd, wk.hist, ntiles, dcount);
      else
        hipLaunchKernelGGL(k_scatter<false>, dim3(ntiles), dim3(kBlockThreads), 0, st, k0, h0, p0, k1, h1, p1,
                           count, shift, wd, wk.hist, ntiles, dcount);
      if (scatter_timer) scatter_timer->end(st);
      std::swap(k0, k1);
      std::swap(h0, h1);
      std::swap(p0, p1);
    }
  }
  if (pass & 1) {  // an odd number of passes leaves the result in the alternate buffers
    hipLaunchKernelGGL(k_copy3, dim3(ntiles), dim3(kBlockThreads), 0, st, k0, h0, p0, keys, key_hi, pay, count, dcount);
  }
  return hipGetLastError();
}

// ---- dense rank within bins (segmented look-back scan over sorted keys) ---------------
// Whether items i - 1 and i lie in different bins: by value (the bin of each value with
// one eb), by key_hi (a biased bin), or, with field bounds, key_hi is the field and the
// bin comes from the value with that field's error bound.
__device__ __forceinline__ bool new_bin(const uint32_t* keys, const uint32_t* key_hi, uint32_t i, float eb, double inv,
                                        const float* feb, const double* finv) {
  if (feb) {
    const uint32_t f = key_hi[i];
    if (f != key_hi[i - 1]) return true;
    const float e = feb[f];
    const double v = finv[f];
    return cp_bin(okey_value(keys[i]), e, v) != cp_bin(okey_value(keys[i - 1]), e, v);
  }
  if (key_hi) return key_hi[i] != key_hi[i - 1];
  return cp_bin(okey_value(keys[i]), eb, inv) != cp_bin(okey_value(keys[i - 1]), eb, inv);
}

__global__ __launch_bounds__(kBlockThreads) void k_rank(const uint32_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ key_hi,
                                                       const uint32_t* __restrict__ pay, uint32_t count, float eb,
                                                       double inv, int32_t* __restrict__ out, uint32_t out_cap,
                                                       uint64_t* st, uint32_t* counter, uint32_t* err,
                                                       const uint32_t* __restrict__ dcount,
                                                       const float* __restrict__ feb, const double* __restrict__ finv) {
  constexpr int NS = kWavesPerBlock * kItemsPerThread;
  __shared__ uint32_t sf[NS], sv[NS];
  __shared__ uint32_t sh_tile;
  __shared__ uint64_t sh_pre;
  const int tid = threadIdx.x, w = tid / kWave, lane = tid & (kWave - 1);
  count = dev_count(count, dcount);
  if (tid == 0) sh_tile = tile_id(counter);
  __syncthreads();
  const uint32_t tile = sh_tile;
  if (tile * kSortTile >= count) return;  // past the items counted on the device
  const uint32_t base = tile * kSortTile + w * (kWave * kItemsPerThread);

  uint32_t F[kItemsPerThread], V[kItemsPerThread];
#pragma unroll
  for (int s = 0; s < kItemsPerThread; s++) {
    const uint32_t i = base + s * kWave + lane;
    uint32_t f = 0, v = 0;
    if (i < count) {
      if (i == 0) {
        f = 1;
      } else {
        f = new_bin(keys, key_hi, i, eb, inv, feb, finv);
        v = f ? 0u : (keys[i] != keys[i - 1]);
      }
    }
    // segmented inclusive scan across the wave: (fa, va) o (fb, vb) = (fa|fb, fb ? vb : va+vb)
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t fo = __shfl_up(f, o, kWave), vo = __shfl_up(v, o, kWave);
      if (lane >= o) {
        if (!f) v += vo;
        f |= fo;
      }
    }
    F[s] = f;
    V[s] = v;
    if (lane == kWave - 1) {
      sf[w * kItemsPerThread + s] = f;
      sv[w * kItemsPerThread + s] = v;
    }
  }
  __syncthreads();
  if (w == 0) {
    // exclusive segmented prefix over the 64 (wave, step) groups, one per lane
    uint32_t f = sf[lane], v = sv[lane];
    uint32_t fi = f, vi = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t fo = __shfl_up(fi, o, kWave), vo = __shfl_up(vi, o, kWave);
      if (lane >= o) {
        if (!fi) vi += vo;
        fi |= fo;
      }
    }
    uint32_t fx = __shfl_up(fi, 1, kWave), vx = __shfl_up(vi, 1, kWave);
    if (lane == 0) fx = vx = 0;
    const uint32_t ft = __shfl(fi, kWave - 1, kWave), vt = __shfl(vi, kWave - 1, kWave);
    sf[lane] = fx;
    sv[lane] = vx;
    uint64_t agg = (uint64_t)vt | (ft ? kSegReset : 0ull);
    uint64_t ex = lookback_seg(st, tile, agg, err);
    if (lane == 0) sh_pre = ex;
  }
  __syncthreads();
  const uint32_t pre = (uint32_t)sh_pre;
#pragma unroll
  for (int s = 0; s < kItemsPerThread; s++) {
    const uint32_t i = base + s * kWave + lane;
    if (i >= count) continue;
    const uint32_t fx = sf[w * kItemsPerThread + s], vx = sv[w * kItemsPerThread + s];
    const uint32_t f1 = fx | F[s];
    const uint32_t v1 = F[s] ? V[s] : vx + V[s];
    const uint32_t rank = f1 ? v1 : pre + v1;
    const uint32_t p = pay[i];
    if (p < out_cap) out[p] = (int32_t)rank;
  }
}

hipError_t launch_rank(const uint32_t* keys, const uint32_t* key_hi, const uint32_t* pay, uint32_t count,
                       float eb, double inv, int32_t* out, uint32_t out_cap, const SortWork& wk, hipStream_t st,
                       const uint32_t* dcount, const float* field_eb, const double* field_inv) {
  if (count == 0) return hipSuccess;
  const uint32_t ntiles = ceil_div(count, kSortTile);
  const uint32_t stiles = ceil_div((uint64_t)256 * ntiles, kSortTile);
  hipLaunchKernelGGL(k_rank, dim3(ntiles), dim3(kBlockThreads), 0, st, keys, key_hi, pay, count, eb, inv, out,
                     out_cap, wk.st + (size_t)8 * stiles, dyn_counter(wk.counters + 8), wk.err, dcount, field_eb,
                     field_inv);
  return hipGetLastError();
}

}  // namespace tszp
