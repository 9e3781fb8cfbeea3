// tszp_post.hip -- topology-restoring post-process and its verification.
//
// The reference restores critical points after decompression with sequential host
// stencils (example/test_cuda_topology.cu:118-326, example/test_cuda_topology_3d.cu:
// 77-434) that leave 2-15 % of them lost.  Here the decoded field d = (float)q * eb is
// turned into y with classify(y) == the transmitted types at every interior point,
// as a constraint repair that keeps every value inside its quantization cell:
//
//   * the interval of a point: [d, d + 0.9 eb] when d > 0, [d - 0.9 eb, d] when d < 0
//     (both on the side of d that holds the original value, so |y - x| < eb), and
//     [-eb 2^-16, eb 2^-16] in the zero cell, whose sign truncation lost
//     (|y - x| < eb (1 + 2^-16)).  Intervals of different cells never overlap, so a
//     comparison between points of different cells is the original's: only the order
//     inside a plateau (a connected set of points with one decoded value) is free,
//     and the original field proves that a solution exists;
//   * start: every point at the middle of its interval; the extrema of a non-zero cell
//     at lo + (hi - lo) (0.05 + 0.9 (sp + 1) / (S + 2)), S the largest sort position,
//     so that the extrema of one cell keep the order of their sort positions (ties
//     stay ties).  These extrema are anchors: the local search does not move them;
//   * local search: points within one step of a mismatch move to the value that
//     minimises the mismatches of their closed neighbourhood.  Points of one colour
//     class are at least three steps apart ((x + 2y) mod 5 in 2D, (x + 2y + 3z) mod 7
//     in 3D), so a class moves in parallel with exact gains, and every accepted move
//     strictly lowers the total.  After a round only points within two steps of a
//     previous mismatch can have changed, and only those are re-checked;
//   * exact window solve for what remains: the plateau around a mismatch (up to
//     kWinMax points, breadth first) is re-ordered by a depth-first search over the
//     weak orders of its free points relative to the fixed values around them, which
//     checks every constraint as soon as the points it reads are placed; a found order
//     becomes values spaced evenly between the fixed ones.  Windows of one 16-wide tile
//     run in order in one thread, tiles of one parity class in parallel, so the result
//     does not depend on scheduling.
// Between the small plateaus and the local rounds, whole plateaus are re-ordered by a
// min-conflicts search (tszp_platmc.hip); behind it the window solve only takes windows
// that hold their whole plateau, and a longer plateau search, then windows over parts of
// plateaus beyond its capacity, handle what is left.
// The full marking pass also yields the verification counts (missing / spurious /
// changed critical points) and the largest |y - d|.
#include <cstdlib>

#include "tszp_device.hpp"
#include "tszp_kernels.h"
#include "tszp_post_geo.hpp"

namespace tszp {

// ---------------------------------------------------------------------------------
// XCD-aware split of a streaming pass.  Workgroups are dispatched round-robin over the
// 8 XCDs (workgroup b runs on XCD b % 8), and each XCD has its own L2.  Giving every
// XCD one contiguous eighth of the items, swept in order by its workgroups, keeps the
// stencil neighbours one row or one plane away in the L2 that loaded them.  Items are
// handled in chunks of `chunk`; the grid is a multiple of 8.
// ---------------------------------------------------------------------------------
constexpr uint32_t kXcds = 8;

// non-temporal 16-byte loads for data streamed once (keeps L2 for reused rows)
typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt(const float4* p) {
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
struct XcdSweep {
  uint32_t begin, end, step;
};
__device__ __forceinline__ XcdSweep xcd_sweep(uint32_t n, uint32_t chunk) {
  const uint32_t x = blockIdx.x % kXcds, j = blockIdx.x / kXcds, per = gridDim.x / kXcds;
  const uint32_t nch = (n + chunk - 1) / chunk, chx = (nch + kXcds - 1) / kXcds;  // chunks per XCD
  XcdSweep r;
  r.begin = (x * chx + j) * chunk;
  r.end = min((x + 1) * chx * chunk, n);
  r.step = per * chunk;
  return r;
}
// one chunk per workgroup, the chunks split into kXcds contiguous ranges (grid: a
// multiple of kXcds from xcd_grid; chunks past the last are empty)
__device__ __forceinline__ uint32_t xcd_chunk() {
  return (blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
}
static inline uint32_t xcd_grid(uint32_t chunks) { return (chunks + kXcds - 1) / kXcds * kXcds; }


// append val to list for every lane whose flag is set (one atomic per wave)
__device__ __forceinline__ void wave_append(bool flag, uint32_t val, uint32_t* __restrict__ list, uint32_t cap,
                                            uint32_t* count) {
  const uint64_t m = __ballot(flag);
  if (!m) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader, kWave);
  if (flag) {
    const uint32_t pos = base + (uint32_t)__popcll(m & lanemask_lt(lane));
    if (pos < cap) list[pos] = val;
  }
}

__device__ __forceinline__ void wave_count(bool flag, uint32_t* count) {
  const uint64_t m = __ballot(flag);
  if (m && (threadIdx.x & (kWave - 1)) == __ffsll((unsigned long long)m) - 1)
    atomicAdd(count, (uint32_t)__popcll(m));
}

// Block-level list staging: lanes append to an LDS buffer (one LDS atomic per wave);
// lds_flush moves it to the global list with one global atomic per block, so that
// millions of appends never queue on one global counter.
__device__ __forceinline__ void lds_append(bool flag, uint32_t val, uint32_t* lbuf, uint32_t* lcnt) {
  const uint64_t m = __ballot(flag);
  if (!m) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(lcnt, (uint32_t)__popcll(m));
  base = __shfl(base, leader, kWave);
  if (flag) lbuf[base + (uint32_t)__popcll(m & lanemask_lt(lane))] = val;
}

// whole block; leaves the LDS list empty
__device__ __forceinline__ void lds_flush(const uint32_t* lbuf, uint32_t* lcnt, uint32_t* sh_base,
                                          uint32_t* __restrict__ list, uint32_t cap, uint32_t* gcount) {
  __syncthreads();
  const uint32_t cnt = *lcnt;
  if (threadIdx.x == 0 && cnt) *sh_base = atomicAdd(gcount, cnt);
  __syncthreads();
  if (cnt) {
    const uint32_t base = *sh_base;
    for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x)
      if (base + j < cap) list[base + j] = lbuf[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) *lcnt = 0;
  __syncthreads();
}

// ---------------------------------------------------------------------------------
// largest sort position
// ---------------------------------------------------------------------------------
// (16-byte loads when sp is aligned; few workgroups, since their atomics on one word
// serialise)
__global__ __launch_bounds__(kBlockThreads) void k_pp_spmax(const int32_t* __restrict__ sp, uint32_t m,
                                                           uint32_t* __restrict__ misc) {
  __shared__ uint32_t red[kWavesPerBlock];
  int32_t v = 0;
  const uint32_t tid = blockIdx.x * kBlockThreads + threadIdx.x, stride = gridDim.x * kBlockThreads;
  uint32_t head = m;
  if ((reinterpret_cast<uintptr_t>(sp) & 15u) == 0) {
    const int4* sp4 = reinterpret_cast<const int4*>(sp);
    const uint32_t m4 = m / 4;
    for (uint32_t i = tid; i < m4; i += stride) {
      const int4 x = sp4[i];
      v = max(v, max(max(x.x, x.y), max(x.z, x.w)));
    }
    head = 4 * m4;
  } else {
    head = 0;
  }
  for (uint32_t i = head + tid; i < m; i += stride) v = max(v, sp[i]);
  const uint32_t vm = seg_max<kWave>((uint32_t)v);  // sort positions are non-negative
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = vm;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t r = red[0];
#pragma unroll
    for (int k = 1; k < kWavesPerBlock; k++) r = red[k] > r ? red[k] : r;
    atomicMax(&misc[kPpSpMax], r);
  }
}

// largest sort position of every field of a batch: one workgroup per kFlatAlign chunk of
// the flat sort-position array (a chunk lies in one field), one atomic per workgroup;
// spm zeroed
__global__ __launch_bounds__(kBlockThreads) void k_pp_spmax_b(GeoB geo, const int32_t* __restrict__ sp,
                                                             uint32_t* __restrict__ spm) {
  __shared__ uint32_t red[kWavesPerBlock];
  const uint32_t c0 = blockIdx.x << kFlatShift;
  const FieldView F = geo.at(c0);
  const uint32_t ne = (uint32_t)*geo.f[F.fid].n_ext;
  const uint32_t k0 = c0 - F.base, k1 = min(ne, k0 + kFlatAlign);
  uint32_t v = 0;
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += kBlockThreads) {
    const int32_t x = sp[F.base + k];
    v = x > (int32_t)v ? (uint32_t)x : v;
  }
  v = seg_max<kWave>(v);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t r = red[0];
#pragma unroll
    for (int k = 1; k < kWavesPerBlock; k++) r = red[k] > r ? red[k] : r;
    if (r) atomicMax(&spm[F.fid], r);
  }
}

// Binomial smoothing along one axis (weights C(8, k) / 256, radius 4, edges clamped):
// the smoothed decoded field is the first guide of the whole-plateau solve.  fflags
// (batch, nullable): only fields whose flag is set are smoothed, chunk by chunk of the
// field map (the others' dst is left as it was).
template <int ND>
__device__ __forceinline__ void blur_point(const FieldView& F, const float* __restrict__ src,
                                           float* __restrict__ dst, uint32_t e, int axis) {
  constexpr float w[9] = {1.f / 256, 8.f / 256, 28.f / 256, 56.f / 256, 70.f / 256,
                          56.f / 256, 28.f / 256, 8.f / 256, 1.f / 256};
  int c[ND];
  coords_of<ND>(F.g, e - F.base, c);
  const int dim = dim_of<ND>(F.g, axis), ca = c[axis];
  const int64_t s = (int64_t)stride_of<ND>(F.g, axis);
  float acc = 0.f;
#pragma unroll
  for (int k = -4; k <= 4; k++) {
    const int q = min(max(ca + k, 0), dim - 1);
    acc += w[k + 4] * src[(int64_t)e + (int64_t)(q - ca) * s];
  }
  dst[e] = acc;
}

template <int ND, class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_blur(GEO geo, const float* __restrict__ src,
                                                          float* __restrict__ dst, uint32_t n, int axis,
                                                          const uint32_t* __restrict__ fflags) {
  if constexpr (GEO::kBatch) {
    if (fflags != nullptr) {
      const uint32_t nch = (n + kFlatAlign - 1) / kFlatAlign;
      for (uint32_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const uint32_t e0 = ch * kFlatAlign;
        const FieldView F = geo.at(e0);  // a chunk of the field map lies in one field
        if (!fflags[F.fid]) continue;
        const uint32_t end = min(e0 + kFlatAlign, F.base + F.n);
        for (uint32_t e = e0 + threadIdx.x; e < end; e += kBlockThreads) blur_point<ND>(F, src, dst, e, axis);
      }
      return;
    }
  }
  for (uint32_t e = blockIdx.x * kBlockThreads + threadIdx.x; e < n; e += gridDim.x * kBlockThreads) {
    const FieldView F = geo.at(e);
    if (e >= F.base + F.n) continue;
    blur_point<ND>(F, src, dst, e, axis);
  }
}

// ---------------------------------------------------------------------------------
// initial values.  A tile of kInitTile points ranks its extrema in row-major order
// (look-back over tiles) so that the k-th extremum takes sort position sp[k].
// ---------------------------------------------------------------------------------
constexpr int kInitPer = 64;
constexpr uint32_t kInitTile = (uint32_t)kBlockThreads * kInitPer;

// extrema per tile of kInitTile points (pass 1; an exclusive scan of these counts gives
// every tile the rank of its first extremum, so the start pass needs no look-back)
// (batch: kInitTile == kFlatAlign, so a tile lies in one field)
static_assert(kInitTile == kFlatAlign, "an initial-value tile lies in one field");
template <class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_ext_count(GEO geo, const uint8_t* __restrict__ t,
                                                               uint32_t* __restrict__ cnt) {
  __shared__ uint32_t red[kWavesPerBlock];
  const uint32_t t0 = blockIdx.x * kInitTile;
  const FieldView F = geo.at(t0);
  const uint32_t n = F.base + F.n;  // end of the tile's field
  const uint32_t* tg = reinterpret_cast<const uint32_t*>(t);
  uint32_t c = 0;
#pragma unroll 4
  for (int j = 0; j < (int)(kInitTile / 4 / kBlockThreads); j++) {
    const uint32_t e = t0 + 4 * (j * kBlockThreads + threadIdx.x);
    uint32_t v = 0;
    if (e + 3 < n) {
      v = tg[e >> 2];
    } else {
#pragma unroll
      for (int b = 0; b < 4; b++)
        if (e + b < n) v |= (uint32_t)t[e + b] << (8 * b);
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const uint32_t tb = (v >> (8 * b)) & 0xffu;
      c += (tb == 1u || tb == 2u) ? 1u : 0u;
    }
  }
  c = seg_sum<kWave>(c);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// tile_base[k]: extrema in the tiles before tile k (the scanned counts); the k-th
// extremum of a field takes sort position sp[base + k]
template <class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_init(GEO geo, const float* __restrict__ d,
                                                          const uint8_t* __restrict__ t,
                                                          const int32_t* __restrict__ sp, uint32_t n_ext_one,
                                                          float* __restrict__ y,
                                                          const uint32_t* __restrict__ tile_base,
                                                          uint32_t* misc) {
  __shared__ uint32_t tw[kInitTile / 4];
  __shared__ uint64_t msk[kBlockThreads];
  __shared__ uint32_t exl[kBlockThreads];
  __shared__ uint32_t red[kWavesPerBlock];
  const int tid = threadIdx.x;
  const uint32_t tile = blockIdx.x;
  const uint32_t t0 = tile * kInitTile;
  const FieldView F = geo.at(t0);
  const uint32_t n = F.base + F.n;  // end of the tile's field
  const float eb = F.eb;
  uint32_t n_ext = n_ext_one;
  if constexpr (GEO::kBatch) n_ext = (uint32_t)*geo.f[F.fid].n_ext;
  // stage the tile's types (4-byte aligned words, coalesced)
  const uint32_t* tg = reinterpret_cast<const uint32_t*>(t);
#pragma unroll 4
  for (int j = 0; j < (int)(kInitTile / 4 / kBlockThreads); j++) {
    const uint32_t w = j * kBlockThreads + tid;
    const uint32_t e = t0 + 4 * w;
    uint32_t v = 0;
    if (e + 3 < n) {
      v = tg[e >> 2];
    } else {
#pragma unroll
      for (int b = 0; b < 4; b++)
        if (e + b < n) v |= (uint32_t)t[e + b] << (8 * b);
    }
    tw[w] = v;
  }
  __syncthreads();
  uint64_t m = 0;
#pragma unroll
  for (int k = 0; k < kInitPer / 4; k++) {
    const uint32_t v = tw[tid * (kInitPer / 4) + k];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const uint32_t tb = (v >> (8 * b)) & 0xffu;
      m |= (uint64_t)(tb == 1u || tb == 2u) << (4 * k + b);
    }
  }
  msk[tid] = m;
  uint32_t tot;
  const uint32_t ex = block_excl_scan((uint32_t)__popcll(m), red, &tot);
  exl[tid] = ex;
  uint32_t base = tile_base[tile];
  if constexpr (GEO::kBatch) {
    base -= tile_base[F.base / kInitTile];  // extrema of the field before this tile
  } else {
    if (tid == 0 && tile == gridDim.x - 1) misc[kPpExt] = base + tot;
  }
  __syncthreads();
  const float sden = (float)geo.spmax(F, misc) + 2.f;
  // float4 groups of 4 points (the tile start is a multiple of 4); all loads of a batch
  // are issued before the values are computed
  constexpr int kGroups = kInitTile / 4 / kBlockThreads;  // 16
  constexpr int kBatch = 2;
  const float4* d4 = reinterpret_cast<const float4*>(d);
  float4* y4 = reinterpret_cast<float4*>(y);
  const bool vec = ((reinterpret_cast<uintptr_t>(d) | reinterpret_cast<uintptr_t>(y)) & 15u) == 0;
#pragma unroll
  for (int j0 = 0; j0 < kGroups; j0 += kBatch) {
    float dv[kBatch][4];
    int32_t sv[kBatch][4];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const uint32_t gidx = (uint32_t)(j0 + u) * kBlockThreads + tid;  // group in the tile
      const uint32_t e0 = t0 + 4 * gidx;
      if (vec && e0 + 3 < n) {
        const float4 q = d4[e0 >> 2];
        dv[u][0] = q.x;
        dv[u][1] = q.y;
        dv[u][2] = q.z;
        dv[u][3] = q.w;
      } else {
#pragma unroll
        for (int c = 0; c < 4; c++) dv[u][c] = e0 + c < n ? d[e0 + c] : 0.f;
      }

      const uint32_t tv = tw[gidx];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const uint32_t i = 4 * gidx + c;
        const uint32_t tb = (tv >> (8 * c)) & 0xffu;
        sv[u][c] = 0;
        if ((tb == 1u || tb == 2u) && sp != nullptr) {
          const uint32_t o = i >> 6, b = i & 63u;
          const uint32_t k = base + exl[o] + (uint32_t)__popcll(msk[o] & ((1ull << b) - 1ull));
          sv[u][c] = k < n_ext ? sp[F.base + k] : 0;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const uint32_t gidx = (uint32_t)(j0 + u) * kBlockThreads + tid;
      const uint32_t e0 = t0 + 4 * gidx;
      const uint32_t tv = tw[gidx];
      float out[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const uint32_t tb = (tv >> (8 * c)) & 0xffu;
        const float dd = dv[u][c];
        float lo, hi;
        cell_of(dd, eb, &lo, &hi);
        const float mid = dd == 0.f ? 0.f : 0.5f * (lo + hi);
        float v = mid;
        if (tb == 1u || tb == 2u) {
          const int32_t s = sv[u][c] < 0 ? 0 : sv[u][c];
          if (sp != nullptr && dd != 0.f) {
            v = lo + (hi - lo) * (0.05f + 0.9f * (((float)s + 1.f) / sden));
          } else {
            const float r = (float)s / ((float)s + 64.f);
            v = tb == 1u ? mid + (hi - mid) * (0.55f + 0.4f * r) : mid - (mid - lo) * (0.95f - 0.4f * r);
          }
        }
        out[c] = v;
      }
      if (vec && e0 + 3 < n) {
        y4[e0 >> 2] = make_float4(out[0], out[1], out[2], out[3]);
      } else {
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (e0 + c < n) y[e0 + c] = out[c];
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// full marking pass: list of mismatched points + verification counts
// ---------------------------------------------------------------------------------
template <int ND>
__global__ __launch_bounds__(kBlockThreads) void k_pp_mark(Grid g, const float* __restrict__ y,
                                                          const uint8_t* __restrict__ t, const float* __restrict__ d,
                                                          uint32_t n, uint32_t* __restrict__ bad, uint32_t cap,
                                                          uint32_t* __restrict__ misc) {
  constexpr uint32_t kStage = 4096;
  __shared__ uint32_t red[4][kWavesPerBlock];
  __shared__ uint32_t lbuf[kStage];
  __shared__ uint32_t lcnt, lbase;
  if (threadIdx.x == 0) lcnt = 0;
  __syncthreads();
  uint32_t miss = 0, spur = 0, chg = 0, dev = 0;
  for (uint32_t b0 = blockIdx.x * kBlockThreads; b0 < n; b0 += gridDim.x * kBlockThreads) {
    const uint32_t e = b0 + threadIdx.x;
    bool isbad = false;
    if (e < n) {
      int c[ND];
      coords_of<ND>(g, e, c);
      const int k = classify_at<ND>(g, y, e, c);
      const int tt = t[e];
      miss += (tt != 0 && k == 0);
      spur += (tt == 0 && k != 0);
      chg += (tt != 0 && k != 0 && k != tt);
      isbad = k != tt;
      if (d != nullptr) {
        const uint32_t dv = __float_as_uint(fabsf(y[e] - d[e]));
        dev = dv > dev ? dv : dev;
      }
    }
    if (bad != nullptr) {
      lds_append(isbad, e, lbuf, &lcnt);
      __syncthreads();
      if (lcnt > kStage - kBlockThreads) lds_flush(lbuf, &lcnt, &lbase, bad, cap, &misc[kPpBad]);
    }
  }
  if (bad != nullptr) lds_flush(lbuf, &lcnt, &lbase, bad, cap, &misc[kPpBad]);
  miss = seg_sum<kWave>(miss);
  spur = seg_sum<kWave>(spur);
  chg = seg_sum<kWave>(chg);
  dev = seg_max<kWave>(dev);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[0][w] = miss;
    red[1][w] = spur;
    red[2][w] = chg;
    red[3][w] = dev;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
      s0 += red[0][k];
      s1 += red[1][k];
      s2 += red[2][k];
      s3 = red[3][k] > s3 ? red[3][k] : s3;
    }
    unsigned long long* vo = reinterpret_cast<unsigned long long*>(misc + kPpVerify);
    if (s0) atomicAdd(&vo[0], (unsigned long long)s0);
    if (s1) atomicAdd(&vo[1], (unsigned long long)s1);
    if (s2) atomicAdd(&vo[2], (unsigned long long)s2);
    if (s3) atomicMax(&vo[3], (unsigned long long)s3);
  }
}

// Same pass, 4 consecutive points per thread along the fastest axis (its length a
// multiple of 4, buffers 16-byte aligned): one float4 per row of the stencil.  Each
// wave stages its mismatches in its own LDS list (its fill level is wave-uniform, so
// no LDS atomics and no workgroup barriers) and flushes it with one global atomic.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t kMarkWaveStage = 1024;  // staged entries per wave
constexpr int kMarkGroups = 1;            // float4 groups per thread and loop step

__device__ __forceinline__ void wave_stage_flush(const uint32_t* lw, uint32_t& wcnt, uint32_t* __restrict__ list,
                                                 uint32_t cap, uint32_t* gcount) {
  const int lane = threadIdx.x & (kWave - 1);
  wave_sync();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(gcount, wcnt);
  base = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(base, 0, kWave));
  for (uint32_t j = lane; j < wcnt; j += kWave)
    if (base + j < cap) list[base + j] = lw[j];
  wave_sync();
  wcnt = 0;
}

// Verification counts go to per-workgroup partials (part[4 * block + k]) that
// k_pp_reduce_verify adds up: thousands of same-address atomics would serialise.
// bm (nullable) receives one bit per point (word e / 32, bit e % 32), 1 where the point
// is mismatched.
template <int ND>
__global__ __launch_bounds__(kBlockThreads) void k_pp_mark4(Grid g, const float* __restrict__ y,
                                                           const uint8_t* __restrict__ t,
                                                           const float* __restrict__ d, uint32_t n4,
                                                           uint32_t* __restrict__ bad, uint32_t cap,
                                                           uint32_t* __restrict__ part,
                                                           uint32_t* __restrict__ bm,
                                                           uint32_t* __restrict__ misc) {
  constexpr int FA = ND - 1;  // fastest axis
  constexpr int U = kMarkGroups;
  __shared__ uint32_t red[4][kWavesPerBlock];
  __shared__ uint32_t lbuf[kWavesPerBlock][kMarkWaveStage];
  __shared__ uint32_t wc[kWavesPerBlock], sh_base;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  uint32_t* lw = lbuf[w];
  uint32_t wcnt = 0;  // wave-uniform fill level of lw
  const float4* y4 = reinterpret_cast<const float4*>(y);
  const float4* d4 = reinterpret_cast<const float4*>(d);
  const uint32_t* t4 = reinterpret_cast<const uint32_t*>(t);
  const int dimz = dim_of<ND>(g, FA);
  uint32_t miss = 0, spur = 0, chg = 0, dev = 0;
  const XcdSweep sw = xcd_sweep(n4, kBlockThreads * U);
  // the loop bound is uniform per workgroup, so every lane takes part in the ballots
  for (uint32_t b0 = sw.begin; b0 < sw.end; b0 += sw.step) {
    float4 v[U], lo4[U][ND - 1], hi4[U][ND - 1];
    float zm[U], zp[U];
    uint32_t tw[U];
    int c[U][ND];
    bool rowin[U], in[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t gi = b0 + u * kBlockThreads + threadIdx.x;
      in[u] = gi < sw.end;
      const uint32_t e0 = 4 * (in[u] ? gi : 0u);
      coords_of<ND>(g, e0, c[u]);
      rowin[u] = in[u];
#pragma unroll
      for (int a = 0; a < FA; a++) rowin[u] = rowin[u] && c[u][a] >= 1 && c[u][a] + 2 <= dim_of<ND>(g, a);
      v[u] = in[u] ? y4[gi] : make_float4(0.f, 0.f, 0.f, 0.f);
      tw[u] = in[u] ? __builtin_nontemporal_load(&t4[gi]) : 0u;  // streamed once: keep L2 for y
      zm[u] = zp[u] = 0.f;
      if (rowin[u]) {
#pragma unroll
        for (int a = 0; a < FA; a++) {
          const uint32_t s4 = stride_of<ND>(g, a) >> 2;
          lo4[u][a] = y4[gi - s4];
          hi4[u][a] = y4[gi + s4];
        }
        if (c[u][FA] > 0) zm[u] = y[e0 - 1];
        if (c[u][FA] + 4 < dimz) zp[u] = y[e0 + 4];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t e0 = 4 * (b0 + u * kBlockThreads + threadIdx.x);
      const float vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
      bool flag[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int z = c[u][FA] + i;
        int k = 0;
        if (rowin[u] && z >= 1 && z + 2 <= dimz) {
          float nb[2 * ND];
#pragma unroll
          for (int a = 0; a < FA; a++) {
            const float l[4] = {lo4[u][a].x, lo4[u][a].y, lo4[u][a].z, lo4[u][a].w};
            const float h[4] = {hi4[u][a].x, hi4[u][a].y, hi4[u][a].z, hi4[u][a].w};
            nb[2 * a] = l[i];
            nb[2 * a + 1] = h[i];
          }
          nb[2 * FA] = i == 0 ? zm[u] : vv[i - 1];
          nb[2 * FA + 1] = i == 3 ? zp[u] : vv[i + 1];
          k = classify_vals<ND>(vv[i], nb);
        }
        const int tt = (int)((tw[u] >> (8 * i)) & 0xffu);
        miss += (in[u] && tt != 0 && k == 0);
        spur += (in[u] && tt == 0 && k != 0);
        chg += (in[u] && tt != 0 && k != 0 && k != tt);
        flag[i] = in[u] && k != tt;
      }
      if (bm != nullptr) {
        // eight lanes' flags make one bit word (32 points; group b0 + ... is a multiple
        // of 8 at lanes 0, 8, ...): lane l's four bits at 4 (l % 8)
        uint32_t v = ((flag[0] ? 1u : 0u) | (flag[1] ? 2u : 0u) | (flag[2] ? 4u : 0u) | (flag[3] ? 8u : 0u))
                     << (4 * (lane & 7));
        v |= (uint32_t)__shfl_xor((int)v, 1, kWave);
        v |= (uint32_t)__shfl_xor((int)v, 2, kWave);
        v |= (uint32_t)__shfl_xor((int)v, 4, kWave);
        const uint32_t gi = b0 + u * kBlockThreads + threadIdx.x;
        if ((lane & 7) == 0 && in[u]) __builtin_nontemporal_store(v, &bm[gi >> 3]);
      }
      if (d != nullptr && in[u]) {
        const float4 dd = ld_nt(&d4[b0 + u * kBlockThreads + threadIdx.x]);
        const float dv[4] = {dd.x, dd.y, dd.z, dd.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const uint32_t q = __float_as_uint(fabsf(vv[i] - dv[i]));
          dev = q > dev ? q : dev;
        }
      }
      if (bad != nullptr) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const uint64_t m = __ballot(flag[i]);
          if (flag[i]) lw[wcnt + (uint32_t)__popcll(m & lanemask_lt(lane))] = e0 + i;
          wcnt += (uint32_t)__popcll(m);
        }
      }
    }
    if (bad != nullptr && wcnt > kMarkWaveStage - 4 * U * kWave) wave_stage_flush(lw, wcnt, bad, cap, &misc[kPpBad]);
  }
  // what is left in the four wave lists goes out with one global atomic per workgroup
  if (bad != nullptr) {
    if (lane == 0) wc[w] = wcnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
#pragma unroll
      for (int k = 0; k < kWavesPerBlock; k++) tot += wc[k];
      sh_base = tot ? atomicAdd(&misc[kPpBad], tot) : 0u;
    }
    __syncthreads();
    uint32_t base = sh_base;
    for (int k = 0; k < w; k++) base += wc[k];
    for (uint32_t j = lane; j < wcnt; j += kWave)
      if (base + j < cap) bad[base + j] = lw[j];
  }
  miss = seg_sum<kWave>(miss);
  spur = seg_sum<kWave>(spur);
  chg = seg_sum<kWave>(chg);
  dev = seg_max<kWave>(dev);
  if (lane == 0) {
    red[0][w] = miss;
    red[1][w] = spur;
    red[2][w] = chg;
    red[3][w] = dev;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
      s0 += red[0][k];
      s1 += red[1][k];
      s2 += red[2][k];
      s3 = red[3][k] > s3 ? red[3][k] : s3;
    }
    part[4 * blockIdx.x] = s0;
    part[4 * blockIdx.x + 1] = s1;
    part[4 * blockIdx.x + 2] = s2;
    part[4 * blockIdx.x + 3] = s3;
  }
}

// The mismatch list in index order from the mismatch bitmap (bit e % 32 of word e / 32):
// one tile of kCompactTile points per workgroup, each thread a contiguous run of
// kCompactTile / 256 points (eight 16-byte words), tiles ordered by a decoupled
// look-back; status: one zeroed word per tile.  The last tile stores the total in
// misc[kPpBad]; entries past cap are dropped.
constexpr uint32_t kCompactTile = 262144;
__device__ __forceinline__ uint32_t bits_below(uint32_t avail) {  // mask of the first min(avail, 32) bits
  return avail >= 32u ? 0xffffffffu : (1u << avail) - 1u;
}
__global__ __launch_bounds__(kBlockThreads) void k_pp_compact_map(const uint4* __restrict__ bm16, uint32_t n,
                                                                 uint64_t* __restrict__ status,
                                                                 uint32_t* __restrict__ out, uint32_t cap,
                                                                 uint32_t* __restrict__ misc) {
  constexpr int V = kCompactTile / 128 / kBlockThreads;  // 16-byte words (128 points each) per thread
  __shared__ uint32_t lds4[kWavesPerBlock];
  __shared__ uint32_t sh_excl;
  const uint32_t tile = blockIdx.x;
  const uint32_t v0 = (tile * kBlockThreads + threadIdx.x) * V;
  const uint32_t n16 = (n + 127) / 128;
  uint4 m[V];
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < V; k++) {
    const uint32_t wi = v0 + k;
    m[k] = wi < n16 ? bm16[wi] : make_uint4(0, 0, 0, 0);
    const uint32_t first = 128u * wi;
    if (first + 128u > n) {  // points at and past n are not part of the map
      const uint32_t avail = first < n ? n - first : 0u;
      m[k].x &= bits_below(avail);
      m[k].y &= bits_below(avail > 32u ? avail - 32u : 0u);
      m[k].z &= bits_below(avail > 64u ? avail - 64u : 0u);
      m[k].w &= bits_below(avail > 96u ? avail - 96u : 0u);
    }
    c += __popc(m[k].x) + __popc(m[k].y) + __popc(m[k].z) + __popc(m[k].w);
  }
  uint32_t tot;
  uint32_t pos = block_excl_scan(c, lds4, &tot);
  if (threadIdx.x < kWave) {
    const uint64_t ex = lookback(status, tile, tot, &misc[kPpErr]);
    if (threadIdx.x == 0) {
      sh_excl = (uint32_t)ex;
      if (tile == gridDim.x - 1) misc[kPpBad] = (uint32_t)ex + tot;
    }
  }
  __syncthreads();
  pos += sh_excl;
#pragma unroll
  for (int k = 0; k < V; k++) {
    const uint32_t w[4] = {m[k].x, m[k].y, m[k].z, m[k].w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      uint32_t bits = w[q];
      while (bits) {
        const uint32_t bit = (uint32_t)__builtin_ctz(bits);
        bits &= bits - 1u;
        if (pos < cap) out[pos] = 128u * (v0 + k) + 32u * q + bit;
        pos++;
      }
    }
  }
}

// adds up the per-workgroup verification partials (one workgroup)
__global__ __launch_bounds__(kBlockThreads) void k_pp_reduce_verify(const uint32_t* __restrict__ part, uint32_t nb,
                                                                   uint32_t* __restrict__ misc) {
  __shared__ unsigned long long red[4][kWavesPerBlock];
  unsigned long long s0 = 0, s1 = 0, s2 = 0;
  uint32_t s3 = 0;
  for (uint32_t k = threadIdx.x; k < nb; k += kBlockThreads) {
    s0 += part[4 * k];
    s1 += part[4 * k + 1];
    s2 += part[4 * k + 2];
    s3 = part[4 * k + 3] > s3 ? part[4 * k + 3] : s3;
  }
  s0 = seg_sum<kWave>(s0);
  s1 = seg_sum<kWave>(s1);
  s2 = seg_sum<kWave>(s2);
  s3 = seg_max<kWave>(s3);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[0][w] = s0;
    red[1][w] = s1;
    red[2][w] = s2;
    red[3][w] = s3;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t0 = 0, t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
      t0 += red[0][k];
      t1 += red[1][k];
      t2 += red[2][k];
      t3 = red[3][k] > t3 ? red[3][k] : t3;
    }
    unsigned long long* vo = reinterpret_cast<unsigned long long*>(misc + kPpVerify);
    vo[0] += t0;
    vo[1] += t1;
    vo[2] += t2;
    vo[3] = t3 > vo[3] ? t3 : vo[3];
  }
}

// sum of n per-workgroup counts into misc[idx] (one workgroup)
__global__ __launch_bounds__(kBlockThreads) void k_pp_reduce_sum(const uint32_t* __restrict__ part, uint32_t n,
                                                                uint32_t* __restrict__ misc, int idx) {
  __shared__ uint32_t red[kWavesPerBlock];
  // independent partial sums, so that several loads are in flight per thread
  uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0;
  uint32_t k = threadIdx.x;
  for (; k + 3 * kBlockThreads < n; k += 4 * kBlockThreads) {
    v0 += part[k];
    v1 += part[k + kBlockThreads];
    v2 += part[k + 2 * kBlockThreads];
    v3 += part[k + 3 * kBlockThreads];
  }
  for (; k < n; k += kBlockThreads) v0 += part[k];
  uint32_t v = seg_sum<kWave>(v0 + v1 + v2 + v3);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) t += red[k];
    misc[idx] += t;
  }
}

// Full marking pass over a batch of fields (any fastest-axis length): workgroups take
// chunks of kMarkChunk points (a chunk lies in one field), list the mismatches through a
// workgroup-staged LDS list (one global atomic per flush), write the mismatch map when
// bm (one bit per point) is set, and leave per-workgroup verification partials
// (k_pp_reduce_verify).
// With fcount, every field's own counts (missing, spurious, changed, max |y - d| bits)
// are added up too, one atomic per chunk and count.
constexpr uint32_t kMarkChunk = 4 * kBlockThreads;
template <int ND>
__global__ __launch_bounds__(kBlockThreads) void k_pp_mark_b(GeoB geo, const float* __restrict__ y,
                                                            const uint8_t* __restrict__ t,
                                                            const float* __restrict__ d, uint32_t nchunks,
                                                            uint32_t* __restrict__ bad, uint32_t cap,
                                                            uint32_t* __restrict__ part, uint32_t* __restrict__ bm,
                                                            unsigned long long* __restrict__ fcount,
                                                            uint32_t* __restrict__ misc) {
  constexpr uint32_t kStage = 4096;
  __shared__ uint32_t red[4][kWavesPerBlock];
  __shared__ uint32_t lbuf[kStage];
  __shared__ uint32_t lcnt, lbase;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (threadIdx.x == 0) lcnt = 0;
  __syncthreads();
  uint32_t miss = 0, spur = 0, chg = 0, dev = 0;
  // 3D: every XCD sweeps one contiguous eighth of the chunks, so the plane neighbours
  // stay in the L2 that loaded them (measured: 1.06 -> 0.78 ms for the 3D fields of the
  // mixed 256-field batch; the same sweep made its 2D fields 2.5x slower, 0.25 -> 0.61 ms)
  XcdSweep sw{blockIdx.x, nchunks, gridDim.x};
  if constexpr (ND == 3) sw = xcd_sweep(nchunks, 1);
  for (uint32_t ch = sw.begin; ch < sw.end; ch += sw.step) {
    const uint32_t c0 = ch * kMarkChunk;
    const FieldView F = geo.at(c0);
    const uint32_t end = F.base + F.n;
    if (c0 >= end) continue;  // padding after the field (uniform over the workgroup)
    uint32_t fm = 0, fs = 0, fc = 0, fd = 0;
    bool flags[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t e = c0 + j * kBlockThreads + threadIdx.x;
      bool isbad = false;
      if (e < end) {
        int c[ND];
        coords_of<ND>(F.g, e - F.base, c);
        const int k = classify_at<ND>(F.g, y, e, c);
        const int tt = t[e];
        fm += (tt != 0 && k == 0);
        fs += (tt == 0 && k != 0);
        fc += (tt != 0 && k != 0 && k != tt);
        isbad = k != tt;
        if (d != nullptr) {
          const uint32_t dv = __float_as_uint(fabsf(y[e] - d[e]));
          fd = dv > fd ? dv : fd;
        }
      }
      flags[j] = isbad;
      if (bad != nullptr) lds_append(isbad, e, lbuf, &lcnt);
    }
    if (bm != nullptr) {  // each wave's 64 points of a step are two bit words (c0: a multiple of 1024)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint64_t m = __ballot(flags[j]);
        const uint32_t wd = (c0 + j * kBlockThreads + w * kWave) >> 5;
        if (lane == 0) {
          bm[wd] = (uint32_t)m;
          bm[wd + 1] = (uint32_t)(m >> 32);
        }
      }
    }
    if (bad != nullptr) {
      __syncthreads();
      if (lcnt > kStage - kMarkChunk) lds_flush(lbuf, &lcnt, &lbase, bad, cap, &misc[kPpBad]);
    }
    if (fcount != nullptr) {  // the field's counts (whole workgroup)
      const uint32_t v[4] = {seg_sum<kWave>(fm), seg_sum<kWave>(fs), seg_sum<kWave>(fc), seg_max<kWave>(fd)};
      __syncthreads();
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) red[k][w] = v[k];
      }
      __syncthreads();
      if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        uint32_t a = 0;
#pragma unroll
        for (int q = 0; q < kWavesPerBlock; q++) a = k == 3 ? (red[k][q] > a ? red[k][q] : a) : a + red[k][q];
        if (a) {
          if (k == 3) atomicMax(&fcount[4 * F.fid + 3], (unsigned long long)a);
          else atomicAdd(&fcount[4 * F.fid + k], (unsigned long long)a);
        }
      }
    }
    miss += fm;
    spur += fs;
    chg += fc;
    dev = fd > dev ? fd : dev;
  }
  if (bad != nullptr) lds_flush(lbuf, &lcnt, &lbase, bad, cap, &misc[kPpBad]);
  miss = seg_sum<kWave>(miss);
  spur = seg_sum<kWave>(spur);
  chg = seg_sum<kWave>(chg);
  dev = seg_max<kWave>(dev);
  __syncthreads();
  if (lane == 0) {
    red[0][w] = miss;
    red[1][w] = spur;
    red[2][w] = chg;
    red[3][w] = dev;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
      s0 += red[0][k];
      s1 += red[1][k];
      s2 += red[2][k];
      s3 = red[3][k] > s3 ? red[3][k] : s3;
    }
    part[4 * blockIdx.x] = s0;
    part[4 * blockIdx.x + 1] = s1;
    part[4 * blockIdx.x + 2] = s2;
    part[4 * blockIdx.x + 3] = s3;
  }
}

// ---------------------------------------------------------------------------------
// active set: the non-anchored points of the closed neighbourhoods of the mismatches
// ---------------------------------------------------------------------------------
template <int ND, class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_active(GEO geo, const uint32_t* __restrict__ bad, uint32_t nbad,
                                                            const float* __restrict__ d,
                                                            const uint8_t* __restrict__ t, uint32_t* bits,
                                                            uint32_t* __restrict__ act, uint32_t cap, int anchor,
                                                            uint32_t* misc) {
  // one list per colour class (act + colour * cap), filled in two passes: block-local
  // offsets from LDS counters, then one global atomic per colour per block
  constexpr int NC = ND == 2 ? 5 : 7;
  __shared__ uint32_t lcnt[NC], lbase[NC];
  if (threadIdx.x < NC) lcnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  const bool valid = i < nbad;
  const uint32_t e = valid ? bad[i] : 0u;
  const FieldView F = geo.at(e);
  const Grid& g = F.g;
  int c[ND];
  coords_of<ND>(g, e - F.base, c);
  uint32_t pk[2 * ND + 1], off[2 * ND + 1];
  int col[2 * ND + 1];
#pragma unroll
  for (int k = 0; k <= 2 * ND; k++) {
    col[k] = -1;
    pk[k] = e;
    off[k] = 0;
    if (!valid) continue;
    int q[ND];
#pragma unroll
    for (int a = 0; a < ND; a++) q[a] = c[a];
    if (k > 0) {
      const int a = (k - 1) >> 1;
      const int s = ((k - 1) & 1) ? 1 : -1;
      q[a] += s;
      pk[k] = s > 0 ? e + stride_of<ND>(g, a) : e - stride_of<ND>(g, a);
    }
    if (k == 0 || inside<ND>(g, q)) {
      const uint32_t p = pk[k];
      if (!(anchor && anchored(t[p], d[p])) && claim_bit(bits, p)) {
        col[k] = colour_of<ND>(q);
        off[k] = atomicAdd(&lcnt[col[k]], 1u);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    const uint32_t n = lcnt[threadIdx.x];
    lbase[threadIdx.x] = n ? atomicAdd(&misc[kPpActC + threadIdx.x], n) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k <= 2 * ND; k++) {
    if (col[k] < 0) continue;
    const uint32_t pos = lbase[col[k]] + off[k];
    if (pos < cap) act[(size_t)col[k] * cap + pos] = pk[k];
  }
}

// ---------------------------------------------------------------------------------
// single-point moves of one colour class
// ---------------------------------------------------------------------------------
template <int M>
__device__ __forceinline__ void sort_small(float (&v)[M]) {
#pragma unroll
  for (int a = 0; a < M; a++)
#pragma unroll
    for (int b = 0; b + 1 < M - a; b++) {
      const float x = v[b], z = v[b + 1];
      v[b] = fminf(x, z);
      v[b + 1] = fmaxf(x, z);
    }
}

template <int ND>
__device__ bool move_point(const Grid& g, uint32_t p, const int (&c)[ND], const float* __restrict__ d,
                           const uint8_t* __restrict__ t, float* __restrict__ y, float eb) {
  constexpr int K = 2 * ND;
  const bool intp = interior<ND>(g, c);
  const int tp = t[p];
  float yq[K], nn[K][K];
  int tq[K];
  bool inq[K], intq[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int a = k >> 1;
    const uint32_t s = stride_of<ND>(g, a);
    int q[ND];
#pragma unroll
    for (int b = 0; b < ND; b++) q[b] = c[b];
    q[a] += (k & 1) ? 1 : -1;
    inq[k] = inside<ND>(g, q);
    intq[k] = inq[k] && interior<ND>(g, q);
    const uint32_t qe = (k & 1) ? p + s : p - s;
    yq[k] = inq[k] ? y[qe] : 0.f;
    tq[k] = inq[k] ? (int)t[qe] : 0;
#pragma unroll
    for (int j = 0; j < K; j++) {
      nn[k][j] = 0.f;
      if (intq[k] && j != (k ^ 1)) {
        const uint32_t sj = stride_of<ND>(g, j >> 1);
        nn[k][j] = y[(j & 1) ? qe + sj : qe - sj];
      }
    }
  }
  auto nbad = [&](float v) -> int {
    int b = intp ? (classify_vals<ND>(v, yq) != tp) : 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (intq[k]) {
        float w[K];
#pragma unroll
        for (int j = 0; j < K; j++) w[j] = (j == (k ^ 1)) ? v : nn[k][j];
        b += classify_vals<ND>(yq[k], w) != tq[k];
      }
    }
    return b;
  };
  const float cur = y[p];
  const int v0 = nbad(cur);
  if (v0 == 0) return false;
  float lo, hi;
  cell_of(d[p], eb, &lo, &hi);
  float vals[K + 2];
  vals[0] = lo;
  vals[1] = hi;
#pragma unroll
  for (int k = 0; k < K; k++) vals[k + 2] = (inq[k] && yq[k] >= lo && yq[k] <= hi) ? yq[k] : INFINITY;
  sort_small<K + 2>(vals);
  int best = v0;
  float bestv = cur;
#pragma unroll
  for (int a = 0; a < K + 2; a++) {
    const float va = vals[a];
    if (!(va <= hi)) continue;
    if (a == 0 || va != vals[a - 1]) {
      const int s = nbad(va);
      if (s < best) {
        best = s;
        bestv = va;
      }
    }
    if (a + 1 < K + 2) {
      const float vb = vals[a + 1];
      if (vb <= hi && vb > va) {
        const float mv = 0.5f * (va + vb);
        if (mv > va && mv < vb) {
          const int s = nbad(mv);
          if (s < best) {
            best = s;
            bestv = mv;
          }
        }
      }
    }
  }
  if (best < v0) {
    y[p] = bestv;
    return true;
  }
  return false;
}

template <int ND, class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_move(GEO geo, const uint32_t* __restrict__ act, uint32_t cap,
                                                          int colour, const float* __restrict__ d,
                                                          const uint8_t* __restrict__ t, float* __restrict__ y,
                                                          uint32_t* misc) {
  __shared__ uint32_t red[kWavesPerBlock];
  const uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  uint32_t cnt = misc[kPpActC + colour];
  cnt = cnt < cap ? cnt : cap;
  bool moved = false;
  if (i < cnt) {
    const uint32_t p = act[(size_t)colour * cap + i];
    const FieldView F = geo.at(p);
    int c[ND];
    coords_of<ND>(F.g, p - F.base, c);
    moved = move_point<ND>(F.g, p, c, d, t, y, F.eb);
  }
  // one atomic per workgroup (same-address atomics serialise)
  const uint32_t nm = (uint32_t)__popcll(__ballot(moved));
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nm;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t tot = red[0] + red[1] + red[2] + red[3];
    if (tot) atomicAdd(&misc[kPpMoves], tot);
  }
}

// The same move with the candidates spread over a 16-lane segment: every lane of the
// segment loads the point's neighbourhood (the same addresses, one request per wave
// instruction) and sorts the candidate values; lane j scores candidate j of the serial
// order above (even j: value j/2 of the sorted list, odd j: the midpoint after it), and
// the lowest score wins, the earliest candidate among equal scores -- the serial
// loop's choice.  One point's candidates cost one candidate's latency instead of
// sixteen.
template <int ND>
__device__ __forceinline__ bool move_point_seg(const Grid& g, uint32_t p, const int (&c)[ND], const float* __restrict__ d,
                                               const uint8_t* __restrict__ t, float* __restrict__ y, float eb, int j,
                                               bool act) {
  constexpr int K = 2 * ND, M = K + 2;
  bool intp = false;
  int tp = 0;
  float yq[K], nn[K][K];
  int tq[K];
  bool inq[K], intq[K];
  float cur = 0.f;
  if (act) {
    intp = interior<ND>(g, c);
    tp = t[p];
    cur = y[p];
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int a = k >> 1;
    const uint32_t s = stride_of<ND>(g, a);
    int q[ND];
#pragma unroll
    for (int b = 0; b < ND; b++) q[b] = c[b];
    q[a] += (k & 1) ? 1 : -1;
    inq[k] = act && inside<ND>(g, q);
    intq[k] = inq[k] && interior<ND>(g, q);
    const uint32_t qe = (k & 1) ? p + s : p - s;
    yq[k] = inq[k] ? y[qe] : 0.f;
    tq[k] = inq[k] ? (int)t[qe] : 0;
#pragma unroll
    for (int jj = 0; jj < K; jj++) {
      nn[k][jj] = 0.f;
      if (intq[k] && jj != (k ^ 1)) {
        const uint32_t sj = stride_of<ND>(g, jj >> 1);
        nn[k][jj] = y[(jj & 1) ? qe + sj : qe - sj];
      }
    }
  }
  auto nbad = [&](float v) -> int {
    int b = intp ? (classify_vals<ND>(v, yq) != tp) : 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (intq[k]) {
        float w[K];
#pragma unroll
        for (int jj = 0; jj < K; jj++) w[jj] = (jj == (k ^ 1)) ? v : nn[k][jj];
        b += classify_vals<ND>(yq[k], w) != tq[k];
      }
    }
    return b;
  };
  const int v0 = act ? nbad(cur) : 0;
  float lo = 0.f, hi = 0.f;
  if (act) cell_of(d[p], eb, &lo, &hi);
  float vals[M];
  vals[0] = lo;
  vals[1] = hi;
#pragma unroll
  for (int k = 0; k < K; k++) vals[k + 2] = (inq[k] && yq[k] >= lo && yq[k] <= hi) ? yq[k] : INFINITY;
  sort_small<M>(vals);
  // candidate j of the serial order
  const int a = j >> 1;
  float va = INFINITY, vp = INFINITY, vb = INFINITY;
#pragma unroll
  for (int k = 0; k < M; k++) {
    va = k == a ? vals[k] : va;
    vp = k + 1 == a ? vals[k] : vp;
    vb = k == a + 1 ? vals[k] : vb;
  }
  bool ok = act && v0 != 0 && a < M && va <= hi;
  float cand = va;
  if (ok) {
    if ((j & 1) == 0) {
      ok = a == 0 || va != vp;
    } else {
      const float mv = 0.5f * (va + vb);
      ok = a + 1 < M && vb <= hi && vb > va && mv > va && mv < vb;
      cand = mv;
    }
  }
  const int sc = ok ? nbad(cand) : 0x7fff;
  // lowest (score, candidate index) over the segment
  int key = (sc << 4) | j;
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) {
    const int other = __shfl_xor(key, o, kWave);
    key = other < key ? other : key;
  }
  const int best = key >> 4, bj = key & 15;
  const float bestv = __shfl(cand, (threadIdx.x & (kWave - 1) & ~15) + bj, kWave);
  if (act && v0 != 0 && best < v0) {
    if (j == 0) y[p] = bestv;
    return j == 0;
  }
  return false;
}

template <int ND, class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_move_seg(GEO geo, const uint32_t* __restrict__ act,
                                                              uint32_t cap, int colour, const float* __restrict__ d,
                                                              const uint8_t* __restrict__ t, float* __restrict__ y,
                                                              uint32_t* misc) {
  __shared__ uint32_t red[kWavesPerBlock];
  const uint32_t gi = blockIdx.x * kBlockThreads + threadIdx.x;
  const uint32_t i = gi >> 4;
  const int j = (int)(gi & 15u);
  uint32_t cnt = misc[kPpActC + colour];
  cnt = cnt < cap ? cnt : cap;
  const bool valid = i < cnt;
  uint32_t p = 0;
  int c[ND];
#pragma unroll
  for (int b = 0; b < ND; b++) c[b] = 0;
  if (valid) p = act[(size_t)colour * cap + i];
  const FieldView F = geo.at(p);
  if (valid) coords_of<ND>(F.g, p - F.base, c);
  const bool moved = move_point_seg<ND>(F.g, p, c, d, t, y, F.eb, j, valid);
  const uint32_t nm = (uint32_t)__popcll(__ballot(moved));
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nm;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t tot = red[0] + red[1] + red[2] + red[3];
    if (tot) atomicAdd(&misc[kPpMoves], tot);
  }
}

// ---------------------------------------------------------------------------------
// re-marking: every point within two steps of a previous mismatch
// ---------------------------------------------------------------------------------
__constant__ const int8_t kRing2D[13][2] = {{0, 0},  {-1, 0}, {1, 0},   {0, -1}, {0, 1},  {-2, 0}, {2, 0},
                                            {0, -2}, {0, 2},  {-1, -1}, {-1, 1}, {1, -1}, {1, 1}};
__constant__ const int8_t kRing3D[25][3] = {
    {0, 0, 0},  {-1, 0, 0},  {1, 0, 0},  {0, -1, 0},  {0, 1, 0},   {0, 0, -1},  {0, 0, 1},
    {-2, 0, 0}, {2, 0, 0},   {0, -2, 0}, {0, 2, 0},   {0, 0, -2},  {0, 0, 2},   {-1, -1, 0},
    {-1, 1, 0}, {1, -1, 0},  {1, 1, 0},  {-1, 0, -1}, {-1, 0, 1},  {1, 0, -1},  {1, 0, 1},
    {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1}};

constexpr int kRemarkChunks = 8;

template <int ND, class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_remark(GEO geo, const uint32_t* __restrict__ old,
                                                            uint32_t nold, const float* __restrict__ y,
                                                            const uint8_t* __restrict__ t, uint32_t* bits,
                                                            uint32_t* __restrict__ bad, uint32_t cap, int rings,
                                                            uint32_t* misc) {
  // rings: 1 re-checks the listed points only, otherwise everything within two steps.
  // One thread per (listed point, ring offset) and kRemarkChunks of them per thread: the
  // claims and stencil loads of a point's ring are in flight together instead of one
  // after the other, and a workgroup appends to the global list once.
  constexpr int RM = ND == 2 ? 13 : 25;
  const uint32_t R = rings == 1 ? 1u : (uint32_t)RM;
  __shared__ uint32_t lbuf[kBlockThreads * kRemarkChunks];
  __shared__ uint32_t lcnt, lbase;
  if (threadIdx.x == 0) lcnt = 0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRemarkChunks; k++) {
    const uint64_t gi = ((uint64_t)blockIdx.x * kRemarkChunks + k) * kBlockThreads + threadIdx.x;
    const uint32_t i = (uint32_t)(gi / R), r = (uint32_t)(gi % R);
    bool isbad = false;
    uint32_t p = 0;
    if (i < nold) {
      const uint32_t e = old[i];
      const FieldView F = geo.at(e);
      const Grid& g = F.g;
      int c[ND], q[ND];
      coords_of<ND>(g, e - F.base, c);
      int64_t off = 0;
#pragma unroll
      for (int a = 0; a < ND; a++) {
        const int o = ND == 2 ? kRing2D[r][a] : kRing3D[r][a];
        q[a] = c[a] + o;
        off += (int64_t)o * (int64_t)stride_of<ND>(g, a);
      }
      if (inside<ND>(g, q)) {
        p = (uint32_t)((int64_t)e + off);
        if (claim_bit(bits, p)) isbad = classify_at<ND>(g, y, p, q) != (int)t[p];
      }
    }
    lds_append(isbad, p, lbuf, &lcnt);
  }
  lds_flush(lbuf, &lcnt, &lbase, bad, cap, &misc[kPpBad]);
}

// ---------------------------------------------------------------------------------
// exact solve of plateaus of two or three points
// ---------------------------------------------------------------------------------

// relation mask of one point's decoded value: for each of its 2*ND neighbours (axis a:
// bit 2a at -1, bit 2a+1 at +1), bit k says the neighbour is inside the field with the
// same value, bit 8+k that it is inside with a smaller one
template <int ND>
__device__ __forceinline__ uint32_t rel_of(const Grid& g, uint32_t base, const float* __restrict__ d, uint32_t e) {
  int c[ND];
  coords_of<ND>(g, e - base, c);
  const float v = d[e];
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 2 * ND; k++) {
    const int a = k >> 1;
    const bool ok = (k & 1) ? c[a] + 1 < dim_of<ND>(g, a) : c[a] >= 1;
    if (!ok) continue;
    const uint32_t s = stride_of<ND>(g, a);
    const float q = d[(k & 1) ? e + s : e - s];
    r |= (q == v ? 1u : 0u) << k;
    r |= (v > q ? 1u : 0u) << (8 + k);
  }
  return r;
}

// Solve of one plateau of two or three points.  Every other neighbour of such a plateau
// lies in another cell, so only the order inside the plateau decides the
// types of its points, and no other point's type depends on it.  Each point's type is
// tabulated for every relation (<, =, >) to each of the other plateau points (a grid
// has no triangles, so a point sees at most two of them); the 3 (two points) or 13
// (three points) weak orders are then table look-ups, and the first that gives every
// point its type is written.
__constant__ const int8_t kWeak3[13][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0},
                                           {1, 0, 1}, {0, 1, 1}, {0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                           {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
__constant__ const int8_t kWeak2[3][2] = {{0, 0}, {0, 1}, {1, 0}};

template <int ND, class GEO>
__device__ __forceinline__ bool tiny_solve_one(const GEO& geo, const uint32_t (&P)[3], int np,
                                               const float* __restrict__ d, const uint32_t (&rms)[3],
                                               const uint8_t* __restrict__ t, float* __restrict__ y, int anchor,
                                               bool write, uint32_t* matched) {
  constexpr int K = 2 * ND;
  bool done = false;
  const float db = d[P[0]];
  const FieldView F = geo.at(P[0]);
  const Grid& g = F.g;
  const float eb = F.eb;
  int tp[3];
  bool intr[3], anc[3];
  float yv[3];
  uint32_t ttab[3];  // type of point j for relations (ra, rb) to its two other points: 2 bits each
#pragma unroll
  for (int j = 0; j < 3; j++) {
    tp[j] = 0;
    intr[j] = false;
    anc[j] = false;
    yv[j] = 0.f;
    ttab[j] = 0;
    if (j >= np) continue;
    const uint32_t e = P[j];
    int c[ND];
    coords_of<ND>(g, e - F.base, c);
    tp[j] = t[e];
    intr[j] = interior<ND>(g, c);
    anc[j] = anchor && anchored(tp[j], db);
    yv[j] = y[e];
    if (!intr[j]) continue;
    // the other two plateau indices of j, in increasing order
    const int oa = j == 0 ? 1 : 0, ob = j == 2 ? 1 : 2;
    int code[K];  // +-1: fixed sign of (point - neighbour); 2: other point oa; 3: other point ob
    const uint32_t rm = rms[j];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const uint32_t s = stride_of<ND>(g, k >> 1);
      const uint32_t qe = (k & 1) ? e + s : e - s;
      // a same-value neighbour of a plateau of at most three points is one of them
      code[k] = (oa < np && qe == P[oa]) ? 2 : (ob < np && qe == P[ob]) ? 3 : (((rm >> (8 + k)) & 1u) ? 1 : -1);
    }
#pragma unroll
    for (int ra = -1; ra <= 1; ra++) {
#pragma unroll
      for (int rb = -1; rb <= 1; rb++) {
        bool all_gt = true, all_lt = true, any_hi = false, any_lo = false;
#pragma unroll
        for (int a = 0; a < ND; a++) {
          const int c0 = code[2 * a], c1 = code[2 * a + 1];
          const int r0 = c0 == 2 ? ra : c0 == 3 ? rb : c0;
          const int r1 = c1 == 2 ? ra : c1 == 3 ? rb : c1;
          const bool hi2 = r0 > 0 && r1 > 0, lo2 = r0 < 0 && r1 < 0;
          all_gt = all_gt && hi2;
          all_lt = all_lt && lo2;
          any_hi = any_hi || hi2;
          any_lo = any_lo || lo2;
        }
        const uint32_t k = all_gt ? 1u : (all_lt ? 2u : ((any_hi && any_lo) ? 3u : 0u));
        ttab[j] |= k << (2 * (3 * (ra + 1) + (rb + 1)));
      }
    }
  }
  float lo, hi;
  cell_of(db, eb, &lo, &hi);
  const int nord = np == 2 ? 3 : 13;
  for (int o = 0; o < nord && !done; o++) {
    int rk[3];
#pragma unroll
    for (int j = 0; j < 3; j++) rk[j] = j < np ? (np == 2 ? kWeak2[o][j] : kWeak3[o][j]) : 0;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j >= np || !intr[j]) continue;
      const int oa = j == 0 ? 1 : 0, ob = j == 2 ? 1 : 2;
      const int ra = rk[j] > rk[oa] ? 1 : (rk[j] < rk[oa] ? -1 : 0);
      const int rb = ob < np ? (rk[j] > rk[ob] ? 1 : (rk[j] < rk[ob] ? -1 : 0)) : 0;
      ok = ok && (int)((ttab[j] >> (2 * (3 * (ra + 1) + (rb + 1)))) & 3u) == tp[j];
    }
    if (!ok) continue;
    // values: classes holding an anchor keep its value, the others are spaced
    // evenly between their neighbours (or the interval ends)
    int nr = 0;
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (j < np) nr = max(nr, rk[j] + 1);
    float cv[3] = {0.f, 0.f, 0.f};
    bool ca[3] = {false, false, false};
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j >= np || !anc[j]) continue;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (k != rk[j]) continue;
        if (ca[k] && cv[k] != yv[j]) ok = false;
        ca[k] = true;
        cv[k] = yv[j];
      }
    }
    if (!ok) continue;
    float fv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      fv[k] = cv[k];
      if (k >= nr || ca[k]) continue;
      int pa = -1, na = nr;  // nearest anchored classes below and above
#pragma unroll
      for (int m = 0; m < 3; m++) {
        if (m < nr && ca[m] && m < k) pa = m;
        if (m < nr && ca[m] && m > k && na == nr) na = m;
      }
      const float L = pa >= 0 ? (pa == 0 ? cv[0] : (pa == 1 ? cv[1] : cv[2])) : lo;
      const float U = na < nr ? (na == 0 ? cv[0] : (na == 1 ? cv[1] : cv[2])) : hi;
      fv[k] = L + (U - L) * ((float)(k - pa) / (float)(na - pa));
    }
#pragma unroll
    for (int k = 0; k + 1 < 3; k++) ok = ok && (k + 1 >= nr || fv[k] < fv[k + 1]);
    if (!ok) continue;
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (write && j < np && !anc[j]) y[P[j]] = rk[j] == 0 ? fv[0] : (rk[j] == 1 ? fv[1] : fv[2]);
    done = true;
  }
  // after a solve every point carries its type: the interior ones by the check above,
  // the others (classified as regular) where that is their type
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < 3; j++) m |= (j < np && (intr[j] || tp[j] == 0)) ? 1u << j : 0u;
  *matched = done ? m : 0u;
  return done;
}

// The plateau of every listed point, given up beyond three points; the thread of the
// plateau's smallest listed point solves it and writes the values.  Plateaus are
// disjoint and a solve writes only its own points' values, so no thread waits for
// another.  Every other listed point of the plateau runs the same (deterministic) solve
// without writing, to learn whether its own point ends up matched: its result depends
// on the decoded values, the types and the anchored points' values, none of which a
// solve writes (the other points' values it loads go unused).  Cells of different
// decoded values do not overlap, so no other point's type changes: keep[i] is the
// listed point if it stays mismatched, else kNoEntry.
constexpr uint32_t kNoEntry = 0xffffffffu;  // an empty slot of a sparse list
template <int ND, class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_tiny(GEO geo, const uint32_t* __restrict__ bad,
                                                          uint32_t nbad, const float* __restrict__ d,
                                                          const uint32_t* __restrict__ bm,
                                                          const uint8_t* __restrict__ t, float* __restrict__ y,
                                                          int anchor, uint32_t* __restrict__ keep,
                                                          uint32_t* __restrict__ part) {
  constexpr int K = 2 * ND;
  __shared__ uint32_t red[kWavesPerBlock];
  // every XCD takes one contiguous eighth of the list, in order: a list in index order
  // then keeps the plateaus' neighbourhoods in the L2 that loaded them
  const uint32_t chunk = xcd_chunk();
  const uint32_t i = chunk * kBlockThreads + threadIdx.x;
  const uint32_t b = i < nbad ? bad[i] : 0u;
  const FieldView F = geo.at(b);  // a plateau lies in one field
  const Grid& g = F.g;
  uint32_t P[3] = {b, b, b};
  uint32_t rms[3] = {0u, 0u, 0u};
  int np = 1;
  bool big = i >= nbad;
#pragma unroll
  for (int u = 0; u < 3; u++) {
    if (u >= np || big) continue;
    rms[u] = rel_of<ND>(g, F.base, d, P[u]);
    const uint32_t eq = rms[u] & ((1u << K) - 1u);  // same-value neighbours
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (!((eq >> k) & 1u)) continue;
      const uint32_t s = stride_of<ND>(g, k >> 1);
      const uint32_t qe = (k & 1) ? P[u] + s : P[u] - s;
      const bool seen = qe == P[0] || (np > 1 && qe == P[1]) || (np > 2 && qe == P[2]);
      if (seen) continue;
      if (np == 3) {
        big = true;
      } else {
        if (np == 1) P[1] = qe;
        else P[2] = qe;
        np++;
      }
    }
  }
  // the plateau goes to the thread of its smallest mismatched point, which solves it
  const bool solvable = !big && np >= 2;
  bool take = solvable;
  if (take && np > 1 && P[1] < b && ((bm[P[1] >> 5] >> (P[1] & 31u)) & 1u)) take = false;
  if (take && np > 2 && P[2] < b && ((bm[P[2] >> 5] >> (P[2] & 31u)) & 1u)) take = false;
  uint32_t matched = 0;
  const bool done = solvable && tiny_solve_one<ND>(geo, P, np, d, rms, t, y, anchor, take, &matched);
  if (i < nbad) keep[i] = done && (matched & 1u) ? kNoEntry : b;  // P[0] is the thread's point
  const uint32_t nd = (uint32_t)__popcll(__ballot(done && take));
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = nd;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}


// ---------------------------------------------------------------------------------
// exact window solve
// ---------------------------------------------------------------------------------
// wave maximum of a non-negative value (whole wave active)
__device__ __forceinline__ int wave_max7(int v) {
  // DPP maximum over each 16-lane row, then over the four rows
  const uint32_t r = seg_max_dpp<16>((uint32_t)v);
  const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)r, 0), b = (uint32_t)__builtin_amdgcn_readlane((int)r, 16);
  const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)r, 32), e = (uint32_t)__builtin_amdgcn_readlane((int)r, 48);
  return (int)max(max(a, b), max(c, e));
}
// value of v in lane `src` (wave-uniform), through a scalar register
__device__ __forceinline__ int rdl(int v, int src) { return __builtin_amdgcn_readlane(v, src); }
__device__ __forceinline__ float rdlf(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
constexpr int kTileShift = 4;  // 16-wide tiles: windows of one parity class never interact
constexpr int kWinMax = 8;
constexpr int kSolveBudget = 1 << 12;
// search nodes of one tile's windows beyond which its remaining windows are not tried: the
// one-launch solve lasts as long as its longest chain of dependent tiles, and the windows
// that exhaust their budgets lie on plateaus the whole-plateau stage orders
constexpr uint32_t kTileNodeCap = 1u << 14;

template <int ND>
__device__ __forceinline__ uint32_t tile_of(const int (&c)[ND], uint32_t ty, uint32_t tz) {
  uint32_t id = 0;
#pragma unroll
  for (int a = 0; a < ND; a++) {
    const uint32_t tc = (uint32_t)c[a] >> kTileShift;
    id = a == 0 ? tc : id * (a == 1 ? ty : tz) + tc;
  }
  return id;
}

template <int ND, class GEO>
__global__ __launch_bounds__(kBlockThreads) void k_pp_tiles(GEO geo, const uint32_t* __restrict__ bad, uint32_t nbad,
                                                           uint32_t* pbits,
                                                           uint32_t* tbits, uint32_t* __restrict__ tiles,
                                                           uint32_t cap, uint32_t* misc) {
  // one tile list per parity phase: tiles + phase * cap
  const uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  bool take = false;
  uint32_t id = 0;
  int ph = 0;
  if (i < nbad) {
    const uint32_t e = bad[i];
    const FieldView F = geo.at(e);
    int c[ND];
    coords_of<ND>(F.g, e - F.base, c);
    atomicOr(&pbits[e >> 5], 1u << (e & 31u));
    id = F.tile0 + tile_of<ND>(c, F.ty, F.tz);
#pragma unroll
    for (int a = 0; a < ND; a++) ph |= (int)(((uint32_t)c[a] >> kTileShift) & 1u) << a;
    take = claim_bit(tbits, id);
  }
#pragma unroll
  for (int p = 0; p < (1 << ND); p++)
    wave_append(take && ph == p, id, tiles + (size_t)p * cap, cap, &misc[kPpTilesP + p]);
}

// Window of one mismatch (LDS, one per wavefront).  Elements: the window's plateau
// points, then the plateau points next to them.  Every interior element is a
// constraint; an attempt frees some window points (every other element keeps its
// value) and checks the constraints that read a free point.
template <int ND>
struct Window {
  static constexpr int K = 2 * ND;
  static constexpr int EM = kWinMax * (K + 1);
  static constexpr int kGt = 125, kLt = 126, kEq = 127;  // fixed relations (centre vs neighbour)
  uint32_t el[EM];
  float val[EM];
  int8_t cnb[EM][K];  // neighbour codes of an interior element: element index or kGt / kLt / kEq
  int8_t typ[EM];     // target type, -1 for a boundary element
  int8_t anch[EM];    // anchored extremum
  int8_t picks[kWinMax];
  float cv[EM];      // realisation: value of each class
  int8_t canchor[EM];
  uint32_t cand[kWave];
  int ne, nw, npick;
  uint32_t base;  // flat offset of the window's field
  float lo, hi, db;
};

// neighbour k of element point e, if it lies in the plateau of value db
template <int ND>
__device__ __forceinline__ uint32_t plateau_nbr(const Grid& g, uint32_t base, const float* __restrict__ d, uint32_t e,
                                                int k, float db) {
  int c[ND];
  coords_of<ND>(g, e - base, c);
  const int a = k >> 1;
  c[a] += (k & 1) ? 1 : -1;
  if (!inside<ND>(g, c)) return 0xffffffffu;
  const uint32_t s = stride_of<ND>(g, a);
  const uint32_t qe = (k & 1) ? e + s : e - s;
  return d[qe] == db ? qe : 0xffffffffu;
}

// Appends the candidates q of lanes [0, ncand) that are neither elements of the window
// (lane i holds element i in el_reg, ne elements) nor repeats of a lower lane's
// candidate, in lane order, up to cap elements; returns the new element count.  The
// same list as appending one candidate at a time with a linear search.
template <int ND>
__device__ __forceinline__ int window_append(Window<ND>& w, uint32_t& el_reg, int ne, uint32_t q, int ncand, int cap,
                                             int lane) {
  bool fresh = lane < ncand && q != 0xffffffffu;
  for (int i = 0; i < ne; i++) fresh = fresh && (uint32_t)rdl((int)el_reg, i) != q;
  for (int j = 0; j < ncand; j++) {
    const uint32_t v = (uint32_t)rdl((int)q, j);
    fresh = fresh && !(j < lane && v == q);
  }
  const uint64_t m = __ballot(fresh);
  const int room = cap - ne;
  const int rank = __popcll(m & lanemask_lt(lane));
  if (fresh && rank < room) w.el[ne + rank] = q;
  __syncthreads();
  const int add = min(__popcll(m), room);
  if (lane >= ne && lane < ne + add) el_reg = w.el[lane];
  return ne + add;
}

// Builds the window of mismatch b (whole wavefront: the lanes load neighbours in
// parallel; lane i keeps element i in a register, so membership is a wave-wide
// compare instead of a list search).
template <int ND, bool COH>
__device__ void build_window(Window<ND>& w, const Grid& g, uint32_t base, const float* __restrict__ d,
                             const uint8_t* __restrict__ t, const float* y, float eb, uint32_t b,
                             bool anchor, int lane) {
  constexpr int K = 2 * ND;
  static_assert(Window<ND>::EM <= kWave, "one window element per lane");
  const float db = d[b];
  uint32_t el_reg = lane == 0 ? b : 0xffffffffu;
  if (lane == 0) {
    w.el[0] = b;
    w.base = base;
    w.db = db;
    cell_of(db, eb, &w.lo, &w.hi);
  }
  __syncthreads();
  // breadth-first levels over the plateau, up to kWinMax points
  int ne = 1, head = 0;
  while (head < ne && ne < kWinMax) {
    const int ne0 = ne;
    const int ncand = (ne0 - head) * K;
    const uint32_t q =
        lane < ncand ? plateau_nbr<ND>(g, base, d, w.el[head + lane / K], lane % K, db) : 0xffffffffu;
    ne = window_append<ND>(w, el_reg, ne, q, ncand, kWinMax, lane);
    head = ne0;
  }
  const int nw = ne;
  // the plateau points next to the window keep their values
  {
    const int ncand = nw * K;
    const uint32_t q = lane < ncand ? plateau_nbr<ND>(g, base, d, w.el[lane / K], lane % K, db) : 0xffffffffu;
    ne = window_append<ND>(w, el_reg, ne, q, ncand, Window<ND>::EM, lane);
  }
  if (lane == 0) {
    w.nw = nw;
    w.ne = ne;
  }
  if (lane < ne) {
    const uint32_t e = el_reg;
    w.val[lane] = y_ld<COH>(y, e);
    const int tt = t[e];
    w.anch[lane] = anchor && anchored(tt, db);
    int c[ND];
    coords_of<ND>(g, e - base, c);
    w.typ[lane] = interior<ND>(g, c) ? (int8_t)tt : (int8_t)-1;
  }
  __syncthreads();
  // neighbour codes: (element, direction) pairs j = lane + kWave * it; every load of
  // every pair is issued before the first is used (one memory round trip, not one per
  // pair and load)
  constexpr int kIt = (Window<ND>::EM * K + kWave - 1) / kWave;
  int jj_r[kIt];
  uint32_t qe_r[kIt];
  float dq_r[kIt], yq_r[kIt];
#pragma unroll
  for (int it = 0; it < kIt; it++) {
    const int j = it * kWave + lane;
    const bool on = j < ne * K;
    const int i = on ? j / K : 0, k = on ? j % K : 0;
    const bool con = on && w.typ[i] >= 0;
    const uint32_t e = w.el[i];
    const uint32_t s = stride_of<ND>(g, k >> 1);
    qe_r[it] = (k & 1) ? e + s : e - s;
    int jj = -1;
    for (int x = 0; x < ne; x++)
      if ((uint32_t)rdl((int)el_reg, x) == qe_r[it] && jj < 0) jj = x;
    jj_r[it] = con ? jj : -2;  // -2: no code to write
  }
#pragma unroll
  for (int it = 0; it < kIt; it++) {
    const bool ld = jj_r[it] == -1;
    dq_r[it] = ld ? d[qe_r[it]] : 0.f;
    yq_r[it] = ld ? y_ld<COH>(y, qe_r[it]) : 0.f;
  }
#pragma unroll
  for (int it = 0; it < kIt; it++) {
    const int j = it * kWave + lane;
    const int i = j / K, k = j % K;
    if (jj_r[it] >= 0) {
      w.cnb[i][k] = (int8_t)jj_r[it];
    } else if (jj_r[it] == -1) {
      // not an element: another cell (the order of the cells) or a point that keeps its
      // value next to an element that keeps its value
      const float dq = dq_r[it];
      const float r = dq != db ? db - dq : w.val[i] - yq_r[it];
      w.cnb[i][k] =
          r > 0.f ? (int8_t)Window<ND>::kGt : (r < 0.f ? (int8_t)Window<ND>::kLt : (int8_t)Window<ND>::kEq);
    }
  }
  __syncthreads();
}

// One attempt (whole wavefront).  mode 0: the anchored extrema keep their values; 1:
// window element `pick` (an anchored extremum) moves too, within (v - delta, v + delta),
// which keeps it between the extrema of the neighbouring sort positions; 2: `pick`
// moves freely; 3: every window point moves.
//
// Depth-first search over the weak orders of the free points: lane i holds the class
// (position in the order) of element i, so inserting a class is one vector compare-
// and-increment, and the constraints that become decidable at a depth are checked by
// their own lanes with K lane shuffles.  A complete order becomes values (classes of
// fixed elements keep their value, runs of free classes are spaced evenly between
// them), is written to the field and verified there by the constraint lanes; the first
// order that holds is kept.  All control flow is uniform over the wavefront.
template <int ND, bool COH>
__device__ bool run_attempt(Window<ND>& w, const Grid& g, const uint8_t* __restrict__ t, float* y,
                            int mode, int pick, float delta, int lane, uint32_t* nodes_out, uint32_t* stats) {
  constexpr int K = 2 * ND;
  const int ne = w.ne, nw = w.nw;
  const bool is_el = lane < ne;
  const float my_val = is_el ? w.val[lane] : 0.f;
  const bool my_free = is_el && lane < nw && (!w.anch[lane] || mode == 3 || lane == pick);
  const uint64_t fmask = __ballot(my_free);
  const int nf = __popcll(fmask);
  if (nf == 0) return false;
  const int my_fidx = my_free ? __popcll(fmask & lanemask_lt(lane)) : -1;
  int felem_reg = 0;  // lane f: element index of free point f
  {
    uint64_t m = fmask;
    for (int f = 0; m; f++) {
      const int i = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      if (lane == f) felem_reg = i;
    }
  }
  // classes of the fixed elements: rank among their distinct values
  const bool my_fixed = is_el && !my_free;
  const uint64_t xmask = __ballot(my_fixed);
  bool my_first = my_fixed;
  for (int j = 0; j < ne; j++) {
    const float vj = rdlf(my_val, j);
    if (((xmask >> j) & 1ull) && j < lane && vj == my_val) my_first = false;
  }
  const uint64_t first_mask = __ballot(my_first);
  int my_cls = -1;
  {
    int rank = 0;
    for (int j = 0; j < ne; j++) {
      const float vj = rdlf(my_val, j);
      if (((first_mask >> j) & 1ull) && vj < my_val) rank++;
    }
    if (my_fixed) my_cls = rank;
  }
  int ncls = __popcll(first_mask);
  const float my_flo = (mode == 1 && lane == pick) ? my_val - delta : -INFINITY;
  const float my_fhi = (mode == 1 && lane == pick) ? my_val + delta : INFINITY;
  // constraints: every interior element, checked at the depth of its last free point
  const bool my_con = is_el && w.typ[lane] >= 0;
  const int my_typ = my_con ? w.typ[lane] : 0;
  int code[K];
#pragma unroll
  for (int k = 0; k < K; k++) code[k] = my_con ? w.cnb[lane][k] : Window<ND>::kGt;
  int my_dep = my_con ? my_fidx : -1;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int src = code[k] < Window<ND>::kGt ? code[k] : lane;
    const int fj = __shfl(my_fidx, src, kWave);
    if (my_con && code[k] < Window<ND>::kGt && fj > my_dep) my_dep = fj;
  }
  const int budget = mode == 0 ? kSolveBudget : kSolveBudget / 4;
  int depth = 0, nodes = 0;
  uint32_t leaves = 0, vfails = 0;
  // lane d: iteration count, new-class flag and starting choice of depth d
  int choice_reg = -1, newcls_reg = 0, home_reg = 0;
  bool found = false;
  float my_fval = 0.f, my_old = 0.f;
  while (true) {
    bool leaf = false;
    const uint64_t tdfs = wall_clock64();
    while (depth >= 0) {
      const int fe = rdl(felem_reg, depth);
      int it = rdl(choice_reg, depth);
      if (it >= 0) {  // undo the previous choice at this depth
        const int pos = rdl(my_cls, fe);
        const int nc = rdl(newcls_reg, depth);
        if (lane == fe) my_cls = -1;
        if (nc) {
          if (my_cls > pos) my_cls--;
          ncls--;
        }
      } else {
        // starting choice: the class of an equal current value, else a new class just
        // above every placed element of smaller current value
        const float vf = rdlf(my_val, fe);
        const int above = wave_max7(my_cls >= 0 && my_val < vf ? my_cls + 1 : 0);
        const int eqc = wave_max7(my_cls >= 0 && my_val == vf ? my_cls + 1 : 0) - 1;
        const int hm = eqc >= 0 ? 2 * eqc + 1 : 2 * above;
        if (lane == depth) home_reg = hm;
      }
      it++;
      if (lane == depth) choice_reg = it;
      const int maxch = 2 * ncls;
      if (it > maxch) {
        if (lane == depth) choice_reg = -1;
        depth--;
        continue;
      }
      if (++nodes > budget) {
        depth = -1;
        break;
      }
      // choices in the order home, home + 1, home - 1, home + 2, ... over [0, 2 ncls]
      int ch;
      {
        const int home = rdl(home_reg, depth);
        const int nlo = home, nhi = maxch - home, mm = min(nlo, nhi);
        if (it <= 2 * mm) ch = (it & 1) ? home + ((it + 1) >> 1) : home - (it >> 1);
        else ch = nhi > nlo ? home + (it - mm) : home - (it - mm);
      }
      const int pos = ch >> 1;
      if ((ch & 1) == 0) {  // a new class at position pos
        if (my_cls >= pos) my_cls++;
        if (lane == fe) my_cls = pos;
        ncls++;
        if (lane == depth) newcls_reg = 1;
      } else {  // join class pos
        if (lane == fe) my_cls = pos;
        if (lane == depth) newcls_reg = 0;
      }
      // every placed constraint: fully decided ones must give their type; partly
      // decided maxima, minima and saddles are refuted as soon as possible
      bool okl = true;
      {
        int r[K];
        bool kn[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int src = code[k] < Window<ND>::kGt ? code[k] : lane;
          const int oc = __shfl(my_cls, src, kWave);
          kn[k] = code[k] >= Window<ND>::kGt || oc >= 0;
          r[k] = code[k] == Window<ND>::kGt ? 1
                 : code[k] == Window<ND>::kLt ? -1
                 : code[k] == Window<ND>::kEq ? 0
                 : (my_cls > oc ? 1 : (my_cls < oc ? -1 : 0));
        }
        if (my_con && my_cls >= 0) {
          bool all_kn = true, all_gt = true, all_lt = true, any_hi = false, any_lo = false;
          bool can_hi[ND], can_lo[ND];
#pragma unroll
          for (int a = 0; a < ND; a++) {
            const int r0 = r[2 * a], r1 = r[2 * a + 1];
            const bool k0 = kn[2 * a], k1 = kn[2 * a + 1];
            all_kn = all_kn && k0 && k1;
            const bool hi = r0 > 0 && r1 > 0, lo = r0 < 0 && r1 < 0;
            all_gt = all_gt && hi;
            all_lt = all_lt && lo;
            any_hi = any_hi || hi;
            any_lo = any_lo || lo;
            can_hi[a] = (!k0 || r0 > 0) && (!k1 || r1 > 0);
            can_lo[a] = (!k0 || r0 < 0) && (!k1 || r1 < 0);
          }
          if (all_kn) {
            const int kk = all_gt ? 1 : (all_lt ? 2 : ((any_hi && any_lo) ? 3 : 0));
            okl = kk == my_typ;
          } else if (my_typ == 1 || my_typ == 2) {
            bool cand = true;
#pragma unroll
            for (int a = 0; a < ND; a++) cand = cand && (my_typ == 1 ? can_hi[a] : can_lo[a]);
            okl = cand;
          } else if (my_typ == 3) {
            bool cand = false;
#pragma unroll
            for (int a = 0; a < ND; a++)
#pragma unroll
              for (int b2 = 0; b2 < ND; b2++) cand = cand || (a != b2 && can_hi[a] && can_lo[b2]);
            okl = cand;
          }
        }
      }
      if (__any(!okl)) continue;
      if (depth + 1 < nf) {
        depth++;
        if (lane == depth) choice_reg = -1;
        continue;
      }
      // a complete order: values of the classes.  Lane k holds class k; classes with a
      // fixed element keep its value, the others are spaced evenly between the nearest
      // fixed classes (or the interval ends).
      leaves++;
      if (lane < ncls) w.canchor[lane] = 0;
      __syncthreads();
      if (my_fixed) {
        w.cv[my_cls] = my_val;
        w.canchor[my_cls] = 1;
      }
      __syncthreads();
      const bool is_cls = lane < ncls;
      const bool anc_k = is_cls && w.canchor[lane];
      float v_k = anc_k ? w.cv[lane] : 0.f;
      const uint64_t am = __ballot(anc_k);
      const uint64_t below = am & lanemask_lt(lane);
      const uint64_t above = lane < kWave - 1 ? am & (~0ull << (lane + 1)) : 0ull;
      const int pa = below ? 63 - __clzll((long long)below) : -1;
      const int na = above ? __ffsll((unsigned long long)above) - 1 : ncls;
      const float vpa = __shfl(v_k, pa >= 0 ? pa : 0, kWave);
      const float vna = __shfl(v_k, na < ncls ? na : 0, kWave);
      if (is_cls && !anc_k) {
        const float L = pa >= 0 ? vpa : w.lo;
        const float U = na < ncls ? vna : w.hi;
        v_k = L + (U - L) * ((float)(lane - pa) / (float)(na - pa));
      }
      const float v_next = __shfl(v_k, lane + 1 < kWave ? lane + 1 : lane, kWave);
      bool vok = !(is_cls && lane + 1 < ncls) || v_k < v_next;
      my_fval = __shfl(v_k, my_cls >= 0 ? my_cls : 0, kWave);
      if (my_free) vok = vok && my_fval > my_flo && my_fval < my_fhi;
      __syncthreads();
      if (!__any(!vok)) {
        leaf = true;
        break;
      }
    }
    if (lane == 0) stats[6] += (uint32_t)(wall_clock64() - tdfs);
    if (!leaf) break;
    // write the free points and check every constraint on the field
    if (my_free) {
      // the window's points hold their values of the build: no other wavefront writes
      // them, and every failed attempt puts them back
      my_old = w.val[lane];
      y_st<COH>(y, w.el[lane], my_fval);
    }
    if constexpr (COH) __builtin_amdgcn_s_waitcnt(0);  // the stores land before other lanes read
    __syncthreads();
    bool good = true;
    if (my_dep >= 0) {
      const uint32_t e = w.el[lane];
      int c[ND];
      coords_of<ND>(g, e - w.base, c);
      good = classify_at_y<ND, COH>(g, y, e, c) == (int)t[e];
    }
    if (!__any(!good)) {
      found = true;
      break;
    }
    if (my_free) y_st<COH>(y, w.el[lane], my_old);
    vfails++;
    if constexpr (COH) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  if (lane == 0) {
    *nodes_out += (uint32_t)nodes;
    stats[0] += leaves;
    stats[1] += vfails;
  }
  __syncthreads();
  return found;
}

// Solve the window of mismatch b; returns the mode that succeeded or -1.  An extremum
// whose original value lies on a cell boundary can carry the sort position of the
// neighbouring bin (the bin rounds x + eb in float, the cell truncates x / eb in
// double), which makes the anchored values of its window contradictory; modes 1-3
// free extrema, the most extreme anchored value first.
template <int ND, bool COH>
__device__ int solve_window(Window<ND>& w, const Grid& g, uint32_t base, const float* __restrict__ d,
                            const uint8_t* __restrict__ t,
                            float* y, float eb, uint32_t b, bool anchor, bool whole_only, float delta_frac,
                            int lane, uint32_t* nodes, uint32_t* stats) {
  const uint64_t t0 = wall_clock64();
  build_window<ND, COH>(w, g, base, d, t, y, eb, b, anchor, lane);
  const uint64_t t1 = wall_clock64();
  stats[4] += (uint32_t)(t1 - t0);
  if (whole_only && w.ne > w.nw) return -1;  // part of a larger plateau: left to the plateau solves
  const bool ok0 = run_attempt<ND, COH>(w, g, t, y, 0, -1, 0.f, lane, nodes, stats);
  stats[5] += (uint32_t)(wall_clock64() - t1);
  if (ok0) return 0;
  if (lane == 0) {
    // anchored window points, the most extreme value first
    const float mid = 0.5f * (w.lo + w.hi);
    w.npick = 0;
    for (int i = 0; i < w.nw; i++) {
      if (!w.anch[i]) continue;
      int j = w.npick++;
      while (j > 0 && fabsf(w.val[w.picks[j - 1]] - mid) < fabsf(w.val[i] - mid)) {
        w.picks[j] = w.picks[j - 1];
        j--;
      }
      w.picks[j] = (int8_t)i;
    }
  }
  __syncthreads();
  const int na = w.npick;
  if (na == 0) return -1;
  const float delta = delta_frac * (w.hi - w.lo);
  for (int r = 0; r < na; r++)
    if (run_attempt<ND, COH>(w, g, t, y, 1, w.picks[r], delta, lane, nodes, stats)) return 1;
  for (int r = 0; r < na; r++)
    if (run_attempt<ND, COH>(w, g, t, y, 2, w.picks[r], 0.f, lane, nodes, stats)) return 2;
  if (na > 1 && run_attempt<ND, COH>(w, g, t, y, 3, -1, 0.f, lane, nodes, stats)) return 3;
  return -1;
}

constexpr int kTileSide = 1 << kTileShift;

template <int ND>
__device__ __forceinline__ void tile_coords(uint32_t id, uint32_t ty, uint32_t tz, int (&tc)[ND]) {
  if constexpr (ND == 2) {
    tc[1] = (int)(id % ty);
    tc[0] = (int)(id / ty);
  } else {
    tc[2] = (int)(id % tz);
    id /= tz;
    tc[1] = (int)(id % ty);
    tc[0] = (int)(id / ty);
  }
}

// parity phase of a tile: tiles of one phase are two or more tiles apart, so their
// windows never meet
template <int ND>
__device__ __forceinline__ int tile_phase(const int (&tc)[ND]) {
  int ph = 0;
#pragma unroll
  for (int a = 0; a < ND; a++) ph |= (tc[a] & 1) << a;
  return ph;
}

template <int ND>
struct SolveShared {
  Window<ND> w;
  uint32_t blist[ND == 2 ? kTileSide * kTileSide : kTileSide * kTileSide * kTileSide];
  uint32_t nb;
  uint32_t tile;
  uint32_t wait[ND == 2 ? 9 : 27];  // the one-launch solve: neighbouring tiles to wait for
};

// One wavefront solves the listed mismatches of tile `id` (of field F, numbered within
// the field) in flat order; spmax is the field's largest sort position.
template <int ND, bool COH>
__device__ void solve_tile(SolveShared<ND>& sh, const FieldView& F, uint32_t id, uint32_t spmax,
                           const uint32_t* __restrict__ pbits, const float* __restrict__ d,
                           const uint8_t* __restrict__ t, float* y, int anchor, uint32_t* misc) {
  constexpr int kT = kTileSide;
  const Grid& g = F.g;
  const float eb = F.eb;
  Window<ND>& w = sh.w;
  uint32_t* blist = sh.blist;
  const int lane = threadIdx.x;
  const uint64_t t_work = wall_clock64();
  int tc[ND];
  tile_coords<ND>(id, F.ty, F.tz, tc);
  const int x0 = tc[0] * kT, x1 = min(x0 + kT, (int)g.d0);
  const int y0 = tc[1] * kT, y1 = min(y0 + kT, (int)g.d1);
  int z0 = 0, z1 = 1;
  if constexpr (ND == 3) {
    z0 = tc[2] * kT;
    z1 = min(z0 + kT, (int)g.d2);
  }
  // the tile's listed mismatches in flat order: row segments along the fastest axis
  constexpr int kSegs = ND == 2 ? kT : kT * kT;
  uint32_t nb = 0;
  for (int s0 = 0; s0 < kSegs; s0 += kWave) {
    const int sgi = s0 + lane;
    uint32_t bits = 0, e0 = 0;
    if (sgi < kSegs) {
      int x, len;
      if constexpr (ND == 2) {
        x = x0 + sgi;
        e0 = F.base + (uint32_t)x * g.d1 + (uint32_t)y0;
        len = y1 - y0;
      } else {
        x = x0 + sgi / kT;
        const int yy = y0 + sgi % kT;
        e0 = F.base + (uint32_t)x * g.plane + (uint32_t)yy * g.d2 + (uint32_t)z0;
        len = z1 - z0;
        if (yy >= y1) x = x1;
      }
      if (x < x1) {
        const uint64_t wv = (uint64_t)pbits[e0 >> 5] | ((uint64_t)pbits[(e0 >> 5) + 1] << 32);
        bits = (uint32_t)(wv >> (e0 & 31u)) & (uint32_t)((1ull << len) - 1ull);
      }
    }
    const uint32_t c = (uint32_t)__popc(bits);
    const uint32_t incl = seg_incl_scan<kWave>(c, lane);
    uint32_t pos = nb + incl - c;
    while (bits) {
      const int j = __ffs(bits) - 1;
      bits &= bits - 1u;
      blist[pos++] = e0 + (uint32_t)j;
    }
    nb += __shfl(incl, kWave - 1, kWave);
  }
  if (lane == 0) sh.nb = nb;
  __syncthreads();
  nb = sh.nb;
  uint32_t solved = 0, failed = 0, freed = 0, nodes = 0;
  uint32_t tst[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // [0] leaves [1] verify fails [4] build ticks [5] attempt ticks
  // extrema of the sort positions s - 1 and s + 1 sit 0.9 (hi - lo) / (S + 2) away
  const float delta_frac = 0.99f * 0.9f / ((float)spmax + 2.f);
  for (uint32_t j = 0; j < nb; j++) {
    const uint32_t e = blist[j];
    int c[ND];
    coords_of<ND>(g, e - F.base, c);
    if (classify_at_y<ND, COH>(g, y, e, c) == (int)t[e]) continue;  // fixed by an earlier window
    // (lane 0 keeps the count; the branch must be uniform over the wave)
    if ((uint32_t)__shfl((int)nodes, 0, kWave) >= kTileNodeCap) {  // left to the whole-plateau stage
      failed++;
      continue;
    }
    const int mode =
        solve_window<ND, COH>(w, g, F.base, d, t, y, eb, e, (anchor & 1) != 0, (anchor & 2) != 0, delta_frac, lane,
                              &nodes, tst);
    solved += mode >= 0;
    failed += mode < 0;
    freed += mode >= 2;
  }
  if (lane == 0) {
    if (solved) atomicAdd(&misc[kPpSolved], solved);
    if (failed) atomicAdd(&misc[kPpFailed], failed);
    if (freed) atomicAdd(&misc[kPpFreed], freed);
    atomicAdd(&misc[kPpStatNodes], nodes);
    atomicMax(&misc[kPpStatMaxNodes], nodes);
    atomicMax(&misc[kPpStatMaxWin], solved + failed);
    atomicMax(&misc[kPpStatMaxTicks], (uint32_t)(wall_clock64() - t_work));
    atomicAdd(&misc[kPpStatLeaves], tst[0]);
    atomicAdd(&misc[kPpStatLeaves + 1], tst[1]);
    atomicMax(&misc[kPpStatLeaves + 2], tst[4]);
    atomicMax(&misc[kPpStatLeaves + 3], tst[5]);
    atomicMax(&misc[kPpStatLeaves + 4], tst[6]);
    atomicAdd(&misc[kPpStatSumBuild], tst[4]);
    atomicAdd(&misc[kPpStatSumAttempt], tst[5]);
    atomicAdd(&misc[kPpStatSumTile], (uint32_t)(wall_clock64() - t_work));
  }
}

// One wavefront per listed tile of one parity phase (a launch per phase: the launch
// boundary makes one phase's values visible to the next).
template <int ND, class GEO>
__global__ __launch_bounds__(kWave) void k_pp_solve(GEO geo, const uint32_t* __restrict__ tiles, uint32_t cap,
                                                    int phase, const uint32_t* __restrict__ pbits,
                                                    const float* __restrict__ d, const uint8_t* __restrict__ t,
                                                    float* __restrict__ y, int anchor, uint32_t* misc) {
  __shared__ SolveShared<ND> sh;
  const uint32_t id = tiles[(size_t)phase * cap + blockIdx.x];
  const FieldView F = geo.of_tile(id);
  solve_tile<ND, false>(sh, F, id - F.tile0, geo.spmax(F, misc), pbits, d, t, y, anchor, misc);
}

// Every phase in one launch, as a dataflow: workgroup i takes the i-th listed tile in
// phase order (workgroups start in id order), and a tile starts once each neighbouring
// listed tile of a lower phase has finished (a done flag per tile, set after the
// tile's field stores have completed).
// The field is read and written with agent-coherent accesses, so no XCD-wide cache
// write-back or invalidation is needed around the flags.  Tiles two or more apart never
// interact, and every pair of neighbouring tiles runs in phase order, so the values
// equal those of one launch per phase; the time is the longest chain of dependent tiles
// instead of the sum of the phases' slowest tiles.  A tile only waits for tiles of lower
// ids, which have started before it, so the waits end; a bounded spin reports an error
// instead of hanging.
template <int ND, class GEO>
__global__ __launch_bounds__(kWave) void k_pp_solve_flow(GEO geo, const uint32_t* __restrict__ tiles, uint32_t cap,
                                                         const uint32_t* __restrict__ tbits, uint32_t* done,
                                                         const uint32_t* __restrict__ pbits,
                                                         const float* __restrict__ d, const uint8_t* __restrict__ t,
                                                         float* __restrict__ y, int anchor, uint32_t* misc,
                                                         uint32_t* counter) {
  __shared__ SolveShared<ND> sh;
  const int lane = threadIdx.x;
  if (lane == 0) sh.tile = tile_id(counter);
  __syncthreads();
  // phase and position of the handed-out index
  uint32_t i = sh.tile, p = 0;
  for (; p < (1u << ND); p++) {
    const uint32_t c = min(misc[kPpTilesP + p], cap);
    if (i < c) break;
    i -= c;
  }
  if (p == (1u << ND)) return;  // more workgroups than listed tiles
  const uint32_t id = tiles[(size_t)p * cap + i];  // over all fields
  const FieldView F = geo.of_tile(id);
  const uint32_t tx = F.tx, ty = F.ty, tz = F.tz, lid = id - F.tile0;
  int tc[ND];
  tile_coords<ND>(lid, ty, tz, tc);
  // lane k finds neighbour k of the 3^ND block (k != centre), if listed and of a lower
  // phase; lane 0 then waits for those one at a time (one poll in flight per waiting
  // tile: thousands of waiting tiles polling every neighbour flag at once flood the
  // memory system that the working tiles need)
  constexpr int kNb = ND == 2 ? 9 : 27;
  uint32_t nid = 0xffffffffu;
  if (lane < kNb && lane != kNb / 2) {
    int nc[ND];
    int k = lane;
    bool in = true;
#pragma unroll
    for (int a = ND - 1; a >= 0; a--) {
      nc[a] = tc[a] + (k % 3) - 1;
      k /= 3;
      const uint32_t lim = a == 0 ? tx : (a == 1 ? ty : tz);
      in = in && nc[a] >= 0 && (uint32_t)nc[a] < lim;
    }
    if (in && (uint32_t)tile_phase<ND>(nc) < p) {
      uint32_t q = (uint32_t)nc[0];
#pragma unroll
      for (int a = 1; a < ND; a++) q = q * (a == 1 ? ty : tz) + (uint32_t)nc[a];
      q += F.tile0;
      if ((tbits[q >> 5] >> (q & 31u)) & 1u) nid = q;
    }
  }
  const uint64_t wm = __ballot(nid != 0xffffffffu);
  if (nid != 0xffffffffu) sh.wait[__popcll(wm & lanemask_lt(lane))] = nid;
  __syncthreads();
  if (lane == 0) {
    const int nw = __popcll(wm);
    for (int q = 0; q < nw; q++) {
      const uint32_t id_q = sh.wait[q];
      uint32_t spins = 0;
      while (__hip_atomic_load(&done[id_q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
        if (++spins > kSpinLimit) {
          atomicOr(&misc[kPpErr], TSZP_ERR_LOOKBACK);
          break;
        }
        if (spins < 8) __builtin_amdgcn_s_sleep(2);
        else __builtin_amdgcn_s_sleep(64);  // about 4000 cycles
      }
    }
  }
  __syncthreads();
  solve_tile<ND, true>(sh, F, lid, geo.spmax(F, misc), pbits, d, t, y, anchor, misc);
  // the flag is a release store at agent scope: every field store of the tile is
  // performed before a neighbouring tile can see it set.  The waiting side reads the
  // field only with agent-coherent loads issued after its spin has seen the flag (the
  // loop waits for the flag's value and the workgroup barrier orders the loads after
  // it), so it needs no acquire, whose L2 invalidation would slow every window build.
  __syncthreads();
  if (lane == 0) __hip_atomic_store(&done[id], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------
// Whole-plateau solve for what the window solve leaves.  A plateau (connected points of
// one decoded value) orders its points freely inside the cell, and no other point's type
// depends on that order (every other neighbour is in another cell), so a plateau is
// solved on its own: its points are inserted one at a time in increasing value, and a
// point is insertable when its type with the inserted neighbours below it and the others
// above equals its target.  The search tries points in the order of a guide (a smoothed
// field, then the current values), and after each insertion rejects it when a
// neighbour could no longer reach its type whatever is inserted before it (forward
// checking); it backtracks within a node budget.  Anchored extrema keep their values and
// are inserted in the order of those values (ties back to back), and the other points
// take values spaced between them, so the sort-position order of a cell's extrema holds.
// One wavefront per listed mismatch; the plateau is claimed by its smallest point.
// ---------------------------------------------------------------------------------
// capacities: the small variant first, the large one (most of a CU's 160 KB of LDS) for
// the plateaus the small one could not hold
constexpr int kPlatSmall = 512, kPlatLarge = 3072;
constexpr int kPlatAnchored = 4, kPlatPasses = 4 + 4;  // insertion passes with anchors, in all
template <int ND, int PM>
struct PlatShared {
  static constexpr int kCap = PM;
  static constexpr int kHash = PM <= 512 ? 1024 : 8192;  // slots, a power of two >= 2 PM
  uint32_t el[PM];
  int16_t nb[PM][2 * ND];  // >= 0: plateau index; -1 lower cell; -2 higher cell or outside
  uint16_t sorted[PM];
  uint16_t order[PM], choice[PM];
  uint16_t alow[PM], aupto[PM];
  float av[PM];
  int8_t tgt[PM];  // target type, -1 off the interior
  int8_t anc[PM];
  uint32_t assigned[PM / 32];
  uint32_t hkey[kHash];  // membership while the plateau is gathered, then the guide keys
  uint16_t hval[kHash];
  uint32_t cand[kWave];
  int ne;
  int abort;
  float lo, hi, db;
};

template <int H>
__device__ __forceinline__ uint32_t plat_hash(uint32_t e) {
  constexpr int bits = H == 1024 ? 10 : 13;
  return (e * 2654435761u) >> (32 - bits);
}

template <class SH>
__device__ __forceinline__ int plat_find(const SH& sh, uint32_t e) {
  constexpr int H = SH::kHash;
  uint32_t h = plat_hash<H>(e);
  for (int k = 0; k < H; k++) {
    const uint32_t key = sh.hkey[h];
    if (key == e) return sh.hval[h];
    if (key == 0xffffffffu) return -1;
    h = (h + 1) & (H - 1);
  }
  return -1;
}

template <class SH>
__device__ __forceinline__ bool plat_assigned(const SH& sh, int i) {
  return (sh.assigned[i >> 5] >> (i & 31)) & 1u;
}

// type from "lower than the centre" flags of the 2*ND neighbours (axis a: 2a, 2a + 1)
template <int ND>
__device__ __forceinline__ int type_from_low(uint32_t low) {
  const uint32_t all = (1u << (2 * ND)) - 1u;
  if (low == all) return 1;
  if (low == 0) return 2;
  bool hi = false, lo = false;
#pragma unroll
  for (int a = 0; a < ND; a++) {
    const uint32_t m = (low >> (2 * a)) & 3u;
    hi = hi || m == 3u;
    lo = lo || m == 0u;
  }
  return (hi && lo) ? 3 : 0;
}

template <int ND, class SH>
__device__ __forceinline__ uint32_t low_now(const SH& sh, int i) {
  uint32_t low = 0;
#pragma unroll
  for (int k = 0; k < 2 * ND; k++) {
    const int q = sh.nb[i][k];
    if (q == -1 || (q >= 0 && plat_assigned(sh, q))) low |= 1u << k;
  }
  return low;
}

// whether element j can still take its type with some of its free neighbours inserted
// before it (whole wave: lane l tries subset l of the free neighbours)
template <int ND, class SH>
__device__ __forceinline__ bool plat_possible(const SH& sh, int j, int lane) {
  if (sh.tgt[j] < 0) return true;
  const uint32_t base = low_now<ND>(sh, j);
  int fk[2 * ND];
  int nf = 0;
#pragma unroll
  for (int k = 0; k < 2 * ND; k++) {
    const int q = sh.nb[j][k];
    fk[k] = -1;
    if (q >= 0 && !plat_assigned(sh, q)) fk[nf++] = k;
  }
  bool ok = false;
  if (lane < (1 << nf)) {
    uint32_t low = base;
    for (int b = 0; b < nf; b++)
      if ((lane >> b) & 1) low |= 1u << fk[b];
    ok = type_from_low<ND>(low) == sh.tgt[j];
  }
  return __ballot(ok) != 0;
}

// Annealing state over the order of a plateau: gk[i] is point i's key (order only; ties
// broken by index), and a point's type is read from which of its neighbours have the
// smaller (key, index).
template <int ND, class SH>
__device__ __forceinline__ bool plat_key_lower(const SH& sh, const float* gk, int q, int i) {
  return gk[q] < gk[i] || (gk[q] == gk[i] && q < i);
}
template <int ND, class SH>
__device__ __forceinline__ int plat_viol(const SH& sh, const float* gk, int i) {
  if (sh.tgt[i] < 0) return 0;
  uint32_t low = 0;
#pragma unroll
  for (int k = 0; k < 2 * ND; k++) {
    const int q = sh.nb[i][k];
    if (q == -1 || (q >= 0 && plat_key_lower<ND>(sh, gk, q, i))) low |= 1u << k;
  }
  return type_from_low<ND>(low) != sh.tgt[i];
}
// violations of point i and of its plateau neighbours: everything a move of i can change
template <int ND, class SH>
__device__ __forceinline__ int plat_local_viol(const SH& sh, const float* gk, int i) {
  int c = plat_viol<ND>(sh, gk, i);
#pragma unroll
  for (int k = 0; k < 2 * ND; k++) {
    const int q = sh.nb[i][k];
    if (q >= 0) c += plat_viol<ND>(sh, gk, q);
  }
  return c;
}
__device__ __forceinline__ uint64_t plat_mix(uint64_t h) {
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
constexpr int kAnnealCycle = 1000, kAnnealSweeps = 6 * kAnnealCycle;

// Annealing over the order of a plateau, when no anchored insertion pass got through.
// A point's type depends only on its order against its 2*ND neighbours, so the search
// moves one point at a time to another gap between its neighbours' keys (Metropolis
// acceptance on the number of violated types of it and its neighbours, temperature
// falling linearly over each cycle of kAnnealCycle sweeps, then reheated). Points of one
// colour, (x + 2y + 3z) mod 7 in 3D and (x + 2y) mod 5 in 2D, are at least 3 apart, so
// none reads what another of its colour writes and a colour moves in parallel. Anchored
// extrema keep their values (keys), so their order holds by construction.
// Returns whether every type holds; then sh.order holds the points in increasing key.
template <int ND, class SH>
__device__ bool plat_anneal(SH& sh, float* gk, int ne, const FieldView& F, const float* __restrict__ guide,
                            float db, int lane, uint32_t* sweeps_out) {
  constexpr int K = 2 * ND;
  const Grid& g = F.g;
  constexpr int NC = ND == 3 ? 7 : 5;
  // per point: the lower-neighbour flags and whether its type is violated, kept in step
  // with the keys (in the hash table's value slots, free once the plateau is gathered)
  uint8_t* msk = reinterpret_cast<uint8_t*>(sh.hval);
  uint8_t* vio = msk + SH::kCap;
  const float inv_eb = 1.f / F.eb;
  for (int i = lane; i < ne; i += kWave) {
    const uint32_t e = sh.el[i];
    const uint64_t hsh = plat_mix(((uint64_t)e << 8) ^ 0x5bd1e995ull);
    const float jit = sh.anc[i] ? 0.f : 1e-4f * ((float)(uint32_t)hsh * 2.3283064e-10f);
    const float v0 = sh.anc[i] || guide == nullptr ? sh.av[i] : guide[e];
    gk[i] = (v0 - db) * inv_eb + jit;
    int c[ND];
    coords_of<ND>(g, e - F.base, c);
    const int col = ND == 3 ? (c[0] + 2 * c[1] + 3 * c[2]) % 7 : (c[0] + 2 * c[1]) % 5;
    sh.choice[i] = (uint16_t)col;
  }
  __syncthreads();
  if (lane == 0) {  // points by colour: sh.sorted, colour starts in sh.alow[0..NC]
    for (int c = 0; c <= NC; c++) sh.alow[c] = 0;
    for (int i = 0; i < ne; i++) sh.alow[sh.choice[i] + 1]++;
    for (int c = 0; c < NC; c++) sh.alow[c + 1] += sh.alow[c];
    for (int c = 0; c < NC; c++) sh.aupto[c] = sh.alow[c];
    for (int i = 0; i < ne; i++) sh.sorted[sh.aupto[sh.choice[i]]++] = (uint16_t)i;
  }
  int cur = 0;
  for (int i = lane; i < ne; i += kWave) {
    uint32_t low = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = sh.nb[i][k];
      if (q == -1 || (q >= 0 && plat_key_lower<ND>(sh, gk, q, i))) low |= 1u << k;
    }
    const int v = sh.tgt[i] >= 0 && type_from_low<ND>(low) != sh.tgt[i];
    msk[i] = (uint8_t)low;
    vio[i] = (uint8_t)v;
    cur += v;
  }
  cur = wave_sum_i(cur);
  __syncthreads();
  int sw = 0;
  for (; sw < kAnnealSweeps && cur > 0; sw++) {
    const float temp = 0.7f * (1.f - (float)(sw % kAnnealCycle) / (float)kAnnealCycle) + 0.02f;
    int dsum = 0;
    for (int c = 0; c < NC; c++) {
      const int s1 = sh.alow[c + 1];
      for (int p = sh.alow[c] + lane; p < s1; p += kWave) {
        const int i = sh.sorted[p];
        if (sh.anc[i]) continue;
        int qk[K];
        float kq[K];
        float nk[K];
        int n = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
          qk[k] = sh.nb[i][k];
          kq[k] = qk[k] >= 0 ? gk[qk[k]] : 0.f;
          if (qk[k] >= 0) nk[n++] = kq[k];
        }
        if (!n) continue;
        for (int a = 1; a < n; a++) {  // insertion sort of at most 6 keys
          const float v = nk[a];
          int b2 = a - 1;
          while (b2 >= 0 && nk[b2] > v) {
            nk[b2 + 1] = nk[b2];
            b2--;
          }
          nk[b2 + 1] = v;
        }
        int gaps = 0;
        for (int a = 0; a + 1 < n; a++) gaps += nk[a + 1] > nk[a];
        const uint64_t r =
            plat_mix(((uint64_t)sh.el[i] << 20) ^ ((uint64_t)sw << 3) ^ (uint64_t)blockIdx.x * 0x9e3779b97f4a7c15ull);
        const float u1 = (float)(uint32_t)(r >> 40) * (1.f / 16777216.f);
        const float u2 = (float)(uint32_t)((r >> 16) & 0xffffffu) * (1.f / 16777216.f);
        const int pick = (int)((uint32_t)r & 0xffffu) % (gaps + 2);
        float cand;
        if (pick == 0) cand = nk[0] - 1e-3f * (0.5f + u1);
        else if (pick == gaps + 1) cand = nk[n - 1] + 1e-3f * (0.5f + u1);
        else {
          int a = 0, seen = 0;
          for (; a + 1 < n; a++)
            if (nk[a + 1] > nk[a] && ++seen == pick) break;
          cand = nk[a] + (nk[a + 1] - nk[a]) * (0.25f + 0.5f * u1);
        }
        // the move's effect on i and on each plateau neighbour q (which sees i in
        // direction k ^ 1)
        uint32_t low = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int q = qk[k];
          if (q == -1 || (q >= 0 && (kq[k] < cand || (kq[k] == cand && q < i)))) low |= 1u << k;
        }
        const int vi = sh.tgt[i] >= 0 && type_from_low<ND>(low) != sh.tgt[i];
        int dl = vi - (int)vio[i];
        uint8_t qm[K], qv[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int q = qk[k];
          qm[k] = 0;
          qv[k] = 0;
          if (q < 0) continue;
          const uint32_t bit = 1u << (k ^ 1);
          const uint32_t mq = ((low >> k) & 1u) ? (msk[q] & ~bit) : (msk[q] | bit);
          const int vq = sh.tgt[q] >= 0 && type_from_low<ND>(mq) != sh.tgt[q];
          qm[k] = (uint8_t)mq;
          qv[k] = (uint8_t)vq;
          dl += vq - (int)vio[q];
        }
        if (dl <= 0 || u2 < __expf(-(float)dl / temp)) {
          dsum += dl;
          gk[i] = cand;
          msk[i] = (uint8_t)low;
          vio[i] = (uint8_t)vi;
#pragma unroll
          for (int k = 0; k < K; k++)
            if (qk[k] >= 0) {
              msk[qk[k]] = qm[k];
              vio[qk[k]] = qv[k];
            }
        }
      }
      __syncthreads();
    }
    cur += wave_sum_i(dsum);
  }
  *sweeps_out = (uint32_t)sw;
  if (cur == 0) {
    int v = 0;  // recount from the keys: the kept flags are exact, this is their check
    for (int i = lane; i < ne; i += kWave) v += plat_viol<ND>(sh, gk, i);
    if (wave_sum_i(v) == 0) {
      for (int i = lane; i < ne; i += kWave) {
        const float gi = gk[i];
        int rk = 0;
        for (int j = 0; j < ne; j++) rk += (gk[j] < gi) || (gk[j] == gi && j < i);
        sh.order[rk] = (uint16_t)i;
      }
      __syncthreads();
      return true;
    }
  }
  return false;
}

template <int ND, class GEO, int PM>
__global__ __launch_bounds__(kWave) void k_pp_plateau(GEO geo, const uint32_t* __restrict__ bad, uint32_t nbad,
                                                      const float* __restrict__ d, const uint8_t* __restrict__ t,
                                                      float* __restrict__ y, const float* __restrict__ guide,
                                                      uint32_t* __restrict__ claim, uint32_t* misc) {
  constexpr int K = 2 * ND;
  using SH = PlatShared<ND, PM>;
  constexpr int H = SH::kHash;
  __shared__ SH sh;
  const int lane = threadIdx.x;
  if (blockIdx.x >= nbad) return;
  const uint32_t b = bad[blockIdx.x];
  const FieldView F = geo.at(b);
  const Grid& g = F.g;
  {
    int c[ND];
    coords_of<ND>(g, b - F.base, c);
    if (classify_at<ND>(g, y, b, c) == (int)t[b]) return;  // solved with another plateau
  }
  // ---- the plateau, breadth first (membership in an LDS hash table) ----------------
  for (int i = lane; i < H; i += kWave) sh.hkey[i] = 0xffffffffu;
  if (lane == 0) {
    sh.ne = 1;
    sh.abort = 0;
    sh.el[0] = b;
    sh.db = d[b];
    cell_of(sh.db, F.eb, &sh.lo, &sh.hi);
    const uint32_t h = plat_hash<H>(b);
    sh.hkey[h] = b;
    sh.hval[h] = 0;
  }
  __syncthreads();
  const float db = sh.db;
  int head = 0;
  while (true) {
    const int ne0 = sh.ne;
    if (head >= ne0 || sh.abort) break;
    for (int c0 = head * K; c0 < ne0 * K; c0 += kWave) {
      const int ci = c0 + lane;
      uint32_t q = 0xffffffffu;
      if (ci < ne0 * K) {
        q = plateau_nbr<ND>(g, F.base, d, sh.el[ci / K], ci % K, db);
        if (q != 0xffffffffu && plat_find(sh, q) >= 0) q = 0xffffffffu;
      }
      sh.cand[lane] = q;
      __syncthreads();
      if (lane == 0) {
        for (int j = 0; j < kWave && !sh.abort; j++) {
          const uint32_t e = sh.cand[j];
          if (e == 0xffffffffu || plat_find(sh, e) >= 0) continue;
          if (sh.ne >= PM) {
            sh.abort = 1;
            break;
          }
          uint32_t h = plat_hash<H>(e);
          while (sh.hkey[h] != 0xffffffffu) h = (h + 1) & (H - 1);
          sh.hkey[h] = e;
          sh.hval[h] = (uint16_t)sh.ne;
          sh.el[sh.ne++] = e;
        }
      }
      __syncthreads();
    }
    head = ne0;
  }
  if (sh.abort) {
    if (lane == 0) atomicAdd(&misc[kPpPlatBig], 1u);
    return;
  }
  const int ne = sh.ne;
  // claimed by its smallest point: one wavefront per plateau
  uint32_t mn = 0xffffffffu;
  for (int i = lane; i < ne; i += kWave) mn = min(mn, sh.el[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, kWave));
  if (lane == 0) sh.cand[0] = claim_bit(claim, mn) ? 1u : 0u;
  __syncthreads();
  if (!sh.cand[0]) return;
  // ---- neighbour table, targets, anchors ---------------------------------------------
  for (int i = lane; i < ne; i += kWave) {
    const uint32_t e = sh.el[i];
    int c[ND];
    coords_of<ND>(g, e - F.base, c);
    const bool in = interior<ND>(g, c);
    const int tt = t[e];
    sh.tgt[i] = in ? (int8_t)tt : (int8_t)-1;
    sh.anc[i] = (tt == 1 || tt == 2) && db != 0.f;
    sh.av[i] = y[e];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int a = k >> 1;
      int q[ND];
#pragma unroll
      for (int x = 0; x < ND; x++) q[x] = c[x];
      q[a] += (k & 1) ? 1 : -1;
      int16_t code = -2;
      if (inside<ND>(g, q)) {
        const uint32_t s = stride_of<ND>(g, a);
        const uint32_t qe = (k & 1) ? e + s : e - s;
        const float dq = d[qe];
        code = dq == db ? (int16_t)plat_find(sh, qe) : (dq < db ? (int16_t)-1 : (int16_t)-2);
      }
      sh.nb[i][k] = code;
    }
  }
  __syncthreads();
  for (int i = lane; i < ne; i += kWave) {
    uint16_t lo = 0, up = 0;
    if (sh.anc[i])
      for (int j = 0; j < ne; j++) {
        if (!sh.anc[j]) continue;
        lo += sh.av[j] < sh.av[i];
        up += sh.av[j] <= sh.av[i];
      }
    sh.alow[i] = lo;
    sh.aupto[i] = up;
  }
  __syncthreads();
  // Passes: the smoothed field and the current values as guides, then restarts on the
  // smoothed field with deterministic noise (a backtracking search's run time is heavy
  // tailed in its guide, so a few short searches beat one long one); first with the
  // anchored extrema in place, then annealing with them in place (what large plateaus
  // need: a backtracking search fails there on a few conflicts deep in its order), then
  // (an extremum on a cell boundary can carry the sort position of the neighbouring bin)
  // insertion passes with every point free.
  float* gk = reinterpret_cast<float*>(sh.hkey);  // the hash table is no longer needed
  const uint64_t t_solve = wall_clock64();
  bool solved = false, annealed = false;
  int pass = 0;
  for (; pass < kPlatPasses && !solved; pass++) {
    // a plateau beyond the small capacity goes to annealing after the two guided passes
    if (pass == 2 && ne > kPlatSmall && pass < kPlatAnchored) pass = kPlatAnchored;
    const int rp = pass < kPlatAnchored ? pass : pass - kPlatAnchored;  // pass within its group
    if (pass == kPlatAnchored) {
      uint32_t sweeps = 0;
      const uint64_t ta = wall_clock64();
      const bool ok = plat_anneal<ND>(sh, gk, ne, F, guide, db, lane, &sweeps);
      if (lane == 0) {
        atomicMax(&misc[kPpPlatMaxAnneal], (uint32_t)(wall_clock64() - ta));
        atomicMax(&misc[kPpPlatMaxSweeps], sweeps);
      }
      if (ok) {
        solved = annealed = true;
        break;
      }
      for (int i = lane; i < ne; i += kWave) sh.anc[i] = 0;
      __syncthreads();
    }
    // ---- guide ranks ------------------------------------------------------------------
    for (int i = lane; i < ne; i += kWave) {
      float gv = rp == 1 || guide == nullptr ? sh.av[i] : guide[sh.el[i]];
      if (rp >= 2) {  // two hashed uniforms: about 0.3 eb of spread
        uint64_t hsh = ((uint64_t)sh.el[i] << 8) ^ (uint64_t)(pass * 0x9e3779b9u);
        hsh ^= hsh >> 33;
        hsh *= 0xff51afd7ed558ccdull;
        hsh ^= hsh >> 33;
        hsh *= 0xc4ceb9fe1a85ec53ull;
        hsh ^= hsh >> 33;
        const float u0 = (float)(uint32_t)hsh * 2.3283064e-10f, u1 = (float)(uint32_t)(hsh >> 32) * 2.3283064e-10f;
        gv += F.eb * 1.2f * (u0 + u1 - 1.f);
      }
      gk[i] = gv;
    }
    __syncthreads();
    for (int i = lane; i < ne; i += kWave) {
      const float gi = gk[i];
      int r = 0;
      for (int j = 0; j < ne; j++) r += (gk[j] < gi) || (gk[j] == gi && j < i);
      sh.sorted[r] = (uint16_t)i;
    }
    for (int i = lane; i < PM / 32; i += kWave) sh.assigned[i] = 0u;
    __syncthreads();
    // ---- insertion search (uniform control flow over the wave) ------------------------
    int depth = 0, pos = -1, nanc = 0;
    uint32_t nodes = 0;
    const uint32_t budget = rp < 2 && ne <= kPlatSmall ? 8u * (uint32_t)ne + 2048u : 4u * (uint32_t)ne + 1024u;
    while (true) {
      if (depth == ne) {
        solved = true;
        break;
      }
      if (++nodes > budget) break;
      // tie pending: the last insertion was an anchor whose value other anchors share
      int tie = -1;
      if (depth > 0) {
        const int last = sh.order[depth - 1];
        if (sh.anc[last] && nanc < sh.aupto[last]) tie = sh.alow[last];
      }
      int found = -1, fpos = -1;
      for (int p0 = pos + 1; p0 < ne && found < 0; p0 += kWave) {
        const int p = p0 + lane;
        bool ok = false;
        int i = 0;
        if (p < ne) {
          i = sh.sorted[p];
          ok = !plat_assigned(sh, i);
          if (ok && tie >= 0) ok = sh.anc[i] && sh.alow[i] == tie;
          if (ok && sh.anc[i]) ok = nanc >= sh.alow[i];
          if (ok && sh.tgt[i] >= 0) ok = type_from_low<ND>(low_now<ND>(sh, i)) == sh.tgt[i];
        }
        const uint64_t m = __ballot(ok);
        if (m) {
          const int l = __ffsll((unsigned long long)m) - 1;
          fpos = p0 + l;
          found = sh.sorted[fpos];
        }
      }
      if (found >= 0) {
        if (lane == 0) sh.assigned[found >> 5] |= 1u << (found & 31);
        __syncthreads();
        bool alive = true;
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int j = sh.nb[found][k];
          if (j >= 0 && !plat_assigned(sh, j) && !plat_possible<ND>(sh, j, lane)) alive = false;
        }
        if (alive) {
          if (lane == 0) {
            sh.order[depth] = (uint16_t)found;
            sh.choice[depth] = (uint16_t)fpos;
          }
          nanc += sh.anc[found] ? 1 : 0;
          depth++;
          pos = -1;
        } else {
          if (lane == 0) sh.assigned[found >> 5] &= ~(1u << (found & 31));
          pos = fpos;
        }
        __syncthreads();
      } else {
        if (depth == 0) break;  // nothing insertable at all
        depth--;
        const int i = sh.order[depth];
        if (lane == 0) sh.assigned[i >> 5] &= ~(1u << (i & 31));
        nanc -= sh.anc[i] ? 1 : 0;
        pos = sh.choice[depth];
        __syncthreads();
      }
    }
  }
  if (lane == 0) atomicMax(&misc[kPpPlatMaxTicks], (uint32_t)(wall_clock64() - t_solve));
  if (!solved) {
    if (lane == 0) atomicAdd(&misc[kPpPlatFailed], 1u);
    return;
  }
  // ---- values: anchors keep theirs, the points between take spaced values -----------
  if (lane == 0) {
    int s0 = 0;
    float prev = sh.lo;
    for (int r = 0; r <= ne; r++) {
      const bool end = r == ne;
      const int i = end ? -1 : sh.order[r];
      if (end || sh.anc[i]) {
        const float nxt = end ? sh.hi : sh.av[i];
        const int cnt = r - s0;
        for (int k = 0; k < cnt; k++) {
          const int j = sh.order[s0 + k];
          y[sh.el[j]] = prev + (nxt - prev) * ((float)(k + 1) / (float)(cnt + 1));
        }
        if (!end) prev = sh.av[i];
        s0 = r + 1;
      }
    }
    atomicAdd(&misc[kPpPlatSolved], 1u);
    if (annealed) atomicAdd(&misc[kPpPlatAnneal], 1u);
    else if (pass > kPlatAnchored) atomicAdd(&misc[kPpFreed], 1u);  // solved with the anchors freed
  }
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
static inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + kBlockThreads - 1) / kBlockThreads); }
static inline unsigned grid_stride_blocks(uint64_t n) {
  const unsigned b = blocks_for(n);
  return b < 4096u ? (b ? b : 1u) : 4096u;
}

uint32_t pp_init_tiles(uint32_t n) { return (n + kInitTile - 1) / kInitTile; }


hipError_t pp_spmax(const PpGeo& geo, const int32_t* sp, uint32_t m, uint32_t* misc, hipStream_t st) {
  if (geo.fields) {  // every field's largest sort position
    hipError_t e = launch_fill_u32(geo.spmax, 0u, geo.nfields, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pp_spmax_b, dim3(geo.n >> kFlatShift), dim3(kBlockThreads), 0, st, geob(geo), sp,
                       geo.spmax);
    return hipGetLastError();
  }
  if (!m) return hipSuccess;
  const unsigned g = blocks_for((m + 3) / 4);
  hipLaunchKernelGGL(k_pp_spmax, dim3(g < 512u ? g : 512u), dim3(kBlockThreads), 0, st, sp, m, misc);
  return hipGetLastError();
}

// work: pp_init_tiles(n) + 1 counts, then scan_status_words(tiles + 1) zeroed u64 status
// words and a zeroed u32 counter and error word (pp_init_work_words in total)
size_t pp_init_work_words(uint32_t n) {
  const uint32_t nt = pp_init_tiles(n) + 1;
  return nt + 1 + 2 * scan_status_words(nt) + 2;
}

hipError_t pp_blur(const PpGeo& geo, const float* src, float* dst, int axis, const uint32_t* fflags,
                   hipStream_t st) {
  PP_GEO_LAUNCH(k_pp_blur, dim3(grid_stride_blocks(geo.n)), dim3(kBlockThreads), st, src, dst, geo.n, axis,
                geo.fields ? fflags : nullptr);
  return hipGetLastError();
}

// flags[f] = 1 for every field f that holds a listed point (flags zeroed first)
__global__ __launch_bounds__(kBlockThreads) void k_pp_field_flags(GeoB geo, const uint32_t* __restrict__ bad,
                                                                 uint32_t nbad, uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  if (i < nbad) flags[geo.at(bad[i]).fid] = 1u;
}

hipError_t pp_field_flags(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, uint32_t* flags, hipStream_t st) {
  if (!geo.fields) return hipErrorInvalidValue;
  hipError_t e = launch_fill_u32(flags, 0u, geo.nfields, st);
  if (e != hipSuccess || !nbad) return e;
  hipLaunchKernelGGL(k_pp_field_flags, dim3(blocks_for(nbad)), dim3(kBlockThreads), 0, st, geob(geo), bad, nbad,
                     flags);
  return hipGetLastError();
}

hipError_t pp_init(const PpGeo& geo, const float* d, const uint8_t* t, const int32_t* sp, uint32_t n_ext, float* y,
                   uint32_t* work, uint32_t* misc, hipStream_t st) {
  const uint32_t n = geo.n;
  const uint32_t tiles = pp_init_tiles(n);
  uint32_t* cnt = work;
  uint64_t* status = reinterpret_cast<uint64_t*>(work + ((tiles + 2) & ~1u));  // 8-byte aligned
  const size_t sw = scan_status_words(tiles + 1);
  uint32_t* ctr = reinterpret_cast<uint32_t*>(status + sw);
  hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(status), 0u, 2 * sw + 2, st);
  if (e != hipSuccess) return e;
  e = launch_fill_u32(cnt + tiles, 0u, 1, st);  // the scan turns it into the total
  if (e != hipSuccess) return e;
  if (geo.fields) hipLaunchKernelGGL(k_pp_ext_count<GeoB>, dim3(tiles), dim3(kBlockThreads), 0, st, geob(geo), t, cnt);
  else hipLaunchKernelGGL(k_pp_ext_count<Geo1>, dim3(tiles), dim3(kBlockThreads), 0, st, geo1(geo), t, cnt);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_scan_excl(cnt, tiles + 1, status, ctr, ctr + 1, st)) != hipSuccess) return e;
  if (geo.fields)
    hipLaunchKernelGGL(k_pp_init<GeoB>, dim3(tiles), dim3(kBlockThreads), 0, st, geob(geo), d, t, sp, n_ext, y, cnt,
                       misc);
  else
    hipLaunchKernelGGL(k_pp_init<Geo1>, dim3(tiles), dim3(kBlockThreads), 0, st, geo1(geo), d, t, sp, n_ext, y, cnt,
                       misc);
  return hipGetLastError();
}

constexpr unsigned kMarkBlocks = 2048;
uint32_t pp_mark_parts() { return kMarkBlocks; }

bool pp_mark_maps(const PpGeo& geo) { return geo.fields || (geo.ndim == 2 ? geo.g.d1 : geo.g.d2) % 4 == 0; }

hipError_t pp_mark(const PpGeo& geo, const float* y, const uint8_t* t, const float* d, uint32_t* bad, uint32_t cap,
                   uint32_t* part, uint32_t* bm, bool per_field, uint32_t* misc, hipStream_t st) {
  const uint32_t n = geo.n;
  const dim3 blk(kBlockThreads);
  if (geo.fields) {
    const uint32_t nchunks = (n + kMarkChunk - 1) / kMarkChunk;
    unsigned long long* fc = per_field ? geo.fcount : nullptr;
    if (fc) {
      hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(fc), 0u, 8 * (size_t)geo.nfields, st);
      if (e != hipSuccess) return e;
    }
    if (geo.ndim == 2)
      hipLaunchKernelGGL(k_pp_mark_b<2>, dim3(kMarkBlocks), blk, 0, st, geob(geo), y, t, d, nchunks, bad, cap, part,
                         bm, fc, misc);
    else
      hipLaunchKernelGGL(k_pp_mark_b<3>, dim3(kMarkBlocks), blk, 0, st, geob(geo), y, t, d, nchunks, bad, cap, part,
                         bm, fc, misc);
    hipLaunchKernelGGL(k_pp_reduce_verify, dim3(1), blk, 0, st, part, kMarkBlocks, misc);
    return hipGetLastError();
  }
  const Grid& g = geo.g;
  const uint32_t fast = g.ndim == 2 ? g.d1 : g.d2;
  const bool vec = fast % 4 == 0 && (reinterpret_cast<uintptr_t>(y) & 15) == 0 &&
                   (reinterpret_cast<uintptr_t>(t) & 3) == 0 && (reinterpret_cast<uintptr_t>(d) & 15) == 0;
  if (vec) {
    unsigned nb = kMarkBlocks;  // a multiple of the XCD count
    const dim3 grid(nb);
    if (g.ndim == 2) hipLaunchKernelGGL(k_pp_mark4<2>, grid, blk, 0, st, g, y, t, d, n / 4, bad, cap, part, bm, misc);
    else hipLaunchKernelGGL(k_pp_mark4<3>, grid, blk, 0, st, g, y, t, d, n / 4, bad, cap, part, bm, misc);
    hipLaunchKernelGGL(k_pp_reduce_verify, dim3(1), blk, 0, st, part, nb, misc);
    return hipGetLastError();
  }
  if (bm) return hipErrorInvalidValue;  // the mismatch map needs the vector layout
  const dim3 grid(grid_stride_blocks(n));
  if (g.ndim == 2) hipLaunchKernelGGL(k_pp_mark<2>, grid, blk, 0, st, g, y, t, d, n, bad, cap, misc);
  else hipLaunchKernelGGL(k_pp_mark<3>, grid, blk, 0, st, g, y, t, d, n, bad, cap, misc);
  return hipGetLastError();
}

hipError_t pp_active(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, const float* d, const uint8_t* t,
                     uint32_t* bits, uint32_t* act, uint32_t cap, int anchor, uint32_t* misc, hipStream_t st) {
  if (!nbad) return hipSuccess;
  PP_GEO_LAUNCH(k_pp_active, dim3(blocks_for(nbad)), dim3(kBlockThreads), st, bad, nbad, d, t, bits, act, cap, anchor,
                misc);
  return hipGetLastError();
}

hipError_t pp_move(const PpGeo& geo, const uint32_t* act, uint32_t max_act, uint32_t cap, int colour, const float* d,
                   const uint8_t* t, float* y, uint32_t* misc, hipStream_t st) {
  const uint32_t m = max_act < cap ? max_act : cap;
  if (!m) return hipSuccess;
  const dim3 blk(kBlockThreads);
  if (getenv("TSZP_MOVE_SERIAL")) {  // one lane per point
    PP_GEO_LAUNCH(k_pp_move, dim3(blocks_for(m)), blk, st, act, cap, colour, d, t, y, misc);
  } else {  // sixteen lanes per point
    PP_GEO_LAUNCH(k_pp_move_seg, dim3(blocks_for((uint64_t)m * 16)), blk, st, act, cap, colour, d, t, y, misc);
  }
  return hipGetLastError();
}

hipError_t pp_remark(const PpGeo& geo, const uint32_t* old, uint32_t nold, const float* y, const uint8_t* t,
                     uint32_t* bits, uint32_t* bad, uint32_t cap, int rings, uint32_t* misc, hipStream_t st) {
  if (!nold) return hipSuccess;
  const uint64_t R = rings == 1 ? 1 : (geo.ndim == 2 ? 13 : 25);
  const uint64_t items = (uint64_t)nold * R, per = (uint64_t)kBlockThreads * kRemarkChunks;
  PP_GEO_LAUNCH(k_pp_remark, dim3((unsigned)((items + per - 1) / per)), dim3(kBlockThreads), st, old, nold, y, t, bits,
                bad, cap, rings, misc);
  return hipGetLastError();
}

hipError_t pp_tiny(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, const float* d, const uint32_t* bm,
                   const uint8_t* t, float* y, int anchor, uint32_t* keep, uint32_t* part, uint32_t* misc,
                   hipStream_t st) {
  if (!nbad) return hipSuccess;
  const dim3 grid(pp_tiny_parts(nbad)), blk(kBlockThreads);
  PP_GEO_LAUNCH(k_pp_tiny, grid, blk, st, bad, nbad, d, bm, t, y, anchor, keep, part);
  hipLaunchKernelGGL(k_pp_reduce_sum, dim3(1), blk, 0, st, part, grid.x, misc, (int)kPpTiny);
  return hipGetLastError();
}

uint32_t pp_tiny_parts(uint32_t nbad) { return xcd_grid(blocks_for(nbad)); }

// mismatch bitmap from a list: bm must be zero; one bit per listed point
__global__ __launch_bounds__(kBlockThreads) void k_pp_badmap(const uint32_t* __restrict__ bad, uint32_t nbad,
                                                            uint32_t* __restrict__ bm) {
  const uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  if (i < nbad) {
    const uint32_t e = bad[i];
    atomicOr(&bm[e >> 5], 1u << (e & 31u));
  }
}

hipError_t pp_badmap(const uint32_t* bad, uint32_t nbad, uint32_t* bm, hipStream_t st) {
  if (!nbad) return hipSuccess;
  hipLaunchKernelGGL(k_pp_badmap, dim3(blocks_for(nbad)), dim3(kBlockThreads), 0, st, bad, nbad, bm);
  return hipGetLastError();
}

// The entries of a sparse list (kNoEntry: empty), compacted per workgroup: a workgroup
// takes kRelistChunk consecutive entries, collects the kept ones in LDS in list order
// and appends them with one atomic (same-address atomics serialise, so one per wave
// over millions of entries would cost more than the filter).
constexpr uint32_t kRelistChunk = 4096;
__global__ __launch_bounds__(kBlockThreads) void k_pp_relist(const uint32_t* __restrict__ bad, uint32_t nbad,
                                                            uint32_t* __restrict__ out, uint32_t cap,
                                                            uint32_t* misc) {
  __shared__ uint32_t kept[kRelistChunk];
  __shared__ uint32_t wcnt[kWavesPerBlock], sh_base, sh_n;
  const int tid = threadIdx.x, w = tid / kWave, lane = tid & (kWave - 1);
  const uint32_t c0 = xcd_chunk() * kRelistChunk;
  uint32_t n = 0;  // kept so far (uniform)
  for (uint32_t r = 0; r < kRelistChunk / kBlockThreads; r++) {
    const uint32_t i = c0 + r * kBlockThreads + tid;
    uint32_t e = 0;
    bool keep = false;
    if (i < nbad) {
      e = bad[i];
      keep = e != kNoEntry;
    }
    const uint64_t m = __ballot(keep);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
      before += k < w ? wcnt[k] : 0u;
      tot += wcnt[k];
    }
    if (keep) kept[n + before + (uint32_t)__popcll(m & lanemask_lt(lane))] = e;
    n += tot;
    __syncthreads();
  }
  if (tid == 0) {
    sh_n = n;
    sh_base = n ? atomicAdd(&misc[kPpBad], n) : 0u;
  }
  __syncthreads();
  const uint32_t base = sh_base;
  for (uint32_t k = tid; k < sh_n; k += kBlockThreads)
    if (base + k < cap) out[base + k] = kept[k];
}

uint32_t pp_compact_tiles(uint32_t n) { return (n + kCompactTile - 1) / kCompactTile; }

hipError_t pp_compact_map(const uint32_t* bm, uint32_t n, uint64_t* status, uint32_t* out, uint32_t cap,
                          uint32_t* misc, hipStream_t st) {
  // the map is read in 16-byte words up to n rounded up to 128 points: its buffer must
  // extend that far (the bits past n are masked off)
  if (reinterpret_cast<uintptr_t>(bm) & 15u) return hipErrorInvalidValue;
  const uint32_t tiles = pp_compact_tiles(n);
  hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(status), 0u, 2 * (size_t)tiles, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pp_compact_map, dim3(tiles), dim3(kBlockThreads), 0, st, reinterpret_cast<const uint4*>(bm), n,
                     status, out, cap, misc);
  return hipGetLastError();
}

hipError_t pp_relist(const uint32_t* sparse, uint32_t n, uint32_t* out, uint32_t cap, uint32_t* misc,
                     hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_pp_relist, dim3(xcd_grid((n + kRelistChunk - 1) / kRelistChunk)), dim3(kBlockThreads), 0, st,
                     sparse, n, out, cap, misc);
  return hipGetLastError();
}

void pp_tile_counts(const Grid& g, uint32_t* tx, uint32_t* ty, uint32_t* tz) {
  const uint32_t s = 1u << kTileShift;
  *tx = (g.d0 + s - 1) / s;
  *ty = (g.d1 + s - 1) / s;
  *tz = g.ndim == 3 ? (g.d2 + s - 1) / s : 1u;
}

hipError_t pp_tiles(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, uint32_t* pbits, uint32_t* tbits,
                    uint32_t* tiles, uint32_t cap, uint32_t* misc, hipStream_t st) {
  if (!nbad) return hipSuccess;
  PP_GEO_LAUNCH(k_pp_tiles, dim3(blocks_for(nbad)), dim3(kBlockThreads), st, bad, nbad, pbits, tbits, tiles, cap, misc);
  return hipGetLastError();
}

hipError_t pp_solve_flow(const PpGeo& geo, const uint32_t* tiles, uint32_t total, uint32_t cap, const uint32_t* tbits,
                         uint32_t* done, const uint32_t* pbits, const float* d, const uint8_t* t, float* y, int anchor,
                         uint32_t* misc, hipStream_t st) {
  if (!total) return hipSuccess;
  PP_GEO_LAUNCH(k_pp_solve_flow, dim3(total), dim3(kWave), st, tiles, cap, tbits, done, pbits, d, t, y, anchor, misc,
                dyn_counter(misc + kPpFlowCounter));
  return hipGetLastError();
}

#define PP_PLAT_LAUNCH(PMV)                                                                                   \
  do {                                                                                                         \
    if (geo.fields) {                                                                                          \
      const GeoB gb_ = geob(geo);                                                                              \
      if (geo.ndim == 2)                                                                                       \
        hipLaunchKernelGGL((k_pp_plateau<2, GeoB, PMV>), dim3(nbad), dim3(kWave), 0, st, gb_, bad, nbad, d, t, y, \
                           guide, claim, misc);                                                                \
      else                                                                                                     \
        hipLaunchKernelGGL((k_pp_plateau<3, GeoB, PMV>), dim3(nbad), dim3(kWave), 0, st, gb_, bad, nbad, d, t, y, \
                           guide, claim, misc);                                                                \
    } else {                                                                                                   \
      const Geo1 g1_ = geo1(geo);                                                                              \
      if (geo.ndim == 2)                                                                                       \
        hipLaunchKernelGGL((k_pp_plateau<2, Geo1, PMV>), dim3(nbad), dim3(kWave), 0, st, g1_, bad, nbad, d, t, y, \
                           guide, claim, misc);                                                                \
      else                                                                                                     \
        hipLaunchKernelGGL((k_pp_plateau<3, Geo1, PMV>), dim3(nbad), dim3(kWave), 0, st, g1_, bad, nbad, d, t, y, \
                           guide, claim, misc);                                                                \
    }                                                                                                          \
  } while (0)

hipError_t pp_plateau(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, const float* d, const uint8_t* t, float* y,
                      const float* guide, uint32_t* claim, bool large, uint32_t* misc, hipStream_t st) {
  if (!nbad) return hipSuccess;
  if (large) PP_PLAT_LAUNCH(kPlatLarge);
  else PP_PLAT_LAUNCH(kPlatSmall);
  return hipGetLastError();
}

hipError_t pp_solve(const PpGeo& geo, const uint32_t* tiles, uint32_t ntiles, uint32_t cap, int phase,
                    const uint32_t* pbits, const float* d, const uint8_t* t, float* y, int anchor, uint32_t* misc,
                    hipStream_t st) {
  if (!ntiles) return hipSuccess;
  PP_GEO_LAUNCH(k_pp_solve, dim3(ntiles), dim3(kWave), st, tiles, cap, phase, pbits, d, t, y, anchor, misc);
  return hipGetLastError();
}

}  // namespace tszp
