// tszp_device.hpp -- device-side building blocks shared by every TopoSZp kernel.
//
// Written for gfx950 (CDNA4): 64-wide wavefronts, 64-bit ballots, LDS atomics.
// All arithmetic that feeds the byte stream reproduces the reference's x86 CPU
// semantics exactly (see cvt_i32 and bitcount below).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tszp_types.h"

namespace tszp {

constexpr int kWave = 64;
constexpr int kBlockThreads = 256;  // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlockThreads / kWave;

// ---------------------------------------------------------------------------
// scalar rules of the reference
// ---------------------------------------------------------------------------

// C cast of a double to int on x86-64 (cvttsd2si): truncation toward zero, and
// 0x80000000 for NaN or out-of-range input.  The reference quantizes with this
// cast (SZp/src/szp_float.cc:1212,1222), so the GPU must produce the same integer
// where a plain v_cvt_i32_f64 would saturate.
__device__ __forceinline__ int cvt_i32(double d) {
  return (d > -2147483649.0 && d < 2147483648.0) ? (int)d : (int)0x80000000u;
}

// Smallest value of binade k whose bit count is k + 2 under the reference rule
// bc = floorf(log2f((float)m)) + 1 (SZp/src/szp_float.cc:1243) with glibc's
// log2f.  Below 2^20 the rule is exact (bc = k + 1); above, the float rounding of
// m and of log2f push the top values of each binade to k + 2.  Measured
// exhaustively over all 2^32 inputs; re-checked by the oracle tests.
__constant__ const uint32_t kBitcountThreshold[32] = {
    2u,         4u,          8u,          16u,        32u,        64u,         128u,
    256u,       512u,        1024u,       2048u,      4096u,      8192u,       16384u,
    32768u,     65536u,      131072u,     262144u,    524288u,    1048576u,    2097151u,
    4194302u,   8388603u,    16777205u,   33554410u,  67108819u,  134217637u,  268435273u,
    536870545u, 1073741089u, 2147482177u, 4294964353u};

__device__ __forceinline__ unsigned bitcount(unsigned m) {
  if (m == 0) return 0;
  const unsigned k = 31u - (unsigned)__clz((int)m);
  // below 2^20 the threshold is 2^(k+1), which m never reaches: no table read (a
  // memory round trip) on the common path
  if (__builtin_expect(m < (1u << 20), 1)) return k + 1u;
  return k + 1u + (m >= kBitcountThreshold[k] ? 1u : 0u);
}

// Order-preserving 32-bit key of a float; -0.0 folds onto +0.0 so that keys are
// equal exactly when the floats compare equal (NaN is not supported).
__device__ __forceinline__ uint32_t okey(float v) {
  uint32_t u = __float_as_uint(v);
  if (u == 0x80000000u) u = 0;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float okey_value(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

// quantized bin of a critical point: float add, then double multiply
// (SZp/src/szp_float.cc:76).
__device__ __forceinline__ int cp_bin(float v, float eb, double inv) {
  float s = __fadd_rn(v, eb);
  return cvt_i32((double)s * inv);
}

// bytes of the magnitude section for n values of width bc
// (Jiajun_save_fixed_length_bits, SZp/src/szp_TypeManager.cc:13-103)
__device__ __forceinline__ unsigned mag_bytes(unsigned n, unsigned bc) {
  unsigned rem = bc & 7u;
  return (bc >> 3) * n + (rem ? (rem * n + 7u) >> 3 : 0u);
}

__device__ __forceinline__ unsigned block_bytes(unsigned cur, unsigned bc, bool topo) {
  unsigned n = cur - 1u;
  unsigned b = 5u;
  if (bc) b += ((n + 7u) >> 3) + mag_bytes(n, bc);
  if (topo) b += (2u * cur + 7u) >> 3;
  return b;
}

// ---------------------------------------------------------------------------
// unsigned division of x < 2^31 by a runtime constant: (x * mul) >> shift with
// mul = ceil(2^(31+l) / d), l = ceil(log2 d), shift = 31 + l (exact for x < 2^31)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t fdiv(uint32_t x, FastDiv d) {
  return (uint32_t)(((uint64_t)x * d.mul) >> d.shift);
}

// ---------------------------------------------------------------------------
// wave-level helpers (segments of S lanes, S a power of two <= 64)
// ---------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ unsigned seg_max(unsigned v) {
#pragma unroll
  for (int o = S / 2; o > 0; o >>= 1) {
    unsigned w = __shfl_xor(v, o, kWave);
    v = w > v ? w : v;
  }
  return v;
}

template <int S, typename T>
__device__ __forceinline__ T seg_sum(T v) {
#pragma unroll
  for (int o = S / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// inclusive prefix sum over the lanes of each S-lane segment
template <int S, typename T>
__device__ __forceinline__ T seg_incl_scan(T v, int lane) {
  const int ls = lane & (S - 1);
#pragma unroll
  for (int o = 1; o < S; o <<= 1) {
    T w = __shfl_up(v, o, kWave);
    if (ls >= o) v += w;
  }
  return v;
}

// DPP lane moves (GFX9 encodings; lane layouts checked on gfx950).  The result of
// lane i is v of the source lane; lanes without a source (or inactive ones) read 0.
// Call with the whole wave active.
constexpr int kDppQuad1032 = 0xB1, kDppQuad2301 = 0x4E, kDppRowHalfMirror = 0x141, kDppRowMirror = 0x140;
constexpr int kDppWaveShr1 = 0x138;  // lane i <- lane i - 1
constexpr int kDppWaveShl1 = 0x130;  // lane i <- lane i + 1
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true);
}

// maximum / sum over aligned segments of S <= 16 lanes, in every lane of the segment
template <int S>
__device__ __forceinline__ uint32_t seg_max_dpp(uint32_t v) {
  static_assert(S >= 1 && S <= 16 && (S & (S - 1)) == 0, "segment of 1..16 lanes");
  if constexpr (S >= 2) v = max(v, dpp_mov<kDppQuad1032>(v));
  if constexpr (S >= 4) v = max(v, dpp_mov<kDppQuad2301>(v));
  if constexpr (S >= 8) v = max(v, dpp_mov<kDppRowHalfMirror>(v));
  if constexpr (S >= 16) v = max(v, dpp_mov<kDppRowMirror>(v));
  return v;
}
template <int S>
__device__ __forceinline__ uint32_t seg_sum_dpp(uint32_t v) {
  static_assert(S >= 1 && S <= 16 && (S & (S - 1)) == 0, "segment of 1..16 lanes");
  if constexpr (S >= 2) v += dpp_mov<kDppQuad1032>(v);
  if constexpr (S >= 4) v += dpp_mov<kDppQuad2301>(v);
  if constexpr (S >= 8) v += dpp_mov<kDppRowHalfMirror>(v);
  if constexpr (S >= 16) v += dpp_mov<kDppRowMirror>(v);
  return v;
}
// sum over the wave (wave-uniform)
__device__ __forceinline__ uint32_t wave_sum_dpp(uint32_t v) {
  v = seg_sum_dpp<16>(v);
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) + (uint32_t)__builtin_amdgcn_readlane((int)v, 16) +
         (uint32_t)__builtin_amdgcn_readlane((int)v, 32) + (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}
// inclusive prefix sum over the wave
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v, int lane) {
  v += dpp_mov<kDppRowShr1>(v);
  v += dpp_mov<kDppRowShr2>(v);
  v += dpp_mov<kDppRowShr4>(v);
  v += dpp_mov<kDppRowShr8>(v);
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
  const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
  const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
  const int row = lane >> 4;
  return v + (row >= 1 ? r0 : 0u) + (row >= 2 ? r1 : 0u) + (row >= 3 ? r2 : 0u);
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// ---------------------------------------------------------------------------
// LDS byte staging: OR `nbytes` little-endian bytes of v at byte address `a`.
// The staging area is zeroed first; every write is an LDS atomic OR so that
// bytes packed by different lanes into one 32-bit word never clobber each other.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lds_or_bytes(uint32_t* words, unsigned a, uint32_t v, unsigned nbytes) {
  uint64_t wide = (uint64_t)(nbytes >= 4 ? v : (v & ((1u << (8 * nbytes)) - 1u))) << (8 * (a & 3u));
  unsigned w = a >> 2;
  uint32_t lo = (uint32_t)wide, hi = (uint32_t)(wide >> 32);
  if (lo) atomicOr(&words[w], lo);
  if (hi) atomicOr(&words[w + 1], hi);
}

// OR a field of `width` (<= 8) bits, MSB first, at bit position `pos` of an
// MSB-first bit stream that starts at byte address `base`.
__device__ __forceinline__ void lds_or_bits_msb(uint32_t* words, unsigned base, unsigned pos,
                                                unsigned value, unsigned width) {
  unsigned byte = base + (pos >> 3);
  unsigned sh = pos & 7u;
  unsigned w16 = (value & ((1u << width) - 1u)) << (16u - width - sh);  // big-endian window
  if (!w16) return;
  unsigned b0 = w16 >> 8, b1 = w16 & 0xffu;
  if (b0) atomicOr(&words[byte >> 2], b0 << (8 * (byte & 3u)));
  if (b1) atomicOr(&words[(byte + 1) >> 2], b1 << (8 * ((byte + 1) & 3u)));
}

__device__ __forceinline__ unsigned lds_byte(const uint32_t* words, unsigned a) {
  return (words[a >> 2] >> (8 * (a & 3u))) & 0xffu;
}

// ---------------------------------------------------------------------------
// decoupled look-back (single-pass prefix over tiles in dynamic tile order)
// status word: [63:62] flag (0 empty, 1 aggregate, 2 inclusive prefix), [61:0] value
// ---------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

// A status word carries its own payload (flag + value in one 64-bit word), so the
// look-back needs no ordering with any other memory: relaxed device-scope atomics
// suffice.  Release/acquire would add an L2 write-back / invalidate per access on
// this multi-XCD part, which serialises the tiles.
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tile of a workgroup in a single-pass look-back kernel.  Workgroup ids are dealt
// round-robin to the XCDs and every XCD dispatches its share in increasing order, so
// when a tile runs, the tile before it has been dispatched on its own XCD and the
// look-back never waits for a workgroup that cannot start.  A dynamic counter (one
// same-address atomic per workgroup: serialised at ~12 ns each on this part, which
// caps a 32 K-tile launch near 0.4 ms) is used when the host passes one.
__device__ __forceinline__ uint32_t tile_id(uint32_t* counter) {
  return counter ? atomicAdd(counter, 1u) : blockIdx.x;
}

// Called by a full wave of tile `tile`.  Publishes `agg`, walks predecessors
// kLookWindow at a time (lanes 0 .. kLookWindow - 1; the nearest inclusive prefix is
// usually a few tiles back, and every examined word is an atomic load: 32 measured
// faster than 8, 16, 64 or 256), publishes the inclusive prefix and returns the
// exclusive prefix (wave-uniform).  Values add component-wise, so two 31-bit counters packed
// into one value are scanned together.  A spin that exceeds kSpinLimit sets
// *err and gives up (the result is then wrong but the kernel still ends).
constexpr int kLookWindow = 32;
__device__ uint64_t lookback(uint64_t* status, uint32_t tile, uint64_t agg, uint32_t* err) {
  const int lane = threadIdx.x & (kWave - 1);
  if (tile == 0) {
    if (lane == 0) st_status(&status[0], kFlagPre | (agg & kValMask));
    return 0;
  }
  if (lane == 0) st_status(&status[tile], kFlagAgg | (agg & kValMask));
  uint64_t excl = 0;
  int64_t pos = (int64_t)tile - 1;
  uint32_t spins = 0;
  while (true) {
    int64_t idx = pos - lane;
    // lanes past the window hold an empty aggregate; a virtual prefix 0 before tile 0
    uint64_t s = lane >= kLookWindow ? kFlagAgg : (idx >= 0 ? ld_status(&status[idx]) : kFlagPre);
    // wait until every examined predecessor has published something
    while (__any((s >> 62) == 0)) {
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(err, TSZP_ERR_LOOKBACK);
        return excl;
      }
      __builtin_amdgcn_s_sleep(1);
      if ((s >> 62) == 0) s = ld_status(&status[idx]);
    }
    uint64_t pre_mask = __ballot((s >> 62) == 2);
    if (pre_mask) {
      int first = __ffsll((unsigned long long)pre_mask) - 1;  // nearest inclusive prefix
      uint64_t v = lane <= first ? (s & kValMask) : 0;
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
      excl += v;
      break;
    }
    uint64_t v = s & kValMask;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    excl += v;
    pos -= kLookWindow;
  }
  if (lane == 0) st_status(&status[tile], kFlagPre | ((excl + agg) & kValMask));
  return excl;
}

// Segmented variant: a value carries kSegReset when a new segment started inside
// its span, and then holds only the sum since that restart.  The walk stops at the
// nearest predecessor that either has an inclusive prefix or restarted.
constexpr uint64_t kSegReset = 1ull << 61;

__device__ uint64_t lookback_seg(uint64_t* status, uint32_t tile, uint64_t agg, uint32_t* err) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t vmask = kSegReset - 1;
  if (tile == 0) {
    if (lane == 0) st_status(&status[0], kFlagPre | (agg & vmask));
    return 0;
  }
  if (lane == 0) st_status(&status[tile], kFlagAgg | (agg & kValMask));
  uint64_t excl = 0;
  int64_t pos = (int64_t)tile - 1;
  uint32_t spins = 0;
  while (true) {
    int64_t idx = pos - lane;
    uint64_t s = idx >= 0 ? ld_status(&status[idx]) : kFlagPre;
    while (__any((s >> 62) == 0)) {
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(err, TSZP_ERR_LOOKBACK);
        return excl;
      }
      __builtin_amdgcn_s_sleep(1);
      if ((s >> 62) == 0) s = ld_status(&status[idx]);
    }
    uint64_t stop = __ballot((s >> 62) == 2 || (s & kSegReset));
    if (stop) {
      int first = __ffsll((unsigned long long)stop) - 1;
      uint64_t v = lane <= first ? (s & vmask) : 0;
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
      excl += v;
      break;
    }
    uint64_t v = s & vmask;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    excl += v;
    pos -= kWave;
  }
  const uint64_t incl = (agg & kSegReset) ? (agg & vmask) : ((excl + (agg & vmask)) & vmask);
  if (lane == 0) st_status(&status[tile], kFlagPre | incl);
  return excl;
}

// exclusive scan of one value per thread over a 256-thread workgroup
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds4, uint32_t* total) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t incl = seg_incl_scan<kWave>(v, lane);
  if (lane == kWave - 1) lds4[w] = incl;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kWavesPerBlock; k++) {
    uint32_t t = lds4[k];
    if (k < w) pre += t;
    tot += t;
  }
  *total = tot;
  __syncthreads();
  return pre + incl - v;
}

}  // namespace tszp
