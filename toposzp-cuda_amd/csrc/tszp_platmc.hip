// tszp_platmc.hip -- whole-plateau ordering by min-conflicts local search.
//
// A plateau (connected points of one decoded value) orders its points freely inside
// their common cell, and a point's type depends only on its order against its 2*ND
// neighbours; neighbours in other cells are fixed below or above the whole cell.  So a
// plateau whose points do not all carry their transmitted types is re-ordered on its
// own, in one wavefront:
//
//   * gather: breadth first over equal decoded values (LDS hash table), claimed by the
//     plateau's smallest point, then sorted by index so that everything after depends
//     on the plateau alone (the result does not depend on which mismatch got there);
//   * keys: one float per point, the order of the restored values (ties broken by
//     index).  Start: a guide -- the smoothed decoded field when the caller passes one,
//     else each point's breadth-first distances to the lower and the higher cells next
//     to the plateau, dl / (dl + dh);
//   * anchored extrema (the sort-position carriers) keep their values, so only their
//     order among themselves is fixed: an anchored point's key stays strictly between
//     the keys of the anchors before and after it in value order, and is otherwise as
//     free as any other point's;
//   * search: min-conflicts.  Each step draws a point whose type is violated; it and its
//     plateau neighbours are the movers, and each mover may go to any of up to 19
//     positions against its own plateau neighbours' keys (far below, just below, just
//     above / middle / just below the top of each gap, just above, far above).  The 64
//     lanes score the 133 (mover, position) pairs by the change of violated types of
//     the mover and its neighbours; the best move is taken (a move made within the last
//     few steps loses ties), or with a fixed probability a random one;
//   * values: in key order, anchored points keep theirs and the points between two of
//     them take values spaced evenly between (below the first: from the cell's low
//     end, above the last: up to its high end), as the other solvers of this file.
// A plateau the search does not finish within its step budget keeps its values and is
// left to the solvers that run after this one.
#include <hip/hip_runtime.h>

#include "tszp_device.hpp"
#include "tszp_kernels.h"
#include "tszp_post_geo.hpp"

namespace tszp {
namespace {

constexpr int kMcGapPos = 3;            // positions inside each gap between neighbour keys
constexpr int kMcPos = 4 + 5 * kMcGapPos;  // far / just below, the gaps, just / far above
constexpr int kMcSlots = 7 * kMcPos;    // the mismatched point and up to 6 plateau neighbours
constexpr int kMcSlotsPerLane = (kMcSlots + 63) / 64;
constexpr uint32_t kMcWalkPct = 15;     // share of random moves, per cent
constexpr uint32_t kMcTabu = 10;        // steps during which a moved point loses ties
constexpr int16_t kNbLow = -1, kNbHigh = -2, kNbOut = -3;
constexpr int kMcTiny = 64, kMcSmall = 512, kMcLarge = 3072;

template <int ND, int PM>
struct McShared {
  static constexpr int kHash = PM <= 64 ? 128 : (PM <= 512 ? 1024 : 8192);  // a power of two >= 2 PM
  uint32_t el[PM];
  int16_t nb[PM][2 * ND];  // >= 0: plateau index; kNbLow / kNbHigh: other cell; kNbOut: off the grid
  float key[PM];
  int8_t tgt[PM];          // target type, -1 off the interior
  uint8_t vio[PM];
  uint16_t vlist[PM], vpos[PM];
  uint16_t apos[PM];       // position in aord, 0xffff for a free point
  uint16_t tabu[PM];       // 1 + step of the point's last move (0: never)
  uint16_t aord[PM];       // anchored points in value order
  uint32_t hkey[kHash];    // membership while gathering; then sort space and the final order
  uint16_t hval[kHash];
  uint32_t cand[kWave];
  // per step: each mover's neighbour records and its own (see mc_nbr_rec / mc_mover_rec)
  uint64_t vbits[(PM + 63) / 64];  // violated types, one bit per point (drawn from uniformly)
  uint64_t tbl[2 * ND + 1][2 * ND];
  uint64_t mrec[2 * ND + 1][2];
  float srt[2 * ND + 1][2 * ND + 1];  // each mover's neighbour keys in order, then the largest
  int ne, na, nv, abort;
  float lo, hi, db;
};
// renormalisation period of the keys (steps): repeated moves next to a neighbour's key
// crowd the keys, and float positions between crowded keys degenerate into ties
constexpr uint32_t kMcRenorm = 128;
constexpr uint32_t kMcPoll = 16;  // steps between a replica's looks at its entry's best finish
static_assert(sizeof(McShared<3, kMcLarge>) <= 160 * 1024, "LDS of the large variant");

__device__ __forceinline__ uint64_t mc_mix(uint64_t h) {
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

constexpr int mc_log2(int v) { return v <= 1 ? 0 : 1 + mc_log2(v >> 1); }
template <int H>
__device__ __forceinline__ uint32_t mc_hash(uint32_t e) {
  return (e * 2654435761u) >> (32 - mc_log2(H));
}

template <class SH>
__device__ __forceinline__ int mc_find(const SH& sh, uint32_t e) {
  constexpr int H = SH::kHash;
  uint32_t h = mc_hash<H>(e);
  for (int k = 0; k < H; k++) {
    const uint32_t key = sh.hkey[h];
    if (key == e) return sh.hval[h];
    if (key == 0xffffffffu) return -1;
    h = (h + 1) & (H - 1);
  }
  return -1;
}

// index of e in the sorted element list (e is an element)
template <class SH>
__device__ __forceinline__ int mc_index(const SH& sh, uint32_t e) {
  int lo = 0, hi = sh.ne - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sh.el[mid] < e) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// whether the point at c has a neighbour in direction k (axis k / 2, side k % 2), and
// the stride of axis a, by selects (no indexed private memory)
template <int ND>
__device__ __forceinline__ bool mc_has_nbr(const Grid& g, const int (&c)[ND], int k) {
  const int a = k >> 1;
  int ca = 0, da = 0;  // one-hot sums: a select chain becomes indexed private memory
#pragma unroll
  for (int x = 0; x < ND; x++) {
    ca += c[x] * (int)(a == x);
    da += (x == 0 ? (int)g.d0 : (x == 1 ? (int)g.d1 : (int)g.d2)) * (int)(a == x);
  }
  return (k & 1) ? ca + 1 < da : ca > 0;
}
template <int ND>
__device__ __forceinline__ uint32_t mc_stride(const Grid& g, int a) {
  if constexpr (ND == 2) return a == 0 ? g.d1 : 1u;
  else return a == 0 ? g.plane : (a == 1 ? g.d2 : 1u);
}

// bitonic sort of p (a power of two) values in LDS by one wavefront
template <class T>
__device__ void wave_bitonic(T* a, int p, int lane) {
  for (int k = 2; k <= p; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < p; i += kWave) {
        const int l = i ^ j;
        if (l > i) {
          const T x = a[i], y = a[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            a[i] = y;
            a[l] = x;
          }
        }
      }
      __syncthreads();
    }
}

// type from "lower than the centre" flags of the 2*ND neighbours (axis a: 2a, 2a + 1)
template <int ND>
__device__ __forceinline__ int mc_type(uint32_t low) {
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

// sort key of (key, index): order-preserving float bits, then the index
__device__ __forceinline__ uint64_t mc_sortkey(float k, int i) {
  uint32_t kb = __float_as_uint(k);
  kb = (kb & 0x80000000u) ? ~kb : (kb | 0x80000000u);
  return ((uint64_t)kb << 32) | (uint32_t)i;
}

// (ka, a) before (kb, b) in key order
__device__ __forceinline__ bool mc_before(float ka, int a, float kb, int b) { return ka < kb || (ka == kb && a < b); }

// whether point i's type is violated when point m sits at key km (m == -1: as it is)
template <int ND, class SH>
__device__ __forceinline__ int mc_viol(const SH& sh, int i, int m, float km) {
  if (sh.tgt[i] < 0) return 0;
  const float ki = i == m ? km : sh.key[i];
  uint32_t low = 0;
#pragma unroll
  for (int k = 0; k < 2 * ND; k++) {
    const int q = sh.nb[i][k];
    bool l = q == kNbLow;
    if (q >= 0) l = mc_before(q == m ? km : sh.key[q], q, ki, i);
    low |= (uint32_t)l << k;
  }
  return mc_type<ND>(low) != sh.tgt[i];
}

__device__ __forceinline__ float mc_near(float x) { return fmaxf(fabsf(x), 1.f) * 9.5367432e-7f; }  // 2^-20

constexpr float kMcNone = 1e30f;  // key of a missing neighbour (sorts last, finite)

// minimum over the wave, wave-uniform (DPP within rows of 16, then the four row results)
__device__ __forceinline__ uint32_t mc_wave_min(uint32_t v) {
  v = min(v, dpp_mov<kDppQuad1032>(v));
  v = min(v, dpp_mov<kDppQuad2301>(v));
  v = min(v, dpp_mov<kDppRowHalfMirror>(v));
  v = min(v, dpp_mov<kDppRowMirror>(v));
  const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0), b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16),
                 c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32), d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
  return min(min(a, b), min(c, d));
}

// position of the k-th set bit of w (k < popcount(w)), by halving
__device__ __forceinline__ int mc_kth_bit(uint64_t w, int k) {
  int pos = 0;
#pragma unroll
  for (int half = 32; half > 0; half >>= 1) {
    const uint64_t lo = w & ((1ull << half) - 1ull);
    const int c = __popcll(lo);
    if (k >= c) {
      k -= c;
      w >>= half;
      pos += half;
    } else {
      w = lo;
    }
  }
  return pos;
}

// Per step, the movers are the drawn point (mover 0) and its plateau neighbours (mover
// 1 + k, direction k).  Neighbour record of mover mi in direction k, 64 bits:
//   bits 0-1 kind (0 plateau point, 1 lower cell, 2 higher cell, 3 off the grid),
//   2 the point's type violated with the mover below it, 3 with the mover above it,
//   4 violated now, 16-31 plateau index, 32-63 key.
// Mover record: [0] bits 0-15 index, 16-18 target + 1, 19 violated, 20 valid, 21 has
// an anchor below, 22 one above, 23 moved within the tabu window, 24-29 directions of
// its neighbours in lower cells, 32-63 key; [1] keys of the anchors below (low half) and
// above (high half).  A mover's candidate positions then cost one compare per
// neighbour: the neighbours' own types are looked up, not recomputed.
__device__ __forceinline__ float mc_hi_f(uint64_t v) { return __uint_as_float((uint32_t)(v >> 32)); }

template <int ND, class SH>
__device__ void mc_relist(const SH& sh, const FieldView& F, const float* y, const uint8_t* __restrict__ t,
                          uint32_t* __restrict__ relist, uint32_t cap, uint32_t* misc, int lane) {
  for (int i0 = 0; i0 < sh.ne; i0 += kWave) {
    const int i = i0 + lane;
    bool bad = false;
    uint32_t e = 0;
    if (i < sh.ne) {
      e = sh.el[i];
      int c[ND];
      coords_of<ND>(F.g, e - F.base, c);
      bad = classify_at<ND>(F.g, y, e, c) != (int)t[e];
    }
    const uint64_t m = __ballot(bad);
    if (!m) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&misc[kPpBad], (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader, kWave);
    if (bad) {
      const uint32_t pos = base + (uint32_t)__popcll(m & lanemask_lt(lane));
      if (pos < cap) relist[pos] = e;
    }
  }
}

// End of a replica (every replica of an entry calls it once, every lane): the last one to
// finish writes the values of the winner -- the fewest steps, then the lowest replica --
// from its staging row (ctl[ent]: (steps << 3) | replica of the best finish, ~0 if none;
// ctl[nent + ent]: finished replicas) and lists what they leave mismatched, or, with no
// winner, the plateau's mismatches as they are.
template <int ND, int PM, class SH>
__device__ void mc_replica_end(SH& sh, const FieldView& F, float* y, const uint8_t* __restrict__ t,
                               const float* stage, uint32_t* ctl, uint32_t nent, uint32_t ent, uint32_t R,
                               uint32_t* __restrict__ relist, uint32_t cap, bool stats, uint32_t* misc, int lane) {
  __threadfence();  // this replica's staged values before its count
  __syncthreads();
  if (lane == 0) sh.cand[0] = atomicAdd(&ctl[nent + ent], 1u) == R - 1u ? 1u : 0u;
  __syncthreads();
  if (!sh.cand[0]) return;
  __threadfence();  // the other replicas' staged values after their counts
  const uint32_t best = __hip_atomic_load(&ctl[ent], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (best != ~0u) {
    const float* src = stage + ((size_t)ent * R + (best & 7u)) * PM;
    for (int i = lane; i < sh.ne; i += kWave) y[sh.el[i]] = src[i];
    __syncthreads();
  }
  if (lane == 0 && stats) atomicAdd(&misc[best != ~0u ? kPpMcSolved : kPpMcFailed], 1u);
  mc_relist<ND>(sh, F, y, t, relist, cap, misc, lane);
}

// R == 0: one search per plateau around each listed mismatch (claimed by the plateau's
// smallest point); a plateau the budget does not finish goes to esc (its smallest point,
// one entry per plateau) when esc is given, else its mismatches to relist.
// R >= 1: entry bad[i] (a smallest point from esc) is searched by R replicas, blocks
// i * R + r, each from the same start keys with a random stream of its own; a replica
// stops once another has finished in fewer steps (polled every kMcPoll steps), a
// finished one stages its values (stage: R * PM floats per entry), and the last replica
// of an entry writes the winner's values (mc_replica_end): the result depends on the
// plateau alone, not on the replicas' timing.
template <int ND, class GEO, int PM>
__global__ __launch_bounds__(kWave, PM <= kMcTiny ? 8 : 1) void k_pp_platmc(GEO geo, const uint32_t* __restrict__ bad, uint32_t nbad,
                                                     const float* __restrict__ d, const uint8_t* __restrict__ t,
                                                     float* __restrict__ y, const float* __restrict__ guide,
                                                     uint32_t* __restrict__ claim, uint32_t budget, bool last,
                                                     uint32_t* __restrict__ relist, uint32_t cap, bool stats,
                                                     uint32_t* misc, uint32_t R, uint32_t* __restrict__ esc,
                                                     float* __restrict__ stage, uint32_t* __restrict__ ctl) {
  constexpr int K = 2 * ND;
  using SH = McShared<ND, PM>;
  constexpr int H = SH::kHash;
  __shared__ SH sh;
  const int lane = threadIdx.x;
  const uint32_t ent = R ? blockIdx.x / R : blockIdx.x;
  const uint32_t rep = R ? blockIdx.x % R : 0u;
  if (ent >= nbad) return;
  const uint32_t b = bad[ent];
  const FieldView F = geo.at(b);
  const Grid& g = F.g;
  if (!R) {
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
    const uint32_t h = mc_hash<H>(b);
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
        const uint32_t e = sh.el[ci / K];
        const int k = ci % K;
        int c[ND];
        coords_of<ND>(g, e - F.base, c);
        if (mc_has_nbr<ND>(g, c, k)) {
          const uint32_t s = mc_stride<ND>(g, k >> 1);
          const uint32_t qe = (k & 1) ? e + s : e - s;
          if (d[qe] == db) q = qe;
        }
      }
      // parallel insertion: a lane whose compare-and-swap claims an empty slot adds a
      // new element, a lane that meets its own key has a duplicate
      bool fresh = false;
      uint32_t slot = 0;
      if (q != 0xffffffffu) {
        uint32_t h = mc_hash<H>(q);
        for (int probe = 0; probe < H; probe++) {
          const uint32_t old = atomicCAS(&sh.hkey[h], 0xffffffffu, q);
          if (old == 0xffffffffu) {
            fresh = true;
            slot = h;
            break;
          }
          if (old == q) break;
          h = (h + 1) & (H - 1);
        }
      }
      const uint64_t m = __ballot(fresh);
      const int base = sh.ne;
      if (fresh) {
        const int idx = base + __popcll(m & lanemask_lt(lane));
        if (idx < PM) {
          sh.el[idx] = q;
          sh.hval[slot] = (uint16_t)idx;
        }
      }
      __syncthreads();
      if (lane == 0) {
        const int nn = base + __popcll(m);
        if (nn > PM) sh.abort = 1;
        sh.ne = nn > PM ? PM : nn;
      }
      __syncthreads();
      if (sh.abort) break;
    }
    head = ne0;
  }
  if (sh.abort) {
    // (a replica's entry was searched by the first launch of the same variant, so it
    // never exceeds the variant; its replicas still count themselves out)
    if (R) {
      __syncthreads();
      if (lane == 0 && atomicAdd(&ctl[nbad + ent], 1u) == R - 1u) {
        const uint32_t pos = atomicAdd(&misc[kPpBad], 1u);
        if (pos < cap) relist[pos] = b;
      }
      return;
    }
    if (lane == 0) {
      atomicAdd(&misc[kPpMcBig], 1u);
      if (last) {  // beyond every variant: the mismatch stays listed
        const uint32_t pos = atomicAdd(&misc[kPpBad], 1u);
        if (pos < cap) relist[pos] = b;
      }
    }
    return;
  }
  const int ne = sh.ne;
  if (!R) {  // claimed by its smallest point: one wavefront per plateau
    uint32_t mn = 0xffffffffu;
    for (int i = lane; i < ne; i += kWave) mn = min(mn, sh.el[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, kWave));
    if (lane == 0) sh.cand[0] = claim_bit(claim, mn) ? 1u : 0u;
    __syncthreads();
    if (!sh.cand[0]) return;
  }
  const uint64_t t_start = wall_clock64();
  // ---- elements in index order (the hash table is no longer needed) -----------------
  int p2 = 1;
  while (p2 < ne) p2 <<= 1;
  uint32_t* srt = sh.hkey;
  for (int i = lane; i < p2; i += kWave) srt[i] = i < ne ? sh.el[i] : 0xffffffffu;
  __syncthreads();
  wave_bitonic(srt, p2, lane);
  for (int i = lane; i < ne; i += kWave) sh.el[i] = srt[i];
  __syncthreads();
  // ---- neighbour codes, targets, anchors --------------------------------------------
  int na = 0;
  for (int i0 = 0; i0 < ne; i0 += kWave) {
    const int i = i0 + lane;
    bool anc = false;
    if (i < ne) {
      const uint32_t e = sh.el[i];
      int c[ND];
      coords_of<ND>(g, e - F.base, c);
      const int tt = t[e];
      sh.tgt[i] = interior<ND>(g, c) ? (int8_t)tt : (int8_t)-1;
      anc = (tt == 1 || tt == 2) && db != 0.f;
      sh.tabu[i] = 0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int a = k >> 1;
        int q[ND];
#pragma unroll
        for (int x = 0; x < ND; x++) q[x] = c[x];
        q[a] += (k & 1) ? 1 : -1;
        int16_t code = kNbOut;
        if (inside<ND>(g, q)) {
          const uint32_t s = stride_of<ND>(g, a);
          const uint32_t qe = (k & 1) ? e + s : e - s;
          const float dq = d[qe];
          code = dq == db ? (int16_t)mc_index(sh, qe) : (dq < db ? kNbLow : kNbHigh);
        }
        sh.nb[i][k] = code;
      }
      sh.apos[i] = 0xffff;
      if (anc) sh.key[i] = y[e];  // its value from the start pass, until the start keys
    }
    const uint64_t m = __ballot(anc);
    if (anc) sh.vlist[na + __popcll(m & lanemask_lt(lane))] = (uint16_t)i;  // anchored, unordered
    na += __popcll(m);
  }
  if (lane == 0) sh.na = na;
  __syncthreads();
  // anchors by value (the start pass wrote their values), ties by index; their values
  // stay in the gather's (now free) hash values, by anchor order
  float* aval = reinterpret_cast<float*>(sh.hval);  // kHash / 2 >= PM floats
  for (int k = lane; k < na; k += kWave) {
    const int i = sh.vlist[k];
    const float vi = sh.key[i];
    int r = 0;
    for (int j = 0; j < na; j++) {
      const int o = sh.vlist[j];
      const float vo = sh.key[o];
      r += vo < vi || (vo == vi && o < i);
    }
    sh.aord[r] = (uint16_t)i;
    sh.apos[i] = (uint16_t)r;
    aval[r] = vi;
  }
  __syncthreads();
  // ---- start keys ---------------------------------------------------------------------
  const float inv_eb = 1.f / F.eb;
  if (guide) {
    for (int i = lane; i < ne; i += kWave) sh.key[i] = (guide[sh.el[i]] - db) * inv_eb;
  } else {
    // breadth-first distances to the lower (vlist) and the higher (vpos) cells
    for (int i = lane; i < ne; i += kWave) {
      uint16_t dl = 0xffff, dh = 0xffff;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int q = sh.nb[i][k];
        if (q == kNbLow) dl = 1;
        if (q == kNbHigh) dh = 1;
      }
      sh.vlist[i] = dl;
      sh.vpos[i] = dh;
    }
    __syncthreads();
    for (int round = 0; round < ne; round++) {
      bool changed = false;
      for (int i = lane; i < ne; i += kWave) {
        uint32_t dl = sh.vlist[i], dh = sh.vpos[i];
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int q = sh.nb[i][k];
          if (q >= 0) {
            dl = min(dl, (uint32_t)sh.vlist[q] + 1u);
            dh = min(dh, (uint32_t)sh.vpos[q] + 1u);
          }
        }
        if (dl < sh.vlist[i]) {
          sh.vlist[i] = (uint16_t)dl;
          changed = true;
        }
        if (dh < sh.vpos[i]) {
          sh.vpos[i] = (uint16_t)dh;
          changed = true;
        }
      }
      __syncthreads();
      if (!__any(changed)) break;
    }
    for (int i = lane; i < ne; i += kWave) {
      const float dl = sh.vlist[i] == 0xffff ? 1000.f : (float)sh.vlist[i];
      const float dh = sh.vpos[i] == 0xffff ? 1000.f : (float)sh.vpos[i];
      const uint64_t h = mc_mix((uint64_t)sh.el[i] * 0x9e3779b97f4a7c15ull);
      sh.key[i] = dl / (dl + dh) + 1e-5f * ((float)(uint32_t)h * 2.3283064e-10f);
    }
  }
  __syncthreads();
  // anchors take their own start keys in anchor order, strictly increasing
  if (na > 0) {
    for (int k = lane; k < na; k += kWave) {
      const int i = sh.aord[k];
      const float ki = sh.key[i];
      int r = 0;
      for (int j = 0; j < na; j++) {
        const int o = sh.aord[j];
        const float ko = sh.key[o];
        r += ko < ki || (ko == ki && j < k);
      }
      reinterpret_cast<float*>(sh.hkey)[r] = ki;
    }
    __syncthreads();
    if (lane == 0) {
      float prev = 0.f;
      for (int k = 0; k < na; k++) {
        float v = reinterpret_cast<float*>(sh.hkey)[k];
        if (k > 0 && !(v > prev)) v = prev + mc_near(prev);
        sh.key[sh.aord[k]] = v;
        prev = v;
      }
    }
    __syncthreads();
  }
  // ---- violated types: a byte per point for the records, a bit per point for the draws ----
  static_assert((PM + 63) / 64 <= kWave, "one bitmap word per lane");
  const int nwords = (ne + 63) / 64;
  int nv = 0;
  for (int i0 = 0; i0 < ne; i0 += kWave) {
    const int i = i0 + lane;
    const int v = i < ne ? mc_viol<ND>(sh, i, -1, 0.f) : 0;
    if (i < ne) sh.vio[i] = (uint8_t)v;
    const uint64_t m = __ballot(v != 0);
    if (lane == 0) sh.vbits[i0 / 64] = m;
    nv += __popcll(m);
  }
  __syncthreads();
  // ---- min-conflicts --------------------------------------------------------------------
  const uint64_t seed = mc_mix((uint64_t)sh.el[0] ^ 0x5851f42d4c957f2dull ^ ((uint64_t)(R ? rep + 1u : 0u) << 40));
  uint32_t step = 0;
  uint32_t polled = ~0u;  // the entry's best finish, loaded at the previous poll
  float far = 0.25f;  // the start keys span about one error bound
  uint64_t* sk = reinterpret_cast<uint64_t*>(sh.hkey);  // p2 <= kHash / 2 entries
  uint64_t tk[4] = {0, 0, 0, 0};  // statistics: ticks per phase
  for (; step < budget && nv > 0; step++) {
    if (R && step % kMcPoll == 0) {  // another replica finished in fewer steps: stop
      if (step > (__builtin_amdgcn_readfirstlane(polled) >> 3)) break;
      polled = __hip_atomic_load(&ctl[ent], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    uint64_t tp = stats ? wall_clock64() : 0;
    if (step > 0 && step % kMcRenorm == 0) {  // keys := ranks (the order is unchanged)
      for (int i = lane; i < p2; i += kWave) sk[i] = i < ne ? mc_sortkey(sh.key[i], i) : ~0ull;
      __syncthreads();
      wave_bitonic(sk, p2, lane);
      for (int rk = lane; rk < ne; rk += kWave) sh.key[(uint32_t)sk[rk]] = (float)rk;
      __syncthreads();
      far = 0.25f * (float)ne;
      if (stats) {
        const uint64_t t2 = wall_clock64();
        tk[3] += t2 - tp;
        tp = t2;
      }
    }
    const uint64_t r = mc_mix(seed + (uint64_t)step * 0x9e3779b97f4a7c15ull);
    // a violated point drawn uniformly: the k-th set bit of the bitmap (word by a prefix
    // sum of the words' counts, bit by halving)
    int p;
    {
      const uint64_t w = lane < nwords ? sh.vbits[lane] : 0ull;
      const uint32_t cnt = (uint32_t)__popcll(w);
      const uint32_t incl = wave_incl_scan_dpp(cnt, lane);
      const uint32_t k = (uint32_t)(r >> 32) % (uint32_t)nv;
      const uint64_t hit = __ballot(incl > k);  // words whose prefix passes k
      const int wi = __ffsll((unsigned long long)hit) - 1;
      const uint64_t word = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(w >> 32), wi) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)w, wi);
      const uint32_t before = (uint32_t)__builtin_amdgcn_readlane((int)(incl - cnt), wi);
      p = wi * 64 + mc_kth_bit(word, (int)(k - before));
    }
    const bool walk = (uint32_t)(r & 0xffffu) % 100u < kMcWalkPct;
    // ---- records: lane mi * K + k, mover mi, direction k --------------------------------
    if (lane < (K + 1) * K) {
      const int mi = lane / K, k = lane % K;
      const int m = mi == 0 ? p : sh.nb[p][mi - 1];
      if (m >= 0) {
        const int c = sh.nb[m][k];
        uint64_t rec;
        if (c >= 0) {
          const float kq = sh.key[c];
          uint32_t base = 0;
#pragma unroll
          for (int k2 = 0; k2 < K; k2++) {
            if (k2 == (k ^ 1)) continue;
            const int c2 = sh.nb[c][k2];
            bool l = c2 == kNbLow;
            if (c2 >= 0) l = mc_before(sh.key[c2], c2, kq, c);
            base |= (uint32_t)l << k2;
          }
          // the neighbour's violation with the mover below it and above it
          const int tq = sh.tgt[c];
          const uint32_t vlo = tq >= 0 && mc_type<ND>(base | (1u << (k ^ 1))) != tq;
          const uint32_t vhi = tq >= 0 && mc_type<ND>(base) != tq;
          rec = ((uint64_t)__float_as_uint(kq) << 32) | ((uint32_t)c << 16) | ((uint32_t)sh.vio[c] << 4) |
                (vhi << 3) | (vlo << 2);
        } else {
          rec = c == kNbLow ? 1u : (c == kNbHigh ? 2u : 3u);
        }
        sh.tbl[mi][k] = rec;
        if (k == 0) {
          const uint32_t tb = sh.tabu[m];
          const bool tabu = tb != 0 && (uint16_t)(step + 1u - tb) < kMcTabu;
          const int ap = sh.apos[m];
          const bool has_lo = ap != 0xffff && ap > 0, has_hi = ap != 0xffff && ap + 1 < sh.na;
          const float alo = has_lo ? sh.key[sh.aord[ap - 1]] : 0.f, ahi = has_hi ? sh.key[sh.aord[ap + 1]] : 0.f;
          sh.mrec[mi][0] = ((uint64_t)__float_as_uint(sh.key[m]) << 32) | ((uint32_t)tabu << 23) |
                           ((uint32_t)has_hi << 22) | ((uint32_t)has_lo << 21) | (1u << 20) |
                           ((uint32_t)sh.vio[m] << 19) | ((uint32_t)(sh.tgt[m] + 1) << 16) | (uint32_t)m;
          sh.mrec[mi][1] = ((uint64_t)__float_as_uint(ahi) << 32) | __float_as_uint(alo);
        }
      } else if (k == 0) {
        sh.mrec[mi][0] = 0;
      }
    }
    __syncthreads();
    // each mover's plateau-neighbour keys in order (lane mi), its last one, and the
    // directions of its neighbours in lower cells
    if (lane <= K && ((sh.mrec[lane][0] >> 20) & 1u)) {
      float nk[K];
      uint32_t fixed_low = 0;
      float last = -kMcNone;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const uint64_t rec = sh.tbl[lane][k];
        const uint32_t kind = (uint32_t)rec & 3u;
        nk[k] = kind == 0 ? mc_hi_f(rec) : kMcNone;
        last = kind == 0 ? fmaxf(last, nk[k]) : last;
        fixed_low |= (uint32_t)(kind == 1) << k;
      }
#pragma unroll
      for (int pass = 0; pass < K; pass++)  // odd-even transposition sort
#pragma unroll
        for (int i = pass & 1; i + 1 < K; i += 2) {
          const float lo = fminf(nk[i], nk[i + 1]), hi = fmaxf(nk[i], nk[i + 1]);
          nk[i] = lo;
          nk[i + 1] = hi;
        }
#pragma unroll
      for (int k = 0; k < K; k++) sh.srt[lane][k] = nk[k];
      sh.srt[lane][K] = last;
      sh.mrec[lane][0] |= (uint64_t)fixed_low << 24;
    }
    __syncthreads();
    if (stats) {
      const uint64_t t2 = wall_clock64();
      tk[0] += t2 - tp;
      tp = t2;
    }
    // ---- candidates: lane + 64 j is slot (mover, position) ----------------------------
    static_assert(kMcSlots <= 256, "slot in the low byte of the key");
    uint32_t best = ~0u;
    float cands[kMcSlotsPerLane];
    uint32_t newv[kMcSlotsPerLane];
#pragma unroll
    for (int j = 0; j < kMcSlotsPerLane; j++) {
      cands[j] = 0.f;
      newv[j] = 0;
      const int s = lane + j * kWave;
      const int mi = s / kMcPos, pos = s % kMcPos;
      if (mi > K) continue;
      const uint64_t mr0 = sh.mrec[mi][0];
      if (!((mr0 >> 20) & 1u)) continue;
      const float first = sh.srt[mi][0];
      if (!(first < kMcNone)) continue;  // no plateau neighbour
      float c;
      if (pos == 0) c = first - far;
      else if (pos == 1) c = first - mc_near(first);
      else if (pos >= kMcPos - 2) {
        const float last = sh.srt[mi][K];
        c = pos == kMcPos - 2 ? last + mc_near(last) : last + far;
      } else {
        const int a = (pos - 2) / kMcGapPos, w = (pos - 2) % kMcGapPos;
        if (a + 1 >= K) continue;  // 2D: three gaps
        const float lo = sh.srt[mi][a], hi = sh.srt[mi][a + 1];
        if (!(hi < kMcNone) || !(hi > lo)) continue;
        const float span = hi - lo;
        c = w == 0 ? lo + span * (1.f / 1024.f)
                   : (w == kMcGapPos - 1 ? hi - span * (1.f / 1024.f) : lo + 0.5f * span);
      }
      if ((mr0 >> 21) & 3u) {  // an anchored point stays strictly between its neighbours in anchor order
        const uint64_t mr1 = sh.mrec[mi][1];
        const bool has_lo = (mr0 >> 21) & 1u, has_hi = (mr0 >> 22) & 1u;
        const float alo = __uint_as_float((uint32_t)mr1), ahi = mc_hi_f(mr1);
        if (has_lo && !(c > alo)) c = has_hi ? alo + (ahi - alo) * (1.f / 1024.f) : alo + mc_near(alo);
        if (has_hi && !(c < ahi)) c = has_lo ? ahi - (ahi - alo) * (1.f / 1024.f) : ahi - mc_near(ahi);
        if ((has_lo && !(c > alo)) || (has_hi && !(c < ahi))) continue;
      }
      cands[j] = c;
      const int m = (int)(mr0 & 0xffffu);
      // the mover's own type, then each plateau neighbour's, with the mover at c
      uint32_t low_m = (uint32_t)(mr0 >> 24) & 63u, nv_bits = 0;
      int dl = 0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const uint64_t rec = sh.tbl[mi][k];
        if (((uint32_t)rec & 3u) != 0) continue;
        const int q = (int)(((uint32_t)rec >> 16) & 0xffffu);
        const bool q_low = mc_before(mc_hi_f(rec), q, c, m);
        low_m |= (uint32_t)q_low << k;
        const int vq = (int)(((uint32_t)rec >> (q_low ? 3 : 2)) & 1u);
        nv_bits |= (uint32_t)vq << (1 + k);
        dl += vq - (int)(((uint32_t)rec >> 4) & 1u);
      }
      const int tm = (int)((mr0 >> 16) & 7u) - 1;
      const int vm = tm >= 0 && mc_type<ND>(low_m) != tm;
      nv_bits |= (uint32_t)vm;
      dl += vm - (int)((mr0 >> 19) & 1u);
      newv[j] = nv_bits;
      uint32_t rs = (uint32_t)r ^ ((uint32_t)s * 0x9e3779b1u);  // tie-break
      rs ^= rs >> 16;
      rs *= 0x7feb352du;
      rs ^= rs >> 15;
      // 32-bit key: (violation change, tabu, random) or random alone, then the slot
      const uint32_t pk = walk ? ((rs >> 8) << 8) | (uint32_t)s
                               : ((uint32_t)(dl + 16) << 27) | ((uint32_t)((mr0 >> 23) & 1u) << 26) |
                                     ((rs >> 14) << 8) | (uint32_t)s;
      best = pk < best ? pk : best;
    }
    best = mc_wave_min(best);
    if (stats) {
      const uint64_t t2 = wall_clock64();
      tk[1] += t2 - tp;
      tp = t2;
    }
    if (best == ~0u) continue;  // nothing can move (every mover pinned between anchors)
    const int s = (int)(best & 0xffu);
    const int js = s / kWave, ms = s / kMcPos;
    float cw = cands[0];
    uint32_t vw = newv[0];
#pragma unroll
    for (int j = 1; j < kMcSlotsPerLane; j++) {
      cw = js == j ? cands[j] : cw;
      vw = js == j ? newv[j] : vw;
    }
    const float c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cw), s % kWave));
    const uint32_t vbits = (uint32_t)__builtin_amdgcn_readlane((int)vw, s % kWave);
    // ---- the move; lane 0 updates the mover, lane 1 + k its neighbour k ----------------
    if (lane <= K) {
      const int m = (int)(sh.mrec[ms][0] & 0xffffu);
      int q = m;
      if (lane > 0) {
        const uint64_t rec = sh.tbl[ms][lane - 1];
        q = ((uint32_t)rec & 3u) == 0 ? (int)(((uint32_t)rec >> 16) & 0xffffu) : -1;
      }
      if (lane == 0) {
        sh.key[m] = c;
        sh.tabu[m] = (uint16_t)(step + 1u);
      }
      if (q >= 0) {
        const int vq = (vbits >> lane) & 1u;
        if (vq != (int)sh.vio[q]) {
          sh.vio[q] = (uint8_t)vq;
          const unsigned long long bit = 1ull << (q & 63);
          if (vq) atomicOr(reinterpret_cast<unsigned long long*>(&sh.vbits[q >> 6]), bit);
          else atomicAnd(reinterpret_cast<unsigned long long*>(&sh.vbits[q >> 6]), ~bit);
        }
      }
    }
    __syncthreads();
    {
      const uint64_t w = lane < nwords ? sh.vbits[lane] : 0ull;
      nv = (int)wave_sum_dpp((uint32_t)__popcll(w));
    }
    if (stats) tk[2] += wall_clock64() - tp;
  }
  // (statistics only on request: one same-address atomic per plateau serialises)
  if (lane == 0 && stats) {
    for (int k = 0; k < 4; k++) atomicAdd(&misc[kPpMcTicks + k], (uint32_t)tk[k]);
    atomicAdd(&misc[kPpMcSteps], step);
    atomicMax(&misc[kPpMcMaxSteps], step);
    atomicMax(&misc[kPpMcMaxTicks], (uint32_t)min<uint64_t>(wall_clock64() - t_start, 0xffffffffull));
    if (nv > 0 && !R && !esc) atomicAdd(&misc[kPpMcFailed], 1u);
  }
  if (nv > 0) {  // not finished: the plateau keeps its values
    if (R) {
      mc_replica_end<ND, PM>(sh, F, y, t, stage, ctl, nbad, ent, R, relist, cap, stats, misc, lane);
    } else if (esc) {  // to the replicas
      if (lane == 0) {
        const uint32_t pos = atomicAdd(&misc[kPpMcEsc], 1u);
        if (pos < nbad) esc[pos] = sh.el[0];
      }
    } else {  // its mismatches are listed again
      mc_relist<ND>(sh, F, y, t, relist, cap, misc, lane);
    }
    return;
  }
  // ---- order by key, then values ----------------------------------------------------------
  for (int i = lane; i < p2; i += kWave) sk[i] = i < ne ? mc_sortkey(sh.key[i], i) : ~0ull;
  __syncthreads();
  wave_bitonic(sk, p2, lane);
  if (lane == 0) {  // values into the key array, in key order
    int s0 = 0;
    float prev = sh.lo;
    for (int r = 0; r <= ne; r++) {
      const bool end = r == ne;
      const int i = end ? -1 : (int)(uint32_t)sk[r];
      if (end || sh.apos[i] != 0xffff) {
        const float nxt = end ? sh.hi : aval[sh.apos[i]];
        const int cnt = r - s0;
        for (int k = 0; k < cnt; k++) {
          const int j = (int)(uint32_t)sk[s0 + k];
          sh.key[j] = prev + (nxt - prev) * ((float)(k + 1) / (float)(cnt + 1));
        }
        if (!end) {
          sh.key[i] = nxt;
          prev = nxt;
        }
        s0 = r + 1;
      }
    }
    if (stats && !R) atomicAdd(&misc[kPpMcSolved], 1u);
  }
  __syncthreads();
  if (R) {  // staged; the entry's last replica writes the winner's
    float* dst = stage + ((size_t)ent * R + rep) * PM;
    for (int i = lane; i < ne; i += kWave) dst[i] = sh.key[i];
    if (lane == 0) atomicMin(&ctl[ent], (step << 3) | rep);
    mc_replica_end<ND, PM>(sh, F, y, t, stage, ctl, nbad, ent, R, relist, cap, stats, misc, lane);
    return;
  }
  for (int i = lane; i < ne; i += kWave) y[sh.el[i]] = sh.key[i];
  __syncthreads();
  // what the written values leave mismatched (equal anchor values with points between
  // them, float spacing) is listed again
  mc_relist<ND>(sh, F, y, t, relist, cap, misc, lane);
}

}  // namespace

#define PP_MC_LAUNCH(PMV)                                                                                        \
  do {                                                                                                           \
    if (geo.fields) {                                                                                            \
      const GeoB gb_ = geob(geo);                                                                                \
      if (geo.ndim == 2)                                                                                         \
        hipLaunchKernelGGL((k_pp_platmc<2, GeoB, PMV>), dim3(grid), dim3(kWave), 0, st, gb_, bad, nbad, d, t, y,   \
                           guide, claim, budget, variant == 2, relist, cap, stats, misc, R, esc, stage, ctl);                                                          \
      else                                                                                                       \
        hipLaunchKernelGGL((k_pp_platmc<3, GeoB, PMV>), dim3(grid), dim3(kWave), 0, st, gb_, bad, nbad, d, t, y,   \
                           guide, claim, budget, variant == 2, relist, cap, stats, misc, R, esc, stage, ctl);                                                          \
    } else {                                                                                                     \
      const Geo1 g1_ = geo1(geo);                                                                                \
      if (geo.ndim == 2)                                                                                         \
        hipLaunchKernelGGL((k_pp_platmc<2, Geo1, PMV>), dim3(grid), dim3(kWave), 0, st, g1_, bad, nbad, d, t, y,   \
                           guide, claim, budget, variant == 2, relist, cap, stats, misc, R, esc, stage, ctl);                                                          \
      else                                                                                                       \
        hipLaunchKernelGGL((k_pp_platmc<3, Geo1, PMV>), dim3(grid), dim3(kWave), 0, st, g1_, bad, nbad, d, t, y,   \
                           guide, claim, budget, variant == 2, relist, cap, stats, misc, R, esc, stage, ctl);                                                          \
    }                                                                                                            \
  } while (0)

hipError_t pp_platmc(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, const float* d, const uint8_t* t, float* y,
                     const float* guide, uint32_t* claim, int variant, uint32_t budget, uint32_t* relist,
                     uint32_t cap, bool stats, uint32_t* misc, uint32_t* esc, hipStream_t st) {
  if (!nbad) return hipSuccess;
  const uint32_t R = 0, grid = nbad;
  float* stage = nullptr;
  uint32_t* ctl = nullptr;
  if (variant == 2) PP_MC_LAUNCH(kMcLarge);
  else if (variant == 1) PP_MC_LAUNCH(kMcSmall);
  else PP_MC_LAUNCH(kMcTiny);
  return hipGetLastError();
}

uint32_t pp_platmc_capacity(int variant) { return variant == 2 ? kMcLarge : (variant == 1 ? kMcSmall : kMcTiny); }

hipError_t pp_platmc_replicas(const PpGeo& geo, const uint32_t* ent, uint32_t nent, const float* d, const uint8_t* t,
                              float* y, const float* guide, int variant, uint32_t budget, uint32_t R, float* stage,
                              uint32_t* ctl, uint32_t* relist, uint32_t cap, bool stats, uint32_t* misc,
                              hipStream_t st) {
  if (!nent) return hipSuccess;
  if (R < 1 || R > 8) return hipErrorInvalidValue;
  const uint32_t* bad = ent;
  const uint32_t nbad = nent, grid = nent * R;
  uint32_t* claim = nullptr;
  uint32_t* esc = nullptr;
  if (variant == 2) PP_MC_LAUNCH(kMcLarge);
  else if (variant == 1) PP_MC_LAUNCH(kMcSmall);
  else PP_MC_LAUNCH(kMcTiny);
  return hipGetLastError();
}

}  // namespace tszp
