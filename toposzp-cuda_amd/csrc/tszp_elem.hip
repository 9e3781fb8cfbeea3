// tszp_elem.hip -- SZp streams partitioned by elements rather than by blocks:
//   * the threadblock codec (SZp/src/szp_float.cc:1445-1590, SZp/src/szp_double.cc:148-318;
//     CUDA SZp/src/szp_cuda_compress.cu:245-354, szp_cuda_decompress.cu:597-665), for
//     float and double data: chunk t holds one 4-byte anchor q[lo] and then blocks of up
//     to bs deltas from lo + 1 on, chained across blocks (each delta is q[e] - q[e - 1]);
//   * the double random-access codec (SZp/src/szp_double.cc:759-920): blocks of bs
//     elements from each chunk start, every block with its own anchor and a bc byte
//     (also for a one-element block).
// Chunk t of T covers [t * (n / T), (t + 1) * (n / T)), the last chunk up to n.
//
// A "unit" is what one SZp block contributes to the stream: an optional anchor, an
// optional bc byte, and the signs and magnitudes of its m deltas.  Units are numbered
// chunk by chunk, so unit u of a T-chunk stream lies at byte blk[u] of the payload.
//
// Compression is one pass: a workgroup owns a tile of consecutive units (S lanes per
// unit, EPL deltas per lane), sizes them, finds the tile's byte offset by a decoupled
// look-back over tiles, packs the tile in LDS and writes it out with word stores.
// Decompression loads a tile's bytes into LDS, decodes every delta, scans them within
// each unit across lanes, then across the tile's units with the anchors as restarts,
// and takes the running value at the tile start from a segmented look-back (needed
// only when a tile starts inside a threadblock chunk).
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

template <int S, int EPL>
struct ElemGeom {
  static constexpr int G = kWave / S;                   // units per wave step
  static constexpr int SPW = EPL >= 16 ? 1 : 16 / EPL;  // steps per wave
  static constexpr int BPT = kWavesPerBlock * SPW * G;  // units per tile
  static constexpr int MMAX = S * EPL;                  // most deltas of a unit
  static constexpr int WORST = 4 + 1 + (MMAX + 7) / 8 + 4 * MMAX;
};

// Unit u: anchor element (== s - 1, when `anchor`), first delta element s, m deltas.
struct Unit {
  uint32_t s, m;
  bool anchor, has_bc;
};

template <bool RA>
__device__ __forceinline__ Unit elem_unit(const ElemArgs& a, uint32_t u) {
  Unit r{0u, 0u, false, false};
  if (u >= a.nu) return r;
  uint32_t t, k;
  const uint32_t full = (a.T - 1u) * a.ukc;  // units of the chunks before the last
  if (u < full) {
    t = fdiv(u, a.div_ukc);
    k = u - t * a.ukc;
  } else {
    t = a.T - 1u;
    k = u - full;
  }
  const uint32_t lo = t * a.ce, hi = t == a.T - 1u ? a.n : lo + a.ce;
  if constexpr (RA) {
    const uint32_t e0 = lo + k * a.bs;
    r.s = e0 + 1u;
    r.m = min(a.bs, hi - e0) - 1u;
    r.anchor = true;
    r.has_bc = true;
  } else {
    r.s = lo + 1u + k * a.bs;
    r.m = r.s < hi ? min(a.bs, hi - r.s) : 0u;
    r.anchor = k == 0u;
    r.has_bc = r.m != 0u;
  }
  return r;
}

__device__ __forceinline__ uint32_t unit_bytes(const Unit& un, unsigned bc) {
  uint32_t b = un.anchor ? 4u : 0u;
  if (un.has_bc) b += 1u + (bc ? ((un.m + 7u) >> 3) + mag_bytes(un.m, bc) : 0u);
  return b;
}

template <typename TF>
__device__ __forceinline__ int elem_quant(TF x, double inv) {
  return cvt_i32((double)x * inv);
}

// ---------------------------------------------------------------------------
// compression
// ---------------------------------------------------------------------------
template <typename TF, bool RA, int S, int EPL>
__global__ __launch_bounds__(kBlockThreads) void k_elem_compress(ElemArgs a) {
  using Gm = ElemGeom<S, EPL>;
  constexpr int G = Gm::G, SPW = Gm::SPW, BPT = Gm::BPT;
  constexpr int STAGE_WORDS = (BPT * Gm::WORST + 3) / 4 + 2;
  __shared__ uint32_t stage[STAGE_WORDS];
  __shared__ uint32_t blk_off[BPT + 1];
  __shared__ uint64_t sh_prefix;
  __shared__ uint32_t sh_tile;

  const int tid = threadIdx.x;
  const int w = tid / kWave, lane = tid & (kWave - 1);
  const int g = lane / S, ls = lane & (S - 1);
  const TF* in = reinterpret_cast<const TF*>(a.in);

  if (tid == 0) sh_tile = tile_id(a.tile_counter);
  for (int i = tid; i < STAGE_WORDS; i += kBlockThreads) stage[i] = 0;
  __syncthreads();
  const uint32_t tile = sh_tile;

  // ---- loads: the lane's deltas of every step, and the element before each unit ----
  TF raw[SPW][EPL];
  TF before[SPW];
  // float data, four deltas per lane: the lane's elements come from two aligned 16-byte
  // loads when both lie inside the array (fewer, wider requests than four 4-byte ones;
  // for double data three 16-byte loads per lane were slower than four 8-byte ones)
  constexpr bool kVecF = sizeof(TF) == 4 && EPL == 4;
  const bool in16 = (reinterpret_cast<uintptr_t>(a.in) & 15u) == 0;
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const Unit un = elem_unit<RA>(a, tile * BPT + (uint32_t)((w * SPW + s) * G + g));
    bool vec = false;
    if constexpr (kVecF) {
      const uint32_t e0 = un.s + (uint32_t)(ls * EPL);
      const uint32_t b4 = e0 & ~3u;
      vec = in16 && (uint32_t)(ls * EPL + EPL) <= un.m && b4 + 8u <= a.n;
      if (vec) {
        const uint4* v4 = reinterpret_cast<const uint4*>(a.in) + (b4 >> 2);
        const uint4 lo = v4[0], hi = v4[1];
        const uint32_t w8[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        const uint32_t m = e0 & 3u;
#pragma unroll
        for (int r = 0; r < EPL; r++) {
          const uint32_t k = m + (uint32_t)r;  // 0..6
          const uint32_t v = k == 0 ? w8[0] : k == 1 ? w8[1] : k == 2 ? w8[2] : k == 3 ? w8[3]
                           : k == 4 ? w8[4] : k == 5 ? w8[5] : w8[6];
          raw[s][r] = __uint_as_float(v);
        }
      }
    }
    if (!vec) {
#pragma unroll
      for (int r = 0; r < EPL; r++) {
        const uint32_t j = (uint32_t)(ls * EPL + r);
        raw[s][r] = j < un.m ? in[un.s + j] : (TF)0;
      }
    }
    before[s] = (ls == 0 && (un.m || un.anchor)) ? in[un.s - 1u] : (TF)0;
  }

  // ---- quantize, deltas, bit counts, unit sizes ------------------------------------
  int dl[SPW][EPL];
  int qa[SPW];  // quantized element before the unit (the anchor where there is one)
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    const Unit un = elem_unit<RA>(a, tile * BPT + lb);
    int q[EPL];
#pragma unroll
    for (int r = 0; r < EPL; r++) q[r] = elem_quant<TF>(raw[s][r], a.inv);
    const int qb = elem_quant<TF>(before[s], a.inv);
    const int up = __shfl_up(q[EPL - 1], 1, kWave);
    const int prev0 = ls == 0 ? qb : up;
    qa[s] = qb;
    unsigned mx = 0;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      const int p = r ? q[r - 1] : prev0;
      const unsigned d = (unsigned)q[r] - (unsigned)p;
      dl[s][r] = j < un.m ? (int)d : 0;
      const unsigned mag = dl[s][r] < 0 ? 0u - (unsigned)dl[s][r] : (unsigned)dl[s][r];
      mx = mag > mx ? mag : mx;
    }
    mx = seg_max<S>(mx);
    if (ls == 0) blk_off[lb] = unit_bytes(un, bitcount(mx));
  }
  __syncthreads();

  // ---- tile scan of the unit sizes and the tile's byte offset (wave 0) ---------------
  if (w == 0) {
    constexpr int PER = (BPT + kWave - 1) / kWave;
    uint32_t loc[PER];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = lane * PER + k;
      loc[k] = i < BPT ? blk_off[i] : 0u;
      sum += loc[k];
    }
    const uint32_t incl = seg_incl_scan<kWave>(sum, lane);
    uint32_t run = incl - sum;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = lane * PER + k;
      if (i < BPT) blk_off[i] = run;
      run += loc[k];
    }
    const uint32_t tile_bytes = __shfl(incl, kWave - 1, kWave);
    if (lane == 0) blk_off[BPT] = tile_bytes;
    const uint64_t pb = lookback(a.st_bytes, tile, tile_bytes, &a.res->error);
    if (lane == 0) {
      sh_prefix = pb;
      if (tile == a.ntiles - 1) a.res->payload_bytes = pb + tile_bytes;
    }
  }
  __syncthreads();

  // ---- pack into the LDS stage ---------------------------------------------------
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    const Unit un = elem_unit<RA>(a, tile * BPT + lb);
    unsigned mags[EPL];
    bool negs[EPL];
    unsigned mx = 0;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      negs[r] = dl[s][r] < 0;
      mags[r] = negs[r] ? 0u - (unsigned)dl[s][r] : (unsigned)dl[s][r];
      mx = mags[r] > mx ? mags[r] : mx;
    }
    mx = seg_max<S>(mx);
    const unsigned bc = bitcount(mx);
    const unsigned off = blk_off[lb];
    const unsigned A = un.anchor ? 4u : 0u;
    if (ls == 0) {
      if (un.anchor) lds_or_bytes(stage, off, (uint32_t)qa[s], 4);
      if (un.has_bc) lds_or_bytes(stage, off + A, bc, 1);
    }
    if (!bc) continue;
    const unsigned m = un.m;
    const unsigned sbase = off + A + 1u;
    const unsigned nfull = bc >> 3, rem = bc & 7u;
    const unsigned fbase = sbase + ((m + 7u) >> 3);
    const unsigned rbase = fbase + nfull * m;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      if (j >= m) continue;
      if (negs[r]) lds_or_bits_msb(stage, sbase, j, 1u, 1u);
      if (nfull) lds_or_bytes(stage, fbase + j * nfull, mags[r] >> rem, nfull);
      if (rem) lds_or_bits_msb(stage, rbase, j * rem, mags[r], rem);
    }
  }
  __syncthreads();

  // ---- stage -> stream (aligned words in the middle), unit offsets ------------------
  const uint64_t g0 = sh_prefix;
  const uint32_t len = blk_off[BPT];
  if (g0 + len > a.cap) {
    if (tid == 0) atomicOr(&a.res->error, TSZP_ERR_OVERFLOW);
  } else if (len) {
    uint8_t* dst = a.payload;
    const uint64_t a0 = (g0 + 3) & ~3ull, e0 = g0 + len;
    const uint64_t a1 = e0 & ~3ull;
    if (a0 >= e0 || a1 <= a0) {
      for (uint32_t i = tid; i < len; i += kBlockThreads) dst[g0 + i] = (uint8_t)lds_byte(stage, i);
    } else {
      const uint32_t head = (uint32_t)(a0 - g0), tail = (uint32_t)(e0 - a1);
      if (tid < (int)head) dst[g0 + tid] = (uint8_t)lds_byte(stage, tid);
      if (tid < (int)tail) dst[a1 + tid] = (uint8_t)lds_byte(stage, (uint32_t)(a1 - g0) + tid);
      const uint32_t nw = (uint32_t)((a1 - a0) >> 2);
      uint32_t* dw = reinterpret_cast<uint32_t*>(dst + a0);
      for (uint32_t i = tid; i < nw; i += kBlockThreads) {
        const uint32_t sb = head + 4 * i;
        dw[i] = lds_byte(stage, sb) | (lds_byte(stage, sb + 1) << 8) | (lds_byte(stage, sb + 2) << 16) |
                (lds_byte(stage, sb + 3) << 24);
      }
    }
  }
  for (int lb = tid; lb < BPT; lb += kBlockThreads) {
    const uint32_t u = tile * BPT + lb;
    if (u < a.nu) a.blk[u] = (uint32_t)(g0 + blk_off[lb]);
  }
  if (tile == a.ntiles - 1 && tid == 0) a.blk[a.nu] = (uint32_t)(g0 + len);
}

// ---------------------------------------------------------------------------
// decompression
// ---------------------------------------------------------------------------
template <typename TF>
__device__ __forceinline__ TF elem_value(int q, TF eb);
template <>
__device__ __forceinline__ float elem_value<float>(int q, float eb) {
  return __fmul_rn(__int2float_rn(q), eb);  // (float)q * eb in float
}
template <>
__device__ __forceinline__ double elem_value<double>(int q, double eb) {
  return __dmul_rn((double)q, eb);
}

template <typename TF, bool RA, int S, int EPL>
__global__ __launch_bounds__(kBlockThreads) void k_elem_decode(ElemDecodeArgs a) {
  using Gm = ElemGeom<S, EPL>;
  constexpr int G = Gm::G, SPW = Gm::SPW, BPT = Gm::BPT;
  constexpr int TILE_WORDS = (BPT * Gm::WORST + 3) / 4 + 2;
  __shared__ uint32_t tw[TILE_WORDS];
  __shared__ uint32_t usum[BPT];    // anchor (if any) + sum of the unit's deltas
  __shared__ uint32_t ubase[BPT];   // value before the unit's first delta
  __shared__ uint8_t ureset[BPT];   // the unit starts with an anchor
  __shared__ uint32_t sh_tile, sh_bad;

  const int tid = threadIdx.x;
  const int w = tid / kWave, lane = tid & (kWave - 1);
  const int g = lane / S, ls = lane & (S - 1);
  const ElemArgs& ge = a.geo;

  if (tid == 0) {
    sh_tile = tile_id(a.tile_counter);
    sh_bad = 0;
  }
  __syncthreads();
  const uint32_t tile = sh_tile;
  const uint32_t u0 = tile * BPT;
  const uint32_t u1 = min(ge.nu, u0 + (uint32_t)BPT);
  const uint32_t t0 = a.blk[u0], t1 = a.blk[u1];
  const uint32_t wbase = t0 & ~3u;
  const uint32_t nwords = ((t1 + 3u) & ~3u) / 4u - wbase / 4u;
  const bool bad_tile = t1 < t0 || nwords > (uint32_t)TILE_WORDS;
  if (!bad_tile) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.payload + wbase);
    for (uint32_t i0 = 0; i0 < nwords; i0 += 8 * kBlockThreads) {
      uint32_t v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint32_t i = i0 + k * kBlockThreads + tid;
        v[k] = i < nwords ? src[i] : 0u;
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint32_t i = i0 + k * kBlockThreads + tid;
        if (i < nwords) tw[i] = v[k];
      }
    }
  } else if (tid == 0) {
    atomicOr(a.err, TSZP_ERR_STREAM);
  }
  __syncthreads();

  // ---- deltas, in-unit prefix sums ----------------------------------------------------
  int part[SPW][EPL];  // value of delta j relative to the value before the unit
  int anc[SPW];
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    Unit un = elem_unit<RA>(ge, u0 + lb);
    if (bad_tile) un = Unit{0u, 0u, false, false};
    uint32_t o = (un.m || un.anchor) ? a.blk[u0 + lb] - wbase : 0u;
    if (o > t1 - wbase) {  // a side-car that does not fit the tile
      if (ls == 0) sh_bad = 1;
      un = Unit{0u, 0u, false, false};
      o = 0;
    }
    int anchor = 0;
    if (un.anchor)
      anchor = (int)(lds_byte(tw, o) | (lds_byte(tw, o + 1) << 8) | (lds_byte(tw, o + 2) << 16) |
                     (lds_byte(tw, o + 3) << 24));
    const unsigned A = un.anchor ? 4u : 0u;
    unsigned bc = un.has_bc ? lds_byte(tw, o + A) : 0u;
    if (bc > 32) {
      if (ls == 0) sh_bad = 1;
      bc = 0;
    }
    const unsigned m = un.m;
    const unsigned sbase = o + A + 1u;
    const unsigned nfull = bc >> 3, rem = bc & 7u;
    const unsigned fbase = sbase + ((m + 7u) >> 3);
    const unsigned rbase = fbase + nfull * m;
    int lane_sum = 0;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      int d = 0;
      if (bc && j < m) {
        const unsigned neg = (lds_byte(tw, sbase + (j >> 3)) >> (7u - (j & 7u))) & 1u;
        unsigned mag = 0;
        for (unsigned kk = 0; kk < nfull; kk++) mag |= lds_byte(tw, fbase + j * nfull + kk) << (8u * kk);
        mag <<= rem;
        if (rem) {
          const unsigned pos = j * rem, by = rbase + (pos >> 3), sh = pos & 7u;
          unsigned win = lds_byte(tw, by) << 8;
          if (sh + rem > 8u) win |= lds_byte(tw, by + 1);
          mag |= (win >> (16u - rem - sh)) & ((1u << rem) - 1u);
        }
        d = (int)(neg ? 0u - mag : mag);
      }
      lane_sum = (int)((unsigned)lane_sum + (unsigned)d);
      part[s][r] = lane_sum;
    }
    const int incl = seg_incl_scan<S>(lane_sum, lane);
    const int excl = (int)((unsigned)incl - (unsigned)lane_sum);
#pragma unroll
    for (int r = 0; r < EPL; r++) part[s][r] = (int)((unsigned)part[s][r] + (unsigned)excl);
    const int total = __shfl(incl, (lane & ~(S - 1)) + S - 1, kWave);
    anc[s] = anchor;
    if (ls == 0) {
      usum[lb] = (uint32_t)anchor + (uint32_t)total;
      ureset[lb] = un.anchor ? 1 : 0;
    }
  }
  __syncthreads();

  // ---- segmented scan over the tile's units, anchors restart (wave 0) -----------------
  if (w == 0) {
    constexpr int PER = (BPT + kWave - 1) / kWave;
    uint32_t v = 0;
    bool f = false;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = lane * PER + k;
      if (i < BPT) {
        const bool rf = ureset[i] != 0;
        v = rf ? usum[i] : v + usum[i];
        f = f || rf;
      }
    }
    // inclusive segmented scan of (f, v) over the lanes
    uint32_t iv = v;
    bool ifl = f;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t pv = __shfl_up(iv, o, kWave);
      const int pf = __shfl_up((int)ifl, o, kWave);
      if (lane >= o) {
        iv = ifl ? iv : pv + iv;
        ifl = ifl || pf;
      }
    }
    const uint32_t tv = __shfl(iv, kWave - 1, kWave);
    const bool tf = __shfl((int)ifl, kWave - 1, kWave) != 0;
    uint32_t ev = __shfl_up(iv, 1, kWave);
    bool ef = __shfl_up((int)ifl, 1, kWave) != 0;
    if (lane == 0) {
      ev = 0;
      ef = false;
    }
    // running value at the tile start: not needed when the tile's first unit restarts
    uint64_t x = 0;
    if (ureset[0]) {
      if (lane == 0) st_status(&a.st[tile], kFlagPre | (uint64_t)tv);
    } else {
      x = lookback_seg(a.st, tile, (tf ? kSegReset : 0ull) | (uint64_t)tv, a.err);
    }
    uint32_t run = ef ? ev : (uint32_t)x + ev;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = lane * PER + k;
      if (i < BPT) {
        const bool rf = ureset[i] != 0;
        // the unit's deltas start from its anchor, or from the running value
        const uint32_t b = rf ? 0u : run;
        ubase[i] = b;
        run = rf ? usum[i] : run + usum[i];
      }
    }
  }
  __syncthreads();

  // ---- values ----------------------------------------------------------------------
  TF* out = reinterpret_cast<TF*>(a.out);
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    Unit un = elem_unit<RA>(ge, u0 + lb);
    if (bad_tile) continue;
    const uint32_t base = un.anchor ? (uint32_t)anc[s] : ubase[lb];
    if (un.anchor && ls == 0) out[un.s - 1u] = elem_value<TF>(anc[s], (TF)a.eb);
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      if (j < un.m) out[un.s + j] = elem_value<TF>((int)(base + (uint32_t)part[s][r]), (TF)a.eb);
    }
  }
  if (tid == 0 && sh_bad) atomicOr(a.err, TSZP_ERR_STREAM);
}

// header: [eb big-endian, 4 or 8 bytes][u64 off[T]], off[t] = byte offset of chunk t's
// first unit
__global__ void k_elem_header(uint8_t* stream, int dsize, unsigned T, uint32_t ukc, const uint32_t* blk,
                              double eb) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    if (dsize == 4) {
      const uint32_t u = __float_as_uint((float)eb);
      for (int k = 0; k < 4; k++) stream[k] = (uint8_t)(u >> (8 * (3 - k)));
    } else {
      const uint64_t u = (uint64_t)__double_as_longlong(eb);
      for (int k = 0; k < 8; k++) stream[k] = (uint8_t)(u >> (8 * (7 - k)));
    }
  }
  if (t >= T) return;
  const uint64_t off = blk[(size_t)t * ukc];
  uint8_t* o = stream + dsize + 8 * (size_t)t;
  for (int k = 0; k < 8; k++) o[k] = (uint8_t)(off >> (8 * k));
}

// q[i] - q[i - 1] with q[-1] = 0 and q = C cast of x * inv: the quantization helpers
// (SZp/src/szp_cuda_compress.cu:360-377); four points per thread, q of the point before
// recomputed rather than exchanged
__global__ __launch_bounds__(kBlockThreads) void k_predict_quant(const float* __restrict__ x, uint32_t n, double inv,
                                                                int32_t* __restrict__ out) {
  const uint32_t groups = (n + 3u) / 4u;
  for (uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x; i < groups; i += gridDim.x * kBlockThreads) {
    const uint32_t e0 = 4u * i;
    int prev = e0 ? cvt_i32((double)x[e0 - 1u] * inv) : 0;
#pragma unroll
    for (uint32_t r = 0; r < 4u; r++) {
      if (e0 + r >= n) break;
      const int q = cvt_i32((double)x[e0 + r] * inv);
      out[e0 + r] = (int)((uint32_t)q - (uint32_t)prev);
      prev = q;
    }
  }
}

hipError_t launch_predict_quant(const float* x, uint32_t n, double inv, int32_t* out, hipStream_t st) {
  uint64_t b = ((uint64_t)n + 4u * kBlockThreads - 1) / (4u * kBlockThreads);
  if (b > 16384) b = 16384;
  hipLaunchKernelGGL(k_predict_quant, dim3(b ? (unsigned)b : 1u), dim3(kBlockThreads), 0, st, x, n, inv, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------
template <typename TF, bool RA>
static hipError_t launch_elem_compress_t(const ElemArgs& a, int S, int EPL, hipStream_t st) {
#define TSZP_CASE(SS, EE)                                                                               \
  if (S == SS && EPL == EE) {                                                                           \
    hipLaunchKernelGGL((k_elem_compress<TF, RA, SS, EE>), dim3(a.ntiles), dim3(kBlockThreads), 0, st, a); \
    return hipGetLastError();                                                                           \
  }
  TSZP_CASE(1, 1) TSZP_CASE(2, 1) TSZP_CASE(4, 1) TSZP_CASE(8, 1) TSZP_CASE(16, 1)
  TSZP_CASE(8, 4) TSZP_CASE(16, 4) TSZP_CASE(64, 2) TSZP_CASE(64, 4) TSZP_CASE(64, 8)
  TSZP_CASE(64, 16) TSZP_CASE(64, 32)
#undef TSZP_CASE
  return hipErrorInvalidValue;
}

template <typename TF, bool RA>
static hipError_t launch_elem_decode_t(const ElemDecodeArgs& a, int S, int EPL, hipStream_t st) {
#define TSZP_CASE(SS, EE)                                                                                 \
  if (S == SS && EPL == EE) {                                                                             \
    hipLaunchKernelGGL((k_elem_decode<TF, RA, SS, EE>), dim3(a.geo.ntiles), dim3(kBlockThreads), 0, st, a); \
    return hipGetLastError();                                                                             \
  }
  TSZP_CASE(1, 1) TSZP_CASE(2, 1) TSZP_CASE(4, 1) TSZP_CASE(8, 1) TSZP_CASE(16, 1)
  TSZP_CASE(8, 4) TSZP_CASE(16, 4) TSZP_CASE(64, 2) TSZP_CASE(64, 4) TSZP_CASE(64, 8)
  TSZP_CASE(64, 16) TSZP_CASE(64, 32)
#undef TSZP_CASE
  return hipErrorInvalidValue;
}

uint32_t elem_units(uint64_t n, uint32_t bs, uint32_t T, bool ra, uint32_t* ukc) {
  const uint64_t ce = n / T, cl = n - (uint64_t)(T - 1) * ce;
  auto units = [&](uint64_t len) -> uint64_t {
    if (!len) return 0;
    if (ra) return (len + bs - 1) / bs;
    return len == 1 ? 1 : (len - 1 + bs - 1) / bs;
  };
  const uint64_t uc = units(ce), total = (uint64_t)(T - 1) * uc + units(cl);
  if (ukc) *ukc = (uint32_t)uc;
  return total > 0xffffffffull ? 0xffffffffu : (uint32_t)total;
}

uint32_t elem_tiles(uint32_t nu, uint32_t bs) {
  int S, EPL;
  uint32_t bpt;
  block_geometry(bs, &S, &EPL, &bpt);
  return (nu + bpt - 1) / bpt;
}

hipError_t launch_elem_compress(int dsize, bool ra, const ElemArgs& a, hipStream_t st) {
  int S, EPL;
  uint32_t bpt;
  block_geometry(a.bs, &S, &EPL, &bpt);
  if (dsize == 4 && !ra) return launch_elem_compress_t<float, false>(a, S, EPL, st);
  if (dsize == 8 && !ra) return launch_elem_compress_t<double, false>(a, S, EPL, st);
  if (dsize == 8 && ra) return launch_elem_compress_t<double, true>(a, S, EPL, st);
  return hipErrorInvalidValue;
}

hipError_t launch_elem_decode(int dsize, bool ra, const ElemDecodeArgs& a, hipStream_t st) {
  int S, EPL;
  uint32_t bpt;
  block_geometry(a.geo.bs, &S, &EPL, &bpt);
  if (dsize == 4 && !ra) return launch_elem_decode_t<float, false>(a, S, EPL, st);
  if (dsize == 8 && !ra) return launch_elem_decode_t<double, false>(a, S, EPL, st);
  if (dsize == 8 && ra) return launch_elem_decode_t<double, true>(a, S, EPL, st);
  return hipErrorInvalidValue;
}

hipError_t launch_elem_header(uint8_t* stream, int dsize, unsigned T, uint32_t ukc, const uint32_t* blk, double eb,
                              hipStream_t st) {
  const unsigned threads = T < 64 ? 64 : 256;
  hipLaunchKernelGGL(k_elem_header, dim3((T + threads - 1) / threads), dim3(threads), 0, st, stream, dsize, T, ukc,
                     blk, eb);
  return hipGetLastError();
}

}  // namespace tszp
