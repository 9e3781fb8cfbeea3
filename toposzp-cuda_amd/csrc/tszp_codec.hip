// tszp_codec.hip -- SZp block codec kernels for gfx950.
//
// One kernel compresses a field in a single pass over HBM:
//   quantize (+ critical-point classification on the fly) -> per-block delta,
//   bit count and byte size -> tile scan -> decoupled look-back for the tile's
//   global byte offset (and critical-point counts) -> bit packing in LDS ->
//   coalesced copy of the tile's bytes to the stream.
// The byte stream is the reference CPU writer's (SZp/src/szp_float.cc:998-1318 for
// topology, :2061-2292 plain, :521-996 sort positions); see the block layout in
// put_block() of oracle/oracle.c.  A side-car of u32 block offsets is written next
// to the stream so that decoding is fully parallel.
//
// Work decomposition: a workgroup (4 waves) owns a tile of BPT consecutive SZp
// blocks.  Within a wave, one "step" covers G = 64 / S blocks, each block spread
// over S lanes with EPL consecutive elements per lane (S * EPL >= blockSize).
#include <cstdlib>

#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

template <int S, int EPL>
struct Geom {
  static constexpr int G = kWave / S;                        // blocks per wave step
  static constexpr int SPW = EPL >= 16 ? 1 : 16 / EPL;       // steps per wave
  static constexpr int BPT = kWavesPerBlock * SPW * G;       // blocks per tile
  static constexpr int BSMAX = S * EPL;                      // largest block size
  static constexpr int worst_block(bool topo) {
    return 5 + (BSMAX - 1 + 7) / 8 + 4 * (BSMAX - 1) + (topo ? (2 * BSMAX + 7) / 8 : 0);
  }
};

// ---------------------------------------------------------------------------
// element access per mode
// ---------------------------------------------------------------------------
template <int MODE>
__device__ __forceinline__ int quantize_raw(const CompressArgs& a, uint32_t raw) {
  if constexpr (MODE == kModeInt || MODE == kModeTopoInt) {
    return (int)raw;
  } else {
    return cvt_i32((double)__uint_as_float(raw) * a.inv);
  }
}

// Interior test and neighbour offsets of element e; the caller loads the
// neighbours (many at once) and calls type_of().
__device__ __forceinline__ bool interior_at(const Grid& g, uint32_t e) {
  if (g.ndim == 2) {
    const uint32_t i = fdiv(e, g.div_row), j = e - i * g.d1;
    return i != 0 && j != 0 && i + 1 < g.d0 && j + 1 < g.d1;
  }
  if (g.ndim == 3) {
    const uint32_t x = fdiv(e, g.div_plane), r = e - x * g.plane;
    const uint32_t y = fdiv(r, g.div_row), z = r - y * g.d2;
    return x != 0 && y != 0 && z != 0 && x + 1 < g.d0 && y + 1 < g.d1 && z + 1 < g.d2;
  }
  return false;
}

// critical-point type from the centre and its neighbours (0 regular, 1 max, 2 min,
// 3 saddle); SZp/src/szp_float.cc:73-90 (4 neighbours: up, down, left, right) and
// SZp/src/szp_topology_3d.cc:79-98 (6 neighbours: -x, +x, -y, +y, -z, +z)
__device__ __forceinline__ int type_of(int ndim, float c, const float* nb) {
  if (ndim == 2) {
    const float up = nb[0], dn = nb[1], lf = nb[2], rt = nb[3];
    if (c > up && c > dn && c > lf && c > rt) return 1;
    if (c < up && c < dn && c < lf && c < rt) return 2;
    if ((c < up && c < dn && c > lf && c > rt) || (c > up && c > dn && c < lf && c < rt)) return 3;
    return 0;
  }
  const float xm = nb[0], xp = nb[1], ym = nb[2], yp = nb[3], zm = nb[4], zp = nb[5];
  if (c > xm && c > xp && c > ym && c > yp && c > zm && c > zp) return 1;
  if (c < xm && c < xp && c < ym && c < yp && c < zm && c < zp) return 2;
  const int hi = (c > xm && c > xp) + (c > ym && c > yp) + (c > zm && c > zp);
  const int lo = (c < xm && c < xp) + (c < ym && c < yp) + (c < zm && c < zp);
  return (hi >= 1 && lo >= 1) ? 3 : 0;
}

// ---------------------------------------------------------------------------
// compression
// ---------------------------------------------------------------------------
// One tile of BPT blocks; `tile_in` is the tile unless a.tile_counter hands tiles out.
template <int MODE, int S, int EPL>
__device__ __forceinline__ void compress_tile(const CompressArgs& a, uint32_t tile_in) {
  using Gm = Geom<S, EPL>;
  constexpr int G = Gm::G, SPW = Gm::SPW, BPT = Gm::BPT;
  constexpr bool TOPO = (MODE == kModeTopoClassify || MODE == kModeTopoMap || MODE == kModeTopoInt);
  constexpr bool CLASSIFY = (MODE == kModeTopoClassify);
  constexpr int STAGE_WORDS = (BPT * Gm::worst_block(TOPO) + 3) / 4 + 2;

  __shared__ uint32_t stage[STAGE_WORDS];
  __shared__ uint32_t blk_off[BPT + 1];
  __shared__ uint32_t step_base[kWavesPerBlock * SPW];  // cp | ext << 16 (tile-local prefix)
  __shared__ uint64_t sh_prefix[2];
  __shared__ uint32_t sh_tile;

  const int tid = threadIdx.x;
  const int w = tid / kWave, lane = tid & (kWave - 1);
  const int g = lane / S, ls = lane & (S - 1);
  const uint32_t bs = a.bs;

  if (tid == 0) sh_tile = a.tile_counter ? atomicAdd(a.tile_counter, 1u) : tile_in;
  for (int i = tid; i < STAGE_WORDS; i += kBlockThreads) stage[i] = 0;
  __syncthreads();
  const uint32_t tile = sh_tile;

  // ---- phase A1: every element load of the tile in flight at once ---------------
  uint32_t raw[SPW][EPL];
  uint32_t tbits = 0;  // 2-bit types of the lane's SPW*EPL elements (topology modes)
  const uint32_t* in32 = reinterpret_cast<const uint32_t*>(a.in);
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t b = tile * BPT + (uint32_t)((w * SPW + s) * G + g);
    const uint32_t base = b * bs;
    const uint32_t cur = b < a.nb ? min(bs, a.n - base) : 0u;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      raw[s][r] = j < cur ? in32[base + j] : 0u;
      if constexpr (MODE == kModeTopoMap || MODE == kModeTopoInt) {
        const uint32_t t = j < cur ? (a.tmap[base + j] & 3u) : 0u;
        tbits |= t << (2 * (s * EPL + r));
      }
    }
  }

  // ---- phase A2: classification, neighbour loads of 4 steps in flight together ---
  if constexpr (CLASSIFY) {
    const float* f = reinterpret_cast<const float*>(a.in);
    const Grid& gr = a.grid;
    const int nn = gr.ndim == 3 ? 6 : 4;
    const uint32_t o1 = gr.ndim == 3 ? gr.plane : gr.d1;  // first axis stride
    const uint32_t o2 = gr.ndim == 3 ? gr.d2 : 1u;        // second axis stride
    constexpr int GS = SPW < 4 ? SPW : 4;
#pragma unroll
    for (int s0 = 0; s0 < SPW; s0 += GS) {
      float nbv[GS][EPL][6];
      bool inner[GS][EPL];
#pragma unroll
      for (int k = 0; k < GS; k++) {
        const int s = s0 + k;
        const uint32_t b = tile * BPT + (uint32_t)((w * SPW + s) * G + g);
        const uint32_t base = b * bs;
        const uint32_t cur = b < a.nb ? min(bs, a.n - base) : 0u;
#pragma unroll
        for (int r = 0; r < EPL; r++) {
          const uint32_t j = (uint32_t)(ls * EPL + r);
          const uint32_t e = base + j;
          inner[k][r] = j < cur && interior_at(gr, e);
          if (inner[k][r]) {
            nbv[k][r][0] = f[e - o1];
            nbv[k][r][1] = f[e + o1];
            nbv[k][r][2] = f[e - o2];
            nbv[k][r][3] = f[e + o2];
            if (nn == 6) {
              nbv[k][r][4] = f[e - 1];
              nbv[k][r][5] = f[e + 1];
            }
          }
        }
      }
#pragma unroll
      for (int k = 0; k < GS; k++) {
#pragma unroll
        for (int r = 0; r < EPL; r++) {
          const int t = inner[k][r] ? type_of(gr.ndim, __uint_as_float(raw[s0 + k][r]), nbv[k][r]) : 0;
          tbits |= (uint32_t)t << (2 * ((s0 + k) * EPL + r));
        }
      }
    }
  }

  // ---- phase A3: quantize, block bit counts and sizes ---------------------------
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    const uint32_t b = tile * BPT + lb;
    const uint32_t base = b * bs;
    const uint32_t cur = b < a.nb ? min(bs, a.n - base) : 0u;
    int q[EPL];
#pragma unroll
    for (int r = 0; r < EPL; r++) q[r] = (uint32_t)(ls * EPL + r) < cur ? quantize_raw<MODE>(a, raw[s][r]) : 0;
    const int nxt = __shfl_down(q[0], 1, kWave);
    unsigned m = 0;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      const int qn = r + 1 < EPL ? q[r + 1] : nxt;
      const unsigned d = (unsigned)qn - (unsigned)q[r];
      const unsigned mag = ((int)d < 0) ? 0u - d : d;
      if (j + 1 < cur && mag > m) m = mag;
    }
    m = seg_max<S>(m);
    const unsigned bc = bitcount(m);
    if (ls == 0) blk_off[lb] = cur ? block_bytes(cur, bc, TOPO) : 0u;
    if constexpr (CLASSIFY) {
      // per-step totals of the whole wave (packed cp | ext << 16); the per-lane
      // counts are recomputed from tbits in phase E
      unsigned step_cp = 0, step_ext = 0;
#pragma unroll
      for (int r = 0; r < EPL; r++) {
        const unsigned t = (tbits >> (2 * (s * EPL + r))) & 3u;
        step_cp += t != 0;
        step_ext += (t == 1 || t == 2);
      }
      unsigned tot = seg_sum<kWave>(step_cp | (step_ext << 16));
      if (lane == 0) step_base[w * SPW + s] = tot;
    }
  }
  __syncthreads();

  // ---- phase B: tile scans + look-back (wave 0) ---------------------------------
  if (w == 0) {
    // exclusive scan of block sizes, BPT / 64 consecutive entries per lane
    constexpr int PER = (BPT + kWave - 1) / kWave;
    uint32_t loc[PER];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = lane * PER + k;
      loc[k] = i < BPT ? blk_off[i] : 0u;
      sum += loc[k];
    }
    uint32_t incl = seg_incl_scan<kWave>(sum, lane);
    uint32_t run = incl - sum;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = lane * PER + k;
      if (i < BPT) blk_off[i] = run;
      run += loc[k];
    }
    const uint32_t tile_bytes = __shfl(incl, kWave - 1, kWave);
    if (lane == 0) blk_off[BPT] = tile_bytes;

    uint64_t cnt_agg = 0;
    if constexpr (CLASSIFY) {
      // exclusive prefix over the (wave, step) totals in element order
      constexpr int NS = kWavesPerBlock * SPW;  // <= 64
      uint32_t v = lane < NS ? step_base[lane] : 0u;
      uint32_t inc = seg_incl_scan<kWave>(v, lane);
      if (lane < NS) step_base[lane] = inc - v;
      uint32_t tot = __shfl(inc, kWave - 1, kWave);
      cnt_agg = (uint64_t)(tot & 0xffffu) | ((uint64_t)(tot >> 16) << 31);
    }
    uint64_t pb = lookback(a.st_bytes, tile, tile_bytes, &a.res->error);
    uint64_t pc = 0;
    if constexpr (CLASSIFY) pc = lookback(a.st_cnt, tile, cnt_agg, &a.res->error);
    if (lane == 0) {
      sh_prefix[0] = pb;
      sh_prefix[1] = pc;
      if (tile == a.ntiles - 1) {
        a.res->payload_bytes = pb + tile_bytes;
        if constexpr (CLASSIFY) {
          uint64_t tc = pc + cnt_agg;
          a.res->cp_count = tc & 0x7fffffffull;
          a.res->ext_count = (tc >> 31) & 0x7fffffffull;
        }
      }
    }
  }
  __syncthreads();

  // ---- phase C: pack every block into the LDS stage -----------------------------
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    const uint32_t b = tile * BPT + lb;
    const uint32_t base = b * bs;
    const uint32_t cur = b < a.nb ? min(bs, a.n - base) : 0u;
    int q[EPL];
#pragma unroll
    for (int r = 0; r < EPL; r++) q[r] = (uint32_t)(ls * EPL + r) < cur ? quantize_raw<MODE>(a, raw[s][r]) : 0;
    const int nxt = __shfl_down(q[0], 1, kWave);
    unsigned mags[EPL];
    bool negs[EPL];
    unsigned m = 0;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      const int qn = r + 1 < EPL ? q[r + 1] : nxt;
      const unsigned d = (unsigned)qn - (unsigned)q[r];
      negs[r] = (int)d < 0;
      mags[r] = negs[r] ? 0u - d : d;
      if (j + 1 >= cur) { mags[r] = 0; negs[r] = false; }
      if (mags[r] > m) m = mags[r];
    }
    m = seg_max<S>(m);
    if (cur == 0) continue;  // uniform per segment; no shuffles below
    const unsigned bc = bitcount(m);
    const unsigned off = blk_off[lb];
    const unsigned n = cur - 1u;
    if (ls == 0) {
      lds_or_bytes(stage, off, (uint32_t)q[0], 4);
      lds_or_bytes(stage, off + 4, bc, 1);
    }
    const unsigned sB = bc ? (n + 7u) >> 3 : 0u;
    const unsigned nfull = bc >> 3, rem = bc & 7u;
    const unsigned fbase = off + 5u + sB;
    const unsigned rbase = fbase + nfull * n;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      if (bc && j < n) {
        if (negs[r]) lds_or_bits_msb(stage, off + 5u, j, 1u, 1u);
        if (nfull) lds_or_bytes(stage, fbase + j * nfull, mags[r] >> rem, nfull);
        if (rem) lds_or_bits_msb(stage, rbase, j * rem, mags[r], rem);
      }
      if constexpr (TOPO) {
        if (j < cur) {
          const unsigned t = (tbits >> (2 * (s * EPL + r))) & 3u;
          const unsigned tbase = off + 5u + (bc ? sB + mag_bytes(n, bc) : 0u);
          if (t) lds_or_bits_msb(stage, tbase, 2u * j, t, 2u);
        }
      }
    }
  }
  __syncthreads();

  // ---- phase D: stage -> stream, block offsets side-car --------------------------
  const uint64_t g0 = sh_prefix[0];
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
    const uint32_t b = tile * BPT + lb;
    if (b < a.nb) a.blk[b] = (uint32_t)(g0 + blk_off[lb]);
  }
  if (tile == a.ntiles - 1 && tid == 0) a.blk[a.nb] = (uint32_t)(g0 + len);

  // ---- phase E: critical-point records in row-major order ------------------------
  if constexpr (CLASSIFY) {
    if (a.cp_key) {
      const uint64_t pc = sh_prefix[1];
      const uint32_t cp0 = (uint32_t)(pc & 0x7fffffffull), ext0 = (uint32_t)((pc >> 31) & 0x7fffffffull);
#pragma unroll
      for (int s = 0; s < SPW; s++) {
        unsigned c = 0, x = 0;
#pragma unroll
        for (int r = 0; r < EPL; r++) {
          const unsigned t = (tbits >> (2 * (s * EPL + r))) & 3u;
          c += t != 0;
          x += (t == 1 || t == 2);
        }
        const unsigned packed = c | (x << 16);
        const unsigned incl = seg_incl_scan<kWave>(packed, lane);
        const unsigned excl = incl - packed;
        const unsigned sb = step_base[w * SPW + s];
        uint32_t pcp = cp0 + (sb & 0xffffu) + (excl & 0xffffu);
        uint32_t pex = ext0 + (sb >> 16) + (excl >> 16);
        if (!c) continue;
#pragma unroll
        for (int r = 0; r < EPL; r++) {
          const unsigned t = (tbits >> (2 * (s * EPL + r))) & 3u;
          if (!t) continue;
          if (pcp < a.cp_cap) {
            a.cp_key[pcp] = okey(__uint_as_float(raw[s][r]));
            a.cp_pay[pcp] = (t == 3) ? 0xffffffffu : pex;
          } else {
            atomicOr(&a.res->error, TSZP_ERR_OVERFLOW);
          }
          pcp++;
          if (t != 3) pex++;
        }
      }
    }
  }
}

template <int MODE, int S, int EPL>
__global__ __launch_bounds__(kBlockThreads) void k_compress(CompressArgs a) {
  compress_tile<MODE, S, EPL>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// decompression (block offsets from the side-car)
// ---------------------------------------------------------------------------
template <int MODE, int S, int EPL>
__device__ __forceinline__ void decode_tile(const DecodeArgs& a, uint32_t tile) {
  using Gm = Geom<S, EPL>;
  constexpr int G = Gm::G, SPW = Gm::SPW, BPT = Gm::BPT;
  constexpr bool TOPO = (MODE == kModeTopoClassify || MODE == kModeTopoMap || MODE == kModeTopoInt);
  constexpr int TILE_WORDS = (BPT * Gm::worst_block(TOPO) + 3) / 4 + 2;
  __shared__ uint32_t tw[TILE_WORDS];
  __shared__ uint32_t sh_blk[BPT];  // the tile's block offsets

  const int tid = threadIdx.x;
  const int w = tid / kWave, lane = tid & (kWave - 1);
  const int g = lane / S, ls = lane & (S - 1);
  const uint32_t bs = a.bs;
  const uint32_t b0 = tile * BPT;
  const uint32_t b1 = min(a.nb, b0 + (uint32_t)BPT);

  for (uint32_t i = tid; i < (uint32_t)BPT; i += kBlockThreads) sh_blk[i] = b0 + i < b1 ? a.blk[b0 + i] : 0u;
  const uint32_t t0 = a.blk[b0], t1 = a.blk[b1];
  const uint32_t wbase = t0 & ~3u;
  const uint32_t nwords = ((t1 + 3u) & ~3u) / 4u - wbase / 4u;
  if (t1 < t0 || nwords > (uint32_t)TILE_WORDS) {
    if (tid == 0) atomicOr(a.err, TSZP_ERR_STREAM);
    return;
  }
  const uint32_t* src = reinterpret_cast<const uint32_t*>(a.payload + wbase);
  // eight loads in flight per thread before the LDS stores
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
  __syncthreads();
  // 16-byte stores of four elements need block starts at multiples of 4 and aligned outputs
  const bool vec = (bs & 3u) == 0 && ((uintptr_t)a.out & 15u) == 0 && ((uintptr_t)a.types8 & 3u) == 0 &&
                   ((uintptr_t)a.types32 & 15u) == 0;

  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    const uint32_t b = b0 + lb;
    const uint32_t base = b * bs;
    const uint32_t cur = b < a.nb ? min(bs, a.n - base) : 0u;
    const uint32_t o = cur ? sh_blk[lb] - wbase : 0u;
    int anchor = 0;
    unsigned bc = 0;
    if (cur) {
      anchor = (int)(lds_byte(tw, o) | (lds_byte(tw, o + 1) << 8) | (lds_byte(tw, o + 2) << 16) |
                     (lds_byte(tw, o + 3) << 24));
      // the topology decoder always reads bc (SZp/src/szpd_float.cc:735); the
      // others only for blocks of more than one element (:487, :958)
      if (TOPO || cur > 1) bc = lds_byte(tw, o + 4);
      if (bc > 32) {
        if (ls == 0) atomicOr(a.err, TSZP_ERR_STREAM);
        bc = 0;
      }
    }
    const unsigned n = cur ? cur - 1u : 0u;
    const unsigned sB = bc ? (n + 7u) >> 3 : 0u;
    const unsigned nfull = bc >> 3, rem = bc & 7u;
    const unsigned fbase = o + 5u + sB;
    const unsigned rbase = fbase + nfull * n;
    int dl[EPL];
    int lane_sum = 0;
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      int d = 0;
      if (bc && j >= 1 && j < cur) {
        const unsigned k = j - 1u;  // delta k leads into element j
        const unsigned neg = (lds_byte(tw, o + 5u + (k >> 3)) >> (7u - (k & 7u))) & 1u;
        unsigned mag = 0;
        for (unsigned kk = 0; kk < nfull; kk++) mag |= lds_byte(tw, fbase + k * nfull + kk) << (8u * kk);
        mag <<= rem;
        if (rem) {
          const unsigned pos = k * rem, by = rbase + (pos >> 3), sh = pos & 7u;
          unsigned win = lds_byte(tw, by) << 8;
          if (sh + rem > 8u) win |= lds_byte(tw, by + 1);
          mag |= (win >> (16u - rem - sh)) & ((1u << rem) - 1u);
        }
        d = (int)(neg ? 0u - mag : mag);
      }
      lane_sum = (int)((unsigned)lane_sum + (unsigned)d);
      dl[r] = lane_sum;
    }
    const int incl = seg_incl_scan<S>(lane_sum, lane);
    const int excl = (int)((unsigned)incl - (unsigned)lane_sum);
    if (cur == 0) continue;
    unsigned tbase = 0;
    if constexpr (TOPO) tbase = o + 5u + (bc ? sB + mag_bytes(n, bc) : 0u);
    if constexpr (EPL == 4) {  // a lane's four whole elements: one 16-byte store (and one of types)
      const uint32_t j0 = (uint32_t)(ls * 4);
      if (vec && j0 + 3 < cur) {
        int qv[4];
#pragma unroll
        for (int r = 0; r < 4; r++) qv[r] = (int)((unsigned)anchor + (unsigned)excl + (unsigned)dl[r]);
        const uint32_t e = base + j0;
        if constexpr (MODE == kModeInt) {
          reinterpret_cast<int4*>(a.out)[e >> 2] = make_int4(qv[0], qv[1], qv[2], qv[3]);
        } else {
          reinterpret_cast<float4*>(a.out)[e >> 2] =
              make_float4(__fmul_rn(__int2float_rn(qv[0]), a.eb), __fmul_rn(__int2float_rn(qv[1]), a.eb),
                          __fmul_rn(__int2float_rn(qv[2]), a.eb), __fmul_rn(__int2float_rn(qv[3]), a.eb));
        }
        if constexpr (TOPO) {
          const unsigned tb = lds_byte(tw, tbase + (j0 >> 2));  // element j0 in the top two bits
          const unsigned t0v = (tb >> 6) & 3u, t1v = (tb >> 4) & 3u, t2v = (tb >> 2) & 3u, t3v = tb & 3u;
          if (a.types8) reinterpret_cast<uint32_t*>(a.types8)[e >> 2] = t0v | t1v << 8 | t2v << 16 | t3v << 24;
          if (a.types32)
            reinterpret_cast<int4*>(a.types32)[e >> 2] = make_int4((int)t0v, (int)t1v, (int)t2v, (int)t3v);
        }
        continue;
      }
    }
#pragma unroll
    for (int r = 0; r < EPL; r++) {
      const uint32_t j = (uint32_t)(ls * EPL + r);
      if (j >= cur) continue;
      const int qv = (int)((unsigned)anchor + (unsigned)excl + (unsigned)dl[r]);
      const uint32_t e = base + j;
      if constexpr (MODE == kModeInt) {
        reinterpret_cast<int*>(a.out)[e] = qv;
      } else {
        reinterpret_cast<float*>(a.out)[e] = __fmul_rn(__int2float_rn(qv), a.eb);
      }
      if constexpr (TOPO) {
        const unsigned t = (lds_byte(tw, tbase + ((2u * j) >> 3)) >> (6u - ((2u * j) & 7u))) & 3u;
        if (a.types8) a.types8[e] = (uint8_t)t;
        if (a.types32) a.types32[e] = (int)t;
      }
    }
  }
}

template <int MODE, int S, int EPL>
__global__ __launch_bounds__(kBlockThreads) void k_decode(DecodeArgs a) {
  decode_tile<MODE, S, EPL>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// compression, fast path for block sizes 32 and 64: S = bs / 4 lanes per block and
// four consecutive elements per lane, loaded as one 16-byte vector; with the fastest
// axis a multiple of 4 long, the classification stencil rows are 16-byte loads too and
// the neighbours along the fastest axis come from the adjacent lanes.  Each lane packs
// its four elements as whole runs: one nibble of sign bits, one run of full magnitude
// bytes, one run of remainder bits and one byte of types, so a lane issues a handful
// of LDS ORs per four elements.  Same stream bytes as k_compress.
// ---------------------------------------------------------------------------
// OR a little-endian run of up to 16 bytes (w0..w3) at byte address a of the stage.
__device__ __forceinline__ void lds_or_run16(uint32_t* words, unsigned a, uint32_t w0, uint32_t w1, uint32_t w2,
                                             uint32_t w3) {
  const unsigned sh = a & 3u, k = a >> 2;
  uint32_t o0 = w0, o1 = w1, o2 = w2, o3 = w3, o4 = 0;
  if (sh) {
    o0 = w0 << (8 * sh);
    o1 = __builtin_amdgcn_alignbyte(w1, w0, 4 - sh);
    o2 = __builtin_amdgcn_alignbyte(w2, w1, 4 - sh);
    o3 = __builtin_amdgcn_alignbyte(w3, w2, 4 - sh);
    o4 = w3 >> (32 - 8 * sh);
  }
  if (o0) atomicOr(&words[k], o0);
  if (o1) atomicOr(&words[k + 1], o1);
  if (o2) atomicOr(&words[k + 2], o2);
  if (o3) atomicOr(&words[k + 3], o3);
  if (o4) atomicOr(&words[k + 4], o4);
}

// critical-point type from the centre and the two neighbours along each axis (plain
// compares: fmaxf / fminf would add NaN canonicalisations)
template <int ND>
__device__ __forceinline__ int type_minmax(float c, const float (&lo)[ND], const float (&hi)[ND]) {
  bool all_gt = true, all_lt = true, any_hi = false, any_lo = false;
#pragma unroll
  for (int k = 0; k < ND; k++) {
    const bool h = c > lo[k] && c > hi[k];
    const bool l = c < lo[k] && c < hi[k];
    all_gt = all_gt && h;
    all_lt = all_lt && l;
    any_hi = any_hi || h;
    any_lo = any_lo || l;
  }
  return all_gt ? 1 : all_lt ? 2 : (any_hi && any_lo) ? 3 : 0;
}

constexpr int kV4TileElems = 4096;  // points per tile of the fast compressor (4 waves x 4 steps x 256)
constexpr int kV4HaloMax = 1024;    // longest row its classification stages in LDS

template <int MODE, int S, int ND>
__device__ __forceinline__ void compress_v4_tile(const CompressArgs& a, uint32_t tile_in) {
  constexpr int EPL = 4, BS = 4 * S;
  using Gm = Geom<S, EPL>;
  constexpr int G = Gm::G, SPW = Gm::SPW, BPT = Gm::BPT;
  constexpr bool TOPO = (MODE == kModeTopoClassify || MODE == kModeTopoMap || MODE == kModeTopoInt);
  constexpr bool CLASSIFY = (MODE == kModeTopoClassify);
  constexpr int STAGE_WORDS = (BPT * Gm::worst_block(TOPO) + 3) / 4 + 8;  // runs may spill 16 zero bytes
  // classification stages the tile's points and the rows next to it in the same LDS
  // before the packing uses it (tile of kV4TileElems points, rows of up to kV4HaloMax)
  static_assert(BPT * BS == kV4TileElems, "a wave step covers 256 points");
  // ... and, after the copy-out, the tile's critical-point records (a key and a payload
  // word per point at most)
  constexpr int CLS_WORDS = kV4TileElems + 2 * kV4HaloMax > 2 * kV4TileElems ? kV4TileElems + 2 * kV4HaloMax
                                                                             : 2 * kV4TileElems;
  constexpr int LDS_WORDS = CLASSIFY && CLS_WORDS > STAGE_WORDS ? CLS_WORDS : STAGE_WORDS;

  __shared__ __align__(16) uint32_t stage[LDS_WORDS];
  __shared__ uint32_t blk_off[BPT + 1];
  __shared__ uint32_t step_base[kWavesPerBlock * SPW];  // cp | ext << 16 (tile-local prefix)
  __shared__ uint64_t sh_prefix[2];
  __shared__ uint64_t sh_cnt_agg;
  __shared__ uint32_t sh_tile;

  const int tid = threadIdx.x;
  const int w = tid / kWave, lane = tid & (kWave - 1);
  const int g = lane / S, ls = lane & (S - 1);
  const uint32_t j0 = 4u * (uint32_t)ls;  // first element of the lane within its block

  if (tid == 0) sh_tile = a.tile_counter ? atomicAdd(a.tile_counter, 1u) : tile_in;
  if constexpr (!CLASSIFY)  // (classification clears it after its staging)
    for (int i = tid; i < STAGE_WORDS; i += kBlockThreads) stage[i] = 0;
  __syncthreads();
  const uint32_t tile = sh_tile;

  auto block_of = [&](int s) { return tile * BPT + (uint32_t)((w * SPW + s) * G + g); };
  auto cur_of = [&](uint32_t b) { return b < a.nb ? min((uint32_t)BS, a.n - b * BS) : 0u; };

  // ---- A1: the lane's 4 elements of every step as 16-byte loads ---------------------
  uint32_t raw[SPW][4];
  uint32_t tbits = 0;
  const uint32_t* in32 = reinterpret_cast<const uint32_t*>(a.in);
  // the full 16-byte loads of every step first (lanes past the end read vector 0), then
  // the element loads of a partial last block: no wait between the steps' loads
  uint4 v4[SPW];
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t b = block_of(s), cur = cur_of(b);
    v4[s] = reinterpret_cast<const uint4*>(a.in)[j0 + 3 < cur ? (b * BS + j0) >> 2 : 0u];
  }
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t b = block_of(s), cur = cur_of(b);
    const uint32_t e0 = b * BS + j0;
    raw[s][0] = v4[s].x;
    raw[s][1] = v4[s].y;
    raw[s][2] = v4[s].z;
    raw[s][3] = v4[s].w;
    if (!(j0 + 3 < cur)) {
      uint32_t p[4];
#pragma unroll
      for (int r = 0; r < 4; r++) p[r] = j0 + r < cur ? in32[e0 + r] : 0u;
#pragma unroll
      for (int r = 0; r < 4; r++) raw[s][r] = p[r];
    }
    if constexpr (MODE == kModeTopoMap || MODE == kModeTopoInt) {
      uint32_t tw = 0;
      if (j0 + 3 < cur) {
        tw = reinterpret_cast<const uint32_t*>(a.tmap)[e0 >> 2];
      } else {
#pragma unroll
        for (int r = 0; r < 4; r++) tw |= (j0 + r < cur ? (uint32_t)a.tmap[e0 + r] : 0u) << (8 * r);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) tbits |= ((tw >> (8 * r)) & 3u) << (2 * (s * 4 + r));
    }
  }

  // ---- A2: classification ------------------------------------------------------------
  // Stencil along the fastest axis: the adjacent lanes (DPP).  Along the next axis (rows
  // of rn points: d2 in 3D, d1 in 2D): 16-byte global loads, which the tile's own loads
  // have mostly brought into L1 / L2; or, with a.lds_rows and rn <= kV4HaloMax, the
  // tile's own points and one row before and after it staged in LDS (the same bytes;
  // measured 2 % slower at 512^3, so not the default).  Along the slowest axis of a 3D
  // grid (planes): 16-byte global loads (a plane apart; the tile 64 tiles back on the
  // same XCD holds them in its L2).
  if constexpr (CLASSIFY) {
    const Grid& gr = a.grid;
    const uint32_t dz = ND == 3 ? gr.d2 : gr.d1;          // fastest axis length (multiple of 4)
    const uint32_t s1 = ND == 3 ? gr.plane : gr.d1;      // stride of the slowest axis
    const uint32_t rn = ND == 3 ? gr.d2 : gr.d1;         // stride of the next axis (a row)
    constexpr int KN = ND - 2;                           // its stencil index
    const uint4* f4 = reinterpret_cast<const uint4*>(a.in);
    const bool near_lds = a.lds_rows && rn <= (uint32_t)kV4HaloMax;  // uniform
    uint4* tl4 = reinterpret_cast<uint4*>(stage);        // point T0 - rn + i at word i
    const uint32_t T0 = tile * (uint32_t)kV4TileElems;
    if (near_lds) {
#pragma unroll
      for (int s = 0; s < SPW; s++) {
        const uint32_t l = (uint32_t)(w * SPW + s) * 256u + 4u * (uint32_t)lane;  // tile-local point
        tl4[(rn + l) >> 2] = make_uint4(raw[s][0], raw[s][1], raw[s][2], raw[s][3]);
      }
      // the row before the tile and the row after it, where they lie in the field
      // (points outside it are never read: a point with a neighbour there is not classified)
      const uint32_t q = rn >> 2;
      for (uint32_t i = tid; i < 2 * q; i += kBlockThreads) {
        const bool hi = i >= q;
        const uint32_t k = hi ? i - q : i;
        const uint64_t e = hi ? (uint64_t)T0 + kV4TileElems + 4u * k : (uint64_t)T0 + 4u * k - rn;
        const bool ok = hi ? e + 3 < a.n : T0 >= rn;
        if (ok) tl4[hi ? (rn + kV4TileElems) / 4 + k : k] = f4[e >> 2];
      }
      __syncthreads();
    }
    // every global stencil load of every step is issued before the first compare: one
    // memory round trip for the tile's classification instead of one per step
    uint4 nlo[SPW][ND - 1], nhi[SPW][ND - 1];
    uint32_t zm[SPW], zp[SPW], zs[SPW];
    bool rowok[SPW];
#pragma unroll
    for (int s = 0; s < SPW; s++) {
      const uint32_t b = block_of(s), cur = cur_of(b);
      const uint32_t e0 = b * BS + j0;
      const bool act = j0 < cur;  // n is a multiple of 4: all four elements or none
      uint32_t z;
      bool ok;
      if constexpr (ND == 3) {
        const uint32_t x = fdiv(e0, gr.div_plane), rr = e0 - x * gr.plane;
        const uint32_t y = fdiv(rr, gr.div_row);
        z = rr - y * gr.d2;
        ok = act && x >= 1 && x + 1 < gr.d0 && y >= 1 && y + 1 < gr.d1;
      } else {
        const uint32_t i = fdiv(e0, gr.div_row);
        z = e0 - i * gr.d1;
        ok = act && i >= 1 && i + 1 < gr.d0;
      }
      zs[s] = z;
      rowok[s] = ok;
      // unconditional 16-byte loads; lanes without a stencil read vector 0 (unused)
      if constexpr (ND == 3) {
        nlo[s][0] = f4[ok ? (e0 - s1) >> 2 : 0u];
        nhi[s][0] = f4[ok ? (e0 + s1) >> 2 : 0u];
      }
      if (!near_lds) {
        nlo[s][KN] = f4[ok ? (e0 - rn) >> 2 : 0u];
        nhi[s][KN] = f4[ok ? (e0 + rn) >> 2 : 0u];
      }
      // neighbours along the fastest axis: the adjacent lanes hold elements e0 - 1 and
      // e0 + 4 (blocks of a step are consecutive); the wave's end lanes load them
      zm[s] = 0u;
      zp[s] = 0u;
      if (lane == 0 && ok && z >= 1) zm[s] = in32[e0 - 1];
      if (lane == kWave - 1 && ok && z + 4 < dz) zp[s] = in32[e0 + 4];
    }
#pragma unroll
    for (int s = 0; s < SPW; s++) {
      const uint32_t z = zs[s];
      const uint32_t zmv = dpp_mov<kDppWaveShr1>(raw[s][3]), zpv = dpp_mov<kDppWaveShl1>(raw[s][0]);
      const uint32_t zm_s = lane == 0 ? zm[s] : zmv, zp_s = lane == kWave - 1 ? zp[s] : zpv;
      if (near_lds) {  // rows from LDS: the point l - rn is at word l, l + rn at word l + 2 rn
        const uint32_t l = (uint32_t)(w * SPW + s) * 256u + 4u * (uint32_t)lane;
        nlo[s][KN] = tl4[l >> 2];
        nhi[s][KN] = tl4[(l + 2 * rn) >> 2];
      }
      if (rowok[s]) {
        const float c[4] = {__uint_as_float(raw[s][0]), __uint_as_float(raw[s][1]), __uint_as_float(raw[s][2]),
                            __uint_as_float(raw[s][3])};
        const float zl[4] = {__uint_as_float(zm_s), c[0], c[1], c[2]};
        const float zh[4] = {c[1], c[2], c[3], __uint_as_float(zp_s)};
#pragma unroll
        for (int r = 0; r < 4; r++) {
          // every element is classified; the ends of a row are masked afterwards
          float lo[ND], hi[ND];
#pragma unroll
          for (int k = 0; k < ND - 1; k++) {
            const uint32_t l4[4] = {nlo[s][k].x, nlo[s][k].y, nlo[s][k].z, nlo[s][k].w};
            const uint32_t h4[4] = {nhi[s][k].x, nhi[s][k].y, nhi[s][k].z, nhi[s][k].w};
            lo[k] = __uint_as_float(l4[r]);
            hi[k] = __uint_as_float(h4[r]);
          }
          lo[ND - 1] = zl[r];
          hi[ND - 1] = zh[r];
          const bool zin = z + r >= 1 && z + r + 1 < dz;
          const uint32_t ty = zin ? (uint32_t)type_minmax<ND>(c[r], lo, hi) : 0u;
          tbits |= ty << (2 * (s * 4 + r));
        }
      }
    }
    // the staging area becomes the packing stage: cleared once every wave has read it
    __syncthreads();
    for (int i = tid; i < STAGE_WORDS; i += kBlockThreads) stage[i] = 0;
  }

  // ---- A3: quantize once; block bit counts, sizes and delta magnitudes ---------------
  uint32_t mg[SPW][4];
  uint32_t nmask = 0;  // sign bits, bit s*4+r
  uint32_t bcs[SPW];
  int q0s[SPW];
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    const uint32_t cur = cur_of(tile * BPT + lb);
    int q[4];
#pragma unroll
    for (int r = 0; r < 4; r++) q[r] = j0 + r < cur ? quantize_raw<MODE>(a, raw[s][r]) : 0;
    const int nxt = (int)dpp_mov<kDppWaveShl1>((uint32_t)q[0]);
    unsigned m = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int qn = r < 3 ? q[r + 1] : nxt;
      const unsigned d = (unsigned)qn - (unsigned)q[r];
      const bool neg = (int)d < 0;
      const bool ok = j0 + r + 1 < cur;
      const unsigned mag = ok ? (neg ? 0u - d : d) : 0u;
      mg[s][r] = mag;
      nmask |= (ok && neg ? 1u : 0u) << (s * 4 + r);
      m = mag > m ? mag : m;
    }
    if constexpr (S > 16) {  // several rows: DPP within each, then across them
      m = seg_max_dpp<16>(m);
#pragma unroll
      for (int o = 16; o < S; o <<= 1) {
        const unsigned v = (unsigned)__shfl_xor((int)m, o, kWave);
        m = v > m ? v : m;
      }
    } else {
      m = seg_max_dpp<S>(m);
    }
    const unsigned bc = bitcount(m);
    bcs[s] = bc;
    q0s[s] = q[0];
    if (ls == 0) blk_off[lb] = cur ? block_bytes(cur, bc, TOPO) : 0u;
    if constexpr (CLASSIFY) {
      unsigned step_cp = 0, step_ext = 0;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const unsigned t = (tbits >> (2 * (s * 4 + r))) & 3u;
        step_cp += t != 0;
        step_ext += (t == 1 || t == 2);
      }
      const unsigned tot = wave_sum_dpp(step_cp | (step_ext << 16));
      if (lane == 0) step_base[w * SPW + s] = tot;
    }
  }
  __syncthreads();

  // ---- B: tile scans (wave 0); the look-back follows, overlapped with the packing -----
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
    const uint32_t incl = wave_incl_scan_dpp(sum, lane);
    uint32_t run = incl - sum;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = lane * PER + k;
      if (i < BPT) blk_off[i] = run;
      run += loc[k];
    }
    const uint32_t tile_bytes = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
    if (lane == 0) blk_off[BPT] = tile_bytes;
    if constexpr (CLASSIFY) {
      constexpr int NS = kWavesPerBlock * SPW;
      const uint32_t v = lane < NS ? step_base[lane] : 0u;
      const uint32_t inc = wave_incl_scan_dpp(v, lane);
      if (lane < NS) step_base[lane] = inc - v;
      const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
      if (lane == 0) sh_cnt_agg = (uint64_t)(tot & 0xffffu) | ((uint64_t)(tot >> 16) << 31);
    }
  }
  __syncthreads();
  // wave 0 finds the tile's stream offset and wave 1 its CP offsets (the two look-backs
  // side by side) while the other waves already pack (the packing uses tile-local
  // offsets only)
  if (w == 0) {
    const uint32_t tile_bytes = blk_off[BPT];
    const uint64_t pb = lookback(a.st_bytes, tile, tile_bytes, &a.res->error);
    if (lane == 0) {
      sh_prefix[0] = pb;
      if (tile == a.ntiles - 1) a.res->payload_bytes = pb + tile_bytes;
    }
  } else if (w == 1) {
    if constexpr (CLASSIFY) {
      const uint64_t cnt_agg = sh_cnt_agg;
      const uint64_t pc = lookback(a.st_cnt, tile, cnt_agg, &a.res->error);
      if (lane == 0) {
        sh_prefix[1] = pc;
        if (tile == a.ntiles - 1) {
          const uint64_t tc = pc + cnt_agg;
          a.res->cp_count = tc & 0x7fffffffull;
          a.res->ext_count = (tc >> 31) & 0x7fffffffull;
        }
      }
    } else if (lane == 0) {
      sh_prefix[1] = 0;
    }
  }

  // ---- C: pack the lane's runs into the LDS stage ---------------------------------------
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    const uint32_t lb = (uint32_t)((w * SPW + s) * G + g);
    const uint32_t cur = cur_of(tile * BPT + lb);
    const unsigned bc = bcs[s];
    // sign nibble of elements j0..j0+3 (MSB first); the even lane of a pair writes the byte
    const unsigned nib = ((nmask >> (s * 4)) & 1u) << 3 | ((nmask >> (s * 4 + 1)) & 1u) << 2 |
                         ((nmask >> (s * 4 + 2)) & 1u) << 1 | ((nmask >> (s * 4 + 3)) & 1u);
    const unsigned nib_next = dpp_mov<kDppWaveShl1>(nib);
    if (cur == 0) continue;  // uniform per block
    const unsigned off = blk_off[lb];
    const unsigned n = cur - 1u;
    if (ls == 0) {
      lds_or_bytes(stage, off, (uint32_t)q0s[s], 4);
      lds_or_bytes(stage, off + 4, bc, 1);
    }
    const unsigned sB = bc ? (n + 7u) >> 3 : 0u;
    if (bc) {
      if ((ls & 1) == 0 && (unsigned)(ls >> 1) < sB) {
        const unsigned byte = (nib << 4) | (ls + 1 < S ? nib_next : 0u);
        if (byte) lds_or_bytes(stage, off + 5u + (ls >> 1), byte, 1);
      }
      const unsigned nfull = bc >> 3, rem = bc & 7u;
      const unsigned fbase = off + 5u + sB;
      if (nfull) {
        const uint32_t v0 = mg[s][0] >> rem, v1 = mg[s][1] >> rem, v2 = mg[s][2] >> rem, v3 = mg[s][3] >> rem;
        uint32_t w0, w1 = 0, w2 = 0, w3 = 0;
        if (nfull == 1) {
          w0 = v0 | v1 << 8 | v2 << 16 | v3 << 24;
        } else if (nfull == 2) {
          w0 = v0 | v1 << 16;
          w1 = v2 | v3 << 16;
        } else if (nfull == 3) {
          w0 = v0 | v1 << 24;
          w1 = v1 >> 8 | v2 << 16;
          w2 = v2 >> 16 | v3 << 8;
        } else {
          w0 = v0;
          w1 = v1;
          w2 = v2;
          w3 = v3;
        }
        if (j0 < n) lds_or_run16(stage, fbase + j0 * nfull, w0, w1, w2, w3);
      }
      if (rem) {
        const uint32_t rm = (1u << rem) - 1u;
        const uint32_t R = ((((mg[s][0] & rm) << rem | (mg[s][1] & rm)) << rem | (mg[s][2] & rm)) << rem) |
                           (mg[s][3] & rm);
        const unsigned P = j0 * rem;                   // bit offset of the run
        const unsigned B = fbase + nfull * n + (P >> 3), sh = P & 7u;
        const uint32_t be = R << (32u - sh - 4u * rem);  // big-endian window at byte B
        if (j0 < n && be) {
          const uint32_t le = __builtin_bswap32(be);
          lds_or_run16(stage, B, le, 0u, 0u, 0u);
        }
      }
    }
    if constexpr (TOPO) {
      const unsigned tb = (tbits >> (8 * s)) & 0xffu;  // types of elements j0..j0+3, 2 bits each
      // MSB-first: element j0 in the top two bits of the lane's byte
      const unsigned byte = (tb & 3u) << 6 | ((tb >> 2) & 3u) << 4 | ((tb >> 4) & 3u) << 2 | ((tb >> 6) & 3u);
      const unsigned tbase = off + 5u + (bc ? sB + mag_bytes(n, bc) : 0u);
      if (byte && j0 < cur) lds_or_bytes(stage, tbase + ls, byte, 1);
    }
  }
  __syncthreads();

  // ---- D: stage -> stream, block offsets side-car ----------------------------------------
  const uint64_t g0 = sh_prefix[0];
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
        // output word i = stage bytes head + 4i .. head + 4i + 3
        const uint32_t lo = stage[i], hi = stage[i + 1];
        dw[i] = head ? __builtin_amdgcn_alignbyte(hi, lo, head) : lo;
      }
    }
  }
  for (int lb = tid; lb < BPT; lb += kBlockThreads) {
    const uint32_t b = tile * BPT + lb;
    if (b < a.nb) a.blk[b] = (uint32_t)(g0 + blk_off[lb]);
  }
  if (tile == a.ntiles - 1 && tid == 0) a.blk[a.nb] = (uint32_t)(g0 + len);

  // ---- E: critical-point records in row-major order ---------------------------------------
  // The tile's records are contiguous in the record arrays: gathered in LDS at their
  // tile-local index, then written with consecutive stores (a lane's own records would be
  // up to four scattered stores per array and step).
  if constexpr (CLASSIFY) {
    if (a.cp_key) {
      const uint64_t pc = sh_prefix[1];
      const uint32_t cp0 = (uint32_t)(pc & 0x7fffffffull), ext0 = (uint32_t)((pc >> 31) & 0x7fffffffull);
      uint32_t* lk = stage;                  // tile-local keys
      uint32_t* lp = stage + kV4TileElems;   // and payloads
      __syncthreads();                       // every wave has copied the stream out of the stage
#pragma unroll
      for (int s = 0; s < SPW; s++) {
        unsigned c = 0, x = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const unsigned t = (tbits >> (2 * (s * 4 + r))) & 3u;
          c += t != 0;
          x += (t == 1 || t == 2);
        }
        const unsigned packed = c | (x << 16);
        const unsigned incl = wave_incl_scan_dpp(packed, lane);
        const unsigned excl = incl - packed;
        const unsigned sb = step_base[w * SPW + s];
        uint32_t lc = (sb & 0xffffu) + (excl & 0xffffu);
        uint32_t pex = ext0 + (sb >> 16) + (excl >> 16);
        if (!c) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const unsigned t = (tbits >> (2 * (s * 4 + r))) & 3u;
          if (!t) continue;
          lk[lc] = okey(__uint_as_float(raw[s][r]));
          lp[lc] = (t == 3) ? 0xffffffffu : pex;
          lc++;
          if (t != 3) pex++;
        }
      }
      __syncthreads();
      const uint32_t ntile = (uint32_t)(sh_cnt_agg & 0x7fffffffull);
      if (tid == 0 && (uint64_t)cp0 + ntile > a.cp_cap) atomicOr(&a.res->error, TSZP_ERR_OVERFLOW);
      for (uint32_t i = tid; i < ntile; i += kBlockThreads) {
        const uint32_t pcp = cp0 + i;
        if (pcp < a.cp_cap) {
          a.cp_key[pcp] = lk[i];
          a.cp_pay[pcp] = lp[i];
        }
      }
    }
  }
}

template <int MODE, int S, int ND>
__global__ __launch_bounds__(kBlockThreads, 4) void k_compress_v4(CompressArgs a) {
  compress_v4_tile<MODE, S, ND>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// batches of plain float fields: every field of one block-size class in one launch.
// Workgroup b takes tile base + b of the batch's tile numbering; tmap names the tile's
// field, whose tiles are contiguous, and the field's look-back status words start at
// its first tile, so a tile only waits on earlier tiles of its own field.  Streams and
// side-cars are those of the one-field codec with one chunk.  With a tile counter
// (nullable) workgroups take their batch tile from it instead of their id, as the
// one-field kernels do under TSZP_TILE_COUNTER.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlockThreads) void k_batch_tilemap(const BatchField* __restrict__ fields,
                                                                uint32_t* __restrict__ tmap) {
  const uint32_t fi = blockIdx.x;  // one workgroup per field
  const uint32_t t0 = fields[fi].tile0, nt = fields[fi].ntiles;
  for (uint32_t t = threadIdx.x; t < nt; t += kBlockThreads) tmap[t0 + t] = fi;
}

// Element count, blocks and tiles of a field in a batch launch: n_dev (when set) holds
// the count on the device, the host numbered tiles for the largest.
template <int S, int EPL>
__device__ __forceinline__ void batch_counts(const BatchField& f, uint32_t* n, uint32_t* nb, uint32_t* ntiles) {
  if (f.n_dev) {
    const uint64_t m = *f.n_dev;
    *n = (uint32_t)(m < f.n ? m : f.n);
    *nb = (*n + f.bs - 1) / f.bs;
    *ntiles = (*nb + Geom<S, EPL>::BPT - 1) / Geom<S, EPL>::BPT;
  } else {
    *n = f.n;
    *nb = f.nb;
    *ntiles = f.ntiles;
  }
}

template <int MODE, int S, int EPL, bool V4, int ND>
__global__ __launch_bounds__(kBlockThreads, V4 ? 4 : 1) void k_compress_batch(const BatchField* __restrict__ fields,
                                                                            const uint32_t* __restrict__ tmap,
                                                                            uint32_t base,
                                                                            uint64_t* __restrict__ status,
                                                                            uint64_t* __restrict__ status_cnt,
                                                                            CompressResult* __restrict__ res,
                                                                            uint32_t* counter) {
  __shared__ uint32_t sh_gt;
  if (threadIdx.x == 0) sh_gt = base + tile_id(counter);
  __syncthreads();
  const uint32_t gt = sh_gt;
  const uint32_t fi = tmap[gt];
  const BatchField& f = fields[fi];
  CompressArgs a{};
  // (the 16-byte path runs S lanes of 4 elements per block: Geom<S, 4> is its tile)
  batch_counts<S, EPL>(f, &a.n, &a.nb, &a.ntiles);
  const uint32_t tile = gt - f.tile0;
  if (tile >= a.ntiles) return;  // past the blocks of a count read on the device
  a.in = f.in;
  a.bs = f.bs;
  a.eb = f.eb;
  a.inv = f.inv;
  a.grid = f.grid;
  a.payload = reinterpret_cast<uint8_t*>(f.out) + f.hdr;
  a.cap = f.cap;
  a.blk = f.blk;
  a.st_bytes = status + f.tile0;
  a.st_cnt = status_cnt ? status_cnt + f.tile0 : nullptr;
  a.cp_key = f.cp_key;
  a.cp_pay = f.cp_pay;
  a.cp_cap = f.cp_cap;
  a.res = res + fi;
  if (tile == 0 && threadIdx.x == 0) {
    uint32_t* h = reinterpret_cast<uint32_t*>(f.out);
    if constexpr (MODE == kModePlain) {  // eb big-endian, then chunk offset 0
      h[0] = __builtin_bswap32(__float_as_uint(f.eb));
      h[1] = 0u;
      h[2] = 0u;
    } else if constexpr (MODE == kModeInt) {  // extremum count, then chunk offset 0
      h[0] = a.n;
      h[1] = 0u;
      h[2] = 0u;
      h[3] = 0u;
    } else {  // chunk offset 0
      h[0] = 0u;
      h[1] = 0u;
    }
  }
  if constexpr (V4) {
    compress_v4_tile<MODE, S, ND>(a, tile);
  } else {
    compress_tile<MODE, S, EPL>(a, tile);
  }
}

template <int MODE, int S, int EPL>
__global__ __launch_bounds__(kBlockThreads) void k_decode_batch(const BatchField* __restrict__ fields,
                                                               const uint32_t* __restrict__ tmap, uint32_t base,
                                                               CompressResult* __restrict__ res) {
  const uint32_t gt = base + blockIdx.x;
  const uint32_t fi = tmap[gt];
  const BatchField& f = fields[fi];
  DecodeArgs a{};
  uint32_t ntiles;
  batch_counts<S, EPL>(f, &a.n, &a.nb, &ntiles);
  const uint32_t tile = gt - f.tile0;
  if (tile >= ntiles) return;
  a.payload = reinterpret_cast<const uint8_t*>(f.in) + f.hdr;
  a.blk = f.blk;
  a.bs = f.bs;
  a.eb = f.eb;
  a.out = f.out;
  a.types8 = f.types8;
  a.err = &res[fi].error;
  decode_tile<MODE, S, EPL>(a, tile);
}

// Payload bytes of sampled windows at each candidate block size: one workgroup per
// field, each wave one window of 128 elements at a time (two per lane), windows spread
// evenly over the field and aligned to 128 elements, so that every candidate's blocks
// inside a window are blocks of the real stream.
template <int L>
__device__ __forceinline__ unsigned tune_block_bytes(unsigned m0, unsigned m1, int lane) {
  // lane holds elements 2 lane, 2 lane + 1; m1 is the delta into 2 lane + 2, which
  // belongs to the block only when 2 lane + 2 does not start the next one
  const unsigned m = ((2 * lane + 2) % (2 * L)) ? (m1 > m0 ? m1 : m0) : m0;
  const unsigned bm = seg_max<L>(m);
  return (lane % L) == 0 ? block_bytes(2 * L, bitcount(bm), false) : 0u;
}

__global__ __launch_bounds__(kBlockThreads) void k_batch_tune(const BatchField* __restrict__ fields,
                                                             uint32_t* __restrict__ est) {
  static_assert(kTuneWindow == 2 * kWave, "two elements per lane");
  static_assert(kTuneSizes[0] == 16 && kTuneSizes[1] == 32 && kTuneSizes[2] == 64 && kTuneSizes[3] == 128,
                "candidate block sizes");
  __shared__ uint32_t red[kWavesPerBlock][kTuneCands];
  const BatchField& f = fields[blockIdx.x];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const uint32_t nwin = f.n / kTuneWindow;
  const uint32_t W = nwin < kTuneMaxWindows ? nwin : kTuneMaxWindows;
  const float* x = reinterpret_cast<const float*>(f.in);
  unsigned acc[kTuneCands] = {0u, 0u, 0u, 0u};
  for (uint32_t k = w; k < W; k += kWavesPerBlock) {
    const uint64_t e0 = ((uint64_t)k * nwin / W) * kTuneWindow + 2u * (uint32_t)lane;
    const int q0 = cvt_i32((double)x[e0] * f.inv);
    const int q1 = cvt_i32((double)x[e0 + 1] * f.inv);
    const int q2 = __shfl_down(q0, 1, kWave);
    const unsigned d0 = (unsigned)q1 - (unsigned)q0, d1 = (unsigned)q2 - (unsigned)q1;
    const unsigned m0 = (int)d0 < 0 ? 0u - d0 : d0;
    const unsigned m1 = lane + 1 < kWave ? ((int)d1 < 0 ? 0u - d1 : d1) : 0u;
    acc[0] += tune_block_bytes<8>(m0, m1, lane);
    acc[1] += tune_block_bytes<16>(m0, m1, lane);
    acc[2] += tune_block_bytes<32>(m0, m1, lane);
    acc[3] += tune_block_bytes<64>(m0, m1, lane);
  }
#pragma unroll
  for (int c = 0; c < kTuneCands; c++) {
    const unsigned v = seg_sum<kWave>(acc[c]);
    if (lane == 0) red[w][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < kTuneCands) {
    uint32_t t = 0;
    for (int k = 0; k < kWavesPerBlock; k++) t += red[k][threadIdx.x];
    est[(size_t)blockIdx.x * kTuneCands + threadIdx.x] = t;
  }
}

// ---------------------------------------------------------------------------
// host-side dispatch on block size
// ---------------------------------------------------------------------------
template <int MODE>
static hipError_t launch_compress_mode(const CompressArgs& a, int S, int EPL, hipStream_t st) {
#define TSZP_CASE(SS, EE)                                                                  \
  if (S == SS && EPL == EE) {                                                              \
    hipLaunchKernelGGL((k_compress<MODE, SS, EE>), dim3(a.ntiles), dim3(kBlockThreads), 0, st, a); \
    return hipGetLastError();                                                              \
  }
  TSZP_CASE(1, 1) TSZP_CASE(2, 1) TSZP_CASE(4, 1) TSZP_CASE(8, 1) TSZP_CASE(16, 1)
  TSZP_CASE(8, 4) TSZP_CASE(16, 4) TSZP_CASE(64, 2) TSZP_CASE(64, 4) TSZP_CASE(64, 8)
  if constexpr (MODE == kModePlain || MODE == kModeInt) {
    TSZP_CASE(64, 16) TSZP_CASE(64, 32)
  }
#undef TSZP_CASE
  return hipErrorInvalidValue;
}

template <int MODE>
static hipError_t launch_decode_mode(const DecodeArgs& a, int S, int EPL, uint32_t ntiles, hipStream_t st) {
#define TSZP_CASE(SS, EE)                                                                 \
  if (S == SS && EPL == EE) {                                                             \
    hipLaunchKernelGGL((k_decode<MODE, SS, EE>), dim3(ntiles), dim3(kBlockThreads), 0, st, a); \
    return hipGetLastError();                                                             \
  }
  TSZP_CASE(1, 1) TSZP_CASE(2, 1) TSZP_CASE(4, 1) TSZP_CASE(8, 1) TSZP_CASE(16, 1)
  TSZP_CASE(8, 4) TSZP_CASE(16, 4) TSZP_CASE(64, 2) TSZP_CASE(64, 4) TSZP_CASE(64, 8)
  if constexpr (MODE == kModePlain || MODE == kModeInt) {
    TSZP_CASE(64, 16) TSZP_CASE(64, 32)
  }
#undef TSZP_CASE
  return hipErrorInvalidValue;
}

void block_geometry(uint32_t bs, int* S, int* EPL, uint32_t* bpt) {
  int s = 1, e = 1;
  if (bs <= 16) {
    while ((uint32_t)s < bs) s <<= 1;
  } else if (bs <= 64) {
    // four consecutive elements per lane: the per-block wave operations (maximum,
    // scans) are shared by four elements
    e = 4;
    while ((uint32_t)(4 * s) < bs) s <<= 1;
  } else {
    s = 64;
    while ((uint32_t)(64 * e) < bs) e <<= 1;
  }
  *S = s;
  *EPL = e;
  const int spw = e >= 16 ? 1 : 16 / e;
  *bpt = (uint32_t)(kWavesPerBlock * spw * (kWave / s));
}

// The fast path (k_compress_v4) takes block sizes 16, 32, 64, 128 and 256 (4 to 64 lanes
// per block; its tile holds as many blocks as the generic kernel's) with a 16-byte
// aligned input; classification also needs the fastest axis a multiple of 4 long, the type map
// 4-byte alignment.
template <int MODE, int S>
static hipError_t launch_v4(const CompressArgs& a, hipStream_t st) {
  if constexpr (MODE == kModeTopoClassify) {
    if (a.grid.ndim == 3) {
      hipLaunchKernelGGL((k_compress_v4<MODE, S, 3>), dim3(a.ntiles), dim3(kBlockThreads), 0, st, a);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_compress_v4<MODE, S, 2>), dim3(a.ntiles), dim3(kBlockThreads), 0, st, a);
  return hipGetLastError();
}

template <int MODE>
static bool try_v4(const CompressArgs& a, hipStream_t st, hipError_t* e) {
  if ((a.bs != 16 && a.bs != 32 && a.bs != 64 && a.bs != 128 && a.bs != 256) ||
      (reinterpret_cast<uintptr_t>(a.in) & 15u))
    return false;
  if (MODE == kModeTopoClassify) {
    const Grid& g = a.grid;
    if (g.ndim != 2 && g.ndim != 3) return false;
    if ((g.ndim == 3 ? g.d2 : g.d1) % 4u) return false;
  }
  if ((MODE == kModeTopoMap || MODE == kModeTopoInt) && (reinterpret_cast<uintptr_t>(a.tmap) & 3u)) return false;
  switch (a.bs) {
    case 16: *e = launch_v4<MODE, 4>(a, st); break;
    case 32: *e = launch_v4<MODE, 8>(a, st); break;
    case 64: *e = launch_v4<MODE, 16>(a, st); break;
    case 128: *e = launch_v4<MODE, 32>(a, st); break;
    default: *e = launch_v4<MODE, 64>(a, st); break;
  }
  return true;
}

hipError_t launch_compress(int mode, const CompressArgs& a, hipStream_t st) {
  int S, EPL;
  uint32_t bpt;
  block_geometry(a.bs, &S, &EPL, &bpt);
  hipError_t e = hipSuccess;
  if (!getenv("TSZP_CODEC_GENERIC")) {
    switch (mode) {
      case kModeTopoClassify: if (try_v4<kModeTopoClassify>(a, st, &e)) return e; break;
      case kModeTopoMap: if (try_v4<kModeTopoMap>(a, st, &e)) return e; break;
      case kModePlain: if (try_v4<kModePlain>(a, st, &e)) return e; break;
      case kModeInt: if (try_v4<kModeInt>(a, st, &e)) return e; break;
      case kModeTopoInt: if (try_v4<kModeTopoInt>(a, st, &e)) return e; break;
    }
  }
  switch (mode) {
    case kModeTopoClassify: return launch_compress_mode<kModeTopoClassify>(a, S, EPL, st);
    case kModeTopoMap: return launch_compress_mode<kModeTopoMap>(a, S, EPL, st);
    case kModePlain: return launch_compress_mode<kModePlain>(a, S, EPL, st);
    case kModeInt: return launch_compress_mode<kModeInt>(a, S, EPL, st);
    case kModeTopoInt: return launch_compress_mode<kModeTopoInt>(a, S, EPL, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_decode(int mode, const DecodeArgs& a, hipStream_t st) {
  int S, EPL;
  uint32_t bpt;
  block_geometry(a.bs, &S, &EPL, &bpt);
  const uint32_t ntiles = (a.nb + bpt - 1) / bpt;
  if (ntiles == 0) return hipSuccess;
  switch (mode) {
    case kModeTopoClassify:
    case kModeTopoMap: return launch_decode_mode<kModeTopoMap>(a, S, EPL, ntiles, st);
    case kModePlain: return launch_decode_mode<kModePlain>(a, S, EPL, ntiles, st);
    case kModeInt: return launch_decode_mode<kModeInt>(a, S, EPL, ntiles, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_tilemap(const BatchField* fields, uint32_t count, uint32_t* tmap, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_batch_tilemap, dim3(count), dim3(kBlockThreads), 0, st, fields, tmap);
  return hipGetLastError();
}

template <int MODE, int S, int EPL, bool V4, int ND>
static hipError_t launch_cb(const BatchField* fields, const uint32_t* tmap, uint32_t base, uint32_t ntiles,
                            uint64_t* status, uint64_t* status_cnt, CompressResult* res, uint32_t* counter,
                            hipStream_t st) {
  hipLaunchKernelGGL((k_compress_batch<MODE, S, EPL, V4, ND>), dim3(ntiles), dim3(kBlockThreads), 0, st, fields,
                     tmap, base, status, status_cnt, res, counter);
  return hipGetLastError();
}

template <int MODE, int ND>
static hipError_t launch_cb_v4(uint32_t bs, const BatchField* fields, const uint32_t* tmap, uint32_t base,
                               uint32_t ntiles, uint64_t* status, uint64_t* status_cnt, CompressResult* res,
                               uint32_t* counter, hipStream_t st) {
  switch (bs) {
    case 16: return launch_cb<MODE, 4, 4, true, ND>(fields, tmap, base, ntiles, status, status_cnt, res, counter, st);
    case 32: return launch_cb<MODE, 8, 4, true, ND>(fields, tmap, base, ntiles, status, status_cnt, res, counter, st);
    case 64: return launch_cb<MODE, 16, 4, true, ND>(fields, tmap, base, ntiles, status, status_cnt, res, counter, st);
    case 128: return launch_cb<MODE, 32, 4, true, ND>(fields, tmap, base, ntiles, status, status_cnt, res, counter, st);
    case 256: return launch_cb<MODE, 64, 4, true, ND>(fields, tmap, base, ntiles, status, status_cnt, res, counter, st);
  }
  return hipErrorInvalidValue;
}

// generic geometries: every block size up to 257 (topology and sort positions), up to
// 1025 (plain)
template <int MODE>
static hipError_t launch_cb_generic(uint32_t bs, const BatchField* fields, const uint32_t* tmap, uint32_t base,
                                    uint32_t ntiles, uint64_t* status, uint64_t* status_cnt, CompressResult* res,
                                    uint32_t* counter, hipStream_t st) {
  int S, EPL;
  uint32_t bpt;
  block_geometry(bs, &S, &EPL, &bpt);
#define TSZP_CASE(SS, EE)                                                                                      \
  if (S == SS && EPL == EE)                                                                                    \
    return launch_cb<MODE, SS, EE, false, 2>(fields, tmap, base, ntiles, status, status_cnt, res, counter, st);
  TSZP_CASE(1, 1) TSZP_CASE(2, 1) TSZP_CASE(4, 1) TSZP_CASE(8, 1) TSZP_CASE(16, 1)
  TSZP_CASE(8, 4) TSZP_CASE(16, 4) TSZP_CASE(64, 2) TSZP_CASE(64, 4) TSZP_CASE(64, 8)
  if constexpr (MODE == kModePlain) {
    TSZP_CASE(64, 16) TSZP_CASE(64, 32)
  }
#undef TSZP_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_compress_batch(int mode, uint32_t bs_class, bool v4, int nd, const BatchField* fields,
                                 const uint32_t* tmap, uint32_t base, uint32_t ntiles, uint64_t* status,
                                 uint64_t* status_cnt, CompressResult* res, uint32_t* counter, hipStream_t st) {
  if (!ntiles) return hipSuccess;
  switch (mode) {
    case kModePlain:
      return v4 ? launch_cb_v4<kModePlain, 2>(bs_class, fields, tmap, base, ntiles, status, nullptr, res, counter, st)
                : launch_cb_generic<kModePlain>(bs_class, fields, tmap, base, ntiles, status, nullptr, res, counter, st);
    case kModeInt:
      return v4 ? launch_cb_v4<kModeInt, 2>(bs_class, fields, tmap, base, ntiles, status, nullptr, res, counter, st)
                : launch_cb_generic<kModeInt>(bs_class, fields, tmap, base, ntiles, status, nullptr, res, counter, st);
    case kModeTopoClassify:
      if (!v4)
        return launch_cb_generic<kModeTopoClassify>(bs_class, fields, tmap, base, ntiles, status, status_cnt, res,
                                                    counter, st);
      return nd == 3 ? launch_cb_v4<kModeTopoClassify, 3>(bs_class, fields, tmap, base, ntiles, status, status_cnt,
                                                          res, counter, st)
                     : launch_cb_v4<kModeTopoClassify, 2>(bs_class, fields, tmap, base, ntiles, status, status_cnt,
                                                          res, counter, st);
  }
  return hipErrorInvalidValue;
}

template <int MODE>
static hipError_t launch_db(uint32_t bs_class, const BatchField* fields, const uint32_t* tmap, uint32_t base,
                            uint32_t ntiles, CompressResult* res, hipStream_t st) {
  int S, EPL;
  uint32_t bpt;
  block_geometry(bs_class, &S, &EPL, &bpt);
#define TSZP_CASE(SS, EE)                                                                                 \
  if (S == SS && EPL == EE) {                                                                             \
    hipLaunchKernelGGL((k_decode_batch<MODE, SS, EE>), dim3(ntiles), dim3(kBlockThreads), 0, st, fields, tmap, \
                       base, res);                                                                        \
    return hipGetLastError();                                                                             \
  }
  TSZP_CASE(1, 1) TSZP_CASE(2, 1) TSZP_CASE(4, 1) TSZP_CASE(8, 1) TSZP_CASE(16, 1)
  TSZP_CASE(8, 4) TSZP_CASE(16, 4) TSZP_CASE(64, 2) TSZP_CASE(64, 4) TSZP_CASE(64, 8)
  if constexpr (MODE == kModePlain) {
    TSZP_CASE(64, 16) TSZP_CASE(64, 32)
  }
#undef TSZP_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_decode_batch(int mode, uint32_t bs_class, const BatchField* fields, const uint32_t* tmap,
                               uint32_t base, uint32_t ntiles, CompressResult* res, hipStream_t st) {
  if (!ntiles) return hipSuccess;
  switch (mode) {
    case kModePlain: return launch_db<kModePlain>(bs_class, fields, tmap, base, ntiles, res, st);
    case kModeTopoMap: return launch_db<kModeTopoMap>(bs_class, fields, tmap, base, ntiles, res, st);
    case kModeInt: return launch_db<kModeInt>(bs_class, fields, tmap, base, ntiles, res, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_tune(const BatchField* fields, uint32_t count, uint32_t* est, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_batch_tune, dim3(count), dim3(kBlockThreads), 0, st, fields, est);
  return hipGetLastError();
}

}  // namespace tszp
