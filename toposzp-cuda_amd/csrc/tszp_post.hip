// tszp_post.hip -- topology-restoring post-process and its verification.
//
// The reference restores critical points after decompression with sequential host
// stencils (example/test_cuda_topology.cu:118-326, example/test_cuda_topology_3d.cu:
// 77-434) that leave 2-15 % of them lost.  Here the decoded field d = (float)q * eb is
// turned into y with classify(y) == the transmitted types at every interior point,
// as a constraint repair that stays inside each quantization cell:
//
//   * allowed interval of a point: [d, d + 0.45 eb] when d > 0, [d - 0.45 eb, d] when
//     d < 0, [-0.45 eb, 0.45 eb] when d == 0.  Points of different cells therefore
//     keep their order (q is monotone in the original value), and |y - x| < eb
//     except in the zero cell (< 1.45 eb);
//   * start: every point at the middle of its interval (equal values inside a
//     plateau make its regular points regular), maxima / minima pushed to the upper
//     / lower part of their interval, ordered by their sort positions;
//   * local search: points within one step of a mismatch move to the value that
//     minimises the mismatches of their closed neighbourhood.  Points of one colour
//     class are at least three apart ((i + 2j) mod 5 in 2D, (x + 2y + 3z) mod 7 in
//     3D), so a whole class moves in parallel with exact gains, and every accepted
//     move strictly lowers the total;
//   * pair moves for what remains: a mismatched point and one neighbour of its
//     plateau move together over a value grid, one point per 8-wide tile of one
//     parity class at a time (tile locks).
// A verification kernel counts missing / spurious / changed critical points.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

// ---------------------------------------------------------------------------------
// geometry helpers
// ---------------------------------------------------------------------------------
struct Coord {
  int c[3];
};

__device__ __forceinline__ Coord coord_of(const Grid& g, uint32_t e) {
  Coord k;
  if (g.ndim == 2) {
    const uint32_t i = fdiv(e, g.div_row);
    k.c[0] = (int)i;
    k.c[1] = (int)(e - i * g.d1);
    k.c[2] = 0;
  } else {
    const uint32_t x = fdiv(e, g.div_plane), r = e - x * g.plane;
    const uint32_t y = fdiv(r, g.div_row);
    k.c[0] = (int)x;
    k.c[1] = (int)y;
    k.c[2] = (int)(r - y * g.d2);
  }
  return k;
}

__device__ __forceinline__ bool inside(const Grid& g, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < (int)g.d0 && y < (int)g.d1 && z < (int)(g.ndim == 3 ? g.d2 : 1u);
}
__device__ __forceinline__ bool interior3(const Grid& g, int x, int y, int z) {
  if (g.ndim == 2) return x > 0 && y > 0 && x + 1 < (int)g.d0 && y + 1 < (int)g.d1;
  return x > 0 && y > 0 && z > 0 && x + 1 < (int)g.d0 && y + 1 < (int)g.d1 && z + 1 < (int)g.d2;
}
__device__ __forceinline__ uint32_t flat_of(const Grid& g, int x, int y, int z) {
  return g.ndim == 2 ? (uint32_t)x * g.d1 + (uint32_t)y : (uint32_t)x * g.plane + (uint32_t)y * g.d2 + (uint32_t)z;
}

// neighbour direction k (2D: up, down, left, right; 3D: -x, +x, -y, +y, -z, +z)
__device__ __forceinline__ void dir_of(int ndim, int k, int* dx, int* dy, int* dz) {
  *dx = *dy = *dz = 0;
  const int s = (k & 1) ? 1 : -1;
  const int axis = k >> 1;
  if (axis == 0) *dx = s;
  else if (axis == 1) *dy = s;
  else *dz = s;
  (void)ndim;
}

// classification from centre + neighbour values in dir_of order
__device__ __forceinline__ int classify_vals(int ndim, float c, const float* nb) {
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

// allowed interval of a decoded value
__device__ __forceinline__ void interval_of(float d, float eb, float* lo, float* hi) {
  const float w = 0.45f * eb;
  if (d > 0.f) {
    *lo = d;
    *hi = __fadd_rn(d, w);
  } else if (d < 0.f) {
    *lo = __fsub_rn(d, w);
    *hi = d;
  } else {
    *lo = -w;
    *hi = w;
  }
}

// classification of point p = (x,y,z) with y[ovr] replaced by v (reads global memory)
__device__ int class_at(const Grid& g, const float* __restrict__ yv, int x, int y, int z, uint32_t ovr_a, float va,
                        uint32_t ovr_b, float vb) {
  if (!interior3(g, x, y, z)) return 0;
  const int K = 2 * g.ndim;
  const uint32_t e = flat_of(g, x, y, z);
  float nb[6];
  for (int k = 0; k < K; k++) {
    int dx, dy, dz;
    dir_of(g.ndim, k, &dx, &dy, &dz);
    const uint32_t f = flat_of(g, x + dx, y + dy, z + dz);
    nb[k] = f == ovr_a ? va : (f == ovr_b ? vb : yv[f]);
  }
  const float c = e == ovr_a ? va : (e == ovr_b ? vb : yv[e]);
  return classify_vals(g.ndim, c, nb);
}

// ---------------------------------------------------------------------------------
// initial values
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlockThreads) void k_pp_init(const float* __restrict__ d, const uint8_t* __restrict__ t,
                                                          float* __restrict__ y, uint32_t n, float eb) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float lo, hi;
  interval_of(d[e], eb, &lo, &hi);
  const float base = 0.5f * (lo + hi);
  const int tt = t[e];
  float v = base;
  if (tt == 1) v = base + 0.7f * (hi - base);
  if (tt == 2) v = base - 0.7f * (base - lo);
  y[e] = v;
}

// extrema ordered by their sort positions inside the upper / lower part of their
// interval (a larger sort position -> a larger value, as in the original field)
__global__ __launch_bounds__(kBlockThreads) void k_pp_extrema(const float* __restrict__ d,
                                                             const uint8_t* __restrict__ t,
                                                             const uint32_t* __restrict__ ext_pos,
                                                             const int32_t* __restrict__ sp, uint32_t n_ext,
                                                             const uint32_t* __restrict__ n_found, uint32_t n,
                                                             float* __restrict__ y, float eb) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_ext || k >= *n_found) return;
  const uint32_t e = ext_pos[k];
  if (e >= n) return;
  float lo, hi;
  interval_of(d[e], eb, &lo, &hi);
  const float base = 0.5f * (lo + hi);
  const float r = 1.f - 1.f / (1.f + (float)max(sp[k], 0) * (1.f / 64.f));  // 0 <= r < 1, increasing
  if (t[e] == 1) y[e] = base + (hi - base) * (0.55f + 0.4f * r);
  else if (t[e] == 2) y[e] = base - (base - lo) * (0.95f - 0.4f * r);
}

// positions of the extrema in row-major order (compaction by a look-back scan)
constexpr int kPpSteps = 16;
constexpr uint32_t kPpTile = kBlockThreads * kPpSteps;

__global__ __launch_bounds__(kBlockThreads) void k_pp_ext_scan(const uint8_t* __restrict__ t, uint32_t n,
                                                              uint32_t* __restrict__ ext_pos, uint32_t cap,
                                                              uint64_t* st, uint32_t* counter, uint32_t* err,
                                                              uint32_t* total) {
  __shared__ uint32_t step_tot[kWavesPerBlock * kPpSteps];
  __shared__ uint32_t sh_tile;
  __shared__ uint64_t sh_pre;
  const int tid = threadIdx.x, w = tid / kWave, lane = tid & (kWave - 1);
  if (tid == 0) sh_tile = atomicAdd(counter, 1u);
  __syncthreads();
  const uint32_t tile = sh_tile;
  const uint32_t base = tile * kPpTile + w * (kWave * kPpSteps);
  uint32_t bits = 0;
#pragma unroll
  for (int s = 0; s < kPpSteps; s++) {
    const uint32_t e = base + s * kWave + lane;
    const int tt = e < n ? t[e] : 0;
    const bool x = tt == 1 || tt == 2;
    bits |= (uint32_t)x << s;
    const uint64_t m = __ballot(x);
    if (lane == 0) step_tot[w * kPpSteps + s] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (w == 0) {
    const uint32_t v = step_tot[lane];
    const uint32_t incl = seg_incl_scan<kWave>(v, lane);
    step_tot[lane] = incl - v;
    const uint32_t tot = __shfl(incl, kWave - 1, kWave);
    const uint64_t ex = lookback(st, tile, tot, err);
    if (lane == 0) {
      sh_pre = ex;
      if (tile == gridDim.x - 1) *total = (uint32_t)(ex + tot);
    }
  }
  __syncthreads();
  const uint64_t lt = lanemask_lt(lane);
#pragma unroll
  for (int s = 0; s < kPpSteps; s++) {
    const bool x = (bits >> s) & 1u;
    const uint64_t m = __ballot(x);
    if (!x) continue;
    const uint32_t pos = (uint32_t)sh_pre + step_tot[w * kPpSteps + s] + (uint32_t)__popcll(m & lt);
    if (pos < cap) ext_pos[pos] = base + s * kWave + lane;
  }
}

// ---------------------------------------------------------------------------------
// marking: mismatches and the active set (points within one step of a mismatch)
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlockThreads) void k_pp_mark(Grid g, const float* __restrict__ yv,
                                                          const uint8_t* __restrict__ t, uint32_t n,
                                                          uint32_t* __restrict__ nbad,
                                                          uint32_t* __restrict__ badlist, uint32_t badcap) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const Coord k = coord_of(g, e);
  const int c = class_at(g, yv, k.c[0], k.c[1], k.c[2], 0xffffffffu, 0.f, 0xffffffffu, 0.f);
  if (c != (int)t[e]) {
    const uint32_t slot = atomicAdd(nbad, 1u);
    if (slot < badcap) badlist[slot] = e;
  }
}

// active points of one colour: every point within one step of a mismatched point
__global__ __launch_bounds__(kBlockThreads) void k_pp_active(Grid g, const uint32_t* __restrict__ badlist,
                                                            uint32_t nbad, uint8_t* __restrict__ act) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbad) return;
  const uint32_t e = badlist[i];
  const Coord k = coord_of(g, e);
  act[e] = 1;
  for (int d = 0; d < 2 * g.ndim; d++) {
    int dx, dy, dz;
    dir_of(g.ndim, d, &dx, &dy, &dz);
    const int x = k.c[0] + dx, y = k.c[1] + dy, z = k.c[2] + dz;
    if (inside(g, x, y, z)) act[flat_of(g, x, y, z)] = 1;
  }
}

__device__ __forceinline__ int colour_of(const Grid& g, const Coord& k) {
  return g.ndim == 2 ? (k.c[0] + 2 * k.c[1]) % 5 : (k.c[0] + 2 * k.c[1] + 3 * k.c[2]) % 7;
}

// mismatches of the closed neighbourhood of (x,y,z) when y[e] = v
__device__ int nbhd_bad(const Grid& g, const float* __restrict__ yv, const uint8_t* __restrict__ t, const Coord& k,
                        uint32_t e, float v, uint32_t e2, float v2) {
  int bad = class_at(g, yv, k.c[0], k.c[1], k.c[2], e, v, e2, v2) != (int)t[e];
  for (int d = 0; d < 2 * g.ndim; d++) {
    int dx, dy, dz;
    dir_of(g.ndim, d, &dx, &dy, &dz);
    const int x = k.c[0] + dx, y = k.c[1] + dy, z = k.c[2] + dz;
    if (!inside(g, x, y, z)) continue;
    const uint32_t f = flat_of(g, x, y, z);
    bad += class_at(g, yv, x, y, z, e, v, e2, v2) != (int)t[f];
  }
  return bad;
}

// single-point moves of one colour class over the active set
__global__ __launch_bounds__(kBlockThreads) void k_pp_move(Grid g, const float* __restrict__ d,
                                                          const uint8_t* __restrict__ t, float* __restrict__ yv,
                                                          uint8_t* __restrict__ act, uint32_t n, float eb, int colour,
                                                          uint32_t* __restrict__ moves) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || !act[e]) return;
  const Coord k = coord_of(g, e);
  if (colour_of(g, k) != colour) return;
  act[e] = 0;
  const float cur = yv[e];
  const int v0 = nbhd_bad(g, yv, t, k, e, cur, 0xffffffffu, 0.f);
  if (v0 == 0) return;
  float lo, hi;
  interval_of(d[e], eb, &lo, &hi);
  // candidate values: the interval ends, every plateau neighbour's value, midpoints
  float vals[8];
  int nv = 0;
  vals[nv++] = lo;
  vals[nv++] = hi;
  for (int q = 0; q < 2 * g.ndim; q++) {
    int dx, dy, dz;
    dir_of(g.ndim, q, &dx, &dy, &dz);
    const int x = k.c[0] + dx, y = k.c[1] + dy, z = k.c[2] + dz;
    if (!inside(g, x, y, z)) continue;
    const float v = yv[flat_of(g, x, y, z)];
    if (v >= lo && v <= hi) vals[nv++] = v;
  }
  for (int a = 1; a < nv; a++) {  // insertion sort
    const float v = vals[a];
    int b = a - 1;
    while (b >= 0 && vals[b] > v) {
      vals[b + 1] = vals[b];
      b--;
    }
    vals[b + 1] = v;
  }
  int best = v0;
  float bestv = cur;
  for (int a = 0; a < nv; a++) {
    if (a > 0 && vals[a] == vals[a - 1]) continue;
    const float c1 = vals[a];
    int s = nbhd_bad(g, yv, t, k, e, c1, 0xffffffffu, 0.f);
    if (s < best) {
      best = s;
      bestv = c1;
    }
    if (a + 1 < nv && vals[a + 1] > vals[a]) {
      const float m = 0.5f * (vals[a] + vals[a + 1]);
      if (m > vals[a] && m < vals[a + 1]) {
        s = nbhd_bad(g, yv, t, k, e, m, 0xffffffffu, 0.f);
        if (s < best) {
          best = s;
          bestv = m;
        }
      }
    }
  }
  if (best < v0) {
    yv[e] = bestv;
    atomicAdd(moves, 1u);
  }
}

// pair moves: a mismatched point with one neighbour of its plateau, over a 9 x 9 grid
// of values.  One point per 8-wide tile of the given parity class at a time.
constexpr int kPairGrid = 9;

__device__ int union_bad(const Grid& g, const float* __restrict__ yv, const uint8_t* __restrict__ t, const Coord& a,
                         uint32_t ea, float va, const Coord& b, uint32_t eb_, float vb) {
  // mismatches over the closed neighbourhoods of a and b (points counted once)
  int bad = nbhd_bad(g, yv, t, a, ea, va, eb_, vb);  // includes b, a neighbour of a
  for (int d = 0; d < 2 * g.ndim; d++) {
    int dx, dy, dz;
    dir_of(g.ndim, d, &dx, &dy, &dz);
    const int x = b.c[0] + dx, y = b.c[1] + dy, z = b.c[2] + dz;
    if (!inside(g, x, y, z)) continue;
    const uint32_t f = flat_of(g, x, y, z);
    if (f == ea) continue;
    // skip points already counted as neighbours of a
    const int md = abs(x - a.c[0]) + abs(y - a.c[1]) + abs(z - a.c[2]);
    if (md <= 1) continue;
    bad += class_at(g, yv, x, y, z, ea, va, eb_, vb) != (int)t[f];
  }
  return bad;
}

__global__ __launch_bounds__(kBlockThreads) void k_pp_pair(Grid g, const float* __restrict__ d,
                                                          const uint8_t* __restrict__ t, float* __restrict__ yv,
                                                          const uint32_t* __restrict__ badlist, uint32_t nbad,
                                                          uint8_t* __restrict__ done, uint32_t* __restrict__ locks,
                                                          uint32_t tiles_x, uint32_t tiles_y, int phase, float eb,
                                                          uint32_t* __restrict__ moves, uint32_t* __restrict__ pending) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbad || done[i]) return;
  const uint32_t e = badlist[i];
  const Coord k = coord_of(g, e);
  const int tx = k.c[0] >> 3, ty = k.c[1] >> 3, tz = k.c[2] >> 3;
  const int ph = (tx & 1) | ((ty & 1) << 1) | ((tz & 1) << 2);
  if (ph != phase) return;
  const uint32_t tile = ((uint32_t)tz * tiles_y + (uint32_t)ty) * tiles_x + (uint32_t)tx;
  if (atomicCAS(&locks[tile], 0u, 1u) != 0u) {
    atomicAdd(pending, 1u);  // another point of this tile goes first
    return;
  }
  done[i] = 1;
  if (class_at(g, yv, k.c[0], k.c[1], k.c[2], 0xffffffffu, 0.f, 0xffffffffu, 0.f) == (int)t[e]) return;
  const float de = d[e];
  float lo, hi;
  interval_of(de, eb, &lo, &hi);
  for (int q = 0; q < 2 * g.ndim; q++) {
    int dx, dy, dz;
    dir_of(g.ndim, q, &dx, &dy, &dz);
    Coord kb;
    kb.c[0] = k.c[0] + dx;
    kb.c[1] = k.c[1] + dy;
    kb.c[2] = k.c[2] + dz;
    if (!inside(g, kb.c[0], kb.c[1], kb.c[2])) continue;
    const uint32_t f = flat_of(g, kb.c[0], kb.c[1], kb.c[2]);
    if (d[f] != de) continue;  // same plateau only
    const float ya = yv[e], yb = yv[f];
    const int base = union_bad(g, yv, t, k, e, ya, kb, f, yb);
    int best = base;
    float ba = ya, bb = yb;
    for (int u = 0; u < kPairGrid; u++) {
      const float va = lo + (hi - lo) * ((float)u + 0.5f) * (1.f / kPairGrid);
      for (int v = 0; v < kPairGrid; v++) {
        const float vb = lo + (hi - lo) * ((float)v + 0.5f) * (1.f / kPairGrid);
        const int s = union_bad(g, yv, t, k, e, va, kb, f, vb);
        if (s < best) {
          best = s;
          ba = va;
          bb = vb;
        }
      }
    }
    if (best < base) {
      yv[e] = ba;
      yv[f] = bb;
      atomicAdd(moves, 1u);
      break;
    }
  }
}

// ---------------------------------------------------------------------------------
// verification: missing / spurious / changed critical points and the largest
// deviation from the decoded field
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlockThreads) void k_pp_verify(Grid g, const float* __restrict__ yv,
                                                            const uint8_t* __restrict__ t,
                                                            const float* __restrict__ d, uint32_t n,
                                                            unsigned long long* __restrict__ out /* 4 */) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t miss = 0, spur = 0, chg = 0;
  uint32_t dev = 0;
  if (e < n) {
    const Coord k = coord_of(g, e);
    const int c = class_at(g, yv, k.c[0], k.c[1], k.c[2], 0xffffffffu, 0.f, 0xffffffffu, 0.f);
    const int tt = t[e];
    miss = (tt != 0 && c == 0);
    spur = (tt == 0 && c != 0);
    chg = (tt != 0 && c != 0 && c != tt);
    if (d) dev = __float_as_uint(fabsf(yv[e] - d[e]));
  }
  miss = seg_sum<kWave>(miss);
  spur = seg_sum<kWave>(spur);
  chg = seg_sum<kWave>(chg);
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t w = __shfl_xor(dev, o, kWave);
    dev = w > dev ? w : dev;
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    if (miss) atomicAdd(&out[0], (unsigned long long)miss);
    if (spur) atomicAdd(&out[1], (unsigned long long)spur);
    if (chg) atomicAdd(&out[2], (unsigned long long)chg);
    if (dev) atomicMax(&out[3], (unsigned long long)dev);
  }
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
static inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + kBlockThreads - 1) / kBlockThreads); }

hipError_t pp_init(const float* d, const uint8_t* t, float* y, uint32_t n, float eb, hipStream_t st) {
  hipLaunchKernelGGL(k_pp_init, dim3(blocks_for(n)), dim3(kBlockThreads), 0, st, d, t, y, n, eb);
  return hipGetLastError();
}

hipError_t pp_ext_scan(const uint8_t* t, uint32_t n, uint32_t* ext_pos, uint32_t cap, uint64_t* st_words,
                       uint32_t* counter, uint32_t* err, uint32_t* total, hipStream_t st) {
  const uint32_t tiles = (n + kPpTile - 1) / kPpTile;
  hipLaunchKernelGGL(k_pp_ext_scan, dim3(tiles), dim3(kBlockThreads), 0, st, t, n, ext_pos, cap, st_words, counter,
                     err, total);
  return hipGetLastError();
}
uint32_t pp_ext_tiles(uint32_t n) { return (n + kPpTile - 1) / kPpTile; }

hipError_t pp_extrema(const float* d, const uint8_t* t, const uint32_t* ext_pos, const int32_t* sp, uint32_t n_ext,
                      const uint32_t* n_found, uint32_t n, float* y, float eb, hipStream_t st) {
  if (!n_ext) return hipSuccess;
  hipLaunchKernelGGL(k_pp_extrema, dim3(blocks_for(n_ext)), dim3(kBlockThreads), 0, st, d, t, ext_pos, sp, n_ext,
                     n_found, n, y, eb);
  return hipGetLastError();
}

hipError_t pp_mark(const Grid& g, const float* y, const uint8_t* t, uint32_t n, uint32_t* nbad, uint32_t* badlist,
                   uint32_t badcap, hipStream_t st) {
  hipLaunchKernelGGL(k_pp_mark, dim3(blocks_for(n)), dim3(kBlockThreads), 0, st, g, y, t, n, nbad, badlist, badcap);
  return hipGetLastError();
}

hipError_t pp_active(const Grid& g, const uint32_t* badlist, uint32_t nbad, uint8_t* act, hipStream_t st) {
  if (!nbad) return hipSuccess;
  hipLaunchKernelGGL(k_pp_active, dim3(blocks_for(nbad)), dim3(kBlockThreads), 0, st, g, badlist, nbad, act);
  return hipGetLastError();
}

hipError_t pp_move(const Grid& g, const float* d, const uint8_t* t, float* y, uint8_t* act, uint32_t n, float eb,
                   int colour, uint32_t* moves, hipStream_t st) {
  hipLaunchKernelGGL(k_pp_move, dim3(blocks_for(n)), dim3(kBlockThreads), 0, st, g, d, t, y, act, n, eb, colour,
                     moves);
  return hipGetLastError();
}

hipError_t pp_pair(const Grid& g, const float* d, const uint8_t* t, float* y, const uint32_t* badlist, uint32_t nbad,
                   uint8_t* done, uint32_t* locks, uint32_t tiles_x, uint32_t tiles_y, int phase, float eb,
                   uint32_t* moves, uint32_t* pending, hipStream_t st) {
  if (!nbad) return hipSuccess;
  hipLaunchKernelGGL(k_pp_pair, dim3(blocks_for(nbad)), dim3(kBlockThreads), 0, st, g, d, t, y, badlist, nbad, done,
                     locks, tiles_x, tiles_y, phase, eb, moves, pending);
  return hipGetLastError();
}

hipError_t pp_verify(const Grid& g, const float* y, const uint8_t* t, const float* d, uint32_t n,
                     unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pp_verify, dim3(blocks_for(n)), dim3(kBlockThreads), 0, st, g, y, t, d, n, out);
  return hipGetLastError();
}

}  // namespace tszp
