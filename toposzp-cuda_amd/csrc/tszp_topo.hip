// tszp_topo.hip -- stand-alone critical-point classification for gfx950.
//
// Classifies every grid point (4 neighbours in 2D, 6 in 3D) and compacts the
// critical points into the reference's AoS record layout in row-major order, i.e.
// the order the reference's finder produces with one thread
// (SZp/src/szp_float.cc:28-120, SZp/src/szp_topology_3d.cc:30-129).  Compaction is
// a single-pass look-back scan of per-tile counts, so the order is deterministic.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

constexpr int kClsSteps = 16;                                   // 64-element steps per wave
constexpr uint32_t kClsTile = kBlockThreads * kClsSteps;        // 4096 elements

uint32_t classify_tiles(uint32_t n) { return (n + kClsTile - 1) / kClsTile; }

__device__ __forceinline__ int classify_point(const float* __restrict__ f, const Grid& g, uint32_t e,
                                              uint32_t* cx, uint32_t* cy, uint32_t* cz) {
  const float c = f[e];
  if (g.ndim == 2) {
    const uint32_t i = fdiv(e, g.div_row), j = e - i * g.d1;
    *cx = i; *cy = j; *cz = 0;
    if (i == 0 || j == 0 || i + 1 >= g.d0 || j + 1 >= g.d1) return 0;
    const float up = f[e - g.d1], dn = f[e + g.d1], lf = f[e - 1], rt = f[e + 1];
    if (c > up && c > dn && c > lf && c > rt) return 1;
    if (c < up && c < dn && c < lf && c < rt) return 2;
    if ((c < up && c < dn && c > lf && c > rt) || (c > up && c > dn && c < lf && c < rt)) return 3;
    return 0;
  }
  const uint32_t x = fdiv(e, g.div_plane), r = e - x * g.plane;
  const uint32_t y = fdiv(r, g.div_row), z = r - y * g.d2;
  *cx = x; *cy = y; *cz = z;
  if (x == 0 || y == 0 || z == 0 || x + 1 >= g.d0 || y + 1 >= g.d1 || z + 1 >= g.d2) return 0;
  const float xm = f[e - g.plane], xp = f[e + g.plane], ym = f[e - g.d2], yp = f[e + g.d2];
  const float zm = f[e - 1], zp = f[e + 1];
  if (c > xm && c > xp && c > ym && c > yp && c > zm && c > zp) return 1;
  if (c < xm && c < xp && c < ym && c < yp && c < zm && c < zp) return 2;
  const int hi = (c > xm && c > xp) + (c > ym && c > yp) + (c > zm && c > zp);
  const int lo = (c < xm && c < xp) + (c < ym && c < yp) + (c < zm && c < zp);
  return (hi >= 1 && lo >= 1) ? 3 : 0;
}

__global__ __launch_bounds__(kBlockThreads) void k_classify(ClassifyArgs a) {
  __shared__ uint32_t step_tot[kWavesPerBlock * kClsSteps];
  __shared__ uint32_t sh_tile;
  __shared__ uint64_t sh_pre;
  const int tid = threadIdx.x, w = tid / kWave, lane = tid & (kWave - 1);
  if (tid == 0) sh_tile = tile_id(a.tile_counter);
  __syncthreads();
  const uint32_t tile = sh_tile;
  const uint32_t base = tile * kClsTile + w * (kWave * kClsSteps);
  const uint64_t lt = lanemask_lt(lane);

  uint32_t tbits = 0;
#pragma unroll
  for (int s = 0; s < kClsSteps; s++) {
    const uint32_t e = base + s * kWave + lane;
    uint32_t x, y, z;
    const int t = e < a.n ? classify_point(a.f, a.grid, e, &x, &y, &z) : 0;
    if (e < a.n && a.types) a.types[e] = (uint8_t)t;
    tbits |= (uint32_t)t << (2 * s);
    const uint64_t m = __ballot(t != 0);
    if (lane == 0) step_tot[w * kClsSteps + s] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (w == 0) {
    const uint32_t v = step_tot[lane];
    const uint32_t incl = seg_incl_scan<kWave>(v, lane);
    step_tot[lane] = incl - v;
    const uint32_t tot = __shfl(incl, kWave - 1, kWave);
    const uint64_t ex = lookback(a.st_cnt, tile, tot, &a.res->error);
    if (lane == 0) {
      sh_pre = ex;
      if (tile == a.ntiles - 1) a.res->cp_count = ex + tot;
    }
  }
  __syncthreads();
  if (!a.cps) return;
  const uint32_t pre = (uint32_t)sh_pre;
#pragma unroll
  for (int s = 0; s < kClsSteps; s++) {
    const int t = (tbits >> (2 * s)) & 3;
    const uint64_t m = __ballot(t != 0);
    if (!t) continue;
    const uint32_t pos = pre + step_tot[w * kClsSteps + s] + (uint32_t)__popcll(m & lt);
    if (pos >= a.cp_cap) {
      atomicOr(&a.res->error, TSZP_ERR_OVERFLOW);
      continue;
    }
    const uint32_t e = base + s * kWave + lane;
    uint32_t x, y, z;
    classify_point(a.f, a.grid, e, &x, &y, &z);  // recovers the coordinates
    const int bin = cp_bin(a.f[e], a.eb, a.inv);
    int32_t* r = a.cps + (size_t)pos * a.rec;
    if (a.rec == 5) {
      r[0] = (int)x; r[1] = (int)y; r[2] = t; r[3] = bin; r[4] = 0;
    } else {
      r[0] = (int)x; r[1] = (int)y; r[2] = (int)z; r[3] = t; r[4] = bin; r[5] = 0;
    }
  }
}

hipError_t launch_classify(const ClassifyArgs& a, hipStream_t st) {
  if (a.ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_classify, dim3(a.ntiles), dim3(kBlockThreads), 0, st, a);
  return hipGetLastError();
}

// Sort keys of a caller's AoS CP list: the bin stored in the record (biased so that
// unsigned order is signed order) is the high key word and the order-preserving
// value key the low word; the payload is the record index.  The drop-in sort
// groups by the stored bin exactly as the reference does
// (SZp/src/szp_float.cc:149-247), whatever the caller put there.
__global__ __launch_bounds__(kBlockThreads) void k_cp_keys(const int32_t* __restrict__ cps, uint32_t count, int rec,
                                                          const float* __restrict__ data, uint64_t data_n,
                                                          uint32_t row, uint64_t plane, uint32_t* keys,
                                                          uint32_t* hi, uint32_t* pay, uint32_t* err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int32_t* r = cps + (size_t)i * rec;
  uint64_t flat;
  int bin;
  if (rec == 5) {
    flat = (uint64_t)(int64_t)r[0] * row + (uint64_t)(int64_t)r[1];
    bin = r[3];
  } else {
    flat = (uint64_t)(int64_t)r[0] * plane + (uint64_t)(int64_t)r[1] * row + (uint64_t)(int64_t)r[2];
    bin = r[4];
  }
  uint32_t k = 0;
  if (flat < data_n) {
    k = okey(data[flat]);
  } else {
    atomicOr(err, TSZP_ERR_STREAM);
  }
  keys[i] = k;
  hi[i] = (uint32_t)bin ^ 0x80000000u;
  pay[i] = i;
}

hipError_t launch_cp_keys(const int32_t* cps, uint32_t count, int rec, const float* data, uint64_t data_n,
                          uint32_t row, uint64_t plane, uint32_t* keys, uint32_t* hi, uint32_t* pay, uint32_t* err,
                          hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_cp_keys, dim3((count + kBlockThreads - 1) / kBlockThreads), dim3(kBlockThreads), 0, st, cps,
                     count, rec, data, data_n, row, plane, keys, hi, pay, err);
  return hipGetLastError();
}

}  // namespace tszp
