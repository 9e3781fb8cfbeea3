// tszp_saddle.hip -- quantization and saddle fix of the double-precision 3D topology
// path (SZp/src/szp_cuda_topology_3d.cu:1065-1188):
//   types  <- critical-point list (mark_critical_types_3d_kernel, :602-618)
//   q      <- C cast of x * (1 / eb) in double (quantize_3d_kernel_double, :641-646)
//   three Jacobi passes of the saddle fix (fix_saddles_3d_kernel, :736-826)
// The reference snapshots the whole quantized field before each pass (a device copy of
// 4 bytes per point) and visits every interior point.  Here the interior saddles are
// listed once (count per workgroup, scan, list), q is never written during the passes,
// and each pass visits only the listed saddles: it reads the accumulated moves of the
// previous pass from a byte map (only saddles move, by at most one step per pass) and
// writes its own map; the moves are added to q once at the end.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

__global__ __launch_bounds__(kBlockThreads) void k_mark_cps3(const int32_t* __restrict__ cps, uint32_t count,
                                                            Grid g, uint8_t* __restrict__ types) {
  const uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  if (i >= count) return;
  const int32_t* r = cps + 6 * (size_t)i;
  const int x = r[0], y = r[1], z = r[2], t = r[3];
  if (x < 0 || y < 0 || z < 0 || (uint32_t)x >= g.d0 || (uint32_t)y >= g.d1 || (uint32_t)z >= g.d2) return;
  if (t >= 1 && t <= 3) types[(size_t)x * g.plane + (size_t)y * g.d2 + z] = (uint8_t)t;
}

__global__ __launch_bounds__(kBlockThreads) void k_quant_f64(const double* __restrict__ x, uint32_t n, double inv,
                                                            int32_t* __restrict__ q) {
  uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  const uint32_t stride = gridDim.x * kBlockThreads;
  for (; i < n; i += stride) q[i] = cvt_i32(x[i] * inv);
}

// new accumulated move of the saddle at e (interior, type 3)
__device__ __forceinline__ int saddle_move(const Grid& g, const int32_t* __restrict__ q,
                                           const uint8_t* __restrict__ types, const int8_t* __restrict__ cd,
                                           uint32_t e) {
  const uint32_t o[6] = {e - g.plane, e + g.plane, e - g.d2, e + g.d2, e - 1u, e + 1u};
  const int m0 = cd ? (int)cd[e] : 0;
  const int c = (int)((uint32_t)q[e] + (uint32_t)m0);
  int v[6];
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = (int)((uint32_t)q[o[k]] + (uint32_t)(cd ? (int)cd[o[k]] : 0));
  int hi = 0, lo = 0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    hi += c > v[2 * a] && c > v[2 * a + 1];
    lo += c < v[2 * a] && c < v[2 * a + 1];
  }
  if (hi >= 1 && lo >= 1) return m0;  // still a saddle
  int nt[6];
#pragma unroll
  for (int k = 0; k < 6; k++) nt[k] = types[o[k]];
  // c - 1, then c + 1 (c itself fails the test above)
#pragma unroll
  for (int step = -1; step <= 1; step += 2) {
    const int t = (int)((uint32_t)c + (uint32_t)step);
    int th = 0, tl = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int mx = v[2 * a] > v[2 * a + 1] ? v[2 * a] : v[2 * a + 1];
      const int mn = v[2 * a] < v[2 * a + 1] ? v[2 * a] : v[2 * a + 1];
      th += t > mx;
      tl += t < mn;
    }
    if (th >= 1 && tl >= 1) {
      bool ok = true;
#pragma unroll
      for (int k = 0; k < 6; k++) {
        if (nt[k] == 1 && v[k] <= t) ok = false;  // a neighbouring maximum stays above
        if (nt[k] == 2 && v[k] >= t) ok = false;  // a neighbouring minimum stays below
      }
      if (ok) return m0 + step;
    }
  }
  return m0;
}

constexpr uint32_t kSadPts = 16;                          // points per thread
constexpr uint32_t kSadBlockPts = kBlockThreads * kSadPts;  // points per workgroup

// the thread's 16 type bytes, bit r of the result set where point e0 + r is an interior
// saddle
__device__ __forceinline__ uint32_t saddle_bits(const Grid& g, const uint8_t* __restrict__ types, uint32_t n,
                                                uint32_t e0) {
  uint32_t w[4] = {0u, 0u, 0u, 0u};
  if (e0 + kSadPts <= n) {
    const uint4 v = *reinterpret_cast<const uint4*>(types + e0);
    w[0] = v.x;
    w[1] = v.y;
    w[2] = v.z;
    w[3] = v.w;
  } else {
    for (uint32_t r = 0; e0 + r < n; r++) w[r >> 2] |= (uint32_t)types[e0 + r] << (8 * (r & 3u));
  }
  uint32_t bits = 0;
#pragma unroll
  for (int r = 0; r < (int)kSadPts; r++) {
    if (((w[r >> 2] >> (8 * (r & 3))) & 0xffu) != 3u) continue;
    const uint32_t e = e0 + r;
    const uint32_t x = fdiv(e, g.div_plane), rr = e - x * g.plane;
    const uint32_t y = fdiv(rr, g.div_row), z = rr - y * g.d2;
    if (x == 0 || y == 0 || z == 0 || x + 1 >= g.d0 || y + 1 >= g.d1 || z + 1 >= g.d2) continue;
    bits |= 1u << r;
  }
  return bits;
}

// interior saddles per workgroup range of kSadBlockPts points
__global__ __launch_bounds__(kBlockThreads) void k_saddle_count(Grid g, const uint8_t* __restrict__ types, uint32_t n,
                                                               uint32_t* __restrict__ cnt) {
  __shared__ uint32_t lds4[kWavesPerBlock];
  const uint32_t e0 = blockIdx.x * kSadBlockPts + threadIdx.x * kSadPts;
  const uint32_t c = e0 < n ? (uint32_t)__popc(saddle_bits(g, types, n, e0)) : 0u;
  uint32_t total;
  block_excl_scan(c, lds4, &total);
  if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}

// list of interior saddles in index order; off: exclusive scan of the counts
__global__ __launch_bounds__(kBlockThreads) void k_saddle_list(Grid g, const uint8_t* __restrict__ types, uint32_t n,
                                                              const uint32_t* __restrict__ off,
                                                              uint32_t* __restrict__ list) {
  __shared__ uint32_t lds4[kWavesPerBlock];
  const uint32_t e0 = blockIdx.x * kSadBlockPts + threadIdx.x * kSadPts;
  uint32_t bits = e0 < n ? saddle_bits(g, types, n, e0) : 0u;
  uint32_t total;
  uint32_t pos = off[blockIdx.x] + block_excl_scan((uint32_t)__popc(bits), lds4, &total);
  while (bits) {
    const int r = __ffs(bits) - 1;
    bits &= bits - 1u;
    list[pos++] = e0 + (uint32_t)r;
  }
}

// one Jacobi pass over the listed saddles (the count is read on the device); the last
// pass also writes the moves in list order (upd), so that applying them reads a dense
// array instead of one scattered byte per saddle
__global__ __launch_bounds__(kBlockThreads) void k_saddle_pass(Grid g, const int32_t* __restrict__ q,
                                                              const uint8_t* __restrict__ types,
                                                              const uint32_t* __restrict__ list,
                                                              const uint32_t* __restrict__ count,
                                                              const int8_t* __restrict__ cd_in,
                                                              int8_t* __restrict__ cd_out, int8_t* __restrict__ upd) {
  const uint32_t m = *count;
  for (uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x; i < m; i += gridDim.x * kBlockThreads) {
    const uint32_t e = list[i];
    const int8_t mv = (int8_t)saddle_move(g, q, types, cd_in, e);
    if (cd_out) cd_out[e] = mv;
    if (upd) upd[i] = mv;
  }
}

// q += move at the listed saddles that moved
__global__ __launch_bounds__(kBlockThreads) void k_saddle_apply(int32_t* __restrict__ q,
                                                               const int8_t* __restrict__ upd,
                                                               const uint32_t* __restrict__ list,
                                                               const uint32_t* __restrict__ count,
                                                               uint32_t* moved) {
  const uint32_t m = *count;
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x; i < m; i += gridDim.x * kBlockThreads) {
    const int d = upd[i];
    if (d) {
      const uint32_t e = list[i];
      q[e] = (int)((uint32_t)q[e] + (uint32_t)d);
      c++;
    }
  }
  c = seg_sum<kWave>(c);
  if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(moved, c);
}

static unsigned grid_for(uint64_t items, unsigned cap) {
  uint64_t b = (items + kBlockThreads - 1) / kBlockThreads;
  if (b > cap) b = cap;
  return b ? (unsigned)b : 1u;
}

hipError_t launch_mark_cps3(const int32_t* cps, uint32_t count, const Grid& g, uint8_t* types, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_mark_cps3, dim3((count + kBlockThreads - 1) / kBlockThreads), dim3(kBlockThreads), 0, st,
                     cps, count, g, types);
  return hipGetLastError();
}

hipError_t launch_quant_f64(const double* x, uint32_t n, double inv, int32_t* q, hipStream_t st) {
  hipLaunchKernelGGL(k_quant_f64, dim3(grid_for(n, 8192)), dim3(kBlockThreads), 0, st, x, n, inv, q);
  return hipGetLastError();
}

uint32_t saddle_blocks(uint32_t n) { return (n + kSadBlockPts - 1) / kSadBlockPts; }

hipError_t launch_saddle_list(const Grid& g, const uint8_t* types, uint32_t n, uint32_t* cnt, uint64_t* scan_st,
                              uint32_t* scan_ctr, uint32_t* err, uint32_t* list, hipStream_t st) {
  const uint32_t nb = saddle_blocks(n);
  hipLaunchKernelGGL(k_saddle_count, dim3(nb), dim3(kBlockThreads), 0, st, g, types, n, cnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // cnt[nb] is zero on entry, so that the scan leaves the total there
  if ((e = launch_scan_excl(cnt, nb + 1, scan_st, scan_ctr, err, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_saddle_list, dim3(nb), dim3(kBlockThreads), 0, st, g, types, n, cnt, list);
  return hipGetLastError();
}

hipError_t launch_saddle_pass(const Grid& g, const int32_t* q, const uint8_t* types, const uint32_t* list,
                              const uint32_t* count, const int8_t* cd_in, int8_t* cd_out, int8_t* upd,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_saddle_pass, dim3(4096), dim3(kBlockThreads), 0, st, g, q, types, list, count, cd_in, cd_out,
                     upd);
  return hipGetLastError();
}

hipError_t launch_saddle_apply(int32_t* q, const int8_t* upd, const uint32_t* list, const uint32_t* count,
                               uint32_t* moved, hipStream_t st) {
  hipLaunchKernelGGL(k_saddle_apply, dim3(4096), dim3(kBlockThreads), 0, st, q, upd, list, count, moved);
  return hipGetLastError();
}

}  // namespace tszp
