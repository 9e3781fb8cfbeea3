// tszp_saddle.hip -- quantization and saddle fix of the double-precision 3D topology
// path (SZp/src/szp_cuda_topology_3d.cu:1065-1188):
//   types  <- critical-point list (mark_critical_types_3d_kernel, :602-618)
//   q      <- C cast of x * (1 / eb) in double (quantize_3d_kernel_double, :641-646)
//   three Jacobi passes of the saddle fix (fix_saddles_3d_kernel, :736-826)
// The reference snapshots the whole quantized field before each pass (a device copy of
// 4 bytes per point).  Here q is never written during the passes: each pass reads the
// moves of the previous pass from a byte map (only saddles move, by at most one step per
// pass) and writes its own map, so a pass costs the type map plus two bytes per point;
// the moves are added to q once at the end.
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

// four points per thread: one 4-byte load of types, one 4-byte store of moves
__global__ __launch_bounds__(kBlockThreads) void k_saddle_pass(Grid g, const int32_t* __restrict__ q,
                                                              const uint8_t* __restrict__ types,
                                                              const int8_t* __restrict__ cd_in,
                                                              int8_t* __restrict__ cd_out, uint32_t n) {
  const uint32_t groups = (n + 3u) / 4u;
  uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  const uint32_t stride = gridDim.x * kBlockThreads;
  for (; i < groups; i += stride) {
    const uint32_t e0 = 4u * i;
    const bool full = e0 + 4u <= n;
    uint32_t tw = 0;
    if (full) {
      tw = reinterpret_cast<const uint32_t*>(types)[i];
    } else {
      for (uint32_t r = 0; e0 + r < n; r++) tw |= (uint32_t)types[e0 + r] << (8 * r);
    }
    uint32_t out = 0;
    if (tw & 0x02020202u) {  // some type is 2 or 3
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (((tw >> (8 * r)) & 0xffu) != 3u) continue;
        const uint32_t e = e0 + r;
        const uint32_t x = fdiv(e, g.div_plane), rr = e - x * g.plane;
        const uint32_t y = fdiv(rr, g.div_row), z = rr - y * g.d2;
        if (x == 0 || y == 0 || z == 0 || x + 1 >= g.d0 || y + 1 >= g.d1 || z + 1 >= g.d2) continue;
        const int mv = saddle_move(g, q, types, cd_in, e);
        out |= (uint32_t)(uint8_t)(int8_t)mv << (8 * r);
      }
    }
    if (full) {
      reinterpret_cast<uint32_t*>(cd_out)[i] = out;
    } else {
      for (uint32_t r = 0; e0 + r < n; r++) cd_out[e0 + r] = (int8_t)(uint8_t)(out >> (8 * r));
    }
  }
}

__global__ __launch_bounds__(kBlockThreads) void k_saddle_apply(int32_t* __restrict__ q,
                                                               const int8_t* __restrict__ cd, uint32_t n,
                                                               uint32_t* moved) {
  const uint32_t groups = (n + 3u) / 4u;
  uint32_t i = blockIdx.x * kBlockThreads + threadIdx.x;
  const uint32_t stride = gridDim.x * kBlockThreads;
  uint32_t cnt = 0;
  for (; i < groups; i += stride) {
    const uint32_t e0 = 4u * i;
    uint32_t w = 0;
    if (e0 + 4u <= n) {
      w = reinterpret_cast<const uint32_t*>(cd)[i];
    } else {
      for (uint32_t r = 0; e0 + r < n; r++) w |= (uint32_t)(uint8_t)cd[e0 + r] << (8 * r);
    }
    if (!w) continue;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int d = (int)(int8_t)(uint8_t)(w >> (8 * r));
      if (d) {
        q[e0 + r] = (int)((uint32_t)q[e0 + r] + (uint32_t)d);
        cnt++;
      }
    }
  }
  cnt = seg_sum<kWave>(cnt);
  if ((threadIdx.x & (kWave - 1)) == 0 && cnt) atomicAdd(moved, cnt);
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

hipError_t launch_saddle_pass(const Grid& g, const int32_t* q, const uint8_t* types, const int8_t* cd_in,
                              int8_t* cd_out, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_saddle_pass, dim3(grid_for((n + 3u) / 4u, 8192)), dim3(kBlockThreads), 0, st, g, q, types,
                     cd_in, cd_out, n);
  return hipGetLastError();
}

hipError_t launch_saddle_apply(int32_t* q, const int8_t* cd, uint32_t n, uint32_t* moved, hipStream_t st) {
  hipLaunchKernelGGL(k_saddle_apply, dim3(grid_for((n + 3u) / 4u, 8192)), dim3(kBlockThreads), 0, st, q, cd, n,
                     moved);
  return hipGetLastError();
}

}  // namespace tszp
