// tszp_util.hip -- small kernels: fills, min/max for relative error bounds, and the
// chunk-offset headers of the three stream kinds.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

// one range: words up to 16-byte alignment, then 16-byte stores, then the tail
__device__ __forceinline__ void fill_range(uint32_t* p, uint32_t v, size_t count, size_t tid, size_t stride) {
  size_t head = ((16u - (reinterpret_cast<uintptr_t>(p) & 15u)) & 15u) / 4u;  // p is 4-byte aligned
  if (head > count) head = count;
  if (tid < head) p[tid] = v;
  uint4* p4 = reinterpret_cast<uint4*>(p + head);
  const size_t n4 = (count - head) / 4;
  const uint4 v4 = make_uint4(v, v, v, v);
  for (size_t i = tid; i < n4; i += stride) p4[i] = v4;
  for (size_t i = head + 4 * n4 + tid; i < count; i += stride) p[i] = v;
}

// (a few hundred workgroups at most: a fill is store-bound, and empty workgroups of a
// large grid still cost their dispatch)
static inline unsigned fill_blocks(size_t count) {
  const size_t b = (count + 4 * kBlockThreads - 1) / (4 * kBlockThreads);
  return (unsigned)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t count) {
  fill_range(p, v, count, (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t count, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(count)), dim3(kBlockThreads), 0, st, p, v, count);
  return hipGetLastError();
}

// several ranges in one launch: blockIdx.y picks the range
__global__ void k_fill_multi(FillRanges r) {
  const uint32_t k = blockIdx.y;
  fill_range(r.p[k], r.v[k], r.count[k], (size_t)blockIdx.x * blockDim.x + threadIdx.x,
             (size_t)gridDim.x * blockDim.x);
}

hipError_t launch_fill_multi(const FillRanges& r, hipStream_t st) {
  if (r.n < 1 || r.n > kFillRanges) return hipErrorInvalidValue;
  size_t most = 0;
  for (int k = 0; k < r.n; k++) most = r.count[k] > most ? r.count[k] : most;
  if (!most) return hipSuccess;
  hipLaunchKernelGGL(k_fill_multi, dim3(fill_blocks(most), (unsigned)r.n), dim3(kBlockThreads), 0, st, r);
  return hipGetLastError();
}

// min / max as order-preserving keys (out2 must hold {0xffffffff, 0} on entry); 16-byte
// loads when f is aligned, one atomic pair per workgroup (a few hundred workgroups:
// atomics on one word serialise)
__global__ __launch_bounds__(kBlockThreads) void k_minmax(const float* __restrict__ f, size_t n, uint32_t* out2) {
  __shared__ uint32_t rlo[kWavesPerBlock], rhi[kWavesPerBlock];
  uint32_t lo = 0xffffffffu, hi = 0;
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t head = 0;
  if ((reinterpret_cast<uintptr_t>(f) & 15u) == 0) {
    const float4* f4 = reinterpret_cast<const float4*>(f);
    const size_t n4 = n / 4;
    for (size_t i = tid; i < n4; i += stride) {
      const float4 v = f4[i];
      const uint32_t k0 = okey(v.x), k1 = okey(v.y), k2 = okey(v.z), k3 = okey(v.w);
      lo = min(lo, min(min(k0, k1), min(k2, k3)));
      hi = max(hi, max(max(k0, k1), max(k2, k3)));
    }
    head = 4 * n4;
  }
  for (size_t i = head + tid; i < n; i += stride) {
    const uint32_t k = okey(f[i]);
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t a = __shfl_xor(lo, o, kWave), b = __shfl_xor(hi, o, kWave);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    rlo[threadIdx.x / kWave] = lo;
    rhi[threadIdx.x / kWave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 1; k < kWavesPerBlock; k++) {
      lo = rlo[k] < lo ? rlo[k] : lo;
      hi = rhi[k] > hi ? rhi[k] : hi;
    }
    atomicMin(&out2[0], lo);
    atomicMax(&out2[1], hi);
  }
}

hipError_t launch_minmax(const float* f, size_t n, uint32_t* out2, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, st, out2, 0xffffffffu, (size_t)1);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, st, out2 + 1, 0u, (size_t)1);
  size_t blocks = (n + kBlockThreads * 16 - 1) / (kBlockThreads * 16);
  if (blocks > 512) blocks = 512;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_minmax, dim3((unsigned)blocks), dim3(kBlockThreads), 0, st, f, n, out2);
  return hipGetLastError();
}

// Chunk partition of the reference's OpenMP writers (SZp/src/szp_float.cc:1148-1176).
__device__ __forceinline__ void chunk_range(uint32_t nb, unsigned T, int adaptive, unsigned t, uint32_t* b0,
                                            uint32_t* b1) {
  uint64_t bpt, s, e;
  if (adaptive && T >= 16 && nb < (uint64_t)T * 8) {
    uint64_t opt = (nb + 7) / 8;
    if (opt < 1) opt = 1;
    bpt = (nb + opt - 1) / opt;
    const uint64_t et = t % opt;
    s = et * bpt;
    e = (et + 1) * bpt;
    if (e > nb) e = nb;
    if (t >= opt) s = e;
  } else {
    bpt = (nb + T - 1) / T;
    s = (uint64_t)t * bpt;
    e = ((uint64_t)t + 1) * bpt;
    if (e > nb) e = nb;
  }
  if (s > e) s = e;
  *b0 = (uint32_t)s;
  *b1 = (uint32_t)e;
}

// kind 0: topology [u64 off[T]]; 1: sort positions [u64 count][u64 off[T]];
// 2: plain [f32 eb big-endian][u64 off[T]]
__global__ void k_header(uint8_t* stream, int kind, unsigned T, int adaptive, uint32_t nb, const uint32_t* blk,
                         uint64_t count, float eb) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  uint8_t* offs = stream + (kind == 1 ? 8 : kind == 2 ? 4 : 0);
  if (t == 0) {
    if (kind == 1) {
      for (int k = 0; k < 8; k++) stream[k] = (uint8_t)(count >> (8 * k));
    } else if (kind == 2) {
      const uint32_t u = __float_as_uint(eb);
      stream[0] = (uint8_t)(u >> 24);
      stream[1] = (uint8_t)(u >> 16);
      stream[2] = (uint8_t)(u >> 8);
      stream[3] = (uint8_t)u;
    }
  }
  if (t >= T) return;
  uint32_t b0, b1;
  chunk_range(nb, T, adaptive, t, &b0, &b1);
  // chunk t starts after every block of the chunks before it
  const uint64_t off = b0 < b1 ? blk[b0] : blk[nb];
  for (int k = 0; k < 8; k++) offs[8 * t + k] = (uint8_t)(off >> (8 * k));
}

hipError_t launch_write_header(uint8_t* stream, int kind, unsigned T, int adaptive, uint32_t nb,
                               const uint32_t* blk, uint64_t count, float eb, hipStream_t st) {
  const unsigned threads = T < 64 ? 64 : 256;
  hipLaunchKernelGGL(k_header, dim3((T + threads - 1) / threads), dim3(threads), 0, st, stream, kind, T,
                     adaptive, nb, blk, count, eb);
  return hipGetLastError();
}

}  // namespace tszp
