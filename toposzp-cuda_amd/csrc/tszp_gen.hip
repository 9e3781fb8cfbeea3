// tszp_gen.hip -- synthetic input fields generated on the device from a seed.
//
// Counter-based hashing (SplitMix64) gives every element and every blob parameter
// its own reproducible random numbers, so a field depends only on (seed, shape).
//   blobs: sum of Gaussian blobs (anisotropic and rotated in 2D, isotropic in 3D),
//          optionally plus sin(kx x) sin(ky y) sin(kz z), plus white noise;
//   fbm:   6-octave value noise (lacunarity 2, gain 0.5) plus white noise.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(seed ^ mix64(a)) ^ b);
}
__device__ __forceinline__ double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
__device__ __forceinline__ float normal(uint64_t seed, uint64_t i) {
  const double u1 = u01(hash3(seed, i, 0x51u)) + 1e-300, u2 = u01(hash3(seed, i, 0x52u));
  return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

struct BlobParams {
  int ndim, nblobs, waves;
  uint32_t d0, d1, d2;
  uint64_t seed;
  float smin, smax, noise;
};

__global__ __launch_bounds__(kBlockThreads) void k_gen_blobs(float* f, BlobParams p) {
  const size_t n = (size_t)p.d0 * p.d1 * (p.ndim == 3 ? p.d2 : 1);
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x0, x1, x2 = 0.f;
  if (p.ndim == 2) {
    x0 = (float)(i / p.d1);
    x1 = (float)(i % p.d1);
  } else {
    const size_t pl = (size_t)p.d1 * p.d2;
    x0 = (float)(i / pl);
    x1 = (float)((i % pl) / p.d2);
    x2 = (float)(i % p.d2);
  }
  float acc = 0.f;
  for (int b = 0; b < p.nblobs; b++) {
    const float amp = (float)(2.0 * u01(hash3(p.seed, b, 1)) - 1.0);
    const float c0 = (float)(u01(hash3(p.seed, b, 2)) * p.d0);
    const float c1 = (float)(u01(hash3(p.seed, b, 3)) * p.d1);
    const float sa = (float)(p.smin + (p.smax - p.smin) * u01(hash3(p.seed, b, 5)));
    float q;
    if (p.ndim == 2) {
      const float sb = (float)(p.smin + (p.smax - p.smin) * u01(hash3(p.seed, b, 6)));
      const float th = (float)(3.141592653589793 * u01(hash3(p.seed, b, 7)));
      const float dy = x0 - c0, dx = x1 - c1;
      const float cs = cosf(th), sn = sinf(th);
      const float u = cs * dx + sn * dy, v = -sn * dx + cs * dy;
      q = (u * u) / (sa * sa) + (v * v) / (sb * sb);
    } else {
      const float c2 = (float)(u01(hash3(p.seed, b, 4)) * p.d2);
      const float d0 = x0 - c0, d1 = x1 - c1, d2 = x2 - c2;
      q = (d0 * d0 + d1 * d1 + d2 * d2) / (sa * sa);
    }
    acc += amp * __expf(-0.5f * q);
  }
  if (p.waves) {
    const float k0 = (float)(0.02 + 0.18 * u01(hash3(p.seed, 1000003, 1)));
    const float k1 = (float)(0.02 + 0.18 * u01(hash3(p.seed, 1000003, 2)));
    const float k2 = (float)(0.02 + 0.18 * u01(hash3(p.seed, 1000003, 3)));
    acc += sinf(k0 * x0) * sinf(k1 * x1) * (p.ndim == 3 ? sinf(k2 * x2) : 1.f);
  }
  f[i] = acc + p.noise * normal(p.seed ^ 0xabcdefull, i);
}

hipError_t launch_gen_blobs(float* f, int ndim, const uint32_t* dims, uint64_t seed, int nblobs, float smin,
                            float smax, float noise, int waves, hipStream_t st) {
  BlobParams p{ndim, nblobs, waves, dims[0], dims[1], ndim == 3 ? dims[2] : 1u, seed, smin, smax, noise};
  const size_t n = (size_t)p.d0 * p.d1 * (ndim == 3 ? p.d2 : 1);
  hipLaunchKernelGGL(k_gen_blobs, dim3((unsigned)((n + kBlockThreads - 1) / kBlockThreads)), dim3(kBlockThreads),
                     0, st, f, p);
  return hipGetLastError();
}

__device__ __forceinline__ float lattice(uint64_t seed, int o, int64_t a, int64_t b, int64_t c) {
  const uint64_t h = hash3(seed + (uint64_t)o * 0x1000193ull, (uint64_t)a * 73856093ull ^ (uint64_t)b * 19349663ull,
                           (uint64_t)c * 83492791ull);
  return (float)(2.0 * u01(h) - 1.0);
}

__global__ __launch_bounds__(kBlockThreads) void k_gen_fbm(float* f, uint32_t d0, uint32_t d1, uint32_t d2,
                                                          uint64_t seed, int octaves, float noise) {
  const size_t n = (size_t)d0 * d1 * d2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t pl = (size_t)d1 * d2;
  const float x = (float)(i / pl), y = (float)((i % pl) / d2), z = (float)(i % d2);
  float acc = 0.f, amp = 1.f, freq = 1.f / 64.f;
  for (int o = 0; o < octaves; o++) {
    const float X = x * freq, Y = y * freq, Z = z * freq;
    const float fx = floorf(X), fy = floorf(Y), fz = floorf(Z);
    const float tx = X - fx, ty = Y - fy, tz = Z - fz;
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    float v = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int a = c >> 2, b = (c >> 1) & 1, e = c & 1;
      const float wgt = (a ? tx : 1.f - tx) * (b ? ty : 1.f - ty) * (e ? tz : 1.f - tz);
      v += wgt * lattice(seed, o, ix + a, iy + b, iz + e);
    }
    acc += amp * v;
    amp *= 0.5f;
    freq *= 2.f;
  }
  f[i] = acc + noise * normal(seed ^ 0x5eedull, i);
}

hipError_t launch_gen_fbm(float* f, const uint32_t* dims, uint64_t seed, int octaves, float noise, hipStream_t st) {
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  hipLaunchKernelGGL(k_gen_fbm, dim3((unsigned)((n + kBlockThreads - 1) / kBlockThreads)), dim3(kBlockThreads), 0,
                     st, f, dims[0], dims[1], dims[2], seed, octaves, noise);
  return hipGetLastError();
}

}  // namespace tszp
