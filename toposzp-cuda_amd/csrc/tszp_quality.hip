// tszp_quality.hip -- topology quality of a decompressed field against the original,
// on the GPU: the metrics the reference's CUDA driver prints
// (example/test_cuda_topology.cu:540-868), for 2D and 3D fields.
//   * critical points of both fields (max / min / saddle rule of the codec), preserved,
//     lost (became regular), changed type, new;
//   * pointwise error: maximum, points above 1.01 eb and above 2 eb;
//   * saddle persistence proxy: min |centre - neighbour| over the stencil, and how many
//     saddles above 0, 0.5, 1, 2, 5, 10 eb keep their type;
//   * local order: adjacent pairs of interior points (forward neighbours along each
//     axis) whose original order is strict, and how many of them change order -- all
//     pairs, pairs with a critical point, pairs of critical points, pairs with an
//     extremum.
// One grid-stride pass; per-workgroup partials (no same-address atomics) added up by a
// one-workgroup kernel.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

constexpr double kPersThr[kQualThresholds] = {0.0, 0.5, 1.0, 2.0, 5.0, 10.0};

template <int ND>
__device__ __forceinline__ void qcoords(const Grid& g, uint32_t e, int (&c)[ND]) {
  if constexpr (ND == 2) {
    const uint32_t x = fdiv(e, g.div_row);
    c[0] = (int)x;
    c[1] = (int)(e - x * g.d1);
  } else {
    const uint32_t x = fdiv(e, g.div_plane), r = e - x * g.plane;
    const uint32_t y = fdiv(r, g.div_row);
    c[0] = (int)x;
    c[1] = (int)y;
    c[2] = (int)(r - y * g.d2);
  }
}

template <int ND>
__device__ __forceinline__ int qdim(const Grid& g, int a) {
  return a == 0 ? (int)g.d0 : (a == 1 ? (int)g.d1 : (int)g.d2);
}

template <int ND>
__device__ __forceinline__ uint32_t qstride(const Grid& g, int a) {
  if constexpr (ND == 2) return a == 0 ? g.d1 : 1u;
  else return a == 0 ? g.plane : (a == 1 ? g.d2 : 1u);
}

// type of an interior point from its centre and its 2*ND neighbours (axis a: v[2a] at
// -1, v[2a+1] at +1): max above all, min below all, saddle when one axis is higher on
// both sides and another lower on both sides
template <int ND>
__device__ __forceinline__ int qtype(float c, const float (&v)[2 * ND]) {
  bool all_gt = true, all_lt = true, any_hi = false, any_lo = false;
#pragma unroll
  for (int a = 0; a < ND; a++) {
    const bool hi = c > v[2 * a] && c > v[2 * a + 1];
    const bool lo = c < v[2 * a] && c < v[2 * a + 1];
    all_gt = all_gt && hi;
    all_lt = all_lt && lo;
    any_hi = any_hi || hi;
    any_lo = any_lo || lo;
  }
  return all_gt ? 1 : (all_lt ? 2 : ((any_hi && any_lo) ? 3 : 0));
}

template <int ND>
__device__ __forceinline__ bool qinterior(const Grid& g, const int (&c)[ND]) {
  bool r = true;
#pragma unroll
  for (int a = 0; a < ND; a++) r = r && c[a] >= 1 && c[a] + 1 < qdim<ND>(g, a);
  return r;
}

// type of point e in field f (0 for boundary points)
template <int ND>
__device__ __forceinline__ int qtype_at(const Grid& g, const float* __restrict__ f, uint32_t e, const int (&c)[ND],
                                        bool inter) {
  if (!inter) return 0;
  float v[2 * ND];
#pragma unroll
  for (int a = 0; a < ND; a++) {
    const uint32_t s = qstride<ND>(g, a);
    v[2 * a] = f[e - s];
    v[2 * a + 1] = f[e + s];
  }
  return qtype<ND>(f[e], v);
}

template <int ND>
__global__ __launch_bounds__(kBlockThreads) void k_quality(Grid g, const float* __restrict__ x,
                                                          const float* __restrict__ y, uint32_t n, float eb,
                                                          uint64_t* __restrict__ part) {
  uint32_t cnt[kQualCounts];
#pragma unroll
  for (int k = 0; k < kQualCounts; k++) cnt[k] = 0;
  double emax = 0.0, pmin = 1e300, pmax = 0.0, psum = 0.0;
  const double e1 = (double)eb * 1.01, e2 = (double)eb * 2.0;
  for (uint32_t e = blockIdx.x * kBlockThreads + threadIdx.x; e < n; e += gridDim.x * kBlockThreads) {
    int c[ND];
    qcoords<ND>(g, e, c);
    const bool inter = qinterior<ND>(g, c);
    const int to = qtype_at<ND>(g, x, e, c, inter), td = qtype_at<ND>(g, y, e, c, inter);
    cnt[kQOrig] += to != 0;
    cnt[kQOrig + 1] += to == 1;
    cnt[kQOrig + 2] += to == 2;
    cnt[kQOrig + 3] += to == 3;
    cnt[kQDec] += td != 0;
    cnt[kQDec + 1] += td == 1;
    cnt[kQDec + 2] += td == 2;
    cnt[kQDec + 3] += td == 3;
    if (to) {
      cnt[kQPreserved] += td == to;
      cnt[kQLost] += td == 0;
      cnt[kQChanged] += td != 0 && td != to;
      // per type, with constant indices (a runtime index would put the counters in
      // scratch memory)
#pragma unroll
      for (int k = 1; k <= 3; k++) {
        cnt[kQLost + k] += to == k && td == 0;
        cnt[kQChanged + k] += to == k && td != 0 && td != to;
      }
    } else {
      cnt[kQNew] += td != 0;
    }
    const float xe = x[e];
    const double err = fabs((double)xe - (double)y[e]);
    emax = err > emax ? err : emax;
    cnt[kQOver1] += err > e1;
    cnt[kQOver2] += err > e2;
    if (to == 3) {  // saddle persistence proxy
      double p = 1e300;
#pragma unroll
      for (int a = 0; a < ND; a++) {
        const uint32_t s = qstride<ND>(g, a);
        p = fmin(p, fmin(fabs((double)xe - x[e - s]), fabs((double)xe - x[e + s])));
      }
      cnt[kQSaddles]++;
      pmin = fmin(pmin, p);
      pmax = fmax(pmax, p);
      psum += p;
#pragma unroll
      for (int k = 0; k < kQualThresholds; k++) {
        const bool above = p > kPersThr[k] * (double)eb;
        cnt[kQAbove + k] += above;
        cnt[kQAboveKept + k] += above && td == 3;
      }
    }
    if (inter) {  // local order with the forward neighbour along each axis
      const float ye = y[e];
#pragma unroll
      for (int a = 0; a < ND; a++) {
        if (c[a] + 2 >= qdim<ND>(g, a)) continue;  // the neighbour must be interior too
        const uint32_t q = e + qstride<ND>(g, a);
        const float xq = x[q];
        if (xe == xq) continue;
        const float yq = y[q];
        const int oo = xe < xq ? -1 : 1;
        const int od = ye < yq ? -1 : (ye > yq ? 1 : 0);
        const bool v = oo != od;
        int cq[ND];
#pragma unroll
        for (int b = 0; b < ND; b++) cq[b] = c[b] + (b == a);
        const int tq = qtype_at<ND>(g, x, q, cq, true);
        const bool cp1 = to || tq, cp2 = to && tq, ex = to == 1 || to == 2 || tq == 1 || tq == 2;
        cnt[kQPairs] += 1;
        cnt[kQViol] += v;
        cnt[kQPairs + 1] += cp1;
        cnt[kQViol + 1] += cp1 && v;
        cnt[kQPairs + 2] += cp2;
        cnt[kQViol + 2] += cp2 && v;
        cnt[kQPairs + 3] += ex;
        cnt[kQViol + 3] += ex && v;
      }
    }
  }
  // workgroup partials: counts summed, error maximum, persistence min / max / sum
  __shared__ uint64_t red[kQualCounts][kWavesPerBlock];
  __shared__ double dred[4][kWavesPerBlock];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int k = 0; k < kQualCounts; k++) {
    const uint32_t s = seg_sum<kWave>(cnt[k]);
    if (lane == 0) red[k][w] = s;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    emax = fmax(emax, __shfl_xor(emax, o, kWave));
    pmin = fmin(pmin, __shfl_xor(pmin, o, kWave));
    pmax = fmax(pmax, __shfl_xor(pmax, o, kWave));
    psum += __shfl_xor(psum, o, kWave);
  }
  if (lane == 0) {
    dred[0][w] = emax;
    dred[1][w] = pmin;
    dred[2][w] = pmax;
    dred[3][w] = psum;
  }
  __syncthreads();
  uint64_t* out = part + (size_t)blockIdx.x * kQualWords;
  if (threadIdx.x < kQualCounts) {
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) s += red[threadIdx.x][k];
    out[threadIdx.x] = s;
  }
  if (threadIdx.x == 0) {
    double m0 = dred[0][0], m1 = dred[1][0], m2 = dred[2][0], m3 = 0.0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; k++) {
      m0 = fmax(m0, dred[0][k]);
      m1 = fmin(m1, dred[1][k]);
      m2 = fmax(m2, dred[2][k]);
      m3 += dred[3][k];
    }
    double* dout = reinterpret_cast<double*>(out + kQualCounts);
    dout[0] = m0;
    dout[1] = m1;
    dout[2] = m2;
    dout[3] = m3;
  }
}

// one workgroup per result word: counts summed, error maximum, persistence min / max / sum
__global__ __launch_bounds__(kBlockThreads) void k_quality_reduce(const uint64_t* __restrict__ part, uint32_t nb,
                                                                 uint64_t* __restrict__ out) {
  __shared__ uint64_t red[kBlockThreads];
  __shared__ double dred[kBlockThreads];
  const int k = blockIdx.x, tid = threadIdx.x;
  if (k < kQualCounts) {
    uint64_t s = 0;
    for (uint32_t b = tid; b < nb; b += kBlockThreads) s += part[(size_t)b * kQualWords + k];
    red[tid] = s;
  } else {
    const int j = k - kQualCounts;
    double v = j == 1 ? 1e300 : 0.0;
    for (uint32_t b = tid; b < nb; b += kBlockThreads) {
      const double u = reinterpret_cast<const double*>(part + (size_t)b * kQualWords + kQualCounts)[j];
      v = j == 0 || j == 2 ? fmax(v, u) : (j == 1 ? fmin(v, u) : v + u);
    }
    dred[tid] = v;
  }
  __syncthreads();
  for (int o = kBlockThreads / 2; o > 0; o >>= 1) {
    if (tid < o) {
      if (k < kQualCounts) {
        red[tid] += red[tid + o];
      } else {
        const int j = k - kQualCounts;
        const double a = dred[tid], b = dred[tid + o];
        dred[tid] = j == 0 || j == 2 ? fmax(a, b) : (j == 1 ? fmin(a, b) : a + b);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (k < kQualCounts) out[k] = red[0];
    else reinterpret_cast<double*>(out + kQualCounts)[k - kQualCounts] = dred[0];
  }
}

hipError_t launch_quality(const Grid& g, const float* x, const float* y, uint32_t n, float eb, uint64_t* part,
                          uint32_t nblocks, uint64_t* out, hipStream_t st) {
  if (g.ndim == 2)
    hipLaunchKernelGGL(k_quality<2>, dim3(nblocks), dim3(kBlockThreads), 0, st, g, x, y, n, eb, part);
  else
    hipLaunchKernelGGL(k_quality<3>, dim3(nblocks), dim3(kBlockThreads), 0, st, g, x, y, n, eb, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_quality_reduce, dim3(kQualCounts + 4), dim3(kBlockThreads), 0, st, part, nblocks, out);
  return hipGetLastError();
}

}  // namespace tszp
