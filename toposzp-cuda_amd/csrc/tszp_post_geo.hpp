// tszp_post_geo.hpp -- geometry shared by the restoration kernels (tszp_post.hip,
// tszp_platmc.hip): coordinates, interior tests, classification, the field of a flat
// index for one field or a batch, the interval of a decoded value.
#pragma once
#include <hip/hip_runtime.h>

#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

// ---------------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------------
template <int ND>
__device__ __forceinline__ void coords_of(const Grid& g, uint32_t e, int (&c)[ND]) {
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
__device__ __forceinline__ int dim_of(const Grid& g, int a) {
  return a == 0 ? (int)g.d0 : (a == 1 ? (int)g.d1 : (int)g.d2);
}

template <int ND>
__device__ __forceinline__ uint32_t stride_of(const Grid& g, int a) {
  if constexpr (ND == 2) return a == 0 ? g.d1 : 1u;
  else return a == 0 ? g.plane : (a == 1 ? g.d2 : 1u);
}

template <int ND>
__device__ __forceinline__ bool interior(const Grid& g, const int (&c)[ND]) {
  bool r = true;
#pragma unroll
  for (int a = 0; a < ND; a++) r = r && c[a] >= 1 && c[a] + 2 <= dim_of<ND>(g, a);
  return r;
}

template <int ND>
__device__ __forceinline__ bool inside(const Grid& g, const int (&c)[ND]) {
  bool r = true;
#pragma unroll
  for (int a = 0; a < ND; a++) r = r && c[a] >= 0 && c[a] < dim_of<ND>(g, a);
  return r;
}

template <int ND>
__device__ __forceinline__ int colour_of(const int (&c)[ND]) {
  if constexpr (ND == 2) return (c[0] + 2 * c[1]) % 5;
  else return (c[0] + 2 * c[1] + 3 * c[2]) % 7;
}

// classification from the centre value and its 2*ND neighbours in axis order
// (axis a: v[2a] at -1, v[2a+1] at +1).  2D: max / min over the 4 neighbours, saddle
// when one axis is higher on both sides and the other lower (example/
// test_cuda_topology.cu:56-68); 3D: saddle when at least one axis is higher and one
// lower (SZp/src/szp_topology_3d.cc:79-98) -- the same rule for two axes.
template <int ND>
__device__ __forceinline__ int classify_vals(float c, const float (&v)[2 * ND]) {
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
__device__ __forceinline__ int classify_at(const Grid& g, const float* __restrict__ y, uint32_t e,
                                           const int (&c)[ND]) {
  if (!interior<ND>(g, c)) return 0;
  float v[2 * ND];
#pragma unroll
  for (int a = 0; a < ND; a++) {
    const uint32_t s = stride_of<ND>(g, a);
    v[2 * a] = y[e - s];
    v[2 * a + 1] = y[e + s];
  }
  return classify_vals<ND>(y[e], v);
}

// Field accesses of the window solver.  COH: agent-coherent loads and stores (they
// bypass the XCD-local caches), for the one-launch solve where neighbouring tiles run
// on other XCDs of the same launch.
template <bool COH>
__device__ __forceinline__ float y_ld(const float* y, uint32_t e) {
  if constexpr (COH) return __hip_atomic_load(const_cast<float*>(y) + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return y[e];
}
template <bool COH>
__device__ __forceinline__ void y_st(float* y, uint32_t e, float v) {
  if constexpr (COH) __hip_atomic_store(y + e, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else y[e] = v;
}
template <int ND, bool COH>
__device__ __forceinline__ int classify_at_y(const Grid& g, const float* y, uint32_t e, const int (&c)[ND]) {
  if (!interior<ND>(g, c)) return 0;
  float v[2 * ND];
#pragma unroll
  for (int a = 0; a < ND; a++) {
    const uint32_t s = stride_of<ND>(g, a);
    v[2 * a] = y_ld<COH>(y, e - s);
    v[2 * a + 1] = y_ld<COH>(y, e + s);
  }
  return classify_vals<ND>(y_ld<COH>(y, e), v);
}

// ---------------------------------------------------------------------------------
// geometry source: the field of a flat index (coordinates, interval width, tiles)
// ---------------------------------------------------------------------------------
struct FieldView {
  Grid g;
  uint32_t base, n, fid;
  float eb;
  uint32_t tile0, tx, ty, tz;
};
// one field at flat offset 0
struct Geo1 {
  Grid g;
  float eb;
  uint32_t n, tx, ty, tz;
  static constexpr bool kBatch = false;
  __device__ __forceinline__ FieldView at(uint32_t) const {
    FieldView v;
    v.g = g;
    v.base = 0;
    v.n = n;
    v.fid = 0;
    v.eb = eb;
    v.tile0 = 0;
    v.tx = tx;
    v.ty = ty;
    v.tz = tz;
    return v;
  }
  __device__ __forceinline__ FieldView of_tile(uint32_t) const { return at(0); }
  __device__ __forceinline__ uint32_t spmax(const FieldView&, const uint32_t* misc) const { return misc[kPpSpMax]; }
};
// a batch of fields in flat arrays (PpField)
struct GeoB {
  const PpField* f;
  const uint32_t* fmap;
  const uint32_t* tmap;
  const uint32_t* spm;
  static constexpr bool kBatch = true;
  __device__ __forceinline__ FieldView view(uint32_t i) const {
    const PpField& p = f[i];
    FieldView v;
    v.g = p.g;
    v.base = p.base;
    v.n = p.n;
    v.fid = i;
    v.eb = p.eb;
    v.tile0 = p.tile0;
    v.tx = p.tx;
    v.ty = p.ty;
    v.tz = p.tz;
    return v;
  }
  __device__ __forceinline__ FieldView at(uint32_t e) const { return view(fmap[e >> kFlatShift]); }
  __device__ __forceinline__ FieldView of_tile(uint32_t tile) const { return view(tmap[tile]); }
  __device__ __forceinline__ uint32_t spmax(const FieldView& v, const uint32_t*) const { return spm[v.fid]; }
};

// interval of the restored value of a point with decoded value d.  The zero cell is
// the one cell whose points may leave the error bound (by at most eb kZeroCell); its
// half-width only has to leave room for the orders inside a plateau, which float
// spacing near zero gives many times over.
constexpr float kZeroCell = 1.f / 65536.f;
__device__ __forceinline__ void cell_of(float d, float eb, float* lo, float* hi) {
  if (d > 0.f) {
    *lo = d;
    *hi = d + 0.9f * eb;
  } else if (d < 0.f) {
    *lo = d - 0.9f * eb;
    *hi = d;
  } else {
    *lo = -eb * kZeroCell;
    *hi = eb * kZeroCell;
  }
}

__device__ __forceinline__ bool anchored(int t, float d) { return (t == 1 || t == 2) && d != 0.f; }

__device__ __forceinline__ bool claim_bit(uint32_t* bits, uint32_t e) {
  const uint32_t b = 1u << (e & 31u);
  return (atomicOr(&bits[e >> 5], b) & b) == 0u;
}

// ---- host: geometry arguments of the kernels ----
inline Geo1 geo1(const PpGeo& G) {
  Geo1 g;
  g.g = G.g;
  g.eb = G.eb;
  g.n = G.n;
  g.tx = G.tx;
  g.ty = G.ty;
  g.tz = G.tz;
  return g;
}
inline GeoB geob(const PpGeo& G) {
  GeoB b;
  b.f = G.fields;
  b.fmap = G.fmap;
  b.tmap = G.tmap;
  b.spm = G.spmax;
  return b;
}

// launches kernel template K<ND, GEO> with the geometry of `geo` as first argument
#define PP_GEO_LAUNCH(K, grid, blk, st, ...)                                                    \
  do {                                                                                          \
    if (geo.fields) {                                                                           \
      const GeoB gb_ = geob(geo);                                                               \
      if (geo.ndim == 2) hipLaunchKernelGGL((K<2, GeoB>), grid, blk, 0, st, gb_, __VA_ARGS__);  \
      else hipLaunchKernelGGL((K<3, GeoB>), grid, blk, 0, st, gb_, __VA_ARGS__);                \
    } else {                                                                                    \
      const Geo1 g1_ = geo1(geo);                                                               \
      if (geo.ndim == 2) hipLaunchKernelGGL((K<2, Geo1>), grid, blk, 0, st, g1_, __VA_ARGS__);  \
      else hipLaunchKernelGGL((K<3, Geo1>), grid, blk, 0, st, g1_, __VA_ARGS__);                \
    }                                                                                           \
  } while (0)

}  // namespace tszp
