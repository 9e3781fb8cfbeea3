// tszp_topo_batch.hip -- kernels of the batched topology pipeline that have no one-field
// counterpart: the per-field critical-point counts turned into offsets of one
// contiguous key set, the gather of every field's critical-point keys into it (with the
// field as the high key word, so that one radix sort orders the whole batch by
// (field, value) and one segmented rank gives every field its sort positions), and the
// sort-position stream headers, whose extremum counts are known on the device only.
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

// One workgroup: exclusive scan of the critical-point counts of the fields in table
// order (res[i].cp_count) into cbase[i], the total into *total; the sort-position
// stream header [u64 E][u64 0] of every field (also for E == 0, so that a decoder can
// always read E from the stream head).
__global__ __launch_bounds__(kBlockThreads) void k_tb_offsets(const TopoBatchDev* __restrict__ tab, uint32_t count,
                                                             const CompressResult* __restrict__ res,
                                                             uint32_t* __restrict__ cbase, uint32_t* __restrict__ total,
                                                             uint32_t* __restrict__ err) {
  __shared__ uint32_t red[kWavesPerBlock];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < count; b0 += kBlockThreads) {
    const uint32_t i = b0 + threadIdx.x;
    uint32_t c = 0;
    if (i < count) {
      const CompressResult& r = res[i];
      c = (uint32_t)r.cp_count;
      uint32_t* h = reinterpret_cast<uint32_t*>(tab[i].sp_stream);
      h[0] = (uint32_t)r.ext_count;
      h[1] = 0u;
      h[2] = 0u;
      h[3] = 0u;
      if (r.cp_count > tab[i].n) atomicOr(err, TSZP_ERR_OVERFLOW);
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan(c, red, &tot);
    if (i < count) cbase[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// One workgroup per (field, chunk of kSortTile key slots): key i < cp_count of field f
// goes to position cbase[f] + i with high word f and payload base_f + extremum ordinal
// (saddles keep 0xffffffff), so that the rank lands in the flat sort-position array.
__global__ __launch_bounds__(kBlockThreads) void k_tb_gather(const TopoBatchDev* __restrict__ tab,
                                                            const uint32_t* __restrict__ chunk_field,
                                                            const CompressResult* __restrict__ res,
                                                            const uint32_t* __restrict__ cbase,
                                                            const uint32_t* __restrict__ cp_key,
                                                            const uint32_t* __restrict__ cp_pay,
                                                            uint32_t* __restrict__ keys, uint32_t* __restrict__ hi,
                                                            uint32_t* __restrict__ pay) {
  const uint32_t f = chunk_field[blockIdx.x];
  const TopoBatchDev& t = tab[f];
  const uint32_t c0 = (blockIdx.x - t.chunk0) * kSortTile;
  const uint32_t cnt = (uint32_t)res[f].cp_count;
  if (c0 >= cnt) return;
  const uint32_t end = min(cnt, c0 + kSortTile);
  const uint32_t dst = cbase[f];
  for (uint32_t i = c0 + threadIdx.x; i < end; i += kBlockThreads) {
    const uint32_t p = cp_pay[t.base + i];
    keys[dst + i] = cp_key[t.base + i];
    hi[dst + i] = f;
    pay[dst + i] = p == 0xffffffffu ? p : (uint32_t)t.base + p;
  }
}

__global__ __launch_bounds__(kBlockThreads) void k_tb_chunkmap(const TopoBatchDev* __restrict__ tab,
                                                              uint32_t* __restrict__ chunk_field) {
  const uint32_t f = blockIdx.x;
  const uint32_t c0 = tab[f].chunk0, nc = tab[f].nchunks;
  for (uint32_t c = threadIdx.x; c < nc; c += kBlockThreads) chunk_field[c0 + c] = f;
}

hipError_t launch_tb_offsets(const TopoBatchDev* tab, uint32_t count, const CompressResult* res, uint32_t* cbase,
                             uint32_t* total, uint32_t* err, hipStream_t st) {
  hipLaunchKernelGGL(k_tb_offsets, dim3(1), dim3(kBlockThreads), 0, st, tab, count, res, cbase, total, err);
  return hipGetLastError();
}

hipError_t launch_tb_chunkmap(const TopoBatchDev* tab, uint32_t count, uint32_t* chunk_field, hipStream_t st) {
  hipLaunchKernelGGL(k_tb_chunkmap, dim3(count), dim3(kBlockThreads), 0, st, tab, chunk_field);
  return hipGetLastError();
}

hipError_t launch_tb_gather(const TopoBatchDev* tab, const uint32_t* chunk_field, uint32_t nchunks,
                            const CompressResult* res, const uint32_t* cbase, const uint32_t* cp_key,
                            const uint32_t* cp_pay, uint32_t* keys, uint32_t* hi, uint32_t* pay, hipStream_t st) {
  if (!nchunks) return hipSuccess;
  hipLaunchKernelGGL(k_tb_gather, dim3(nchunks), dim3(kBlockThreads), 0, st, tab, chunk_field, res, cbase, cp_key,
                     cp_pay, keys, hi, pay);
  return hipGetLastError();
}

}  // namespace tszp
