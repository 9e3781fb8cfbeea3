// tszp_parse.hip -- block offsets of a stream that arrives without a side-car.
//
// A block's byte size follows from its element count and its bit-count byte (the
// fifth byte of the block), so finding the blocks of a stream is a sequential walk
// (SZp/src/szp_cuda_decompress.cu:391-477 does it on the host).  Here it runs in
// parallel as a composition of per-segment transition tables:
//   1. the payload is cut into segments of kParseSeg bytes.  For every possible entry
//      offset e into a segment (0 <= e < largest block size) a workgroup finds where
//      the walk that starts at e leaves the segment and how many blocks it passes, by
//      one walk per entry over the segment's bytes in LDS (k_parse_seg_walk) or, when
//      those walks would be long, by pointer doubling over all byte positions
//      (k_parse_seg); blocks are taken as full blocks here -- only the last block of
//      the field is shorter, and nothing after it matters;
//   2. tables of consecutive segments are composed in groups (k_parse_compose), twice,
//      so that a single thread can walk the top level;
//   3. the entry offset and the first block index of every segment come back down the
//      levels (k_parse_down);
//   4. one thread per segment walks its blocks from its entry, with the true size of
//      the last block, checks every bit count, and writes the block offsets
//      (k_parse_final).
// The chunk table at the head of the stream is checked against the parsed block starts
// (k_parse_check_offsets).
#include "tszp_device.hpp"
#include "tszp_kernels.h"

namespace tszp {

static_assert(kParseSeg == 8192, "segment geometry");

// ---- 1. per-segment transition tables -------------------------------------------------
// tab[seg * maxb + e] = {exit offset into the next segment, blocks passed}
__global__ __launch_bounds__(kBlockThreads) void k_parse_seg(ParseArgs a, uint2* __restrict__ tab) {
  __shared__ uint8_t bytes[kParseSeg + kParseMaxBlock];
  __shared__ uint16_t J[2][kParseSeg], Cn[2][kParseSeg];
  const uint64_t s0 = (uint64_t)blockIdx.x * kParseSeg;
  // bytes of the segment and of the largest block beyond it (zero past the payload)
  for (uint32_t i = threadIdx.x; i < kParseSeg + a.maxb; i += kBlockThreads)
    bytes[i] = s0 + i < a.len ? a.p[s0 + i] : (uint8_t)0;
  __syncthreads();
  // one step from every position: bit count at +bc_off, size of a full block
  for (uint32_t q = threadIdx.x; q < kParseSeg; q += kBlockThreads) {
    const uint32_t bc = a.has_bc ? bytes[q + a.bc_off] : 0u;
    J[0][q] = (uint16_t)(q + a.sz[bc <= 32 ? bc : 33]);  // >= kParseSeg: the walk has left
    Cn[0][q] = 1;
  }
  __syncthreads();
  // pointer doubling: after round r every position knows where 2^r steps take it
  int cur = 0;
  for (uint32_t r = 0; r < a.rounds; r++) {
    for (uint32_t q = threadIdx.x; q < kParseSeg; q += kBlockThreads) {
      const uint32_t j = J[cur][q];
      if (j < kParseSeg) {
        J[cur ^ 1][q] = J[cur][j];
        Cn[cur ^ 1][q] = (uint16_t)(Cn[cur][q] + Cn[cur][j]);
      } else {
        J[cur ^ 1][q] = (uint16_t)j;
        Cn[cur ^ 1][q] = Cn[cur][q];
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  for (uint32_t e = threadIdx.x; e < a.maxb; e += kBlockThreads) {
    const uint32_t j = J[cur][e];
    // j < kParseSeg cannot happen after enough rounds; report it as a corrupt table
    tab[(size_t)blockIdx.x * a.maxb + e] = make_uint2(j >= kParseSeg ? j - kParseSeg : 0xffffffffu, Cn[cur][e]);
  }
}

// The same tables by walking from each entry offset (one thread per entry, the bytes in
// LDS): maxb walks of at most kParseSeg / (smallest block) steps.  With blocks of tens of
// bytes that is far less work than doubling over every byte position, and the small LDS
// footprint lets many segments share a CU; the launcher picks it when the walks are
// short enough.
__global__ __launch_bounds__(kBlockThreads) void k_parse_seg_walk(ParseArgs a, uint2* __restrict__ tab) {
  __shared__ uint8_t bytes[kParseSeg + kParseMaxBlock];
  __shared__ uint16_t sz[34];
  const uint64_t s0 = (uint64_t)blockIdx.x * kParseSeg;
  for (uint32_t i = threadIdx.x; i < kParseSeg + a.maxb; i += kBlockThreads)
    bytes[i] = s0 + i < a.len ? a.p[s0 + i] : (uint8_t)0;
  if (threadIdx.x < 34) sz[threadIdx.x] = a.sz[threadIdx.x];
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < a.maxb; e += kBlockThreads) {
    uint32_t q = e, c = 0;
    while (q < kParseSeg) {  // every block is at least one byte
      const uint32_t bc = a.has_bc ? bytes[q + a.bc_off] : 0u;
      q += sz[bc <= 32 ? bc : 33];
      c++;
    }
    tab[(size_t)blockIdx.x * a.maxb + e] = make_uint2(q - kParseSeg, c);
  }
}

// ---- 2. composition of `group` consecutive tables ---------------------------------------
__global__ __launch_bounds__(kBlockThreads) void k_parse_compose(const uint2* __restrict__ tin, uint32_t n_in,
                                                                uint32_t group, uint32_t maxb,
                                                                uint2* __restrict__ tout) {
  const uint32_t g = blockIdx.x;
  const uint32_t k0 = g * group, k1 = min(k0 + group, n_in);
  for (uint32_t e = threadIdx.x; e < maxb; e += kBlockThreads) {
    uint32_t st = e, cnt = 0;
    for (uint32_t k = k0; k < k1 && st < maxb; k++) {
      const uint2 v = tin[(size_t)k * maxb + st];
      st = v.x;
      cnt += v.y;
    }
    tout[(size_t)g * maxb + e] = make_uint2(st < maxb ? st : 0xffffffffu, cnt);
  }
}

// ---- 3. entries back down: out[k] = {entry offset, blocks before} for each table k of
//         a group, from the group's own entry
__global__ __launch_bounds__(kBlockThreads) void k_parse_down(const uint2* __restrict__ tab, uint32_t n,
                                                             uint32_t group, uint32_t maxb,
                                                             const uint2* __restrict__ gentry,
                                                             uint2* __restrict__ out, uint32_t* err) {
  const uint32_t g = blockIdx.x * kBlockThreads + threadIdx.x;
  const uint32_t k0 = g * group;
  if (k0 >= n) return;
  const uint32_t k1 = min(k0 + group, n);
  uint2 en = gentry ? gentry[g] : make_uint2(0u, 0u);
  for (uint32_t k = k0; k < k1; k++) {
    out[k] = en;
    if (en.x >= maxb) {  // a walk that never left a segment: the stream is not a block sequence
      atomicOr(err, TSZP_ERR_STREAM);
      en = make_uint2(0xffffffffu, en.y);
      continue;
    }
    const uint2 v = tab[(size_t)k * maxb + en.x];
    en = make_uint2(v.x, en.y + v.y);
  }
}

// ---- 4. the real walk of every segment -------------------------------------------------
__global__ __launch_bounds__(kBlockThreads) void k_parse_final(ParseArgs a, const uint2* __restrict__ entry,
                                                              uint32_t nseg, uint64_t n, uint32_t bs, uint32_t nb,
                                                              uint32_t* __restrict__ blk, uint32_t* err) {
  const uint32_t s = blockIdx.x * kBlockThreads + threadIdx.x;
  if (s >= nseg) return;
  const uint2 en = entry[s];
  if (en.x >= a.maxb) return;  // reported by k_parse_down
  uint64_t pos = (uint64_t)s * kParseSeg + en.x;
  const uint64_t end = (uint64_t)(s + 1) * kParseSeg;
  uint32_t idx = en.y;
  while (pos < end && idx < nb) {
    blk[idx] = (uint32_t)pos + a.out_base;
    const uint64_t i = (uint64_t)idx * bs;
    const uint32_t cur = i + bs > n ? (uint32_t)(n - i) : bs;
    const uint32_t dm = a.dm1 ? cur - 1u : cur;  // deltas of the block
    uint64_t q = pos + (a.anchor ? 4u : 0u);
    if (a.topo || a.bc_always || dm > 0) {
      if (q >= a.len) {
        atomicOr(err, TSZP_ERR_STREAM);
        return;
      }
      const uint32_t bc = a.p[q++];
      if (bc > 32) {
        atomicOr(err, TSZP_ERR_STREAM);
        return;
      }
      if (bc) q += mag_bytes(dm, bc) + ((dm + 7u) >> 3);
    }
    if (a.topo) q += (2u * cur + 7u) >> 3;
    if (q > a.len) {
      atomicOr(err, TSZP_ERR_STREAM);
      return;
    }
    pos = q;
    idx++;
    if (idx == nb) {
      blk[nb] = (uint32_t)pos + a.out_base;
      if (a.end_out) *a.end_out = (uint32_t)pos + a.out_base;
      err[1] = 1u;  // the last block was reached
    }
  }
}

// every chunk offset off[1..T) must be a block start (or the end)
__global__ __launch_bounds__(kBlockThreads) void k_parse_check_offsets(const uint8_t* __restrict__ head, uint32_t T,
                                                                      const uint32_t* __restrict__ blk, uint32_t nb,
                                                                      uint32_t* err) {
  const uint32_t k = blockIdx.x * kBlockThreads + threadIdx.x + 1;
  if (k >= T) return;
  uint64_t o = 0;
  for (int b = 0; b < 8; b++) o |= (uint64_t)head[8 * k + b] << (8 * b);
  uint32_t lo = 0, hi = nb;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (blk[mid] < o) lo = mid + 1;
    else hi = mid;
  }
  if (o > 0xffffffffull || blk[lo] != (uint32_t)o) atomicOr(err, TSZP_ERR_STREAM);
}

// ---------------------------------------------------------------------------------
// launcher: tables, two composition levels, a single-thread top, entries down, walk
// ---------------------------------------------------------------------------------
size_t parse_work_bytes(uint64_t len, uint32_t maxb) {
  const uint64_t nseg = (len + kParseSeg - 1) / kParseSeg + 1;
  const uint64_t n1 = (nseg + kParseGroup - 1) / kParseGroup, n2 = (n1 + kParseGroup - 1) / kParseGroup;
  return (size_t)((nseg + n1 + n2) * maxb * 8 + (nseg + n1 + n2 + 8) * 8 + 64);
}

hipError_t parse_blocks(const ParseArgs& a, uint64_t n, uint32_t bs, uint32_t nb, uint32_t* blk, void* work,
                        uint32_t* err, hipStream_t st) {
  const uint32_t nseg = (uint32_t)((a.len + kParseSeg - 1) / kParseSeg);
  if (!nseg) return hipErrorInvalidValue;
  const uint32_t n1 = (nseg + kParseGroup - 1) / kParseGroup, n2 = (n1 + kParseGroup - 1) / kParseGroup;
  if (n2 > kParseGroup) return hipErrorInvalidValue;  // payload beyond what two levels cover
  uint2* t0 = reinterpret_cast<uint2*>(work);
  uint2* t1 = t0 + (size_t)nseg * a.maxb;
  uint2* t2 = t1 + (size_t)n1 * a.maxb;
  uint2* e2 = t2 + (size_t)n2 * a.maxb;  // entries of the level-2 tables
  uint2* e1 = e2 + n2;                   // entries of the level-1 tables
  uint2* e0 = e1 + n1;                   // entries of the segments
  // walks when they stay short (steps per lane), else pointer doubling
  uint32_t minb = 0xffffu;
  for (int b = 0; b < 34; b++) minb = a.sz[b] < minb ? a.sz[b] : minb;
  const uint64_t walk_steps = (uint64_t)a.maxb * (kParseSeg / (minb ? minb : 1u)) / kBlockThreads;
  if (walk_steps <= 16384) hipLaunchKernelGGL(k_parse_seg_walk, dim3(nseg), dim3(kBlockThreads), 0, st, a, t0);
  else hipLaunchKernelGGL(k_parse_seg, dim3(nseg), dim3(kBlockThreads), 0, st, a, t0);
  hipLaunchKernelGGL(k_parse_compose, dim3(n1), dim3(kBlockThreads), 0, st, t0, nseg, kParseGroup, a.maxb, t1);
  hipLaunchKernelGGL(k_parse_compose, dim3(n2), dim3(kBlockThreads), 0, st, t1, n1, kParseGroup, a.maxb, t2);
  // top: one thread walks the level-2 tables from the payload start
  hipLaunchKernelGGL(k_parse_down, dim3(1), dim3(kBlockThreads), 0, st, t2, n2, kParseGroup, a.maxb, nullptr, e2, err);
  hipLaunchKernelGGL(k_parse_down, dim3((n2 + kBlockThreads - 1) / kBlockThreads), dim3(kBlockThreads), 0, st, t1, n1,
                     kParseGroup, a.maxb, e2, e1, err);
  hipLaunchKernelGGL(k_parse_down, dim3((n1 + kBlockThreads - 1) / kBlockThreads), dim3(kBlockThreads), 0, st, t0,
                     nseg, kParseGroup, a.maxb, e1, e0, err);
  hipLaunchKernelGGL(k_parse_final, dim3((nseg + kBlockThreads - 1) / kBlockThreads), dim3(kBlockThreads), 0, st, a,
                     e0, nseg, n, bs, nb, blk, err);
  return hipGetLastError();
}

// ---- element-partitioned streams: chunk starts and the chunk-end check -----------------
__device__ __forceinline__ uint64_t rd_u64(const uint8_t* p) {
  uint64_t o = 0;
  for (int b = 0; b < 8; b++) o |= (uint64_t)p[b] << (8 * b);
  return o;
}

__global__ __launch_bounds__(kBlockThreads) void k_parse_elem_finish(const uint8_t* __restrict__ tab, uint32_t T,
                                                                    uint64_t n, uint32_t ukc, int ra, uint64_t avail,
                                                                    uint32_t* __restrict__ blk,
                                                                    const uint32_t* __restrict__ ends,
                                                                    uint32_t* err) {
  const uint32_t t = blockIdx.x * kBlockThreads + threadIdx.x;
  if (t >= T) return;
  const uint64_t ce = n / T, lo = (uint64_t)t * ce, hi = t == T - 1 ? n : lo + ce;
  const uint64_t o = rd_u64(tab + 8 * (size_t)t);
  uint64_t e = ends[t];
  if (hi - lo == 0) e = o;                     // no units
  else if (!ra && hi - lo == 1) e = o + 4;     // the anchor alone
  if (!ra && hi > lo) blk[(size_t)t * ukc] = (uint32_t)o;  // the first unit holds the anchor
  if (t + 1 < T) {
    if (e != rd_u64(tab + 8 * (size_t)(t + 1))) atomicOr(err, TSZP_ERR_STREAM);
  } else {
    if (e == 0xffffffffull || e > avail) atomicOr(err, TSZP_ERR_STREAM);
    if (hi - lo <= (ra ? 0u : 1u)) blk[(size_t)t * ukc + (hi > lo ? 1u : 0u)] = (uint32_t)e;  // the end entry
  }
}

hipError_t parse_elem_finish(const uint8_t* tab, uint32_t T, uint64_t n, uint32_t ukc, bool ra, uint64_t avail,
                             uint32_t* blk, uint32_t* ends, uint32_t* err, hipStream_t st) {
  hipLaunchKernelGGL(k_parse_elem_finish, dim3((T + kBlockThreads - 1) / kBlockThreads), dim3(kBlockThreads), 0, st,
                     tab, T, n, ukc, ra ? 1 : 0, avail, blk, ends, err);
  return hipGetLastError();
}

hipError_t parse_check_offsets(const uint8_t* head, uint32_t T, const uint32_t* blk, uint32_t nb, uint32_t* err,
                               hipStream_t st) {
  if (T <= 1) return hipSuccess;
  hipLaunchKernelGGL(k_parse_check_offsets, dim3((T + kBlockThreads - 1) / kBlockThreads), dim3(kBlockThreads), 0, st,
                     head, T, blk, nb, err);
  return hipGetLastError();
}

}  // namespace tszp
