// tszp_noalloc.hip -- plain random-access compression that allocates nothing.
//
// tszp_plain_compress keeps its scratch (the look-back status words of every tile, the
// tile counter and the result words) in the library's grow-only workspace, which calls
// hipMalloc the first time a size is reached.  Callers that may not allocate -- code
// under stream capture, tight in-situ loops, harnesses that forbid it -- call
// tszp_plain_compress_noalloc instead: the same kernels, with that scratch in the
// file-scope device array below.  Each device holds its own copy of the array until
// the process ends; every call on a device shares it, so calls that overlap on one
// device (other threads, other streams) race, and each call clears what must start at
// zero with a fill kernel (no hipMemsetAsync).
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/toposzp_hip.h"
#include "tszp_kernels.h"

namespace tszp {
namespace {

// Largest tile count of one call: 2^17 tiles of at least 2048 elements each (the
// smallest tile of any block size) covers every field of up to 2^28 elements.
constexpr uint32_t kNoAllocMaxTiles = 1u << 17;
// two status words per tile, the tile counter, then the result words
constexpr size_t kNoAllocWords = 2 * (size_t)kNoAllocMaxTiles + 2 + sizeof(CompressResult) / 8;

__device__ __align__(256) uint64_t scratch[kNoAllocWords];

// Device address of `scratch` on each device, looked up once per device with that
// device current (the array's name in host code is only a stand-in).
constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t* g_ptr[kMaxDevices] = {};

int scratch_of(int dev, uint64_t** out) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(TSZP_EINVAL, "plain_compress_noalloc: device id out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return set_error(TSZP_EHIP, "plain_compress_noalloc: hipGetSymbolAddress failed");
    g_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_ptr[dev];
  return TSZP_OK;
}

}  // namespace
}  // namespace tszp

using namespace tszp;

extern "C" int tszp_plain_compress_noalloc(const float* d_in, size_t n, float eb, int bs, unsigned T, uint8_t* d_out,
                                           size_t cap, uint32_t* d_blk, uint64_t* out_bytes, void* stream) {
  clear_error();
  if (!d_in || !d_out || !d_blk || !out_bytes || T < 1 || T > 4096)
    return set_error(TSZP_EINVAL, "plain_compress_noalloc: bad arguments");
  if (!(eb > 0)) return set_error(TSZP_EINVAL, "plain_compress_noalloc: error bound must be positive");
  if (bs < 2 || bs > 1025) return set_error(TSZP_EINVAL, "plain_compress_noalloc: blockSize out of range [2, 1025]");
  if (n == 0 || n >= 0x7fffffffull)
    return set_error(TSZP_EINVAL, "plain_compress_noalloc: element count out of range [1, 2^31)");
  int S, EPL;
  uint32_t bpt;
  block_geometry((uint32_t)bs, &S, &EPL, &bpt);
  const uint32_t nb = (uint32_t)((n + bs - 1) / bs);
  const uint32_t ntiles = (nb + bpt - 1) / bpt;
  if (ntiles > kNoAllocMaxTiles)
    return set_error(TSZP_EINVAL, "plain_compress_noalloc: field needs more tiles than the static scratch holds "
                                  "(at most 2^17 tiles: 2^28 elements at any block size)");
  const size_t hdr = 4 + 8 * (size_t)T;
  if (((uintptr_t)(d_out + hdr)) & 3u) return set_error(TSZP_EINVAL, "plain_compress_noalloc: stream payload must be 4-byte aligned");
  if (cap < hdr + 4) return set_error(TSZP_ECAP, "plain_compress_noalloc: stream capacity too small");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(TSZP_EHIP, "plain_compress_noalloc: hipGetDevice failed");
  uint64_t* base;
  int r = scratch_of(dev, &base);
  if (r) return r;
  hipStream_t st = (hipStream_t)stream;
  uint64_t* st_bytes = base;
  uint64_t* st_cnt = base + ntiles;
  uint32_t* counter = reinterpret_cast<uint32_t*>(base + 2 * (size_t)ntiles);
  CompressResult* res = reinterpret_cast<CompressResult*>(base + 2 * (size_t)kNoAllocMaxTiles + 2);
  FillRanges fr{};
  fr.add(reinterpret_cast<uint32_t*>(base), 0u, (2 * (size_t)ntiles + 2) * 2);
  fr.add(reinterpret_cast<uint32_t*>(res), 0u, sizeof(CompressResult) / 4);
  if (launch_fill_multi(fr, st) != hipSuccess) return set_error(TSZP_EHIP, "plain_compress_noalloc: fill launch failed");
  // SZp/src/szp_float.cc:2125: the reciprocal is an int / float division, then widened
  volatile float finv = 1 / eb;
  CompressArgs a{};
  a.in = d_in;
  a.n = (uint32_t)n;
  a.bs = (uint32_t)bs;
  a.nb = nb;
  a.ntiles = ntiles;
  a.eb = eb;
  a.inv = (double)finv;
  a.payload = d_out + hdr;
  const uint64_t pc = cap - hdr - 4;  // keep 4 readable bytes past the end
  a.cap = pc > 0xffffffffull ? 0xffffffffull : pc;
  a.blk = d_blk;
  a.st_bytes = st_bytes;
  a.st_cnt = st_cnt;
  a.tile_counter = dyn_counter(counter);
  a.res = res;
  if (launch_compress(kModePlain, a, st) != hipSuccess)
    return set_error(TSZP_EHIP, "plain_compress_noalloc: codec launch failed");
  if (launch_write_header(d_out, 2, T, 1, nb, d_blk, 0, eb, st) != hipSuccess)
    return set_error(TSZP_EHIP, "plain_compress_noalloc: header launch failed");
  CompressResult h;
  if (hipMemcpyAsync(&h, res, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_error(TSZP_EHIP, "plain_compress_noalloc: result copy failed");
  if (h.error) return set_error(TSZP_EDEVICE, "plain_compress_noalloc: device error (look-back timeout or capacity)");
  *out_bytes = hdr + h.payload_bytes;
  return TSZP_OK;
}
