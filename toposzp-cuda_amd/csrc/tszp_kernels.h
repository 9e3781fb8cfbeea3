// tszp_kernels.h -- launch interface between the host API (tszp_api.cpp) and the
// HIP kernels.  Host-only callers see plain structs and launcher functions.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdlib>
#include <stdint.h>

#include "tszp_types.h"
#include "../../include/toposzp_hip.h"

namespace tszp {

enum Mode {
  kModeTopoClassify = 0,  // topology stream, types classified from the field itself
  kModeTopoMap = 1,       // topology stream, types from a uint8 map (a caller's CP list)
  kModePlain = 2,         // plain random-access float stream
  kModeInt = 3,           // int stream (sort positions)
  kModeTopoInt = 4,       // topology stream of pre-quantized ints, types from a uint8 map
};

struct CompressArgs {
  const void* in;         // float[n] or int[n]
  uint32_t n, bs, nb, ntiles;
  float eb;
  double inv;             // reciprocal used for quantization (mode dependent)
  Grid grid;              // classify mode
  const uint8_t* tmap;    // map mode
  uint8_t* payload;       // stream bytes after the chunk-offset header (4-byte aligned)
  uint64_t cap;           // payload capacity in bytes
  uint32_t* blk;          // nb + 1 block offsets relative to payload
  uint64_t* st_bytes;     // look-back status words, ntiles each
  uint64_t* st_cnt;
  uint32_t* tile_counter;
  uint32_t* cp_key;       // classify mode: order-preserving value keys, row-major
  uint32_t* cp_pay;       // extremum ordinal, or 0xffffffff for saddles
  uint32_t cp_cap;
  uint32_t lds_rows;      // classify mode, fast path: stage the rows next to each tile in LDS
  CompressResult* res;
};

struct DecodeArgs {
  const uint8_t* payload; // 4-byte aligned; readable up to the next multiple of 4 past the end
  const uint32_t* blk;    // nb + 1 block offsets relative to payload
  uint32_t n, bs, nb;
  float eb;
  void* out;              // float[n] or int[n]
  uint8_t* types8;        // topology: uint8 types (nullable)
  int32_t* types32;       // topology: int types, the reference FN layout (nullable)
  uint32_t* err;
};

void block_geometry(uint32_t bs, int* S, int* EPL, uint32_t* blocks_per_tile);
hipError_t launch_compress(int mode, const CompressArgs& a, hipStream_t st);
hipError_t launch_decode(int mode, const DecodeArgs& a, hipStream_t st);

// ---- batches of fields (tszp_codec.hip) ----
// One entry per field; the fields of one launch share a block-size class (block_geometry)
// and number their tiles contiguously from tile0 in field order.  Streams have one
// chunk: plain [f32 eb big-endian][u64 0], topology [u64 0], sort positions
// [u64 E][u64 0] (the header bytes before the payload are `hdr`).
constexpr uint32_t kBatchHdr = 12;      // plain: [f32 eb big-endian][u64 off[1] = 0]
constexpr uint32_t kBatchTopoHdr = 8;   // topology: [u64 off[1] = 0]
constexpr uint32_t kBatchSpHdr = 16;    // sort positions: [u64 E][u64 off[1] = 0]
struct BatchField {
  const void* in;         // compress: float[n] (int[n] for sort positions); decode: the stream
  void* out;              // compress: the stream (4-byte aligned); decode: float[n] (int[n])
  uint32_t* blk;          // nb + 1 block offsets relative to the payload
  uint64_t cap;           // compress: payload capacity in bytes
  uint32_t n, bs, nb, ntiles, tile0;  // n, nb, ntiles: the largest when n_dev is set
  float eb;
  double inv;             // compress: reciprocal used for quantization
  uint32_t hdr;           // header bytes before the payload
  const uint64_t* n_dev;  // nullable: element count read on the device (<= n); the
                          // tiles past its blocks exit at once
  Grid grid;              // topology compression: the field's geometry
  uint32_t* cp_key;       // topology compression: critical-point keys and payloads, row-major
  uint32_t* cp_pay;
  uint32_t cp_cap;
  uint8_t* types8;        // topology decode: uint8 types (nullable)
};
hipError_t launch_batch_tilemap(const BatchField* fields, uint32_t count, uint32_t* tmap, hipStream_t st);
// tiles [base, base + ntiles) of the batch's numbering, all of one block-size class;
// mode kModePlain, kModeTopoClassify (status_cnt: second look-back words, for the
// critical-point counts) or kModeInt; v4: the 16-byte path (block size 16, 32, 64, 128
// or 256, every input 16-byte aligned, classification: every fastest axis a multiple
// of 4 and every field of the class `nd`-dimensional); counter (nullable, zeroed):
// tiles handed out from it instead of workgroup ids
hipError_t launch_compress_batch(int mode, uint32_t bs_class, bool v4, int nd, const BatchField* fields,
                                 const uint32_t* tmap, uint32_t base, uint32_t ntiles, uint64_t* status,
                                 uint64_t* status_cnt, CompressResult* res, uint32_t* counter, hipStream_t st);
// mode kModePlain, kModeTopoMap (values and types) or kModeInt
hipError_t launch_decode_batch(int mode, uint32_t bs_class, const BatchField* fields, const uint32_t* tmap,
                               uint32_t base, uint32_t ntiles, CompressResult* res, hipStream_t st);
// Estimated payload bytes of every field at each block size of kTuneSizes, from
// sampled windows of kTuneWindow elements (est: count * kTuneCands words).
constexpr int kTuneCands = 4;
constexpr uint32_t kTuneSizes[kTuneCands] = {16, 32, 64, 128};
constexpr uint32_t kTuneWindow = 128;
constexpr uint32_t kTuneMaxWindows = 512;
hipError_t launch_batch_tune(const BatchField* fields, uint32_t count, uint32_t* est, hipStream_t st);

// tszp_last_error() of the calling thread, for entry points outside tszp_api.cpp
int set_error(int code, const char* msg);
void clear_error();
// the library's per-stage timers (tszp_prof_*), for entry points outside tszp_api.cpp
enum : int { kProfTopoBatch = 15, kProfTopoBatchDecode = 16, kProfTopoBatchRestore = 17, kProfTopoBatchSort = 18 };
bool prof_on();
void prof_add(int kind, double ms);
Grid grid_of(int ndim, const uint32_t* dims);

// ---- batched topology pipeline (tszp_topo_batch.hip) ----
// Fields of a batch live at aligned offsets of flat arrays (critical-point keys, sort
// positions, decoded values, types, restored values): field f covers
// [base, base + n) of every one of them.
constexpr uint32_t kFlatAlign = 16384;
struct TopoBatchDev {
  uint64_t base;        // flat element offset (a multiple of kFlatAlign)
  uint32_t n;           // elements
  uint32_t chunk0;      // first gather chunk (kSortTile key slots each)
  uint32_t nchunks;
  uint8_t* sp_stream;   // the field's sort-position stream
};
// table order = the batch's field order; res[i] the topology compression result of field i
hipError_t launch_tb_offsets(const TopoBatchDev* tab, uint32_t count, const CompressResult* res, uint32_t* cbase,
                             uint32_t* total, uint32_t* err, hipStream_t st);
hipError_t launch_tb_chunkmap(const TopoBatchDev* tab, uint32_t count, uint32_t* chunk_field, hipStream_t st);
hipError_t launch_tb_gather(const TopoBatchDev* tab, const uint32_t* chunk_field, uint32_t nchunks,
                            const CompressResult* res, const uint32_t* cbase, const uint32_t* cp_key,
                            const uint32_t* cp_pay, uint32_t* keys, uint32_t* hi, uint32_t* pay, hipStream_t st);
// ---- geometry of a restoration (tszp_post.hip) ----
// One field, or a batch of fields in flat arrays: field f covers [base, base + n) of the
// decoded values, types, sort positions (its k-th extremum at base + k) and restored
// values, with base a multiple of kFlatAlign, so that the field of flat index e is
// fmap[e / kFlatAlign].  Indices stay flat (a field's points are contiguous, so e +-
// stride is a neighbour); coordinates are the field's own.
constexpr int kFlatShift = 14;
static_assert((1u << kFlatShift) == kFlatAlign, "flat alignment");
struct PpField {
  Grid g;
  uint32_t base, n;
  float eb;
  uint32_t tile0, tx, ty, tz;  // window-solve tiles of the field: [tile0, tile0 + tx * ty * tz)
  const uint64_t* n_ext;       // extremum count of the field (device)
};
struct PpGeo {
  int ndim;
  uint32_t n;                  // points (one field) or flat size (batch)
  Grid g;                      // one field
  float eb;                    // one field
  const PpField* fields;       // batch: device table, null for one field
  const uint32_t* fmap;        // batch: field of each kFlatAlign points
  const uint32_t* tmap;        // batch: field of each window-solve tile
  uint32_t* spmax;             // batch: largest sort position of each field
  unsigned long long* fcount;  // batch: missing / spurious / changed / max |y - d| bits per field
  uint32_t nfields;
  uint32_t ntiles;             // window-solve tiles (all fields)
  uint32_t tx, ty, tz;         // one field's solve tiles
};

// Restoration of a batch of fields laid out in flat arrays (tszp_api.cpp): decoded
// values, types, sort positions (the k-th extremum of field f at base + k), restored
// values; one entry per field in flat order.  The workspace is the library's per-device
// one; info receives each field's missing / spurious / changed counts and largest
// |y - d| (moves, windows and rounds are the batch's).
struct RestoreBatchField {
  int ndim;
  uint32_t dims[3];
  float eb;
  uint64_t base;
  const uint64_t* n_ext_dev;  // device: the field's extremum count
};
// ws: the restoration's workspace (tszp_workspace_create), or null for the device's
// own one (shared by the library's calls on that device, one at a time)
int restore_batch(const RestoreBatchField* fields, uint32_t count, uint64_t flat, const float* d, const uint8_t* t,
                  const int32_t* sp, float* y, int max_rounds, tszp_restore_info* info, tszp_workspace* ws,
                  hipStream_t st);

// Tile counter to pass to a look-back kernel: null (tiles = workgroup ids) unless
// TSZP_TILE_COUNTER is set, which hands tiles out from the counter instead.
inline uint32_t* dyn_counter(uint32_t* counter) { return getenv("TSZP_TILE_COUNTER") ? counter : nullptr; }

// ---- element-partitioned streams (tszp_elem.hip): threadblock codec, double RA ----
struct ElemArgs {
  const void* in;         // float[n] or double[n]
  uint32_t n, bs, T;
  uint32_t ce;            // n / T: elements of every chunk but the last
  uint32_t ukc;           // units of every chunk but the last
  uint32_t nu, ntiles;    // units, tiles
  FastDiv div_ukc;
  double inv;             // reciprocal used for quantization
  uint8_t* payload;       // stream bytes after the header (4-byte aligned)
  uint64_t cap;           // payload capacity in bytes
  uint32_t* blk;          // nu + 1 unit offsets relative to payload
  uint64_t* st_bytes;     // look-back status words, ntiles
  uint32_t* tile_counter;
  CompressResult* res;
};
struct ElemDecodeArgs {
  ElemArgs geo;           // geometry fields only (n, bs, T, ce, ukc, nu, ntiles, div_ukc)
  const uint8_t* payload; // 4-byte aligned; readable up to the next multiple of 4 past the end
  const uint32_t* blk;
  double eb;
  void* out;
  uint64_t* st;           // segmented look-back status words, ntiles
  uint32_t* tile_counter;
  uint32_t* err;
};
uint32_t elem_units(uint64_t n, uint32_t bs, uint32_t T, bool ra, uint32_t* units_per_chunk);
uint32_t elem_tiles(uint32_t nu, uint32_t bs);
hipError_t launch_elem_compress(int dsize, bool ra, const ElemArgs& a, hipStream_t st);
hipError_t launch_elem_decode(int dsize, bool ra, const ElemDecodeArgs& a, hipStream_t st);
hipError_t launch_elem_header(uint8_t* stream, int dsize, unsigned T, uint32_t ukc, const uint32_t* blk, double eb,
                              hipStream_t st);
// out[i] = q[i] - q[i - 1], q[-1] = 0, q = C cast of x * inv
hipError_t launch_predict_quant(const float* x, uint32_t n, double inv, int32_t* out, hipStream_t st);

// ---- double 3D topology: quantization and saddle fix (tszp_saddle.hip) ----
// types: zero-filled uint8[n], then marked from AoS CP records of 6 ints (x, y, z, type, ..)
hipError_t launch_mark_cps3(const int32_t* cps, uint32_t count, const Grid& g, uint8_t* types, hipStream_t st);
hipError_t launch_quant_f64(const double* x, uint32_t n, double inv, int32_t* q, hipStream_t st);
// interior saddles of the type map listed in index order: cnt holds saddle_blocks(n) + 1
// zeroed words (the total lands in cnt[saddle_blocks(n)]), scan_st / scan_ctr zeroed
// look-back words of launch_scan_excl over that many entries
uint32_t saddle_blocks(uint32_t n);
hipError_t launch_saddle_list(const Grid& g, const uint8_t* types, uint32_t n, uint32_t* cnt, uint64_t* scan_st,
                              uint32_t* scan_ctr, uint32_t* err, uint32_t* list, hipStream_t st);
// one Jacobi pass over the listed saddles: the accumulated move of saddle list[i] given
// cd_in (nullable = all zero) on top of q goes to cd_out[list[i]] (nullable; must be zero
// away from saddles) and to upd[i] (nullable)
hipError_t launch_saddle_pass(const Grid& g, const int32_t* q, const uint8_t* types, const uint32_t* list,
                              const uint32_t* count, const int8_t* cd_in, int8_t* cd_out, int8_t* upd,
                              hipStream_t st);
// q[list[i]] += upd[i]; *moved += the saddles that moved
hipError_t launch_saddle_apply(int32_t* q, const int8_t* upd, const uint32_t* list, const uint32_t* count,
                               uint32_t* moved, hipStream_t st);

// ---- topology helpers (tszp_topo.hip) ----
struct ClassifyArgs {
  const float* f;
  uint32_t n;
  Grid grid;
  float eb;
  double inv;
  uint32_t ntiles;
  uint64_t* st_cnt;
  uint32_t* tile_counter;
  uint8_t* types;          // nullable: type per element
  int32_t* cps;            // nullable: AoS records of `rec` ints: x, y[, z], type, bin, 0
  int rec;                 // 5 (2D) or 6 (3D)
  uint32_t cp_cap;
  CompressResult* res;
};
hipError_t launch_classify(const ClassifyArgs& a, hipStream_t st);
uint32_t classify_tiles(uint32_t n);
hipError_t launch_cp_keys(const int32_t* cps, uint32_t count, int rec, const float* data, uint64_t data_n,
                          uint32_t row, uint64_t plane, uint32_t* keys, uint32_t* hi, uint32_t* pay, uint32_t* err,
                          hipStream_t st);

// ---- helpers (tszp_util.hip) ----
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t count, hipStream_t st);
// up to kFillRanges (pointer, value, count) ranges in one launch
constexpr int kFillRanges = 8;
struct FillRanges {
  uint32_t* p[kFillRanges];
  uint32_t v[kFillRanges];
  size_t count[kFillRanges];
  int n;
  void add(uint32_t* ptr, uint32_t val, size_t cnt) {
    p[n] = ptr;
    v[n] = val;
    count[n] = cnt;
    n++;
  }
};
hipError_t launch_fill_multi(const FillRanges& r, hipStream_t st);
hipError_t launch_minmax(const float* f, size_t n, uint32_t* out2 /* okey min, okey max */,
                         hipStream_t st);
hipError_t launch_write_header(uint8_t* stream, int kind, unsigned T, int adaptive, uint32_t nb,
                               const uint32_t* blk, uint64_t count, float eb, hipStream_t st);

// ---- sort positions (tszp_sort.hip) ----
constexpr uint32_t kSortTile = 4096;  // items per sort / scan / rank tile
struct SortWork {
  uint32_t* keys_alt;      // count
  uint32_t* hi_alt;        // count (only with a second key word)
  uint32_t* pay_alt;       // count
  uint32_t* hist;          // 256 * ceil(count / kSortTile)
  uint64_t* st;            // look-back words: 8 passes * ceil(256 * tiles / kSortTile) + rank tiles
  uint32_t* counters;      // 16 tile counters
  uint32_t* err;
};
size_t sort_hist_words(uint32_t count);
size_t sort_status_words(uint32_t count);
// In-place exclusive scan of a[0..count) in one launch (decoupled look-back);
// status: scan_status_words(count) zeroed u64 words, counter: one zeroed u32.
size_t scan_status_words(uint32_t count);
hipError_t launch_scan_excl(uint32_t* a, uint32_t count, uint64_t* status, uint32_t* counter, uint32_t* err,
                            hipStream_t st);
// Optional per-launch timing: begin/end are called around a launch on its stream.
struct LaunchTimer {
  virtual void begin(hipStream_t st) = 0;
  virtual void end(hipStream_t st) = 0;
  virtual ~LaunchTimer() = default;
};

// Stable LSD radix sort of (key, payload) pairs, 8-bit digits; key_hi (nullable)
// is a more significant key word.  Results land back in keys / key_hi / pay.
// scatter_timer (nullable) times every scatter launch.
// dcount (nullable): the item count on the device, at most `count` (the launches cover
// `count`); hi_bits: significant low bits of key_hi (passes of 8 bits over them).
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* key_hi, uint32_t* pay, uint32_t count,
                            const SortWork& wk, hipStream_t st, LaunchTimer* scatter_timer = nullptr,
                            const uint32_t* dcount = nullptr, int hi_bits = 32);
// Dense rank of distinct values inside quantized bins over the sorted keys
// (SZp/src/szp_float.cc:344-358): rank resets at each new bin and grows at each new
// value.  Bins come from the float value (by_value: bin = (int)((v + eb) * inv)) or
// from key_hi (biased bin).  Writes out[pay] = rank for every pay < out_cap
// (saddles carry pay = 0xffffffff).
// With field_eb / field_inv (device, per field), key_hi is the field of each key and
// its bins come from the value with that field's error bound (sorted by (field, value)).
hipError_t launch_rank(const uint32_t* keys, const uint32_t* key_hi, const uint32_t* pay, uint32_t count,
                       float eb, double inv, int32_t* out, uint32_t out_cap, const SortWork& wk, hipStream_t st,
                       const uint32_t* dcount = nullptr, const float* field_eb = nullptr,
                       const double* field_inv = nullptr);

// ---- sort positions by per-bin bitmaps (tszp_binrank.hip) ----
// Ranks by value (keys are order-preserving float keys, bins come from the values)
// without a full sort: bucket the keys by bin, mark each bin's distinct keys in a
// bitmap over the bin's key span, and read each rank as a prefix popcount.  Bins
// whose span exceeds kBinWideBits (the bin around zero) are ranked by the radix
// sort above, on their buckets only.  Keys and payloads are read as uint4: both arrays
// must be 16-byte aligned and readable up to the next multiple of 4 entries.
constexpr uint32_t kBinMax = 4096;             // bins per call (more: radix path)
constexpr uint32_t kBinSlices = 512;           // key slices (one workgroup each) for count and scatter
constexpr uint32_t kBinChunkWords = 16384;     // bitmap words per workgroup
constexpr uint32_t kBinMaxChunks = 256;        // chunks per bin (spans up to 2^27 keys: every bin but the zero cell at usual bounds)
constexpr uint64_t kBinWideBits = (uint64_t)kBinChunkWords * 32 * kBinMaxChunks;
struct BinWork {
  uint32_t* mm;       // 4: min key, ~max key, min bin ^ 2^31, ~(max bin ^ 2^31)
  uint32_t* part;     // range partials: min and max key per workgroup
  uint32_t* klo;      // kBinMax + 1: lowest key of each bin's span
  uint32_t* boff;     // kBinMax + 1: bucket offsets
  uint4* bt;          // kBinMax + 1: per bin {klo, word offset, first chunk, wide}
  uint32_t* cbin;     // kBinMax * kBinMaxChunks: bin of each chunk
  uint32_t* ctot;     // kBinMax * kBinMaxChunks: distinct keys per chunk
  uint32_t* info;     // 16: words, chunks, wide bins first/last, their bucket range, scan error
  uint64_t* scan_st;  // scan status words for the histogram
  uint32_t* scan_ctr; // scan tile counter
  uint32_t* hist;     // kBinMax * kBinSlices + 1: per (bin, slice) counts, bin-major, then scanned
  uint32_t* bkey;     // count: keys bucketed by bin
  uint32_t* bpay;     // count: payloads (written for the wide-bin range only)
};
size_t bin_meta_words();
void bin_work_carve(uint32_t* meta, BinWork* w);  // meta: bin_meta_words() u32, 16-byte aligned
// slices for `count` keys: *nslice workgroups of *slice keys each
void bin_slices(uint32_t count, uint32_t* nslice, uint32_t* slice);
hipError_t bin_range(const uint32_t* key, uint32_t count, float eb, double inv, const BinWork& w, hipStream_t st);
hipError_t bin_count(const uint32_t* key, uint32_t count, float eb, double inv, int bmin, uint32_t B,
                     const BinWork& w, hipStream_t st);
hipError_t bin_plan(uint32_t B, uint32_t nslice, const BinWork& w, hipStream_t st);
hipError_t bin_scatter(const uint32_t* key, const uint32_t* pay, uint32_t count, float eb, double inv, int bmin,
                       uint32_t B, const BinWork& w, hipStream_t st);
hipError_t bin_bitmap(uint32_t nchunks, const BinWork& w, uint2* gw, hipStream_t st);
// chunk totals -> per-bin exclusive prefix over the bin's chunks (before bin_query)
hipError_t bin_chunk_scan(uint32_t B, const BinWork& w, hipStream_t st);
hipError_t bin_query(const uint32_t* key, const uint32_t* pay, uint32_t count, float eb, double inv, int bmin,
                     uint32_t B, const BinWork& w, const uint2* gw, int32_t* out, uint32_t out_cap, hipStream_t st);

// ---- block offsets of streams without a side-car (tszp_parse.hip) ----
constexpr uint32_t kParseSeg = 8192;       // payload bytes per segment
constexpr uint32_t kParseMaxBlock = 2048;  // largest block the device parse handles
constexpr uint32_t kParseGroup = 64;       // tables per composition group
struct ParseArgs {
  const uint8_t* p;  // payload (after the chunk table)
  uint64_t len;      // payload bytes readable
  uint16_t sz[34];   // size of a full block by bit count; [33]: an impossible bit count
  uint32_t maxb;     // largest full-block size (entry offsets into a segment)
  uint32_t rounds;   // pointer-doubling rounds: 2^rounds >= kParseSeg / smallest size + 1
  int has_bc;        // full blocks carry a bit-count byte
  int topo;          // blocks carry 2-bit types
  // block layout (SZp blocks: anchor, bit count at +4, cur - 1 deltas); the units of the
  // element-partitioned streams differ (tszp_elem.hip)
  uint32_t bc_off;   // offset of the bit-count byte in a block
  int anchor;        // a block starts with a 4-byte anchor
  int dm1;           // a block of cur elements holds cur - 1 deltas (else cur)
  int bc_always;     // the bit-count byte is there also without deltas
  uint32_t out_base; // added to every offset written
  uint32_t* end_out; // nullable: receives the end offset (+ out_base) of the last block
};
size_t parse_work_bytes(uint64_t len, uint32_t maxb);
// err: 2 zeroed u32 (error bits, "last block reached")
hipError_t parse_blocks(const ParseArgs& a, uint64_t n, uint32_t bs, uint32_t nb, uint32_t* blk, void* work,
                        uint32_t* err, hipStream_t st);
// element-partitioned streams: blk[t * ukc] = off[t] for every threadblock chunk (whose
// first unit holds the chunk's anchor) with its end in ends[t] when the chunk has one
// element or none, then err |= TSZP_ERR_STREAM unless every chunk's walk ended where the
// next chunk starts (ends[t] == off[t + 1]) and the last one within `avail` bytes; tab:
// the stream's u64 chunk offsets (device)
hipError_t parse_elem_finish(const uint8_t* tab, uint32_t T, uint64_t n, uint32_t ukc, bool ra, uint64_t avail,
                             uint32_t* blk, uint32_t* ends, uint32_t* err, hipStream_t st);
hipError_t parse_check_offsets(const uint8_t* head, uint32_t T, const uint32_t* blk, uint32_t nb, uint32_t* err,
                               hipStream_t st);

// ---- quality metrics (tszp_quality.hip); word layout of tszp_quality ----
constexpr int kQualThresholds = 6;
enum : int {
  kQOrig = 0, kQDec = 4, kQPreserved = 8, kQNew = 9, kQLost = 10, kQChanged = 14, kQOver1 = 18, kQOver2 = 19,
  kQSaddles = 20, kQAbove = 21, kQAboveKept = 27, kQPairs = 33, kQViol = 37, kQualCounts = 41,
  kQualWords = 48  // counts, then 4 doubles (max error, persistence min / max / sum), padding
};
hipError_t launch_quality(const Grid& g, const float* x, const float* y, uint32_t n, float eb, uint64_t* part,
                          uint32_t nblocks, uint64_t* out, hipStream_t st);

// ---- topology restoration (tszp_post.hip) ----
// Control block: uint32 words (kPpMiscWords) shared by the restoration kernels.
constexpr int kPpMiscWords = 80;
enum : int {
  kPpBad = 0, kPpMoves = 1, kPpAct = 2, kPpTiles = 3, kPpVerify = 4, kPpErr = 12, kPpExt = 13, kPpSpMax = 14,
  kPpSolved = 15, kPpFailed = 16, kPpCounter = 17, kPpFreed = 18, kPpPlatAnneal = 19,
  kPpActC = 20,      // 7 colour-list counts
  kPpMcSolved = 27, kPpMcFailed = 28, kPpMcBig = 29, kPpMcMaxTicks = 30, kPpMcMaxSteps = 31,  // min-conflicts plateau solves
  kPpTilesP = 32,    // 8 parity-phase tile-list counts
  kPpStatNodes = 40, kPpStatMaxNodes = 41, kPpStatMaxWin = 44, kPpStatMaxTicks = 45, kPpStatMaxWait = 46, kPpStatLeaves = 47, kPpTiny = 42,
  kPpSolveCounter = 43,
  kPpPlatMaxTicks = 52, kPpPlatMaxAnneal = 53, kPpPlatMaxSweeps = 54,  // whole-plateau solve statistics
  kPpStatSumBuild = 55, kPpStatSumAttempt = 56, kPpStatSumTile = 57,    // window solve: summed ticks
  kPpFlowCounter = 60,  // tile counter of the one-launch solve
  kPpPlatSolved = 61, kPpPlatFailed = 62, kPpPlatBig = 63,  // whole-plateau solves
  kPpMcTicks = 64,  // min-conflicts statistics: summed ticks of draw + records, candidates, move, renormalisation
  kPpMcSteps = 68,  // ... and summed steps
  kPpMcEsc = 72     // plateaus a first search left to the replicas
};
// mismatches left after a local round below which the window solve runs at once
constexpr uint32_t kPpSolveFew = 16384;
constexpr uint32_t kPpLocalFew = 64;  // listed mismatches solved by windows without a local round first
uint32_t pp_init_tiles(uint32_t n);
// largest sort position: one field into misc[kPpSpMax], a batch into geo.spmax
hipError_t pp_spmax(const PpGeo& geo, const int32_t* sp, uint32_t m, uint32_t* misc, hipStream_t st);
// start values; work: pp_init_work_words(geo.n) u32 (8-byte aligned), no initialisation
// needed; n_ext: extrema of one field (a batch reads each field's count on the device)
size_t pp_init_work_words(uint32_t n);
hipError_t pp_init(const PpGeo& geo, const float* d, const uint8_t* t, const int32_t* sp, uint32_t n_ext, float* y,
                   uint32_t* work, uint32_t* misc, hipStream_t st);
// whole-plateau solves around the listed mismatches (one wavefront each); guide: the
// field whose order is tried first (nullable); claim: zeroed bits, one per point
// binomial smoothing of src along one axis into dst (radius 4, edges clamped); fflags
// (batch, nullable): one word per field, only flagged fields are smoothed
hipError_t pp_blur(const PpGeo& geo, const float* src, float* dst, int axis, const uint32_t* fflags,
                   hipStream_t st);
// a batch's field flags: 1 for the fields of the listed points, 0 for the others
hipError_t pp_field_flags(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, uint32_t* flags, hipStream_t st);
// full marking pass; part: 4 * pp_mark_parts() words of per-workgroup partials;
// bm (nullable) receives one bit per point (word e / 32, bit e % 32), 1 where mismatched
// (where pp_mark_maps; a batch's padding between fields is left unwritten);
// per_field: a batch's per-field counts into geo.fcount
hipError_t pp_mark(const PpGeo& geo, const float* y, const uint8_t* t, const float* d, uint32_t* bad, uint32_t cap,
                   uint32_t* part, uint32_t* bm, bool per_field, uint32_t* misc, hipStream_t st);
uint32_t pp_mark_parts();
bool pp_mark_maps(const PpGeo& geo);
hipError_t pp_active(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, const float* d, const uint8_t* t,
                     uint32_t* bits, uint32_t* act, uint32_t cap, int anchor, uint32_t* misc, hipStream_t st);
hipError_t pp_move(const PpGeo& geo, const uint32_t* act, uint32_t max_act, uint32_t cap, int colour, const float* d,
                   const uint8_t* t, float* y, uint32_t* misc, hipStream_t st);
hipError_t pp_remark(const PpGeo& geo, const uint32_t* old, uint32_t nold, const float* y, const uint8_t* t,
                     uint32_t* bits, uint32_t* bad, uint32_t cap, int rings, uint32_t* misc, hipStream_t st);
// plateaus of two or three listed points solved exactly (relations of the decoded
// values d computed on the fly; bm: the mismatch bitmap, set at the listed points);
// solved counts go to misc[kPpTiny]; keep[i] (nbad words): bad[i] if it stays mismatched, else
// 0xffffffff -- after the solves, which change no other point's type, the exact list
hipError_t pp_tiny(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, const float* d, const uint32_t* bm,
                   const uint8_t* t, float* y, int anchor, uint32_t* keep, uint32_t* part, uint32_t* misc,
                   hipStream_t st);
uint32_t pp_tiny_parts(uint32_t nbad);  // partial-count words pp_tiny needs
// the mismatch list in index order from the mismatch bitmap of n points (buffer 16-byte
// aligned and readable up to n rounded up to 128 points); status: pp_compact_tiles(n)
// words of 8 bytes; the count goes to misc[kPpBad]
uint32_t pp_compact_tiles(uint32_t n);
hipError_t pp_compact_map(const uint32_t* bm, uint32_t n, uint64_t* status, uint32_t* out, uint32_t cap,
                          uint32_t* misc, hipStream_t st);
// mismatch bitmap (one bit per point) from a list; the bitmap must be zeroed first
hipError_t pp_badmap(const uint32_t* bad, uint32_t nbad, uint32_t* bm, hipStream_t st);
// the entries of a sparse list (0xffffffff: empty) appended to out (count misc[kPpBad])
hipError_t pp_relist(const uint32_t* sparse, uint32_t n, uint32_t* out, uint32_t cap, uint32_t* misc,
                     hipStream_t st);
void pp_tile_counts(const Grid& g, uint32_t* tx, uint32_t* ty, uint32_t* tz);
hipError_t pp_tiles(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, uint32_t* pbits, uint32_t* tbits,
                    uint32_t* tiles, uint32_t cap, uint32_t* misc, hipStream_t st);
// (large: the variant for plateaus of up to 3072 points, 150 KB of LDS per wavefront)
hipError_t pp_plateau(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, const float* d, const uint8_t* t, float* y,
                      const float* guide, uint32_t* claim, bool large, uint32_t* misc, hipStream_t st);
// whole-plateau min-conflicts solves around the listed mismatches (tszp_platmc.hip), one
// wavefront each: guide (nullable) the start order, else breadth-first distances inside
// the plateau; claim: one bit per point, zeroed before the first variant and kept across
// the variants, so that a plateau one variant took is not searched again; variant 0, 1,
// 2: plateaus of up to 64, 512, 3072 points (4, 23, 150 KB of LDS per wavefront; a
// larger plateau counts in misc[kPpMcBig] for the next variant); budget: steps per
// plateau; relist (cap entries, counter misc[kPpBad]): every point of a searched
// plateau that is mismatched afterwards, and with variant 2 the listed mismatches of
// plateaus beyond it -- with all variants run over an exact list, relist is exact;
// stats: solved / failed counts and the largest step count and time in misc; esc
// (nullable, nbad entries): a plateau the budget does not finish is not relisted but
// its smallest point goes to esc (count misc[kPpMcEsc]), once per plateau
hipError_t pp_platmc(const PpGeo& geo, const uint32_t* bad, uint32_t nbad, const float* d, const uint8_t* t, float* y,
                     const float* guide, uint32_t* claim, int variant, uint32_t budget, uint32_t* relist,
                     uint32_t cap, bool stats, uint32_t* misc, uint32_t* esc, hipStream_t st);
uint32_t pp_platmc_capacity(int variant);  // points of the variant's largest plateau
// the plateaus of ent (smallest points from pp_platmc's esc, searched there by the same
// variant) searched again by R (1..8) replicas each, budget steps, the first to finish
// -- fewest steps, then lowest replica -- writing the values; stage: nent * R *
// pp_platmc_capacity(variant) floats; ctl: 2 * nent words, the first nent ~0u and the
// others 0; what is left mismatched goes to relist as with pp_platmc
hipError_t pp_platmc_replicas(const PpGeo& geo, const uint32_t* ent, uint32_t nent, const float* d, const uint8_t* t,
                              float* y, const float* guide, int variant, uint32_t budget, uint32_t R, float* stage,
                              uint32_t* ctl, uint32_t* relist, uint32_t cap, bool stats, uint32_t* misc,
                              hipStream_t st);
// window solves (one launch per parity phase, or every phase in one launch below);
// anchor: bit 0 keeps the anchored extrema in place, bit 1 solves only windows that hold
// their whole plateau (the larger ones are left to the plateau solves)
hipError_t pp_solve(const PpGeo& geo, const uint32_t* tiles, uint32_t ntiles, uint32_t cap, int phase,
                    const uint32_t* pbits, const float* d, const uint8_t* t, float* y, int anchor, uint32_t* misc,
                    hipStream_t st);
// every phase in one launch (dataflow over neighbouring tiles): `total` listed tiles,
// tbits the listed-tile bitmap, done one zeroed word per tile, misc[kPpFlowCounter] zeroed
hipError_t pp_solve_flow(const PpGeo& geo, const uint32_t* tiles, uint32_t total, uint32_t cap, const uint32_t* tbits,
                         uint32_t* done, const uint32_t* pbits, const float* d, const uint8_t* t, float* y, int anchor,
                         uint32_t* misc, hipStream_t st);

// ---- synthetic fields (tszp_gen.hip) ----
hipError_t launch_gen_blobs(float* f, int ndim, const uint32_t* dims, uint64_t seed, int nblobs,
                            float smin, float smax, float noise, int waves, hipStream_t st);
hipError_t launch_gen_fbm(float* f, const uint32_t* dims, uint64_t seed, int octaves, float noise,
                          hipStream_t st);

}  // namespace tszp
