/*
 * toposzp_hip.h -- native (device-pointer) API of libtoposzp_hip.so for MI355X.
 *
 * The reference has no counterpart: its CUDA entry points take host pointers and
 * copy every buffer over PCIe (SZp/src/szp_cuda_topology.cu:93-121).  These entry
 * points keep fields, streams and metadata resident in HBM and run on a caller's
 * HIP stream; the drop-in host-pointer API (szp_cuda_*.h) is built on them.
 *
 * Conventions
 *   - d_* arguments are device pointers, 4-byte aligned.  Stream buffers must stay
 *     readable up to the next multiple of 4 bytes past their end (size them with
 *     the *_bound functions, which include that padding).
 *   - `stream` is a hipStream_t (NULL = the null stream).  Calls are asynchronous
 *     except where a result size has to reach the host (compress calls return
 *     after their sizes are known).
 *   - Every call returns 0 on success or a negative TSZP_E* code; the message is
 *     in tszp_last_error() (per host thread).
 *   - Block-offset side-car: d_*_blk[b] is the byte offset of block b from the
 *     first byte after the chunk-offset header, for b = 0..blocks, entry [blocks]
 *     being the payload size.  The compressors write it; the decompressors read it,
 *     so decoding needs no serial walk.
 */
#ifndef TOPOSZP_HIP_H
#define TOPOSZP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSZP_OK 0
#define TSZP_EINVAL (-1)   /* invalid argument */
#define TSZP_EHIP (-2)     /* HIP runtime error */
#define TSZP_ECAP (-3)     /* output buffer too small */
#define TSZP_EDEVICE (-4)  /* a kernel reported a corrupt stream or a look-back timeout */

typedef struct {
  uint64_t topo_bytes;     /* [u64 off[T]][blocks], the reference CPU topology stream */
  uint64_t sortpos_bytes;  /* [u64 E][u64 off[T]][blocks], the CPU sort-position stream; 0 if E == 0 */
  uint64_t cp_count;       /* critical points */
  uint64_t ext_count;      /* maxima + minima (E) */
} tszp_topo_info;

const char *tszp_last_error(void);
int tszp_version(void);

/* worst-case stream sizes (including header and read padding) */
size_t tszp_topo_bound(size_t n, int block_size, unsigned nchunks);
size_t tszp_sortpos_bound(size_t n_extrema, int block_size, unsigned nchunks);
size_t tszp_plain_bound(size_t n, int block_size, unsigned nchunks);
size_t tszp_blocks(size_t n, int block_size);

/* relative -> absolute error bound exactly as SZp/include/szp_errbound.h:34-51:
 * eb = rel * (max - min) in float, FLT_EPSILON when the range is not positive. */
int tszp_rel_to_abs(const float *d_field, size_t n, float rel, float *eb, float *range, void *stream);

/* Full topology compression of a 2D (dims = rows, cols) or 3D (d1, d2, d3) field:
 * critical points, per-bin sort positions, the sort-position stream and the
 * topology stream, byte-identical to the reference CPU path with nchunks threads
 * (SZp/src/szp_float.cc:28-1318; SZp/src/szp_topology_3d.cc:30-632).
 * d_sp_values (nullable, >= ext_count ints) receives the extrema sort positions.
 * info holds the sizes on return; the streams are complete once `stream` reaches the
 * point of the call (the last header bytes are written in stream order). */
int tszp_topo_compress(const float *d_field, int ndim, const uint32_t *dims, float eb, int block_size,
                       unsigned nchunks, uint8_t *d_topo, size_t topo_cap, uint32_t *d_topo_blk,
                       uint8_t *d_sortpos, size_t sortpos_cap, uint32_t *d_sortpos_blk,
                       int32_t *d_sp_values, tszp_topo_info *info, void *stream);

/* Topology stream -> (float)q * eb and the per-point types (0..3, uint8). */
int tszp_topo_decompress(const uint8_t *d_topo, const uint32_t *d_topo_blk, unsigned nchunks, size_t n,
                         float eb, int block_size, float *d_out, uint8_t *d_types, void *stream);

/* Sort-position stream (with its u64 count header) -> int sort positions. */
int tszp_sortpos_decompress(const uint8_t *d_sortpos, const uint32_t *d_sortpos_blk, unsigned nchunks,
                            size_t n_extrema, int block_size, int32_t *d_out, void *stream);

/* Plain random-access SZp (SZp/src/szp_float.cc:2061-2292, szpd_float.cc:431-524):
 * [f32 eb big-endian][u64 off[T]][blocks]. */
int tszp_plain_compress(const float *d_in, size_t n, float eb, int block_size, unsigned nchunks,
                        uint8_t *d_out, size_t cap, uint32_t *d_blk, uint64_t *out_bytes, void *stream);
int tszp_plain_decompress(const uint8_t *d_stream, const uint32_t *d_blk, unsigned nchunks, size_t n,
                          float eb, int block_size, float *d_out, void *stream);

/* tszp_plain_compress without any allocation, for callers that cannot allocate device
 * memory (code under stream capture, tight in-situ loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  Same arguments, bytes and side-car as
 * tszp_plain_compress; its scratch (look-back words, tile counter, results) lives in a
 * static device array of about 2 MiB per device, held until the process ends.  Fields
 * that need more than 2^17 tiles (more than 2^28 elements at the smallest tiles) are
 * rejected with TSZP_EINVAL.  Every call on a device shares that array, so calls that
 * overlap on one device race; the call returns after its stream has finished it. */
int tszp_plain_compress_noalloc(const float *d_in, size_t n, float eb, int block_size, unsigned nchunks,
                                uint8_t *d_out, size_t cap, uint32_t *d_blk, uint64_t *out_bytes, void *stream);

/* Batches of plain float fields: every field of a batch is compressed (decompressed) by
 * one launch per block-size class instead of one call per field; each field's stream
 * and side-car are byte-identical to tszp_plain_compress / tszp_plain_decompress with
 * nchunks 1: [f32 eb big-endian][u64 0][blocks].  No counterpart in the reference,
 * whose callers loop over szp_cuda_float_compress_randomaccess
 * (SZp/src/szp_cuda_compress.cu:718-723) one field at a time.
 * d_field: compression input / decompression output, n floats; d_stream: 4-byte
 * aligned, stream_cap >= tszp_plain_bound(n, block_size, 1) for compression; d_blk:
 * tszp_blocks(n, block_size) + 1 block offsets.  block_size 2 .. 1025; 0 asks
 * tszp_batch_tune to choose.  Errors name the first failing field (caller's index). */
typedef struct {
  float *d_field;
  uint64_t n;
  float eb;               /* absolute error bound */
  int block_size;
  uint8_t *d_stream;
  uint64_t stream_cap;
  uint32_t *d_blk;
  uint64_t stream_bytes;  /* set by tszp_plain_compress_batch */
} tszp_batch_field;
int tszp_plain_compress_batch(tszp_batch_field *fields, uint32_t count, void *stream);
int tszp_plain_decompress_batch(const tszp_batch_field *fields, uint32_t count, void *stream);
/* Block-size autotuner: for every field with block_size 0, estimates the stream size at
 * block sizes 16, 32, 64 and 128 from up to 512 windows of 128 elements spread evenly
 * over the field (one launch for all fields) and sets block_size to the smallest
 * estimate (64 on a tie or for fields shorter than 128). */
int tszp_batch_tune(tszp_batch_field *fields, uint32_t count, void *stream);

/* Batch plans: the layout of a batch (field table, tile map, look-back words, results)
 * kept on the device for repeated runs over the same fields (pointers, sizes, error
 * bounds and block sizes fixed; contents may change between runs).  The compress and
 * decompress calls only launch kernels on `stream` -- no allocation, no host copy, no
 * synchronisation -- so a caller may capture them in a hipGraph and replay it;
 * tszp_batch_plan_results waits for the stream and returns the per-field stream sizes
 * (caller's order) and the first device error of the last runs.  A plan is used by one
 * thread at a time, on the device that was current when it was created.  No
 * counterpart in the reference. */
typedef struct tszp_batch_plan tszp_batch_plan;
/* d_outs (nullable): decompression output of each field (n floats); NULL decompresses
 * into d_field */
int tszp_batch_plan_create(const tszp_batch_field *fields, uint32_t count, float *const *d_outs,
                           tszp_batch_plan **plan, void *stream);
int tszp_batch_plan_destroy(tszp_batch_plan *plan);
int tszp_batch_plan_compress(tszp_batch_plan *plan, void *stream);
int tszp_batch_plan_decompress(tszp_batch_plan *plan, void *stream);
int tszp_batch_plan_results(tszp_batch_plan *plan, uint64_t *stream_bytes, void *stream);

/* Request loop over the batch entry points (no counterpart in the reference).  Requests
 * from any host thread are admitted in arrival order while their streams fit the HBM
 * budget (one device arena of hbm_budget bytes, allocated at creation); every request
 * waiting when a batch finishes joins the next one, up to max_batch.  Inputs must be
 * complete on the device when submitted and stay valid until the ticket finishes;
 * outputs are valid after tszp_server_wait.  A compress ticket's stream and side-car
 * stay in the arena until tszp_server_release; a compression that does not fit waits
 * for releases (one larger than the budget fails).  Release every ticket when done
 * with it.  Errors: tszp_server_last_error(). */
typedef struct tszp_server tszp_server;
typedef struct {
  uint64_t stream_bytes;  /* compress tickets: [f32 eb big-endian][u64 0][blocks] */
  double queue_us;        /* submission to the start of its batch */
  double latency_us;      /* submission to the end of its batch */
} tszp_ticket_info;
typedef struct {
  uint64_t batches, requests, queued, hbm_in_use, hbm_peak, hbm_budget;
} tszp_server_stats_t;
const char *tszp_server_last_error(void);
int tszp_server_create(uint64_t hbm_budget, uint32_t max_batch, tszp_server **out);
/* serves what is queued, then stops; call it once no other thread uses the server */
int tszp_server_destroy(tszp_server *server);
/* block_size 0: chosen by the tuner of tszp_batch_tune when the request runs; the block
 * size used is returned by tszp_server_stream */
int tszp_server_compress(tszp_server *server, const float *d_field, uint64_t n, float eb, int block_size,
                         uint64_t *ticket);
int tszp_server_decompress(tszp_server *server, uint64_t compress_ticket, float *d_out, uint64_t *ticket);
int tszp_server_wait(tszp_server *server, uint64_t ticket, tszp_ticket_info *info);
int tszp_server_stream(tszp_server *server, uint64_t compress_ticket, const uint8_t **d_stream,
                       const uint32_t **d_blk, uint64_t *stream_bytes, int *block_size);
/* copies a finished compress ticket's stream into d_dst (cap bytes), ordered on `stream` */
int tszp_server_copy_stream(tszp_server *server, uint64_t compress_ticket, uint8_t *d_dst, uint64_t cap,
                            void *stream);
int tszp_server_release(tszp_server *server, uint64_t ticket);
int tszp_server_stats(tszp_server *server, tszp_server_stats_t *out);

/* Streams partitioned by elements: chunk t of nchunks covers [t * (n / nchunks), ...),
 * the last chunk up to n; header [eb big-endian, elem_bytes wide][u64 off[nchunks]].
 * d_blk receives tszp_elem_units(n, block_size, nchunks, ra) + 1 unit offsets relative
 * to the payload (the side-car the decoders take).  block_size 1 .. 1025, nchunks
 * 1 .. 4096, elem_bytes 4 (float data) or 8 (double data).
 *
 * Threadblock (non-random-access) codec: per chunk one i32 anchor, then blocks of up to
 * block_size deltas chained across blocks.  Replaces szp_float_openmp_threadblock /
 * szp_double_openmp_threadblock (SZp/src/szp_float.cc:1445, szp_double.cc:148; CUDA
 * szp_cuda_{float,double}_compress_threadblock, SZp/src/szp_cuda_compress.cu:752-782)
 * and their decoders (SZp/src/szpd_float.cc:28, szpd_double.cc:6; CUDA
 * szp_cuda_{float,double}_decompress_threadblock, szp_cuda_decompress.cu:1032-1060). */
size_t tszp_elem_units(size_t n, int block_size, unsigned nchunks, int ra);
/* stream capacity that always suffices */
size_t tszp_elem_bound(size_t n, int block_size, unsigned nchunks, int elem_bytes, int ra);
int tszp_threadblock_compress(const void *d_in, int elem_bytes, size_t n, double eb, int block_size,
                              unsigned nchunks, uint8_t *d_out, size_t cap, uint32_t *d_blk,
                              uint64_t *out_bytes, void *stream);
int tszp_threadblock_decompress(const uint8_t *d_stream, const uint32_t *d_blk, unsigned nchunks, size_t n,
                                double eb, int block_size, int elem_bytes, void *d_out, void *stream);

/* Double random-access codec: blocks of block_size elements from each chunk start, each
 * with its own i32 anchor and a bc byte (also for a one-element block).  Replaces
 * szp_double_openmp_threadblock_randomaccess (SZp/src/szp_double.cc:759) and decodes
 * what it writes (SZp/src/szpd_double.cc:427 reads it back for nchunks = 1). */
int tszp_plain_compress_f64(const double *d_in, size_t n, double eb, int block_size, unsigned nchunks,
                            uint8_t *d_out, size_t cap, uint32_t *d_blk, uint64_t *out_bytes, void *stream);
int tszp_plain_decompress_f64(const uint8_t *d_stream, const uint32_t *d_blk, unsigned nchunks, size_t n,
                              double eb, int block_size, double *d_out, void *stream);

/* Quantize and predict: d_out[i] = q[i] - q[i - 1] with q[-1] = 0 and
 * q = C cast of x * (1 / (double)eb) (the quantization helpers of
 * SZp/src/szp_cuda_compress.cu:360-377, :664-710; int32 wrap-around differences). */
int tszp_predict_quantization(const float *d_in, size_t n, float eb, int32_t *d_out, void *stream);

/* Double-precision 3D topology stream [u64 0][blocks] with the saddle fix
 * (szp_cuda_double_compress_topology_3d, SZp/src/szp_cuda_topology_3d.cu:1065-1188):
 * types marked from the critical-point records d_cps (cp_count records of 6 ints:
 * x, y, z, type, bin, sort position -- CriticalPoint3D), q = C cast of x * (1 / eb),
 * three Jacobi passes moving broken saddles by one step, then the topology block codec
 * on q with the marked types.  d_blk receives tszp_blocks(n, block_size) + 1 offsets;
 * d_q (nullable) the fixed quantized field; *moved (nullable) the saddles moved. */
int tszp_topo_compress_f64_3d(const double *d_field, const uint32_t *dims, double eb, int block_size,
                              const int32_t *d_cps, uint64_t cp_count, uint8_t *d_out, size_t cap,
                              uint32_t *d_blk, uint64_t *out_bytes, int32_t *d_q, uint64_t *moved,
                              void *stream);

/* Per-point critical-point types of a field (0 regular / boundary, 1 max, 2 min,
 * 3 saddle); *cp_count (nullable, host) gets the number of critical points. */
int tszp_classify(const float *d_field, int ndim, const uint32_t *dims, uint8_t *d_types,
                  uint64_t *cp_count, void *stream);

/* Topology restoration (no counterpart in the reference library: its post-process is
 * host code in example/test_cuda_topology.cu:118-326 and test_cuda_topology_3d.cu:77-434).
 * d_decoded holds the decoded (float)q * eb; the result goes to d_field, which may be
 * the same buffer (then a copy of the decoded values is kept internally).  Every point
 * stays in the interval of its cell -- [d, d + 0.9 eb] for d > 0, [d - 0.9 eb, d] for
 * d < 0 (error < eb), [-eb 2^-16, eb 2^-16] for d == 0 (error < eb (1 + 2^-16)) -- while
 * a colour-class local search and an exact solve of small plateau windows make the
 * critical-point type of every interior point equal to d_types.  d_sp (nullable):
 * extrema sort positions in row-major extremum order; with them the extrema of one
 * non-zero cell keep the order of their sort positions.  d_types must be 4-byte
 * aligned.  The counts returned are those of tszp_topo_verify on the result. */
typedef struct {
  uint64_t missing;        /* type != 0, restored field regular */
  uint64_t spurious;       /* type == 0, restored field critical */
  uint64_t changed;        /* both critical, different types */
  uint64_t moves;          /* accepted single-point moves */
  uint64_t windows;        /* plateau windows solved exactly */
  uint64_t windows_reordered; /* of which moved an anchored extremum freely (the order of
                                 its sort position may break: a boundary value whose
                                 sort position belongs to the neighbouring bin) */
  uint32_t rounds;         /* local-search rounds */
  float max_dev_over_eb;   /* max |y - d| / eb */
  uint32_t solve_tiles;    /* tiles of the last exact window solve */
  uint32_t solve_chain;    /* its longest chain of dependent tiles (one field, set while
                              the stage timers run: tszp_prof_enable) */
} tszp_restore_info;

int tszp_topo_restore(const float *d_decoded, float *d_field, const uint8_t *d_types, const int32_t *d_sp,
                      uint64_t n_extrema, int ndim, const uint32_t *dims, float eb, int max_rounds,
                      tszp_restore_info *info, void *stream);

/* Critical-point mismatches of a field against reference types:
 * counts[0] missing, counts[1] spurious, counts[2] changed type. */
int tszp_topo_verify(const float *d_field, const uint8_t *d_types, int ndim, const uint32_t *dims,
                     uint64_t *counts, void *stream);

/* Batched topology pipeline: many 2D / 3D fields of any sizes, error bounds and block
 * sizes through one launch per stage (per block-size class) instead of one call chain
 * per field.  No counterpart in the reference, whose callers run the per-field chain
 * szp_cuda_find_critical_points -> szp_cuda_sort_critical_points_by_original_data ->
 * szp_cuda_compress_sort_positions -> szp_cuda_float_compress_randomaccess_topology_
 * preserved (SZp/src/szp_cuda_topology.cu:77-772, example/test_cuda_topology.cu:374-530)
 * once per field.  Every field's streams and side-cars are byte-identical to
 * tszp_topo_compress with nchunks 1: topology [u64 0][blocks], sort positions
 * [u64 E][u64 0][blocks] (the header is written also when E == 0, when sortpos_bytes
 * is 0).  A batch object keeps the layout and the workspace on the device: compress and
 * decompress are kernel launches only (no allocation, host copy or synchronisation), so
 * a caller may capture them in a hipGraph; restore synchronises a few times per batch
 * (not per field); results reads the sizes and counts back.  Decoded values, types,
 * sort positions and restored values live in the batch's own device arrays
 * (tszp_topo_batch_output).  A batch is used by one thread at a time, on the device
 * that was current when it was created.  Fields need |x / eb| < 2^31 (the
 * one-field call also takes larger values).  Errors name the first failing field by its
 * index in the caller's array. */
typedef struct {
  const float *d_field;      /* compression input (may be NULL for decompress-only batches) */
  int ndim;                  /* 2 or 3 */
  uint32_t dims[3];          /* dims[0] slowest */
  float eb;                  /* absolute error bound */
  int block_size;            /* 1 .. 257; 0: chosen as tszp_batch_tune chooses */
  uint8_t *d_topo;           /* topology stream, topo_cap >= tszp_topo_bound(n, bs, 1), 4-byte aligned */
  uint64_t topo_cap;
  uint32_t *d_topo_blk;      /* tszp_blocks(n, bs) + 1 block offsets */
  uint8_t *d_sortpos;        /* sort-position stream, sortpos_cap >= tszp_sortpos_bound(n, bs, 1) */
  uint64_t sortpos_cap;
  uint32_t *d_sortpos_blk;   /* tszp_blocks(n, bs) + 1 block offsets */
  tszp_topo_info info;       /* set by tszp_topo_batch_results after a compression */
  tszp_restore_info restore; /* set by tszp_topo_batch_results after a restoration (per field:
                                missing / spurious / changed, max_dev_over_eb; the batch's
                                moves, windows and rounds in every entry) */
} tszp_topo_batch_field;
typedef struct tszp_topo_batch tszp_topo_batch;
/* fields[i].block_size 0 is replaced by the chosen block size */
int tszp_topo_batch_create(tszp_topo_batch_field *fields, uint32_t count, tszp_topo_batch **batch, void *stream);
int tszp_topo_batch_destroy(tszp_topo_batch *batch);
int tszp_topo_batch_compress(tszp_topo_batch *batch, void *stream);
/* topology + sort-position streams -> decoded values (float)q * eb, types, sort positions */
int tszp_topo_batch_decompress(tszp_topo_batch *batch, void *stream);
/* tszp_topo_restore of every decoded field into the batch's restored values */
int tszp_topo_batch_restore(tszp_topo_batch *batch, int max_rounds, void *stream);
/* waits for `stream`; fills info (after compress) and restore (after restore) of every field */
int tszp_topo_batch_results(tszp_topo_batch *batch, tszp_topo_batch_field *fields, void *stream);
/* device arrays of field i (n elements each): decoded values, restored values, types,
 * sort positions (info.ext_count of them); any may be NULL */
int tszp_topo_batch_output(tszp_topo_batch *batch, uint32_t i, float **d_decoded, float **d_restored,
                           uint8_t **d_types, int32_t **d_sortpos);
/* copies one of those arrays of field i into d_dst, ordered on `stream`: which 0 decoded
 * values (n floats), 1 restored values (n floats), 2 types (n bytes), 3 sort positions
 * (ext_count ints; `count` elements are copied, at most n) */
int tszp_topo_batch_copy(tszp_topo_batch *batch, uint32_t i, int which, void *d_dst, uint64_t count, void *stream);
/* The same batch in the caller's device memory (256-byte aligned, mem_bytes of it): no
 * allocation on the device, so nothing synchronises the device when the batch comes and
 * goes (what a request loop that keeps several batches in flight needs).  Fails with
 * TSZP_ECAP when the layout needs more than mem_bytes; tszp_topo_batch_mem_bound gives
 * enough for any batch of up to `count` fields and `elements` elements in all.  The
 * memory must outlive the batch. */
int tszp_topo_batch_create_mem(tszp_topo_batch_field *fields, uint32_t count, void *d_mem, uint64_t mem_bytes,
                               tszp_topo_batch **batch, void *stream);
uint64_t tszp_topo_batch_mem_bound(uint64_t elements, uint32_t count);

/* A restoration workspace of the caller's own (device buffers that grow on demand and
 * stay until destroyed).  The library's calls otherwise share one workspace per device
 * and run one at a time; restorations with different workspaces, on different streams,
 * run concurrently.  A workspace serves one call at a time, on the device that was
 * current when it was created. */
typedef struct tszp_workspace tszp_workspace;
int tszp_workspace_create(tszp_workspace **ws);
int tszp_workspace_destroy(tszp_workspace *ws);
/* tszp_topo_batch_restore with the given workspace (NULL: the device's own) */
int tszp_topo_batch_restore_ws(tszp_topo_batch *batch, int max_rounds, tszp_workspace *ws, void *stream);

/* Request loop of the topology pipeline (no counterpart in the reference, whose callers
 * run the per-field chain once per field: example/test_cuda_topology.cu:374-530).
 * Compress requests (2D / 3D fields) and decompress requests (decode, then the topology
 * restoration, of a finished compress ticket's streams) from any host thread are
 * admitted in arrival order while their streams fit the HBM budget; `lanes` worker
 * threads, each with its own stream, batch memory for up to max_batch requests and
 * max_batch_elements elements, and restoration workspace, take turns at the queue, so
 * that the batch formed from what arrived during the previous one runs on the device
 * beside it.  The budget covers the lanes' working memory (reserved at creation) and one
 * device arena for the compress tickets' streams and side-cars (tszp_topo_bound and
 * tszp_sortpos_bound of their block size, or of the largest candidate with block size
 * 0).  A request larger than a batch, or than the arena, fails; one that does not fit
 * yet waits for releases.  Streams are byte-identical to tszp_topo_compress with
 * nchunks 1 and decoded values to tszp_topo_decompress; restored values meet the same
 * invariants as tszp_topo_restore (the restoration's rounds depend on the batch).
 * Inputs must be complete on the device when submitted and stay valid until the ticket
 * finishes; outputs are valid after tszp_topo_server_wait.  Release every ticket when
 * done with it.  Errors: tszp_topo_server_last_error(). */
typedef struct tszp_topo_server tszp_topo_server;
const char *tszp_topo_server_last_error(void);
int tszp_topo_server_create(uint64_t hbm_budget, uint32_t max_batch, uint64_t max_batch_elements, uint32_t lanes,
                            tszp_topo_server **out);
/* serves what is queued, then stops; call it once no other thread uses the server */
int tszp_topo_server_destroy(tszp_topo_server *server);
int tszp_topo_server_compress(tszp_topo_server *server, const float *d_field, int ndim, const uint32_t *dims,
                              float eb, int block_size, uint64_t *ticket);
/* d_out: n floats, restored values (restore != 0) or decoded values; d_types: n bytes
 * of types, or NULL */
int tszp_topo_server_decompress(tszp_topo_server *server, uint64_t compress_ticket, float *d_out, uint8_t *d_types,
                                int restore, uint64_t *ticket);
/* stream_bytes of a compress ticket: topology + sort-position stream bytes */
int tszp_topo_server_wait(tszp_topo_server *server, uint64_t ticket, tszp_ticket_info *info);
/* a finished ticket's compression results (of its source, for a decompress ticket), its
 * restoration counts (decompress tickets) and block size; any may be NULL */
int tszp_topo_server_result(tszp_topo_server *server, uint64_t ticket, tszp_topo_info *info,
                            tszp_restore_info *restore, int *block_size);
/* device pointers into the arena of a finished compress ticket's streams and side-cars */
int tszp_topo_server_streams(tszp_topo_server *server, uint64_t compress_ticket, const uint8_t **d_topo,
                             const uint32_t **d_topo_blk, const uint8_t **d_sortpos,
                             const uint32_t **d_sortpos_blk);
int tszp_topo_server_release(tszp_topo_server *server, uint64_t ticket);
/* hbm figures include the lanes' reserved working memory */
int tszp_topo_server_stats(tszp_topo_server *server, tszp_server_stats_t *out);

/* Topology quality of a decompressed field y against the original x, as the
 * reference's CUDA driver reports it (example/test_cuda_topology.cu:540-868), for 2D
 * and 3D fields, in one GPU pass.  Critical points follow the codec's max / min /
 * saddle rule on interior points. */
typedef struct {
  uint64_t orig_cp[4];          /* original field: total, maxima, minima, saddles */
  uint64_t dec_cp[4];           /* decompressed field: the same */
  uint64_t preserved;           /* original CPs with the same type in y */
  uint64_t new_cps;             /* CPs of y at regular points of x */
  uint64_t lost[4];             /* original CPs that became regular: total, by type 1..3 */
  uint64_t changed[4];          /* original CPs of another type in y: total, by type */
  uint64_t over_1eb, over_2eb;  /* points with |x - y| > 1.01 eb, > 2 eb (in double) */
  uint64_t saddles;             /* saddles of x */
  uint64_t saddles_above[6];    /* persistence proxy min |x_c - x_nb| > 0, .5, 1, 2, 5, 10 eb */
  uint64_t saddles_kept_above[6]; /* of those, saddles that are saddles in y */
  uint64_t pairs[4];            /* adjacent interior pairs with x_a != x_b: all, >= 1 CP,
                                   both CPs, >= 1 extremum (forward neighbour per axis) */
  uint64_t violations[4];       /* of those, pairs whose order in y differs */
  double max_err;               /* max |x - y| */
  double pers_min, pers_max, pers_sum; /* persistence proxy over the saddles of x */
} tszp_quality;

int tszp_quality_metrics(const float *d_x, const float *d_y, int ndim, const uint32_t *dims, float eb,
                         tszp_quality *q, void *stream);

/* Block offsets of a device-resident stream that has no side-car (a stream read from a
 * file written by the reference), parsed on the device: replaces the host walk of
 * SZp/src/szp_cuda_decompress.cu:391-477 and detects the chunk count T as
 * SZp/include/szp_detect_threads.h:229-309 (first T whose offsets are non-decreasing,
 * whose blocks walk within the buffer, and whose chunk offsets fall on block starts).
 * kind: 0 topology stream [u64 off[T]][blocks], 1 sort-position stream
 * [u64 E][u64 off[T]][blocks] (n = E), 2 plain stream [f32 eb][u64 off[T]][blocks].
 * len: readable bytes of d_stream.  d_blk receives tszp_blocks(n, bs) + 1 offsets
 * relative to the payload, ready for the decoders above; *T the chunk count.
 * Block sizes whose largest block exceeds 2048 bytes are not supported (TSZP_EINVAL). */
int tszp_parse_blocks(const uint8_t *d_stream, size_t len, size_t n, int block_size, int kind, uint32_t *d_blk,
                      unsigned *nchunks, void *stream);

/* Unit offsets of a device-resident element-partitioned stream (threadblock codec, or
 * double random access with ra = 1) that arrives without its side-car, parsed on the
 * device: replaces the host walk of the reference's threadblock decoders
 * (SZp/src/szp_cuda_decompress.cu:480-525, :907-970) and detects the chunk count as the
 * drop-in readers do (the largest T in 2..128 whose chunk offsets are non-decreasing
 * and whose chunks walk, each ending where the next one starts; else 1).  d_stream: the
 * whole stream [eb, elem_bytes wide][u64 off[T]][payload], len readable bytes of it;
 * d_blk receives tszp_elem_units(n, block_size, T, ra) + 1 unit offsets relative to the
 * payload, what tszp_threadblock_decompress / tszp_plain_decompress_f64 take (room for
 * tszp_elem_units(n, block_size, 1, ra) + 130 serves every T up to 128); *T the chunk
 * count.  Units whose largest size exceeds 2048 bytes are not supported. */
int tszp_elem_parse(const uint8_t *d_stream, size_t len, size_t n, int block_size, int elem_bytes, int ra,
                    uint32_t *d_blk, unsigned *T, void *stream);

/* .SZp container written by the reference's CPU compress driver and read by its
 * decompress driver: [topology stream][sort-position stream][u64 topology bytes]
 * [u64 sort-position bytes] (example/testfloat_compress_fastmode1_topology.cc:105-139,
 * example/testfloat_decompress_fastmode1_topology.cc:722-816).  The topology stream is
 * the native [u64 off[T]][blocks], the sort-position stream the native
 * [u64 E][u64 off[T]][blocks].  Buffers may be host or device memory.
 * tszp_szp_pack needs cap >= topo_bytes + sp_bytes + 16 and returns after the copies.
 * tszp_szp_unpack validates the footer as the reference does (otherwise the whole
 * buffer is one topology stream) and reads E from the sort-position stream. */
int tszp_szp_pack(const uint8_t *topo, size_t topo_bytes, const uint8_t *sp, size_t sp_bytes, uint8_t *out,
                  size_t cap, size_t *out_bytes, void *stream);
int tszp_szp_unpack(const uint8_t *buf, size_t len, size_t *topo_bytes, size_t *sp_offset, size_t *sp_bytes,
                    uint64_t *n_extrema);

/* Kernel timing with HIP events on the launch stream (off by default).  Kinds:
 * 0 topology codec (classify + pack), 1 int codec (sort positions), 2 plain codec,
 * 3 decoders, 4 sort-position bucketing (bins, scatter, bitmaps; or the radix sort),
 * 5 rank queries (and the radix sort of the bins next to zero; or the rank scan),
 * 6 topology restoration (whole call), 7 each scatter launch (bin scatter or radix
 * pass), 8 each window-solve launch, 9 each bin rank-query launch, 10 threadblock /
 * double random-access compression, 11 their decoder, 12 double 3D quantization and
 * saddle fix, 13 batch compression, 14 batch decompression.  tszp_prof_read
 * returns the summed milliseconds and the number of timed launches since
 * tszp_prof_reset. */
void tszp_prof_enable(int on);
void tszp_prof_reset(void);
int tszp_prof_read(int kind, double *ms, uint64_t *launches);

/* Synthetic fields from a seed (device generated). */
int tszp_gen_blobs(float *d_field, int ndim, const uint32_t *dims, uint64_t seed, int nblobs,
                   float sigma_min, float sigma_max, float noise, int waves, void *stream);
int tszp_gen_fbm(float *d_field, const uint32_t *dims, uint64_t seed, int octaves, float noise,
                 void *stream);

#ifdef __cplusplus
}
#endif

#endif
