/*
 * oracle.h -- CPU restatement of the reference TopoSZp OpenMP path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls this code;
 * it exists so that tests/ and bench.py's cpu_baseline leg can check the HIP path
 * against an independent CPU statement of the reference algorithm.  Every function
 * cites the reference file:line it restates.
 *
 * Semantics are those of the reference run with OMP_NUM_THREADS = T (T = 1 gives the
 * canonical byte stream: a single zero chunk offset and row-major critical-point order).
 */
#ifndef TOPOSZP_ORACLE_H
#define TOPOSZP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layouts as CriticalPoint (SZp/include/szp_float.h:20-25) and
 * CriticalPoint3D (SZp/include/szp_topology_3d.h:22-27). */
typedef struct { int x, y, type, quantized_bin, sort_position; } or_cp2;
typedef struct { int x, y, z, type, quantized_bin, sort_position; } or_cp3;

/* bit count of a block's largest |delta|: floorf(log2f((float)m)) + 1, 0 for m == 0. */
unsigned or_bitcount(unsigned maxv);
/* Checks bc(x) == 32 - clz(x) + (x >= thr[31 - clz(x)]) for every x in [lo, hi). */
uint64_t or_check_bitcount_table(const uint32_t *thr, uint64_t lo, uint64_t hi);

/* x86 cvttsd2si semantics: truncation toward zero, INT_MIN when out of range / NaN. */
int or_cvt_i32(double d);

/* Relative -> absolute error bound (SZp/include/szp_errbound.h:34-51). */
int or_rel_to_abs(const float *data, size_t n, float rel, float *eb, float *range);

/* Critical points (row-major order, T = 1).  `out` holds (rows-2)*(cols-2) entries. */
size_t or_find_cps_2d(const float *data, int rows, int cols, float eb, or_cp2 *out);
size_t or_find_cps_3d(const float *data, int d1, int d2, int d3, float eb, or_cp3 *out);

/* Per-bin dense rank of original values (sort_position), CP array order unchanged. */
void or_sort_positions_2d(or_cp2 *cps, size_t n, const float *data, int cols);
void or_sort_positions_3d(or_cp3 *cps, size_t n, const float *data, int d2, int d3);

/* Chunk partition used by the OpenMP writers: blocks [*b0, *b1) of chunk t. */
void or_chunk_range(size_t nb, unsigned T, int adaptive, unsigned t, size_t *b0, size_t *b1);

/* Worst-case byte sizes of the streams below. */
size_t or_topo_bound(size_t n, int bs, unsigned T);
size_t or_plain_bound(size_t n, int bs, unsigned T);
size_t or_sortpos_bound(size_t e, int bs, unsigned T);

/* Topology stream [u64 off[T]][blocks] (SZp/src/szp_float.cc:998-1318). */
size_t or_topo_compress(const float *data, size_t n, float eb, int bs,
                        const unsigned char *types, unsigned T, int adaptive,
                        unsigned char *out);
/* Type map from a CP list (SZp/src/szp_float.cc:1040-1064). */
void or_mark_types_2d(const or_cp2 *cps, size_t n, int rows, int cols, unsigned char *types);
void or_mark_types_3d(const or_cp3 *cps, size_t n, int d1, int d2, int d3, unsigned char *types);

/* Sort-position stream [u64 E][u64 off[T]][blocks] (SZp/src/szp_float.cc:521-996). */
size_t or_compress_sortpos(const int *sp, size_t e, int bs, unsigned T, int adaptive,
                           unsigned char *out);
/* Extrema sort positions in CP-array order (SZp/src/szp_float.cc:529-582). */
size_t or_extrema_sortpos_2d(const or_cp2 *cps, size_t n, int *sp);
size_t or_extrema_sortpos_3d(const or_cp3 *cps, size_t n, int *sp);

/* Plain random-access stream [f32 eb big-endian][u64 off[T]][blocks]
 * (SZp/src/szp_float.cc:2061-2292). */
size_t or_plain_compress(const float *data, size_t n, float eb, int bs, unsigned T,
                         int adaptive, unsigned char *out);

/* Chunk-count detection from the offset table (SZp/include/szp_detect_threads.h). */
unsigned or_detect_T_topo(const unsigned char *cmp, size_t n, int bs);
unsigned or_detect_T_ra(const unsigned char *cmp, size_t n, int bs, size_t elem);

/* Decoders (SZp/src/szpd_float.cc). `cmp` starts at the chunk offset table. */
void or_topo_decompress(const unsigned char *cmp, size_t n, float eb, int bs,
                        float *out, int *types);
void or_sortpos_decompress(const unsigned char *cmp, size_t e, int bs, int *out);
void or_plain_decompress(const unsigned char *cmp, size_t n, float eb, int bs, float *out);
void or_plain_decompress_T(const unsigned char *cmp, size_t n, float eb, int bs, unsigned T, float *out);

/* Byte offset (from payload start) of every block of a T = 1 stream: a serial walk. */
int or_block_offsets(const unsigned char *payload, size_t n, int bs, int topo,
                     int bc_always, uint64_t *offs);

/* Classification of a reconstructed field (example/test_cuda_topology.cu:56-68,
 * example/test_cuda_topology_3d.cu:52-73); boundary points get 0. */
void or_classify_2d(const float *data, int rows, int cols, unsigned char *types);
void or_classify_3d(const float *data, int d1, int d2, int d3, unsigned char *types);

/* Element-partitioned streams [eb big-endian][u64 off[T]][chunks], chunk t covering
 * [t * (n / T), ...): ra = 0 the threadblock codec (SZp/src/szp_float.cc:1445-1590,
 * SZp/src/szp_double.cc:148-318), ra = 1 the double random-access codec
 * (SZp/src/szp_double.cc:759-920). dsize 4 = float data, 8 = double data. */
size_t or_elem_bound(size_t n, int bs, unsigned T, int dsize, int ra);
size_t or_elem_compress(const void *data, int dsize, size_t n, double eb, int bs, unsigned T, int ra,
                        unsigned char *out);
int or_elem_decompress(const unsigned char *cmp, size_t n, double eb, int bs, unsigned T, int ra, int dsize,
                       void *out);

/* Double 3D topology stream with the saddle fix (SZp/src/szp_cuda_topology_3d.cu:736-826,
 * :1065-1188). */
size_t or_fix_saddles_3d_pass(const int *src, int *dst, const unsigned char *types, int d1, int d2, int d3);
size_t or_topo3d_f64_compress(const double *data, int d1, int d2, int d3, double eb, int bs,
                              const unsigned char *types, unsigned char *out, int *q_out);

#ifdef __cplusplus
}
#endif
#endif
