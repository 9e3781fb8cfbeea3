/*
 * szp_cuda_decompress.h -- decompression entry points of libtoposzp_hip.so.
 * Drop-in for SZp/include/szp_cuda_decompress.cuh:23-72.
 *
 * Streams carry no length and no block offsets: the chunk count T is detected from
 * the offset table as the reference does (SZp/include/szp_detect_threads.h) and the
 * block offsets are found by walking the block headers on the host, as in
 * SZp/src/szp_cuda_decompress.cu:233-477; the blocks then decode in parallel on
 * the GPU.
 */
#ifndef TOPOSZP_SZP_CUDA_DECOMPRESS_H
#define TOPOSZP_SZP_CUDA_DECOMPRESS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces szp_cuda_float_decompress_randomaccess, .cuh:23-25.  cmpBytes points
 * after the 4-byte error-bound header (at the chunk-offset table). */
float *szp_cuda_float_decompress_randomaccess(
    size_t nbEle, float absErrBound, int blockSize,
    unsigned char *cmpBytes);

/* Replaces szp_cuda_float_decompress_randomaccess_arg, .cuh:27-29. */
void szp_cuda_float_decompress_randomaccess_arg(
    float *newData, size_t nbEle, float absErrBound, int blockSize,
    unsigned char *cmpBytes);

/* Replaces szp_cuda_float_decompress_randomaccess_topology_preserved, .cuh:61-63.
 * *newData (malloc) gets (float)q * eb, *FN (calloc) the 2-bit types as ints. */
void szp_cuda_float_decompress_randomaccess_topology_preserved(
    float **newData, size_t nbEle, float absErrBound, int blockSize,
    unsigned char *cmpBytes, int **FN);

/* Replaces szp_cuda_decompress_sort_positions, .cuh:69-70.  cmpBytes points at the
 * chunk-offset table ([u64 off[T]][blocks]). */
int *szp_cuda_decompress_sort_positions(
    unsigned char *cmpBytes, size_t critical_count, int blockSize);

/* Replace szp_cuda_double_decompress_randomaccess[_arg] (.cuh:31-37) and
 * szp_cuda_{float,double}_decompress_threadblock[_arg] (.cuh:43-57).  cmpBytes points
 * after the error-bound header (at the chunk-offset table); streams of the CPU writers
 * with any chunk count, which is detected (candidates as in
 * SZp/include/szp_detect_threads.h:196-222: the largest T in 2..128 with monotone
 * offsets whose chunks walk, each ending where the next starts, else 1; the reference
 * takes the smallest, which small chunks make ambiguous). */
double *szp_cuda_double_decompress_randomaccess(
    size_t nbEle, double absErrBound, int blockSize,
    unsigned char *cmpBytes);
void szp_cuda_double_decompress_randomaccess_arg(
    double *newData, size_t nbEle, double absErrBound, int blockSize,
    unsigned char *cmpBytes);
float *szp_cuda_float_decompress_threadblock(
    size_t nbEle, float absErrBound, int blockSize,
    unsigned char *cmpBytes);
void szp_cuda_float_decompress_threadblock_arg(
    float *newData, size_t nbEle, float absErrBound, int blockSize,
    unsigned char *cmpBytes);
double *szp_cuda_double_decompress_threadblock(
    size_t nbEle, double absErrBound, int blockSize,
    unsigned char *cmpBytes);
void szp_cuda_double_decompress_threadblock_arg(
    double *newData, size_t nbEle, double absErrBound, int blockSize,
    unsigned char *cmpBytes);

#ifdef __cplusplus
}
#endif

#endif
