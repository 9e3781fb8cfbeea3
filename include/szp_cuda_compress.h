/*
 * szp_cuda_compress.h -- plain (no topology) random-access SZp compression of
 * libtoposzp_hip.so.  Drop-in for the float random-access entries of
 * SZp/include/szp_cuda_compress.cuh:22-28.
 *
 * Stream = [f32 eb, big-endian][u64 0][blocks]: the reference CPU writer's bytes
 * (SZp/src/szp_float.cc:2061-2292, reciprocal 1/eb taken in float), which the
 * reference CUDA writer approximates with a native-endian header and a double
 * reciprocal (SZp/src/szp_cuda_compress.cu:393,459).
 */
#ifndef TOPOSZP_SZP_CUDA_COMPRESS_H
#define TOPOSZP_SZP_CUDA_COMPRESS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces szp_cuda_float_compress_randomaccess, .cuh:22-24.  blockSize in [2, 1025]. */
unsigned char *szp_cuda_float_compress_randomaccess(
    float *oriData, size_t *outSize, float absErrBound,
    size_t nbEle, int blockSize);

/* Replaces szp_cuda_float_compress_randomaccess_arg, .cuh:26-28.  `output` must hold
 * szp_cuda_float_compress_randomaccess_bound(nbEle, blockSize) bytes. */
void szp_cuda_float_compress_randomaccess_arg(
    unsigned char *output, float *oriData, size_t *outSize,
    float absErrBound, size_t nbEle, int blockSize);

/* Worst-case stream size (no counterpart in the reference, which sizes its buffers
 * at 4 * nbEle + 4 bytes and can overrun them on incompressible data). */
size_t szp_cuda_float_compress_randomaccess_bound(size_t nbEle, int blockSize);

/* Double random access, replaces szp_cuda_double_compress_randomaccess[_arg]
 * (.cuh:30-40): the CPU writer's stream with one chunk
 * (SZp/src/szp_double.cc:759-920): [f64 eb big-endian][u64 0][blocks], each block
 * [i32 anchor][bc][signs][magnitudes], the bc byte also for a one-element block.
 * blockSize in [1, 1025]; `_arg` outputs must hold tszp_elem_bound(nbEle, blockSize,
 * 1, 8, 1) bytes (include/toposzp_hip.h). */
unsigned char *szp_cuda_double_compress_randomaccess(
    double *oriData, size_t *outSize, double absErrBound,
    size_t nbEle, int blockSize);
void szp_cuda_double_compress_randomaccess_arg(
    unsigned char *output, double *oriData, size_t *outSize,
    double absErrBound, size_t nbEle, int blockSize);

/* Threadblock (non-random-access) codec, replaces
 * szp_cuda_{float,double}_compress_threadblock[_arg] (.cuh:42-60): the CPU writers'
 * stream with one chunk (SZp/src/szp_float.cc:1445-1590, szp_double.cc:148-318):
 * [eb big-endian][u64 0][i32 anchor][blocks of blockSize deltas chained across
 * blocks, each [bc][signs][magnitudes]].  `_arg` outputs must hold
 * tszp_elem_bound(nbEle, blockSize, 1, sizeof(T), 0) bytes. */
unsigned char *szp_cuda_float_compress_threadblock(
    float *oriData, size_t *outSize, float absErrBound,
    size_t nbEle, int blockSize);
void szp_cuda_float_compress_threadblock_arg(
    unsigned char *output, float *oriData, size_t *outSize,
    float absErrBound, size_t nbEle, int blockSize);
unsigned char *szp_cuda_double_compress_threadblock(
    double *oriData, size_t *outSize, double absErrBound,
    size_t nbEle, int blockSize);
void szp_cuda_double_compress_threadblock_arg(
    unsigned char *output, double *oriData, size_t *outSize,
    double absErrBound, size_t nbEle, int blockSize);

/* Replaces szp_cuda_float_{direct,threadblock}_predict_quantization (.cuh:62-64):
 * malloc'ed int[nbEle] of q[i] - q[i - 1] (q[-1] = 0), q = C cast of
 * x * (1 / (double)absErrBound); outSize and blockSize are not used, as there. */
int *szp_cuda_float_direct_predict_quantization(
    float *oriData, size_t *outSize, float absErrBound,
    size_t nbEle, int blockSize);
int *szp_cuda_float_threadblock_predict_quantization(
    float *oriData, size_t *outSize, float absErrBound,
    size_t nbEle, int blockSize);

#ifdef __cplusplus
}
#endif

#endif
