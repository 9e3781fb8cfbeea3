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

#ifdef __cplusplus
}
#endif

#endif
