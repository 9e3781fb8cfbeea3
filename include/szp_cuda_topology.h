/*
 * szp_cuda_topology.h -- 2D TopoSZp topology entry points of libtoposzp_hip.so
 * (MI355X / gfx950).  Drop-in for the reference's SZp/include/szp_cuda_topology.cuh:
 * same names, argument meaning, host-pointer ownership and NULL-on-error results.
 * Returned buffers are malloc'ed; the caller frees them.
 */
#ifndef TOPOSZP_SZP_CUDA_TOPOLOGY_H
#define TOPOSZP_SZP_CUDA_TOPOLOGY_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef SZP_CRITICAL_POINT_DEFINED
#define SZP_CRITICAL_POINT_DEFINED
/* Record layout of CriticalPoint, SZp/include/szp_float.h:20-25. */
typedef struct {
    int x, y;
    int type;           /* 1 max, 2 min, 3 saddle */
    int quantized_bin;  /* (int)((v + eb) * (1.0 / eb)), float add then double multiply */
    int sort_position;  /* dense rank of v among the distinct CP values of its bin */
} CriticalPoint;
#endif

/* Replaces szp_cuda_find_critical_points, SZp/include/szp_cuda_topology.cuh:23-24.
 * Interior points with strict 4-neighbour tests.  Records come in row-major order
 * (the reference's CUDA finder returns them in atomicAdd order); sort_position = 0.
 * rows <= 2 or cols <= 2 -> NULL with *outCount = 0. */
CriticalPoint *szp_cuda_find_critical_points(float *data, size_t *outCount,
                                              int rows, int cols, float absErrBound);

/* Replaces szp_cuda_sort_critical_points_by_original_data,
 * SZp/include/szp_cuda_topology.cuh:30-32.  Fills sort_position in place; the
 * record order is unchanged. */
void szp_cuda_sort_critical_points_by_original_data(CriticalPoint *critical_points,
                                                     size_t critical_count,
                                                     float *data, int cols);

/* Replaces szp_cuda_compress_sort_positions, SZp/include/szp_cuda_topology.cuh:38-40.
 * Stream = [u64 0][blocks] (single chunk), the sort positions of the maxima and
 * minima in record order.  NULL with *outSize = 0 when there is no extremum. */
unsigned char *szp_cuda_compress_sort_positions(CriticalPoint *critical_points,
                                                 size_t critical_count,
                                                 size_t *outSize, int blockSize);

/* Replaces szp_cuda_float_compress_randomaccess_topology_preserved,
 * SZp/include/szp_cuda_topology.cuh:46-50.  Stream = [u64 0][blocks with 2-bit
 * types], byte-identical to the reference CPU writer run with one thread
 * (SZp/src/szp_float.cc:998-1318), including the bit-count byte of a trailing
 * one-element block. */
unsigned char *szp_cuda_float_compress_randomaccess_topology_preserved(
    float *oriData, size_t *outSize, float absErrBound,
    size_t nbEle, int blockSize,
    CriticalPoint *critical_points, int critical_count,
    int rows, int cols);

#ifdef __cplusplus
}
#endif

#endif
