/*
 * szp_cuda_topology_3d.h -- 3D (6-connected) TopoSZp entry points of
 * libtoposzp_hip.so.  Drop-in for SZp/include/szp_cuda_topology_3d.cuh:16-38.
 * Flat index = x * d2 * d3 + y * d3 + z.
 */
#ifndef TOPOSZP_SZP_CUDA_TOPOLOGY_3D_H
#define TOPOSZP_SZP_CUDA_TOPOLOGY_3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef SZP_CRITICAL_POINT_3D_DEFINED
#define SZP_CRITICAL_POINT_3D_DEFINED
/* Record layout of CriticalPoint3D, SZp/include/szp_topology_3d.h:22-27. */
typedef struct {
    int x, y, z;
    int type;
    int quantized_bin;
    int sort_position;
} CriticalPoint3D;
#endif

/* Replaces szp_cuda_find_critical_points_3d (.cuh:16-18); row-major order. */
CriticalPoint3D *szp_cuda_find_critical_points_3d(float *data, size_t *outCount,
                                                    int d1, int d2, int d3,
                                                    float absErrBound);

/* Replaces szp_cuda_sort_critical_points_3d (.cuh:20-22). */
void szp_cuda_sort_critical_points_3d(CriticalPoint3D *critical_points,
                                       size_t critical_count,
                                       float *data, int d2, int d3);

/* Replaces szp_cuda_compress_sort_positions_3d (.cuh:24-26); [u64 0][blocks]. */
unsigned char *szp_cuda_compress_sort_positions_3d(CriticalPoint3D *critical_points,
                                                     size_t critical_count,
                                                     size_t *outSize, int blockSize);

/* Replaces szp_cuda_float_compress_topology_3d (.cuh:28-32); [u64 0][blocks],
 * byte-identical to SZp/src/szp_topology_3d.cc:433-632 run with one thread. */
unsigned char *szp_cuda_float_compress_topology_3d(
    float *oriData, size_t *outSize, float absErrBound,
    size_t nbEle, int blockSize,
    CriticalPoint3D *critical_points, int critical_count,
    int d1, int d2, int d3);

/* Replaces szp_cuda_double_compress_topology_3d (.cuh:34-38,
 * SZp/src/szp_cuda_topology_3d.cu:1065-1188): types from the critical points, q = C
 * cast of x * (1 / eb), three Jacobi passes of the saddle fix, then [u64 0][blocks] of
 * the topology codec on q.  blockSize in [1, 257], every dimension at least 3. */
unsigned char *szp_cuda_double_compress_topology_3d(
    double *oriData, size_t *outSize, double absErrBound,
    size_t nbEle, int blockSize,
    CriticalPoint3D *critical_points, int critical_count,
    int d1, int d2, int d3);

#ifdef __cplusplus
}
#endif

#endif
