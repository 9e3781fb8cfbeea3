/* tszp_types.h -- plain data shared by host code and kernels. */
#ifndef TSZP_TYPES_H
#define TSZP_TYPES_H

#include <stdint.h>

/* error bits a kernel can raise in its device error word */
#define TSZP_ERR_LOOKBACK 1u   /* a look-back spin exceeded its bound */
#define TSZP_ERR_STREAM 2u     /* a decoder met an impossible bit count / offset */
#define TSZP_ERR_OVERFLOW 4u   /* an output would exceed its capacity */

#ifdef __cplusplus
namespace tszp {
#endif

/* x / d for x < 2^31 as (x * mul) >> shift */
typedef struct {
  uint32_t mul;
  uint32_t shift;
} FastDiv;

/* Grid geometry of a field: dims[0] is the slowest axis.  ndim = 1 means a flat
 * array with no topology. */
typedef struct {
  int ndim;
  uint32_t d0, d1, d2;   /* 2D: rows = d0, cols = d1; 3D: d0 x d1 x d2 */
  uint32_t plane;        /* d1 * d2 (3D) */
  FastDiv div_row;       /* by cols (2D) or d2 (3D) */
  FastDiv div_plane;     /* by d1 * d2 (3D) */
} Grid;

/* device-side result words of one compression launch */
typedef struct {
  uint64_t payload_bytes;  /* bytes after the chunk-offset header */
  uint64_t cp_count;       /* critical points (classify modes) */
  uint64_t ext_count;      /* extrema among them */
  uint32_t error;          /* TSZP_ERR_* bits */
  uint32_t pad;
} CompressResult;

#ifdef __cplusplus
}  /* namespace tszp */
#endif

#endif
