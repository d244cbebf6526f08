/*
 * q_types.h -- scalar types, error codes and tile geometry of the Orkan C API.
 *
 * Interface facts kept identical to the reference so that callers compiled
 * against it link against this library unchanged:
 *   qs_error_t values           include/q_types.h:29-38
 *   qubit_t / cplx_t / idx_t    include/q_types.h:46-48
 *   TILE_DIM / TILE_SIZE        include/q_types.h:56-57 (LOG_TILE_DIM comes from
 *                               the compile line and is part of the ABI)
 *   POW2                        include/q_types.h:73
 */
#ifndef QTYPES_H
#define QTYPES_H

#include <stdint.h>
#ifndef __cplusplus
#include <complex.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifndef LOG_TILE_DIM
#define LOG_TILE_DIM 5
#endif

typedef enum {
    QS_OK         =  0,
    QS_ERR_NULL   = -1,
    QS_ERR_OOM    = -2,
    QS_ERR_QUBIT  = -3,
    QS_ERR_TYPE   = -4,
    QS_ERR_FILE   = -5,
    QS_ERR_FORMAT = -6,
    QS_ERR_PARAM  = -7
} qs_error_t;

typedef unsigned char qubit_t;   /* qubit position, 0 = least significant bit */
#ifdef __cplusplus
typedef struct { double re, im; } cplx_t;   /* layout-identical to double _Complex */
#else
typedef double _Complex cplx_t;
#endif
typedef uint64_t idx_t;

#define TILE_DIM  (1 << LOG_TILE_DIM)
#define TILE_SIZE (TILE_DIM * TILE_DIM)

#ifndef MIN
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#endif
#define SWAP(a, b, T) do { T swap_tmp_ = *(a); *(a) = *(b); *(b) = swap_tmp_; } while (0)
#define POW2(a, T) (((T)1) << (a))

#ifdef __cplusplus
}
#endif
#endif /* QTYPES_H */
