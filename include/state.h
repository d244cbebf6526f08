/*
 * state.h -- quantum state container (statevector or Hermitian density matrix).
 *
 * Same type layout and entry points as the reference (include/state.h:31-143).
 * In this library `data` is a DEVICE pointer (HBM of the current HIP device):
 * read and write it only through state_get / state_set / state_cp /
 * state_print / orkan_state_download / orkan_state_upload.
 */
#ifndef STATE_H
#define STATE_H

#include "q_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/state.h:31-35 */
typedef enum state_type {
    PURE,          /* statevector, 2^n amplitudes */
    MIXED_PACKED,  /* lower triangle, column-major packed (LAPACK uplo='L') */
    MIXED_TILED    /* lower triangle of TILE_DIM x TILE_DIM row-major tiles */
} state_type_t;

/* include/state.h:40-44 */
typedef struct state {
    state_type_t type;
    cplx_t *data;
    qubit_t qubits;
} state_t;

/* include/state.h:53 */
typedef void (*qop_t)(state_t *state);

idx_t  state_len(const state_t *state);                              /* include/state.h:67  */
void   state_init(state_t *state, qubit_t qubits, cplx_t **data);    /* include/state.h:79  */
void   state_free(state_t *state);                                   /* include/state.h:88  */
void   state_plus(state_t *state, qubit_t qubits);                   /* include/state.h:100 */
state_t state_cp(const state_t *state);                              /* include/state.h:108 */
void   state_print(const state_t *state);                            /* include/state.h:115 */
cplx_t state_get(const state_t *state, idx_t row, idx_t col);        /* include/state.h:128 */
void   state_set(state_t *state, idx_t row, idx_t col, cplx_t val);  /* include/state.h:141 */

#ifdef __cplusplus
}
#endif
#endif /* STATE_H */
