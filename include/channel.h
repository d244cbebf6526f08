/*
 * channel.h -- Kraus channels through their superoperator.
 * Types and entry points identical to the reference, include/channel.h:28-67.
 */
#ifndef CHANNEL_H
#define CHANNEL_H

#include "state.h"

#ifdef __cplusplus
extern "C" {
#endif

/* n_terms Kraus matrices of size 2^n x 2^n, concatenated, each row-major. */
typedef struct {
    qubit_t n_qubits;
    uint64_t n_terms;
    cplx_t *data;
} kraus_t;

/* 4^n x 4^n superoperator, row-major, acting on vec(rho) = (rho_00, rho_01, ...). */
typedef struct {
    qubit_t n_qubits;
    cplx_t *data;
} superop_t;

/* S[(i,j),(k,l)] = sum_a K_a[i,k] conj(K_a[j,l]).  Host memory from calloc: the
 * caller free()s sop.data. */
superop_t kraus_to_superop(const kraus_t *kraus);

/* rho -> S(rho) on one qubit; mixed states only. */
void channel_1q(state_t *state, const superop_t *sop, qubit_t target);

#ifdef __cplusplus
}
#endif
#endif /* CHANNEL_H */
