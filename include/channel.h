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

/* Extension, no counterpart in the reference (whose channel_1q rejects n_qubits != 1,
 * src/channel/channel.c:71): rho -> S(rho) on the qubit pair (q1, q2) for a 2-qubit
 * superoperator from kraus_to_superop, whose 4 x 4 Kraus matrices use the local index
 * a = x_q1 + 2 x_q2 (q1 is the low bit).  Mixed states only; exits on a pure state, equal or
 * out-of-range qubits, or n_qubits != 2, like channel_1q. */
void channel_2q(state_t *state, const superop_t *sop, qubit_t q1, qubit_t q2);

#ifdef __cplusplus
}
#endif
#endif /* CHANNEL_H */
