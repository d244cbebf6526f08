/*
 * orc_channel.c -- CPU oracle of kraus_to_superop and channel_1q.
 * TEST INFRASTRUCTURE only.
 *
 * Follows:
 *   kraus_to_superop  S[(i,j),(k,l)] = sum_a K_a[i,k] conj(K_a[j,l]), row-major,
 *                     calloc'd (caller frees)            src/channel/channel.c:21-50
 *   channel_1q validation messages and exit()             src/channel/channel.c:8-13, 67-85
 *   channel_packed_1q (lower / upper-conj / diagonal block) src/channel/channel_packed.c:48-120
 *   channel_tiled_1q  (within-tile; cross-tile a, b, c)    src/channel/channel_tiled.c:52-185
 * The superoperator step itself is op1_apply(OP_SOP) in orc_ops.h.
 *
 * channel_2q has no counterpart in the reference (channel_1q rejects n_qubits != 1,
 * src/channel/channel.c:71; kraus_to_superop already builds n-qubit superoperators,
 * src/channel/channel.c:24-47): every 4x4 block of the qubit pair, read and written through
 * the Hermitian state_get / state_set, gets out = S vec(block).  With no reference function to
 * compare against, it is pinned by tests/test_oracle_math.py against sum_a K_a rho K_a^dagger
 * on dense matrices (every ordered pair, both layouts and tile sizes).
 */
#include "channel.h"
#include "orc_internal.h"

#include <stdio.h>
#include <stdlib.h>

#define CV(cond, msg) do { if (!(cond)) { fprintf(stderr, "orkan: channel: %s\n", (msg)); exit(EXIT_FAILURE); } } while (0)

superop_t kraus_to_superop(const kraus_t *kraus) {
    CV(kraus && kraus->data, "kraus_to_superop: null kraus or data");
    const idx_t N = (idx_t)1 << kraus->n_qubits, M = N * N;
    superop_t out;
    out.n_qubits = kraus->n_qubits;
    out.data = calloc(M * M, sizeof(cplx_t));
    CV(out.data, "kraus_to_superop: allocation failed");
    for (uint64_t a = 0; a < kraus->n_terms; ++a) {
        const cplx_t *K = kraus->data + a * M;
        for (idx_t i = 0; i < N; ++i)
            for (idx_t j = 0; j < N; ++j)
                for (idx_t k = 0; k < N; ++k)
                    for (idx_t l = 0; l < N; ++l)
                        out.data[(i * N + j) * M + k * N + l] += K[i * N + k] * conj(K[j * N + l]);
    }
    return out;
}

void channel_packed_1q(state_t *state, const cplx_t *sop, const qubit_t target) {
    op1_t op = op1_make(OP_SOP, 0.0);
    op.sop = sop;
    orc_packed_1q(state, target, &op);
}

void channel_tiled_1q(state_t *state, const cplx_t *sop, const qubit_t target) {
    op1_t op = op1_make(OP_SOP, 0.0);
    op.sop = sop;
    orc_tiled_1q(state, target, &op);
}

void channel_1q(state_t *state, const superop_t *sop, const qubit_t target) {
    CV(state && state->data, "channel_1q: null state or data pointer");
    CV(target < state->qubits, "channel_1q: target qubit out of range");
    CV(sop && sop->data, "channel_1q: null superoperator or data pointer");
    CV(sop->n_qubits == 1, "channel_1q: superoperator is not 1-local");
    switch (state->type) {
    case PURE:
        CV(0, "channel_1q: not available for pure states");
        break;
    case MIXED_PACKED:
        channel_packed_1q(state, sop->data, target);
        break;
    case MIXED_TILED:
        channel_tiled_1q(state, sop->data, target);
        break;
    default:
        CV(0, "channel_1q: unknown state type");
    }
}

void channel_2q(state_t *state, const superop_t *sop, const qubit_t q1, const qubit_t q2) {
    CV(state && state->data, "channel_2q: null state or data pointer");
    CV(q1 < state->qubits && q2 < state->qubits, "channel_2q: qubit out of range");
    CV(q1 != q2, "channel_2q: qubits must differ");
    CV(sop && sop->data, "channel_2q: null superoperator or data pointer");
    CV(sop->n_qubits == 2, "channel_2q: superoperator is not 2-local");
    CV(state->type == PURE || state->type == MIXED_PACKED || state->type == MIXED_TILED,
       "channel_2q: unknown state type");
    CV(state->type != PURE, "channel_2q: not available for pure states");
    orc_dm_sop2(state, q1, q2, sop->data);
}
