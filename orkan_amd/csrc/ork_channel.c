/*
 * ork_channel.c -- Kraus channels through their superoperator.
 *
 * kraus_to_superop runs on the host and returns calloc'd memory the caller
 * frees, exactly as the reference (src/channel/channel.c:21-50):
 *   S[(i,j),(k,l)] = sum_a K_a[i,k] conj(K_a[j,l]),  row-major, 4^n x 4^n.
 * channel_1q validates like src/channel/channel.c:67-85 and applies the 4x4
 * superoperator on the device (streamed kernels, tile groups where they do not
 * apply).  channel_2q is an extension: the reference's kraus_to_superop already
 * builds 2-qubit superoperators (src/channel/channel.c:24-47) but channel_1q
 * accepts only 1-qubit ones (src/channel/channel.c:71).
 */
#include "ork_host.h"
#include "orkan_backend.h"

#include <string.h>

#define CV(cond, msg) ORK_VALIDATE("channel", cond, msg)

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

static void apply_sop(state_t *state, const cplx_t *sop, qubit_t target) {
    double S[32];
    for (int i = 0; i < 16; ++i) {
        S[2 * i] = creal(sop[i]);
        S[2 * i + 1] = cimag(sop[i]);
    }
    ork_mono_t m;
    memset(&m, 0, sizeof(m));
    m.k = 1;
    m.qs[0] = target;
    ORK_HIP("channel_1q", ork_dm_group(ork_dev()->stream, ork_prec(state), ork_layout(state), LOG_TILE_DIM,
                                       state->data, state->qubits, ORK_DM_SOP, &m, S));
}

void channel_packed_1q(state_t *state, const cplx_t *sop, qubit_t target) { apply_sop(state, sop, target); }
void channel_tiled_1q(state_t *state, const cplx_t *sop, qubit_t target) { apply_sop(state, sop, target); }

void channel_1q(state_t *state, const superop_t *sop, qubit_t target) {
    CV(state && state->data, "channel_1q: null state or data pointer");
    CV(target < state->qubits, "channel_1q: target qubit out of range");
    CV(sop && sop->data, "channel_1q: null superoperator or data pointer");
    CV(sop->n_qubits == 1, "channel_1q: superoperator is not 1-local");
    CV(ork_type_ok(state), "channel_1q: unknown state type");
    switch (ork_layout(state)) {
    case PURE:
        CV(0, "channel_1q: not available for pure states");
        break;
    case MIXED_PACKED:
        channel_packed_1q(state, sop->data, target);
        break;
    default:
        channel_tiled_1q(state, sop->data, target);
        break;
    }
}

/* rho -> S(rho) on the qubit pair (q1, q2): S is 16 x 16, row-major on vec of the 4 x 4 block
 * whose local index is a = x_q1 + 2 x_q2 (q1 is the low bit). */
void channel_2q(state_t *state, const superop_t *sop, qubit_t q1, qubit_t q2) {
    CV(state && state->data, "channel_2q: null state or data pointer");
    CV(q1 < state->qubits && q2 < state->qubits, "channel_2q: qubit out of range");
    CV(q1 != q2, "channel_2q: qubits must differ");
    CV(sop && sop->data, "channel_2q: null superoperator or data pointer");
    CV(sop->n_qubits == 2, "channel_2q: superoperator is not 2-local");
    CV(ork_type_ok(state), "channel_2q: unknown state type");
    CV(ork_layout(state) != PURE, "channel_2q: not available for pure states");
    ork_mono_t m;
    memset(&m, 0, sizeof(m));
    m.k = 2;
    m.qs[0] = q1;
    m.qs[1] = q2;
    ork_lock();
    const double *S = ork_dev_doubles((const double *)sop->data, 2 * 256);
    const int e = ork_dm_group(ork_dev()->stream, ork_prec(state), ork_layout(state), LOG_TILE_DIM, state->data,
                               state->qubits, ORK_DM_SOP2, &m, S);
    ork_unlock();
    ORK_HIP("channel_2q", e);
}
