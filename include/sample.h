/*
 * sample.h -- sample_matrix (extension).  The reference specifies this operation but does not
 * implement it (docs/SAMPLE_MATRIX.md:1-80): for an ansatz |psi> = K_L ... K_1 |iota>,
 * K_l = sum_m w_l[m] M_l[m], and a real diagonal H = diag(obs), the sample matrix at channel l*
 *   S_ab = <psi_a| H |psi_b>,  |psi_a> = K_L ... K_{l*+1} M_{l*}[a] K_{l*-1} ... K_1 |iota>.
 */
#ifndef SAMPLE_H
#define SAMPLE_H

#include <stddef.h>

#include "state.h"

#ifdef __cplusplus
extern "C" {
#endif

/* one channel of the ansatz: K = sum_m weights[m] ops[m] */
typedef struct {
    uint64_t n_ops;          /* K_l >= 1 */
    const qop_t *ops;        /* M_l[0 .. n_ops): operations on a PURE state (qop_t, include/state.h) */
    const cplx_t *weights;   /* w_l[0 .. n_ops); may be NULL at the channel the matrix is taken at */
} ansatz_channel_t;

/* S (caller-allocated, K x K column-major, K = channels[lstar].n_ops) = the sample matrix at the
 * channel of 0-based index lstar.  iota: a PURE complex128 state, CONSUMED (it ends holding
 * K_{lstar-1} ... K_1 |iota>; pass a copy to keep it).  obs: 2^n doubles (host or device).
 * workspace_bytes caps the state vectors the call holds besides iota: two for the linear
 * combinations and 2B leaves (B = the leaf block, at most 64), so at least 4 x 16 x 2^n bytes.
 * Both triangles are filled, S_ba = conj(S_ab) exactly, the diagonal is real.  The result does not
 * depend on the block size.  Returns QS_OK, QS_ERR_NULL, QS_ERR_TYPE (iota not PURE complex128),
 * QS_ERR_PARAM (lstar out of range, a channel without operations or, other than at lstar, without
 * weights, workspace below the minimum), QS_ERR_OOM. */
qs_error_t sample_matrix(state_t *iota, const double *obs, const ansatz_channel_t *channels, uint64_t n_channels,
                         uint64_t lstar, size_t workspace_bytes, cplx_t *S);

#ifdef __cplusplus
}
#endif
#endif /* SAMPLE_H */
