/*
 * meas.h -- diagonal-observable reductions.  Identical to the reference,
 * include/meas.h:38 and include/meas.h:57.  `obs` is a caller-owned HOST array
 * of 2^n doubles.
 */
#ifndef MEAS_H
#define MEAS_H

#include "state.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sum_x obs[x] |psi_x|^2 (PURE) or sum_x obs[x] Re rho_xx (mixed). */
double mean(const state_t *state, const double *obs);

/* sum_x obs[x] conj(bra_x) ket_x; both states PURE. */
cplx_t matel_diag(const state_t *bra, const state_t *ket, const double *obs);

#ifdef __cplusplus
}
#endif
#endif /* MEAS_H */
