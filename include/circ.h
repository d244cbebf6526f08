/*
 * circ.h -- diagonal evolution U = exp(-i D t).  Identical to the reference,
 * include/circ.h:41.  `diag` is a caller-owned HOST array of 2^n doubles.
 */
#ifndef CIRC_H
#define CIRC_H

#include "state.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PURE: psi_x *= exp(-i diag[x] t);  mixed: rho_rc *= exp(-i (diag[r]-diag[c]) t).
 * t == 0 returns without touching the state. */
void exp_diag(state_t *state, const double *diag, double t);

#ifdef __cplusplus
}
#endif
#endif /* CIRC_H */
