/*
 * orkan_backend.h -- per-layout backend entry points.
 *
 * The reference builds without symbol visibility control
 * (cmake/CompilerFlags.cmake:12-17), so its backend functions are exported
 * although its headers do not declare them, and its tests and profilers link
 * against them (test/gate/test_gate_tiled_1q.c:39, test/gate/test_gate_tiled_2q_mat.c:45,
 * profile/src/profile_channel_tiled.c:30-35).  This library exports the same
 * names with the same meaning; like the reference they do not validate their
 * arguments.
 */
#ifndef ORKAN_BACKEND_H
#define ORKAN_BACKEND_H

#include "orkan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/gate/tiled/gate_tiled.h:32 -- rho -> U rho U^dagger, mat = 2x2 column-major */
void single_from_mat(state_t *state, qubit_t target, const cplx_t *mat);
/* src/gate/tiled/gate_tiled_2q.c:3 -- mat = 4x4 column-major in the (hi, lo) basis of
 * the sorted pair, so the argument order does not matter */
void two_from_mat(state_t *state, qubit_t q1, qubit_t q2, const cplx_t *mat);

/* src/channel/channel.c:57-58 -- sop = 16 entries, row-major */
void channel_packed_1q(state_t *state, const cplx_t *sop, qubit_t target);
void channel_tiled_1q(state_t *state, const cplx_t *sop, qubit_t target);

/* src/gate/gate.c:31-110 -- <gate>_{pure,packed,tiled} for all 18 gates */
#define ORKAN_BACKEND_1Q(name)                                  \
    void name##_pure(state_t *state, qubit_t target);           \
    void name##_packed(state_t *state, qubit_t target);         \
    void name##_tiled(state_t *state, qubit_t target);
#define ORKAN_BACKEND_ROT(name)                                         \
    void name##_pure(state_t *state, qubit_t target, double theta);     \
    void name##_packed(state_t *state, qubit_t target, double theta);   \
    void name##_tiled(state_t *state, qubit_t target, double theta);
#define ORKAN_BACKEND_2Q(name)                                          \
    void name##_pure(state_t *state, qubit_t a, qubit_t b);             \
    void name##_packed(state_t *state, qubit_t a, qubit_t b);           \
    void name##_tiled(state_t *state, qubit_t a, qubit_t b);
ORKAN_BACKEND_1Q(x) ORKAN_BACKEND_1Q(y) ORKAN_BACKEND_1Q(z)
ORKAN_BACKEND_1Q(h) ORKAN_BACKEND_1Q(s) ORKAN_BACKEND_1Q(sdg)
ORKAN_BACKEND_1Q(t) ORKAN_BACKEND_1Q(tdg) ORKAN_BACKEND_1Q(hy)
ORKAN_BACKEND_ROT(rx) ORKAN_BACKEND_ROT(ry) ORKAN_BACKEND_ROT(rz) ORKAN_BACKEND_ROT(p)
ORKAN_BACKEND_2Q(cx) ORKAN_BACKEND_2Q(cy) ORKAN_BACKEND_2Q(cz) ORKAN_BACKEND_2Q(swap)
void ccx_pure(state_t *state, qubit_t c1, qubit_t c2, qubit_t target);
void ccx_packed(state_t *state, qubit_t c1, qubit_t c2, qubit_t target);
void ccx_tiled(state_t *state, qubit_t c1, qubit_t c2, qubit_t target);

/* src/meas/meas.c:36-39, src/circ/circ.c:35-37 */
double mean_pure(const state_t *state, const double *obs);
double mean_packed(const state_t *state, const double *obs);
double mean_tiled(const state_t *state, const double *obs);
cplx_t matel_diag_pure(const state_t *bra, const state_t *ket, const double *obs);
void exp_diag_pure(state_t *state, const double *diag, double t);
void exp_diag_packed(state_t *state, const double *diag, double t);
void exp_diag_tiled(state_t *state, const double *diag, double t);

#ifdef __cplusplus
}
#endif
#endif /* ORKAN_BACKEND_H */
