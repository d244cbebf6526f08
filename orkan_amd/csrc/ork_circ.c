/*
 * ork_circ.c -- exp_diag, U = exp(-i D t), validated like src/circ/circ.c:23-65
 * (t == 0 returns without touching the state).  PURE: one sincos per
 * amplitude; mixed: per-index phases (cos, sin of d_x t) into a device scratch
 * array, then an elementwise pass over the stored entries with
 * e^{-i(d_r - d_c)t} = (cr - i sr)(cc + i sc), diagonal entries untouched
 * (src/circ/circ_packed.c:28-71, src/circ/circ_tiled.c:25-105).
 */
#include "ork_host.h"
#include "orkan_backend.h"

#include <hip/hip_runtime_api.h>

#define XV(cond, msg) ORK_VALIDATE("circ", cond, msg)

static void exp_any(state_t *state, const double *diag, double t) {
    const idx_t dim = (idx_t)1 << state->qubits;
    ork_lock();
    ork_dev_t *d = ork_dev();
    const double *dd = ork_dev_doubles(diag, dim);
    int e;
    if (ork_layout(state) == PURE) {
        e = ork_sv_exp_diag(d->stream, ork_prec(state), state->data, state->qubits, dd, t);
    } else {
        const size_t need = (size_t)dim * 16;
        if (d->phase_bytes < need) {
            if (d->phase) {
                ORK_HIP("exp_diag", hipStreamSynchronize((hipStream_t)d->stream));
                ORK_HIP("exp_diag", hipFree(d->phase));
            }
            d->phase = NULL;
            d->phase_bytes = 0;
            ORK_HIP("exp_diag scratch", hipMalloc(&d->phase, need));
            d->phase_bytes = need;
        }
        e = ork_dm_exp_diag(d->stream, ork_prec(state), ork_layout(state), LOG_TILE_DIM, state->data, state->qubits,
                            dd, t, d->phase);
    }
    /* the staged copy and the phase scratch are reused by later calls on the same
     * stream, which orders them after this kernel */
    ork_unlock();
    ORK_HIP("exp_diag", e);
}

void exp_diag_pure(state_t *state, const double *diag, double t) { exp_any(state, diag, t); }
void exp_diag_packed(state_t *state, const double *diag, double t) { exp_any(state, diag, t); }
void exp_diag_tiled(state_t *state, const double *diag, double t) { exp_any(state, diag, t); }

void exp_diag(state_t *state, const double *diag, double t) {
    XV(state && state->data, "exp_diag: null state or data pointer");
    XV(diag, "exp_diag: null diagonal pointer");
    if (t == 0.0) return;
    XV(ork_type_ok(state), "exp_diag: invalid state type");
    exp_any(state, diag, t);
}

/* Allocation-free exp_diag (include/orkan_amd.h): the same kernels as exp_diag on the library
 * stream, with the density-matrix phase scratch in the fixed device array of k_noalloc.hip, the
 * diagonal taken from device memory only (no staging), and nothing allocated or freed. */
_Static_assert(ORK_NOALLOC_MAX_QUBITS == ORKAN_NOALLOC_MAX_QUBITS, "scratch size of the header and the kernels");

int orkan_exp_diag_noalloc(state_t *state, const double *diag_dev, double t) {
    if (!state || !state->data || !diag_dev) return QS_ERR_NULL;
    if (!ork_type_ok(state)) return QS_ERR_TYPE;
    if (ork_layout(state) != PURE && state->qubits > ORK_NOALLOC_MAX_QUBITS) {
        fprintf(stderr,
                "orkan: circ: orkan_exp_diag_noalloc: a density matrix of %d qubits exceeds the fixed phase scratch "
                "(%d qubits, %llu B)\n",
                (int)state->qubits, ORK_NOALLOC_MAX_QUBITS, (unsigned long long)16 << ORK_NOALLOC_MAX_QUBITS);
        return QS_ERR_QUBIT;
    }
    if (!ork_is_dev_ptr(diag_dev)) {
        fprintf(stderr, "orkan: circ: orkan_exp_diag_noalloc: the diagonal must be device memory\n");
        return QS_ERR_PARAM;
    }
    if (t == 0.0) return QS_OK;
    ork_dev_t *d = ork_dev_existing();
    if (!d || d->noalloc_err) {
        fprintf(stderr, "orkan: circ: orkan_exp_diag_noalloc: library runtime not set up on this device (%s)\n",
                d ? hipGetErrorString((hipError_t)d->noalloc_err) : "no state created on it yet");
        return QS_ERR_PARAM;
    }
    const int e = ork_exp_diag_noalloc(d->stream, ork_prec(state), ork_layout(state), LOG_TILE_DIM, state->data,
                                       state->qubits, diag_dev, t);
    if (e) {
        fprintf(stderr, "orkan: circ: orkan_exp_diag_noalloc: %s\n", hipGetErrorString((hipError_t)e));
        return QS_ERR_PARAM;
    }
    return QS_OK;
}
