/*
 * ork_meas.c -- mean and matel_diag: two-stage deterministic reductions on the
 * device (per-workgroup partial sums in a fixed order, then one workgroup),
 * validated like src/meas/meas.c:23-74.  `obs` may be a host array (staged into
 * HBM) or a device array from orkan_dev_alloc (used in place).
 */
#include "ork_host.h"
#include "orkan_backend.h"

#define MV(cond, msg) ORK_VALIDATE("meas", cond, msg)

static double mean_any(const state_t *state, const double *obs) {
    const idx_t dim = (idx_t)1 << state->qubits;
    double out = 0.0;
    ork_lock();
    ork_dev_t *d = ork_dev();
    const double *o = ork_dev_doubles(obs, dim);
    int e;
    if (ork_layout(state) == PURE)
        e = ork_sv_mean(d->stream, ork_prec(state), state->data, state->qubits, o, d->partials, &out);
    else
        e = ork_dm_mean(d->stream, ork_prec(state), ork_layout(state), LOG_TILE_DIM, state->data, state->qubits, o,
                        d->partials, &out);
    ork_unlock();
    ORK_HIP("mean", e);
    return out;
}

double mean_pure(const state_t *state, const double *obs) { return mean_any(state, obs); }
double mean_packed(const state_t *state, const double *obs) { return mean_any(state, obs); }
double mean_tiled(const state_t *state, const double *obs) { return mean_any(state, obs); }

double mean(const state_t *state, const double *obs) {
    MV(state && state->data, "mean: null state or data pointer");
    MV(obs, "mean: null observable pointer");
    MV(ork_type_ok(state), "mean: invalid state type");
    return mean_any(state, obs);
}

cplx_t matel_diag_pure(const state_t *bra, const state_t *ket, const double *obs) {
    const idx_t dim = (idx_t)1 << bra->qubits;
    double out[2] = {0.0, 0.0};
    ork_lock();
    ork_dev_t *d = ork_dev();
    const double *o = ork_dev_doubles(obs, dim);
    int e;
    if (ork_prec(bra) == ork_prec(ket)) {
        e = ork_sv_matel(d->stream, ork_prec(bra), bra->data, ket->data, bra->qubits, o, d->partials, out);
    } else {
        /* mixed precisions: compare in complex128 */
        state_t b2 = ork_prec(bra) == ORK_F64 ? *bra : orkan_state_convert(bra, 0);
        state_t k2 = ork_prec(ket) == ORK_F64 ? *ket : orkan_state_convert(ket, 0);
        e = ork_sv_matel(d->stream, ORK_F64, b2.data, k2.data, bra->qubits, o, d->partials, out);
        if (b2.data != bra->data) state_free(&b2);
        if (k2.data != ket->data) state_free(&k2);
    }
    ork_unlock();
    ORK_HIP("matel_diag", e);
    return CMPLX(out[0], out[1]);
}

cplx_t matel_diag(const state_t *bra, const state_t *ket, const double *obs) {
    MV(bra && bra->data, "matel_diag: null bra state or data pointer");
    MV(ket && ket->data, "matel_diag: null ket state or data pointer");
    MV(obs, "matel_diag: null observable pointer");
    MV(bra->qubits == ket->qubits, "matel_diag: qubit count mismatch");
    MV(ork_layout(bra) == PURE, "matel_diag: bra must be PURE (mixed not supported)");
    MV(ork_layout(ket) == PURE, "matel_diag: ket must be PURE (mixed not supported)");
    return matel_diag_pure(bra, ket, obs);
}
