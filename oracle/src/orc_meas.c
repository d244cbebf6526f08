/*
 * orc_meas.c -- CPU oracle of mean and matel_diag.  TEST INFRASTRUCTURE only.
 *
 * Follows:
 *   validation messages and exit()  src/meas/meas.c:23-28, 48-74
 *   mean_pure   sum obs[x] (re^2 + im^2)                 src/meas/meas_pure.c:20-30
 *   mean_packed diagonal walk, stride dim, dim-1, ...     src/meas/meas_packed.c:23-51
 *   mean_tiled  diagonal tiles (k,k), stride TILE_DIM+1   src/meas/meas_tiled.c:29-69
 *   matel_diag_pure sum obs[x] conj(bra_x) ket_x          src/meas/meas_pure.c:32-46
 * The sums are plain left-to-right accumulations per OpenMP chunk; their
 * rounding depends on the thread count, exactly as in the reference.
 */
#include "meas.h"
#include "orc_index.h"

#include <stdio.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MV(cond, msg) do { if (!(cond)) { fprintf(stderr, "orkan: meas: %s\n", (msg)); exit(EXIT_FAILURE); } } while (0)

static double mean_pure_o(const state_t *st, const double *obs) {
    const idx_t dim = (idx_t)1 << st->qubits;
    double sum = 0.0;
    #pragma omp parallel for reduction(+ : sum) schedule(static) if (dim >= OMP_THRESHOLD_PURE)
    for (idx_t x = 0; x < dim; ++x) {
        const cplx_t a = st->data[x];
        sum += obs[x] * (creal(a) * creal(a) + cimag(a) * cimag(a));
    }
    return sum;
}

static double mean_packed_o(const state_t *st, const double *obs) {
    const idx_t dim = (idx_t)1 << st->qubits;
    double sum = 0.0;
    #pragma omp parallel reduction(+ : sum) if (dim >= OMP_THRESHOLD_MIXED)
    {
        idx_t nth = 1, tid = 0;
#ifdef _OPENMP
        nth = (idx_t)omp_get_num_threads();
        tid = (idx_t)omp_get_thread_num();
#endif
        const idx_t chunk = dim / nth;
        const idx_t x0 = tid * chunk, x1 = (tid == nth - 1) ? dim : x0 + chunk;
        idx_t at = pk_idx(dim, x0, x0), stride = dim - x0;
        for (idx_t x = x0; x < x1; ++x) {
            sum += obs[x] * creal(st->data[at]);
            at += stride--;
        }
    }
    return sum;
}

static double mean_tiled_o(const state_t *st, const double *obs) {
    const idx_t dim = (idx_t)1 << st->qubits;
    const idx_t nt = (dim + TILE_DIM - 1) >> LOG_TILE_DIM;
    const idx_t td = dim < (idx_t)TILE_DIM ? dim : (idx_t)TILE_DIM;
    double sum = 0.0;
    #pragma omp parallel reduction(+ : sum) if (dim >= OMP_THRESHOLD_MIXED)
    {
        idx_t nth = 1, tid = 0;
#ifdef _OPENMP
        nth = (idx_t)omp_get_num_threads();
        tid = (idx_t)omp_get_thread_num();
#endif
        const idx_t chunk = nt / nth;
        const idx_t k0 = tid * chunk, k1 = (tid == nth - 1) ? nt : k0 + chunk;
        for (idx_t k = k0; k < k1; ++k) {
            const cplx_t *T = st->data + tl_tile(k, k);
            double dg[TILE_DIM];
            for (idx_t j = 0; j < td; ++j) dg[j] = creal(T[j * (TILE_DIM + 1)]);
            for (idx_t j = 0; j < td; ++j) sum += obs[k * td + j] * dg[j];
        }
    }
    return sum;
}

double mean(const state_t *state, const double *obs) {
    MV(state && state->data, "mean: null state or data pointer");
    MV(obs, "mean: null observable pointer");
    switch (state->type) {
    case PURE: return mean_pure_o(state, obs);
    case MIXED_PACKED: return mean_packed_o(state, obs);
    case MIXED_TILED: return mean_tiled_o(state, obs);
    default: MV(0, "mean: invalid state type"); return 0.0;
    }
}

cplx_t matel_diag(const state_t *bra, const state_t *ket, const double *obs) {
    MV(bra && bra->data, "matel_diag: null bra state or data pointer");
    MV(ket && ket->data, "matel_diag: null ket state or data pointer");
    MV(obs, "matel_diag: null observable pointer");
    MV(bra->qubits == ket->qubits, "matel_diag: qubit count mismatch");
    MV(bra->type == PURE, "matel_diag: bra must be PURE (mixed not supported)");
    MV(ket->type == PURE, "matel_diag: ket must be PURE (mixed not supported)");
    const idx_t dim = (idx_t)1 << bra->qubits;
    double sre = 0.0, sim = 0.0;
    #pragma omp parallel for reduction(+ : sre, sim) schedule(static) if (dim >= OMP_THRESHOLD_PURE)
    for (idx_t x = 0; x < dim; ++x) {
        const double ar = creal(bra->data[x]), ai = cimag(bra->data[x]);
        const double br = creal(ket->data[x]), bi = cimag(ket->data[x]);
        sre += obs[x] * (ar * br + ai * bi);
        sim += obs[x] * (ar * bi - ai * br);
    }
    return CMPLX(sre, sim);
}
