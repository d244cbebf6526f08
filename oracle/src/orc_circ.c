/*
 * orc_circ.c -- CPU oracle of exp_diag.  TEST INFRASTRUCTURE only.
 *
 * Follows:
 *   validation, t == 0 early return          src/circ/circ.c:23-28, 46-65
 *   exp_diag_pure  psi_x *= e^{-i d_x t}      src/circ/circ_pure.c:21-34
 *   exp_diag_packed per-index sincos, then the product identity
 *                  e^{-i(d_r-d_c)t} = (cr - i sr)(cc + i sc), strict lower triangle
 *                                            src/circ/circ_packed.c:28-71
 *   exp_diag_tiled  same factors; diagonal tiles update the strict lower
 *                  triangle and mirror it    src/circ/circ_tiled.c:25-105
 */
#define _GNU_SOURCE
#include "circ.h"
#include "orc_index.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define XV(cond, msg) do { if (!(cond)) { fprintf(stderr, "orkan: circ: %s\n", (msg)); exit(EXIT_FAILURE); } } while (0)

static void phases(idx_t dim, const double *diag, double t, double **cs, double **sn) {
    *cs = malloc(dim * sizeof(double));
    *sn = malloc(dim * sizeof(double));
    if (!*cs || !*sn) {
        free(*cs);
        free(*sn);
        fprintf(stderr, "orkan: circ: exp_diag: allocation failed\n");
        exit(EXIT_FAILURE);
    }
    double *c = *cs, *s = *sn;
    #pragma omp parallel for schedule(static) if (dim >= OMP_THRESHOLD_MIXED)
    for (idx_t x = 0; x < dim; ++x) sincos(diag[x] * t, &s[x], &c[x]);
}

static inline cplx_t rot(cplx_t v, double cr, double sr, double cc, double sc) {
    const double fre = cr * cc + sr * sc;
    const double fim = -sr * cc + cr * sc;
    const double re = creal(v), im = cimag(v);
    return CMPLX(fre * re - fim * im, fre * im + fim * re);
}

void exp_diag(state_t *state, const double *diag, double t) {
    XV(state && state->data, "exp_diag: null state or data pointer");
    XV(diag, "exp_diag: null diagonal pointer");
    if (t == 0.0) return;
    const idx_t dim = (idx_t)1 << state->qubits;
    cplx_t *d = state->data;
    if (state->type == PURE) {
        #pragma omp parallel for schedule(static) if (dim >= OMP_THRESHOLD_PURE)
        for (idx_t x = 0; x < dim; ++x) {
            const double a = diag[x] * t, c = cos(a), s = sin(a);
            const double re = creal(d[x]), im = cimag(d[x]);
            d[x] = CMPLX(c * re + s * im, c * im - s * re);
        }
        return;
    }
    XV(state->type == MIXED_PACKED || state->type == MIXED_TILED, "exp_diag: invalid state type");
    double *cs, *sn;
    phases(dim, diag, t, &cs, &sn);
    if (state->type == MIXED_PACKED) {
        #pragma omp parallel for schedule(guided) if (dim >= OMP_THRESHOLD_MIXED)
        for (idx_t c = 0; c < dim; ++c) {
            idx_t at = pk_col(dim, c) + 1;
            for (idx_t r = c + 1; r < dim; ++r, ++at) d[at] = rot(d[at], cs[r], sn[r], cs[c], sn[c]);
        }
    } else {
        const idx_t nt = (dim + TILE_DIM - 1) >> LOG_TILE_DIM;
        const idx_t td = dim < (idx_t)TILE_DIM ? dim : (idx_t)TILE_DIM;
        #pragma omp parallel for schedule(dynamic, 1) if (dim >= OMP_THRESHOLD_MIXED)
        for (idx_t tr = 0; tr < nt; ++tr) {
            for (idx_t tc = 0; tc <= tr; ++tc) {
                cplx_t *T = d + tl_tile(tr, tc);
                for (idx_t lr = 0; lr < td; ++lr) {
                    const idx_t r = tr * td + lr;
                    const idx_t lc_end = (tr == tc) ? lr : td;
                    for (idx_t lc = 0; lc < lc_end; ++lc) {
                        const idx_t c = tc * td + lc;
                        T[tl_elem(lr, lc)] = rot(T[tl_elem(lr, lc)], cs[r], sn[r], cs[c], sn[c]);
                        if (tr == tc) T[tl_elem(lc, lr)] = conj(T[tl_elem(lr, lc)]);
                    }
                }
            }
        }
    }
    free(cs);
    free(sn);
}
