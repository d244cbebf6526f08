/*
 * A caller of the Orkan C API written the way the reference's own users write one
 * (reference README.md:73-96, benchmark/src/bench_orkan.c:24-36): designated initialisers for
 * state_t / kraus_t, complex values returned by value (state_get, matel_diag), the caller's buffer
 * handed over with state_init, kraus_to_superop memory freed by the caller.  The same source is
 * linked against every library that exports the API; the tests compare their printed results.
 *
 * Output: one line per result, "<name> <re> <im>" with 17 significant digits.
 */
#include "orkan.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static void put(const char *name, idx_t i, cplx_t v) {
    printf("%s[%llu] %.17g %.17g\n", name, (unsigned long long)i, creal(v), cimag(v));
}

int main(void) {
    /* ---- the reference's quick start: 3-qubit pure state from |+>^3 ---- */
    state_t state = {.type = PURE, .data = NULL, .qubits = 0};
    state_plus(&state, 3);
    h(&state, 0);
    cx(&state, 0, 1);
    rz(&state, 2, 0.5);
    for (idx_t i = 0; i < state_len(&state); i++) put("amp", i, state_get(&state, i, 0));

    /* matel_diag and mean on a second state, exp_diag */
    double obs[8];
    for (int i = 0; i < 8; ++i) obs[i] = 0.25 * i - 0.8;
    state_t ket = state_cp(&state);
    ry(&ket, 1, 0.3);
    sdg(&ket, 2);
    put("matel", 0, matel_diag(&state, &ket, obs));
    put("mean", 0, mean(&state, obs));
    exp_diag(&state, obs, 0.7);
    for (idx_t i = 0; i < state_len(&state); i++) put("exp", i, state_get(&state, i, 0));

    /* a caller-owned 64-byte aligned buffer handed to state_init (ownership moves, *buf = NULL) */
    cplx_t *buf = aligned_alloc(64, 16 * sizeof(cplx_t));
    for (int i = 0; i < 16; ++i) buf[i] = (0.1 * i - 0.7) + I * (0.05 * (15 - i));
    state_t own = {.type = PURE, .data = NULL, .qubits = 0};
    state_init(&own, 4, &buf);
    printf("buf_after_init %s\n", buf == NULL ? "null" : "kept");
    swap_gate(&own, 0, 3);
    ccx(&own, 1, 2, 0);
    state_set(&own, 5, 0, 0.25 - 0.5 * I);
    for (idx_t i = 0; i < state_len(&own); i++) put("own", i, state_get(&own, i, 0));

    /* ---- mixed states: 6 qubits, tiled and packed ---- */
    cplx_t K[8] = {1.0, 0.0, 0.0, sqrt(1.0 - 0.2),   /* amplitude damping, gamma 0.2, row-major */
                   0.0, sqrt(0.2), 0.0, 0.0};
    kraus_t kraus = {.n_qubits = 1, .n_terms = 2, .data = K};
    superop_t sop = kraus_to_superop(&kraus);
    for (int layout = 0; layout < 2; ++layout) {
        state_t rho = {.type = layout ? MIXED_PACKED : MIXED_TILED, .data = NULL, .qubits = 0};
        state_plus(&rho, 6);
        ry(&rho, 0, 0.4);
        h(&rho, 5);
        cx(&rho, 0, 5);
        rz(&rho, 2, 0.9);
        ccx(&rho, 5, 0, 3);
        cy(&rho, 4, 1);
        channel_1q(&rho, &sop, 3);
        channel_1q(&rho, &sop, 5);
        double o64[64];
        for (int i = 0; i < 64; ++i) o64[i] = sin(0.3 * i);
        exp_diag(&rho, o64, 0.25);
        const char *nm = layout ? "packed" : "tiled";
        /* lower and upper triangle (returned as conj of the stored entry) */
        for (idx_t r = 0; r < 64; r += 7)
            for (idx_t c = 0; c < 64; c += 9) {
                printf("%s(%llu,%llu) ", nm, (unsigned long long)r, (unsigned long long)c);
                cplx_t v = state_get(&rho, r, c);
                printf("%.17g %.17g\n", creal(v), cimag(v));
            }
        printf("%s_mean %.17g 0\n", nm, mean(&rho, o64));
        state_free(&rho);
        printf("%s_freed %s %d\n", nm, rho.data == NULL ? "null" : "kept", (int)rho.qubits);
    }
    free(sop.data);
    state_free(&own);
    state_free(&ket);
    state_free(&state);
    state_free(&state);   /* idempotent */
    printf("done\n");
    return 0;
}
