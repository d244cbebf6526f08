/*
 * Drives the oracle's restatement of the reference CPU path (oracle/src) through every entry
 * point on all three layouts, at both tile sizes the build uses and at the small sizes where the
 * padded tile, the diagonal tiles and the cross-tile cases all occur, so that an AddressSanitizer /
 * UndefinedBehaviorSanitizer build of it reports any out-of-bounds access, use after free or
 * undefined operation (the reference builds its tests the same way with ENABLE_ASAN,
 * reference CMakeLists.txt:59, :83-91).  Values are not checked here: parity is the other tests'
 * business.  Exit status 0 when every call returned.
 */
#include "orkan.h"
#include "orkan_backend.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t rng_state = 0x4f524b414e000099ull;
static double urand(void) {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * 0x1.0p-53;
}

static void fill(state_t *st) {
    const idx_t dim = (idx_t)1 << st->qubits;
    if (st->type == PURE) {
        for (idx_t i = 0; i < dim; ++i) state_set(st, i, 0, (urand() - 0.5) + I * (urand() - 0.5));
        return;
    }
    for (idx_t r = 0; r < dim; ++r)
        for (idx_t c = 0; c <= r; ++c) state_set(st, r, c, r == c ? urand() : (urand() - 0.5) + I * (urand() - 0.5));
}

int main(void) {
    long calls = 0;
    cplx_t K[3 * 16];
    for (int i = 0; i < 3 * 16; ++i) K[i] = (urand() - 0.5) + I * (urand() - 0.5);
    kraus_t k1 = {.n_qubits = 1, .n_terms = 3, .data = K}, k2 = {.n_qubits = 2, .n_terms = 3, .data = K};
    superop_t s1 = kraus_to_superop(&k1), s2 = kraus_to_superop(&k2);
    cplx_t m2[4], m4[16];
    for (int i = 0; i < 4; ++i) m2[i] = urand() + I * urand();
    for (int i = 0; i < 16; ++i) m4[i] = urand() + I * urand();
    for (int type = 0; type < 3; ++type) {
        for (qubit_t n = 1; n <= 7; ++n) {
            state_t st = {.type = (state_type_t)type, .data = NULL, .qubits = 0};
            state_init(&st, n, NULL);
            fill(&st);
            double *obs = malloc(sizeof(double) << n);
            for (idx_t i = 0; i < ((idx_t)1 << n); ++i) obs[i] = urand() - 0.5;
            for (qubit_t q = 0; q < n; ++q) {
                x(&st, q); y(&st, q); z(&st, q); h(&st, q);
                s(&st, q); sdg(&st, q); t(&st, q); tdg(&st, q); hy(&st, q);
                rx(&st, q, 0.3); ry(&st, q, 1.1); rz(&st, q, -0.7); p(&st, q, 2.2);
                calls += 13;
                if (type != PURE) {
                    channel_1q(&st, &s1, q);
                    calls += 1;
                }
                if (type == MIXED_TILED) {
                    single_from_mat(&st, q, m2);
                    calls += 1;
                }
            }
            for (qubit_t a = 0; a < n; ++a)
                for (qubit_t b = 0; b < n; ++b) {
                    if (a == b) continue;
                    cx(&st, a, b); cy(&st, a, b); cz(&st, a, b); swap_gate(&st, a, b);
                    calls += 4;
                    if (type != PURE) {
                        channel_2q(&st, &s2, a, b);
                        calls += 1;
                    }
                    if (type == MIXED_TILED) {
                        two_from_mat(&st, a, b, m4);
                        calls += 1;
                    }
                    for (qubit_t c = 0; c < n; ++c)
                        if (c != a && c != b) {
                            ccx(&st, a, b, c);
                            ++calls;
                        }
                }
            (void)mean(&st, obs);
            exp_diag(&st, obs, 0.4);
            state_t c = state_cp(&st);
            if (type == PURE) (void)matel_diag(&st, &c, obs);
            (void)state_get(&c, ((idx_t)1 << n) - 1, 0);
            state_free(&c);
            state_free(&st);
            state_free(&st);
            free(obs);
            calls += 6;
        }
    }
    free(s1.data);
    free(s2.data);
    printf("oracle calls %ld\n", calls);
    return 0;
}
