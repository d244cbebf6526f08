/*
 * orc_state.c -- CPU oracle: state allocation, |+> initialisation and
 * Hermitian element access for the three layouts.  TEST INFRASTRUCTURE only.
 *
 * Follows:
 *   state_len / per-layout lengths   src/state/state.c:43-57, state_pure.c:25-27,
 *                                    state_packed.c:36-54, state_tiled.c:68-81
 *   state_init (64 B aligned, zeroed, ownership transfer)  src/state/state.c:26-99
 *   state_free                       src/state/state.c:102-108
 *   state_plus (every entry 1/sqrt(N) or 1/D; tiled n<LOG_TILE_DIM fills the
 *               D x D corner of the single padded tile)   state.c:111-127,
 *                                    state_pure.c:30-43, state_packed.c:57-76,
 *                                    state_tiled.c:84-114
 *   state_cp                         src/state/state.c:130-148
 *   state_get / state_set (upper triangle through conj; tiled set mirrors
 *               inside diagonal tiles) state.c:170-205, state_packed.c:79-104,
 *                                    state_tiled.c:117-150
 */
#include "state.h"
#include "orc_index.h"

#include <assert.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define ORC_ALIGN 64

static cplx_t *orc_alloc(idx_t len) {
    size_t bytes = (size_t)len * sizeof(cplx_t);
    bytes = (bytes + ORC_ALIGN - 1) & ~(size_t)(ORC_ALIGN - 1);
    cplx_t *p = aligned_alloc(ORC_ALIGN, bytes);
    if (p) memset(p, 0, bytes);
    return p;
}

static idx_t len_of(state_type_t type, qubit_t n) {
    const idx_t dim = (idx_t)1 << n;
    switch (type) {
    case PURE:
        return dim;
    case MIXED_PACKED:
        return n == 0 ? 1 : (dim >> 1) * (dim + 1);
    case MIXED_TILED: {
        const idx_t nt = (dim + TILE_DIM - 1) >> LOG_TILE_DIM;
        return (nt * (nt + 1) >> 1) * (idx_t)TILE_SIZE;
    }
    default:
        assert(0 && "state_len(): invalid state type");
        return 0;
    }
}

idx_t state_len(const state_t *state) {
    assert(state != NULL);
    return len_of(state->type, state->qubits);
}

void state_init(state_t *state, qubit_t qubits, cplx_t **data) {
    assert(state != NULL);
    free(state->data);
    state->data = NULL;
    state->qubits = qubits;
    const idx_t len = state_len(state);
    if (data && *data) {
        if (((uintptr_t)*data & (ORC_ALIGN - 1)) == 0) {
            state->data = *data;
        } else {
            state->data = orc_alloc(len);
            if (!state->data) {
                fprintf(stderr, "state_init(): aligned realloc failed for %zu elements\n", (size_t)len);
                state->qubits = 0;
                free(*data);
                *data = NULL;
                return;
            }
            memcpy(state->data, *data, (size_t)len * sizeof(cplx_t));
            free(*data);
        }
        *data = NULL;
        return;
    }
    state->data = orc_alloc(len);
    if (!state->data) {
        fprintf(stderr, "state_init(): allocation failed for %zu elements\n", (size_t)len);
        state->qubits = 0;
    }
}

void state_free(state_t *state) {
    if (!state) return;
    free(state->data);
    state->data = NULL;
    state->qubits = 0;
}

void state_plus(state_t *state, qubit_t qubits) {
    assert(state != NULL);
    state_init(state, qubits, NULL);
    if (!state->data) return;
    const idx_t dim = (idx_t)1 << qubits;
    const idx_t len = state_len(state);
    cplx_t v;
    if (state->type == PURE) v = 1.0 / sqrt((double)dim);
    else v = 1.0 / (double)dim;
    if (state->type == MIXED_TILED && qubits < LOG_TILE_DIM) {
        for (idx_t r = 0; r < dim; ++r)
            for (idx_t c = 0; c < dim; ++c) state->data[r * TILE_DIM + c] = v;
        return;
    }
    for (idx_t i = 0; i < len; ++i) state->data[i] = v;
}

state_t state_cp(const state_t *state) {
    assert(state != NULL && state->data != NULL);
    state_t out = *state;
    const idx_t len = state_len(state);
    out.data = orc_alloc(len);
    if (!out.data) {
        fprintf(stderr, "state_cp(): allocation failed\n");
        out.qubits = 0;
        return out;
    }
    memcpy(out.data, state->data, (size_t)len * sizeof(cplx_t));
    return out;
}

cplx_t state_get(const state_t *state, idx_t row, idx_t col) {
    assert(state != NULL && state->data != NULL);
    const idx_t dim = (idx_t)1 << state->qubits;
    switch (state->type) {
    case PURE:
        assert(col == 0 && row < dim);
        return state->data[row];
    case MIXED_PACKED:
        assert(row < dim && col < dim);
        return row >= col ? state->data[pk_idx(dim, row, col)]
                          : conj(state->data[pk_idx(dim, col, row)]);
    case MIXED_TILED:
        assert(row < dim && col < dim);
        return row >= col ? state->data[tl_idx(row, col)]
                          : conj(state->data[tl_idx(col, row)]);
    default:
        assert(0 && "state_get(): invalid state type");
        return 0.0;
    }
}

void state_set(state_t *state, idx_t row, idx_t col, cplx_t val) {
    assert(state != NULL && state->data != NULL);
    const idx_t dim = (idx_t)1 << state->qubits;
    switch (state->type) {
    case PURE:
        assert(col == 0 && row < dim);
        state->data[row] = val;
        break;
    case MIXED_PACKED:
        assert(row < dim && col < dim);
        if (row >= col) state->data[pk_idx(dim, row, col)] = val;
        else state->data[pk_idx(dim, col, row)] = conj(val);
        break;
    case MIXED_TILED: {
        assert(row < dim && col < dim);
        const int same_tile = (row >> LOG_TILE_DIM) == (col >> LOG_TILE_DIM);
        if (row >= col) {
            state->data[tl_idx(row, col)] = val;
            if (same_tile && row != col) state->data[tl_idx(col, row)] = conj(val);
        } else {
            state->data[tl_idx(col, row)] = conj(val);
            if (same_tile) state->data[tl_idx(row, col)] = val;
        }
        break;
    }
    default:
        assert(0 && "state_set(): invalid state type");
    }
}

static void pr_cplx(cplx_t v) {
    printf("%.7f", creal(v));
    if (cimag(v) >= 0) printf("+i%.7f", cimag(v));
    else printf("-i%.7f", fabs(cimag(v)));
}

void state_print(const state_t *state) {
    assert(state != NULL);
    static const char *names[] = {"PURE", "MIXED_PACKED", "MIXED_TILED"};
    printf("Type: %s\nQubits: %u\n", names[state->type], state->qubits);
    const idx_t dim = (idx_t)1 << state->qubits;
    if (state->type == PURE) {
        printf("[");
        for (idx_t i = 0; i < dim; ++i) {
            pr_cplx(state->data[i]);
            if (i + 1 < dim) printf(", ");
        }
        printf("]\n");
        return;
    }
    printf("[");
    for (idx_t r = 0; r < dim; ++r) {
        if (r) printf(" ");
        printf("[");
        for (idx_t c = 0; c < dim; ++c) {
            pr_cplx(state_get(state, r, c));
            if (c + 1 < dim) printf(", ");
        }
        printf("]");
        if (r + 1 < dim) printf("\n");
    }
    printf("]\n");
}
