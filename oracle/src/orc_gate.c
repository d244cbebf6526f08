/*
 * orc_gate.c -- CPU oracle of the 18 named gates on all three layouts, the
 * generic tiled single_from_mat / two_from_mat, and the 1-qubit traversals
 * shared with the channel.  TEST INFRASTRUCTURE only.
 *
 * Follows:
 *   validation messages and exit()        src/gate/gate.c:19-24, 115-219
 *   pure 1q pair traversal and ops        src/gate/gate_pure.c:33-267
 *   pure cx/cy/cz/swap/ccx                src/gate/gate_pure.c:282-434
 *   packed 2x2 block traversal            src/gate/packed/gate_packed_1q.c:65-104
 *   tiled within-tile / cross-tile a,b,c  src/gate/tiled/gate_tiled_h.c:28-225
 *   packed/tiled cx, cy, cz, swap, ccx    src/gate/packed/gate_packed_{cx,cy,cz,swap,ccx}.c,
 *                                         src/gate/tiled/gate_tiled_{cx,cy,cz,swap,ccx}.c
 *     These gates only move entries and multiply them by +-1 or +-i, so the
 *     oracle states them once, as a permutation-with-phase of each group of
 *     4^k entries (the reference spells out every placement case; the values
 *     it produces are the same moves).
 *   single_from_mat                       src/gate/tiled/gate_tiled_1q.c:3
 *   two_from_mat (4x4 column-major U in the hi(x)lo basis of the sorted pair)
 *                                         src/gate/tiled/gate_tiled_2q.c:3-80
 */
#include "gate.h"
#include "orc_index.h"
#include "orc_ops.h"
#include "orc_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GV(cond, msg) do { if (!(cond)) { fprintf(stderr, "orkan: gate: %s\n", (msg)); exit(EXIT_FAILURE); } } while (0)

/* ------------------------------------------------------------------------------------------------
 * PURE
 * ---------------------------------------------------------------------------------------------- */

static void pure_1q(state_t *st, unsigned q, const op1_t *op_dm, int kind, double c, double s) {
    (void)op_dm;
    const idx_t dim = (idx_t)1 << st->qubits;
    const idx_t stride = (idx_t)1 << q;
    const idx_t step = stride << 1;
    cplx_t *d = st->data;
    #pragma omp parallel for collapse(2) if (dim >= OMP_THRESHOLD_PURE)
    for (idx_t i = 0; i < dim; i += step) {
        for (idx_t j = 0; j < stride; ++j) {
            cplx_t *a = d + i + j, *b = a + stride;
            const double ar = creal(*a), ai = cimag(*a), br = creal(*b), bi = cimag(*b);
            switch (kind) {
            case OP_X: { cplx_t tmp = *a; *a = *b; *b = tmp; break; }
            case OP_Y: *a = CMPLX(bi, -br); *b = CMPLX(-ai, ar); break;
            case OP_Z: *b = -*b; break;
            case OP_H: { cplx_t sm = *a + *b, df = *a - *b; *a = sm * M_SQRT1_2; *b = df * M_SQRT1_2; break; }
            case OP_HY: { cplx_t df = *a - *b, sm = *a + *b; *a = df * M_SQRT1_2; *b = sm * M_SQRT1_2; break; }
            case OP_S: *b = CMPLX(-bi, br); break;
            case OP_SDG: *b = CMPLX(bi, -br); break;
            case OP_T: *b = CMPLX((br - bi) * M_SQRT1_2, (br + bi) * M_SQRT1_2); break;
            case OP_TDG: *b = CMPLX((br + bi) * M_SQRT1_2, (bi - br) * M_SQRT1_2); break;
            case OP_P: *b = CMPLX(c * br - s * bi, s * br + c * bi); break;
            case OP_RX: {
                cplx_t na = c * (*a) + CMPLX(s * bi, -s * br);
                cplx_t nb = CMPLX(s * ai, -s * ar) + c * (*b);
                *a = na; *b = nb; break;
            }
            case OP_RY: {
                cplx_t na = c * (*a) - s * (*b);
                cplx_t nb = s * (*a) + c * (*b);
                *a = na; *b = nb; break;
            }
            case -1: /* RZ on a statevector: e^{-i theta/2} on |0>, e^{+i theta/2} on |1> */
                *a = CMPLX(c * ar + s * ai, c * ai - s * ar);
                *b = CMPLX(c * br - s * bi, c * bi + s * br);
                break;
            default: break;
            }
        }
    }
}

/* Permutation-with-phase of one k-qubit subspace: basis index x (bit j <-> qs[j])
 * goes to perm[x] with factor ph[x] in {1, -1, i, -i}. */
typedef struct {
    int k;
    unsigned qs[3];
    int perm[8];
    cplx_t ph[8];
} mono_t;

static void mono_identity(mono_t *m, int k) {
    m->k = k;
    for (int x = 0; x < 8; ++x) { m->perm[x] = x; m->ph[x] = 1.0; }
}

static idx_t spread(int x, const unsigned *qs, int k) {
    idx_t v = 0;
    for (int j = 0; j < k; ++j) if (x >> j & 1) v |= (idx_t)1 << qs[j];
    return v;
}

static idx_t ins_sorted(idx_t v, const unsigned *qs, int k) {
    unsigned srt[3];
    memcpy(srt, qs, sizeof(unsigned) * (size_t)k);
    for (int i = 0; i < k; ++i)
        for (int j = i + 1; j < k; ++j)
            if (srt[j] < srt[i]) { unsigned tq = srt[i]; srt[i] = srt[j]; srt[j] = tq; }
    for (int i = 0; i < k; ++i) v = ins0(v, srt[i]);
    return v;
}

static void pure_mono(state_t *st, const mono_t *m) {
    const idx_t dim = (idx_t)1 << st->qubits;
    const int K = 1 << m->k;
    const idx_t nb = dim >> m->k;
    cplx_t *d = st->data;
    #pragma omp parallel for if (dim >= OMP_THRESHOLD_PURE)
    for (idx_t b = 0; b < nb; ++b) {
        const idx_t base = ins_sorted(b, m->qs, m->k);
        cplx_t in[8];
        for (int x = 0; x < K; ++x) in[x] = d[base | spread(x, m->qs, m->k)];
        for (int x = 0; x < K; ++x) {
            if (m->perm[x] == x && m->ph[x] == 1.0) continue;
            const cplx_t f = m->ph[x];
            cplx_t v = in[x];
            if (f == -1.0) v = -v;
            else if (f == I) v = CMPLX(-cimag(v), creal(v));
            else if (f == -I) v = CMPLX(cimag(v), -creal(v));
            d[base | spread(m->perm[x], m->qs, m->k)] = v;
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * MIXED: 1-qubit traversals (shared with channel_1q)
 * ---------------------------------------------------------------------------------------------- */

void orc_packed_1q(state_t *st, unsigned q, const op1_t *op) {
    const idx_t dim = (idx_t)1 << st->qubits;
    const idx_t incr = (idx_t)1 << q;
    const idx_t half = dim >> 1;
    cplx_t *d = st->data;
    #pragma omp parallel for schedule(static) if (dim >= OMP_THRESHOLD_MIXED)
    for (idx_t bc = 0; bc < half; ++bc) {
        const idx_t c0 = ins0(bc, q), c1 = c0 | incr;
        for (idx_t br = bc; br < half; ++br) {
            const idx_t r0 = ins0(br, q), r1 = r0 | incr;
            const idx_t i00 = pk_idx(dim, r0, c0), i10 = pk_idx(dim, r1, c0), i11 = pk_idx(dim, r1, c1);
            const int lower01 = r0 >= c1;
            const idx_t i01 = lower01 ? pk_idx(dim, r0, c1) : pk_idx(dim, c1, r0);
            const int diag = (br == bc);
            cplx_t v[4], n[4];
            v[0] = d[i00]; v[2] = d[i10]; v[3] = d[i11];
            v[1] = diag ? conj(v[2]) : (lower01 ? d[i01] : conj(d[i01]));
            op1_apply(op, v, n);
            d[i00] = n[0]; d[i11] = n[3]; d[i10] = n[2];
            if (!diag) d[i01] = lower01 ? n[1] : conj(n[1]);
        }
    }
}

void orc_tiled_1q(state_t *st, unsigned q, const op1_t *op) {
    const idx_t dim = (idx_t)1 << st->qubits;
    const idx_t nt = (dim + TILE_DIM - 1) >> LOG_TILE_DIM;
    const idx_t td = dim < (idx_t)TILE_DIM ? dim : (idx_t)TILE_DIM;
    cplx_t *d = st->data;
    if (q < LOG_TILE_DIM) {
        const idx_t incr = (idx_t)1 << q, nb = td >> 1;
        #pragma omp parallel for schedule(static, 1) if (dim >= OMP_THRESHOLD_MIXED)
        for (idx_t tr = 0; tr < nt; ++tr) {
            for (idx_t tc = 0; tc <= tr; ++tc) {
                cplx_t *T = d + tl_tile(tr, tc);
                for (idx_t br = 0; br < nb; ++br) {
                    const idx_t r0 = ins0(br, q), r1 = r0 | incr;
                    for (idx_t bc = 0; bc < nb; ++bc) {
                        const idx_t c0 = ins0(bc, q), c1 = c0 | incr;
                        cplx_t v[4], n[4];
                        v[0] = T[tl_elem(r0, c0)]; v[1] = T[tl_elem(r0, c1)];
                        v[2] = T[tl_elem(r1, c0)]; v[3] = T[tl_elem(r1, c1)];
                        op1_apply(op, v, n);
                        T[tl_elem(r0, c0)] = n[0]; T[tl_elem(r0, c1)] = n[1];
                        T[tl_elem(r1, c0)] = n[2]; T[tl_elem(r1, c1)] = n[3];
                    }
                }
            }
        }
        return;
    }
    const unsigned qt = q - LOG_TILE_DIM;
    const idx_t it = (idx_t)1 << qt, nb = nt >> 1;
    #pragma omp parallel for schedule(static, 1) if (dim >= OMP_THRESHOLD_MIXED)
    for (idx_t btr = 0; btr < nb; ++btr) {
        const idx_t tr0 = ins0(btr, qt), tr1 = tr0 | it;
        for (idx_t btc = 0; btc <= btr; ++btc) {
            const idx_t tc0 = ins0(btc, qt), tc1 = tc0 | it;
            cplx_t *t00 = d + tl_tile(tr0, tc0), *t10 = d + tl_tile(tr1, tc0), *t11 = d + tl_tile(tr1, tc1);
            if (tr0 > tc1) {                    /* (a) all four tiles stored */
                cplx_t *t01 = d + tl_tile(tr0, tc1);
                for (idx_t e = 0; e < (idx_t)TILE_SIZE; ++e) {
                    cplx_t v[4] = {t00[e], t01[e], t10[e], t11[e]}, n[4];
                    op1_apply(op, v, n);
                    t00[e] = n[0]; t01[e] = n[1]; t10[e] = n[2]; t11[e] = n[3];
                }
            } else if (btr != btc) {            /* (b) T(tr0,tc1) stored transposed at T(tc1,tr0) */
                cplx_t *tx = d + tl_tile(tc1, tr0);
                for (idx_t lr = 0; lr < (idx_t)TILE_DIM; ++lr) {
                    for (idx_t lc = 0; lc < (idx_t)TILE_DIM; ++lc) {
                        const idx_t e = tl_elem(lr, lc), et = tl_elem(lc, lr);
                        cplx_t v[4] = {t00[e], conj(tx[et]), t10[e], t11[e]}, n[4];
                        op1_apply(op, v, n);
                        t00[e] = n[0]; t11[e] = n[3]; t10[e] = n[2]; tx[et] = conj(n[1]);
                    }
                }
            } else {                            /* (c) diagonal block: t01 shares storage with t10 */
                for (idx_t lr = 0; lr < (idx_t)TILE_DIM; ++lr) {
                    for (idx_t lc = 0; lc <= lr; ++lc) {
                        const idx_t e = tl_elem(lr, lc), et = tl_elem(lc, lr);
                        cplx_t v[4], n[4];
                        v[0] = t00[e]; v[2] = t10[e]; v[3] = t11[e];
                        v[1] = (lr == lc) ? conj(v[2]) : conj(t10[et]);
                        op1_apply(op, v, n);
                        t00[e] = n[0]; t11[e] = n[3];
                        if (lr != lc) { t00[et] = conj(n[0]); t11[et] = conj(n[3]); }
                        t10[e] = n[2];
                        if (lr != lc) t10[et] = conj(n[1]);
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * MIXED: k-qubit groups (cx, cy, cz, swap, ccx, two_from_mat)
 *
 * A group is the 2^k x 2^k block of entries (base_r | spread(a), base_c | spread(b)) whose
 * bases have the gate bits cleared.  Groups with base_r >= base_c cover every stored entry
 * once (the mirror of a group with base_r < base_c is its conjugate transpose).  Entries are
 * read through Hermitian access and written back with state_set semantics, which also keeps
 * the upper triangle of diagonal tiles mirrored.
 * ---------------------------------------------------------------------------------------------- */

typedef void (*grp_fn)(const cplx_t *in, cplx_t *out, int K, const void *ctx);

static void dm_groups(state_t *st, const unsigned *qs, int k, grp_fn fn, const void *ctx) {
    const idx_t dim = (idx_t)1 << st->qubits;
    const idx_t nb = dim >> k;
    const int K = 1 << k;
    idx_t off[8];
    for (int a = 0; a < K; ++a) off[a] = spread(a, qs, k);
    #pragma omp parallel for schedule(dynamic, 1) if (dim >= OMP_THRESHOLD_MIXED)
    for (idx_t br = 0; br < nb; ++br) {
        const idx_t r0 = ins_sorted(br, qs, k);
        for (idx_t bc = 0; bc <= br; ++bc) {
            const idx_t c0 = ins_sorted(bc, qs, k);
            cplx_t in[64], out[64];
            for (int a = 0; a < K; ++a)
                for (int b = 0; b < K; ++b) in[a * K + b] = state_get(st, r0 | off[a], c0 | off[b]);
            fn(in, out, K, ctx);
            for (int a = 0; a < K; ++a)
                for (int b = 0; b < K; ++b) {
                    const idx_t r = r0 | off[a], c = c0 | off[b];
                    if (br == bc && r < c) continue;   /* mirror of an entry of the same group */
                    state_set(st, r, c, out[a * K + b]);
                }
        }
    }
}

static cplx_t mul_ph(cplx_t v, cplx_t f) {
    if (f == 1.0) return v;
    if (f == -1.0) return -v;
    if (f == I) return CMPLX(-cimag(v), creal(v));
    return CMPLX(cimag(v), -creal(v));   /* f == -i */
}

static void grp_mono(const cplx_t *in, cplx_t *out, int K, const void *ctx) {
    const mono_t *m = ctx;
    for (int a = 0; a < K; ++a)
        for (int b = 0; b < K; ++b) {
            /* rho'(perm a, perm b) = ph_a conj(ph_b) rho(a, b) */
            const cplx_t v = mul_ph(mul_ph(in[a * K + b], m->ph[a]), conj(m->ph[b]));
            out[m->perm[a] * K + m->perm[b]] = v;
        }
}

static void grp_dense(const cplx_t *in, cplx_t *out, int K, const void *ctx) {
    const cplx_t *U = ctx;   /* row-major K x K */
    for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c) {
            cplx_t acc = 0.0;
            for (int a = 0; a < K; ++a)
                for (int b = 0; b < K; ++b) acc += U[r * K + a] * in[a * K + b] * conj(U[c * K + b]);
            out[r * K + c] = acc;
        }
}

static void apply_mono(state_t *st, const mono_t *m) {
    if (st->type == PURE) pure_mono(st, m);
    else dm_groups(st, m->qs, m->k, grp_mono, m);
}

/* 2-qubit superoperator on every 4x4 block (channel_2q): local index bit 0 <-> q1, bit 1 <-> q2;
 * S row-major on vec(block), out[(a,b)] = sum_(c,d) S[(a,b),(c,d)] in[(c,d)] */
static void grp_sop2(const cplx_t *in, cplx_t *out, int K, const void *ctx) {
    const cplx_t *S = ctx;
    (void)K;
    for (int i = 0; i < 16; ++i) {
        cplx_t acc = 0.0;
        for (int j = 0; j < 16; ++j) acc += S[i * 16 + j] * in[j];
        out[i] = acc;
    }
}

void orc_dm_sop2(state_t *st, unsigned q1, unsigned q2, const cplx_t *S) {
    const unsigned qs[2] = {q1, q2};
    dm_groups(st, qs, 2, grp_sop2, S);
}

/* ------------------------------------------------------------------------------------------------
 * Public gates
 * ---------------------------------------------------------------------------------------------- */

static void apply_1q(state_t *st, unsigned q, op_kind_t kind, double theta, int rz) {
    if (st->type == PURE) {
        double c = 0.0, s = 0.0;
        if (kind == OP_P) { c = cos(theta); s = sin(theta); }
        if (kind == OP_RX || kind == OP_RY || rz) { c = cos(theta / 2.0); s = sin(theta / 2.0); }
        pure_1q(st, q, NULL, rz ? -1 : (int)kind, c, s);
        return;
    }
    const op1_t op = op1_make(kind, theta);
    if (st->type == MIXED_PACKED) orc_packed_1q(st, q, &op);
    else orc_tiled_1q(st, q, &op);
}

#define CHECK_1Q(name, st, tq) \
    GV((st) && (st)->data, #name ": null state or data pointer"); \
    GV((tq) < (st)->qubits, #name ": target qubit out of range"); \
    GV((st)->type == PURE || (st)->type == MIXED_PACKED || (st)->type == MIXED_TILED, #name ": unknown state type")

void x(state_t *st, const qubit_t q) { CHECK_1Q(x, st, q); apply_1q(st, q, OP_X, 0, 0); }
void y(state_t *st, const qubit_t q) { CHECK_1Q(y, st, q); apply_1q(st, q, OP_Y, 0, 0); }
void z(state_t *st, const qubit_t q) { CHECK_1Q(z, st, q); apply_1q(st, q, OP_Z, 0, 0); }
void h(state_t *st, const qubit_t q) { CHECK_1Q(h, st, q); apply_1q(st, q, OP_H, 0, 0); }
void s(state_t *st, const qubit_t q) { CHECK_1Q(s, st, q); apply_1q(st, q, OP_S, 0, 0); }
void sdg(state_t *st, const qubit_t q) { CHECK_1Q(sdg, st, q); apply_1q(st, q, OP_SDG, 0, 0); }
void t(state_t *st, const qubit_t q) { CHECK_1Q(t, st, q); apply_1q(st, q, OP_T, 0, 0); }
void tdg(state_t *st, const qubit_t q) { CHECK_1Q(tdg, st, q); apply_1q(st, q, OP_TDG, 0, 0); }
void hy(state_t *st, const qubit_t q) { CHECK_1Q(hy, st, q); apply_1q(st, q, OP_HY, 0, 0); }
void rx(state_t *st, const qubit_t q, const double th) { CHECK_1Q(rx, st, q); apply_1q(st, q, OP_RX, th, 0); }
void ry(state_t *st, const qubit_t q, const double th) { CHECK_1Q(ry, st, q); apply_1q(st, q, OP_RY, th, 0); }
void rz(state_t *st, const qubit_t q, const double th) {
    CHECK_1Q(rz, st, q);
    if (st->type == PURE) apply_1q(st, q, OP_P, th, 1);
    else apply_1q(st, q, OP_P, th, 0);        /* rz_packed delegates to p_packed */
}
void p(state_t *st, const qubit_t q, const double th) { CHECK_1Q(p, st, q); apply_1q(st, q, OP_P, th, 0); }

#define CHECK_2Q(name, st, a, b, an, bn) \
    GV((st) && (st)->data, #name ": null state or data pointer"); \
    GV((a) < (st)->qubits, #name ": " an " qubit out of range"); \
    GV((b) < (st)->qubits, #name ": " bn " qubit out of range"); \
    GV((a) != (b), #name ": " an " and " bn " must differ"); \
    GV((st)->type == PURE || (st)->type == MIXED_PACKED || (st)->type == MIXED_TILED, #name ": unknown state type")

void cx(state_t *st, const qubit_t c, const qubit_t tq) {
    CHECK_2Q(cx, st, c, tq, "control", "target");
    mono_t m; mono_identity(&m, 2); m.qs[0] = c; m.qs[1] = tq;
    m.perm[1] = 3; m.perm[3] = 1;
    apply_mono(st, &m);
}
void cy(state_t *st, const qubit_t c, const qubit_t tq) {
    CHECK_2Q(cy, st, c, tq, "control", "target");
    mono_t m; mono_identity(&m, 2); m.qs[0] = c; m.qs[1] = tq;
    m.perm[1] = 3; m.ph[1] = I;     /* |c=1,t=0> -> i |c=1,t=1> */
    m.perm[3] = 1; m.ph[3] = -I;    /* |c=1,t=1> -> -i |c=1,t=0> */
    apply_mono(st, &m);
}
void cz(state_t *st, const qubit_t c, const qubit_t tq) {
    CHECK_2Q(cz, st, c, tq, "control", "target");
    mono_t m; mono_identity(&m, 2); m.qs[0] = c; m.qs[1] = tq;
    m.ph[3] = -1.0;
    apply_mono(st, &m);
}
void swap_gate(state_t *st, const qubit_t q1, const qubit_t q2) {
    GV(st && st->data, "swap: null state or data pointer");
    GV(q1 < st->qubits, "swap: q1 qubit out of range");
    GV(q2 < st->qubits, "swap: q2 qubit out of range");
    GV(q1 != q2, "swap: q1 and q2 must differ");
    GV(st->type == PURE || st->type == MIXED_PACKED || st->type == MIXED_TILED, "swap: unknown state type");
    mono_t m; mono_identity(&m, 2); m.qs[0] = q1; m.qs[1] = q2;
    m.perm[1] = 2; m.perm[2] = 1;
    apply_mono(st, &m);
}
void ccx(state_t *st, const qubit_t c1, const qubit_t c2, const qubit_t tq) {
    GV(st && st->data, "ccx: null state or data pointer");
    GV(c1 < st->qubits, "ccx: ctrl1 qubit out of range");
    GV(c2 < st->qubits, "ccx: ctrl2 qubit out of range");
    GV(tq < st->qubits, "ccx: target qubit out of range");
    GV(c1 != c2, "ccx: ctrl1 and ctrl2 must differ");
    GV(c1 != tq, "ccx: ctrl1 and target must differ");
    GV(c2 != tq, "ccx: ctrl2 and target must differ");
    GV(st->type == PURE || st->type == MIXED_PACKED || st->type == MIXED_TILED, "ccx: unknown state type");
    mono_t m; mono_identity(&m, 3); m.qs[0] = c1; m.qs[1] = c2; m.qs[2] = tq;
    m.perm[3] = 7; m.perm[7] = 3;
    apply_mono(st, &m);
}

/* ------------------------------------------------------------------------------------------------
 * Generic matrices (exported backend symbols of the reference, tiled layout)
 * ---------------------------------------------------------------------------------------------- */

void single_from_mat(state_t *st, const qubit_t target, const cplx_t *mat) {
    op1_t op = op1_make(OP_MAT, 0.0);
    for (int i = 0; i < 4; ++i) op.m[i] = mat[i];
    op.m00_2 = pow(cabs(mat[0]), 2);
    op.m01_2 = pow(cabs(mat[2]), 2);
    op.m10_2 = pow(cabs(mat[1]), 2);
    op.m11_2 = pow(cabs(mat[3]), 2);
    if (st->type == MIXED_PACKED) orc_packed_1q(st, target, &op);
    else orc_tiled_1q(st, target, &op);
}

void two_from_mat(state_t *st, const qubit_t q1, const qubit_t q2, const cplx_t *mat) {
    /* mat is column-major in the basis (hi bit, lo bit) of the sorted pair */
    const unsigned lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
    const unsigned qs[2] = {lo, hi};
    cplx_t U[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) U[r * 4 + c] = mat[r + 4 * c];
    dm_groups(st, qs, 2, grp_dense, U);
}
