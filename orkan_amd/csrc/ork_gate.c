/*
 * ork_gate.c -- the 18 named gates, the generic single_from_mat/two_from_mat,
 * and their per-layout backends, dispatched onto the gfx950 kernels.
 *
 * Validation and messages follow src/gate/gate.c:19-219 (exit(EXIT_FAILURE)
 * after "orkan: gate: <name>: ...").  Gate constants (cos/sin of the angle)
 * are evaluated here on the host with libm, as the reference does, so the
 * kernels receive bit-identical coefficients.
 *
 * Kernel choice per gate:
 *   PURE         pair kernel (1 qubit) or permutation-with-phase kernel (cx, cy,
 *                cz, swap, ccx): src/gate/gate_pure.c:33-434
 *   mixed, diagonal gates (z, s, sdg, t, tdg, p, rz, cz): elementwise phase
 *                kernel, only the entries whose factor is not 1 are touched
 *   mixed, x / y: streamed superoperator kernels (exact moves: entries 0, +-1, +-i)
 *   mixed, cx / cy / swap / ccx: tile-group kernel, permutation with phases
 *                applied during the store; entries the gate leaves unchanged are
 *                neither read nor written
 *   mixed, h / hy / rx / ry / single_from_mat: tile-group kernel with the
 *                gate's 4x4 superoperator; its entries are the coefficients of
 *                the reference's closed forms (src/gate/packed/gate_packed_1q.c:205-594)
 *   two_from_mat: tile-group kernel, dense 4x4 U rho U^dagger
 */
#include "ork_host.h"
#include "orkan_backend.h"

#include <math.h>
#include <string.h>

#ifndef M_SQRT1_2
#define M_SQRT1_2 0.70710678118654752440
#endif

static void *stream(void) { return ork_dev()->stream; }

static int L_of(void) { return LOG_TILE_DIM; }

/* ---- statevector -------------------------------------------------------------------------- */

static void sv1(state_t *s, qubit_t q, int op, double c, double sn) {
    double coef[8] = {c, sn, 0, 0, 0, 0, 0, 0};
    ORK_HIP("gate", ork_sv_1q(stream(), ork_prec(s), s->data, s->qubits, q, op, coef));
}

static void mono(state_t *s, int k, const int *qs, const int *perm, const int *ph) {
    ork_mono_t m;
    memset(&m, 0, sizeof(m));
    m.k = k;
    for (int j = 0; j < k; ++j) m.qs[j] = qs[j];
    for (int a = 0; a < (1 << k); ++a) {
        m.perm[a] = perm[a];
        m.ph[a] = ph[a];
    }
    if (ork_layout(s) == PURE) {
        ORK_HIP("gate", ork_sv_mono(stream(), ork_prec(s), s->data, s->qubits, &m));
        return;
    }
    if (k == 1) {
        /* x, y: the superoperator S[(i,j),(k,l)] = U[i][k] conj(U[j][l]) of U[perm a][a] = i^ph[a] has
         * entries 0, +-1, +-i only, so the streamed superoperator kernels move every entry exactly */
        static const double IRE[4] = {1.0, 0.0, -1.0, 0.0}, IIM[4] = {0.0, 1.0, 0.0, -1.0};
        double ure[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, uim[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        for (int a = 0; a < 2; ++a) {
            ure[perm[a]][a] = IRE[ph[a] & 3];
            uim[perm[a]][a] = IIM[ph[a] & 3];
        }
        double S[32];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int kk = 0; kk < 2; ++kk)
                    for (int l = 0; l < 2; ++l) {
                        /* U[i][kk] * conj(U[j][l]) */
                        const double ar = ure[i][kk], ai = uim[i][kk], br = ure[j][l], bi = -uim[j][l];
                        S[2 * ((2 * i + j) * 4 + 2 * kk + l)] = ar * br - ai * bi;
                        S[2 * ((2 * i + j) * 4 + 2 * kk + l) + 1] = ar * bi + ai * br;
                    }
        ORK_HIP("gate", ork_dm_group(stream(), ork_prec(s), ork_layout(s), L_of(), s->data, s->qubits, ORK_DM_SOP, &m, S));
        return;
    }
    ORK_HIP("gate", ork_dm_group(stream(), ork_prec(s), ork_layout(s), L_of(), s->data, s->qubits, ORK_DM_MONO, &m, NULL));
}

/* ---- density matrix ------------------------------------------------------------------------ */

/* rho -> diag(f) rho diag(f)^dagger on qubits qs (f: 2^k complex, interleaved re/im) */
static void dm_phase(state_t *s, int k, const int *qs, const double *f) {
    ORK_HIP("gate", ork_dm_phase(stream(), ork_prec(s), ork_layout(s), L_of(), s->data, s->qubits, k, qs, f));
}

static void dm_phase1(state_t *s, qubit_t q, double re, double im) {
    const int qs[1] = {q};
    const double f[4] = {1.0, 0.0, re, im};
    dm_phase(s, 1, qs, f);
}

/* 4x4 superoperator, 16 complex row-major (re, im interleaved) */
static void dm_sop(state_t *s, qubit_t q, const double *S) {
    ork_mono_t m;
    memset(&m, 0, sizeof(m));
    m.k = 1;
    m.qs[0] = q;
    ORK_HIP("gate", ork_dm_group(stream(), ork_prec(s), ork_layout(s), L_of(), s->data, s->qubits, ORK_DM_SOP, &m, S));
}

/* set row r of a 4x4 real-coefficient superoperator; entries given as (re, im) pairs */
#define ROW(S, r, a0, b0, a1, b1, a2, b2, a3, b3) \
    do { double *R_ = (S) + 8 * (r); R_[0] = a0; R_[1] = b0; R_[2] = a1; R_[3] = b1; \
         R_[4] = a2; R_[5] = b2; R_[6] = a3; R_[7] = b3; } while (0)

static void sop_h(double *S) {
    /* new00 = (v00+v01+v10+v11)/2, new01 = (v00-v01+v10-v11)/2,
     * new10 = (v00+v01-v10-v11)/2, new11 = (v00-v01-v10+v11)/2 */
    ROW(S, 0, .5, 0, .5, 0, .5, 0, .5, 0);
    ROW(S, 1, .5, 0, -.5, 0, .5, 0, -.5, 0);
    ROW(S, 2, .5, 0, .5, 0, -.5, 0, -.5, 0);
    ROW(S, 3, .5, 0, -.5, 0, -.5, 0, .5, 0);
}

static void sop_hy(double *S) {
    ROW(S, 0, .5, 0, -.5, 0, -.5, 0, .5, 0);
    ROW(S, 1, .5, 0, .5, 0, -.5, 0, -.5, 0);
    ROW(S, 2, .5, 0, -.5, 0, .5, 0, -.5, 0);
    ROW(S, 3, .5, 0, .5, 0, .5, 0, .5, 0);
}

static void sop_rx(double *S, double theta) {
    const double c = cos(theta / 2.0), s = sin(theta / 2.0);
    const double c2 = c * c, s2 = s * s, cs = c * s;
    /* new00 = c2 v00 + s2 v11 + i cs (v01 - v10); new01 = i cs (v00 - v11) + c2 v01 + s2 v10
     * new10 = -i cs (v00 - v11) + c2 v10 + s2 v01; new11 = s2 v00 + c2 v11 - i cs (v01 - v10) */
    ROW(S, 0, c2, 0, 0, cs, 0, -cs, s2, 0);
    ROW(S, 1, 0, cs, c2, 0, s2, 0, 0, -cs);
    ROW(S, 2, 0, -cs, s2, 0, c2, 0, 0, cs);
    ROW(S, 3, s2, 0, 0, -cs, 0, cs, c2, 0);
}

static void sop_ry(double *S, double theta) {
    const double c = cos(theta / 2.0), s = sin(theta / 2.0);
    const double c2 = c * c, s2 = s * s, cs = c * s;
    ROW(S, 0, c2, 0, -cs, 0, -cs, 0, s2, 0);
    ROW(S, 1, cs, 0, c2, 0, -s2, 0, -cs, 0);
    ROW(S, 2, cs, 0, -s2, 0, c2, 0, -cs, 0);
    ROW(S, 3, s2, 0, cs, 0, cs, 0, c2, 0);
}

/* S[(i,j),(k,l)] = U[i][k] conj(U[j][l]) for a 2x2 U given column-major (single_from_mat) */
static void sop_from_mat(double *S, const cplx_t *m) {
    cplx_t U[2][2] = {{m[0], m[2]}, {m[1], m[3]}};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                for (int l = 0; l < 2; ++l) {
                    const cplx_t v = U[i][k] * conj(U[j][l]);
                    S[2 * ((2 * i + j) * 4 + 2 * k + l)] = creal(v);
                    S[2 * ((2 * i + j) * 4 + 2 * k + l) + 1] = cimag(v);
                }
}

/* ---- backends ------------------------------------------------------------------------------ */

static const int ID2[2] = {0, 1};
static const int PH0[8] = {0, 0, 0, 0, 0, 0, 0, 0};

#define DEF_PURE1(name, OP) \
    void name##_pure(state_t *s, qubit_t q) { sv1(s, q, OP, 0.0, 0.0); }
DEF_PURE1(x, ORK_SV_X)
DEF_PURE1(y, ORK_SV_Y)
DEF_PURE1(z, ORK_SV_Z)
DEF_PURE1(h, ORK_SV_H)
DEF_PURE1(s, ORK_SV_S)
DEF_PURE1(sdg, ORK_SV_SDG)
DEF_PURE1(t, ORK_SV_T)
DEF_PURE1(tdg, ORK_SV_TDG)
DEF_PURE1(hy, ORK_SV_HY)
void rx_pure(state_t *s, qubit_t q, double th) { sv1(s, q, ORK_SV_RX, cos(th / 2.0), sin(th / 2.0)); }
void ry_pure(state_t *s, qubit_t q, double th) { sv1(s, q, ORK_SV_RY, cos(th / 2.0), sin(th / 2.0)); }
void rz_pure(state_t *s, qubit_t q, double th) { sv1(s, q, ORK_SV_RZ, cos(th / 2.0), sin(th / 2.0)); }
void p_pure(state_t *s, qubit_t q, double th) { sv1(s, q, ORK_SV_P, cos(th), sin(th)); }

static void dm_x(state_t *s, qubit_t q) {
    const int qs[1] = {q}, perm[2] = {1, 0};
    mono(s, 1, qs, perm, PH0);
}
static void dm_y(state_t *s, qubit_t q) {
    /* Y|0> = i|1>, Y|1> = -i|0> */
    const int qs[1] = {q}, perm[2] = {1, 0}, ph[2] = {1, 3};
    mono(s, 1, qs, perm, ph);
}
static void dm_h(state_t *s, qubit_t q) { double S[32]; sop_h(S); dm_sop(s, q, S); }
static void dm_hy(state_t *s, qubit_t q) { double S[32]; sop_hy(S); dm_sop(s, q, S); }
static void dm_rx(state_t *s, qubit_t q, double th) { double S[32]; sop_rx(S, th); dm_sop(s, q, S); }
static void dm_ry(state_t *s, qubit_t q, double th) { double S[32]; sop_ry(S, th); dm_sop(s, q, S); }
static void dm_z(state_t *s, qubit_t q) { dm_phase1(s, q, -1.0, 0.0); }
static void dm_s(state_t *s, qubit_t q) { dm_phase1(s, q, 0.0, 1.0); }
static void dm_sdg(state_t *s, qubit_t q) { dm_phase1(s, q, 0.0, -1.0); }
static void dm_t(state_t *s, qubit_t q) { dm_phase1(s, q, M_SQRT1_2, M_SQRT1_2); }
static void dm_tdg(state_t *s, qubit_t q) { dm_phase1(s, q, M_SQRT1_2, -M_SQRT1_2); }
/* p and rz coincide on a density matrix: off-diagonals times e^{+-i theta}
 * (src/gate/packed/gate_packed_1q.c:698-700) */
static void dm_p(state_t *s, qubit_t q, double th) { dm_phase1(s, q, cos(th), sin(th)); }

#define DEF_DM1(name)                                                   \
    void name##_packed(state_t *s, qubit_t q) { dm_##name(s, q); }      \
    void name##_tiled(state_t *s, qubit_t q) { dm_##name(s, q); }
DEF_DM1(x) DEF_DM1(y) DEF_DM1(z) DEF_DM1(h) DEF_DM1(s) DEF_DM1(sdg) DEF_DM1(t) DEF_DM1(tdg) DEF_DM1(hy)
void rx_packed(state_t *s, qubit_t q, double th) { dm_rx(s, q, th); }
void rx_tiled(state_t *s, qubit_t q, double th) { dm_rx(s, q, th); }
void ry_packed(state_t *s, qubit_t q, double th) { dm_ry(s, q, th); }
void ry_tiled(state_t *s, qubit_t q, double th) { dm_ry(s, q, th); }
void rz_packed(state_t *s, qubit_t q, double th) { dm_p(s, q, th); }
void rz_tiled(state_t *s, qubit_t q, double th) { dm_p(s, q, th); }
void p_packed(state_t *s, qubit_t q, double th) { dm_p(s, q, th); }
void p_tiled(state_t *s, qubit_t q, double th) { dm_p(s, q, th); }

/* two- and three-qubit gates: local index bit j <-> argument j */
static void g_cx(state_t *s, qubit_t c, qubit_t t) {
    const int qs[2] = {c, t}, perm[4] = {0, 3, 2, 1};
    mono(s, 2, qs, perm, PH0);
}
static void g_cy(state_t *s, qubit_t c, qubit_t t) {
    /* control set: |t=0> -> i|t=1>, |t=1> -> -i|t=0> (src/gate/gate_pure.c:316-337) */
    const int qs[2] = {c, t}, perm[4] = {0, 3, 2, 1}, ph[4] = {0, 1, 0, 3};
    mono(s, 2, qs, perm, ph);
}
static void g_cz(state_t *s, qubit_t c, qubit_t t) {
    const int qs[2] = {c, t}, perm[4] = {0, 1, 2, 3}, ph[4] = {0, 0, 0, 2};
    if (ork_layout(s) == PURE) {
        mono(s, 2, qs, perm, ph);
        return;
    }
    const double f[8] = {1, 0, 1, 0, 1, 0, -1, 0};
    dm_phase(s, 2, qs, f);
}
static void g_swap(state_t *s, qubit_t a, qubit_t b) {
    const int qs[2] = {a, b}, perm[4] = {0, 2, 1, 3};
    mono(s, 2, qs, perm, PH0);
}
static void g_ccx(state_t *s, qubit_t c1, qubit_t c2, qubit_t t) {
    const int qs[3] = {c1, c2, t}, perm[8] = {0, 1, 2, 7, 4, 5, 6, 3};
    mono(s, 3, qs, perm, PH0);
}
#define DEF_2Q(name)                                                            \
    void name##_pure(state_t *s, qubit_t a, qubit_t b) { g_##name(s, a, b); }   \
    void name##_packed(state_t *s, qubit_t a, qubit_t b) { g_##name(s, a, b); } \
    void name##_tiled(state_t *s, qubit_t a, qubit_t b) { g_##name(s, a, b); }
DEF_2Q(cx) DEF_2Q(cy) DEF_2Q(cz) DEF_2Q(swap)
void ccx_pure(state_t *s, qubit_t a, qubit_t b, qubit_t c) { g_ccx(s, a, b, c); }
void ccx_packed(state_t *s, qubit_t a, qubit_t b, qubit_t c) { g_ccx(s, a, b, c); }
void ccx_tiled(state_t *s, qubit_t a, qubit_t b, qubit_t c) { g_ccx(s, a, b, c); }

/* single_from_mat / two_from_mat are the reference's internal tiled backends
 * (src/gate/tiled/gate_tiled_1q.c:3, src/gate/tiled/gate_tiled_2q.c:3), which validate nothing; here
 * they also serve the packed layout, and a bad argument stops the program with a message like every
 * other gate instead of reaching the device */
#define MV_MIXED(name, st)                                                                        \
    ORK_VALIDATE("gate", (st) && (st)->data, name ": null state or data pointer");                \
    ORK_VALIDATE("gate", ork_type_ok(st) && ork_layout(st) != PURE,                               \
                 name ": state must be MIXED_PACKED or MIXED_TILED")

void single_from_mat(state_t *s, qubit_t target, const cplx_t *mat) {
    MV_MIXED("single_from_mat", s);
    ORK_VALIDATE("gate", target < s->qubits, "single_from_mat: target qubit out of range");
    ORK_VALIDATE("gate", mat, "single_from_mat: null matrix pointer");
    double S[32];
    sop_from_mat(S, mat);
    dm_sop(s, target, S);
}

void two_from_mat(state_t *s, qubit_t q1, qubit_t q2, const cplx_t *mat) {
    MV_MIXED("two_from_mat", s);
    ORK_VALIDATE("gate", q1 < s->qubits, "two_from_mat: q1 qubit out of range");
    ORK_VALIDATE("gate", q2 < s->qubits, "two_from_mat: q2 qubit out of range");
    ORK_VALIDATE("gate", q1 != q2, "two_from_mat: q1 and q2 must differ");
    ORK_VALIDATE("gate", mat, "two_from_mat: null matrix pointer");
    const qubit_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
    ork_mono_t m;
    memset(&m, 0, sizeof(m));
    m.k = 2;
    m.qs[0] = lo;   /* local index = lo_bit + 2 hi_bit, the (hi, lo) basis */
    m.qs[1] = hi;
    double U[32];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            U[2 * (r * 4 + c)] = creal(mat[r + 4 * c]);
            U[2 * (r * 4 + c) + 1] = cimag(mat[r + 4 * c]);
        }
    ORK_HIP("two_from_mat", ork_dm_group(stream(), ork_prec(s), ork_layout(s), L_of(), s->data, s->qubits,
                                         ORK_DM_DENSE, &m, U));
}

/* ---- public API with validation ------------------------------------------------------------- */

#define GV(cond, msg) ORK_VALIDATE("gate", cond, msg)

#define DISPATCH_1Q(name, st, q)                                                        \
    GV((st) && (st)->data, #name ": null state or data pointer");                       \
    GV((q) < (st)->qubits, #name ": target qubit out of range");                        \
    GV(ork_type_ok(st), #name ": unknown state type");                                  \
    switch (ork_layout(st)) {                                                           \
    case PURE: name##_pure(st, q); break;                                               \
    case MIXED_PACKED: name##_packed(st, q); break;                                     \
    default: name##_tiled(st, q); break;                                                \
    }

#define DISPATCH_ROT(name, st, q, th)                                                   \
    GV((st) && (st)->data, #name ": null state or data pointer");                       \
    GV((q) < (st)->qubits, #name ": target qubit out of range");                        \
    GV(ork_type_ok(st), #name ": unknown state type");                                  \
    switch (ork_layout(st)) {                                                           \
    case PURE: name##_pure(st, q, th); break;                                           \
    case MIXED_PACKED: name##_packed(st, q, th); break;                                 \
    default: name##_tiled(st, q, th); break;                                            \
    }

#define DISPATCH_2Q(name, fn, st, a, b, an, bn)                                         \
    GV((st) && (st)->data, #name ": null state or data pointer");                       \
    GV((a) < (st)->qubits, #name ": " an " qubit out of range");                        \
    GV((b) < (st)->qubits, #name ": " bn " qubit out of range");                        \
    GV((a) != (b), #name ": " an " and " bn " must differ");                            \
    GV(ork_type_ok(st), #name ": unknown state type");                                  \
    g_##fn(st, a, b);

void x(state_t *st, const qubit_t q) { DISPATCH_1Q(x, st, q) }
void y(state_t *st, const qubit_t q) { DISPATCH_1Q(y, st, q) }
void z(state_t *st, const qubit_t q) { DISPATCH_1Q(z, st, q) }
void h(state_t *st, const qubit_t q) { DISPATCH_1Q(h, st, q) }
void s(state_t *st, const qubit_t q) { DISPATCH_1Q(s, st, q) }
void sdg(state_t *st, const qubit_t q) { DISPATCH_1Q(sdg, st, q) }
void t(state_t *st, const qubit_t q) { DISPATCH_1Q(t, st, q) }
void tdg(state_t *st, const qubit_t q) { DISPATCH_1Q(tdg, st, q) }
void hy(state_t *st, const qubit_t q) { DISPATCH_1Q(hy, st, q) }
void rx(state_t *st, const qubit_t q, const double th) { DISPATCH_ROT(rx, st, q, th) }
void ry(state_t *st, const qubit_t q, const double th) { DISPATCH_ROT(ry, st, q, th) }
void rz(state_t *st, const qubit_t q, const double th) { DISPATCH_ROT(rz, st, q, th) }
void p(state_t *st, const qubit_t q, const double th) { DISPATCH_ROT(p, st, q, th) }
void cx(state_t *st, const qubit_t c, const qubit_t t) { DISPATCH_2Q(cx, cx, st, c, t, "control", "target") }
void cy(state_t *st, const qubit_t c, const qubit_t t) { DISPATCH_2Q(cy, cy, st, c, t, "control", "target") }
void cz(state_t *st, const qubit_t c, const qubit_t t) { DISPATCH_2Q(cz, cz, st, c, t, "control", "target") }
void swap_gate(state_t *st, const qubit_t a, const qubit_t b) { DISPATCH_2Q(swap, swap, st, a, b, "q1", "q2") }

void ccx(state_t *st, const qubit_t c1, const qubit_t c2, const qubit_t t) {
    GV(st && st->data, "ccx: null state or data pointer");
    GV(c1 < st->qubits, "ccx: ctrl1 qubit out of range");
    GV(c2 < st->qubits, "ccx: ctrl2 qubit out of range");
    GV(t < st->qubits, "ccx: target qubit out of range");
    GV(c1 != c2, "ccx: ctrl1 and ctrl2 must differ");
    GV(c1 != t, "ccx: ctrl1 and target must differ");
    GV(c2 != t, "ccx: ctrl2 and target must differ");
    GV(ork_type_ok(st), "ccx: unknown state type");
    g_ccx(st, c1, c2, t);
}

/* ---- fused layer of single-qubit gates (no reference counterpart) ---------------------------- */
qs_error_t orkan_layer_1q(state_t *s, const cplx_t *mats) {
    if (!s || !s->data || !mats) return QS_ERR_NULL;
    if (!ork_type_ok(s) || ork_layout(s) != PURE) return QS_ERR_TYPE;
    const int n = s->qubits;
    double *m = malloc((size_t)8 * (n ? n : 1) * sizeof(double));
    if (!m) return QS_ERR_OOM;
    for (int q = 0; q < n; ++q) {
        const cplx_t *c = mats + 4 * q;   /* column-major: c[0] = m00, c[1] = m10, c[2] = m01, c[3] = m11 */
        const cplx_t rm[4] = {c[0], c[2], c[1], c[3]};
        for (int i = 0; i < 4; ++i) {
            m[8 * q + 2 * i] = creal(rm[i]);
            m[8 * q + 2 * i + 1] = cimag(rm[i]);
        }
    }
    const int e = n ? ork_sv_layer(stream(), ork_prec(s), s->data, n, m) : 0;
    free(m);
    ORK_HIP("orkan_layer_1q", e);
    return QS_OK;
}
