/*
 * orc_ops.h -- per-gate 2x2 block formulas of the density-matrix update
 * rho -> U rho U^dagger, plus the 4x4 superoperator step of a 1-qubit channel.
 * TEST INFRASTRUCTURE (CPU oracle) only.
 *
 * Block convention (src/gate/packed/gate_packed_1q.c:8-16): for rows r0 < r1
 * and columns c0 < c1 that differ only in the target bit,
 *   v[0] = rho(r0,c0), v[1] = rho(r0,c1), v[2] = rho(r1,c0), v[3] = rho(r1,c1).
 * The closed forms follow the reference block macros:
 *   X  gate_packed_1q.c:117-134     Y   :150-167    Z   :178-186
 *   H  gate_packed_1q.c:205-225     HY  :240-258    S   :272-293
 *   SDG gate_packed_1q.c:306-327    T   :347-371    TDG :384-408
 *   P / RZ (off-diagonals times e^{+-i theta}) gate_packed_1q.c:436-468, 698-700
 *   RX gate_packed_1q.c:491-539     RY  :552-594
 *   generic 2x2 U (column-major mat) src/gate/tiled/gate_tiled_1q.c:3-60
 *   superoperator (row-major 4x4 on vec(rho)) src/channel/channel_packed.c:8-46
 */
#ifndef ORC_OPS_H
#define ORC_OPS_H

#include "q_types.h"
#include <complex.h>
#include <math.h>

#ifndef M_SQRT1_2
#define M_SQRT1_2 0.70710678118654752440
#endif

typedef enum {
    OP_X, OP_Y, OP_Z, OP_H, OP_HY, OP_S, OP_SDG, OP_T, OP_TDG,
    OP_P, OP_RX, OP_RY, OP_MAT, OP_SOP
} op_kind_t;

typedef struct {
    op_kind_t kind;
    double c, s, c2, s2, cs;   /* trig constants of the rotation / phase */
    cplx_t m[4];               /* OP_MAT: column-major U; derived products below */
    cplx_t m00_2, m01_2, m10_2, m11_2;
    const cplx_t *sop;         /* OP_SOP: 16 entries, row-major */
} op1_t;

static inline void op1_apply(const op1_t *op, const cplx_t v[4], cplx_t n[4]) {
    const cplx_t v00 = v[0], v01 = v[1], v10 = v[2], v11 = v[3];
    n[0] = v00; n[1] = v01; n[2] = v10; n[3] = v11;
    switch (op->kind) {
    case OP_X:
        n[0] = v11; n[1] = v10; n[2] = v01; n[3] = v00;
        break;
    case OP_Y:
        n[0] = v11; n[1] = -v10; n[2] = -v01; n[3] = v00;
        break;
    case OP_Z:
        n[0] = v00; n[1] = -v01; n[2] = -v10; n[3] = v11;
        break;
    case OP_H:
        n[0] = 0.5 * (v00 + v01 + v10 + v11);
        n[3] = 0.5 * (v00 - v01 - v10 + v11);
        n[1] = 0.5 * (v00 - v01 + v10 - v11);
        n[2] = 0.5 * (v00 + v01 - v10 - v11);
        break;
    case OP_HY:
        n[0] = 0.5 * (v00 - v01 - v10 + v11);
        n[3] = 0.5 * (v00 + v01 + v10 + v11);
        n[1] = 0.5 * (v00 + v01 - v10 - v11);
        n[2] = 0.5 * (v00 - v01 + v10 - v11);
        break;
    case OP_S:    /* rho10 *= i, rho01 *= -i */
        n[0] = v00; n[3] = v11;
        n[2] = CMPLX(-cimag(v10), creal(v10));
        n[1] = CMPLX(cimag(v01), -creal(v01));
        break;
    case OP_SDG:  /* rho10 *= -i, rho01 *= i */
        n[0] = v00; n[3] = v11;
        n[2] = CMPLX(cimag(v10), -creal(v10));
        n[1] = CMPLX(-cimag(v01), creal(v01));
        break;
    case OP_T: {  /* rho10 *= e^{i pi/4}, rho01 *= e^{-i pi/4} */
        const double a = creal(v10), b = cimag(v10), e = creal(v01), f = cimag(v01);
        n[0] = v00; n[3] = v11;
        n[2] = CMPLX((a - b) * M_SQRT1_2, (a + b) * M_SQRT1_2);
        n[1] = CMPLX((e + f) * M_SQRT1_2, (f - e) * M_SQRT1_2);
        break;
    }
    case OP_TDG: {
        const double a = creal(v10), b = cimag(v10), e = creal(v01), f = cimag(v01);
        n[0] = v00; n[3] = v11;
        n[2] = CMPLX((a + b) * M_SQRT1_2, (b - a) * M_SQRT1_2);
        n[1] = CMPLX((e - f) * M_SQRT1_2, (e + f) * M_SQRT1_2);
        break;
    }
    case OP_P: {  /* rho10 *= e^{i theta}, rho01 *= e^{-i theta} */
        const double c = op->c, s = op->s;
        const double a = creal(v10), b = cimag(v10), e = creal(v01), f = cimag(v01);
        n[0] = v00; n[3] = v11;
        n[2] = CMPLX(a * c - b * s, b * c + a * s);
        n[1] = CMPLX(e * c + f * s, f * c - e * s);
        break;
    }
    case OP_RX: {
        const double cs = op->cs, c2 = op->c2, s2 = op->s2;
        const cplx_t d1 = v01 - v10, d2 = v00 - v11;
        const cplx_t i1 = CMPLX(-cs * cimag(d1), cs * creal(d1));
        const cplx_t i2 = CMPLX(-cs * cimag(d2), cs * creal(d2));
        n[0] = c2 * v00 + s2 * v11 + i1;
        n[3] = s2 * v00 + c2 * v11 - i1;
        n[1] = i2 + c2 * v01 + s2 * v10;
        n[2] = -i2 + c2 * v10 + s2 * v01;
        break;
    }
    case OP_RY: {
        const double cs = op->cs, c2 = op->c2, s2 = op->s2;
        const cplx_t sm = v01 + v10, d2 = v00 - v11;
        n[0] = c2 * v00 + s2 * v11 - cs * sm;
        n[3] = s2 * v00 + c2 * v11 + cs * sm;
        n[1] = cs * d2 + c2 * v01 - s2 * v10;
        n[2] = cs * d2 + c2 * v10 - s2 * v01;
        break;
    }
    case OP_MAT: {
        const cplx_t *m = op->m;
        n[0] = v00 * op->m00_2 + v01 * m[0] * conj(m[2]) + v10 * m[2] * conj(m[0]) + v11 * op->m01_2;
        n[1] = v00 * m[0] * conj(m[1]) + v01 * m[0] * conj(m[3]) + v10 * m[2] * conj(m[1]) + v11 * m[2] * conj(m[3]);
        n[2] = v00 * m[1] * conj(m[0]) + v01 * m[1] * conj(m[2]) + v10 * m[3] * conj(m[0]) + v11 * m[3] * conj(m[2]);
        n[3] = v00 * op->m10_2 + v01 * m[1] * conj(m[3]) + v10 * m[3] * conj(m[1]) + v11 * op->m11_2;
        break;
    }
    case OP_SOP: {
        const cplx_t *S = op->sop;
        for (int i = 0; i < 4; ++i) {
            double re = 0.0, im = 0.0;
            for (int j = 0; j < 4; ++j) {
                re += creal(S[4 * i + j]) * creal(v[j]);
                re -= cimag(S[4 * i + j]) * cimag(v[j]);
                im += creal(S[4 * i + j]) * cimag(v[j]);
                im += cimag(S[4 * i + j]) * creal(v[j]);
            }
            n[i] = CMPLX(re, im);
        }
        break;
    }
    }
}

/* Constants of a rotation/phase op (the reference evaluates cos/sin of theta/2
 * for RX/RY, and of theta for P/RZ on mixed states). */
static inline op1_t op1_make(op_kind_t kind, double theta) {
    op1_t op;
    op.kind = kind;
    op.c = op.s = op.c2 = op.s2 = op.cs = 0.0;
    op.sop = 0;
    if (kind == OP_RX || kind == OP_RY) {
        op.c = cos(theta / 2.0);
        op.s = sin(theta / 2.0);
        op.c2 = op.c * op.c;
        op.s2 = op.s * op.s;
        op.cs = op.c * op.s;
    } else if (kind == OP_P) {
        op.c = cos(theta);
        op.s = sin(theta);
    }
    return op;
}

#endif /* ORC_OPS_H */
