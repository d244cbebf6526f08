/*
 * ork_kernels.h -- C-ABI launchers of the gfx950 kernels, called by the C17
 * host layer of the library.  Plain pointers and sizes only.
 *
 * Every launcher enqueues on `stream` and returns a hipError_t value (0 on
 * success).  Pointers named *_dev are device (HBM) addresses.
 */
#ifndef ORK_KERNELS_H
#define ORK_KERNELS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element storage of a state: complex128 (the Orkan API) or complex64 (the
 * FP32-storage variant) */
enum { ORK_F64 = 0, ORK_F32 = 1 };

/* state layouts, numerically equal to state_type_t */
enum { ORK_PURE = 0, ORK_PACKED = 1, ORK_TILED = 2 };

/* statevector single-qubit op classes */
enum {
    ORK_SV_MAT = 0,    /* dense 2x2: a' = m00 a + m01 b, b' = m10 a + m11 b */
    ORK_SV_H = 1,      /* a' = (a+b)/sqrt2, b' = (a-b)/sqrt2 */
    ORK_SV_HY = 2,     /* a' = (a-b)/sqrt2, b' = (a+b)/sqrt2 */
    ORK_SV_X = 3,
    ORK_SV_Y = 4,
    ORK_SV_RX = 5,     /* coef: c, s */
    ORK_SV_RY = 6,     /* coef: c, s */
    ORK_SV_RZ = 7,     /* coef: c, s: a *= (c - i s), b *= (c + i s) */
    ORK_SV_P = 8,      /* coef: c, s: b *= (c + i s) */
    ORK_SV_Z = 9,
    ORK_SV_S = 10,
    ORK_SV_SDG = 11,
    ORK_SV_T = 12,
    ORK_SV_TDG = 13
};

/* density-matrix group ops */
enum {
    ORK_DM_SOP = 0,    /* 1 qubit, 4x4 superoperator (row-major on (r00, r01, r10, r11)) */
    ORK_DM_MONO = 1,   /* k <= 3 qubits, basis permutation with phases in {1, i, -1, -i} */
    ORK_DM_DENSE = 2,  /* k <= 3 qubits, dense 2^k x 2^k unitary, rho -> U rho U^dagger */
    ORK_DM_SOP2 = 3    /* 2 qubits, 16x16 superoperator on the 4x4 block (row-major on vec(block),
                        * local index bit j <-> qs[j]); `mat` is then a DEVICE pointer to 256 complex */
};

/* Monomial (permutation-with-phase) action of a k-qubit gate:
 * U |a> = i^ph[a] |perm[a]>, local index bit j <-> qubit qs[j]. */
typedef struct {
    int k;
    int qs[3];
    int perm[8];
    int ph[8];
} ork_mono_t;

int ork_fill(void *stream, int prec, void *data_dev, uint64_t len, double re, double im);
int ork_fill_tiled_corner(void *stream, int prec, void *data_dev, int n, int log_tile, double v);

int ork_sv_1q(void *stream, int prec, void *psi_dev, int n, int q, int op, const double *coef);
/* every qubit's 2x2 matrix (mats: n x 8 doubles, row-major, re/im) in LDS-tiled sweeps */
int ork_sv_layer(void *stream, int prec, void *psi_dev, int n, const double *mats);
int ork_sv_mono(void *stream, int prec, void *psi_dev, int n, const ork_mono_t *m);
int ork_sv_exp_diag(void *stream, int prec, void *psi_dev, int n, const double *diag_dev, double t);
int ork_sv_mean(void *stream, int prec, const void *psi_dev, int n, const double *obs_dev,
                double *partials_dev, double *out_host);
int ork_sv_matel(void *stream, int prec, const void *bra_dev, const void *ket_dev, int n,
                 const double *obs_dev, double *partials_dev, double *out_host);

/* rho -> op(rho) on the packed or tiled layout */
int ork_dm_group(void *stream, int prec, int layout, int log_tile, void *rho_dev, int n,
                 int op, const ork_mono_t *m, const double *mat);
/* rho_rc *= f[r_Q] conj(f[c_Q]) for a diagonal k-qubit gate (f: 2^k complex) */
int ork_dm_phase(void *stream, int prec, int layout, int log_tile, void *rho_dev, int n,
                 int k, const int *qs, const double *f);
/* rho_rc *= e^{-i (d_r - d_c) t};  phase_dev: scratch of 2^n complex128 */
int ork_dm_exp_diag(void *stream, int prec, int layout, int log_tile, void *rho_dev, int n,
                    const double *diag_dev, double t, void *phase_dev);
int ork_dm_mean(void *stream, int prec, int layout, int log_tile, const void *rho_dev, int n,
                const double *obs_dev, double *partials_dev, double *out_host);

/* Allocation-free exp_diag: the phase scratch of a density matrix is a fixed device array of
 * 2^ORK_NOALLOC_MAX_QUBITS complex128 (k_noalloc.hip).  ork_noalloc_setup takes that array's
 * address on the current device (once per device, at runtime creation); ork_exp_diag_noalloc
 * launches exp_diag with it (hipErrorInvalidValue for a density matrix above the maximum,
 * hipErrorNotInitialized before the set-up).  Calls that overlap on one GPU share the array. */
#define ORK_NOALLOC_MAX_QUBITS 17
int ork_noalloc_setup(void);
int ork_exp_diag_noalloc(void *stream, int prec, int layout, int log_tile, void *data_dev, int n,
                         const double *diag_dev, double t);

/* launch shape of the density-matrix kernels: which 0 = tile-group kernel, 1 = phase kernel;
 * xcd: keep consecutive groups on one XCD; lds_bytes: LDS per workgroup (caps residency) */
int ork_dm_set_shape(int which, int flags, int lds_bytes);
/* residency cap (dynamic LDS bytes per workgroup) of the statevector mean / matel reduction */
int ork_sv_set_red_lds(int flags, int lds_bytes);
int ork_sv_set_exp_shape(int flags, int lds_bytes);
int ork_sv_set_mono_lds32(int lds_bytes);
int ork_sv_set_perm_shape(int flags, int lds_bytes);
int ork_mlp_gemm_select(int variant);
int ork_sv_set_pair_shape(int flags, int lds_bytes);
int ork_sv_set_nt_off(int mask, int min_bytes);
int ork_sv_set_layer_variant(int v);
unsigned long long ork_nt_min_bytes(void);
void ork_dm_reset_shapes(void);
void ork_sv_reset_tuning(void);
void ork_mlp_gemm_reset(void);

/* number of doubles the *_mean / matel launchers need in partials_dev */
uint64_t ork_partials_len(void);

/* dst (packed or tiled, len stored entries) <- the logical density matrix of src (pure: |psi><psi|,
 * packed or tiled), same precision */
int ork_relayout(void *stream, int prec, const void *src, int src_layout, void *dst, int dst_layout, int n,
                 int log_tile, uint64_t len);

/* complex128 <-> complex64 conversion of a whole storage array */
int ork_convert(void *stream, const void *src_dev, int src_prec, void *dst_dev, int dst_prec, uint64_t len);

#ifdef __cplusplus
}
#endif
#endif /* ORK_KERNELS_H */
