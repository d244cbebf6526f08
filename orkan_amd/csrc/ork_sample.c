/*
 * ork_sample.c -- sample_matrix on the device (no reference implementation; the operation is
 * specified in the reference's docs/SAMPLE_MATRIX.md:1-80, its recommended algorithm in §6 there).
 *
 *   1. channels 0 .. lstar-1 are applied to iota in place (prefix shared by every leaf);
 *   2. leaves |psi_a> are built in blocks of B: copy of the prefix, M_lstar[a], then the
 *      channels after lstar, each as sum_m w[m] M[m] psi through two scratch vectors;
 *   3. for every block pair (i >= j) the block G = Phi_i^dagger diag(obs) Phi_j is one product
 *      kernel; entries with a >= b are kept and mirrored, the diagonal's imaginary part zeroed.
 * Peak: iota + 2 scratch + 2B leaves (B from the workspace budget, at most 64); every allocation
 * happens at entry.  The operations are the caller's qop_t, called on PURE states whose data are
 * the library's device buffers, so they run through this library's gates.
 */
#include "ork_host.h"
#include "sample.h"

#include <string.h>
#include <hip/hip_runtime_api.h>

unsigned ork_sm_max_block(void);
unsigned ork_sm_parts(uint64_t N);
int ork_sm_axpy(void *stream, void *acc, const void *v, double wr, double wi, uint64_t n, int first);
int ork_sm_block(void *stream, const void *P, int na, const void *Q, int nb, const double *obs, uint64_t N,
                 void *part, void *G);

typedef struct {
    void *stream;
    qubit_t n;
    idx_t N;
    cplx_t *tmp, *acc;       /* linear-combination scratch */
} lcu_t;

static void run_op(qop_t op, cplx_t *data, qubit_t n) {
    state_t s;
    s.type = PURE;
    s.data = data;
    s.qubits = n;
    op(&s);
}

/* psi <- sum_m w[m] M[m] psi */
static int apply_channel(lcu_t *c, cplx_t *psi, const ansatz_channel_t *ch) {
    const size_t bytes = (size_t)c->N * sizeof(cplx_t);
    for (uint64_t m = 0; m < ch->n_ops; ++m) {
        int e = (int)hipMemcpyAsync(c->tmp, psi, bytes, hipMemcpyDeviceToDevice, (hipStream_t)c->stream);
        if (e) return e;
        run_op(ch->ops[m], c->tmp, c->n);
        e = ork_sm_axpy(c->stream, c->acc, c->tmp, creal(ch->weights[m]), cimag(ch->weights[m]), c->N, m == 0);
        if (e) return e;
    }
    return (int)hipMemcpyAsync(psi, c->acc, bytes, hipMemcpyDeviceToDevice, (hipStream_t)c->stream);
}

/* leaves a0 .. a0 + cnt - 1 into blk (cnt vectors of N) */
static int build_leaves(lcu_t *c, cplx_t *blk, const cplx_t *prefix, const ansatz_channel_t *chs, uint64_t nch,
                        uint64_t lstar, uint64_t a0, int cnt) {
    const size_t bytes = (size_t)c->N * sizeof(cplx_t);
    for (int i = 0; i < cnt; ++i) {
        cplx_t *leaf = blk + (size_t)i * c->N;
        int e = (int)hipMemcpyAsync(leaf, prefix, bytes, hipMemcpyDeviceToDevice, (hipStream_t)c->stream);
        if (e) return e;
        run_op(chs[lstar].ops[a0 + (uint64_t)i], leaf, c->n);
        for (uint64_t l = lstar + 1; l < nch; ++l)
            if ((e = apply_channel(c, leaf, &chs[l])) != 0) return e;
    }
    return 0;
}

qs_error_t sample_matrix(state_t *iota, const double *obs, const ansatz_channel_t *channels, uint64_t n_channels,
                         uint64_t lstar, size_t workspace_bytes, cplx_t *S) {
    if (!iota || !iota->data || !obs || !channels || !S) return QS_ERR_NULL;
    if (iota->type != PURE) return QS_ERR_TYPE;
    if (lstar >= n_channels) return QS_ERR_PARAM;
    for (uint64_t l = 0; l < n_channels; ++l) {
        if (channels[l].n_ops == 0 || !channels[l].ops) return QS_ERR_PARAM;
        if (l != lstar && !channels[l].weights) return QS_ERR_PARAM;
        for (uint64_t m = 0; m < channels[l].n_ops; ++m)
            if (!channels[l].ops[m]) return QS_ERR_NULL;
    }
    const qubit_t n = iota->qubits;
    const idx_t N = (idx_t)1 << n;
    const uint64_t K = channels[lstar].n_ops;
    const size_t vec = (size_t)N * sizeof(cplx_t);
    const size_t nvec = workspace_bytes / vec;
    if (nvec < 4) return QS_ERR_PARAM;
    uint64_t B = (nvec - 2) / 2;
    if (B > ork_sm_max_block()) B = ork_sm_max_block();
    if (B > K) B = K;

    /* no library lock is held while the caller's operations run (they may call exp_diag or mean,
     * which stage their arrays in the shared buffer): obs gets a buffer of its own */
    ork_dev_t *d = ork_dev();
    const size_t part_bytes = (size_t)ork_sm_parts(N) * 64 * 64 * sizeof(cplx_t);
    cplx_t *tmp = NULL, *acc = NULL, *phj = NULL, *phi = NULL, *part = NULL, *G = NULL;
    double *o = NULL;
    cplx_t *Gh = malloc((size_t)B * B * sizeof(cplx_t));
    qs_error_t rc = Gh ? QS_OK : QS_ERR_OOM;
    if (rc == QS_OK && (hipMalloc((void **)&tmp, vec) != hipSuccess || hipMalloc((void **)&acc, vec) != hipSuccess ||
                        hipMalloc((void **)&phj, vec * B) != hipSuccess || hipMalloc((void **)&phi, vec * B) != hipSuccess ||
                        hipMalloc((void **)&part, part_bytes) != hipSuccess ||
                        hipMalloc((void **)&G, (size_t)B * B * sizeof(cplx_t)) != hipSuccess ||
                        hipMalloc((void **)&o, (size_t)N * sizeof(double)) != hipSuccess)) {
        (void)hipGetLastError();
        rc = QS_ERR_OOM;
    }
    int e = 0;
    if (rc == QS_OK)
        e = (int)hipMemcpyAsync(o, obs, (size_t)N * sizeof(double), hipMemcpyDefault, (hipStream_t)d->stream);
    lcu_t c = {d->stream, n, N, tmp, acc};
    /* 1. prefix */
    for (uint64_t l = 0; rc == QS_OK && !e && l < lstar; ++l) e = apply_channel(&c, iota->data, &channels[l]);
    /* 2-3. blocks */
    for (uint64_t j0 = 0; rc == QS_OK && !e && j0 < K; j0 += B) {
        const int nj = (int)(K - j0 < B ? K - j0 : B);
        e = build_leaves(&c, phj, iota->data, channels, n_channels, lstar, j0, nj);
        for (uint64_t i0 = j0; !e && i0 < K; i0 += B) {
            const int ni = (int)(K - i0 < B ? K - i0 : B);
            const cplx_t *P = phj;
            if (i0 != j0) {
                e = build_leaves(&c, phi, iota->data, channels, n_channels, lstar, i0, ni);
                P = phi;
            }
            if (!e) e = ork_sm_block(d->stream, P, ni, phj, nj, o, N, part, G);
            if (!e) e = (int)hipMemcpyAsync(Gh, G, (size_t)ni * nj * sizeof(cplx_t), hipMemcpyDeviceToHost,
                                            (hipStream_t)d->stream);
            if (!e) e = (int)hipStreamSynchronize((hipStream_t)d->stream);
            for (int a = 0; !e && a < ni; ++a)
                for (int b = 0; b < nj; ++b) {
                    const uint64_t ia = i0 + (uint64_t)a, jb = j0 + (uint64_t)b;
                    if (ia < jb) continue;
                    cplx_t v = Gh[a * nj + b];
                    if (ia == jb) v = CMPLX(creal(v), 0.0);
                    S[ia + jb * K] = v;
                    S[jb + ia * K] = conj(v);
                }
        }
    }
    (void)hipStreamSynchronize((hipStream_t)d->stream);
    void *bufs[] = {tmp, acc, phj, phi, part, G, o};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); ++i)
        if (bufs[i]) (void)hipFree(bufs[i]);
    free(Gh);
    if (e) ork_hip_fail("sample_matrix", e);   /* a failed HIP call exits like every other entry point */
    return rc;
}
