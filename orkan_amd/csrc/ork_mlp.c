/*
 * ork_mlp.c -- host side of the QAOA energy surrogate: an MLP
 * in_dim -> hidden x nhidden (GELU) -> 1 whose forward and backward passes run
 * on the BF16 MFMA kernels of k_mlp.hip, trained with Adam on fp32 master
 * weights.  No reference counterpart (the learned screening stage).
 *
 * Layer l (0..nhidden-1) maps width K_l -> hidden: W_l is kept as fp32 [K_l][hidden]
 * (master), bf16 [K_l][hidden] (for dX = dY W^T) and bf16 [hidden][K_l] (forward).
 * The head is fp32: y = h . w_out + b_out.
 */
#include "ork_host.h"

#include <math.h>
#include <string.h>
#include <hip/hip_runtime_api.h>

int ork_mlp_gemm(void *stream, int epi, const void *A, int lda, const void *B, int ldb, int M, int N, int K,
                 float *c32, int ldc, const float *bias, float *pre, void *act);
int ork_mlp_gemm_splitk(void *stream, const void *A, int lda, const void *B, int ldb, int M, int N, int K, int slices,
                        float *work, float *c32, int ldc);
int ork_mlp_transpose(void *stream, const void *in, void *out, int R, int Cc);
int ork_mlp_to_bf16(void *stream, const float *in, void *out, long n);
int ork_mlp_head_forward(void *stream, const void *h, const float *w, float b, int M, int K, float *y,
                         const float *target, float *dy, float *sqerr, int valid, float scale);
int ork_mlp_head_backward(void *stream, const float *dy, const float *w, const float *pre, void *dpre, int M, int K);
int ork_mlp_colsum(void *stream, const void *a, const float *s, int M, int K, float *out, float *part);
int ork_mlp_sum(void *stream, const float *v, int n, float *out);
int ork_mlp_first_layer_graph(void *stream, const float *gf, int ng, const void *w, int N, float *u);
int ork_mlp_first_layer_rows(void *stream, const float *angles, int na, long r0, int rows, int Mp, const float *u,
                             const void *wa, const float *bias, int N, void *act);
int ork_mlp_adam(void *stream, float *w, const float *g, float *m, float *v, long n, int ncols, float lr, float b1,
                 float b2, float eps, float c1, float c2, void *wb, void *wtb, int nrows);

#define MLP_MAX_LAYERS 4
#define SPLITK_MAX 32   /* K slices of the weight-gradient products */
#define TILE 128

typedef struct {
    int K, N;
    float *w, *g, *m, *v;        /* fp32 [K][N] */
    float *b, *gb, *mb, *vb;     /* fp32 [N] */
    void *wb, *wtb;              /* bf16 [K][N] and [N][K] */
} layer_t;

struct orkan_mlp {
    int in_dim, hidden, nl, max_batch;
    layer_t L[MLP_MAX_LAYERS];
    float *w_out, *g_out, *m_out, *v_out;   /* [hidden] */
    float b_out, mb_out, vb_out;
    float *gb_out_d;                        /* device scalar: d loss / d b_out */
    /* activations (max_batch rows, padded to TILE) */
    float *x32;                 /* [M][in_dim] staged input */
    void *x16;                  /* bf16 [M][in_dim] */
    float *pre[MLP_MAX_LAYERS]; /* fp32 [M][hidden] */
    void *act[MLP_MAX_LAYERS];  /* bf16 [M][hidden] */
    void *dpre, *dpre2;         /* bf16 [M][hidden] ping-pong */
    void *tA, *tB;              /* bf16 transposes [hidden][M] */
    float *y, *dy, *sq, *tgt, *scal;
    float *colpart;             /* 64 x hidden partial column sums */
    float *splitk;              /* SPLITK_MAX x max(in_dim, hidden) x hidden split-K partial products */
    float *gf_dev, *u1;         /* screening: graph features [in_dim], their layer-1 share [hidden] */
    long step;
};

static void *dmalloc(size_t bytes, int *err) {
    void *p = NULL;
    if (*err) return NULL;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
        *err = 1;
        return NULL;
    }
    return p;
}

static int pad_rows(int m) { return (m + TILE - 1) / TILE * TILE; }

orkan_mlp_t *orkan_mlp_create(int in_dim, int hidden, int nhidden, int max_batch) {
    if (in_dim % 32 || in_dim <= 0 || hidden % TILE || hidden <= 0 || nhidden < 1 || nhidden > MLP_MAX_LAYERS ||
        max_batch <= 0)
        return NULL;
    orkan_mlp_t *q = calloc(1, sizeof(*q));
    if (!q) return NULL;
    ork_dev();
    int err = 0;
    q->in_dim = in_dim;
    q->hidden = hidden;
    q->nl = nhidden;
    q->max_batch = pad_rows(max_batch);
    const size_t M = (size_t)q->max_batch, H = (size_t)hidden;
    for (int l = 0; l < nhidden; ++l) {
        layer_t *L = &q->L[l];
        L->K = l == 0 ? in_dim : hidden;
        L->N = hidden;
        const size_t n = (size_t)L->K * L->N;
        L->w = dmalloc(n * 4, &err); L->g = dmalloc(n * 4, &err);
        L->m = dmalloc(n * 4, &err); L->v = dmalloc(n * 4, &err);
        L->b = dmalloc(H * 4, &err); L->gb = dmalloc(H * 4, &err);
        L->mb = dmalloc(H * 4, &err); L->vb = dmalloc(H * 4, &err);
        L->wb = dmalloc(n * 2, &err); L->wtb = dmalloc(n * 2, &err);
        q->pre[l] = dmalloc(M * H * 4, &err);
        q->act[l] = dmalloc(M * H * 2, &err);
    }
    q->w_out = dmalloc(H * 4, &err); q->g_out = dmalloc(H * 4, &err);
    q->m_out = dmalloc(H * 4, &err); q->v_out = dmalloc(H * 4, &err);
    q->gb_out_d = dmalloc(16, &err);
    q->x32 = dmalloc(M * in_dim * 4, &err);
    q->x16 = dmalloc(M * in_dim * 2, &err);
    q->dpre = dmalloc(M * H * 2, &err);
    q->dpre2 = dmalloc(M * H * 2, &err);
    q->tA = dmalloc(M * H * 2, &err);
    q->tB = dmalloc(M * H * 2, &err);
    q->y = dmalloc(M * 4, &err); q->dy = dmalloc(M * 4, &err);
    q->sq = dmalloc(M * 4, &err); q->tgt = dmalloc(M * 4, &err);
    q->scal = dmalloc(64, &err);
    q->colpart = dmalloc((size_t)64 * H * 4, &err);
    q->splitk = dmalloc((size_t)SPLITK_MAX * (size_t)(in_dim > H ? in_dim : H) * H * 4, &err);
    q->gf_dev = dmalloc((size_t)in_dim * 4, &err);
    q->u1 = dmalloc(H * 4, &err);
    if (err) {
        orkan_mlp_destroy(q);
        return NULL;
    }
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    for (int l = 0; l < nhidden; ++l) {
        const size_t n = (size_t)q->L[l].K * q->L[l].N;
        (void)hipMemsetAsync(q->L[l].m, 0, n * 4, st);
        (void)hipMemsetAsync(q->L[l].v, 0, n * 4, st);
        (void)hipMemsetAsync(q->L[l].mb, 0, H * 4, st);
        (void)hipMemsetAsync(q->L[l].vb, 0, H * 4, st);
    }
    (void)hipMemsetAsync(q->m_out, 0, H * 4, st);
    (void)hipMemsetAsync(q->v_out, 0, H * 4, st);
    (void)hipMemsetAsync(q->x32, 0, M * in_dim * 4, st);
    (void)hipStreamSynchronize(st);
    return q;
}

void orkan_mlp_destroy(orkan_mlp_t *q) {
    if (!q) return;
    (void)hipStreamSynchronize((hipStream_t)ork_dev()->stream);
    for (int l = 0; l < MLP_MAX_LAYERS; ++l) {
        layer_t *L = &q->L[l];
        void *p[] = {L->w, L->g, L->m, L->v, L->b, L->gb, L->mb, L->vb, L->wb, L->wtb, q->pre[l], q->act[l]};
        for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); ++i)
            if (p[i]) (void)hipFree(p[i]);
    }
    void *p[] = {q->w_out, q->g_out, q->m_out, q->v_out, q->gb_out_d, q->x32, q->x16, q->dpre, q->dpre2,
                 q->tA, q->tB, q->y, q->dy, q->sq, q->tgt, q->scal, q->colpart, q->splitk, q->gf_dev, q->u1};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); ++i)
        if (p[i]) (void)hipFree(p[i]);
    free(q);
}

static int refresh_bf16(orkan_mlp_t *q, int l) {
    /* an Adam step with lr = 0 only rewrites the bf16 copies */
    layer_t *L = &q->L[l];
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    const long n = (long)L->K * L->N;
    int e = (int)hipMemsetAsync(L->g, 0, (size_t)n * 4, st);
    if (e) return e;
    return ork_mlp_adam(st, L->w, L->g, L->m, L->v, n, L->N, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f, L->wb, L->wtb, L->K);
}

/* W[l]: fp32 [K_l][hidden] row-major; b[l]: [hidden]; w_out: [hidden] */
int orkan_mlp_set_weights(orkan_mlp_t *q, const float *const *W, const float *const *b, const float *w_out,
                          float b_out) {
    if (!q) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    int e = 0;
    for (int l = 0; l < q->nl && !e; ++l) {
        layer_t *L = &q->L[l];
        e |= (int)hipMemcpyAsync(L->w, W[l], (size_t)L->K * L->N * 4, hipMemcpyHostToDevice, st);
        e |= (int)hipMemcpyAsync(L->b, b[l], (size_t)L->N * 4, hipMemcpyHostToDevice, st);
        if (!e) e = refresh_bf16(q, l);
    }
    if (!e) e = (int)hipMemcpyAsync(q->w_out, w_out, (size_t)q->hidden * 4, hipMemcpyHostToDevice, st);
    q->b_out = b_out;
    if (!e) e = (int)hipStreamSynchronize(st);
    return e;
}

int orkan_mlp_get_weights(orkan_mlp_t *q, float *const *W, float *const *b, float *w_out, float *b_out) {
    if (!q) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    int e = 0;
    for (int l = 0; l < q->nl; ++l) {
        layer_t *L = &q->L[l];
        if (W && W[l]) e |= (int)hipMemcpyAsync(W[l], L->w, (size_t)L->K * L->N * 4, hipMemcpyDeviceToHost, st);
        if (b && b[l]) e |= (int)hipMemcpyAsync(b[l], L->b, (size_t)L->N * 4, hipMemcpyDeviceToHost, st);
    }
    if (w_out) e |= (int)hipMemcpyAsync(w_out, q->w_out, (size_t)q->hidden * 4, hipMemcpyDeviceToHost, st);
    if (b_out) *b_out = q->b_out;
    e |= (int)hipStreamSynchronize(st);
    return e;
}

/* forward over rows [0, Mp) of the staged bf16 input; keep_pre stores the pre-activations */
static int forward_rows(orkan_mlp_t *q, int Mp, int keep_pre) {
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    const void *in = q->x16;
    int ld = q->in_dim;
    for (int l = 0; l < q->nl; ++l) {
        layer_t *L = &q->L[l];
        int e = ork_mlp_gemm(st, 1, in, ld, L->wtb, L->K, Mp, L->N, L->K, NULL, L->N, L->b,
                             keep_pre ? q->pre[l] : NULL, q->act[l]);
        if (e) return e;
        in = q->act[l];
        ld = L->N;
    }
    return 0;
}

static int stage_input(orkan_mlp_t *q, const float *x, int rows, int Mp) {
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    int e = (int)hipMemcpyAsync(q->x32, x, (size_t)rows * q->in_dim * 4, hipMemcpyDefault, st);
    if (!e && Mp > rows)
        e = (int)hipMemsetAsync(q->x32 + (size_t)rows * q->in_dim, 0, (size_t)(Mp - rows) * q->in_dim * 4, st);
    if (!e) e = ork_mlp_to_bf16(st, q->x32, q->x16, (long)Mp * q->in_dim);
    return e;
}

/* y[i] for batch rows of x (host or device pointer, [batch][in_dim] fp32); any batch size */
int orkan_mlp_forward(orkan_mlp_t *q, int batch, const float *x, float *y) {
    if (!q || batch < 0) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    for (int r0 = 0; r0 < batch; r0 += q->max_batch) {
        const int rows = batch - r0 < q->max_batch ? batch - r0 : q->max_batch;
        const int Mp = pad_rows(rows);
        int e = stage_input(q, x + (size_t)r0 * q->in_dim, rows, Mp);
        if (!e) e = forward_rows(q, Mp, 0);
        if (!e) e = ork_mlp_head_forward(st, q->act[q->nl - 1], q->w_out, q->b_out, Mp, q->hidden, q->y, NULL, NULL,
                                         NULL, rows, 1.0f);
        if (!e) e = (int)hipMemcpyAsync(y + r0, q->y, (size_t)rows * 4, hipMemcpyDefault, st);
        if (e) return e;
    }
    return (int)hipStreamSynchronize(st);
}

/* Screening: rows r = 0..ncand-1 are [gf (ng floats), angles_dev[r] (in_dim - ng floats)].
 * The graph features' share of layer 1 is computed once (first_layer_graph); each
 * chunk then builds layer 1's activations from the angles alone and runs the
 * remaining layers.  y_dev receives ncand predictions. */
int orkan_mlp_forward_candidates(orkan_mlp_t *q, const float *gf_host, int ng, const float *angles_dev, long ncand,
                                 float *y_dev) {
    if (!q || ng <= 0 || ng >= q->in_dim || ncand < 0) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    const int na = q->in_dim - ng, H = q->hidden;
    const layer_t *L0 = &q->L[0];
    int e = (int)hipMemcpyAsync(q->gf_dev, gf_host, (size_t)ng * 4, hipMemcpyHostToDevice, st);
    if (!e) e = ork_mlp_first_layer_graph(st, q->gf_dev, ng, L0->wb, H, q->u1);
    for (long r0 = 0; r0 < ncand && !e; r0 += q->max_batch) {
        const int rows = (int)(ncand - r0 < q->max_batch ? ncand - r0 : q->max_batch);
        const int Mp = pad_rows(rows);
        e = ork_mlp_first_layer_rows(st, angles_dev, na, r0, rows, Mp, q->u1, (const char *)L0->wb + (size_t)ng * H * 2,
                                     L0->b, H, q->act[0]);
        const void *in = q->act[0];
        for (int l = 1; l < q->nl && !e; ++l) {
            layer_t *L = &q->L[l];
            e = ork_mlp_gemm(st, 1, in, L->K, L->wtb, L->K, Mp, L->N, L->K, NULL, L->N, L->b, NULL, q->act[l]);
            in = q->act[l];
        }
        if (!e) e = ork_mlp_head_forward(st, q->act[q->nl - 1], q->w_out, q->b_out, Mp, H, y_dev + r0, NULL, NULL,
                                         NULL, rows, 1.0f);
    }
    if (!e) e = (int)hipStreamSynchronize(st);
    return e;
}

/* loss = mean squared error; gradients of all parameters into the g buffers */
static int backward(orkan_mlp_t *q, int rows, const float *target, float *loss) {
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    const int Mp = pad_rows(rows), H = q->hidden;
    int e = (int)hipMemcpyAsync(q->tgt, target, (size_t)rows * 4, hipMemcpyDefault, st);
    if (!e) e = forward_rows(q, Mp, 1);
    if (!e) e = ork_mlp_head_forward(st, q->act[q->nl - 1], q->w_out, q->b_out, Mp, H, q->y, q->tgt, q->dy, q->sq,
                                     rows, 1.0f / (float)rows);
    if (!e) e = ork_mlp_sum(st, q->sq, Mp, q->scal);          /* sum of squared errors */
    if (!e) e = ork_mlp_sum(st, q->dy, Mp, q->gb_out_d);      /* d loss / d b_out */
    if (!e) e = ork_mlp_colsum(st, q->act[q->nl - 1], q->dy, Mp, H, q->g_out, q->colpart);
    if (!e) e = ork_mlp_head_backward(st, q->dy, q->w_out, q->pre[q->nl - 1], q->dpre, Mp, H);
    void *dcur = q->dpre, *dnext = q->dpre2;
    for (int l = q->nl - 1; l >= 0 && !e; --l) {
        layer_t *L = &q->L[l];
        const void *xin = l == 0 ? q->x16 : q->act[l - 1];
        /* dW_l = X^T dP: transposes to [K][M] and [N][M], then C = A B^T over the batch */
        e = ork_mlp_transpose(st, xin, q->tA, Mp, L->K);
        if (!e) e = ork_mlp_transpose(st, dcur, q->tB, Mp, L->N);
        if (!e) {
            /* K = the batch (thousands) but only (K_l / 128) x (N / 128) output tiles: split the batch
             * into up to SPLITK_MAX slices of >= 128 rows until ~512 workgroups run */
            const int tiles = (L->K / 128) * (L->N / 128);
            int slices = 1;
            while (slices < SPLITK_MAX && tiles * slices < 512 && Mp % (slices * 2 * 128) == 0) slices *= 2;
            e = slices > 1 ? ork_mlp_gemm_splitk(st, q->tA, Mp, q->tB, Mp, L->K, L->N, Mp, slices, q->splitk, L->g, L->N)
                           : ork_mlp_gemm(st, 0, q->tA, Mp, q->tB, Mp, L->K, L->N, Mp, L->g, L->N, NULL, NULL, NULL);
        }
        if (!e) e = ork_mlp_colsum(st, dcur, NULL, Mp, L->N, L->gb, q->colpart);
        if (!e && l > 0) {
            /* dP_{l-1} = (dP_l W_l^T) * gelu'(pre_{l-1}) */
            e = ork_mlp_gemm(st, 2, dcur, L->N, L->wb, L->N, Mp, L->K, L->N, NULL, L->K, NULL, q->pre[l - 1], dnext);
            void *t = dcur;
            dcur = dnext;
            dnext = t;
        }
    }
    if (!e && loss) {
        float s = 0.0f;
        e = (int)hipMemcpyAsync(&s, q->scal, 4, hipMemcpyDeviceToHost, st);
        if (!e) e = (int)hipStreamSynchronize(st);
        *loss = s / (float)rows;
    }
    return e;
}

int orkan_mlp_train_step(orkan_mlp_t *q, int batch, const float *x, const float *target, float lr, float *loss) {
    if (!q || batch <= 0 || batch > q->max_batch) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    const int Mp = pad_rows(batch);
    int e = stage_input(q, x, batch, Mp);
    if (!e) e = backward(q, batch, target, loss);
    if (e) return e;
    q->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)q->step), c2 = 1.0f - powf(b2, (float)q->step);
    for (int l = 0; l < q->nl && !e; ++l) {
        layer_t *L = &q->L[l];
        e = ork_mlp_adam(st, L->w, L->g, L->m, L->v, (long)L->K * L->N, L->N, lr, b1, b2, eps, c1, c2, L->wb, L->wtb,
                         L->K);
        if (!e) e = ork_mlp_adam(st, L->b, L->gb, L->mb, L->vb, L->N, L->N, lr, b1, b2, eps, c1, c2, NULL, NULL, 1);
    }
    if (!e) e = ork_mlp_adam(st, q->w_out, q->g_out, q->m_out, q->v_out, q->hidden, q->hidden, lr, b1, b2, eps, c1,
                             c2, NULL, NULL, 1);
    if (!e) {
        float gb = 0.0f;
        e = (int)hipMemcpyAsync(&gb, q->gb_out_d, 4, hipMemcpyDeviceToHost, st);
        if (!e) e = (int)hipStreamSynchronize(st);
        q->mb_out = b1 * q->mb_out + (1 - b1) * gb;
        q->vb_out = b2 * q->vb_out + (1 - b2) * gb * gb;
        q->b_out -= lr * (q->mb_out / c1) / (sqrtf(q->vb_out / c2) + eps);
    }
    return e;
}

/* gradients without an update (for checking against a CPU computation) */
int orkan_mlp_grads(orkan_mlp_t *q, int batch, const float *x, const float *target, float *const *dW,
                    float *const *db, float *dw_out, float *db_out, float *loss) {
    if (!q || batch <= 0 || batch > q->max_batch) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    int e = stage_input(q, x, batch, pad_rows(batch));
    if (!e) e = backward(q, batch, target, loss);
    for (int l = 0; l < q->nl && !e; ++l) {
        if (dW && dW[l]) e |= (int)hipMemcpyAsync(dW[l], q->L[l].g, (size_t)q->L[l].K * q->L[l].N * 4,
                                                   hipMemcpyDeviceToHost, st);
        if (db && db[l]) e |= (int)hipMemcpyAsync(db[l], q->L[l].gb, (size_t)q->L[l].N * 4, hipMemcpyDeviceToHost, st);
    }
    if (!e && dw_out) e = (int)hipMemcpyAsync(dw_out, q->g_out, (size_t)q->hidden * 4, hipMemcpyDeviceToHost, st);
    if (!e && db_out) e = (int)hipMemcpyAsync(db_out, q->gb_out_d, 4, hipMemcpyDeviceToHost, st);
    if (!e) e = (int)hipStreamSynchronize(st);
    return e;
}

unsigned long ork_topk_state_bytes(void);
int ork_topk(void *stream, const float *y, long n, int k, int *idx, float *vals, void *st_dev);

/* indices of the k largest entries of the device array y_dev (n floats), largest first, ties
 * by ascending index; values optional.  Selection on the device (radix select), the k
 * survivors ordered on the host. */
int orkan_topk_f32(const float *y_dev, long n, int k, int *idx_host, float *val_host) {
    if (!y_dev || !idx_host || k <= 0 || k > n || k > 65536) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    void *scratch = NULL;
    int e = (int)hipMalloc(&scratch, ork_topk_state_bytes() + (size_t)k * (sizeof(int) + sizeof(float)));
    float *vals = malloc((size_t)k * sizeof(float));
    int *idx = malloc((size_t)k * sizeof(int));
    int *ord = malloc((size_t)k * sizeof(int));
    if (!vals || !idx || !ord) e = (int)hipErrorOutOfMemory;
    int *idx_d = (int *)((char *)scratch + ork_topk_state_bytes());
    float *val_d = (float *)(idx_d + k);
    if (!e) e = ork_topk(st, y_dev, n, k, idx_d, val_d, scratch);
    if (!e) e = (int)hipMemcpyAsync(idx, idx_d, (size_t)k * sizeof(int), hipMemcpyDeviceToHost, st);
    if (!e) e = (int)hipMemcpyAsync(vals, val_d, (size_t)k * sizeof(float), hipMemcpyDeviceToHost, st);
    if (!e) e = (int)hipStreamSynchronize(st);
    if (!e) {
        /* insertion sort of k positions: value descending, then index ascending */
        for (int i = 0; i < k; ++i) ord[i] = i;
        for (int a = 1; a < k; ++a) {
            const int x = ord[a];
            int b = a - 1;
            while (b >= 0 && (vals[ord[b]] < vals[x] || (vals[ord[b]] == vals[x] && idx[ord[b]] > idx[x]))) {
                ord[b + 1] = ord[b];
                --b;
            }
            ord[b + 1] = x;
        }
        for (int i = 0; i < k; ++i) {
            idx_host[i] = idx[ord[i]];
            if (val_host) val_host[i] = vals[ord[i]];
        }
    }
    free(vals);
    free(idx);
    free(ord);
    if (scratch) (void)hipFree(scratch);
    return e;
}

/* C = A B^T for bf16 A [M][K], B [N][K] (device pointers), fp32 C [M][N]: exposed for testing the MFMA GEMM */
int orkan_mlp_gemm_test(int M, int N, int K, const void *A, const void *B, float *C) {
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    int e = ork_mlp_gemm(st, 0, A, K, B, K, M, N, K, C, N, NULL, NULL, NULL);
    if (!e) e = (int)hipStreamSynchronize(st);
    return e;
}
