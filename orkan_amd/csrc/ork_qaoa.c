/*
 * ork_qaoa.c -- host side of the batched QAOA launcher: device arenas for the
 * cost diagonals, the evaluation list and the per-workgroup state scratch; a
 * hipGraph holding the counter reset and the persistent kernel; and the
 * autotuner that picks the LDS tile and workgroup size.
 *
 * No reference counterpart: the reference evaluates one circuit per call
 * sequence (state_plus, exp_diag, rx, mean).  Each evaluation here computes
 * exactly that sequence.
 */
#include "ork_host.h"

#include <math.h>
#include <string.h>
#include <hip/hip_runtime_api.h>

int ork_qaoa_launch(void *stream, int lt, int threads, int grid, const void *evals, const int *order,
                    const int *nevals, const void *graphs, const double *angles, const double *diag, void *states,
                    uint64_t slot_elems, double *out, unsigned *counter);
int ork_qaoa_occupancy(int lt, int threads, int *blocks_per_cu);
int ork_maxcut_diag(void *stream, double *out, int n, int ne, const int *ei, const int *ej, const double *w);

typedef struct { int graph, p, angle_off, pad; } q_eval_t;
typedef struct { uint64_t diag_off, state_off; int n, pad; } q_graph_t;

struct orkan_qaoa {
    int ngraphs, max_evals, max_p, nmax;
    q_graph_t *graphs_h;
    void *graphs_d;
    double *diag_d;
    void *evals_d;
    int *order_d, *nevals_d;
    double *angles_d, *out_d;
    unsigned *counter_d;
    void *states_d;
    size_t states_bytes;
    /* launch configuration */
    int lt, threads, grid;
    /* the graph is captured on this private stream, never on the library stream that other
     * threads' gates and reductions use, and replayed on the library stream */
    hipStream_t capture_stream;
    int nevals_h;          /* host source of the nevals_d upload */
    hipGraph_t graph;
    hipGraphExec_t exec;
    int graph_ok;
    /* host staging */
    q_eval_t *evals_h;
    int *order_h;
    double *angles_h;
};

static int cus(void) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
    return n;
}

static int ensure_states(orkan_qaoa_t *q) {
    /* scratch for states larger than the LDS tile: one slot of 2^nmax per workgroup */
    size_t need = 16;
    if (q->nmax > q->lt) need = (size_t)q->grid * ((size_t)1 << q->nmax) * 16;
    if (q->states_bytes >= need) return 0;
    if (q->states_d) {
        (void)hipStreamSynchronize((hipStream_t)ork_dev()->stream);
        (void)hipFree(q->states_d);
    }
    q->states_d = NULL;
    q->states_bytes = 0;
    int e = (int)hipMalloc(&q->states_d, need);
    if (e) return e;
    q->states_bytes = need;
    q->graph_ok = 0;
    return 0;
}

static void drop_graph(orkan_qaoa_t *q) {
    if (q->graph_ok) {
        (void)hipGraphExecDestroy(q->exec);
        (void)hipGraphDestroy(q->graph);
    }
    q->graph_ok = 0;
}

/* caller holds ork_lock (the state scratch may be freed and the graph dropped) */
static int set_config(orkan_qaoa_t *q, int lt, int threads, int blocks_per_cu) {
    if (2 * lt < q->nmax) return (int)hipErrorInvalidValue;     /* pass B needs n <= 2 LT */
    int occ = 0;
    int e = ork_qaoa_occupancy(lt, threads, &occ);
    if (e) return e;
    if (occ <= 0) return (int)hipErrorInvalidConfiguration;
    if (blocks_per_cu <= 0 || blocks_per_cu > occ) blocks_per_cu = occ;
    drop_graph(q);
    q->lt = lt;
    q->threads = threads;
    q->grid = cus() * blocks_per_cu;
    return ensure_states(q);
}

int orkan_qaoa_set_config(orkan_qaoa_t *q, int lt, int threads, int blocks_per_cu) {
    if (!q) return (int)hipErrorInvalidValue;
    ork_lock();
    const int e = set_config(q, lt, threads, blocks_per_cu);
    ork_unlock();
    return e;
}

orkan_qaoa_t *orkan_qaoa_create(int ngraphs, const unsigned char *nq, const double *const *diags, int max_evals,
                                int max_p) {
    if (ngraphs <= 0 || max_evals <= 0 || max_p <= 0 || !nq || !diags) return NULL;
    orkan_qaoa_t *q = calloc(1, sizeof(*q));
    if (!q) return NULL;
    ork_dev_t *d = ork_dev();
    hipStream_t st = (hipStream_t)d->stream;
    q->ngraphs = ngraphs;
    q->max_evals = max_evals;
    q->max_p = max_p;
    q->graphs_h = calloc((size_t)ngraphs, sizeof(q_graph_t));
    q->evals_h = calloc((size_t)max_evals, sizeof(q_eval_t));
    q->order_h = calloc((size_t)max_evals, sizeof(int));
    q->angles_h = calloc((size_t)max_evals * max_p * 3, sizeof(double));
    if (!q->graphs_h || !q->evals_h || !q->order_h || !q->angles_h) goto fail;
    if (hipStreamCreateWithFlags(&q->capture_stream, hipStreamNonBlocking) != hipSuccess) {
        q->capture_stream = NULL;
        goto fail;
    }
    uint64_t total = 0;
    for (int g = 0; g < ngraphs; ++g) {
        if (nq[g] < 1 || nq[g] > 26) goto fail;
        q->graphs_h[g].n = nq[g];
        q->graphs_h[g].diag_off = total;
        total += (uint64_t)1 << nq[g];
        if (nq[g] > q->nmax) q->nmax = nq[g];
    }
    if (hipMalloc((void **)&q->diag_d, total * sizeof(double)) != hipSuccess) goto fail;
    for (int g = 0; g < ngraphs; ++g)
        if (hipMemcpyAsync(q->diag_d + q->graphs_h[g].diag_off, diags[g], ((size_t)1 << nq[g]) * sizeof(double),
                           hipMemcpyDefault, st) != hipSuccess)
            goto fail;
    if (hipMalloc(&q->graphs_d, (size_t)ngraphs * sizeof(q_graph_t)) != hipSuccess) goto fail;
    if (hipMemcpyAsync(q->graphs_d, q->graphs_h, (size_t)ngraphs * sizeof(q_graph_t), hipMemcpyHostToDevice, st) !=
        hipSuccess)
        goto fail;
    if (hipMalloc(&q->evals_d, (size_t)max_evals * sizeof(q_eval_t)) != hipSuccess) goto fail;
    if (hipMalloc((void **)&q->order_d, (size_t)max_evals * sizeof(int)) != hipSuccess) goto fail;
    if (hipMalloc((void **)&q->nevals_d, sizeof(int)) != hipSuccess) goto fail;
    if (hipMalloc((void **)&q->angles_d, (size_t)max_evals * max_p * 3 * sizeof(double)) != hipSuccess) goto fail;
    if (hipMalloc((void **)&q->out_d, (size_t)max_evals * sizeof(double)) != hipSuccess) goto fail;
    if (hipMalloc((void **)&q->counter_d, sizeof(unsigned)) != hipSuccess) goto fail;
    if (hipStreamSynchronize(st) != hipSuccess) goto fail;
    {
        /* default tile: 4096 amplitudes (64 KB of LDS), 512 threads; 8192 when the
         * largest state needs it */
        int lt = 12;
        while (2 * lt < q->nmax) ++lt;
        ork_lock();
        const int e = set_config(q, lt, lt >= 13 ? 1024 : 512, 0);
        ork_unlock();
        if (e) goto fail;
    }
    return q;
fail:
    orkan_qaoa_destroy(q);
    return NULL;
}

void orkan_qaoa_destroy(orkan_qaoa_t *q) {
    if (!q) return;
    (void)hipStreamSynchronize((hipStream_t)ork_dev()->stream);
    drop_graph(q);
    void *bufs[] = {q->graphs_d, q->diag_d, q->evals_d, q->order_d, q->nevals_d,
                    q->angles_d, q->out_d, q->counter_d, q->states_d};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); ++i)
        if (bufs[i]) (void)hipFree(bufs[i]);
    free(q->graphs_h);
    free(q->evals_h);
    free(q->order_h);
    free(q->angles_h);
    if (q->capture_stream) (void)hipStreamDestroy(q->capture_stream);
    free(q);
}

static int capture(orkan_qaoa_t *q) {
    hipStream_t st = q->capture_stream;
    drop_graph(q);
    int e = (int)hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e) return e;
    int le = ork_qaoa_launch(st, q->lt, q->threads, q->grid, q->evals_d, q->order_d, q->nevals_d, q->graphs_d,
                             q->angles_d, q->diag_d, q->states_d, ((uint64_t)1 << q->nmax), q->out_d, q->counter_d);
    e = (int)hipStreamEndCapture(st, &q->graph);
    if (le) return le;
    if (e) return e;
    e = (int)hipGraphInstantiate(&q->exec, q->graph, NULL, NULL, 0);
    if (e) {
        (void)hipGraphDestroy(q->graph);
        return e;
    }
    q->graph_ok = 1;
    return 0;
}

static uint64_t cost(const orkan_qaoa_t *q, const q_eval_t *ev) {
    return ((uint64_t)1 << q->graphs_h[ev->graph].n) * (uint64_t)(ev->p > 0 ? ev->p : 1);
}

static const orkan_qaoa_t *g_sort_q;
static int by_cost_desc(const void *a, const void *b) {
    const q_eval_t *E = g_sort_q->evals_h;
    const uint64_t ca = cost(g_sort_q, &E[*(const int *)a]), cb = cost(g_sort_q, &E[*(const int *)b]);
    if (ca != cb) return ca < cb ? 1 : -1;
    return *(const int *)a - *(const int *)b;
}

/* upload the evaluation list; angles: rows of max_p gammas and betas */
static int stage(orkan_qaoa_t *q, int nevals, const int *graph, const int *p, const double *gamma,
                 const double *beta) {
    if (nevals < 0 || nevals > q->max_evals) return (int)hipErrorInvalidValue;
    for (int e = 0; e < nevals; ++e) {
        if (graph[e] < 0 || graph[e] >= q->ngraphs || p[e] < 1 || p[e] > q->max_p) return (int)hipErrorInvalidValue;
        q->evals_h[e].graph = graph[e];
        q->evals_h[e].p = p[e];
        q->evals_h[e].angle_off = e * q->max_p * 3;
        for (int l = 0; l < p[e]; ++l) {
            /* rx(q, 2 beta): the reference evaluates cos(theta / 2) and sin(theta / 2) */
            const double theta = 2.0 * beta[(size_t)e * q->max_p + l];
            q->angles_h[(size_t)e * q->max_p * 3 + 3 * l] = gamma[(size_t)e * q->max_p + l];
            q->angles_h[(size_t)e * q->max_p * 3 + 3 * l + 1] = cos(theta / 2.0);
            q->angles_h[(size_t)e * q->max_p * 3 + 3 * l + 2] = sin(theta / 2.0);
        }
        q->order_h[e] = e;
    }
    g_sort_q = q;
    qsort(q->order_h, (size_t)nevals, sizeof(int), by_cost_desc);
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    int e = 0;
    if (nevals > 0) {
        e |= (int)hipMemcpyAsync(q->evals_d, q->evals_h, (size_t)nevals * sizeof(q_eval_t), hipMemcpyHostToDevice, st);
        e |= (int)hipMemcpyAsync(q->order_d, q->order_h, (size_t)nevals * sizeof(int), hipMemcpyHostToDevice, st);
        e |= (int)hipMemcpyAsync(q->angles_d, q->angles_h, (size_t)nevals * q->max_p * 3 * sizeof(double),
                                 hipMemcpyHostToDevice, st);
    }
    q->nevals_h = nevals;
    e |= (int)hipMemcpyAsync(q->nevals_d, &q->nevals_h, sizeof(int), hipMemcpyHostToDevice, st);
    e |= (int)hipStreamSynchronize(st);
    return e;
}

/* caller holds ork_lock */
int orkan_qaoa_launch_staged(orkan_qaoa_t *q) {
    if (!q->graph_ok) {
        int e = capture(q);
        if (e) return e;
    }
    return (int)hipGraphLaunch(q->exec, (hipStream_t)ork_dev()->stream);
}

int orkan_qaoa_eval(orkan_qaoa_t *q, int nevals, const int *graph, const int *p, const double *gamma,
                    const double *beta, double *out) {
    if (!q) return (int)hipErrorInvalidValue;
    ork_lock();
    int e = stage(q, nevals, graph, p, gamma, beta);
    if (!e) e = orkan_qaoa_launch_staged(q);
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    if (!e && nevals > 0 && out)
        e = (int)hipMemcpyAsync(out, q->out_d, (size_t)nevals * sizeof(double), hipMemcpyDeviceToHost, st);
    if (!e) e = (int)hipStreamSynchronize(st);
    ork_unlock();
    return e;
}

int orkan_qaoa_get_config(const orkan_qaoa_t *q, int *lt, int *threads, int *grid) {
    if (!q) return (int)hipErrorInvalidValue;
    if (lt) *lt = q->lt;
    if (threads) *threads = q->threads;
    if (grid) *grid = q->grid;
    return 0;
}

/* Autotuner: time every (tile, workgroup size) candidate the largest state
 * allows on the caller's evaluation list (one untimed launch, then the median
 * of `reps` launches), keep the fastest. */
int orkan_qaoa_autotune(orkan_qaoa_t *q, int nevals, const int *graph, const int *p, const double *gamma,
                        const double *beta, int reps, double *best_ms) {
    static const int cand[][2] = {{11, 256}, {11, 512}, {12, 256}, {12, 512}, {12, 1024}, {13, 512}, {13, 1024}};
    if (!q) return (int)hipErrorInvalidValue;
    if (reps < 1) reps = 1;
    if (reps > 15) reps = 15;
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return (int)hipErrorOutOfMemory;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return (int)hipErrorOutOfMemory;
    }
    /* the whole search holds the library lock: staging, every reconfiguration and launch */
    ork_lock();
    int e = stage(q, nevals, graph, p, gamma, beta);
    if (e) {
        ork_unlock();
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return e;
    }
    double best = 1e30;
    int best_i = -1;
    for (size_t i = 0; i < sizeof(cand) / sizeof(cand[0]); ++i) {
        if (set_config(q, cand[i][0], cand[i][1], 0)) continue;
        if (orkan_qaoa_launch_staged(q)) continue;
        float t[15];
        int ok = 1;
        for (int r = 0; r < reps && ok; ++r) {
            ok &= hipEventRecord(a, st) == hipSuccess;
            ok &= orkan_qaoa_launch_staged(q) == 0;
            ok &= hipEventRecord(b, st) == hipSuccess;
            ok &= hipEventSynchronize(b) == hipSuccess;
            ok &= hipEventElapsedTime(&t[r], a, b) == hipSuccess;
        }
        if (!ok) continue;
        for (int x = 0; x < reps; ++x)
            for (int y = x + 1; y < reps; ++y)
                if (t[y] < t[x]) { float tmp = t[x]; t[x] = t[y]; t[y] = tmp; }
        const double med = t[reps / 2];
        if (med < best) {
            best = med;
            best_i = (int)i;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (best_i < 0) {
        ork_unlock();
        return (int)hipErrorInvalidConfiguration;
    }
    if (best_ms) *best_ms = best;
    e = set_config(q, cand[best_i][0], cand[best_i][1], 0);
    ork_unlock();
    return e;
}

int orkan_maxcut_diag_dev(int n, int nedges, const int *ei, const int *ej, const double *w, double *out_dev) {
    ork_dev_t *d = ork_dev();
    hipStream_t st = (hipStream_t)d->stream;
    int *ed = NULL;
    double *wd = NULL;
    int e = (int)hipMalloc((void **)&ed, (size_t)(2 * nedges + 1) * sizeof(int));
    if (!e) e = (int)hipMalloc((void **)&wd, (size_t)(nedges + 1) * sizeof(double));
    if (!e && nedges > 0) {
        e |= (int)hipMemcpyAsync(ed, ei, (size_t)nedges * sizeof(int), hipMemcpyHostToDevice, st);
        e |= (int)hipMemcpyAsync(ed + nedges, ej, (size_t)nedges * sizeof(int), hipMemcpyHostToDevice, st);
        e |= (int)hipMemcpyAsync(wd, w, (size_t)nedges * sizeof(double), hipMemcpyHostToDevice, st);
    }
    if (!e) e = ork_maxcut_diag(st, out_dev, n, nedges, ed, ed + nedges, wd);
    if (!e) e = (int)hipStreamSynchronize(st);
    if (ed) (void)hipFree(ed);
    if (wd) (void)hipFree(wd);
    return e;
}
