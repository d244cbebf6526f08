/*
 * ork_runtime.c -- per-device runtime of the library: one HIP stream per
 * device (default priority), the reduction partials, and the staging buffers
 * that bring caller-owned host arrays (observables, diagonals) into HBM.
 */
#include "ork_host.h"

#include <pthread.h>
#include <string.h>
#include <hip/hip_runtime_api.h>

#define ORK_MAX_DEV 64

static ork_dev_t g_dev[ORK_MAX_DEV];
static int g_ready[ORK_MAX_DEV];
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;        /* ork_lock: shared staging, partials */
static pthread_mutex_t g_init_mu = PTHREAD_MUTEX_INITIALIZER;   /* per-device runtime creation, so that
                                                                 * ork_dev() may be called under ork_lock */

void ork_hip_fail(const char *where, int err) {
    fprintf(stderr, "orkan: hip: %s: %s\n", where, hipGetErrorString((hipError_t)err));
    exit(EXIT_FAILURE);
}

void ork_lock(void) { pthread_mutex_lock(&g_mu); }
void ork_unlock(void) { pthread_mutex_unlock(&g_mu); }

ork_dev_t *ork_dev(void) {
    int dev = 0;
    ORK_HIP("hipGetDevice", hipGetDevice(&dev));
    if (dev < 0 || dev >= ORK_MAX_DEV) ork_hip_fail("device index", (int)hipErrorInvalidDevice);
    if (!__atomic_load_n(&g_ready[dev], __ATOMIC_ACQUIRE)) {
        pthread_mutex_lock(&g_init_mu);
        if (!__atomic_load_n(&g_ready[dev], __ATOMIC_ACQUIRE)) {
            ork_dev_t *d = &g_dev[dev];
            memset(d, 0, sizeof(*d));
            hipStream_t st;
            ORK_HIP("hipStreamCreateWithFlags", hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            d->stream = (void *)st;
            ORK_HIP("hipMalloc(partials)", hipMalloc((void **)&d->partials, ork_partials_len() * sizeof(double)));
            /* device address of the allocation-free exp_diag's scratch array on this device; a failure
             * leaves it unset, and only orkan_exp_diag_noalloc reports it */
            d->noalloc_err = ork_noalloc_setup();
            __atomic_store_n(&g_ready[dev], 1, __ATOMIC_RELEASE);
        }
        pthread_mutex_unlock(&g_init_mu);
    }
    return &g_dev[dev];
}

ork_dev_t *ork_dev_existing(void) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= ORK_MAX_DEV) return NULL;
    return __atomic_load_n(&g_ready[dev], __ATOMIC_ACQUIRE) ? &g_dev[dev] : NULL;
}

int ork_is_dev_ptr(const void *p) {
    hipPointerAttribute_t at;
    memset(&at, 0, sizeof(at));
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return at.type == hipMemoryTypeDevice;
}

const double *ork_dev_doubles(const double *src, size_t count) {
    if (ork_is_dev_ptr(src)) return src;
    ork_dev_t *d = ork_dev();
    const size_t bytes = count * sizeof(double);
    if (d->stage_bytes < bytes) {
        if (d->stage) {
            ORK_HIP("hipStreamSynchronize", hipStreamSynchronize((hipStream_t)d->stream));
            ORK_HIP("hipFree(stage)", hipFree(d->stage));
        }
        d->stage = NULL;
        d->stage_bytes = 0;
        ORK_HIP("hipMalloc(stage)", hipMalloc(&d->stage, bytes));
        d->stage_bytes = bytes;
    }
    ORK_HIP("hipMemcpyAsync(stage)", hipMemcpyAsync(d->stage, src, bytes, hipMemcpyHostToDevice, (hipStream_t)d->stream));
    /* the caller may reuse its array as soon as the API call returns */
    ORK_HIP("hipStreamSynchronize", hipStreamSynchronize((hipStream_t)d->stream));
    return (const double *)d->stage;
}

/* ------------------------------------------------------------------------------------------------
 * extension API: stream access, device arrays, timing
 * ---------------------------------------------------------------------------------------------- */

void *orkan_stream(void) { return ork_dev()->stream; }

int orkan_sync(void) { return (int)hipStreamSynchronize((hipStream_t)ork_dev()->stream); }

double *orkan_dev_alloc(idx_t count) {
    void *p = NULL;
    if (hipMalloc(&p, (size_t)count * sizeof(double)) != hipSuccess) return NULL;
    return (double *)p;
}

int orkan_dev_upload(double *dst_dev, const double *src_host, idx_t count) {
    ork_dev_t *d = ork_dev();
    int e = (int)hipMemcpyAsync(dst_dev, src_host, (size_t)count * sizeof(double), hipMemcpyHostToDevice,
                                (hipStream_t)d->stream);
    if (e) return e;
    return (int)hipStreamSynchronize((hipStream_t)d->stream);
}

int orkan_set_launch_shape(int kernel, int flags, int lds_bytes) {
    if (kernel == 10) return ork_sv_set_red_lds(flags, lds_bytes);
    if (kernel == 11) return ork_mlp_gemm_select(flags);
    if (kernel == 12) return ork_sv_set_pair_shape(flags, lds_bytes);
    if (kernel == 13) return ork_sv_set_nt_off(flags, lds_bytes);
    if (kernel == 17) return ork_sv_set_layer_variant(flags);
    if (kernel == 19) return ork_sv_set_exp_shape(flags, lds_bytes);
    if (kernel == 20) return ork_sv_set_mono_lds32(lds_bytes);
    if (kernel == 21) return ork_sv_set_perm_shape(flags, lds_bytes);
    if (kernel == 14) {
        ork_dm_reset_shapes();
        ork_sv_reset_tuning();
        ork_mlp_gemm_reset();
        return 0;
    }
    return ork_dm_set_shape(kernel, flags, lds_bytes);
}

int orkan_copy_to_host(void *dst_host, const void *src_dev, size_t bytes) {
    ork_dev_t *d = ork_dev();
    int e = (int)hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, (hipStream_t)d->stream);
    if (e) return e;
    return (int)hipStreamSynchronize((hipStream_t)d->stream);
}

int orkan_copy_to_device(void *dst_dev, const void *src_host, size_t bytes) {
    ork_dev_t *d = ork_dev();
    int e = (int)hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, (hipStream_t)d->stream);
    if (e) return e;
    return (int)hipStreamSynchronize((hipStream_t)d->stream);
}

void orkan_dev_free(void *p) {
    if (!p) return;
    ork_dev();
    (void)hipDeviceSynchronize();
    (void)hipFree(p);
}

#define ORK_MAX_EVENTS 65536
static hipEvent_t *g_ev[ORK_MAX_DEV];

int orkan_event_record(int slot) {
    int dev = 0;
    if (slot < 0 || slot >= ORK_MAX_EVENTS) return (int)hipErrorInvalidValue;
    ork_dev_t *d = ork_dev();
    ORK_HIP("hipGetDevice", hipGetDevice(&dev));
    if (!g_ev[dev]) {
        g_ev[dev] = calloc(ORK_MAX_EVENTS, sizeof(hipEvent_t));
        if (!g_ev[dev]) return (int)hipErrorOutOfMemory;
    }
    if (!g_ev[dev][slot]) {
        int e = (int)hipEventCreate(&g_ev[dev][slot]);
        if (e) return e;
    }
    return (int)hipEventRecord(g_ev[dev][slot], (hipStream_t)d->stream);
}

double orkan_event_ms(int a, int b) {
    int dev = 0;
    float ms = -1.0f;
    if (a < 0 || b < 0 || a >= ORK_MAX_EVENTS || b >= ORK_MAX_EVENTS) return -1.0;
    ORK_HIP("hipGetDevice", hipGetDevice(&dev));
    if (!g_ev[dev] || !g_ev[dev][a] || !g_ev[dev][b]) return -1.0;
    if (hipEventSynchronize(g_ev[dev][b]) != hipSuccess) return -1.0;
    if (hipEventElapsedTime(&ms, g_ev[dev][a], g_ev[dev][b]) != hipSuccess) return -1.0;
    return (double)ms;
}
