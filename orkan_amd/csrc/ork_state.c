/*
 * ork_state.c -- state_t lifecycle and element access with the storage in HBM.
 *
 * Same entry points and semantics as the reference (src/state/state.c:43-205):
 * zero-initialised storage, ownership transfer in state_init, idempotent
 * state_free, |+> initialisation (PURE: 1/sqrt(N) everywhere; mixed: every
 * entry 1/D, tiled n < LOG_TILE_DIM fills only the D x D corner of the padded
 * tile), deep copy, and Hermitian get/set (upper triangle through conj, tiled
 * set mirrors inside diagonal tiles).  `data` is a device pointer.
 */
#include "ork_host.h"

#include <assert.h>
#include <math.h>
#include <string.h>
#include <hip/hip_runtime_api.h>

idx_t ork_len(int layout, qubit_t n) {
    const idx_t dim = (idx_t)1 << n;
    switch (layout) {
    case PURE:
        return dim;
    case MIXED_PACKED:
        return n == 0 ? 1 : (dim >> 1) * (dim + 1);
    case MIXED_TILED: {
        const idx_t nt = (dim + TILE_DIM - 1) >> LOG_TILE_DIM;
        return (nt * (nt + 1) >> 1) * (idx_t)TILE_SIZE;
    }
    default:
        return 0;
    }
}

idx_t state_len(const state_t *state) {
    assert(state != NULL);
    assert(ork_type_ok(state) && "state_len(): invalid state type");
    return ork_len(ork_layout(state), state->qubits);
}

static void *dev_zeros(size_t bytes) {
    void *p = NULL;
    if (bytes == 0) bytes = 16;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return NULL;
    }
    ork_dev_t *d = ork_dev();
    if (hipMemsetAsync(p, 0, bytes, (hipStream_t)d->stream) != hipSuccess) {
        (void)hipFree(p);
        return NULL;
    }
    return p;
}

static int is_dev(const void *p) {
    hipPointerAttribute_t at;
    memset(&at, 0, sizeof(at));
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return at.type == hipMemoryTypeDevice;
}

static void dev_release(void *p) {
    if (!p) return;
    (void)hipStreamSynchronize((hipStream_t)ork_dev()->stream);
    (void)hipFree(p);
}

void state_init(state_t *state, qubit_t qubits, cplx_t **data) {
    assert(state != NULL);
    ork_dev();
    if (state->data) {
        dev_release(state->data);
        state->data = NULL;
    }
    state->qubits = qubits;
    const idx_t len = state_len(state);
    const size_t bytes = (size_t)len * ork_elem_bytes(state);
    if (data && *data) {
        if (is_dev(*data)) {
            /* a device buffer becomes the state's storage as it is */
            state->data = *data;
        } else {
            /* a host buffer: copied into HBM, then released like the reference's
             * re-aligning copy (src/state/state.c:72-87) */
            state->data = dev_zeros(bytes);
            if (!state->data) {
                fprintf(stderr, "state_init(): aligned realloc failed for %zu elements\n", (size_t)len);
                state->qubits = 0;
                free(*data);
                *data = NULL;
                return;
            }
            /* on the library stream, after the zero fill that dev_zeros enqueued there */
            hipStream_t st = (hipStream_t)ork_dev()->stream;
            ORK_HIP("state_init upload", hipMemcpyAsync(state->data, *data, bytes, hipMemcpyHostToDevice, st));
            ORK_HIP("state_init upload", hipStreamSynchronize(st));
            free(*data);
        }
        *data = NULL;
        return;
    }
    state->data = dev_zeros(bytes);
    if (!state->data) {
        fprintf(stderr, "state_init(): allocation failed for %zu elements\n", (size_t)len);
        state->qubits = 0;
    }
}

void state_free(state_t *state) {
    if (!state) return;
    dev_release(state->data);
    state->data = NULL;
    state->qubits = 0;
}

void state_plus(state_t *state, qubit_t qubits) {
    assert(state != NULL);
    assert(ork_type_ok(state) && "state_plus(): invalid state type");
    state_init(state, qubits, NULL);
    if (!state->data) return;
    ork_dev_t *d = ork_dev();
    const int layout = ork_layout(state), prec = ork_prec(state);
    const idx_t dim = (idx_t)1 << qubits;
    if (layout == PURE) {
        ORK_HIP("state_plus", ork_fill(d->stream, prec, state->data, dim, 1.0 / sqrt((double)dim), 0.0));
    } else if (layout == MIXED_TILED && qubits < LOG_TILE_DIM) {
        ORK_HIP("state_plus", ork_fill_tiled_corner(d->stream, prec, state->data, qubits, LOG_TILE_DIM, 1.0 / (double)dim));
    } else {
        ORK_HIP("state_plus", ork_fill(d->stream, prec, state->data, state_len(state), 1.0 / (double)dim, 0.0));
    }
}

state_t state_cp(const state_t *state) {
    assert(state != NULL && state->data != NULL);
    state_t out = *state;
    const size_t bytes = (size_t)state_len(state) * ork_elem_bytes(state);
    out.data = dev_zeros(bytes);
    if (!out.data) {
        fprintf(stderr, "state_cp(): allocation failed\n");
        out.qubits = 0;
        return out;
    }
    ORK_HIP("state_cp", hipMemcpyAsync(out.data, state->data, bytes, hipMemcpyDeviceToDevice,
                                       (hipStream_t)ork_dev()->stream));
    return out;
}

/* ---- single-element access ---------------------------------------------------------------- */

static idx_t tiled_index(idx_t r, idx_t c) {
    const idx_t tr = r >> LOG_TILE_DIM, tc = c >> LOG_TILE_DIM;
    return (tr * (tr + 1) / 2 + tc) * (idx_t)TILE_SIZE + (r & (TILE_DIM - 1)) * TILE_DIM + (c & (TILE_DIM - 1));
}

static idx_t packed_index(idx_t dim, idx_t r, idx_t c) { return c * (2 * dim - c + 1) / 2 + (r - c); }

/* one element through the library stream: ordered after every enqueued kernel */
static cplx_t elem_read(const state_t *s, idx_t at) {
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    if (ork_prec(s) == ORK_F32) {
        float v[2];
        ORK_HIP("state_get", hipMemcpyAsync(v, (char *)s->data + at * 8, 8, hipMemcpyDeviceToHost, st));
        ORK_HIP("state_get", hipStreamSynchronize(st));
        return CMPLX((double)v[0], (double)v[1]);
    }
    double v[2];
    ORK_HIP("state_get", hipMemcpyAsync(v, (char *)s->data + at * 16, 16, hipMemcpyDeviceToHost, st));
    ORK_HIP("state_get", hipStreamSynchronize(st));
    return CMPLX(v[0], v[1]);
}

static void elem_write(state_t *s, idx_t at, cplx_t val) {
    hipStream_t st = (hipStream_t)ork_dev()->stream;
    if (ork_prec(s) == ORK_F32) {
        float v[2] = {(float)creal(val), (float)cimag(val)};
        ORK_HIP("state_set", hipMemcpyAsync((char *)s->data + at * 8, v, 8, hipMemcpyHostToDevice, st));
        ORK_HIP("state_set", hipStreamSynchronize(st));
        return;
    }
    double v[2] = {creal(val), cimag(val)};
    ORK_HIP("state_set", hipMemcpyAsync((char *)s->data + at * 16, v, 16, hipMemcpyHostToDevice, st));
    ORK_HIP("state_set", hipStreamSynchronize(st));
}

cplx_t state_get(const state_t *state, idx_t row, idx_t col) {
    assert(state != NULL && state->data != NULL);
    const idx_t dim = (idx_t)1 << state->qubits;
    switch (ork_layout(state)) {
    case PURE:
        assert(col == 0 && "col must be 0 for PURE states");
        assert(row < dim);
        return elem_read(state, row);
    case MIXED_PACKED:
        assert(row < dim && col < dim);
        return row >= col ? elem_read(state, packed_index(dim, row, col))
                          : conj(elem_read(state, packed_index(dim, col, row)));
    case MIXED_TILED:
        assert(row < dim && col < dim);
        return row >= col ? elem_read(state, tiled_index(row, col)) : conj(elem_read(state, tiled_index(col, row)));
    default:
        assert(0 && "state_get(): invalid state type");
        return 0.0;
    }
}

void state_set(state_t *state, idx_t row, idx_t col, cplx_t val) {
    assert(state != NULL && state->data != NULL);
    const idx_t dim = (idx_t)1 << state->qubits;
    switch (ork_layout(state)) {
    case PURE:
        assert(col == 0 && "col must be 0 for PURE states");
        assert(row < dim);
        elem_write(state, row, val);
        break;
    case MIXED_PACKED:
        assert(row < dim && col < dim);
        if (row >= col) elem_write(state, packed_index(dim, row, col), val);
        else elem_write(state, packed_index(dim, col, row), conj(val));
        break;
    case MIXED_TILED: {
        assert(row < dim && col < dim);
        const int same = (row >> LOG_TILE_DIM) == (col >> LOG_TILE_DIM);
        if (row >= col) {
            elem_write(state, tiled_index(row, col), val);
            if (same && row != col) elem_write(state, tiled_index(col, row), conj(val));
        } else {
            elem_write(state, tiled_index(col, row), conj(val));
            if (same) elem_write(state, tiled_index(row, col), val);
        }
        break;
    }
    default:
        assert(0 && "state_set(): invalid state type");
    }
}

/* ---- printing (format of src/internal/utils.h:44-143) --------------------------------------- */

static void print_cplx(cplx_t x) {
    printf("%.7f", creal(x));
    if (cimag(x) >= 0) printf("+i%.7f", cimag(x));
    else printf("-i%.7f", fabs(cimag(x)));
}

void state_print(const state_t *state) {
    assert(state != NULL);
    static const char *names[] = {"PURE", "MIXED_PACKED", "MIXED_TILED"};
    const int layout = ork_layout(state);
    printf("Type: %s\nQubits: %u\n", names[layout], state->qubits);
    const idx_t dim = (idx_t)1 << state->qubits;
    const idx_t len = state_len(state);
    cplx_t *h = malloc((size_t)len * sizeof(cplx_t));
    if (!h) return;
    if (ork_prec(state) == ORK_F32) {
        float *f = malloc((size_t)len * 8);
        if (!f) { free(h); return; }
        orkan_state_download(state, f);
        for (idx_t i = 0; i < len; ++i) h[i] = CMPLX((double)f[2 * i], (double)f[2 * i + 1]);
        free(f);
    } else {
        orkan_state_download(state, h);
    }
    if (layout == PURE) {
        printf("[");
        for (idx_t i = 0; i < dim; ++i) {
            print_cplx(h[i]);
            if (i + 1 < dim) printf(", ");
        }
        printf("]\n");
        free(h);
        return;
    }
    printf("[");
    for (idx_t r = 0; r < dim; ++r) {
        if (r) printf(" ");
        printf("[");
        for (idx_t c = 0; c < dim; ++c) {
            const idx_t lo_r = r >= c ? r : c, lo_c = r >= c ? c : r;
            const idx_t at = layout == MIXED_PACKED ? packed_index(dim, lo_r, lo_c) : tiled_index(lo_r, lo_c);
            print_cplx(r >= c ? h[at] : conj(h[at]));
            if (c + 1 < dim) printf(", ");
        }
        printf("]");
        if (r + 1 < dim) printf("\n");
    }
    printf("]\n");
    free(h);
}

/* ---- bulk transfer and precision conversion ------------------------------------------------- */

int orkan_state_upload(state_t *state, const void *host) {
    if (!state || !state->data || !host) return (int)hipErrorInvalidValue;
    ork_dev_t *d = ork_dev();
    const size_t bytes = (size_t)state_len(state) * ork_elem_bytes(state);
    int e = (int)hipMemcpyAsync(state->data, host, bytes, hipMemcpyHostToDevice, (hipStream_t)d->stream);
    if (e) return e;
    return (int)hipStreamSynchronize((hipStream_t)d->stream);
}

int orkan_state_download(const state_t *state, void *host) {
    if (!state || !state->data || !host) return (int)hipErrorInvalidValue;
    ork_dev_t *d = ork_dev();
    const size_t bytes = (size_t)state_len(state) * ork_elem_bytes(state);
    int e = (int)hipMemcpyAsync(host, state->data, bytes, hipMemcpyDeviceToHost, (hipStream_t)d->stream);
    if (e) return e;
    return (int)hipStreamSynchronize((hipStream_t)d->stream);
}

state_t orkan_state_convert(const state_t *state, int to_fp32) {
    assert(state != NULL && state->data != NULL);
    state_t out = *state;
    out.type = (state_type_t)(((int)state->type & ~ORKAN_FP32) | (to_fp32 ? ORKAN_FP32 : 0));
    const idx_t len = state_len(state);
    out.data = dev_zeros((size_t)len * ork_elem_bytes(&out));
    if (!out.data) {
        out.qubits = 0;
        return out;
    }
    ORK_HIP("orkan_state_convert",
            ork_convert(ork_dev()->stream, state->data, ork_prec(state), out.data, ork_prec(&out), len));
    return out;
}

/* ------------------------------------------------------------------------------------------------
 * Extensions without a reference counterpart: layout conversion on the device and a state file
 * format that uses the reference's reserved QS_ERR_FILE / QS_ERR_FORMAT codes
 * (include/q_types.h:35-36).
 * ---------------------------------------------------------------------------------------------- */

qs_error_t orkan_state_relayout(const state_t *src, state_t *dst, state_type_t layout) {
    if (!src || !src->data || !dst) return QS_ERR_NULL;
    if (!ork_type_ok(src)) return QS_ERR_TYPE;
    if (layout != MIXED_PACKED && layout != MIXED_TILED) return QS_ERR_TYPE;
    state_t out;
    out.type = (state_type_t)((int)layout | ((int)src->type & ORKAN_FP32));
    out.qubits = src->qubits;
    const idx_t len = ork_len((int)layout, src->qubits);
    out.data = dev_zeros((size_t)len * ork_elem_bytes(&out));
    if (!out.data) return QS_ERR_OOM;
    ork_dev_t *d = ork_dev();
    ORK_HIP("orkan_state_relayout", ork_relayout(d->stream, ork_prec(src), src->data, ork_layout(src), out.data,
                                                 (int)layout, src->qubits, LOG_TILE_DIM, len));
    ORK_HIP("orkan_state_relayout", hipStreamSynchronize((hipStream_t)d->stream));
    *dst = out;
    return QS_OK;
}

/* File: a 64-byte little-endian header, then the storage array as the state holds it. */
#define ORK_FILE_MAGIC "ORKAN-ST"
#define ORK_FILE_VERSION 1u
#define ORK_FILE_CHUNK ((size_t)64 << 20)

typedef struct {
    char magic[8];
    uint32_t version;
    uint32_t type;        /* state->type, ORKAN_FP32 flag included */
    uint32_t qubits;
    uint32_t log_tile;    /* LOG_TILE_DIM of the writer (tiled layout only must match) */
    uint64_t length;      /* stored elements */
    uint32_t elem_bytes;  /* 16 (complex128) or 8 (complex64) */
    uint32_t byte_order;  /* 0x01020304 as written by the host */
    uint8_t reserved[24];
} ork_file_header_t;

_Static_assert(sizeof(ork_file_header_t) == 64, "state file header is 64 bytes");

qs_error_t orkan_state_save(const state_t *state, const char *path) {
    if (!state || !state->data || !path) return QS_ERR_NULL;
    if (!ork_type_ok(state)) return QS_ERR_TYPE;
    ork_file_header_t h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, ORK_FILE_MAGIC, 8);
    h.version = ORK_FILE_VERSION;
    h.type = (uint32_t)state->type;
    h.qubits = state->qubits;
    h.log_tile = LOG_TILE_DIM;
    h.length = state_len(state);
    h.elem_bytes = (uint32_t)ork_elem_bytes(state);
    h.byte_order = 0x01020304u;
    FILE *f = fopen(path, "wb");
    if (!f) return QS_ERR_FILE;
    qs_error_t rc = fwrite(&h, sizeof(h), 1, f) == 1 ? QS_OK : QS_ERR_FILE;
    const size_t total = (size_t)h.length * h.elem_bytes;
    char *buf = rc == QS_OK ? malloc(total < ORK_FILE_CHUNK ? (total ? total : 1) : ORK_FILE_CHUNK) : NULL;
    if (rc == QS_OK && !buf) rc = QS_ERR_OOM;
    ork_dev_t *d = ork_dev();
    for (size_t off = 0; rc == QS_OK && off < total; off += ORK_FILE_CHUNK) {
        const size_t n = total - off < ORK_FILE_CHUNK ? total - off : ORK_FILE_CHUNK;
        if (hipMemcpyAsync(buf, (const char *)state->data + off, n, hipMemcpyDeviceToHost,
                           (hipStream_t)d->stream) != hipSuccess ||
            hipStreamSynchronize((hipStream_t)d->stream) != hipSuccess) {
            (void)hipGetLastError();
            rc = QS_ERR_FILE;
            break;
        }
        if (fwrite(buf, 1, n, f) != n) rc = QS_ERR_FILE;
    }
    free(buf);
    if (fclose(f) != 0 && rc == QS_OK) rc = QS_ERR_FILE;
    return rc;
}

/* stored bytes of a (layout, n, element size) claimed by a file header, in 128-bit arithmetic so that
 * no header can wrap them; 0 when the layout is unknown or the bytes exceed ORK_FILE_MAX_BYTES */
#define ORK_FILE_MAX_BYTES ((unsigned __int128)1 << 56)
static size_t file_bytes(int layout, unsigned n, uint64_t length, uint32_t elem) {
    if (n > 63) return 0;
    const unsigned __int128 dim = (unsigned __int128)1 << n;
    unsigned __int128 len;
    if (layout == PURE) len = dim;
    else if (layout == MIXED_PACKED) len = n == 0 ? 1 : (dim >> 1) * (dim + 1);
    else if (layout == MIXED_TILED) {
        const unsigned __int128 nt = (dim + TILE_DIM - 1) >> LOG_TILE_DIM;
        len = (nt * (nt + 1) >> 1) * (unsigned __int128)TILE_SIZE;
    } else return 0;
    const unsigned __int128 bytes = len * elem;
    if (len != length || bytes == 0 || bytes > ORK_FILE_MAX_BYTES) return 0;
    return (size_t)bytes;
}

qs_error_t orkan_state_load(state_t *state, const char *path) {
    if (!state || !path) return QS_ERR_NULL;
    FILE *f = fopen(path, "rb");
    if (!f) return QS_ERR_FILE;
    ork_file_header_t h;
    qs_error_t rc = QS_OK;
    if (fread(&h, sizeof(h), 1, f) != 1) rc = ferror(f) ? QS_ERR_FILE : QS_ERR_FORMAT;
    state_t probe;
    if (rc == QS_OK) {
        probe.type = (state_type_t)h.type;
        probe.qubits = (qubit_t)h.qubits;
        probe.data = NULL;
        const int layout = (int)h.type & 0xff;
        if (memcmp(h.magic, ORK_FILE_MAGIC, 8) != 0 || h.version != ORK_FILE_VERSION || h.byte_order != 0x01020304u ||
            h.qubits > 63 || !ork_type_ok(&probe) || h.elem_bytes != ork_elem_bytes(&probe) ||
            !file_bytes(layout, h.qubits, h.length, h.elem_bytes) ||
            (layout == MIXED_TILED && h.log_tile != LOG_TILE_DIM))
            rc = QS_ERR_FORMAT;
    }
    const size_t total = rc == QS_OK ? file_bytes((int)h.type & 0xff, h.qubits, h.length, h.elem_bytes) : 0;
    void *dev = NULL;
    if (rc == QS_OK) {
        /* the payload must be exactly the header's length */
        const long here = ftell(f);
        if (here < 0 || fseek(f, 0, SEEK_END) != 0) rc = QS_ERR_FILE;
        else {
            const long end = ftell(f);
            if (end < 0 || fseek(f, here, SEEK_SET) != 0) rc = QS_ERR_FILE;
            else if ((size_t)(end - here) != total) rc = QS_ERR_FORMAT;
        }
    }
    if (rc == QS_OK) {
        dev = dev_zeros(total);
        if (!dev) rc = QS_ERR_OOM;
    }
    char *buf = rc == QS_OK ? malloc(total < ORK_FILE_CHUNK ? (total ? total : 1) : ORK_FILE_CHUNK) : NULL;
    if (rc == QS_OK && !buf) rc = QS_ERR_OOM;
    ork_dev_t *d = rc == QS_OK ? ork_dev() : NULL;   /* no device use for a rejected file */
    for (size_t off = 0; rc == QS_OK && off < total; off += ORK_FILE_CHUNK) {
        const size_t n = total - off < ORK_FILE_CHUNK ? total - off : ORK_FILE_CHUNK;
        if (fread(buf, 1, n, f) != n) {
            rc = ferror(f) ? QS_ERR_FILE : QS_ERR_FORMAT;
            break;
        }
        if (hipMemcpyAsync((char *)dev + off, buf, n, hipMemcpyHostToDevice, (hipStream_t)d->stream) != hipSuccess ||
            hipStreamSynchronize((hipStream_t)d->stream) != hipSuccess) {
            (void)hipGetLastError();
            rc = QS_ERR_FILE;
        }
    }
    free(buf);
    fclose(f);
    if (rc != QS_OK) {
        if (dev) dev_release(dev);
        return rc;
    }
    state_free(state);
    state->type = (state_type_t)h.type;
    state->qubits = (qubit_t)h.qubits;
    state->data = dev;
    return QS_OK;
}
