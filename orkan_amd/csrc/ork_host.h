/*
 * ork_host.h -- internals of the C17 host layer: per-device runtime (stream,
 * reduction partials, staging buffers), error reporting, and the helpers the
 * API modules share.
 */
#ifndef ORK_HOST_H
#define ORK_HOST_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "orkan.h"
#include "orkan_amd.h"
#include "ork_kernels.h"

/* Fail-fast reporting in the reference's format: "orkan: <module>: <message>"
 * then exit(EXIT_FAILURE) (src/gate/gate.c:19-24). */
#define ORK_VALIDATE(mod, cond, msg)                               \
    do {                                                           \
        if (!(cond)) {                                             \
            fprintf(stderr, "orkan: " mod ": %s\n", (msg));        \
            exit(EXIT_FAILURE);                                    \
        }                                                          \
    } while (0)

/* a failed HIP call or launch: report and exit like a validation failure */
void ork_hip_fail(const char *where, int err);
#define ORK_HIP(where, call)                                       \
    do {                                                           \
        int e_ = (int)(call);                                      \
        if (e_ != 0) ork_hip_fail((where), e_);                    \
    } while (0)

typedef struct {
    void *stream;
    double *partials;      /* reduction partial sums */
    void *stage;           /* device copy of a caller's obs / diag array */
    size_t stage_bytes;
    void *phase;           /* exp_diag per-index phases */
    size_t phase_bytes;
    int noalloc_err;       /* result of ork_noalloc_setup() on this device */
} ork_dev_t;

/* runtime of the current device (created on first use) */
ork_dev_t *ork_dev(void);
/* runtime of the current device if it has been created, else NULL (creates nothing) */
ork_dev_t *ork_dev_existing(void);
/* does p address device memory? */
int ork_is_dev_ptr(const void *p);
void ork_lock(void);
void ork_unlock(void);

/* device pointer to `count` doubles equal to the caller's array: the pointer
 * itself when it already addresses device memory, else a staged copy.  Call
 * under ork_lock(). */
const double *ork_dev_doubles(const double *host_or_dev, size_t count);

/* layout and storage precision encoded in state->type */
static inline int ork_layout(const state_t *s) { return (int)s->type & 0xff; }
static inline int ork_prec(const state_t *s) { return ((int)s->type & ORKAN_FP32) ? ORK_F32 : ORK_F64; }
static inline size_t ork_elem_bytes(const state_t *s) { return ork_prec(s) == ORK_F32 ? 8 : 16; }
static inline int ork_type_ok(const state_t *s) {
    const int t = (int)s->type & ~ORKAN_FP32;
    return t == PURE || t == MIXED_PACKED || t == MIXED_TILED;
}

idx_t ork_len(int layout, qubit_t n);

#endif /* ORK_HOST_H */
