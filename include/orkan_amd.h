/*
 * orkan_amd.h -- additions of the MI355X library on top of the Orkan C API.
 *
 * None of these has a counterpart in the reference's public headers; each
 * comment names the reference interface it complements where there is one.
 */
#ifndef ORKAN_AMD_H
#define ORKAN_AMD_H

#include "orkan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------------
 * FP32 storage.  OR this flag into state_t.type (PURE | ORKAN_FP32, ...): the
 * state is then stored as complex64 (8 B per element, half the HBM traffic)
 * while every kernel computes in FP64 registers.  All API functions accept
 * such states; state_get/state_set convert to and from cplx_t.
 * ---------------------------------------------------------------------------------------------- */
#define ORKAN_FP32 0x100

/* New state with the same layout and contents in the other precision
 * (to_fp32 = 1: complex128 -> complex64; 0: back).  data == NULL on failure. */
state_t orkan_state_convert(const state_t *state, int to_fp32);

/* ------------------------------------------------------------------------------------------------
 * Bulk transfer of the storage array (state_len elements in the state's own
 * precision: cplx_t, or float pairs for ORKAN_FP32).  Complements state_get /
 * state_set (reference include/state.h:128,141), which move one element.
 * Return 0 on success, a hipError_t value otherwise.
 * ---------------------------------------------------------------------------------------------- */
int orkan_state_upload(state_t *state, const void *host);
int orkan_state_download(const state_t *state, void *host);

/* ------------------------------------------------------------------------------------------------
 * Device-resident arrays for mean / matel_diag / exp_diag.  Those functions
 * accept either a host array (copied to HBM on every call) or a pointer from
 * orkan_dev_alloc (used in place, no copy).
 * ---------------------------------------------------------------------------------------------- */
double *orkan_dev_alloc(idx_t count);
int orkan_dev_upload(double *dst_dev, const double *src_host, idx_t count);
void orkan_dev_free(void *p);

/* ------------------------------------------------------------------------------------------------
 * Stream and timing.  All kernels of the library run in order on one stream
 * per device; gates return as soon as their kernel is enqueued, and every call
 * that returns host data waits for it.
 * ---------------------------------------------------------------------------------------------- */
void *orkan_stream(void);                 /* the library's hipStream_t on the current device */
int orkan_sync(void);                     /* wait for all enqueued work */
int orkan_event_record(int slot);         /* record event `slot` (0..63) on the stream */
double orkan_event_ms(int a, int b);      /* elapsed ms between two recorded events */

#ifdef __cplusplus
}
#endif
#endif /* ORKAN_AMD_H */
