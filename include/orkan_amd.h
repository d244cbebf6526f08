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
 * Layout conversion and state files.  No counterpart in the reference, which has neither a
 * conversion nor a file format but reserves QS_ERR_FILE / QS_ERR_FORMAT for one
 * (include/q_types.h:35-36).
 * ---------------------------------------------------------------------------------------------- */
/* *dst = new state (device memory, the source's precision) holding the logical density matrix of
 * src in `layout` (MIXED_PACKED or MIXED_TILED); a PURE source gives |psi><psi|.  dst's previous
 * contents are overwritten, not freed.  QS_ERR_NULL, QS_ERR_TYPE (pure or unknown destination,
 * unknown source type), QS_ERR_OOM. */
qs_error_t orkan_state_relayout(const state_t *src, state_t *dst, state_type_t layout);
/* File = 64-byte header ("ORKAN-ST", version 1, type incl. ORKAN_FP32, qubits, LOG_TILE_DIM,
 * stored length, element bytes, byte-order mark 0x01020304; little-endian), then the storage
 * array exactly as the state holds it.  save: QS_ERR_NULL, QS_ERR_TYPE, QS_ERR_FILE (open, write
 * or device copy failed), QS_ERR_OOM.  load: frees the state's old data (state_free) and
 * replaces type, qubits and data; QS_ERR_FILE (open or read failed), QS_ERR_FORMAT (bad magic,
 * version, byte order or type, length not matching type and qubits, tiled file of another
 * LOG_TILE_DIM, payload not exactly the stated length), QS_ERR_OOM; on error the state is left
 * unchanged. */
qs_error_t orkan_state_save(const state_t *state, const char *path);
qs_error_t orkan_state_load(state_t *state, const char *path);

/* ------------------------------------------------------------------------------------------------
 * Fused layer of single-qubit gates (no counterpart in the reference, which applies one gate per
 * call; an addition for circuits that apply a layer of rotations on every qubit, reported apart
 * from the per-gate figures).  Applies mats[4 q .. 4 q + 3], the 2x2 matrix of qubit q in column-major order (the
 * single_from_mat convention), to every qubit q of a PURE state: the same state as calling the gates
 * one by one (they act on distinct qubits and commute), to rounding.  The statevector moves through
 * HBM in 1 + ceil((n - 12) / 8) LDS-tiled sweeps (default tile) instead of n passes.  QS_ERR_NULL, QS_ERR_TYPE (not
 * PURE).
 * ---------------------------------------------------------------------------------------------- */
qs_error_t orkan_layer_1q(state_t *state, const cplx_t *mats);

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
/* byte copies between host memory and device memory from orkan_dev_alloc (stream-ordered, then waited for) */
int orkan_copy_to_host(void *dst_host, const void *src_dev, size_t bytes);
int orkan_copy_to_device(void *dst_dev, const void *src_host, size_t bytes);

/* ------------------------------------------------------------------------------------------------
 * Allocation-free exp_diag.  No counterpart in the reference (whose exp_diag, include/circ.h:41,
 * is declared in circ.h here unchanged): an entry point for callers that cannot allocate device
 * memory -- code under HIP graph capture of the library stream (orkan_stream), tight serving
 * loops, harnesses that forbid hipMalloc.  Same result as exp_diag(state, diag_dev, t) (the same
 * kernels on the library stream, enqueued and not waited for), but it allocates and frees nothing:
 * the per-index phase scratch of a density matrix is a fixed device array of the library,
 * 2^ORKAN_NOALLOC_MAX_QUBITS complex128 (2 MiB, held on every GPU the library runs on until the
 * process ends), and diag_dev must already be device memory (orkan_dev_alloc), never staged.
 * Returns QS_OK (also for t == 0, which leaves the state untouched); QS_ERR_NULL; QS_ERR_TYPE;
 * QS_ERR_QUBIT for a density matrix of more than ORKAN_NOALLOC_MAX_QUBITS qubits (PURE states use
 * no scratch and have no such limit); QS_ERR_PARAM (with a stderr line) when diag_dev is not
 * device memory, when no state has been created on the current device yet, or when the launch
 * fails.  Every call on one GPU uses the same scratch array, in stream order on the library
 * stream: calls that overlap on one GPU -- graphs captured from it and replayed concurrently on
 * other streams -- race.
 * ---------------------------------------------------------------------------------------------- */
#define ORKAN_NOALLOC_MAX_QUBITS 17
int orkan_exp_diag_noalloc(state_t *state, const double *diag_dev, double t);

/* ------------------------------------------------------------------------------------------------
 * Stream and timing.  All kernels of the library run in order on one stream
 * per device; gates return as soon as their kernel is enqueued, and every call
 * that returns host data waits for it.
 * ---------------------------------------------------------------------------------------------- */
void *orkan_stream(void);                 /* the library's hipStream_t on the current device */
int orkan_sync(void);                     /* wait for all enqueued work */
int orkan_event_record(int slot);         /* record event `slot` (0..65535) on the stream */
double orkan_event_ms(int a, int b);      /* elapsed ms between two recorded events */
/* tuning hook: launch shape of the density-matrix kernels (kernel 0: tile groups of both layouts,
 * 1: phases of both layouts, 2 / 3: tile groups of the packed / tiled layout, 4 / 5: phases of the
 * packed / tiled layout, 6 / 7: streamed tiled 1-qubit superoperator with the target inside / above
 * the tile, 8: streamed superoperator on (lds_bytes != 0) or off, in favour of the tile groups,
 * 9: streamed packed superoperator, 10: statevector mean / matel_diag reduction (flags bit 0: 8 loads in flight
 * per thread, else 4; bits 1-3: k, at most 2048 >> k workgroups, k <= 4; bit 4: the shape of
 * complex64 states, else of complex128 ones),
 * 11: surrogate GEMM tile where the shape allows (flags only: 0 128 x 128, 1 256 x 256 staged
 * through registers, 2 256 x 256 streamed into LDS by DMA, the default), 12: statevector 1-qubit
 * kernels for every target qubit (flags bit 4 on: bit 0 XCD remap, bit 1 nontemporal loads and
 * stores, bits 2-3 log2 of the pairs per thread, 0..2, in complex128 terms (complex64 storage runs
 * twice the pairs, at most 4, under half the LDS cap); bit 4 off: back to the tuned per-qubit tables),
 * 13: nontemporal (streaming) accesses, used on states of at least lds_bytes (0: 512 MiB, the
 * default) by the statevector kernels and the streamed tiled density-matrix kernels; flags switch
 * the statevector ones off per family (bit 0 1-qubit gates, bit 1 cx / cz / swap / ccx ..., bit 2
 * exp_diag, bit 3 mean / matel_diag; 0, the default, leaves them on where the tuned shapes use
 * them), 14: every
 * launch shape and tuning choice above back to the built-in defaults (flags and lds_bytes unused),
 * 15 / 16: as 6 / 7 for complex64 storage (kernels 6 and 7 then apply to complex128 only),
 * 17: tile and workgroup size of the orkan_layer_1q sweeps (flags only: 0 2^12 amplitudes / 256
 * threads, 1 2^12 / 512, the default, 2 2^11 / 256, 3 2^11 / 512)), 18: as 9 for complex64 storage
 * (kernel 9 then applies to complex128 only), 19: statevector exp_diag (flags bits 0-1: log2 of the
 * amplitudes per thread, 0..2, bit 2: the shape of complex64 states, else of complex128 ones;
 * lds_bytes as below), 20: permutation / phase kernels (cx, cz, swap, ...) on complex64 states
 * (lds_bytes only), 21: statevector permutations (cx, cy, swap, ccx, ...): flags bits 0-1 log2 of
 * the lane-exchange kernel's amplitudes per thread (1..3), bit 2 whole-line writes, bit 3 lane-exchange
 * kernel off, bit 4 lane-exchange kernel for every gate whose qubits are all below 6 (else only where
 * one of them lies inside a 128 B line), bit 5 paired complex64 kernel (two amplitudes per 16 B access,
 * no gate qubit at bit 0) off, bit 6 the amplitudes per thread and lds_bytes are those of complex64
 * states, bit 7 lds_bytes is the paired kernel's cap instead; 22: density-matrix permutations whose qubits
 * all lie at or above the tile (32 x 32 blocks for packed), which move whole tiles: flags bits 0 / 5 as
 * below, bit 1 sends them (and those whose qubits all lie below 5, which permute each 32 x 32 tile
 * within itself, and cx / cy / ccx, which otherwise run through the streamed kernels on their target)
 * to the tile-group kernel instead, bit 2 nontemporal accesses (tiled layout), bits 3-4 log4 of the
 * consecutive tiles each workgroup walks (0..2), bit 6 moves packed cx / cy / ccx with every qubit at
 * or above the 32 x 32 blocks as whole blocks (else they stream where their unchanged entries can be
 * left alone), bit 7 (kernel 22 only) sends swap and the other permutations whose lowest qubit lies
 * below the tile and highest at or above it to the tile-group kernel (else they move virtual tiles of
 * 2^q x 2^q entries, q the lowest qubit, where q >= 3 for complex128 and q >= 4 for complex64, with
 * bit 2's nontemporal accesses for complex64 only);
 * 23 / 24 / 25: cx and cy on packed states / controlled
 * gates on tiled states / ccx on packed states leave the entries they do not change unread and
 * unwritten when their lowest control qubit is at least the bits of a 128 B line (3 complex128, 4
 * complex64) plus flags (0..63, 63: never; defaults 2 / 0 / 0; lds_bytes must be 0);
 * flags bit 0 keeps consecutive groups on one XCD, bit 1 walks the groups column by column, bit 2
 * (phase kernels 1, 4, 5 and streamed kernels 6, 7, 9, 15, 16, 18) uses nontemporal loads and stores,
 * bits 3-4 (kernels 7, 16 with 32 x 32 tiles; 9, 18) give log2 of the workgroups that share one tile quad
 * or packed block (0..2), and (phase kernels 1, 4, 5) log2 of the multiple of 64 B each thread has in
 * flight (0..2; 0, the default: 1 x), bit 3 of kernels 6, 15 (with bit 2) selects the lane-exchange kernel for
 * in-tile targets whose quad columns share 128 B lines, bit 5 hands each XCD runs of 32 consecutive
 * groups instead of the contiguous ranges of bit 0, bit 6 (phase kernels 1, 4, 5) writes every 128 B
 * line holding a changed entry in full where the changed entries do not fill lines;
 * lds_bytes caps workgroups per CU (0: no cap).  The settings are process-wide and unsynchronised:
 * change them while no other thread is calling into the library.  Returns 0, or hipErrorInvalidValue
 * for a kernel id or value outside the ranges above. */
int orkan_set_launch_shape(int kernel, int flags, int lds_bytes);

/* ------------------------------------------------------------------------------------------------
 * Batched QAOA evaluation.  One evaluation is the reference call sequence
 *   state_plus(PURE, n); p x { exp_diag(C, gamma_l); rx(q, 2 beta_l) for q = 0..n-1 }; mean(C)
 * for a graph's cost diagonal C (2^n doubles).  All evaluations of one call
 * run in a single persistent kernel launch replayed from a hipGraph; states of
 * up to 2^LT amplitudes stay in LDS, larger ones (n <= 2 LT) stream through HBM
 * in p + 1 LDS-tiled sweeps.
 * ---------------------------------------------------------------------------------------------- */
typedef struct orkan_qaoa orkan_qaoa_t;

/* diags[g]: host or device array of 2^nq[g] doubles, copied at creation (1 <= nq <= 26) */
orkan_qaoa_t *orkan_qaoa_create(int ngraphs, const unsigned char *nq, const double *const *diags, int max_evals,
                                int max_p);
void orkan_qaoa_destroy(orkan_qaoa_t *q);
/* evaluation e: graph[e], p[e] layers (1..max_p), gamma/beta rows of max_p values
 * (row e, first p[e] used); out[e] = <C> after the circuit.  0 on success. */
int orkan_qaoa_eval(orkan_qaoa_t *q, int nevals, const int *graph, const int *p, const double *gamma,
                    const double *beta, double *out);
/* launch configuration: LDS tile 2^lt amplitudes (10..13), workgroup size, workgroups per CU (0: max) */
int orkan_qaoa_set_config(orkan_qaoa_t *q, int lt, int threads, int blocks_per_cu);
int orkan_qaoa_get_config(const orkan_qaoa_t *q, int *lt, int *threads, int *grid);
/* time every candidate configuration on this evaluation list and keep the fastest */
int orkan_qaoa_autotune(orkan_qaoa_t *q, int nevals, const int *graph, const int *p, const double *gamma,
                        const double *beta, int reps, double *best_ms);
/* MaxCut cost diagonal C(x) = sum_e w_e [x_ei != x_ej] computed on the device into out_dev (2^n doubles) */
int orkan_maxcut_diag_dev(int n, int nedges, const int *ei, const int *ej, const double *w, double *out_dev);

/* ------------------------------------------------------------------------------------------------
 * Energy surrogate: an MLP in_dim -> hidden (x nhidden, GELU) -> 1 trained and
 * evaluated with BF16 MFMA GEMMs (fp32 accumulation) and Adam on fp32 master
 * weights.  Inputs are fp32 rows of in_dim features (host or device memory).
 * in_dim must be a multiple of 32 and hidden of 128.
 * ---------------------------------------------------------------------------------------------- */
typedef struct orkan_mlp orkan_mlp_t;

orkan_mlp_t *orkan_mlp_create(int in_dim, int hidden, int nhidden, int max_batch);
void orkan_mlp_destroy(orkan_mlp_t *m);
/* W[l]: fp32 [K_l][hidden] row-major (K_0 = in_dim, else hidden); b[l]: [hidden]; w_out: [hidden] */
int orkan_mlp_set_weights(orkan_mlp_t *m, const float *const *W, const float *const *b, const float *w_out,
                          float b_out);
int orkan_mlp_get_weights(orkan_mlp_t *m, float *const *W, float *const *b, float *w_out, float *b_out);
int orkan_mlp_forward(orkan_mlp_t *m, int batch, const float *x, float *y);
/* screening: predictions for ncand rows [gf (ng floats), angles_dev[r] (in_dim - ng floats)], with
 * the rows assembled on the device; angles_dev and y_dev are device arrays (ng <= max_batch) */
int orkan_mlp_forward_candidates(orkan_mlp_t *m, const float *gf_host, int ng, const float *angles_dev, long ncand,
                                 float *y_dev);
/* one Adam step on the mean squared error of the batch (batch <= max_batch) */
int orkan_mlp_train_step(orkan_mlp_t *m, int batch, const float *x, const float *target, float lr, float *loss);
/* gradients of the mean squared error without updating the weights */
int orkan_mlp_grads(orkan_mlp_t *m, int batch, const float *x, const float *target, float *const *dW,
                    float *const *db, float *dw_out, float *db_out, float *loss);
/* indices (and values) of the k largest entries of a device float array, largest first */
int orkan_topk_f32(const float *y_dev, long n, int k, int *idx_host, float *val_host);
/* fp32 C [M][N] = A [M][K] B [N][K]^T for bf16 device arrays (M, N multiples of 128, K of 32) */
int orkan_mlp_gemm_test(int M, int N, int K, const void *A, const void *B, float *C);

#ifdef __cplusplus
}
#endif
#endif /* ORKAN_AMD_H */
