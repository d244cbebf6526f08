/* orc_internal.h -- oracle-internal traversals shared by gate and channel code.
 * TEST INFRASTRUCTURE only. */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H
#include "state.h"
#include "orc_ops.h"
void orc_packed_1q(state_t *st, unsigned q, const op1_t *op);
void orc_tiled_1q(state_t *st, unsigned q, const op1_t *op);
/* 2-qubit superoperator S (16 x 16 row-major, local index bit 0 <-> q1) on every 4x4 block */
void orc_dm_sop2(state_t *st, unsigned q1, unsigned q2, const cplx_t *S);
#endif
