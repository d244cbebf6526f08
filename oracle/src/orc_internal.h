/* orc_internal.h -- oracle-internal traversals shared by gate and channel code.
 * TEST INFRASTRUCTURE only. */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H
#include "state.h"
#include "orc_ops.h"
void orc_packed_1q(state_t *st, unsigned q, const op1_t *op);
void orc_tiled_1q(state_t *st, unsigned q, const op1_t *op);
#endif
