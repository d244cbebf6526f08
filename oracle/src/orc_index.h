/*
 * orc_index.h -- index enumeration and storage offsets of the three state
 * layouts, restated for the CPU oracle.
 *
 * TEST INFRASTRUCTURE: this directory is the CPU oracle that the tests and the
 * CPU-baseline leg of the benchmark check the HIP library against.  Nothing in
 * the shipped library includes, links or calls it.
 *
 * Restates src/internal/index.h:28-88 of the reference:
 *   ins0      -- insertBit0       (index.h:28): insert a 0 bit at `pos`
 *   ins0x2    -- insertBits2_0    (index.h:39): insert 0 bits at lo < hi
 *   pk_col    -- col_off          (index.h:66): first packed element of a column
 *   pk_idx    -- pack_idx         (index.h:57): packed (r >= c) offset
 *   tl_tile   -- tile_off         (index.h:79): first element of tile (tr >= tc)
 *   tl_elem   -- elem_off         (index.h:86): row-major offset inside a tile
 */
#ifndef ORC_INDEX_H
#define ORC_INDEX_H

#include "q_types.h"

static inline idx_t ins0(idx_t v, unsigned pos) {
    const idx_t low = v & ((((idx_t)1) << pos) - 1);
    return low | ((v ^ low) << 1);
}

static inline idx_t ins0x2(idx_t v, unsigned lo, unsigned hi) {
    return ins0(ins0(v, lo), hi);
}

static inline idx_t pk_col(idx_t dim, idx_t c) {
    return c * (2 * dim - c + 1) / 2;
}

static inline idx_t pk_idx(idx_t dim, idx_t r, idx_t c) {
    return pk_col(dim, c) + (r - c);
}

static inline idx_t tl_tile(idx_t tr, idx_t tc) {
    return (tr * (tr + 1) / 2 + tc) * (idx_t)TILE_SIZE;
}

static inline idx_t tl_elem(idx_t lr, idx_t lc) {
    return lr * (idx_t)TILE_DIM + lc;
}

/* Offset of (r, c), r >= c, in the tiled layout (src/state/state_tiled.c:55-61). */
static inline idx_t tl_idx(idx_t r, idx_t c) {
    return tl_tile(r >> LOG_TILE_DIM, c >> LOG_TILE_DIM)
         + tl_elem(r & (TILE_DIM - 1), c & (TILE_DIM - 1));
}

#ifndef OMP_THRESHOLD_PURE
#define OMP_THRESHOLD_PURE 4096
#endif
#ifndef OMP_THRESHOLD_MIXED
#define OMP_THRESHOLD_MIXED 512
#endif

#endif /* ORC_INDEX_H */
