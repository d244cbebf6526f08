// k_dm.hip -- density-matrix kernels on the packed and tiled Hermitian layouts.
//
// One workgroup owns one "tile group": the W x W block of logical entries
// (rows R-group x columns C-group) that a k-qubit operation mixes.  The group is
// made of 2^g x 2^g virtual tiles of B x B entries (B = 2^LB); qubits below LB
// act inside a virtual tile, the g qubits at or above LB pick tiles of the
// group.  Only groups with R >= C are launched: a group with R < C is the
// conjugate transpose of the group (C, R), so every stored entry is owned by
// exactly one workgroup and no two workgroups touch the same entry.
//
// load  : HBM -> LDS in the layout's memory order (consecutive lanes read
//         consecutive addresses).  A virtual tile that is not stored (it lies
//         above the diagonal) is read from its stored mirror and transposed
//         and conjugated on the way into LDS; the unstored upper half of a
//         packed diagonal tile is filled from its lower half.
// apply : 1-qubit superoperator on 2x2 blocks, or a dense 2^k x 2^k unitary as
//         a row pass (U rho) then a column pass (. U^dagger), in LDS;
//         permutations with phases are applied as a gather during the store.
// store : LDS -> HBM, again in memory order, writing every stored entry of the
//         group once (and the conjugate transpose into a stored mirror tile).
//
// The reference walks the same 2x2 / 4x4 / 8x8 blocks with one OpenMP loop nest
// per gate and case (src/gate/packed/gate_packed_1q.c:65-104,
// src/gate/tiled/gate_tiled_h.c:28-225, src/channel/channel_tiled.c:52-185,
// src/gate/tiled/gate_tiled_2q.c:3).
#include "k_common.hpp"

namespace {

constexpr int DM_THREADS = 256;
constexpr int LW_MAX = 5;                  // groups of up to 32 x 32 entries
constexpr int W_MAX = 1 << LW_MAX;
constexpr int LDS_STRIDE = W_MAX + 1;      // padded row: column walks hit distinct banks

struct DmPlan {
    u64 dim;
    u64 ngroups;
    u64 nt;         // group rows (dim >> LW)
    int xcd;        // remap group ids so consecutive groups run on one XCD
    int col;        // group order: 0 row by row (R outer), 1 column by column (C outer)
    int L;          // storage tile log (tiled layout)
    int LB;         // virtual tile log
    int g;          // gate qubits at or above LB
    int LW;         // LB + g
    int gpos[3];    // tile-index bit of each such qubit (ascending)
};

struct DmOp {
    int k;
    int lbit[3];    // LDS coordinate bit of gate argument j
    int srt[3];     // lbit sorted ascending
    int pinv[8];    // MONO: local index a' <- pinv[a']
    int ph[8];      // MONO: U|a> = i^ph[a] |perm a>
    cd m[64];       // SOP: 4x4 row-major; DENSE: 2^k x 2^k row-major
    const cd *s2;   // SOP2: 16x16 row-major superoperator in device memory
    u64 still;      // MONO: bit a * 8 + b set when entries with gate bits (a, b) keep their value
    // controlled single-target permutation run by the streamed 1-qubit kernels (cx, cy, ccx): the quad's
    // rows go through U when the row index has every bit of cmask, its columns when the column index
    // has; U = [[0, i^pu01], [i^pu10, 0]] on the target.  skip: quads with neither (which keep their
    // value) are neither read nor written -- set where every control bit lies above the bits of a
    // 128 B line, so that the quads left out cover whole lines
    int ctl, pu01, pu10, skip;
    u64 cmask;
};

__device__ __forceinline__ u64 spread_bits(int a, const int *pos, int k) {
    u64 v = 0;
    for (int j = 0; j < k; ++j)
        if (a >> j & 1) v |= (u64)1 << pos[j];
    return v;
}
__device__ __forceinline__ int gather_bits(u64 x, const int *pos, int k) {
    int a = 0;
    for (int j = 0; j < k; ++j) a |= (int)((x >> pos[j]) & 1) << j;
    return a;
}

// with a grid of one workgroup per group, blocks are dispatched round-robin over
// the 8 XCDs; the remap hands consecutive groups (neighbouring tiles) to one XCD
// (xcd 1); xcd 2 hands each XCD runs of 32 consecutive groups in turn, so that neighbouring groups
// mostly share an XCD while all XCDs still work on nearby groups at any time
__device__ __forceinline__ u64 remap_group(u64 w, u64 n, int xcd) {
    if (!xcd || gridDim.x < n) return w;
    if (xcd == 2) {
        constexpr u64 G = 32, SPAN = 8 * G;
        if (w < (n / SPAN) * SPAN) {
            const u64 x = w & 7, k = w >> 3;
            return (k / G) * SPAN + x * G + (k % G);
        }
        return w;
    }
    const u64 per = n >> 3;
    if (w < (per << 3)) return (w & 7) * per + (w >> 3);
    return w;
}

// group id -> (R, C), R >= C.  Column order walks a column of groups (same C,
// R ascending), which for the column-major packed layout makes neighbouring
// groups share the cache lines at the ends of their column runs.
__device__ __forceinline__ void block_rc(u64 w0, u64 n, u64 nt, int xcd, int col, u64 &R, u64 &C) {
    const u64 w = remap_group(w0, n, xcd);
    if (!col) {
        tri_decode(w, R, C);
    } else {
        u64 rr, cc;
        tri_decode(n - 1 - w, rr, cc);
        R = nt - 1 - cc;
        C = nt - 1 - rr;
    }
}
__device__ __forceinline__ void group_rc(const DmPlan &p, u64 w0, u64 &R, u64 &C) {
    block_rc(w0, p.ngroups, p.nt, p.xcd, p.col, R, C);
}

template <int LAYOUT>
__device__ __forceinline__ u64 elem_off(const DmPlan &p, u64 r, u64 c) {
    if constexpr (LAYOUT == ORK_PACKED) return packed_off(p.dim, r, c);
    else return tiled_off(p.L, r, c);
}

// is virtual tile (TR, TC) stored in full (tiled: its storage tile is lower or
// diagonal; packed: strictly below the diagonal)?  Packed diagonal tiles
// (TR == TC) are stored in their lower half only.
template <int LAYOUT>
__device__ __forceinline__ bool tile_stored(const DmPlan &p, u64 TR, u64 TC) {
    if constexpr (LAYOUT == ORK_PACKED) return TR >= TC;
    else return (TR >> (p.L - p.LB)) >= (TC >> (p.L - p.LB));
}

// Permutations: per LDS coordinate x (one row or column index of the group), a table entry built once per
// workgroup: bits 0-7 the source coordinate (x with its gate bits a replaced by pinv[a]), bits 8-9 the
// phase exponent ph[pinv a] of that source, bits 12-14 the gate bits a.  (Computing these per entry read
// pinv[] and ph[] from the kernel arguments with a per-lane index, a memory load waited for per entry.)
__device__ __forceinline__ int mono_entry(const DmOp &op, int x) {
    const int k = op.k;
    const u64 mask = spread_bits((1 << k) - 1, op.lbit, k);
    const int a = gather_bits((u64)x, op.lbit, k), ai = op.pinv[a];
    const int xs = (int)(((u64)x & ~mask) | spread_bits(ai, op.lbit, k));
    return xs | (op.ph[ai] & 3) << 8 | a << 12;
}

// does LDS entry (x, y) change?  Only a permutation leaves entries as they are.
template <int OP>
__device__ __forceinline__ bool moved(const DmOp &op, const int *mt, int x, int y) {
    if constexpr (OP == ORK_DM_MONO) {
        const int a = (mt[x] >> 12) & 7, b = (mt[y] >> 12) & 7;
        return !((op.still >> (a * 8 + b)) & 1);
    } else {
        return true;
    }
}

// LDS entry (x, y) of the updated group; permutations with phases are applied
// here as a gather: rho'(perm a, perm b) = i^(ph[a] - ph[b]) rho(a, b)
template <int OP>
__device__ __forceinline__ cd lds_get(const cd *lds, const int *mt, int x, int y) {
    if constexpr (OP == ORK_DM_MONO) {
        const int ex = mt[x], ey = mt[y];
        return cipow(lds[(ex & 0xff) * LDS_STRIDE + (ey & 0xff)], ((ex >> 8) & 3) - ((ey >> 8) & 3));
    } else {
        return lds[x * LDS_STRIDE + y];
    }
}

// HOIST: every load of a thread in flight before the first LDS write (4 per thread, at the cost of
// registers), else one load after the other at the occupancy the smaller register count allows
// NTA: nontemporal loads and stores (tiled layout with virtual tiles of whole 128 B rows)
template <typename S, int LAYOUT, int OP, int HOIST, int NTA>
__global__ __launch_bounds__(DM_THREADS) void dm_group_kernel(S *__restrict__ rho, DmPlan p, DmOp op) {
    __shared__ cd lds[W_MAX * LDS_STRIDE];
    // SOP2: the 16 x 16 superoperator staged in LDS once (every lane reads the same entry: broadcasts),
    // instead of 256 dependent loads from device memory per 4 x 4 block and thread
    __shared__ cd s2l[OP == ORK_DM_SOP2 ? 256 : 1];
    constexpr int EPT = W_MAX * W_MAX / DM_THREADS;   // entries of a group per thread
    const int LB = p.LB, g = p.g, LW = p.LW;
    const int B = 1 << LB, W = 1 << LW, NT = 1 << g;
    const int total = W * W;
    const int tid = threadIdx.x;
    // per-workgroup tables: the tile-index bits of each group tile i (the gate qubits at or above LB spread
    // to their positions), and the permutation entries of each LDS coordinate
    __shared__ u64 gsp[8];
    __shared__ int mt[OP == ORK_DM_MONO ? W_MAX : 1];
    if (tid < 8) gsp[tid] = tid < NT ? spread_bits(tid, p.gpos, g) : 0;
    if constexpr (OP == ORK_DM_MONO) {
        if (tid < W) mt[tid] = mono_entry(op, tid);
    }
    if (OP == ORK_DM_MONO || g) __syncthreads();
    // SOP2: the 16 x 16 superoperator, read only after the first group's load barrier (staged after the
    // barrier above: waiting there for these loads cost each workgroup, most of which take one group, ~2 %)
    if constexpr (OP == ORK_DM_SOP2) {
        for (int i = threadIdx.x; i < 256; i += DM_THREADS) s2l[i] = op.s2[i];
    }

    for (u64 w0 = blockIdx.x; w0 < p.ngroups; w0 += gridDim.x) {
        u64 R, C;
        group_rc(p, w0, R, C);
        const u64 TRb = ins0k(R, p.gpos, g), TCb = ins0k(C, p.gpos, g);
        const bool diag_group = (R == C);

        // ---------------- load ----------------
        // for a permutation, entries that keep their value are neither loaded nor stored (nothing reads
        // them: the sources of the moved entries are the moved entries)
        if constexpr (HOIST) {
            cd lv[EPT];
            int lx[EPT], ly[EPT], lk[EPT];   // LDS position; 0 skip, 1 direct, 2 direct + transposed copy, 3 mirror
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                const int t = tid + u * DM_THREADS;
                lk[u] = 0;
                lx[u] = ly[u] = 0;
                lv[u] = cmk(0.0, 0.0);
                if (t >= total) continue;
                const int ti = t >> (2 * LB), e = t & (B * B - 1);
                const int i = ti >> g, j = ti & (NT - 1);
                const u64 TR = TRb | (g ? gsp[i] : 0), TC = TCb | (g ? gsp[j] : 0);
                const int sa = e >> LB, sb = e & (B - 1);                  // slow, fast
                const int sr = (LAYOUT == ORK_TILED) ? sa : sb;            // row-major vs column-major
                const int sc = (LAYOUT == ORK_TILED) ? sb : sa;
                if (tile_stored<LAYOUT>(p, TR, TC)) {
                    if (LAYOUT == ORK_PACKED && TR == TC && sr < sc) continue;
                    lx[u] = i * B + sr;
                    ly[u] = j * B + sc;
                    if (!moved<OP>(op, mt, lx[u], ly[u])) continue;
                    lk[u] = (LAYOUT == ORK_PACKED && TR == TC && sr > sc) ? 2 : 1;
                    lv[u] = ldx<NTA>(rho + elem_off<LAYOUT>(p, TR * B + sr, TC * B + sc));
                } else {
                    // stored mirror tile (TC, TR): its entry (sr, sc) is conj of ours at (sc, sr)
                    lx[u] = i * B + sc;
                    ly[u] = j * B + sr;
                    if (!moved<OP>(op, mt, lx[u], ly[u])) continue;
                    lk[u] = 3;
                    lv[u] = ldx<NTA>(rho + elem_off<LAYOUT>(p, TC * B + sr, TR * B + sc));
                }
            }
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                if (!lk[u]) continue;
                lds[lx[u] * LDS_STRIDE + ly[u]] = lk[u] == 3 ? cconj(lv[u]) : lv[u];
                if (lk[u] == 2) lds[ly[u] * LDS_STRIDE + lx[u]] = cconj(lv[u]);
            }
        } else {
            for (int t = tid; t < total; t += DM_THREADS) {
                const int ti = t >> (2 * LB), e = t & (B * B - 1);
                const int i = ti >> g, j = ti & (NT - 1);
                const u64 TR = TRb | (g ? gsp[i] : 0), TC = TCb | (g ? gsp[j] : 0);
                const int sa = e >> LB, sb = e & (B - 1);
                const int sr = (LAYOUT == ORK_TILED) ? sa : sb;
                const int sc = (LAYOUT == ORK_TILED) ? sb : sa;
                if (tile_stored<LAYOUT>(p, TR, TC)) {
                    if (LAYOUT == ORK_PACKED && TR == TC && sr < sc) continue;
                    if (!moved<OP>(op, mt, i * B + sr, j * B + sc)) continue;
                    const cd v = ldx<NTA>(rho + elem_off<LAYOUT>(p, TR * B + sr, TC * B + sc));
                    lds[(i * B + sr) * LDS_STRIDE + j * B + sc] = v;
                    if (LAYOUT == ORK_PACKED && TR == TC && sr > sc)
                        lds[(i * B + sc) * LDS_STRIDE + j * B + sr] = cconj(v);
                } else {
                    if (!moved<OP>(op, mt, i * B + sc, j * B + sr)) continue;
                    const cd v = ldx<NTA>(rho + elem_off<LAYOUT>(p, TC * B + sr, TR * B + sc));
                    lds[(i * B + sc) * LDS_STRIDE + j * B + sr] = cconj(v);
                }
            }
        }
        __syncthreads();

        // ---------------- apply ----------------
        if constexpr (OP == ORK_DM_SOP) {
            const int m = op.lbit[0];
            const u64 mb = (u64)1 << m;
            const int half = W >> 1;
            for (int b = tid; b < (total >> 2); b += DM_THREADS) {
                const int x0 = (int)ins0((u64)(b / half), m), y0 = (int)ins0((u64)(b % half), m);
                const int x1 = x0 | (int)mb, y1 = y0 | (int)mb;
                cd v[4];
                v[0] = lds[x0 * LDS_STRIDE + y0];
                v[1] = lds[x0 * LDS_STRIDE + y1];
                v[2] = lds[x1 * LDS_STRIDE + y0];
                v[3] = lds[x1 * LDS_STRIDE + y1];
                cd o[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    o[r] = cmk(0.0, 0.0);
#pragma unroll
                    for (int c = 0; c < 4; ++c) cfma(o[r], op.m[4 * r + c], v[c]);
                }
                lds[x0 * LDS_STRIDE + y0] = o[0];
                lds[x0 * LDS_STRIDE + y1] = o[1];
                lds[x1 * LDS_STRIDE + y0] = o[2];
                lds[x1 * LDS_STRIDE + y1] = o[3];
            }
            __syncthreads();
        } else if constexpr (OP == ORK_DM_SOP2) {
            // 2-qubit superoperator on each 4x4 block: out[(a,b)] = sum_(c,d) S[(a,b),(c,d)] in[(c,d)]
            // four threads per block (at most 64 blocks: one item per thread), each computing 4 of its 16
            // outputs; the block's inputs are all read before the barrier, the outputs written after it.
            // (One thread per block left three of the four waves idle through 256 complex FMAs.)
            const int LW2 = LW - 2, nb = total >> 4;
            const bool act = tid < 4 * nb;
            const int b = act ? tid % nb : 0, qd = act ? tid / nb : 0;   // block, quarter of its outputs
            const int x0 = (int)ins0k((u64)(b >> LW2), op.srt, 2);
            const int y0 = (int)ins0k((u64)(b & ((1 << LW2) - 1)), op.srt, 2);
            cd v[16];
            if (act) {
                int xo[4], yo[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    xo[a] = (x0 | (int)spread_bits(a, op.lbit, 2)) * LDS_STRIDE;
                    yo[a] = y0 | (int)spread_bits(a, op.lbit, 2);
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c) v[a * 4 + c] = lds[xo[a] + yo[c]];
            }
            __syncthreads();
            if (act) {
                // output rows 4 qd .. 4 qd + 3 of the superoperator (with 64 blocks a whole wave shares them,
                // so their LDS reads are broadcasts); a rolled loop keeps the rows out of registers
#pragma unroll 1
                for (int i = 4 * qd; i < 4 * qd + 4; ++i) {
                    cd o = cmk(0.0, 0.0);
#pragma unroll
                    for (int j = 0; j < 16; ++j) cfma(o, s2l[i * 16 + j], v[j]);
                    lds[(x0 | (int)spread_bits(i >> 2, op.lbit, 2)) * LDS_STRIDE + (y0 | (int)spread_bits(i & 3, op.lbit, 2))] = o;
                }
            }
            __syncthreads();
        } else if constexpr (OP == ORK_DM_DENSE) {
            const int k = op.k, K = 1 << k;
            const int LWk = LW - k;
            // rows: rho <- U rho on the row index
            for (int tk = tid; tk < (total >> k); tk += DM_THREADS) {
                const int y = tk & (W - 1);
                const int x0 = (int)ins0k((u64)(tk >> LW), op.srt, k);
                cd in[8], out[8];
#pragma unroll
                for (int a = 0; a < 8; ++a)
                    if (a < K) in[a] = lds[(x0 | (int)spread_bits(a, op.lbit, k)) * LDS_STRIDE + y];
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    out[a] = cmk(0.0, 0.0);
#pragma unroll
                    for (int b = 0; b < 8; ++b)
                        if (a < K && b < K) cfma(out[a], op.m[a * K + b], in[b]);
                }
#pragma unroll
                for (int a = 0; a < 8; ++a)
                    if (a < K) lds[(x0 | (int)spread_bits(a, op.lbit, k)) * LDS_STRIDE + y] = out[a];
            }
            __syncthreads();
            // columns: rho <- rho U^dagger on the column index
            for (int tk = tid; tk < (total >> k); tk += DM_THREADS) {
                const int x = tk >> LWk;
                const int y0 = (int)ins0k((u64)(tk & ((1 << LWk) - 1)), op.srt, k);
                cd in[8], out[8];
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    if (b < K) in[b] = lds[x * LDS_STRIDE + (y0 | (int)spread_bits(b, op.lbit, k))];
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    out[a] = cmk(0.0, 0.0);
#pragma unroll
                    for (int b = 0; b < 8; ++b)
                        if (a < K && b < K) cfma(out[a], in[b], cconj(op.m[a * K + b]));
                }
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    if (b < K) lds[x * LDS_STRIDE + (y0 | (int)spread_bits(b, op.lbit, k))] = out[b];
            }
            __syncthreads();
        }

        // ---------------- store ----------------
        for (int t = tid; t < total; t += DM_THREADS) {
            const int ti = t >> (2 * LB), e = t & (B * B - 1);
            const int i = ti >> g, j = ti & (NT - 1);
            const u64 TR = TRb | (g ? gsp[i] : 0), TC = TCb | (g ? gsp[j] : 0);
            const int sa = e >> LB, sb = e & (B - 1);
            const int sr = (LAYOUT == ORK_TILED) ? sa : sb;
            const int sc = (LAYOUT == ORK_TILED) ? sb : sa;
            const bool direct = tile_stored<LAYOUT>(p, TR, TC) && !(LAYOUT == ORK_PACKED && TR == TC && sr < sc) &&
                                moved<OP>(op, mt, i * B + sr, j * B + sc);
            const bool mirror = !diag_group && tile_stored<LAYOUT>(p, TC, TR) && moved<OP>(op, mt, i * B + sc, j * B + sr);
            if (direct) stx<NTA>(rho + elem_off<LAYOUT>(p, TR * B + sr, TC * B + sc), lds_get<OP>(lds, mt, i * B + sr, j * B + sc));
            if (mirror) stx<NTA>(rho + elem_off<LAYOUT>(p, TC * B + sr, TR * B + sc), cconj(lds_get<OP>(lds, mt, i * B + sc, j * B + sr)));
        }
        __syncthreads();
    }
}

// ---- tiled layout, 1-qubit superoperator, streamed ---------------------------------------
// The quad (r0, r1) x (c0, c1), r1 = r0 | 2^q, c1 = c0 | 2^q, goes from HBM to
// registers and back with no LDS round trip wherever its four entries are
// stored as themselves:
//   q <  L: the quad lies inside one stored tile (diagonal tiles hold both
//           triangles), so the tile array is one flat stream of quads;
//   q >= L: the quad takes the same in-tile entry of the four tiles
//           (TR|a, TC|c), a, c in {0, b}, b = 2^(q-L); one workgroup per
//           canonical tile quad (TR, TC with bit q-L clear, TR >= TC).  When TR
//           and TC agree above that bit, (TR, TC|b) lies above the diagonal: it
//           is the conjugate transpose of the stored tile X = (TC|b, TR), staged
//           through LDS so that its reads and writes stay coalesced.  For
//           TR == TC, X is the quad's own (TR|b, TC) tile: every quad reads X from
//           the LDS copy and writes only its direct (r1, c0) entry, the mirror of
//           the (r0, c1) entry another quad of the workgroup computes.
struct StreamPlan {
    u64 nitems;   // quads (q < L) or canonical tile quads (q >= L)
    u64 nblocks;  // workgroups of the launch
    int L, m;     // tile log; in-tile bit q (q < L) or tile-index bit q - L
    int xcd;
};

__device__ __forceinline__ void sop4(const DmOp &op, cd v0, cd v1, cd v2, cd v3, cd &o0, cd &o1, cd &o2, cd &o3) {
    o0 = cmk(0.0, 0.0); o1 = o0; o2 = o0; o3 = o0;
    cfma(o0, op.m[0], v0); cfma(o0, op.m[1], v1); cfma(o0, op.m[2], v2); cfma(o0, op.m[3], v3);
    cfma(o1, op.m[4], v0); cfma(o1, op.m[5], v1); cfma(o1, op.m[6], v2); cfma(o1, op.m[7], v3);
    cfma(o2, op.m[8], v0); cfma(o2, op.m[9], v1); cfma(o2, op.m[10], v2); cfma(o2, op.m[11], v3);
    cfma(o3, op.m[12], v0); cfma(o3, op.m[13], v1); cfma(o3, op.m[14], v2); cfma(o3, op.m[15], v3);
}

__device__ __forceinline__ u64 tile_base(u64 TR, u64 TC, int L) { return (TR * (TR + 1) / 2 + TC) << (2 * L); }

// the quad (r0, r0 | 2^q) x (c0, c0 | 2^q) through the superoperator, or through the controlled
// permutation A V B^dagger (A = U when r0 has every control bit, else I; B likewise for c0), which
// moves entries exactly (multiplications by i^k only)
__device__ __forceinline__ void apply4(const DmOp &op, u64 r0, u64 c0, cd v0, cd v1, cd v2, cd v3, cd &o0, cd &o1,
                                       cd &o2, cd &o3) {
    if (!op.ctl) {
        sop4(op, v0, v1, v2, v3, o0, o1, o2, o3);
        return;
    }
    cd w0 = v0, w1 = v1, w2 = v2, w3 = v3;
    if ((r0 & op.cmask) == op.cmask) {   // rows: (U V)[0] = u01 V[1], (U V)[1] = u10 V[0]
        w0 = cipow(v2, op.pu01);
        w1 = cipow(v3, op.pu01);
        w2 = cipow(v0, op.pu10);
        w3 = cipow(v1, op.pu10);
    }
    if ((c0 & op.cmask) == op.cmask) {   // columns: (M U^dagger)[., 0] = M[., 1] conj u01, [., 1] = M[., 0] conj u10
        const cd a0 = w0, a2 = w2;
        w0 = cipow(w1, -op.pu01);
        w1 = cipow(a0, -op.pu10);
        w2 = cipow(w3, -op.pu01);
        w3 = cipow(a2, -op.pu10);
    }
    o0 = w0; o1 = w1; o2 = w2; o3 = w3;
}

// the quad keeps its value under the controlled permutation, and the operation leaves such quads alone;
// the streamed kernels take SKIP as a template argument, so that their superoperator forms carry no
// trace of the test (SKIP is set exactly when op.skip is)
template <int SKIP>
__device__ __forceinline__ bool quad_still(const DmOp &op, u64 r0, u64 c0) {
    if constexpr (!SKIP) return false;
    else return (r0 & op.cmask) != op.cmask && (c0 & op.cmask) != op.cmask;
}

template <typename S, int U, int NT, int SKIP>
__global__ __launch_bounds__(DM_THREADS) void dm_sop_intile_kernel(S *__restrict__ rho, StreamPlan sp, DmOp op) {
    const int L = sp.L, q = sp.m, hl = L - 1;
    const int mb = 1 << q;
    const u64 w = remap_group(blockIdx.x, sp.nblocks, sp.xcd);
    S *t[U];
    int a00[U];
    cd v[U][4];
    S raw[U][4];   // complex64: loaded without a condition, converted after the batch
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 Q = (w * U + u) * DM_THREADS + threadIdx.x;
        ok[u] = Q < sp.nitems;
        const int i = (int)(Q & (((u64)1 << (2 * hl)) - 1));
        const int r0 = (int)ins0((u64)(i >> hl), q), c0 = (int)ins0((u64)(i & ((1 << hl) - 1)), q);
        t[u] = ok[u] ? rho + ((Q >> (2 * hl)) << (2 * L)) : rho;
        a00[u] = (r0 << L) | c0;
        if (SKIP && ok[u]) {
            u64 TR, TC;
            tri_decode(Q >> (2 * hl), TR, TC);
            ok[u] = !quad_still<SKIP>(op, (TR << L) | (u64)r0, (TC << L) | (u64)c0);
        }
        if constexpr (sizeof(S) == 8) {
            // a quad without work reads rho[0]: a complex64 load inside the condition waited for its own
            // conversion there, one quad in flight per thread
            S *tp = ok[u] ? t[u] : rho;
            const int aa = ok[u] ? a00[u] : 0, mm = ok[u] ? mb : 0;
            raw[u][0] = ldr<NT>(tp + aa);
            raw[u][1] = ldr<NT>(tp + (aa | mm));
            raw[u][2] = ldr<NT>(tp + (aa | (mm << L)));
            raw[u][3] = ldr<NT>(tp + (aa | (mm << L) | mm));
        } else if (ok[u]) {
            v[u][0] = ldx<NT>(t[u] + a00[u]);
            v[u][1] = ldx<NT>(t[u] + (a00[u] | mb));
            v[u][2] = ldx<NT>(t[u] + (a00[u] | (mb << L)));
            v[u][3] = ldx<NT>(t[u] + (a00[u] | (mb << L) | mb));
        }
    }
    if constexpr (sizeof(S) == 8) {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) hold(raw[u][j]);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[u][j] = to_cd(raw[u][j]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        cd o0, o1, o2, o3;
        u64 r0 = 0, c0 = 0;
        if (op.ctl) {   // full row and column of the quad's base entry
            const u64 Q = (w * U + u) * DM_THREADS + threadIdx.x;
            u64 TR, TC;
            tri_decode(Q >> (2 * hl), TR, TC);
            r0 = (TR << L) | (u64)(a00[u] >> L);
            c0 = (TC << L) | (u64)(a00[u] & ((1 << L) - 1));
        }
        apply4(op, r0, c0, v[u][0], v[u][1], v[u][2], v[u][3], o0, o1, o2, o3);
        stx<NT>(t[u] + a00[u], o0);
        stx<NT>(t[u] + (a00[u] | mb), o1);
        stx<NT>(t[u] + (a00[u] | (mb << L)), o2);
        stx<NT>(t[u] + (a00[u] | (mb << L) | mb), o3);
    }
}

// PARTS > 1 (32 x 32 tiles only): PARTS consecutive workgroups share one tile quad, each taking
// 32 / PARTS rows of A, C, D (and of B, or the matching columns of its stored mirror); a diagonal
// tile quad (TR == TC) stays whole in its first workgroup, since there the mirror is C itself.

// In-tile targets whose quad columns c0, c0 | 2^q share 128 B lines: one thread per (row pair, column),
// lanes walking the columns of a row so that every load and store instruction covers whole lines;
// the partner column's two entries come from lane ^ 2^q.  Both lanes of a pair form the same four
// products as the quad kernel above and each stores its own column.
template <typename S, int U, int NT, int SKIP>
__global__ __launch_bounds__(DM_THREADS) void dm_sop_intile_lane_kernel(S *__restrict__ rho, StreamPlan sp, DmOp op) {
    const int L = sp.L, q = sp.m, T = 1 << L, mb = 1 << q;
    const u64 w = remap_group(blockIdx.x, sp.nblocks, sp.xcd);
    const int per_tile_log = 2 * L - 1;   // T / 2 row pairs x T columns per tile
    S *pa[U], *pb[U];
    cd a[U], b[U];
    S ra[U], rb[U];   // complex64: loaded without a condition, converted after the batch
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 Q = (w * U + u) * DM_THREADS + threadIdx.x;
        ok[u] = Q < sp.nitems;
        const int within = (int)(Q & (((u64)1 << per_tile_log) - 1));
        const int r0 = (int)ins0((u64)(within >> L), q), c = within & (T - 1);
        S *tile = ok[u] ? rho + ((Q >> per_tile_log) << (2 * L)) : rho;
        pa[u] = tile + ((r0 << L) | c);
        pb[u] = tile + (((r0 | mb) << L) | c);
        a[u] = cmk(0.0, 0.0);
        b[u] = a[u];
        if (SKIP && ok[u]) {   // both lanes of a pair decide alike: their columns differ in bit q only
            u64 TR, TC;
            tri_decode(Q >> per_tile_log, TR, TC);
            ok[u] = !quad_still<SKIP>(op, (TR << L) | (u64)r0, (TC << L) | (u64)(c & ~mb));
        }
        if constexpr (sizeof(S) == 8) {   // a pair without work reads rho[0] (as in dm_sop_intile_kernel)
            ra[u] = ldr<NT>(ok[u] ? pa[u] : rho);
            rb[u] = ldr<NT>(ok[u] ? pb[u] : rho);
        } else if (ok[u]) {
            a[u] = ldx<NT>(pa[u]);
            b[u] = ldx<NT>(pb[u]);
        }
    }
    if constexpr (sizeof(S) == 8) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            hold(ra[u]);
            hold(rb[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[u] = to_cd(ra[u]);
            b[u] = to_cd(rb[u]);
        }
    }
    // bit q of the column is bit q of the lane (blocks start on multiples of 256 items)
    const bool lo = !(threadIdx.x & mb);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const cd a2 = shfl_xor_cd(a[u], mb), b2 = shfl_xor_cd(b[u], mb);   // every lane takes part
        if (!ok[u]) continue;
        // component-wise selects (a select of the vector type itself goes through memory)
        const cd v0 = sel(lo, a[u], a2), v1 = sel(lo, a2, a[u]), v2 = sel(lo, b[u], b2), v3 = sel(lo, b2, b[u]);
        cd o0, o1, o2, o3;
        u64 r0 = 0, c0 = 0;
        if (op.ctl) {   // full row and column of the quad's base entry (column bit q cleared)
            const u64 Q = (w * U + u) * DM_THREADS + threadIdx.x;
            const int within = (int)(Q & (((u64)1 << per_tile_log) - 1));
            u64 TR, TC;
            tri_decode(Q >> per_tile_log, TR, TC);
            r0 = (TR << L) | ins0((u64)(within >> L), q);
            c0 = (TC << L) | (u64)((within & (T - 1)) & ~mb);
        }
        apply4(op, r0, c0, v0, v1, v2, v3, o0, o1, o2, o3);
        stx<NT>(pa[u], sel(lo, o0, o1));
        stx<NT>(pb[u], sel(lo, o2, o3));
    }
}

template <typename S, int NT, int PARTS, int SKIP>
__global__ __launch_bounds__(DM_THREADS) void dm_sop_tiles_kernel(S *__restrict__ rho, StreamPlan sp, DmOp op) {
    __shared__ cd xs[W_MAX * LDS_STRIDE];
    const int L = sp.L, T = 1 << L, TT = T * T, tid = threadIdx.x;
    const u64 g = remap_group(blockIdx.x, sp.nblocks, sp.xcd);
    const u64 w = g / PARTS;
    const int part = (int)(g % PARTS);
    if (w >= sp.nitems) return;
    u64 Rr, Cr;
    tri_decode(w, Rr, Cr);
    const u64 bb = (u64)1 << sp.m;
    const u64 TR = ins0(Rr, sp.m), TC = ins0(Cr, sp.m);
    // does the quad at in-tile (i, j) of the tile quad change?  (always, but for skipped controlled quads)
    auto live = [&](int i, int j) { return !quad_still<SKIP>(op, (TR << L) | (u64)i, (TC << L) | (u64)j); };
    S *A = rho + tile_base(TR, TC, L);
    S *C = rho + tile_base(TR | bb, TC, L);
    S *D = rho + tile_base(TR | bb, TC | bb, L);
    if ((TR >> (sp.m + 1)) != (TC >> (sp.m + 1))) {
        // all four tiles stored as themselves: with 32 x 32 tiles every thread loads its four
        // in-tile entries of all four tiles at once (64 KB of a workgroup in flight)
        S *B = rho + tile_base(TR, TC | bb, L);
        if (TT == 4 * DM_THREADS) {
            constexpr int E = 4 / PARTS;
            const int o0 = part * (4 * DM_THREADS / PARTS);
            cd va[E], vb[E], vc[E], vd[E];
            bool lv[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int o = o0 + tid + e * DM_THREADS;
                lv[e] = live(o >> L, o & (T - 1));
                if (lv[e]) { va[e] = ldx<NT>(A + o); vb[e] = ldx<NT>(B + o); vc[e] = ldx<NT>(C + o); vd[e] = ldx<NT>(D + o); }
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (!lv[e]) continue;
                const int o = o0 + tid + e * DM_THREADS;
                cd p0, p1, p2, p3;
                apply4(op, (TR << L) | (u64)(o >> L), (TC << L) | (u64)(o & (T - 1)), va[e], vb[e], vc[e], vd[e], p0, p1,
                       p2, p3);
                stx<NT>(A + o, p0); stx<NT>(B + o, p1); stx<NT>(C + o, p2); stx<NT>(D + o, p3);
            }
            return;
        }
        for (int o = tid; o < TT; o += 2 * DM_THREADS) {
            const int o2 = o + DM_THREADS;
            const bool k0 = live(o >> L, o & (T - 1)), k1 = o2 < TT && live(o2 >> L, o2 & (T - 1));
            cd a0 = cmk(0.0, 0.0), b0 = a0, c0 = a0, d0 = a0, a1 = a0, b1 = a0, c1 = a0, d1 = a0;
            if (k0) { a0 = ldx<NT>(A + o); b0 = ldx<NT>(B + o); c0 = ldx<NT>(C + o); d0 = ldx<NT>(D + o); }
            if (k1) { a1 = ldx<NT>(A + o2); b1 = ldx<NT>(B + o2); c1 = ldx<NT>(C + o2); d1 = ldx<NT>(D + o2); }
            cd p0, p1, p2, p3;
            if (k0) {
                apply4(op, (TR << L) | (u64)(o >> L), (TC << L) | (u64)(o & (T - 1)), a0, b0, c0, d0, p0, p1, p2, p3);
                stx<NT>(A + o, p0); stx<NT>(B + o, p1); stx<NT>(C + o, p2); stx<NT>(D + o, p3);
            }
            if (k1) {
                apply4(op, (TR << L) | (u64)(o2 >> L), (TC << L) | (u64)(o2 & (T - 1)), a1, b1, c1, d1, p0, p1, p2, p3);
                stx<NT>(A + o2, p0); stx<NT>(B + o2, p1); stx<NT>(C + o2, p2); stx<NT>(D + o2, p3);
            }
        }
        return;
    }
    // X: the stored mirror of B = (TR, TC | bb); its entry (j, i) is the conjugate of B's (i, j), which
    // belongs to the quad at (i, j)
    const bool diag = (TR == TC);
    S *X = diag ? C : rho + tile_base(TC | bb, TR, L);
    if (PARTS > 1 && TT == 4 * DM_THREADS && !diag) {
        // rows [i0, i0 + R) of A, C, D; the matching columns [i0, i0 + R) of the stored mirror X
        constexpr int E = 4 / PARTS, R = 32 / PARTS;
        const int i0 = part * R;
        cd va[E], vc[E], vd[E], vx[E];
        bool lq[E], lx[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int o = i0 * 32 + tid + e * DM_THREADS;
            const int idx = tid + e * DM_THREADS, xj = idx / R, xi = i0 + idx % R;
            lq[e] = live(o >> 5, o & 31);
            lx[e] = live(xi, xj);
            if (lq[e]) { va[e] = ldx<NT>(A + o); vc[e] = ldx<NT>(C + o); vd[e] = ldx<NT>(D + o); }
            if (lx[e]) vx[e] = ldx<NT>(X + xj * 32 + xi);
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int idx = tid + e * DM_THREADS, xj = idx / R, xi = i0 + idx % R;
            if (lx[e]) xs[xj * LDS_STRIDE + xi] = vx[e];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (!lq[e]) continue;
            const int o = i0 * 32 + tid + e * DM_THREADS, i = o >> 5, j = o & 31;
            // (j, i) of the LDS copy is read and then rewritten by this thread alone
            const cd v1 = cconj(xs[j * LDS_STRIDE + i]);
            cd o0, o1, o2, o3;
            apply4(op, (TR << L) | (u64)i, (TC << L) | (u64)j, va[e], v1, vc[e], vd[e], o0, o1, o2, o3);
            stx<NT>(A + o, o0);
            stx<NT>(C + o, o2);
            stx<NT>(D + o, o3);
            xs[j * LDS_STRIDE + i] = cconj(o1);
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int idx = tid + e * DM_THREADS, xj = idx / R, xi = i0 + idx % R;
            if (lx[e]) stx<NT>(X + xj * 32 + xi, xs[xj * LDS_STRIDE + xi]);
        }
        return;
    }
    if (part != 0) return;   // a diagonal tile quad is handled whole by its first workgroup
    if (TT == 4 * DM_THREADS) {
        // 32 x 32 tiles: every load of the workgroup (A, D, X and, off the diagonal, C) in flight at
        // once.  On the diagonal X is C itself, so a thread's own C entry comes from its registers
        // and only the transposed one from LDS (whether a quad changes is then symmetric in (i, j)).
        cd va[4], vc[4], vd[4], vx[4];
        bool lq[4], lx[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int o = tid + e * DM_THREADS, i = o >> L, j = o & (T - 1);
            lq[e] = live(i, j);
            lx[e] = live(j, i);   // X's (i, j) belongs to the quad at (j, i)
            if (lq[e]) {
                va[e] = ldx<NT>(A + o); vd[e] = ldx<NT>(D + o);
                if (!diag) vc[e] = ldx<NT>(C + o);
            }
            if (lx[e]) vx[e] = ldx<NT>(X + o);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int o = tid + e * DM_THREADS;
            if (lx[e]) xs[(o >> L) * LDS_STRIDE + (o & (T - 1))] = vx[e];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (!lq[e]) continue;
            const int o = tid + e * DM_THREADS, i = o >> L, j = o & (T - 1);
            // (j, i) of the LDS copy is read and then rewritten by this thread alone
            const cd v1 = cconj(xs[j * LDS_STRIDE + i]);
            const cd v2 = diag ? vx[e] : vc[e];
            cd o0, o1, o2, o3;
            apply4(op, (TR << L) | (u64)i, (TC << L) | (u64)j, va[e], v1, v2, vd[e], o0, o1, o2, o3);
            stx<NT>(A + o, o0);
            stx<NT>(C + o, o2);
            stx<NT>(D + o, o3);
            if (!diag) xs[j * LDS_STRIDE + i] = cconj(o1);
        }
        if (!diag) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int o = tid + e * DM_THREADS;
                if (lx[e]) stx<NT>(X + o, xs[(o >> L) * LDS_STRIDE + (o & (T - 1))]);
            }
        }
        return;
    }
    for (int o = tid; o < TT; o += DM_THREADS)
        if (live(o & (T - 1), o >> L)) xs[(o >> L) * LDS_STRIDE + (o & (T - 1))] = ldx<NT>(X + o);
    __syncthreads();
    for (int o = tid; o < TT; o += DM_THREADS) {
        const int i = o >> L, j = o & (T - 1);
        if (!live(i, j)) continue;
        const cd v0 = ldx<NT>(A + o), v3 = ldx<NT>(D + o);
        const cd v1 = cconj(xs[j * LDS_STRIDE + i]);
        const cd v2 = diag ? xs[i * LDS_STRIDE + j] : ldx<NT>(C + o);
        cd o0, o1, o2, o3;
        apply4(op, (TR << L) | (u64)i, (TC << L) | (u64)j, v0, v1, v2, v3, o0, o1, o2, o3);
        stx<NT>(A + o, o0);
        stx<NT>(C + o, o2);
        stx<NT>(D + o, o3);
        if (!diag) xs[j * LDS_STRIDE + i] = cconj(o1);
    }
    if (!diag) {
        __syncthreads();
        for (int o = tid; o < TT; o += DM_THREADS)
            if (live(o & (T - 1), o >> L)) stx<NT>(X + o, xs[(o >> L) * LDS_STRIDE + (o & (T - 1))]);
    }
}

// ---- packed layout, 1-qubit superoperator, streamed --------------------------------------
// Quad bases (r0, c0) = (ins0(a, q), ins0(b, q)).  A quad is canonical when a >= b
// (r0 >= c0); it owns its stored entries and, when (r0, c1) lies above the
// diagonal (a and b agree above bit q, a != b), the mirror entry (c1, r0), which
// belongs to no canonical quad.  For a diagonal quad (a == b) the mirror of
// (r0, c1) is the quad's own (r1, c0).  One workgroup owns a 32 x 32 block (A, B),
// A >= B, of (a, b); lanes walk a, i.e. consecutive rows of one column of the
// column-major layout.  A block's mirror entries lie in 32 columns r0 and 32 rows
// c1 and are staged through LDS so that their reads and writes also walk rows.
struct PackedStreamPlan {
    u64 dim, nblocks, nt;   // nt: block rows (dim / 64)
    int q, xcd, col;
};

__device__ __forceinline__ u64 col_base(u64 D, u64 c) { return c * (2 * D - c + 1) / 2; }   // offset of (c, c)

// Is the 128 B line of stored entry (r, c) entirely inside the rows the workgroup of block row A owns in
// column c?  Those rows are ins0(32 A .. 32 A + 31, q) and the same | 2^q: two runs of 32 rows for
// q >= 5, one run of 64 for q < 5, cut at the diagonal.  A line that crosses a run end is shared
// with another workgroup (the neighbouring block, or the previous column's last rows).
template <typename S>
__device__ __forceinline__ bool packed_line_owned(u64 D, int q, u64 A, u64 r, u64 c) {
    constexpr int LG = sizeof(S) == 8 ? 4 : 3;   // log2 of the entries per line
    u64 rs, re;
    if (q >= 5) {
        const u64 R1 = ins0(A * 32, q);
        rs = r >= R1 + 32 ? R1 + ((u64)1 << q) : R1;
        re = rs + 32;
    } else {
        rs = A * 64;
        re = rs + 64;
    }
    if (rs < c) rs = c;
    const u64 base = col_base(D, c) - c;   // offset of (r, c) is base + r
    const u64 l0 = ((base + r) >> LG) << LG;
    return l0 >= base + rs && l0 + ((u64)1 << LG) <= base + re;
}

// NT 2: nontemporal accesses on the lines a workgroup owns alone, cached ones on lines it shares (the
// ends of its column runs) and on the mirror entries, so that the shared lines meet in L2 instead of
// going to HBM twice.
template <int NT, typename S> __device__ __forceinline__ cd ldh(const S *p, bool own) {
    if constexpr (NT == 2) return own ? ldx<1>(p) : ldx<0>(p);
    else return ldx<NT>(p);
}
template <int NT, typename S> __device__ __forceinline__ S ldhr(const S *p, bool own) {
    if constexpr (NT == 2) return own ? ldr<1>(p) : ldr<0>(p);
    else return ldr<NT>(p);
}
template <int NT, typename S> __device__ __forceinline__ void sth(S *p, cd v, bool own) {
    if constexpr (NT == 2) {
        if (own) stx<1>(p, v);
        else stx<0>(p, v);
    } else {
        stx<NT>(p, v);
    }
}

// PARTS consecutive workgroups (after the XCD remap) share one 32 x 32 block, each taking 32 / PARTS
// of its columns b (their quads and their mirror entries): no entry is touched by two of them.
template <typename S, int NT, int PARTS, int SKIP>
__global__ __launch_bounds__(DM_THREADS) void dm_sop_packed_kernel(S *__restrict__ rho, PackedStreamPlan pp, DmOp op) {
    __shared__ cd ms[32 * LDS_STRIDE];
    const int q = pp.q, tid = threadIdx.x;
    const u64 m = (u64)1 << q, D = pp.dim;
    const u64 g = remap_group(blockIdx.x, pp.nblocks * PARTS, pp.xcd);
    const int part = (int)(g % PARTS);
    u64 A, Bk;
    block_rc(g / PARTS, pp.nblocks, pp.nt, 0, pp.col, A, Bk);
    const bool has_mirror = q >= 5 ? ((A >> (q - 5)) == (Bk >> (q - 5))) : (A == Bk);
    constexpr int QPT = 1024 / DM_THREADS / PARTS, YW = 32 / PARTS;
    const int y0 = part * YW;
    // quad (a, b) holds a mirror entry (c1, r0), and changes it
    auto mirror = [&](u64 a, u64 b) {
        return a > b && (a >> q) == (b >> q) && !quad_still<SKIP>(op, ins0(a, q), ins0(b, q));
    };
    // the block's mirror entries, loaded together with the quads below and staged through LDS
    // once every load is in flight
    cd mv[QPT];
    if (has_mirror) {
#pragma unroll
        for (int h = 0; h < QPT; ++h) {
            const int t = tid + DM_THREADS * h, x = t / YW, y = y0 + t % YW;   // lanes walk y: rows c1 of column r0(x)
            const u64 a = A * 32 + x, b = Bk * 32 + y;
            if (mirror(a, b)) mv[h] = ldx<NT == 1>(rho + col_base(D, ins0(a, q)) + ((ins0(b, q) | m) - ins0(a, q)));
        }
    }
    // all quads of a thread in flight: t = tid + 256 h.  Complex64 values are loaded raw for every quad
    // (a quad left alone reads rho[0]) and converted after the batch: a load inside the quad's condition
    // had the compiler wait for its conversion there, one quad's loads in flight per thread
    {
        cd v[QPT][4];
        S raw[QPT][4];
        S *p0[QPT], *p1[QPT], *p2[QPT], *p3[QPT];
        int kind[QPT];   // 0 skip, 1 direct, 2 mirror, 3 diagonal quad
        int own[QPT];    // bit j: the line of p_j is the workgroup's alone (NT 2)
#pragma unroll
        for (int h = 0; h < QPT; ++h) {
            const int tt = tid + DM_THREADS * h, x = tt & 31, y = y0 + (tt >> 5);   // lanes walk x: rows r0 of column c0(y)
            const u64 a = A * 32 + x, b = Bk * 32 + y;
            kind[h] = a < b ? 0 : a == b ? 3 : ((a >> q) == (b >> q) ? 2 : 1);
            if (kind[h] && quad_still<SKIP>(op, ins0(a, q), ins0(b, q))) kind[h] = 0;
            p0[h] = p1[h] = p2[h] = p3[h] = rho;
            own[h] = 0;
            if (kind[h]) {
                const u64 r0 = ins0(a, q), c0 = ins0(b, q), r1 = r0 | m, c1 = c0 | m;
                const u64 k0 = col_base(D, c0), k1 = col_base(D, c1);
                p0[h] = rho + k0 + (r0 - c0);
                p2[h] = rho + k0 + (r1 - c0);
                p3[h] = rho + k1 + (r1 - c1);
                if (kind[h] == 1) p1[h] = rho + k1 + (r0 - c1);   // (r0, c1) is stored only here
                if constexpr (NT == 2) {
                    own[h] = (int)packed_line_owned<S>(D, q, A, r0, c0) | (int)packed_line_owned<S>(D, q, A, r1, c0) << 2 |
                             (int)packed_line_owned<S>(D, q, A, r1, c1) << 3;
                    if (kind[h] == 1) own[h] |= (int)packed_line_owned<S>(D, q, A, r0, c1) << 1;
                }
            }
            if constexpr (sizeof(S) != 8) {   // complex128: loaded only where used, no conversion to wait for
#pragma unroll
                for (int j = 0; j < 4; ++j) raw[h][j].x = raw[h][j].y = 0.0;
            }
            if (sizeof(S) == 8 || kind[h]) {
                raw[h][0] = ldhr<NT>(p0[h], own[h] & 1);
                raw[h][2] = ldhr<NT>(p2[h], (own[h] >> 2) & 1);
                raw[h][3] = ldhr<NT>(p3[h], (own[h] >> 3) & 1);
                if (sizeof(S) == 8 || kind[h] == 1) raw[h][1] = ldhr<NT>(p1[h], (own[h] >> 1) & 1);
            }
        }
        if constexpr (sizeof(S) == 8) {
#pragma unroll
            for (int h = 0; h < QPT; ++h)
#pragma unroll
                for (int j = 0; j < 4; ++j) hold(raw[h][j]);
        }
#pragma unroll
        for (int h = 0; h < QPT; ++h)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[h][j] = to_cd(raw[h][j]);
        if (has_mirror) {
#pragma unroll
            for (int h = 0; h < QPT; ++h) {
                const int t = tid + DM_THREADS * h, x = t / YW, y = y0 + t % YW;
                const u64 a = A * 32 + x, b = Bk * 32 + y;
                if (mirror(a, b)) ms[x * LDS_STRIDE + y] = mv[h];
            }
            __syncthreads();
        }
#pragma unroll
        for (int h = 0; h < QPT; ++h) {
            if (!kind[h]) continue;
            const int tt = tid + DM_THREADS * h, x = tt & 31, y = y0 + (tt >> 5);
            if (kind[h] == 2) v[h][1] = cconj(ms[x * LDS_STRIDE + y]);
            else if (kind[h] == 3) v[h][1] = cconj(v[h][2]);
            cd o0, o1, o2, o3;
            apply4(op, ins0(A * 32 + (u64)x, q), ins0(Bk * 32 + (u64)y, q), v[h][0], v[h][1], v[h][2], v[h][3], o0, o1,
                   o2, o3);
            sth<NT>(p0[h], o0, own[h] & 1);
            sth<NT>(p2[h], o2, (own[h] >> 2) & 1);
            sth<NT>(p3[h], o3, (own[h] >> 3) & 1);
            if (kind[h] == 1) sth<NT>(p1[h], o1, (own[h] >> 1) & 1);
            else if (kind[h] == 2) ms[x * LDS_STRIDE + y] = cconj(o1);
        }
    }
    if (has_mirror) {
        __syncthreads();
        for (int t = tid; t < 1024 / PARTS; t += DM_THREADS) {
            const int x = t / YW, y = y0 + t % YW;
            const u64 a = A * 32 + x, b = Bk * 32 + y;
            if (mirror(a, b))
                stx<NT == 1>(rho + col_base(D, ins0(a, q)) + ((ins0(b, q) | m) - ins0(a, q)), ms[x * LDS_STRIDE + y]);
        }
    }
}

// ---- permutations of whole tiles: every gate qubit at or above the tile -----------------------
// When every qubit of a permutation gate lies at or above the tile (tiled: the storage tile, q >= L;
// packed: 32 x 32 blocks, q >= 5), the gate maps whole tiles onto whole tiles: the new tile (TR, TC)
// is i^(ph[pinv a] - ph[pinv b]) times the old logical tile (P TR, P TC), with a, b the gate bits of
// TR, TC and P the permutation of tile indices; in-tile positions stay.  For the involutions of the
// gate set (cx, cy, swap, ccx) the stored tiles pair up: s and its source p = fold(P TR, P TC), the
// stored tile holding the source (its conjugate transpose when (P TR, P TC) lies above the diagonal).
// One workgroup per stored tile s; it works only when s is not after p in storage order, then swaps
// both tiles as two coalesced streams (through LDS for a transpose).  Diagonal tiles map onto diagonal
// tiles.  Tiles whose value does not change are neither read nor written.  (Emulated on the CPU against
// U rho U^dagger for every gate and qubit order above the tile, both layouts, every entry written at
// most once, before the first GPU run.)
struct PermTilePlan {
    u64 ntiles;   // stored tiles (tile rows nt: nt (nt + 1) / 2)
    u64 D;        // dimension (packed offsets)
    int L;        // tile log: storage tile (tiled), 5 (packed)
    int k;
    int pos[3];   // tile-index bit of gate argument j (q_j - L)
    int pinv[8];  // gate bits a take the tile whose gate bits are pinv[a] ...
    int ph[8];    // ... times i^(ph[pinv a] - ph[pinv b])
    int xcd;
    int tpw;      // consecutive tiles (slot sets of virtual tiles) per workgroup
    u64 nwg;      // workgroups of the launch
    int V;        // virtual tile log (dm_perm_subtiles_kernel; pos[] then counts from V)
    unsigned pinv_bits, ph_bits;   // pinv[a] at bits 3a..3a+2, ph[s] at bits 2s, 2s+1
};

__device__ __forceinline__ u64 perm_src_tile(const PermTilePlan &p, u64 t, int &a_out) {
    int a = 0;
    for (int j = 0; j < p.k; ++j) a |= (int)((t >> p.pos[j]) & 1) << j;
    const int s = p.pinv[a];
    for (int j = 0; j < p.k; ++j)
        if (((a ^ s) >> j) & 1) t ^= (u64)1 << p.pos[j];
    a_out = s;
    return t;
}

// entry o (memory order) of tile (TR, TC): its in-tile position (i, j), offset, and whether it is stored
template <int LAYOUT>
__device__ __forceinline__ void perm_tile_entry(const PermTilePlan &p, u64 TR, u64 TC, int o, int &i, int &j, u64 &off,
                                                bool &st) {
    if constexpr (LAYOUT == ORK_TILED) {
        i = o >> p.L;
        j = o & ((1 << p.L) - 1);
        off = tile_base(TR, TC, p.L) + (u64)o;
        st = true;
    } else {
        j = o >> 5;
        i = o & 31;
        const u64 r = TR * 32 + (u64)i, c = TC * 32 + (u64)j;
        st = r >= c;
        off = st ? col_base(p.D, c) + (r - c) : 0;
    }
}

// One workgroup per stored tile s; pairs whose source is a conjugate transpose go through one LDS tile
// used twice (stage p, write s; stage s, write p), the whole pair in one workgroup: split, one
// workgroup would read entries another had rewritten.  (A second launch for the transposed pairs alone
// spent most of its time dispatching the workgroups of the other tiles.)
#ifndef ORK_PERMT_COLUMNS
#define ORK_PERMT_COLUMNS 1
#endif
template <typename S, int LAYOUT, int NTA>
__global__ __launch_bounds__(DM_THREADS) void dm_perm_tiles_kernel(S *__restrict__ rho, PermTilePlan p) {
    __shared__ cd ts[W_MAX * LDS_STRIDE];
    constexpr int E = W_MAX * W_MAX / DM_THREADS;   // entries of each tile per thread
    const int T = LAYOUT == ORK_TILED ? (1 << p.L) : 32, TT = T * T, tid = threadIdx.x;
    // p.tpw consecutive tiles per workgroup: most tiles are left at once (a non-owner of its pair, or a
    // tile that keeps its value), which costs a loop iteration here instead of a workgroup dispatch
    const u64 g = remap_group(blockIdx.x, p.nwg, p.xcd);
    for (int tw = 0; tw < p.tpw; ++tw) {
        const u64 w = g * (u64)p.tpw + (u64)tw;
        if (w >= p.ntiles) break;
        u64 TR, TC;
        // the packed layout walks down block columns (a block's column runs share their end lines with the
        // blocks above and below); ownership follows the row-major index wr whatever the walk
        if (LAYOUT == ORK_PACKED && ORK_PERMT_COLUMNS) {
            u64 rr, cc;
            tri_decode(p.ntiles - 1 - w, rr, cc);
            const u64 nt = p.D >> p.L;
            TR = nt - 1 - cc;
            TC = nt - 1 - rr;
        } else {
            tri_decode(w, TR, TC);
        }
        const u64 wr = TR * (TR + 1) / 2 + TC;
        int sa, sb;
        const u64 PR = perm_src_tile(p, TR, sa), PC = perm_src_tile(p, TC, sb);
        const bool trans = PR < PC;
        const u64 pr = trans ? PC : PR, pc = trans ? PR : PC;
        const u64 wp = pr * (pr + 1) / 2 + pc;
        if (wp < wr) continue;   // the workgroup of p swaps both
        const int phs = (p.ph[sa] - p.ph[sb]) & 3;
        if (wp == wr && !trans && phs == 0) continue;   // the tile keeps its value
        int pa, pb;
        (void)perm_src_tile(p, pr, pa);
        (void)perm_src_tile(p, pc, pb);
        const int php = (p.ph[pa] - p.ph[pb]) & 3;
        const bool pair = wp != wr;
        cd vs[E], vp[E];
        S rs[E], rp[E];   // raw values: converted once every load of the tile is in flight
        u64 os[E], op[E];
        bool ok[E];
        int ii[E], jj[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int o = tid + e * DM_THREADS;
            bool st = false;
            int i = 0, j = 0;
            u64 off = 0, offp = 0;
            if (o < TT) perm_tile_entry<LAYOUT>(p, TR, TC, o, i, j, off, st);
            ok[e] = st;
            ii[e] = i;
            jj[e] = j;
            os[e] = off;
            if (st && pair) {
                bool stp;
                int i2, j2;
                perm_tile_entry<LAYOUT>(p, pr, pc, o, i2, j2, offp, stp);
            }
            op[e] = offp;
            // loads without a condition (a conditional complex64 load waits for its own conversion):
            // an entry that is not stored reads rho[0], a tile without a partner its own entry again
            rs[e] = ldr<NTA>(rho + off);
            rp[e] = ldr<NTA>(rho + ((st && pair) ? offp : off));
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            vs[e] = to_cd(rs[e]);
            vp[e] = to_cd(rp[e]);
        }
        if (!trans) {
            // same in-tile positions: s takes p's entries and p takes s's (or s its own, phase only)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (!ok[e]) continue;
                stx<NTA>(rho + os[e], cipow(sel(pair, vp[e], vs[e]), phs));
                if (pair) stx<NTA>(rho + op[e], cipow(vs[e], php));
            }
            continue;
        }
        // new s (i, j) = i^phs conj(old p (j, i)), new p (i, j) = i^php conj(old s (j, i)); off the
        // diagonal every entry of both tiles is stored, and every old value is already in registers
#pragma unroll
        for (int e = 0; e < E; ++e)
            if (ok[e]) ts[ii[e] * LDS_STRIDE + jj[e]] = sel(pair, vp[e], vs[e]);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e)
            if (ok[e]) stx<NTA>(rho + os[e], cipow(cconj(ts[jj[e] * LDS_STRIDE + ii[e]]), phs));
        __syncthreads();
        if (pair) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (ok[e]) ts[ii[e] * LDS_STRIDE + jj[e]] = vs[e];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (ok[e]) stx<NTA>(rho + op[e], cipow(cconj(ts[jj[e] * LDS_STRIDE + ii[e]]), php));
            __syncthreads();   // before the next tile's LDS writes
        }
    }
}

// ---- permutations of virtual tiles smaller than the storage tile -----------------------------------
// A gate whose lowest qubit V lies below the tile and whose highest at or above it (swap(4, 5) with
// 32 x 32 tiles) maps virtual tiles of 2^V x 2^V entries onto virtual tiles: the rule above with tile
// indices r >> V.  The rows of a virtual tile are runs of 2^V entries (V >= 3: whole 128 B lines for
// complex128, V >= 4 for complex64), where the tile-group kernel moved 8-entry runs.  On the tiled layout a
// virtual tile lies in the storage tile (TR >> d, TC >> d), d = L - V; the enumerated tiles are the lower
// virtual tiles (TR >= TC), every source is read from a lower one, and the upper virtual tiles of diagonal
// storage tiles (stored, never read here) are rewritten as the conjugate transpose of their new lower
// mirror by the owner of that mirror.  Each workgroup iteration takes 64 consecutive candidate tiles; its
// first wave lists those that own a pair or change alone (3/8 of them for swap), and slot sets of 1024
// entries walk the list: 2^(10 - 2V) virtual tiles per set, slot e of a thread holding entry
// (tid + 256 e) mod 2^2V of listed tile (tid + 256 e) >> 2V, so that every slot has work (walked tile by
// tile, 5/8 of the slots stood idle: 3.5 TB/s for swap(4, 5), 14 qubits, against 4.7 listed and 2.45 in
// the tile-group kernel).  (Emulated on the CPU against U rho U^dagger for every gate of the set, V from 1
// to the tile, both layouts, every qubit order: every position written at most once, no workgroup reading
// what another writes.)
__device__ __forceinline__ u64 sub_src_tile(const PermTilePlan &p, u64 t, int &s_out) {
    int a = (int)((t >> p.pos[0]) & 1);
    if (p.k > 1) a |= (int)((t >> p.pos[1]) & 1) << 1;
    if (p.k > 2) a |= (int)((t >> p.pos[2]) & 1) << 2;
    const int s = (int)((p.pinv_bits >> (3 * a)) & 7), x = a ^ s;
    s_out = s;
    return t ^ ((u64)(x & 1) << p.pos[0]) ^ ((u64)((x >> 1) & 1) << p.pos[1]) ^ ((u64)((x >> 2) & 1) << p.pos[2]);
}
__device__ __forceinline__ int sub_ph(const PermTilePlan &p, int s) { return (int)((p.ph_bits >> (2 * s)) & 3); }

// entry o (memory order) of virtual tile (TR, TC): in-tile position, offset, stored
template <int LAYOUT>
__device__ __forceinline__ void sub_tile_entry(const PermTilePlan &p, u64 TR, u64 TC, int o, u64 &off, bool &st) {
    const int V = p.V;
    if constexpr (LAYOUT == ORK_TILED) {
        const int d = p.L - V;
        const u64 m = ((u64)1 << d) - 1;
        off = tile_base(TR >> d, TC >> d, p.L) + (((((TR & m) << V) + (u64)(o >> V))) << p.L) + ((TC & m) << V) +
              (u64)(o & ((1 << V) - 1));
        st = true;
    } else {
        const u64 r = (TR << V) + (u64)(o & ((1 << V) - 1)), c = (TC << V) + (u64)(o >> V);
        st = r >= c;
        off = st ? col_base(p.D, c) + (r - c) : 0;
    }
}

#ifndef ORK_SUB_COLUMNS
#define ORK_SUB_COLUMNS 1
#endif
// candidate c -> virtual tile (TR, TC): on the packed layout down the columns of virtual tiles (a virtual
// tile's column runs share their end lines with the tiles above and below), on the tiled layout along rows
template <int LAYOUT>
__device__ __forceinline__ void sub_tile_at(const PermTilePlan &p, u64 c, u64 &TR, u64 &TC) {
    if (LAYOUT == ORK_PACKED && ORK_SUB_COLUMNS) {
        u64 rr, cc;
        tri_decode(p.ntiles - 1 - c, rr, cc);
        const u64 nt = p.D >> p.V;
        TR = nt - 1 - cc;
        TC = nt - 1 - rr;
    } else {
        tri_decode(c, TR, TC);
    }
}

// does virtual tile (TR, TC) own its pair (or change alone)?  Ownership follows the row-major index
// w = TR (TR + 1) / 2 + TC whatever the walk; its partner (pr, pc) and the phase exponent ps of its source
__device__ __forceinline__ bool sub_owner(const PermTilePlan &p, u64 TR, u64 TC, u64 &pr, u64 &pc, bool &tr, int &ps) {
    const u64 w = TR * (TR + 1) / 2 + TC;
    int sa, sb;
    const u64 PR = sub_src_tile(p, TR, sa), PC = sub_src_tile(p, TC, sb);
    tr = PR < PC;
    pr = tr ? PC : PR;
    pc = tr ? PR : PC;
    const u64 wp = pr * (pr + 1) / 2 + pc;
    ps = (sub_ph(p, sa) - sub_ph(p, sb)) & 3;
    return wp >= w && (wp != w || tr || ps != 0);
}

template <typename S, int LAYOUT, int NTA>
__global__ __launch_bounds__(DM_THREADS) void dm_perm_subtiles_kernel(S *__restrict__ rho, PermTilePlan p) {
    __shared__ cd ts[W_MAX * LDS_STRIDE];
    // listed tiles of an iteration (offsets from its first candidate), two buffers: the first wave writes the
    // next iteration's list while the other waves may still read this one's count
    __shared__ int act[2][64];
    __shared__ int nact[2];
    constexpr int E = W_MAX * W_MAX / DM_THREADS;
    const int V = p.V, T = 1 << V, tid = threadIdx.x;
    const int lsub = LW_MAX - V;          // virtual tiles per LDS row of tiles: 2^lsub
    const int d = LAYOUT == ORK_TILED ? p.L - V : 0;
    const u64 dm = ((u64)1 << d) - 1;
    // offset from an entry of a lower virtual tile of a diagonal storage tile to the same in-tile position of
    // the upper one, per unit of (TC & dm) - (TR & dm)
    const long long mstep = LAYOUT == ORK_TILED ? ((1ll << (V + p.L)) - (1ll << V)) : 0;
    const u64 g = remap_group(blockIdx.x, p.nwg, p.xcd);
    for (int tw = 0; tw < p.tpw; ++tw) {
        const u64 base = (g * (u64)p.tpw + (u64)tw) * 64;
        if (base >= p.ntiles) break;
        if (tid < 64) {   // the first wave lists the tiles with work
            const u64 c = base + (u64)tid;
            bool a = false;
            if (c < p.ntiles) {
                u64 TR, TC, pr, pc;
                bool tr;
                int ps;
                sub_tile_at<LAYOUT>(p, c, TR, TC);
                a = sub_owner(p, TR, TC, pr, pc, tr, ps);
            }
            const u64 bal = __ballot(a);
            if (a) act[tw & 1][__popcll(bal & ((1ull << tid) - 1))] = tid;
            if (tid == 0) nact[tw & 1] = __popcll(bal);
        }
        __syncthreads();
        const int cnt = nact[tw & 1];
        for (int c0 = 0; c0 < cnt; c0 += 1 << (2 * lsub)) {
        S rs[E], rp[E];
        u64 os[E], op[E];
        int ii[E], jj[E], lo[E], ms[E], mp[E], phs[E], php[E];
        bool ok[E], pair[E], trans[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int oa = tid + e * DM_THREADS, t = oa >> (2 * V), o = oa & (T * T - 1);
            if constexpr (LAYOUT == ORK_TILED) { ii[e] = o >> V; jj[e] = o & (T - 1); }
            else { jj[e] = o >> V; ii[e] = o & (T - 1); }
            lo[e] = (t >> lsub) * T * LDS_STRIDE + (t & ((1 << lsub) - 1)) * T;
            ok[e] = pair[e] = trans[e] = false;
            ms[e] = mp[e] = phs[e] = php[e] = 0;
            u64 off = 0, offp = 0;
            if (c0 + t < cnt) {
                u64 TR, TC, pr, pc;
                bool tr;
                int ps;
                sub_tile_at<LAYOUT>(p, base + (u64)act[tw & 1][c0 + t], TR, TC);
                (void)sub_owner(p, TR, TC, pr, pc, tr, ps);
                {
                    bool st;
                    sub_tile_entry<LAYOUT>(p, TR, TC, o, off, st);
                    if (st) {
                        ok[e] = true;
                        trans[e] = tr;
                        pair[e] = pr != TR || pc != TC;
                        phs[e] = ps;
                        if (LAYOUT == ORK_TILED && TR != TC && (TR >> d) == (TC >> d))
                            ms[e] = (int)(TC & dm) - (int)(TR & dm);
                        if (pair[e]) {
                            bool stp;
                            sub_tile_entry<LAYOUT>(p, pr, pc, o, offp, stp);
                            int pa, pb;
                            (void)sub_src_tile(p, pr, pa);
                            (void)sub_src_tile(p, pc, pb);
                            php[e] = (sub_ph(p, pa) - sub_ph(p, pb)) & 3;
                            if (LAYOUT == ORK_TILED && pr != pc && (pr >> d) == (pc >> d))
                                mp[e] = (int)(pc & dm) - (int)(pr & dm);
                        }
                    }
                }
            }
            os[e] = off;
            op[e] = offp;
            // loads without a condition: a slot with nothing to do reads rho[0], a tile without a partner
            // its own entry again
            rs[e] = ldr<NTA>(rho + off);
            rp[e] = ldr<NTA>(rho + (pair[e] ? offp : off));
        }
        cd vs[E], vp[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            vs[e] = to_cd(rs[e]);
            vp[e] = to_cd(rp[e]);
        }
        // new s (i, j) = i^phs src(i, j), src = old p (old s without a partner), or i^phs conj(src(j, i))
        // for a transposed pair; the upper mirror of a new tile X takes conj(X(j, i)).  Outputs without a
        // transpose go straight from registers.
        bool needa = false, needb = false;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (!ok[e]) continue;
            const cd src = sel(pair[e], vp[e], vs[e]);
            if (!trans[e]) stx<NTA>(rho + os[e], cipow(src, phs[e]));
            else if (ms[e]) stx<NTA>(rho + os[e] + ms[e] * mstep, cipow(src, -phs[e]));
            if (pair[e]) {
                if (!trans[e]) stx<NTA>(rho + op[e], cipow(vs[e], php[e]));
                else if (mp[e]) stx<NTA>(rho + op[e] + mp[e] * mstep, cipow(vs[e], -php[e]));
            }
            needa |= trans[e] || ms[e] != 0;
            needb |= pair[e] && (trans[e] || mp[e] != 0);
        }
        // transposed outputs through LDS, one virtual tile per T x T block (each barrier also keeps the
        // previous round's LDS reads ahead of the next round's writes)
        if (__syncthreads_or(needa)) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (ok[e] && (trans[e] || ms[e])) ts[lo[e] + ii[e] * LDS_STRIDE + jj[e]] = sel(pair[e], vp[e], vs[e]);
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (!ok[e] || !(trans[e] || ms[e])) continue;
                const cd x = cconj(ts[lo[e] + jj[e] * LDS_STRIDE + ii[e]]);
                if (trans[e]) stx<NTA>(rho + os[e], cipow(x, phs[e]));
                else stx<NTA>(rho + os[e] + ms[e] * mstep, cipow(x, -phs[e]));
            }
        }
        if (__syncthreads_or(needb)) {
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (ok[e] && pair[e] && (trans[e] || mp[e])) ts[lo[e] + ii[e] * LDS_STRIDE + jj[e]] = vs[e];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (!ok[e] || !pair[e] || !(trans[e] || mp[e])) continue;
                const cd x = cconj(ts[lo[e] + jj[e] * LDS_STRIDE + ii[e]]);
                if (trans[e]) stx<NTA>(rho + op[e], cipow(x, php[e]));
                else stx<NTA>(rho + op[e] + mp[e] * mstep, cipow(x, -php[e]));
            }
        }
        // the next slot set writes LDS only after its first barrier, which follows these reads
        }
    }
}

// ---- permutations inside 32 x 32 tiles: every gate qubit below 5 -------------------------------
// When every qubit of cx, cy, swap or ccx lies below 5 (32 x 32 storage tiles; packed: 32 x 32 blocks),
// each tile maps onto itself: new (i, j) = i^(pha[i] - pha[j]) old (P i, P j) with P and pha 32-entry
// tables of the in-tile index (P: the source's in-tile index, pha: the phase exponent of the source's
// gate bits).  One workgroup per stored tile: its changed entries (the sources of changed entries are
// changed entries) are loaded with every load of a thread in flight, staged in LDS (a packed diagonal
// block fills its upper half from the conjugated lower half), and stored from the table lookup.
struct PermInPlan {
    u64 ntiles, D;
    int L;              // tiled: 5; packed: 5 (blocks)
    int xcd;
    int k, q0, q1, q2;  // gate arguments (in-tile bits)
    unsigned pinv_bits; // bits 3a..3a+2: gate bits a take the amplitude with gate bits pinv[a] ...
    unsigned ph_bits;   // bits 2s, 2s+1: ... times i^ph[s] of that source s
};

// in-tile index i -> in-tile index of its source, and the phase exponent of that source, computed from the
// packed gate tables (a table lookup per entry in the kernel arguments waited for each load)
__device__ __forceinline__ int intile_src(const PermInPlan &p, int i, int &ph) {
    int a = (i >> p.q0) & 1;
    if (p.k > 1) a |= ((i >> p.q1) & 1) << 1;
    if (p.k > 2) a |= ((i >> p.q2) & 1) << 2;
    const int sa = (int)((p.pinv_bits >> (3 * a)) & 7), d = a ^ sa;
    ph = (int)((p.ph_bits >> (2 * sa)) & 3);
    return i ^ ((d & 1) << p.q0) ^ (((d >> 1) & 1) << p.q1) ^ (((d >> 2) & 1) << p.q2);
}

template <typename S, int LAYOUT, int NTA>
__global__ __launch_bounds__(DM_THREADS) void dm_perm_intile_kernel(S *__restrict__ rho, PermInPlan p) {
    __shared__ cd t[32 * LDS_STRIDE];
    const int tid = threadIdx.x;
    const u64 w = remap_group(blockIdx.x, p.ntiles, p.xcd);
    if (w >= p.ntiles) return;
    u64 TR, TC;
    if (LAYOUT == ORK_PACKED && ORK_PERMT_COLUMNS) {
        // down block columns: a block's column runs share their end lines with the blocks above and below
        u64 rr, cc;
        tri_decode(p.ntiles - 1 - w, rr, cc);
        const u64 nt = p.D >> 5;
        TR = nt - 1 - cc;
        TC = nt - 1 - rr;
    } else {
        tri_decode(w, TR, TC);
    }
    const bool dg = TR == TC;
    cd v[4];
    S raw[4];
    u64 off[4];
    int si[4], sj[4], pd[4];   // sources of row and column, phase difference
    bool ld[4];
    int ii[4], jj[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int o = tid + e * DM_THREADS;
        int i, j;
        if constexpr (LAYOUT == ORK_TILED) {
            i = o >> 5;
            j = o & 31;
            off[e] = tile_base(TR, TC, 5) + (u64)o;
        } else {
            j = o >> 5;
            i = o & 31;
            const u64 r = TR * 32 + (u64)i, c = TC * 32 + (u64)j;
            off[e] = r >= c ? col_base(p.D, c) + (r - c) : 0;
        }
        ii[e] = i;
        jj[e] = j;
        const bool stored = LAYOUT == ORK_TILED || !dg || i >= j;
        // changed: moves or takes a phase (in a packed diagonal block, an entry is also needed as the
        // mirror of its upper-half partner, which changes exactly when it does)
        int phi, phj;
        si[e] = intile_src(p, i, phi);
        sj[e] = intile_src(p, j, phj);
        pd[e] = (phi - phj) & 3;
        const bool chg = si[e] != i || sj[e] != j || pd[e];
        ld[e] = stored && chg;
        raw[e].x = 0;
        raw[e].y = 0;
        if (ld[e]) raw[e] = ldr<NTA>(rho + off[e]);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if constexpr (sizeof(S) == 8) hold(raw[e]);   // complex64: every load issued before the first conversion
        v[e] = to_cd(raw[e]);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if (!ld[e]) continue;
        t[ii[e] * LDS_STRIDE + jj[e]] = v[e];
        if (LAYOUT == ORK_PACKED && dg && ii[e] > jj[e]) t[jj[e] * LDS_STRIDE + ii[e]] = cconj(v[e]);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if (!ld[e]) continue;
        const int i = ii[e], j = jj[e];
        stx<NTA>(rho + off[e], cipow(t[si[e] * LDS_STRIDE + sj[e]], pd[e]));
    }
}

// ---- elementwise phase kernels: rho_rc *= F(r) conj(F(c)) --------------------------------
// The stored entries are streamed in memory order as one flat array: workgroup w takes the
// E * 256 consecutive entries from w * E * 256 on, thread t the entries w * E * 256 + t + 256 j
// (j < E), so every load and store instruction covers consecutive addresses.  Each thread decodes
// (r, c) of its entries from their flat index (tiled: tile index, then row-major inside the tile;
// packed: the column-major triangle), issues all its loads, then multiplies and stores.
// MODE 0: gate phases from a 2^k table with a per-pair skip mask -- entries whose factor is exactly 1
// are neither read nor written; MODE 1: per-index phases (exp_diag).  Diagonal entries are never
// changed (src/circ/circ_packed.c:28-71, src/gate/tiled/gate_tiled_cz.c:32-525 touch only the
// entries whose factor differs from 1).  WL (whole lines): a 128 B line holding any changed entry is
// read and written in full, its other entries written back exactly as read, so that HBM sees
// whole-line writes where the changed entries do not fill lines (a gate qubit below 3 in complex128).
struct PhaseArgs {
    int k;
    int qs[3];
    unsigned skip[8];   // bit b of skip[a]: entries with row index a, column index b unchanged
    u64 skipm;          // the same as one mask, bit a * 8 + b (tested without a memory load)
    cd f[64];           // f[a * 8 + b] = F[a] conj(F[b])
};

// gate bits of index x (bit j <- qubit qj) from qubits held in registers: the kernels read qs once
// instead of once per entry from the kernel arguments
__device__ __forceinline__ int gate_bits(u64 x, int k, int q0, int q1, int q2) {
    int v = (int)((x >> q0) & 1);
    if (k > 1) v |= (int)((x >> q1) & 1) << 1;
    if (k > 2) v |= (int)((x >> q2) & 1) << 2;
    return v;
}

struct PhasePlan {
    u64 dim;       // D = 2^n
    u64 len;       // stored entries
    u64 nchunks;   // workgroups of the launch
    int L;         // tiled: tiles of 2^L x 2^L (one zero-padded tile when n < L)
    int xcd;
    int tile_exit; // tiled, whole tiles per workgroup: leave at once when none holds a changed entry
    int tpc;       // tiles per workgroup (tile_exit)
    u64 tchg;      // bit a * 8 + b: a tile whose gate bits above the tile are (a, b) holds a changed entry
};

// (r, c) of flat index x; the cache holds the last tile (tiled: key = tile, P, Q = TR, TC) or column
// (packed: key = offset of the next column, P = column, Q = offset of its diagonal entry)
template <int LAYOUT>
__device__ __forceinline__ void flat_rc(const PhasePlan &pp, u64 x, u64 &key, u64 &P, u64 &Q, u64 &r, u64 &c) {
    if constexpr (LAYOUT == ORK_TILED) {
        const int L = pp.L;
        const u64 t = x >> (2 * L), o = x & ((((u64)1) << (2 * L)) - 1);
        if (t != key) {
            tri_decode(t, P, Q);
            key = t;
        }
        r = (P << L) + (o >> L);
        c = (Q << L) + (o & ((((u64)1) << L) - 1));
    } else {
        const u64 D = pp.dim;
        if (x >= key || x < Q) {
            // column c: c (2D - c + 1) / 2 <= x < (c + 1) (2D - c) / 2
            const double b = 2.0 * (double)D + 1.0;
            u64 cc = (u64)((b - sqrt(b * b - 8.0 * (double)x)) * 0.5);
            if (cc >= D) cc = D - 1;
            while (cc > 0 && col_base(D, cc) > x) --cc;
            while (cc + 1 < D && col_base(D, cc + 1) <= x) ++cc;
            P = cc;
            Q = col_base(D, cc);
            key = Q + (D - cc);
        }
        c = P;
        r = P + (x - Q);
    }
}

template <typename S, int LAYOUT, int MODE, int NT, int WL, int E>
__global__ __launch_bounds__(DM_THREADS) void dm_phase_kernel(S *__restrict__ rho, PhasePlan pp, PhaseArgs a,
                                                              const double2 *__restrict__ cs) {
    constexpr int G = 128 / (int)sizeof(S);   // entries per 128 B line: lanes lane & ~(G - 1) .. + G - 1
    const u64 w = remap_group(blockIdx.x, pp.nchunks, pp.xcd);
    const u64 x0 = w * (u64)(E * DM_THREADS) + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int gk = a.k, q0 = a.qs[0], q1 = a.qs[1], q2 = a.qs[2];
    const u64 skipm = a.skipm;
    u64 key = ~(u64)0, P = 0, Q = 0;
    if (LAYOUT == ORK_PACKED) key = 0;
    if constexpr (MODE == 0 && LAYOUT == ORK_TILED) {
        if (pp.tile_exit) {
            // the workgroup's tiles w tpc ..: their gate bits above the tile alone decide whether any entry
            // changes
            bool any = false;
            for (int u = 0; u < pp.tpc; ++u) {
                const u64 t = w * (u64)pp.tpc + (u64)u;
                if ((t << (2 * pp.L)) >= pp.len) break;
                tri_decode(t, P, Q);
                const int at = gate_bits(P << pp.L, gk, q0, q1, q2), bt = gate_bits(Q << pp.L, gk, q0, q1, q2);
                any = any || ((pp.tchg >> (at * 8 + bt)) & 1);
            }
            if (!any) return;
            if (pp.tpc == 1) key = w;   // P, Q hold the decode of tile w; else flat_rc decodes per tile
        }
    }
    cd f[E];
    double2 pr[E], pc[E];   // exp_diag: the row's and the column's phase, loaded for every entry first
    bool chg[E], act[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const u64 x = x0 + (u64)j * DM_THREADS;
        bool ok = x < pp.len;
        u64 r = 0, c = 0;
        if (ok) flat_rc<LAYOUT>(pp, x, key, P, Q, r, c);
        ok = ok && r < pp.dim && c < pp.dim && r != c;   // padding of a one-tile state; the diagonal
        f[j] = cmk(1.0, 0.0);
        if constexpr (MODE == 0) {
            const int ar = gate_bits(r, gk, q0, q1, q2), ac = gate_bits(c, gk, q0, q1, q2);
            ok = ok && !((skipm >> (ar * 8 + ac)) & 1);
            if (ok) f[j] = a.f[ar * 8 + ac];
        } else {
            // without a condition (an entry left alone reads cs[0]), all in flight before the first product:
            // loaded and used inside the entry's condition, each pair was waited for in turn
            pr[j] = cs[ok ? r : 0];
            pc[j] = cs[ok ? c : 0];
        }
        chg[j] = ok;
        if constexpr (WL) {
            const unsigned long long m = __ballot(ok);
            act[j] = ((m >> (lane & ~(G - 1))) & ((1ull << G) - 1)) != 0;
        } else {
            act[j] = ok;
        }
    }
    S raw[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
        raw[j].x = 0;
        raw[j].y = 0;
        if (act[j]) raw[j] = ldr<NT>(rho + x0 + (u64)j * DM_THREADS);
    }
    if constexpr (MODE != 0) {
        // e^{-i(d_r - d_c)t} = (cr - i sr)(cc + i sc)   (src/circ/circ_packed.c:54-60)
#pragma unroll
        for (int j = 0; j < E; ++j)
            if (chg[j]) f[j] = cmk(pr[j].x * pc[j].x + pr[j].y * pc[j].y, -pr[j].y * pc[j].x + pr[j].x * pc[j].y);
    }
    cd v[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
        if constexpr (sizeof(S) == 8) hold(raw[j]);   // complex64: every load issued before the first conversion
        v[j] = to_cd(raw[j]);
    }
#pragma unroll
    for (int j = 0; j < E; ++j) {
        if (!act[j]) continue;
        cd o = v[j];
        if (chg[j]) {
            if constexpr (MODE == 0) o = cmul(f[j], v[j]);
            else o = cmk(f[j].x * v[j].x - f[j].y * v[j].y, f[j].x * v[j].y + f[j].y * v[j].x);
        }
        stx<NT>(rho + x0 + (u64)j * DM_THREADS, o);
    }
}

__global__ __launch_bounds__(DM_THREADS) void sincos_kernel(double2 *__restrict__ cs, const double *__restrict__ d,
                                                            u64 dim, double t) {
    const u64 x = (u64)blockIdx.x * DM_THREADS + threadIdx.x;
    if (x >= dim) return;
    double s, c;
    sincos(d[x] * t, &s, &c);
    cs[x] = cmk(c, s);
}

// ---- mean: sum_x obs[x] Re rho_xx ---------------------------------------------------------
constexpr int RED_BLOCKS = 2048;   // same partials layout as the statevector reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <typename S, int LAYOUT>
__global__ __launch_bounds__(DM_THREADS) void dm_mean_kernel(const S *__restrict__ rho, DmPlan p,
                                                             const double *__restrict__ obs, double *__restrict__ part) {
    __shared__ double red[DM_THREADS / 64];
    double acc = 0.0;
    for (u64 x = (u64)blockIdx.x * DM_THREADS + threadIdx.x; x < p.dim; x += (u64)gridDim.x * DM_THREADS)
        acc += obs[x] * ldc(rho + elem_off<LAYOUT>(p, x, x)).x;
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < DM_THREADS / 64; ++i) s += red[i];
        part[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(DM_THREADS) void dm_finish_kernel(double *__restrict__ part, int nblk) {
    __shared__ double red[DM_THREADS / 64];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += DM_THREADS) acc += part[b];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < DM_THREADS / 64; ++i) s += red[i];
        part[RED_BLOCKS] = s;
    }
}

// ---- host-side planning -------------------------------------------------------------------
// Largest virtual tile B = 2^LB (at most the storage tile for the tiled layout,
// 32 for packed) such that the group, B * 2^g, still fits W_MAX.
DmPlan make_plan(int layout, int L, int n, const int *qs, int k) {
    DmPlan p;
    p.xcd = 0;
    p.col = 0;
    p.dim = (u64)1 << n;
    p.L = L;
    const int cap = (layout == ORK_TILED) ? (L < n ? L : n) : (LW_MAX < n ? LW_MAX : n);
    int LB = cap, g = 0;
    for (; LB >= 0; --LB) {
        g = 0;
        for (int j = 0; j < k; ++j) g += qs[j] >= LB;
        if (LB + g <= LW_MAX && LB + g <= n) break;
    }
    p.LB = LB;
    p.g = g;
    p.LW = LB + g;
    int gq[3], ng = 0;
    for (int j = 0; j < k; ++j)
        if (qs[j] >= LB) gq[ng++] = qs[j] - LB;
    for (int a = 0; a < ng; ++a)
        for (int b = a + 1; b < ng; ++b)
            if (gq[b] < gq[a]) { int t = gq[a]; gq[a] = gq[b]; gq[b] = t; }
    for (int j = 0; j < 3; ++j) p.gpos[j] = j < ng ? gq[j] : 0;
    const u64 T = p.dim >> p.LW;
    p.nt = T;
    p.ngroups = T * (T + 1) / 2;
    return p;
}

// Launch shapes (residency cap through unused dynamic LDS, XCD remap, group
// order); set through orkan_set_launch_shape, defaults from a sweep over every
// target qubit of a 14-qubit density matrix on MI355X.  Unlike the statevector
// streams these kernels want every resident workgroup they can get.  Packed:
// column order plus the XCD remap lets the groups that share the cache lines at
// the ends of a column run meet in one L2 (without it 1.33x the algorithmic
// bytes are fetched); tiled: row order with the remap.
// nt: nontemporal loads and stores; lp: log2 of the workgroups sharing one tile quad (cross-tile kernel)
// wl: whole-line writes of the phase kernels (flags bit 6)
// nosub (whole-tile permutations only): straddling gates to the tile-group kernel instead of virtual tiles
struct Shape { int xcd, col, lds, nt, lp, wl, nosub; };
constexpr Shape kGroupShape[2] = {{1, 1, 0, 0}, {1, 0, 0, 1}};   // [packed, tiled]
// phase kernels (cz, exp_diag, ...): flat streams of E entries per thread, XCD runs of 32 workgroups,
// nontemporal accesses, whole-line writes where the changed entries do not fill lines
constexpr Shape kPhaseShape[2] = {{2, 0, 0, 1, 0, 1}, {2, 0, 0, 1, 0, 1}};
Shape g_group_shape[2] = {kGroupShape[0], kGroupShape[1]};
Shape g_phase_shape[2] = {kPhaseShape[0], kPhaseShape[1]};

unsigned grid_for(u64 ngroups) {
    const u64 cap = (u64)1 << 30;
    return (unsigned)(ngroups < cap ? ngroups : cap);
}

template <typename S, int LAYOUT>
int launch_group(hipStream_t st, void *rho, DmPlan p, int opk, const DmOp &op) {
    const unsigned grid = grid_for(p.ngroups);
    const Shape sh = g_group_shape[LAYOUT == ORK_TILED ? 1 : 0];
    p.xcd = sh.xcd;
    p.col = sh.col;
    const int lds = sh.lds > (int)sizeof(cd) * W_MAX * LDS_STRIDE ? sh.lds - (int)sizeof(cd) * W_MAX * LDS_STRIDE : 0;
    // bit 6 of the group shape flags (the wl field): every load of a thread in flight at once; bit 2
    // (nt): nontemporal accesses on the tiled layout where virtual tiles are whole 128 B rows, on states
    // of at least the nontemporal minimum
    const u64 tl = p.dim >= ((u64)1 << p.L) ? p.dim >> p.L : 1;
    const bool nt = LAYOUT == ORK_TILED && sh.nt && ((u64)sizeof(S) << p.LB) >= 128 &&
                    ((tl * (tl + 1) / 2) << (2 * p.L)) * sizeof(S) >= ork_nt_min_bytes();
#define ORK_GROUP_LAUNCH(OPC)                                                                                    \
    do {                                                                                                         \
        if (nt) {                                                                                                \
            if (sh.wl) dm_group_kernel<S, LAYOUT, OPC, 1, 1><<<grid, DM_THREADS, lds, st>>>((S *)rho, p, op);   \
            else dm_group_kernel<S, LAYOUT, OPC, 0, 1><<<grid, DM_THREADS, lds, st>>>((S *)rho, p, op);         \
        } else {                                                                                                 \
            if (sh.wl) dm_group_kernel<S, LAYOUT, OPC, 1, 0><<<grid, DM_THREADS, lds, st>>>((S *)rho, p, op);   \
            else dm_group_kernel<S, LAYOUT, OPC, 0, 0><<<grid, DM_THREADS, lds, st>>>((S *)rho, p, op);         \
        }                                                                                                        \
    } while (0)
    switch (opk) {
    case ORK_DM_SOP: ORK_GROUP_LAUNCH(ORK_DM_SOP); break;
    case ORK_DM_MONO: ORK_GROUP_LAUNCH(ORK_DM_MONO); break;
    case ORK_DM_DENSE: ORK_GROUP_LAUNCH(ORK_DM_DENSE); break;
    case ORK_DM_SOP2: ORK_GROUP_LAUNCH(ORK_DM_SOP2); break;
    default: return (int)hipErrorInvalidValue;
    }
#undef ORK_GROUP_LAUNCH
    ORK_LAUNCH_CHECK();
    return 0;
}

// streamed tiled superoperator; shape [q < L, q >= L].  A later sweep of every target (h on 14 qubits)
// moved the complex128 defaults to XCD runs of 32 workgroups under an 80 KB cap for in-tile targets 3-4
// (6.74 against 6.58 TB/s) and a 64 KB cap for the split tile quads (6.36 against 6.23).  Measured on a 14-qubit density matrix
// (32 x 32 tiles): in-tile targets stream best like the statevector kernels, with consecutive
// chunks on one XCD and residency capped by 64 KB of LDS per workgroup (6.1 TB/s against 5.6 TB/s
// uncapped); tile quads best split over 4 workgroups (8 rows of each tile, 16 KB per workgroup), in
// plain order with a 48 KB cap (6.18 TB/s against 5.94 whole).  Nontemporal
// accesses lift the tile quads from 5.4 to 6.0 TB/s and in-tile targets 3 and 4 from 6.2 to 6.55;
// where a quad's two columns share 128 B lines (targets below 3 in complex128, below 4 in
// complex64) they cost 1-2 % and refetch lines (complex64 target 3: 1.1 x the reads), so stay off.
// The packed kernel's column runs share lines between workgroups: nontemporal there loses 20 %.
// [in-tile, above the tile] x [complex128, complex64].  Complex64 moves half the bytes per workgroup:
// tile quads split over 2 workgroups (not 4) keep 16 KB each (6.34 TB/s against 5.24 with 4), and
// in-tile targets run best in plain order (5.76 against 5.64 TB/s with the remap).
// Bit 3 (lp) of an in-tile shape enables the lane-exchange kernel for targets whose quad columns
// share 128 B lines; that kernel runs with kLaneShape (in plain order, 48 KB cap for 8 entries per
// thread: 6.41 against 6.01 TB/s for complex128 targets 0-2, 6.18 against 5.54 for complex64 0-3).
constexpr Shape kStreamShape[4] = {{2, 0, 81920, 1, 1}, {0, 0, 65536, 1, 2}, {0, 0, 65536, 1, 1}, {0, 0, 49152, 1, 1}};
constexpr Shape kLaneShape = {0, 0, 49152, 1, 0};
Shape g_stream_shape[4] = {kStreamShape[0], kStreamShape[1], kStreamShape[2], kStreamShape[3]};
int g_stream_on = 1;
// controlled gates in the streamed kernels leave their unchanged quads alone when the lowest control
// qubit is at least the line bits (3 for complex128, 4 for complex64) plus a margin: [packed cx / cy,
// tiled, packed ccx].  Packed column runs start anywhere in a line, so each 2^c-row run of skipped
// quads leaves partial lines at its ends: on 14 qubits that costs cx (1/4 of the quads skipped) 2-8 %
// at controls 3-4, while ccx (9/16 skipped) gains 25-55 % there; tiled rows are line-aligned.
constexpr int kSkipMargin[3] = {2, 0, 0};
int g_skip_margin[3] = {kSkipMargin[0], kSkipMargin[1], kSkipMargin[2]};

template <typename S, int SKIP>
int launch_stream_t(hipStream_t st, void *rho, int L, int n, int q, const DmOp &op) {
    const u64 nt = (u64)1 << (n - L);
    StreamPlan sp;
    sp.L = L;
    Shape sh = g_stream_shape[(q < L ? 0 : 1) + (sizeof(S) == 8 ? 2 : 0)];
    if (((nt * (nt + 1) / 2) << (2 * L)) * sizeof(S) < ork_nt_min_bytes()) sh.nt = 0;
    sp.xcd = sh.xcd;
    const int lds = sh.lds > (int)sizeof(cd) * W_MAX * LDS_STRIDE ? sh.lds - (int)sizeof(cd) * W_MAX * LDS_STRIDE : 0;
    if (q < L && sh.lp && sh.nt && ((u64)sizeof(S) << q) < 128) {
        // quad columns share lines: the lane-exchange kernel keeps whole lines per instruction, so
        // nontemporal accesses apply here too; 4 (complex128) or 8 (complex64) entries per thread
        constexpr int U = sizeof(S) == 8 ? 4 : 2;
        sp.m = q;
        sp.xcd = kLaneShape.xcd;
        sp.nitems = (nt * (nt + 1) / 2) << (2 * L - 1);
        sp.nblocks = (sp.nitems + DM_THREADS * U - 1) / (DM_THREADS * U);
        if (sp.nblocks > 0x7fffffffu) return (int)hipErrorInvalidValue;
        dm_sop_intile_lane_kernel<S, U, 1, SKIP><<<(unsigned)sp.nblocks, DM_THREADS, kLaneShape.lds / (U / 2), st>>>(
            (S *)rho, sp, op);
    } else if (q < L) {
        // complex64 storage: two quads per thread and half the LDS cap keep the bytes in flight
        constexpr int U = sizeof(S) == 8 ? 2 : 1;
        sp.m = q;
        sp.nitems = (nt * (nt + 1) / 2) << (2 * L - 2);
        sp.nblocks = (sp.nitems + DM_THREADS * U - 1) / (DM_THREADS * U);
        if (sp.nblocks > 0x7fffffffu) return (int)hipErrorInvalidValue;
        if (sh.nt && ((u64)sizeof(S) << q) >= 128)   // the quad's columns in different 128 B lines
            dm_sop_intile_kernel<S, U, 1, SKIP><<<(unsigned)sp.nblocks, DM_THREADS, sh.lds / U, st>>>((S *)rho, sp, op);
        else
            dm_sop_intile_kernel<S, U, 0, SKIP><<<(unsigned)sp.nblocks, DM_THREADS, sh.lds / U, st>>>((S *)rho, sp, op);
    } else {
        sp.m = q - L;
        const u64 h = nt >> 1;
        const int lp = L == 5 ? sh.lp : 0;   // split tile quads: 32 x 32 tiles only
        sp.nitems = h * (h + 1) / 2;
        sp.nblocks = sp.nitems << lp;
        if (sp.nblocks > 0x7fffffffu) return (int)hipErrorInvalidValue;
        const unsigned nb = (unsigned)sp.nblocks;
        S *r = (S *)rho;
        if (sh.nt) {
            if (lp == 2) dm_sop_tiles_kernel<S, 1, 4, SKIP><<<nb, DM_THREADS, lds, st>>>(r, sp, op);
            else if (lp == 1) dm_sop_tiles_kernel<S, 1, 2, SKIP><<<nb, DM_THREADS, lds, st>>>(r, sp, op);
            else dm_sop_tiles_kernel<S, 1, 1, SKIP><<<nb, DM_THREADS, lds, st>>>(r, sp, op);
        } else {
            if (lp == 2) dm_sop_tiles_kernel<S, 0, 4, SKIP><<<nb, DM_THREADS, lds, st>>>(r, sp, op);
            else if (lp == 1) dm_sop_tiles_kernel<S, 0, 2, SKIP><<<nb, DM_THREADS, lds, st>>>(r, sp, op);
            else dm_sop_tiles_kernel<S, 0, 1, SKIP><<<nb, DM_THREADS, lds, st>>>(r, sp, op);
        }
    }
    ORK_LAUNCH_CHECK();
    return 0;
}

template <typename S>
int launch_stream(hipStream_t st, void *rho, int L, int n, int q, const DmOp &op) {
    return op.skip ? launch_stream_t<S, 1>(st, rho, L, n, q, op) : launch_stream_t<S, 0>(st, rho, L, n, q, op);
}

// streamed packed superoperator: column order and an 80 KB residency cap; each 32 x 32 block split
// over 4 workgroups of 8 columns (16 KB each), each XCD taking runs of 32 workgroups (8 blocks) in
// turn.  Measured on 14 qubits: 5.37 TB/s (plain dispatch order 5.24, whole blocks with contiguous
// XCD ranges 5.06, the tile groups 4.7).  Complex64 storage keeps whole
// blocks (see launch_pstream).
constexpr Shape kPstreamShape = {2, 1, 81920, 0, 2};
Shape g_pstream_shape = kPstreamShape;
// complex64: half a block per workgroup, XCD runs, at most 3 workgroups per CU.  Measured per
// superoperator on 14 qubits: 5.22 TB/s against 4.12 for whole blocks with the contiguous XCD
// remap and 81920 B (the shape kept before the runs existed), 5.11-5.14 for quarter blocks.
constexpr Shape kPstreamShape32 = {2, 1, 40960, 0, 1};
Shape g_pstream_shape32 = kPstreamShape32;

template <typename S, int SKIP>
int launch_pstream_t(hipStream_t st, void *rho, int n, int q, const DmOp &op) {
    PackedStreamPlan pp;
    pp.dim = (u64)1 << n;
    pp.q = q;
    pp.nt = pp.dim >> 6;
    pp.nblocks = pp.nt * (pp.nt + 1) / 2;
    const Shape &sh = sizeof(S) == 8 ? g_pstream_shape32 : g_pstream_shape;
    pp.xcd = sh.xcd;
    pp.col = sh.col;
    if (pp.nblocks > 0x7fffffffu) return (int)hipErrorInvalidValue;
    const int lds = sh.lds > (int)sizeof(cd) * 32 * LDS_STRIDE ? sh.lds - (int)sizeof(cd) * 32 * LDS_STRIDE : 0;
    const int lp = sh.lp;
    if ((pp.nblocks << lp) > 0x7fffffffu) return (int)hipErrorInvalidValue;
    const unsigned nb = (unsigned)(pp.nblocks << lp);
    S *r = (S *)rho;
    // bit 6 (the wl field) with nontemporal accesses: only on the lines a workgroup owns alone
    const bool hybrid = sh.nt && sh.wl;
    if (hybrid) {
        if (lp == 2) dm_sop_packed_kernel<S, 2, 4, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
        else if (lp == 1) dm_sop_packed_kernel<S, 2, 2, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
        else dm_sop_packed_kernel<S, 2, 1, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
    } else if (sh.nt) {
        if (lp == 2) dm_sop_packed_kernel<S, 1, 4, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
        else if (lp == 1) dm_sop_packed_kernel<S, 1, 2, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
        else dm_sop_packed_kernel<S, 1, 1, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
    } else {
        if (lp == 2) dm_sop_packed_kernel<S, 0, 4, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
        else if (lp == 1) dm_sop_packed_kernel<S, 0, 2, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
        else dm_sop_packed_kernel<S, 0, 1, SKIP><<<nb, DM_THREADS, lds, st>>>(r, pp, op);
    }
    ORK_LAUNCH_CHECK();
    return 0;
}

template <typename S>
int launch_pstream(hipStream_t st, void *rho, int n, int q, const DmOp &op) {
    return op.skip ? launch_pstream_t<S, 1>(st, rho, n, q, op) : launch_pstream_t<S, 0>(st, rho, n, q, op);
}

// whole-tile permutations: XCD runs of 32 workgroups, one workgroup per tile pair (splitting a pair over 2
// or 4 workgroups lost 4-25 % on 14 qubits), nontemporal accesses on the tiled layout; `col` set switches
// the kernel off (tile groups instead); set through orkan_set_launch_shape 22
constexpr Shape kPermTileShape = {2, 0, 0, 1, 0, 0};
Shape g_permt_shape = kPermTileShape;

// the virtual tile log on which the whole-tile kernels run the gate, or -1: an involution whose qubits all
// lie at or above the tile (log2 T: the storage tile, or 32 x 32 blocks for packed) runs on whole tiles
// (V = T); with `sub`, one whose lowest qubit lies below the tile runs on virtual tiles of that log where
// their rows fill 128 B lines (V >= 3 complex128, V >= 4 complex64)
int perm_tiles_v(int layout, int L, int n, int prec, const ork_mono_t *m, bool sub) {
    if (g_permt_shape.col) return -1;
    const int T = layout == ORK_TILED ? L : 5;
    if (n <= T) return -1;
    int lo = T;
    for (int j = 0; j < m->k; ++j) lo = m->qs[j] < lo ? m->qs[j] : lo;
    for (int a = 0; a < (1 << m->k); ++a)
        if (m->perm[m->perm[a]] != a) return -1;
    if (lo == T) return T;
    return (sub && !g_permt_shape.nosub && lo >= (prec == ORK_F32 ? 4 : 3)) ? lo : -1;
}
bool perm_tiles_ok(int layout, int L, int n, const ork_mono_t *m) {
    return perm_tiles_v(layout, L, n, ORK_F64, m, false) >= 0;
}

// Is the gate U on one target qubit controlled by the others (cx, cy, ccx)?  Then tq = the target,
// U = [[0, i^pu01], [i^pu10, 0]] and cmask the control qubits as bits of the row / column index.
bool controlled_single_target(const ork_mono_t *m, int &tq, int &pu01, int &pu10, u64 &cmask) {
    const int k = m->k, K = 1 << k;
    int flip = 0;
    for (int a = 0; a < K; ++a) flip |= a ^ m->perm[a];
    if (flip == 0 || (flip & (flip - 1))) return false;   // exactly one target bit moves
    int tj = 0;
    while (!((flip >> tj) & 1)) ++tj;
    const int cl = (K - 1) & ~flip;   // control bits (local)
    pu01 = pu10 = 0;
    for (int a = 0; a < K; ++a) {
        if ((a & cl) == cl) {
            if (m->perm[a] != (a ^ flip)) return false;
            if ((a >> tj) & 1) pu01 = m->ph[a] & 3;   // U |1> = i^pu01 |0>
            else pu10 = m->ph[a] & 3;                 // U |0> = i^pu10 |1>
        } else if (m->perm[a] != a || (m->ph[a] & 3)) {
            return false;
        }
    }
    tq = m->qs[tj];
    cmask = 0;
    for (int j = 0; j < k; ++j)
        if (j != tj) cmask |= (u64)1 << m->qs[j];
    return true;
}

// can the gate run through dm_perm_intile_kernel?  Every qubit below 5, 32 x 32 tiles (tiled with
// LOG_TILE_DIM 5, or packed blocks) and at least one whole tile
bool perm_intile_ok(int layout, int L, int n, const ork_mono_t *m) {
    if (g_permt_shape.col || n < 5 || (layout == ORK_TILED && L != 5)) return false;
    for (int j = 0; j < m->k; ++j)
        if (m->qs[j] >= 5) return false;
    return true;
}

template <typename S, int LAYOUT>
int launch_perm_intile(hipStream_t st, void *rho, int n, const ork_mono_t *m) {
    PermInPlan p;
    p.D = (u64)1 << n;
    p.L = 5;
    const u64 nt = p.D >> 5;
    p.ntiles = nt * (nt + 1) / 2;
    if (p.ntiles > 0x7fffffffu) return (int)hipErrorInvalidValue;
    int pinv[8];
    for (int a = 0; a < 8; ++a) pinv[a] = a;
    for (int a = 0; a < (1 << m->k); ++a) pinv[m->perm[a]] = a;
    p.k = m->k;
    p.q0 = m->qs[0];
    p.q1 = m->k > 1 ? m->qs[1] : 0;
    p.q2 = m->k > 2 ? m->qs[2] : 0;
    p.pinv_bits = p.ph_bits = 0;
    for (int a = 0; a < 8; ++a) {
        p.pinv_bits |= (unsigned)pinv[a] << (3 * a);
        p.ph_bits |= (unsigned)(a < (1 << m->k) ? (m->ph[a] & 3) : 0) << (2 * a);
    }
    const Shape sh = g_permt_shape;
    p.xcd = sh.xcd;
    const bool ntm = LAYOUT == ORK_TILED && sh.nt && (p.ntiles << 10) * sizeof(S) >= ork_nt_min_bytes();
    const int stat = (int)sizeof(cd) * 32 * LDS_STRIDE;
    const int lds = sh.lds > stat ? sh.lds - stat : 0;
    if (ntm) dm_perm_intile_kernel<S, LAYOUT, 1><<<(unsigned)p.ntiles, DM_THREADS, lds, st>>>((S *)rho, p);
    else dm_perm_intile_kernel<S, LAYOUT, 0><<<(unsigned)p.ntiles, DM_THREADS, lds, st>>>((S *)rho, p);
    ORK_LAUNCH_CHECK();
    return 0;
}

template <typename S, int LAYOUT>
int launch_perm_tiles(hipStream_t st, void *rho, int n, int L, const ork_mono_t *m, int V) {
    PermTilePlan p;
    p.D = (u64)1 << n;
    p.L = LAYOUT == ORK_TILED ? L : 5;
    p.V = V;
    const u64 nt = p.D >> V;
    p.ntiles = nt * (nt + 1) / 2;
    if (p.ntiles > 0x7fffffffu || V > p.L || V > LW_MAX) return (int)hipErrorInvalidValue;
    p.k = m->k;
    for (int j = 0; j < 3; ++j) p.pos[j] = j < m->k ? m->qs[j] - V : 0;
    for (int a = 0; a < 8; ++a) { p.pinv[a] = a; p.ph[a] = 0; }
    for (int a = 0; a < (1 << m->k); ++a) {
        p.pinv[m->perm[a]] = a;
        p.ph[a] = m->ph[a] & 3;
    }
    p.pinv_bits = p.ph_bits = 0;
    for (int a = 0; a < 8; ++a) {
        p.pinv_bits |= (unsigned)p.pinv[a] << (3 * a);
        p.ph_bits |= (unsigned)p.ph[a] << (2 * a);
    }
    const Shape sh = g_permt_shape;
    p.xcd = sh.xcd;
    p.tpw = 1 << (2 * sh.lp);   // flags bits 3-4: log4 of the tiles per workgroup
    S *r = (S *)rho;
    const int stat = (int)sizeof(cd) * W_MAX * LDS_STRIDE;
    const int lds = sh.lds > stat ? sh.lds - stat : 0;
    if (V < p.L) {
        // virtual tiles below the storage tile: 64 candidate tiles per workgroup iteration
        const u64 per = ((u64)p.tpw) * 64;
        p.nwg = (p.ntiles + per - 1) / per;
        // nontemporal accesses for complex64 only (complex128: 4-9 % slower with them, 14 qubits)
        const bool ntm = LAYOUT == ORK_TILED && sh.nt && sizeof(S) == 8 &&
                         (p.ntiles << (2 * V)) * sizeof(S) >= ork_nt_min_bytes();
        if (ntm) dm_perm_subtiles_kernel<S, LAYOUT, 1><<<(unsigned)p.nwg, DM_THREADS, lds, st>>>(r, p);
        else dm_perm_subtiles_kernel<S, LAYOUT, 0><<<(unsigned)p.nwg, DM_THREADS, lds, st>>>(r, p);
        ORK_LAUNCH_CHECK();
        return 0;
    }
    p.nwg = (p.ntiles + p.tpw - 1) / p.tpw;
    // nontemporal accesses (flags bit 2) on the tiled layout, whose tiles are whole lines, on states of at
    // least the nontemporal minimum; the packed layout's block columns share lines with their neighbours
    const bool ntm = LAYOUT == ORK_TILED && sh.nt && (p.ntiles << (2 * p.L)) * sizeof(S) >= ork_nt_min_bytes() &&
                     ((u64)sizeof(S) << p.L) >= 128;
    if (ntm) dm_perm_tiles_kernel<S, LAYOUT, 1><<<(unsigned)p.nwg, DM_THREADS, lds, st>>>(r, p);
    else dm_perm_tiles_kernel<S, LAYOUT, 0><<<(unsigned)p.nwg, DM_THREADS, lds, st>>>(r, p);
    ORK_LAUNCH_CHECK();
    return 0;
}

// entries per thread: 64 B of each thread in flight, times 1, 2 or 4 (phase shape flags bits 3-4)
template <typename S, int LAYOUT, int MODE, int NT, int WL>
void launch_phase_t(hipStream_t st, unsigned grid, int lds, void *rho, const PhasePlan &pp, const PhaseArgs &a,
                    const double2 *cs, int em) {
    constexpr int E = sizeof(S) == 8 ? 8 : 4;
    if (em == 4) dm_phase_kernel<S, LAYOUT, MODE, NT, WL, 4 * E><<<grid, DM_THREADS, lds, st>>>((S *)rho, pp, a, cs);
    else if (em == 2) dm_phase_kernel<S, LAYOUT, MODE, NT, WL, 2 * E><<<grid, DM_THREADS, lds, st>>>((S *)rho, pp, a, cs);
    else dm_phase_kernel<S, LAYOUT, MODE, NT, WL, E><<<grid, DM_THREADS, lds, st>>>((S *)rho, pp, a, cs);
}

template <typename S, int MODE, int NT, int WL>
void launch_phase_l(hipStream_t st, int layout, unsigned grid, int lds, void *rho, const PhasePlan &pp,
                    const PhaseArgs &a, const double2 *cs, int em) {
    if (layout == ORK_PACKED) launch_phase_t<S, ORK_PACKED, MODE, NT, WL>(st, grid, lds, rho, pp, a, cs, em);
    else launch_phase_t<S, ORK_TILED, MODE, NT, WL>(st, grid, lds, rho, pp, a, cs, em);
}

template <int MODE>
int launch_phase(hipStream_t st, int prec, int layout, int L, int n, void *rho, const PhaseArgs &a,
                 const double2 *cs) {
    PhasePlan pp;
    pp.dim = (u64)1 << n;
    pp.L = L;
    if (layout == ORK_PACKED) {
        pp.len = pp.dim * (pp.dim + 1) / 2;
    } else {
        const u64 nt = n >= L ? pp.dim >> L : 1;
        pp.len = (nt * (nt + 1) / 2) << (2 * L);
    }
    Shape sh = g_phase_shape[layout == ORK_TILED ? 1 : 0];
    const u64 elem = prec == ORK_F32 ? 8 : 16;
    const int em = 1 << sh.lp;
    const u64 per = (prec == ORK_F32 ? 8 : 4) * (u64)em * DM_THREADS;
    pp.nchunks = (pp.len + per - 1) / per;
    if (pp.nchunks > 0x7fffffffu) return (int)hipErrorInvalidValue;
    pp.xcd = sh.xcd;
    if (pp.len * elem < ork_nt_min_bytes()) sh.nt = 0;
    int wl = 0;
    if (MODE == 0) {
        // tiled: the lowest gate qubit decides whether changed entries fill 128 B lines (tile rows are
        // line-aligned); runs of 64 B, aligned halves of lines, write as well without the rest of the line;
        // packed: column runs start anywhere in a line, so any line may be partial
        int lo = 7;     // past a 128 B line for either storage
        for (int i = 0; i < a.k; ++i) lo = a.qs[i] < lo ? a.qs[i] : lo;
        if ((elem << lo) < 64 || layout == ORK_PACKED) {
            wl = sh.wl;
            if (!wl) sh.nt = 0;   // partial-line writes stay cached
        }
    }
    // tiled, one tile per workgroup: a mask over the gate bits above the tile of the tiles that hold a
    // changed entry (any in-tile combination of the other gate bits not skipped)
    pp.tile_exit = 0;
    pp.tpc = 1;
    pp.tchg = 0;
    const u64 tsz = (u64)1 << (2 * L);
    if (MODE == 0 && layout == ORK_TILED && n >= L && (tsz == per || (em > 1 && tsz * (u64)em == per))) {
        pp.tpc = (int)(per / tsz);
        int hi = 0;   // gate arguments at or above the tile
        for (int j = 0; j < a.k; ++j) hi |= (a.qs[j] >= L) << j;
        for (int at = 0; at < (1 << a.k); ++at)
            for (int bt = 0; bt < (1 << a.k); ++bt) {
                if ((at & ~hi) || (bt & ~hi)) continue;
                bool any = false;
                for (int ai = 0; ai < (1 << a.k); ++ai)
                    for (int bi = 0; bi < (1 << a.k); ++bi) {
                        if ((ai & hi) || (bi & hi)) continue;
                        if (!((a.skip[at | ai] >> (bt | bi)) & 1)) any = true;
                    }
                if (any) pp.tchg |= (u64)1 << (at * 8 + bt);
            }
        pp.tile_exit = 1;
    }
    const unsigned grid = (unsigned)pp.nchunks;
    const int lds = sh.lds;
    if (prec == ORK_F32) {
        if (sh.nt) {
            if (wl) launch_phase_l<float2, MODE, 1, 1>(st, layout, grid, lds, rho, pp, a, cs, em);
            else launch_phase_l<float2, MODE, 1, 0>(st, layout, grid, lds, rho, pp, a, cs, em);
        } else {
            if (wl) launch_phase_l<float2, MODE, 0, 1>(st, layout, grid, lds, rho, pp, a, cs, em);
            else launch_phase_l<float2, MODE, 0, 0>(st, layout, grid, lds, rho, pp, a, cs, em);
        }
    } else {
        if (sh.nt) {
            if (wl) launch_phase_l<double2, MODE, 1, 1>(st, layout, grid, lds, rho, pp, a, cs, em);
            else launch_phase_l<double2, MODE, 1, 0>(st, layout, grid, lds, rho, pp, a, cs, em);
        } else {
            if (wl) launch_phase_l<double2, MODE, 0, 1>(st, layout, grid, lds, rho, pp, a, cs, em);
            else launch_phase_l<double2, MODE, 0, 0>(st, layout, grid, lds, rho, pp, a, cs, em);
        }
    }
    ORK_LAUNCH_CHECK();
    return 0;
}

// elementwise plan: one workgroup per stored tile of the storage (tiled) or of 32 x 32 (packed)
DmPlan tile_plan(int layout, int L, int n) {
    DmPlan p = make_plan(layout, L, n, nullptr, 0);
    return p;
}

}  // namespace

// which: 0 group kernel (both layouts), 1 phase kernel (both layouts), 2 group kernel packed,
// 3 group kernel tiled, 4 phase kernel packed, 5 phase kernel tiled, 6 / 7 streamed tiled
// superoperator with the target qubit inside / above the tile, 8 streamed superoperators on (lds_bytes
// != 0) or off, 9 streamed packed superoperator; flags bit 0: XCD remap, bit 1: column order of groups,
// bit 2: nontemporal accesses (phase and streamed kernels), bits 3-4: log2 of the workgroups per tile
// quad (kernels 7, 16); 15 / 16: as 6 / 7 for complex64 storage
extern "C" void ork_dm_reset_shapes(void) {
    for (int i = 0; i < 2; ++i) {
        g_group_shape[i] = kGroupShape[i];
        g_phase_shape[i] = kPhaseShape[i];
    }
    for (int i = 0; i < 4; ++i) g_stream_shape[i] = kStreamShape[i];
    g_stream_on = 1;
    for (int i = 0; i < 3; ++i) g_skip_margin[i] = kSkipMargin[i];
    g_pstream_shape = kPstreamShape;
    g_pstream_shape32 = kPstreamShape32;
    g_permt_shape = kPermTileShape;
}

extern "C" int ork_dm_set_shape(int which, int flags, int lds_bytes) {
    if (which >= 23 && which <= 25) {   // skip margin of controlled gates (63: never skip)
        if (flags < 0 || flags > 63 || lds_bytes != 0) return (int)hipErrorInvalidValue;
        g_skip_margin[which - 23] = flags;
        return 0;
    }
    if (lds_bytes < 0 || lds_bytes > 160 * 1024 || flags < 0 || flags > (which == 22 ? 255 : 127) ||
        ((flags >> 3) & 3) > 2)
        return (int)hipErrorInvalidValue;
    // bit 5: XCD runs of 32 groups instead of contiguous ranges; bit 6: whole-line writes (phase kernels)
    const Shape sh{(flags & 32) ? 2 : (flags & 1), (flags >> 1) & 1, lds_bytes, (flags >> 2) & 1, (flags >> 3) & 3,
                   (flags >> 6) & 1, (flags >> 7) & 1};
    switch (which) {
    case 0: g_group_shape[0] = g_group_shape[1] = sh; break;
    case 1: g_phase_shape[0] = g_phase_shape[1] = sh; break;
    case 2: g_group_shape[0] = sh; break;
    case 3: g_group_shape[1] = sh; break;
    case 4: g_phase_shape[0] = sh; break;
    case 5: g_phase_shape[1] = sh; break;
    case 6: g_stream_shape[0] = sh; break;
    case 7: g_stream_shape[1] = sh; break;
    case 8: g_stream_on = lds_bytes != 0; break;
    case 9: g_pstream_shape = sh; break;
    case 18: g_pstream_shape32 = sh; break;
    case 22: g_permt_shape = sh; break;
    case 15: g_stream_shape[2] = sh; break;
    case 16: g_stream_shape[3] = sh; break;
    default: return (int)hipErrorInvalidValue;
    }
    return 0;
}

extern "C" int ork_dm_group(void *stream, int prec, int layout, int log_tile, void *rho, int n, int opk,
                            const ork_mono_t *m, const double *mat) {
    const int k = m->k;
    // only the two density-matrix layouts, with distinct gate qubits inside the state
    if ((layout != ORK_PACKED && layout != ORK_TILED) || k < 1 || k > 3 || n < 1 || n > 31 || log_tile < 1 ||
        log_tile > LW_MAX)
        return (int)hipErrorInvalidValue;
    for (int j = 0; j < k; ++j) {
        if (m->qs[j] < 0 || m->qs[j] >= n) return (int)hipErrorInvalidValue;
        for (int i = 0; i < j; ++i)
            if (m->qs[i] == m->qs[j]) return (int)hipErrorInvalidValue;
    }
    const DmPlan p = make_plan(layout, log_tile, n, m->qs, k);
    DmOp op;
    op.k = k;
    // LDS coordinate bit of each gate argument
    for (int j = 0; j < 3; ++j) op.lbit[j] = 0;
    for (int j = 0; j < k; ++j) {
        const int q = m->qs[j];
        if (q < p.LB) op.lbit[j] = q;
        else {
            int rank = 0;
            for (int i = 0; i < k; ++i) rank += (m->qs[i] >= p.LB && m->qs[i] < q);
            op.lbit[j] = p.LB + rank;
        }
    }
    for (int j = 0; j < 3; ++j) op.srt[j] = op.lbit[j];
    for (int a = 0; a < k; ++a)
        for (int b = a + 1; b < k; ++b)
            if (op.srt[b] < op.srt[a]) { int t = op.srt[a]; op.srt[a] = op.srt[b]; op.srt[b] = t; }
    for (int a = 0; a < 8; ++a) { op.pinv[a] = a; op.ph[a] = 0; }
    op.still = 0;
    op.ctl = op.pu01 = op.pu10 = op.skip = 0;
    op.cmask = 0;
    if (opk == ORK_DM_MONO) {
        for (int a = 0; a < (1 << k); ++a) {
            if (m->perm[a] < 0 || m->perm[a] >= (1 << k)) return (int)hipErrorInvalidValue;
            op.pinv[m->perm[a]] = a;
            op.ph[a] = m->ph[a] & 3;
        }
        // entry (a, b) of a 2^k block receives rho(pinv a, pinv b) i^(ph[pinv a] - ph[pinv b])
        for (int a = 0; a < (1 << k); ++a)
            for (int b = 0; b < (1 << k); ++b)
                if (op.pinv[a] == a && op.pinv[b] == b && ((op.ph[a] - op.ph[b]) & 3) == 0)
                    op.still |= (u64)1 << (a * 8 + b);
    }
    const int nm = (opk == ORK_DM_SOP) ? 16 : (opk == ORK_DM_DENSE ? (1 << (2 * k)) : 0);
    for (int i = 0; i < 64; ++i) op.m[i] = make_double2(0.0, 0.0);
    for (int i = 0; i < nm; ++i) op.m[i] = make_double2(mat[2 * i], mat[2 * i + 1]);
    op.s2 = opk == ORK_DM_SOP2 ? (const cd *)mat : nullptr;
    if (opk == ORK_DM_SOP2 && k != 2) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    // a controlled gate with one target (cx, cy, ccx) runs through the streamed 1-qubit kernels on that
    // target.  Where every control qubit lies far enough above the bits of a 128 B line they move only
    // the quads that change (cx 3/4 of the entries, ccx 7/16; 14 qubits: tiled cx 5.0-5.8 TB/s, packed
    // cx 4.4-4.9, ccx 3.5-4.3, on those bytes), elsewhere every entry once, which still beats the
    // in-tile and tile-group kernels.  With every qubit at or above the tile, whole tiles move faster on
    // the tiled layout (ccx 4.6-5.6 against 3.6-4.7 TB/s, cx +2 %, complex64 cx +11 %); on the packed
    // layout whole blocks lose to streaming (cx 3.9 against 4.3 TB/s) unless its quads cannot be skipped
    // (shape 22 flag bit 6 sends packed cx / cy / ccx to whole blocks anyway)
    int tq, pu01, pu10;
    u64 cm;
    if (opk == ORK_DM_MONO && g_stream_on && !g_permt_shape.col && controlled_single_target(m, tq, pu01, pu10, cm) &&
        ((layout == ORK_PACKED && n >= 6) || (layout == ORK_TILED && n >= log_tile))) {
        DmOp c = op;
        c.ctl = 1;
        c.cmask = cm;
        c.pu01 = pu01;
        c.pu10 = pu10;
        const int line_bits = prec == ORK_F32 ? 4 : 3, lowc = cm ? __builtin_ctzll(cm) : 64;
        c.skip = lowc >= line_bits + g_skip_margin[layout == ORK_TILED ? 1 : (m->k == 3 ? 2 : 0)];
        const bool tiles = perm_tiles_ok(layout, log_tile, n, m) &&
                           (layout == ORK_TILED || !c.skip || g_permt_shape.wl != 0);
        if (!tiles) {
            if (layout == ORK_PACKED)
                return prec == ORK_F32 ? launch_pstream<float2>(st, rho, n, tq, c)
                                       : launch_pstream<double2>(st, rho, n, tq, c);
            return prec == ORK_F32 ? launch_stream<float2>(st, rho, log_tile, n, tq, c)
                                   : launch_stream<double2>(st, rho, log_tile, n, tq, c);
        }
    }
    if (opk == ORK_DM_MONO && perm_intile_ok(layout, log_tile, n, m)) {
        if (prec == ORK_F32)
            return layout == ORK_PACKED ? launch_perm_intile<float2, ORK_PACKED>(st, rho, n, m)
                                        : launch_perm_intile<float2, ORK_TILED>(st, rho, n, m);
        return layout == ORK_PACKED ? launch_perm_intile<double2, ORK_PACKED>(st, rho, n, m)
                                    : launch_perm_intile<double2, ORK_TILED>(st, rho, n, m);
    }
    const int pv = opk == ORK_DM_MONO ? perm_tiles_v(layout, log_tile, n, prec, m, true) : -1;
    if (pv >= 0) {
        if (prec == ORK_F32)
            return layout == ORK_PACKED ? launch_perm_tiles<float2, ORK_PACKED>(st, rho, n, log_tile, m, pv)
                                        : launch_perm_tiles<float2, ORK_TILED>(st, rho, n, log_tile, m, pv);
        return layout == ORK_PACKED ? launch_perm_tiles<double2, ORK_PACKED>(st, rho, n, log_tile, m, pv)
                                    : launch_perm_tiles<double2, ORK_TILED>(st, rho, n, log_tile, m, pv);
    }
    if (layout == ORK_PACKED && opk == ORK_DM_SOP && g_stream_on && n >= 6)
        return prec == ORK_F32 ? launch_pstream<float2>(st, rho, n, m->qs[0], op)
                               : launch_pstream<double2>(st, rho, n, m->qs[0], op);
    if (layout == ORK_TILED && opk == ORK_DM_SOP && g_stream_on && n >= log_tile && log_tile <= LW_MAX &&
        log_tile >= 1)
        return prec == ORK_F32 ? launch_stream<float2>(st, rho, log_tile, n, m->qs[0], op)
                               : launch_stream<double2>(st, rho, log_tile, n, m->qs[0], op);
    if (prec == ORK_F32)
        return layout == ORK_PACKED ? launch_group<float2, ORK_PACKED>(st, rho, p, opk, op)
                                    : launch_group<float2, ORK_TILED>(st, rho, p, opk, op);
    return layout == ORK_PACKED ? launch_group<double2, ORK_PACKED>(st, rho, p, opk, op)
                                : launch_group<double2, ORK_TILED>(st, rho, p, opk, op);
}

extern "C" int ork_dm_phase(void *stream, int prec, int layout, int log_tile, void *rho, int n, int k, const int *qs,
                            const double *f) {
    if ((layout != ORK_PACKED && layout != ORK_TILED) || k < 1 || k > 3 || n < 1 || n > 31 || log_tile < 1 ||
        log_tile > LW_MAX)
        return (int)hipErrorInvalidValue;
    PhaseArgs a;
    a.k = k;
    for (int j = 0; j < 3; ++j) a.qs[j] = j < k ? qs[j] : 0;
    for (int x = 0; x < 8; ++x) {
        a.skip[x] = 0;
        for (int y = 0; y < 8; ++y) {
            if (x >= (1 << k) || y >= (1 << k)) { a.f[x * 8 + y] = make_double2(1.0, 0.0); a.skip[x] |= 1u << y; continue; }
            const double ar = f[2 * x], ai = f[2 * x + 1], br = f[2 * y], bi = -f[2 * y + 1];
            const double re = ar * br - ai * bi, im = ar * bi + ai * br;
            a.f[x * 8 + y] = make_double2(re, im);
            if (re == 1.0 && im == 0.0) a.skip[x] |= 1u << y;
        }
    }
    a.skipm = 0;
    for (int x = 0; x < 8; ++x)
        for (int y = 0; y < 8; ++y)
            if ((a.skip[x] >> y) & 1) a.skipm |= (u64)1 << (x * 8 + y);
    return launch_phase<0>((hipStream_t)stream, prec, layout, log_tile, n, rho, a, nullptr);
}

extern "C" int ork_dm_exp_diag(void *stream, int prec, int layout, int log_tile, void *rho, int n,
                               const double *diag, double t, void *phase) {
    if ((layout != ORK_PACKED && layout != ORK_TILED) || n < 1 || n > 31 || log_tile < 1 || log_tile > LW_MAX)
        return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    const u64 dim = (u64)1 << n;
    const unsigned g1 = (unsigned)((dim + DM_THREADS - 1) / DM_THREADS);
    sincos_kernel<<<g1, DM_THREADS, 0, st>>>((double2 *)phase, diag, dim, t);
    ORK_LAUNCH_CHECK();
    PhaseArgs a;
    a.k = 0;
    a.qs[0] = a.qs[1] = a.qs[2] = 0;   // per-index phases (mode 1): no gate qubits, no skip table
    a.skipm = 0;
    return launch_phase<1>(st, prec, layout, log_tile, n, rho, a, (const double2 *)phase);
}

extern "C" int ork_dm_mean(void *stream, int prec, int layout, int log_tile, const void *rho, int n,
                           const double *obs, double *part, double *out) {
    const DmPlan p = tile_plan(layout, log_tile, n);
    hipStream_t st = (hipStream_t)stream;
    u64 nb = (p.dim + DM_THREADS - 1) / DM_THREADS;
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    if (prec == ORK_F32) {
        if (layout == ORK_PACKED) dm_mean_kernel<float2, ORK_PACKED><<<(unsigned)nb, DM_THREADS, 0, st>>>((const float2 *)rho, p, obs, part);
        else dm_mean_kernel<float2, ORK_TILED><<<(unsigned)nb, DM_THREADS, 0, st>>>((const float2 *)rho, p, obs, part);
    } else {
        if (layout == ORK_PACKED) dm_mean_kernel<double2, ORK_PACKED><<<(unsigned)nb, DM_THREADS, 0, st>>>((const double2 *)rho, p, obs, part);
        else dm_mean_kernel<double2, ORK_TILED><<<(unsigned)nb, DM_THREADS, 0, st>>>((const double2 *)rho, p, obs, part);
    }
    ORK_LAUNCH_CHECK();
    dm_finish_kernel<<<1, DM_THREADS, 0, st>>>(part, (int)nb);
    ORK_LAUNCH_CHECK();
    hipError_t e = hipMemcpyAsync(out, part + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return (int)e;
    return (int)hipStreamSynchronize(st);
}
