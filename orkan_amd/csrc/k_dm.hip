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
    int xcd;        // remap group ids so consecutive groups run on one XCD
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
__device__ __forceinline__ u64 remap_group(u64 w, u64 n, int xcd) {
    if (!xcd || gridDim.x < n) return w;
    const u64 per = n >> 3;
    if (w < (per << 3)) return (w & 7) * per + (w >> 3);
    return w;
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

// LDS entry (x, y) of the updated group; permutations with phases are applied
// here as a gather: rho'(perm a, perm b) = i^(ph[a] - ph[b]) rho(a, b)
template <int OP>
__device__ __forceinline__ cd lds_get(const cd *lds, const DmOp &op, int x, int y) {
    if constexpr (OP == ORK_DM_MONO) {
        const int k = op.k;
        const u64 mask = spread_bits((1 << k) - 1, op.lbit, k);
        const int ai = op.pinv[gather_bits((u64)x, op.lbit, k)], bi = op.pinv[gather_bits((u64)y, op.lbit, k)];
        const int xs = (int)(((u64)x & ~mask) | spread_bits(ai, op.lbit, k));
        const int ys = (int)(((u64)y & ~mask) | spread_bits(bi, op.lbit, k));
        return cipow(lds[xs * LDS_STRIDE + ys], op.ph[ai] - op.ph[bi]);
    } else {
        return lds[x * LDS_STRIDE + y];
    }
}

template <typename S, int LAYOUT, int OP>
__global__ __launch_bounds__(DM_THREADS) void dm_group_kernel(S *__restrict__ rho, DmPlan p, DmOp op) {
    __shared__ cd lds[W_MAX * LDS_STRIDE];
    const int LB = p.LB, g = p.g, LW = p.LW;
    const int B = 1 << LB, W = 1 << LW, NT = 1 << g;
    const int total = W * W;
    const int tid = threadIdx.x;

    for (u64 w0 = blockIdx.x; w0 < p.ngroups; w0 += gridDim.x) {
        const u64 w = remap_group(w0, p.ngroups, p.xcd);
        u64 R, C;
        tri_decode(w, R, C);
        const u64 TRb = ins0k(R, p.gpos, g), TCb = ins0k(C, p.gpos, g);
        const bool diag_group = (R == C);

        // ---------------- load ----------------
        for (int t = tid; t < total; t += DM_THREADS) {
            const int ti = t >> (2 * LB), e = t & (B * B - 1);
            const int i = ti >> g, j = ti & (NT - 1);
            const u64 TR = TRb | spread_bits(i, p.gpos, g), TC = TCb | spread_bits(j, p.gpos, g);
            const int sa = e >> LB, sb = e & (B - 1);                  // slow, fast
            const int sr = (LAYOUT == ORK_TILED) ? sa : sb;            // row-major vs column-major
            const int sc = (LAYOUT == ORK_TILED) ? sb : sa;
            if (tile_stored<LAYOUT>(p, TR, TC)) {
                if (LAYOUT == ORK_PACKED && TR == TC && sr < sc) continue;
                const cd v = ldc(rho + elem_off<LAYOUT>(p, TR * B + sr, TC * B + sc));
                lds[(i * B + sr) * LDS_STRIDE + j * B + sc] = v;
                if (LAYOUT == ORK_PACKED && TR == TC && sr > sc)
                    lds[(i * B + sc) * LDS_STRIDE + j * B + sr] = cconj(v);
            } else {
                // stored mirror tile (TC, TR): its entry (sr, sc) is conj of ours at (sc, sr)
                const cd v = ldc(rho + elem_off<LAYOUT>(p, TC * B + sr, TR * B + sc));
                lds[(i * B + sc) * LDS_STRIDE + j * B + sr] = cconj(v);
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
            const u64 TR = TRb | spread_bits(i, p.gpos, g), TC = TCb | spread_bits(j, p.gpos, g);
            const int sa = e >> LB, sb = e & (B - 1);
            const int sr = (LAYOUT == ORK_TILED) ? sa : sb;
            const int sc = (LAYOUT == ORK_TILED) ? sb : sa;
            const bool direct = tile_stored<LAYOUT>(p, TR, TC) && !(LAYOUT == ORK_PACKED && TR == TC && sr < sc);
            const bool mirror = !diag_group && tile_stored<LAYOUT>(p, TC, TR);
            if (direct) stc(rho + elem_off<LAYOUT>(p, TR * B + sr, TC * B + sc), lds_get<OP>(lds, op, i * B + sr, j * B + sc));
            if (mirror) stc(rho + elem_off<LAYOUT>(p, TC * B + sr, TR * B + sc), cconj(lds_get<OP>(lds, op, i * B + sc, j * B + sr)));
        }
        __syncthreads();
    }
}

// ---- elementwise phase kernels: rho_rc *= F(r) conj(F(c)) --------------------------------
// Each workgroup takes one stored B x B tile (TR >= TC).  `mode` 0: gate phases
// from a 2^k table with a per-pair skip mask; mode 1: per-index phases (exp_diag),
// diagonal entries untouched.
struct PhaseArgs {
    int k;
    int qs[3];
    unsigned skip[8];   // bit b of skip[a]: entries with row index a, column index b unchanged
    cd f[64];           // f[a * 8 + b] = F[a] conj(F[b])
};

template <typename S, int LAYOUT, int MODE>
__global__ __launch_bounds__(DM_THREADS) void dm_phase_kernel(S *__restrict__ rho, DmPlan p, PhaseArgs a,
                                                              const double2 *__restrict__ cs) {
    const int LB = p.LB, B = 1 << LB;
    for (u64 w0 = blockIdx.x; w0 < p.ngroups; w0 += gridDim.x) {
        u64 TR, TC;
        tri_decode(remap_group(w0, p.ngroups, p.xcd), TR, TC);
        for (int e = threadIdx.x; e < B * B; e += DM_THREADS) {
            const int sa = e >> LB, sb = e & (B - 1);
            const int sr = (LAYOUT == ORK_TILED) ? sa : sb;
            const int sc = (LAYOUT == ORK_TILED) ? sb : sa;
            const u64 r = TR * B + sr, c = TC * B + sc;
            if (LAYOUT == ORK_PACKED && r < c) continue;
            if (r == c) continue;
            S *ptr = rho + elem_off<LAYOUT>(p, r, c);
            if constexpr (MODE == 0) {
                const int ar = gather_bits(r, a.qs, a.k), ac = gather_bits(c, a.qs, a.k);
                if (a.skip[ar] >> ac & 1) continue;
                stc(ptr, cmul(a.f[ar * 8 + ac], ldc(ptr)));
            } else {
                // e^{-i(d_r - d_c)t} = (cr - i sr)(cc + i sc)   (src/circ/circ_packed.c:54-60)
                const double2 pr = cs[r], pc = cs[c];
                const double fre = pr.x * pc.x + pr.y * pc.y;
                const double fim = -pr.y * pc.x + pr.x * pc.y;
                const cd v = ldc(ptr);
                stc(ptr, cmk(fre * v.x - fim * v.y, fre * v.y + fim * v.x));
            }
        }
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
    p.ngroups = T * (T + 1) / 2;
    return p;
}

// Launch shapes (residency cap through unused dynamic LDS, XCD remap); set
// through orkan_set_launch_shape, defaults from measurements on MI355X
// (measured: unlike the statevector streams these kernels want every resident
// workgroup they can get; the XCD remap helps the tiled layout only)
struct Shape { int xcd, lds; };
Shape g_group_shape[2] = {{0, 0}, {1, 0}};   // [packed, tiled]
Shape g_phase_shape = {1, 0};

unsigned grid_for(u64 ngroups) {
    const u64 cap = (u64)1 << 30;
    return (unsigned)(ngroups < cap ? ngroups : cap);
}

template <typename S, int LAYOUT>
int launch_group(hipStream_t st, void *rho, DmPlan p, int opk, const DmOp &op) {
    const unsigned grid = grid_for(p.ngroups);
    const Shape sh = g_group_shape[LAYOUT == ORK_TILED ? 1 : 0];
    p.xcd = sh.xcd;
    const int lds = sh.lds > (int)sizeof(cd) * W_MAX * LDS_STRIDE ? sh.lds - (int)sizeof(cd) * W_MAX * LDS_STRIDE : 0;
    switch (opk) {
    case ORK_DM_SOP: dm_group_kernel<S, LAYOUT, ORK_DM_SOP><<<grid, DM_THREADS, lds, st>>>((S *)rho, p, op); break;
    case ORK_DM_MONO: dm_group_kernel<S, LAYOUT, ORK_DM_MONO><<<grid, DM_THREADS, lds, st>>>((S *)rho, p, op); break;
    case ORK_DM_DENSE: dm_group_kernel<S, LAYOUT, ORK_DM_DENSE><<<grid, DM_THREADS, lds, st>>>((S *)rho, p, op); break;
    default: return (int)hipErrorInvalidValue;
    }
    ORK_LAUNCH_CHECK();
    return 0;
}

template <int MODE>
int launch_phase(hipStream_t st, int prec, int layout, void *rho, DmPlan p, const PhaseArgs &a,
                 const double2 *cs) {
    const unsigned grid = grid_for(p.ngroups);
    p.xcd = g_phase_shape.xcd;
    const int lds = g_phase_shape.lds;
    if (prec == ORK_F32) {
        if (layout == ORK_PACKED) dm_phase_kernel<float2, ORK_PACKED, MODE><<<grid, DM_THREADS, lds, st>>>((float2 *)rho, p, a, cs);
        else dm_phase_kernel<float2, ORK_TILED, MODE><<<grid, DM_THREADS, lds, st>>>((float2 *)rho, p, a, cs);
    } else {
        if (layout == ORK_PACKED) dm_phase_kernel<double2, ORK_PACKED, MODE><<<grid, DM_THREADS, lds, st>>>((double2 *)rho, p, a, cs);
        else dm_phase_kernel<double2, ORK_TILED, MODE><<<grid, DM_THREADS, lds, st>>>((double2 *)rho, p, a, cs);
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

extern "C" int ork_dm_set_shape(int which, int xcd, int lds_bytes) {
    if (lds_bytes < 0 || lds_bytes > 160 * 1024) return (int)hipErrorInvalidValue;
    if (which == 0) g_group_shape[0] = g_group_shape[1] = Shape{xcd ? 1 : 0, lds_bytes};
    else if (which == 1) g_phase_shape = Shape{xcd ? 1 : 0, lds_bytes};
    else return (int)hipErrorInvalidValue;
    return 0;
}

extern "C" int ork_dm_group(void *stream, int prec, int layout, int log_tile, void *rho, int n, int opk,
                            const ork_mono_t *m, const double *mat) {
    const int k = m->k;
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
    if (opk == ORK_DM_MONO) {
        for (int a = 0; a < (1 << k); ++a) {
            op.pinv[m->perm[a]] = a;
            op.ph[a] = m->ph[a] & 3;
        }
    }
    const int nm = (opk == ORK_DM_SOP) ? 16 : (opk == ORK_DM_DENSE ? (1 << (2 * k)) : 0);
    for (int i = 0; i < 64; ++i) op.m[i] = make_double2(0.0, 0.0);
    for (int i = 0; i < nm; ++i) op.m[i] = make_double2(mat[2 * i], mat[2 * i + 1]);
    hipStream_t st = (hipStream_t)stream;
    if (prec == ORK_F32)
        return layout == ORK_PACKED ? launch_group<float2, ORK_PACKED>(st, rho, p, opk, op)
                                    : launch_group<float2, ORK_TILED>(st, rho, p, opk, op);
    return layout == ORK_PACKED ? launch_group<double2, ORK_PACKED>(st, rho, p, opk, op)
                                : launch_group<double2, ORK_TILED>(st, rho, p, opk, op);
}

extern "C" int ork_dm_phase(void *stream, int prec, int layout, int log_tile, void *rho, int n, int k, const int *qs,
                            const double *f) {
    const DmPlan p = tile_plan(layout, log_tile, n);
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
    return launch_phase<0>((hipStream_t)stream, prec, layout, rho, p, a, nullptr);
}

extern "C" int ork_dm_exp_diag(void *stream, int prec, int layout, int log_tile, void *rho, int n,
                               const double *diag, double t, void *phase) {
    const DmPlan p = tile_plan(layout, log_tile, n);
    hipStream_t st = (hipStream_t)stream;
    const unsigned g1 = (unsigned)((p.dim + DM_THREADS - 1) / DM_THREADS);
    sincos_kernel<<<g1, DM_THREADS, 0, st>>>((double2 *)phase, diag, p.dim, t);
    ORK_LAUNCH_CHECK();
    PhaseArgs a;
    a.k = 0;
    return launch_phase<1>(st, prec, layout, rho, p, a, (const double2 *)phase);
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
