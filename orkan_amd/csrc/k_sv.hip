// k_sv.hip -- statevector kernels: single-qubit pair updates, multi-qubit
// permutation-with-phase updates, the exp_diag phase sweep, and the
// diagonal-observable reductions (mean, matel_diag).
//
// Formulas per gate follow the reference's pair operations
// (src/gate/gate_pure.c:66-267) so that results agree to the last bits; the
// memory traffic is the HBM stream every gate needs: the touched amplitudes are
// read once and written once, 16 B per complex128 amplitude.
#include "k_common.hpp"

namespace {

constexpr int SV_THREADS = 256;
constexpr int SV_UNROLL = 2;      // bases per thread in the permutation kernel (doubled for complex64)
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 2048;  // room for first-stage partials; their count depends only on the size and shape

struct Coef {
    double v[8];
};

template <int OP>
__device__ __forceinline__ void pair_op(cd &a, cd &b, const Coef &k) {
    if constexpr (OP == ORK_SV_MAT) {
        const cd m00 = cmk(k.v[0], k.v[1]), m01 = cmk(k.v[2], k.v[3]);
        const cd m10 = cmk(k.v[4], k.v[5]), m11 = cmk(k.v[6], k.v[7]);
        cd na = cmul(m00, a), nb = cmul(m10, a);
        cfma(na, m01, b);
        cfma(nb, m11, b);
        a = na; b = nb;
    } else if constexpr (OP == ORK_SV_H) {
        const cd sm = cadd(a, b), df = csub(a, b);
        a = cscale(sm, ORK_SQRT1_2); b = cscale(df, ORK_SQRT1_2);
    } else if constexpr (OP == ORK_SV_HY) {
        const cd df = csub(a, b), sm = cadd(a, b);
        a = cscale(df, ORK_SQRT1_2); b = cscale(sm, ORK_SQRT1_2);
    } else if constexpr (OP == ORK_SV_X) {
        const cd t = a; a = b; b = t;
    } else if constexpr (OP == ORK_SV_Y) {
        const cd na = cmk(b.y, -b.x), nb = cmk(-a.y, a.x);
        a = na; b = nb;
    } else if constexpr (OP == ORK_SV_RX) {
        const double c = k.v[0], s = k.v[1];
        const cd na = cmk(fma(c, a.x, s * b.y), fma(c, a.y, -(s * b.x)));
        const cd nb = cmk(fma(s, a.y, c * b.x), fma(-s, a.x, c * b.y));
        a = na; b = nb;
    } else if constexpr (OP == ORK_SV_RY) {
        const double c = k.v[0], s = k.v[1];
        const cd na = cmk(fma(c, a.x, -(s * b.x)), fma(c, a.y, -(s * b.y)));
        const cd nb = cmk(fma(s, a.x, c * b.x), fma(s, a.y, c * b.y));
        a = na; b = nb;
    } else if constexpr (OP == ORK_SV_RZ) {
        const double c = k.v[0], s = k.v[1];
        const cd na = cmk(fma(c, a.x, s * a.y), fma(c, a.y, -(s * a.x)));
        const cd nb = cmk(fma(c, b.x, -(s * b.y)), fma(c, b.y, s * b.x));
        a = na; b = nb;
    }
}

// ops that change only the |1> amplitude of the pair
template <int OP>
__device__ __forceinline__ cd one_op(cd b, const Coef &k) {
    if constexpr (OP == ORK_SV_P) {
        const double c = k.v[0], s = k.v[1];
        return cmk(fma(c, b.x, -(s * b.y)), fma(s, b.x, c * b.y));
    } else if constexpr (OP == ORK_SV_Z) {
        return cneg(b);
    } else if constexpr (OP == ORK_SV_S) {
        return cmk(-b.y, b.x);
    } else if constexpr (OP == ORK_SV_SDG) {
        return cmk(b.y, -b.x);
    } else if constexpr (OP == ORK_SV_T) {
        return cmk((b.x - b.y) * ORK_SQRT1_2, (b.x + b.y) * ORK_SQRT1_2);
    } else {
        return cmk((b.x + b.y) * ORK_SQRT1_2, (b.y - b.x) * ORK_SQRT1_2);
    }
}

// Launch shape of the streaming kernels.  Measured on MI355X (28-qubit
// states): in-place read-modify-write streams run fastest with fewer
// workgroups resident per CU than the hardware allows (the surplus of requests
// thrashes DRAM rows) and with consecutive blocks kept on one XCD.  Residency
// is capped by requesting unused dynamic LDS.
struct SvShape {
    int xcd;   // remap block ids so consecutive chunks run on the same XCD
    int u;     // pairs per thread (1 or 2; doubled for complex64 storage)
    int lds;   // dynamic LDS bytes per workgroup (0: no cap)
    int nt;    // nontemporal loads and stores
};

// best shape per target qubit of a 2^28-amplitude state (sweep of xcd x u x LDS cap x nontemporal).
// Nontemporal accesses gain 6-10 % from target 3 up; below that a pair shares its 128 B line with
// the next pair, which the cache then has to supply twice, so targets 0-2 take them through the
// lane-exchange kernel (6.68 against 6.36 TB/s).
static const int kPairShapeHost[28][4] = {
    {1, 1, 32768, 1}, {1, 1, 32768, 1}, {1, 1, 32768, 1}, {1, 1, 53248, 1}, {1, 1, 53248, 1}, {1, 1, 53248, 1},
    {1, 1, 53248, 1}, {1, 1, 53248, 1}, {1, 1, 53248, 1}, {1, 1, 53248, 1}, {1, 1, 53248, 1}, {1, 1, 53248, 1},
    {1, 1, 53248, 1}, {0, 1, 53248, 1}, {0, 1, 53248, 1}, {0, 1, 53248, 1}, {0, 1, 53248, 1}, {0, 1, 53248, 1},
    {1, 1, 40960, 1}, {1, 1, 40960, 1}, {1, 1, 53248, 1}, {1, 1, 53248, 1}, {1, 1, 40960, 1}, {0, 1, 53248, 1},
    {0, 1, 53248, 1}, {1, 1, 53248, 1}, {1, 1, 40960, 1}, {1, 1, 40960, 1}};

// complex64 storage (pairs per thread and LDS cap as launched): twice the pairs and half the cap of the
// complex128 shape, except where a sweep of complex64 shapes on 2^28 amplitudes found better
// (several targets above 17: 2 pairs per thread in plain order, 77.7 -> 79.1 % over the 28; targets
// 0-3 through the lane-exchange kernel, uncapped: 6.35 against 5.5-5.85 TB/s)
static const int kPairShapeF32[28][4] = {
    {0, 2, 0, 1}, {0, 2, 0, 1}, {0, 2, 0, 1}, {0, 2, 0, 1}, {1, 2, 26624, 1}, {1, 2, 26624, 1},
    {1, 2, 26624, 1}, {1, 2, 26624, 1}, {1, 2, 26624, 1}, {1, 2, 26624, 1}, {1, 2, 26624, 1}, {1, 2, 26624, 1},
    {1, 2, 26624, 1}, {0, 2, 26624, 1}, {0, 2, 26624, 1}, {0, 2, 26624, 1}, {0, 2, 26624, 1}, {0, 2, 26624, 1},
    {0, 2, 20480, 1}, {0, 2, 20480, 1}, {0, 2, 20480, 1}, {0, 4, 20480, 1}, {1, 2, 20480, 1}, {1, 2, 20480, 1},
    {0, 2, 26624, 1}, {0, 2, 20480, 1}, {0, 4, 32768, 1}, {0, 2, 20480, 1}};

// tuning override of the table for every target qubit (used when g_pair_override.u != 0)
static SvShape g_pair_override{0, 0, 0, 0};
// nontemporal accesses switched off per kernel family (tuning hook): bit 0 1-qubit pair kernels,
// bit 1 permutations (cx ...), bit 2 exp_diag, bit 3 reductions
static int g_nt_off = 0;
// states below this size may stay in the 256 MB last-level cache between gates, where streaming
// (cache-bypassing) accesses would only evict them
constexpr u64 kNtMinBytes = (u64)512 << 20;
static u64 g_nt_min_bytes = kNtMinBytes;

// launch shape for target q; f32: complex64 storage
SvShape pair_shape(int q, bool f32) {
    SvShape s{1, 1, 53248, 0};
    if (g_pair_override.u || q < 0 || q >= 28) {
        if (g_pair_override.u) s = g_pair_override;
        // complex64 storage moves half the bytes per pair: twice the pairs per thread (4 at most,
        // the largest instantiated shape) and twice the resident workgroups keep the bytes in flight
        if (f32) {
            s.u = s.u < 4 ? 2 * s.u : 4;
            s.lds /= 2;
        }
    } else if (f32) {
        s = SvShape{kPairShapeF32[q][0], kPairShapeF32[q][1], kPairShapeF32[q][2], kPairShapeF32[q][3]};
    } else {
        s = SvShape{kPairShapeHost[q][0], kPairShapeHost[q][1], kPairShapeHost[q][2], kPairShapeHost[q][3]};
    }
    if (g_nt_off & 1) s.nt = 0;
    return s;
}

// block id -> chunk id; with xcd, consecutive chunks go to one XCD (blocks are
// dispatched round-robin over the 8 XCDs)
__device__ __forceinline__ u64 chunk_of(int xcd) {
    const unsigned b = blockIdx.x;
    if (xcd) {
        const unsigned per = gridDim.x >> 3;
        if (b < (per << 3)) return (u64)(b & 7) * per + (b >> 3);
    }
    return b;
}

template <typename S, int OP, int U, int XCD, int NT>
__global__ __launch_bounds__(SV_THREADS) void sv_pair_kernel(S *__restrict__ psi, u64 npairs, int q, Coef k) {
    const u64 base = chunk_of(XCD) * (SV_THREADS * U) + threadIdx.x;
    const u64 stride = (u64)1 << q;
    cd a[U], b[U];
    u64 i0[U];
    if constexpr (sizeof(S) == 8) {
        // complex64: raw loads without a condition (past the end: psi[0]), converted after the batch; a load
        // inside the condition waited for its own conversion, one pair in flight per thread
        S ra[U], rb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 kk = base + (u64)u * SV_THREADS;
            i0[u] = ins0(kk, q);
            const u64 o = kk < npairs ? i0[u] : 0;
            ra[u] = ldr<NT>(psi + o);
            rb[u] = ldr<NT>(psi + (kk < npairs ? o + stride : 0));
        }
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
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 kk = base + (u64)u * SV_THREADS;
            i0[u] = ins0(kk, q);
            if (kk < npairs) {
                a[u] = ldx<NT>(psi + i0[u]);
                b[u] = ldx<NT>(psi + i0[u] + stride);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 kk = base + (u64)u * SV_THREADS;
        if (kk < npairs) {
            pair_op<OP>(a[u], b[u], k);
            stx<NT>(psi + i0[u], a[u]);
            stx<NT>(psi + i0[u] + stride, b[u]);
        }
    }
}

// Targets whose pair shares a 128 B line: one amplitude per lane, lanes contiguous so that every
// instruction covers whole lines (nontemporal); the partner amplitude comes from lane ^ 2^q.  Both
// lanes of a pair form the same update and each stores its own amplitude.
template <typename S, int OP, int U, int XCD>
__global__ __launch_bounds__(SV_THREADS) void sv_pair_lane_kernel(S *__restrict__ psi, u64 dim, int q, Coef k) {
    const u64 base = chunk_of(XCD) * (SV_THREADS * U) + threadIdx.x;
    const int mb = 1 << q;
    const bool lo = !(threadIdx.x & mb);   // blocks start on multiples of 256: bit q of x is bit q of the lane
    cd v[U];
    if constexpr (sizeof(S) == 8) {   // complex64: raw loads, converted after the batch (as in sv_pair_kernel)
        S r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 x = base + (u64)u * SV_THREADS;
            r[u] = ldr<1>(psi + (x < dim ? x : 0));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) hold(r[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = to_cd(r[u]);
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 x = base + (u64)u * SV_THREADS;
            v[u] = x < dim ? ldx<1>(psi + x) : cmk(0.0, 0.0);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 x = base + (u64)u * SV_THREADS;
        const cd w = shfl_xor_cd(v[u], mb);   // every lane takes part
        cd a = sel(lo, v[u], w), b = sel(lo, w, v[u]);
        pair_op<OP>(a, b, k);
        if (x < dim) stx<1>(psi + x, sel(lo, a, b));
    }
}

template <typename S, int OP, int U, int XCD, int NT>
__global__ __launch_bounds__(SV_THREADS) void sv_one_kernel(S *__restrict__ psi, u64 npairs, int q, Coef k) {
    const u64 base = chunk_of(XCD) * (SV_THREADS * U) + threadIdx.x;
    const u64 stride = (u64)1 << q;
    cd b[U];
    u64 i1[U];
    if constexpr (sizeof(S) == 8) {   // complex64: raw loads (past the end: psi[0]), converted after the batch
        S r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 kk = base + (u64)u * SV_THREADS;
            i1[u] = ins0(kk, q) + stride;
            r[u] = ldr<NT>(psi + (kk < npairs ? i1[u] : 0));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) hold(r[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) b[u] = to_cd(r[u]);
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 kk = base + (u64)u * SV_THREADS;
            i1[u] = ins0(kk, q) + stride;
            if (kk < npairs) b[u] = ldx<NT>(psi + i1[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 kk = base + (u64)u * SV_THREADS;
        if (kk < npairs) stx<NT>(psi + i1[u], one_op<OP>(b[u], k));
    }
}

// ---- multi-qubit permutation with phases -------------------------------------------------
struct MonoArgs {
    int k, nt;
    int pos[3];       // ascending gate qubits
    u64 src[8];       // offsets of the touched amplitudes inside a group
    u64 dst[8];       // where each touched amplitude goes
    int ph[8];        // times i^ph
};

template <typename S, int U, int NT>
__global__ __launch_bounds__(SV_THREADS) void sv_mono_kernel(S *__restrict__ psi, u64 nbase, MonoArgs m) {
    const u64 b0 = chunk_of(1) * (SV_THREADS * U) + threadIdx.x;
    // fixed-bound loops with guards keep v[][] in registers (no dynamic indexing)
    cd v[U][8];
    u64 base[U];
    if constexpr (sizeof(S) == 8) {
        // complex64: raw loads (past the end: psi[0]), converted after the batch; a load inside the
        // per-lane condition waited for its own conversion, one amplitude in flight per thread
        S r[U][8];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 b = b0 + (u64)u * SV_THREADS;
            base[u] = ins0k(b, m.pos, m.k);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                r[u][j].x = r[u][j].y = 0.0f;
                if (j < m.nt) r[u][j] = ldr<NT>(psi + (b < nbase ? (base[u] | m.src[j]) : 0));
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < 8; ++j) hold(r[u][j]);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < 8; ++j) v[u][j] = to_cd(r[u][j]);
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 b = b0 + (u64)u * SV_THREADS;
            base[u] = ins0k(b, m.pos, m.k);
            if (b < nbase) {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (j < m.nt) v[u][j] = ldx<NT>(psi + (base[u] | m.src[j]));
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 b = b0 + (u64)u * SV_THREADS;
        if (b < nbase) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < m.nt) stx<NT>(psi + (base[u] | m.dst[j]), cipow(v[u][j], m.ph[j]));
        }
    }
}

// Every gate qubit below 6: one amplitude per lane in memory order, so that every instruction covers
// whole 128 B lines.  Lane l's new amplitude is i^ph[s] times the old amplitude of lane l ^ d, where the
// gate bits a of l pick the source s = pinv[a] and d spreads a ^ s over the gate qubits: the gate bits
// are lane bits because workgroups start on multiples of 256 amplitudes.  A line in which no amplitude
// changes is neither read nor written (the sources of changed amplitudes are changed amplitudes); with
// WL the other lines are written whole, unchanged amplitudes as read, else only changed amplitudes.
struct PermLaneArgs {
    int k;
    int qs[3];
    int pinv[8];   // gate bits a take the amplitude whose gate bits are pinv[a]
    int ph[8];     // ... times i^ph[pinv[a]]
    // the same tables packed (pinv[a] at bits 3a..3a+2, ph[s] at bits 2s, 2s+1): read per lane from the
    // arrays, each lookup was a memory load waited for
    unsigned pinv_bits, ph_bits;
};

template <typename S, int U, int XCD, int NT, int WL>
__global__ __launch_bounds__(SV_THREADS) void sv_perm_lane_kernel(S *__restrict__ psi, u64 dim, PermLaneArgs m) {
    constexpr int G = 128 / (int)sizeof(S);   // amplitudes per line: lanes (lane & ~(G - 1)) .. + G - 1
    const u64 base = chunk_of(XCD) * (SV_THREADS * U) + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int a = 0;
    for (int j = 0; j < m.k; ++j) a |= ((lane >> m.qs[j]) & 1) << j;
    const int s = (int)((m.pinv_bits >> (3 * a)) & 7);
    int d = 0;
    for (int j = 0; j < m.k; ++j) d |= (((a ^ s) >> j) & 1) << m.qs[j];
    const int src = lane ^ d, ph = (int)((m.ph_bits >> (2 * s)) & 3);
    const bool chg = (s != a) || ph != 0;
    const unsigned long long cm = __ballot(chg);   // depends on the lane only, the same for every chunk
    const bool line = ((cm >> (lane & ~(G - 1))) & ((1ull << G) - 1)) != 0;
    const bool wr = WL ? line : chg;
    // raw loads without a condition (a lane outside the changed lines reads psi[0], a value no lane uses),
    // converted after the batch: a complex64 load inside the condition waited for its own conversion, one
    // load in flight per thread
    S r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 x = base + (u64)u * SV_THREADS;
        r[u] = ldr<NT>(psi + ((line && x < dim) ? x : 0));
    }
    if constexpr (sizeof(S) == 8) {
#pragma unroll
        for (int u = 0; u < U; ++u) hold(r[u]);
    }
    cd v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = to_cd(r[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 x = base + (u64)u * SV_THREADS;
        const cd w = shfl_cd(v[u], src);   // every lane takes part
        if (wr && x < dim) stx<NT>(psi + x, chg ? cipow(w, ph) : v[u]);
    }
}

// Complex64 storage with no gate qubit at bit 0: amplitudes 2p and 2p + 1 move together, so the
// permutation runs on 16 B elements (two amplitudes each, gate qubits one lower) -- one 16 B access
// per two amplitudes instead of two 8 B accesses; the phases i^k act on each half exactly.
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v cipow2(f4v a, int k) {
    f4v r = a;
    switch (k & 3) {
    case 1: r.x = -a.y; r.y = a.x; r.z = -a.w; r.w = a.z; break;
    case 2: r = -a; break;
    case 3: r.x = a.y; r.y = -a.x; r.z = a.w; r.w = -a.z; break;
    default: break;
    }
    return r;
}
template <int NT> __device__ __forceinline__ f4v ld4(const f4v *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <int NT> __device__ __forceinline__ void st4(f4v *p, f4v v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int U, int NT>
__global__ __launch_bounds__(SV_THREADS) void sv_mono_x2_kernel(f4v *__restrict__ psi2, u64 nbase, MonoArgs m) {
    const u64 b0 = chunk_of(1) * (SV_THREADS * U) + threadIdx.x;
    f4v v[U][8];
    u64 base[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 b = b0 + (u64)u * SV_THREADS;
        base[u] = ins0k(b, m.pos, m.k);
        if (b < nbase) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < m.nt) v[u][j] = ld4<NT>(psi2 + (base[u] | m.src[j]));
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const u64 b = b0 + (u64)u * SV_THREADS;
        if (b < nbase) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < m.nt) st4<NT>(psi2 + (base[u] | m.dst[j]), cipow2(v[u][j], m.ph[j]));
        }
    }
}

// ---- exp_diag: psi_x *= exp(-i diag[x] t) ------------------------------------------------
template <typename S, int NT, int U>
__global__ __launch_bounds__(SV_THREADS) void sv_exp_diag_kernel(S *__restrict__ psi, const double *__restrict__ diag,
                                                                 u64 dim, double t) {
    if constexpr (U == 1) {
        // one amplitude per thread: the diagonal entry's load, then its sincos while psi's load is
        // in flight
        const u64 x = (u64)blockIdx.x * SV_THREADS + threadIdx.x;
        if (x >= dim) return;
        double s, c;
        const double d = NT ? __builtin_nontemporal_load(diag + x) : diag[x];
        sincos(d * t, &s, &c);
        const cd v = ldx<NT>(psi + x);
        stx<NT>(psi + x, cmk(c * v.x + s * v.y, c * v.y - s * v.x));
    } else {
        const u64 base = (u64)blockIdx.x * (SV_THREADS * U) + threadIdx.x;
        double d[U];
        cd v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 x = base + (u64)u * SV_THREADS;
            if (x < dim) {
                d[u] = NT ? __builtin_nontemporal_load(diag + x) : diag[x];
                v[u] = ldx<NT>(psi + x);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 x = base + (u64)u * SV_THREADS;
            if (x < dim) {
                double s, c;
                sincos(d[u] * t, &s, &c);
                stx<NT>(psi + x, cmk(c * v[u].x + s * v[u].y, c * v[u].y - s * v[u].x));
            }
        }
    }
}

// ---- reductions ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block-wide sum of NV values; result valid in thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double red[NV][RED_THREADS / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        v[i] = wave_sum(v[i]);
        if (lane == 0) red[i][wv] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int w = 0; w < RED_THREADS / 64; ++w) s += red[i][w];
            v[i] = s;
        }
    }
}

// U independent accumulators per thread: x, x + G, ..., x + (U-1)G (G = grid stride) are loaded
// together, so each thread keeps U loads in flight; they are added in that fixed order.
template <int U>
__device__ __forceinline__ double sum_fixed(const double (&v)[U]) {
    double s = v[0];
#pragma unroll
    for (int u = 1; u < U; ++u) s += v[u];
    return s;
}

template <typename S, int NT, int U>
__global__ __launch_bounds__(RED_THREADS) void sv_mean_kernel(const S *__restrict__ psi, const double *__restrict__ obs,
                                                              u64 dim, double *__restrict__ part) {
    double a4[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a4[u] = 0.0;
    const u64 G = (u64)gridDim.x * RED_THREADS;
    for (u64 x0 = (u64)blockIdx.x * RED_THREADS + threadIdx.x; x0 < dim; x0 += G * U) {
        cd a[U];
        double o[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 x = x0 + (u64)u * G;
            const bool in = x < dim;
            a[u] = in ? ldx<NT>(psi + x) : cmk(0.0, 0.0);
            o[u] = in ? (NT ? __builtin_nontemporal_load(obs + x) : obs[x]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) a4[u] += o[u] * (a[u].x * a[u].x + a[u].y * a[u].y);
    }
    double acc[1] = {sum_fixed<U>(a4)};
    block_sum<1>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

template <typename S, int NT, int U>
__global__ __launch_bounds__(RED_THREADS) void sv_matel_kernel(const S *__restrict__ bra, const S *__restrict__ ket,
                                                               const double *__restrict__ obs, u64 dim,
                                                               double *__restrict__ part) {
    double re4[U], im4[U];
#pragma unroll
    for (int u = 0; u < U; ++u) re4[u] = im4[u] = 0.0;
    const u64 G = (u64)gridDim.x * RED_THREADS;
    for (u64 x0 = (u64)blockIdx.x * RED_THREADS + threadIdx.x; x0 < dim; x0 += G * U) {
        cd a[U], b[U];
        double o[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 x = x0 + (u64)u * G;
            const bool in = x < dim;
            a[u] = in ? ldx<NT>(bra + x) : cmk(0.0, 0.0);
            b[u] = in ? ldx<NT>(ket + x) : cmk(0.0, 0.0);
            o[u] = in ? (NT ? __builtin_nontemporal_load(obs + x) : obs[x]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            re4[u] += o[u] * (a[u].x * b[u].x + a[u].y * b[u].y);
            im4[u] += o[u] * (a[u].x * b[u].y - a[u].y * b[u].x);
        }
    }
    double acc[2] = {sum_fixed<U>(re4), sum_fixed<U>(im4)};
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = acc[0];
        part[2 * blockIdx.x + 1] = acc[1];
    }
}

// second stage: nv interleaved sums over nblk partials, written after the partials
template <int NV>
__global__ __launch_bounds__(RED_THREADS) void finish_kernel(double *__restrict__ part, int nblk) {
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += RED_THREADS)
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[i] += part[NV * b + i];
    block_sum<NV>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) part[NV * RED_BLOCKS + i] = acc[i];
}

template <typename S, int OP, int U, int XCD, int NT>
void launch_shape(hipStream_t st, void *psi, u64 npairs, int q, const Coef &k, int lds) {
    const u64 per = (u64)SV_THREADS * U;
    const unsigned grid = (unsigned)((npairs + per - 1) / per);
    if (OP == ORK_SV_P || OP == ORK_SV_Z || OP == ORK_SV_S || OP == ORK_SV_SDG || OP == ORK_SV_T ||
        OP == ORK_SV_TDG)
        sv_one_kernel<S, OP, U, XCD, NT><<<grid, SV_THREADS, lds, st>>>((S *)psi, npairs, q, k);
    else
        sv_pair_kernel<S, OP, U, XCD, NT><<<grid, SV_THREADS, lds, st>>>((S *)psi, npairs, q, k);
}

template <typename S, int OP, int U, int XCD>
void launch_nt(hipStream_t st, void *psi, u64 npairs, int q, const Coef &k, int lds, int nt) {
    if (nt) launch_shape<S, OP, U, XCD, 1>(st, psi, npairs, q, k, lds);
    else launch_shape<S, OP, U, XCD, 0>(st, psi, npairs, q, k, lds);
}

template <typename S, int OP, int XCD>
void launch_lane(hipStream_t st, void *psi, u64 dim, int q, const Coef &k, int u, int lds) {
    const u64 per = (u64)SV_THREADS * u;
    const unsigned grid = (unsigned)((dim + per - 1) / per);
    if (u >= 8) sv_pair_lane_kernel<S, OP, 8, XCD><<<grid, SV_THREADS, lds, st>>>((S *)psi, dim, q, k);
    else if (u >= 4) sv_pair_lane_kernel<S, OP, 4, XCD><<<grid, SV_THREADS, lds, st>>>((S *)psi, dim, q, k);
    else sv_pair_lane_kernel<S, OP, 2, XCD><<<grid, SV_THREADS, lds, st>>>((S *)psi, dim, q, k);
}

template <typename S, int OP>
void launch_pair(hipStream_t st, void *psi, u64 npairs, int q, const Coef &k) {
    SvShape sh = pair_shape(q, sizeof(S) == 8);
    // the remap needs a grid of at least 8 blocks; small states use the plain order
    if (npairs < (u64)SV_THREADS * sh.u * 8) sh.xcd = 0;
    // nontemporal only on large states
    if (2 * npairs * sizeof(S) < g_nt_min_bytes) sh.nt = 0;
    constexpr bool one = OP == ORK_SV_P || OP == ORK_SV_Z || OP == ORK_SV_S || OP == ORK_SV_SDG || OP == ORK_SV_T ||
                         OP == ORK_SV_TDG;
    if (((u64)sizeof(S) << q) < 128) {
        // a pair's two amplitudes share a 128 B line (target below 3 in complex128, below 4 in
        // complex64): nontemporal accesses only through the lane-exchange kernel, with the same bytes
        // in flight per thread (2 u amplitudes)
        if (sh.nt && !one) {
            if (sh.xcd) launch_lane<S, OP, 1>(st, psi, 2 * npairs, q, k, 2 * sh.u, sh.lds);
            else launch_lane<S, OP, 0>(st, psi, 2 * npairs, q, k, 2 * sh.u, sh.lds);
            return;
        }
        sh.nt = 0;
    }
    if (sh.u == 4) {
        if (sh.xcd) launch_nt<S, OP, 4, 1>(st, psi, npairs, q, k, sh.lds, sh.nt);
        else launch_nt<S, OP, 4, 0>(st, psi, npairs, q, k, sh.lds, sh.nt);
    } else if (sh.u == 2) {
        if (sh.xcd) launch_nt<S, OP, 2, 1>(st, psi, npairs, q, k, sh.lds, sh.nt);
        else launch_nt<S, OP, 2, 0>(st, psi, npairs, q, k, sh.lds, sh.nt);
    } else {
        if (sh.xcd) launch_nt<S, OP, 1, 1>(st, psi, npairs, q, k, sh.lds, sh.nt);
        else launch_nt<S, OP, 1, 0>(st, psi, npairs, q, k, sh.lds, sh.nt);
    }
}

template <typename S>
int sv_1q(hipStream_t st, void *psi, int n, int q, int op, const Coef &k) {
    const u64 npairs = ((u64)1 << n) >> 1;
    switch (op) {
    case ORK_SV_MAT: launch_pair<S, ORK_SV_MAT>(st, psi, npairs, q, k); break;
    case ORK_SV_H: launch_pair<S, ORK_SV_H>(st, psi, npairs, q, k); break;
    case ORK_SV_HY: launch_pair<S, ORK_SV_HY>(st, psi, npairs, q, k); break;
    case ORK_SV_X: launch_pair<S, ORK_SV_X>(st, psi, npairs, q, k); break;
    case ORK_SV_Y: launch_pair<S, ORK_SV_Y>(st, psi, npairs, q, k); break;
    case ORK_SV_RX: launch_pair<S, ORK_SV_RX>(st, psi, npairs, q, k); break;
    case ORK_SV_RY: launch_pair<S, ORK_SV_RY>(st, psi, npairs, q, k); break;
    case ORK_SV_RZ: launch_pair<S, ORK_SV_RZ>(st, psi, npairs, q, k); break;
    case ORK_SV_P: launch_pair<S, ORK_SV_P>(st, psi, npairs, q, k); break;
    case ORK_SV_Z: launch_pair<S, ORK_SV_Z>(st, psi, npairs, q, k); break;
    case ORK_SV_S: launch_pair<S, ORK_SV_S>(st, psi, npairs, q, k); break;
    case ORK_SV_SDG: launch_pair<S, ORK_SV_SDG>(st, psi, npairs, q, k); break;
    case ORK_SV_T: launch_pair<S, ORK_SV_T>(st, psi, npairs, q, k); break;
    case ORK_SV_TDG: launch_pair<S, ORK_SV_TDG>(st, psi, npairs, q, k); break;
    default: return (int)hipErrorInvalidValue;
    }
    ORK_LAUNCH_CHECK();
    return 0;
}


// ---- fused layer of single-qubit gates ----------------------------------------------------
// One sweep loads tiles of 2^tb amplitudes into LDS, applies the 2x2 matrices of the qubits the
// tile spans and writes the tile back.  Tile element e maps to the amplitude
//   x = low | (tmid << r) | ((e >> r) << lo) | (thigh << (lo + g)),   low = e & (2^r - 1),
// so a tile holds 2^r consecutive amplitudes (whole lines) times the 2^g values of qubits
// lo..lo+g-1; the first sweep (r = tb, g = 0) takes qubits 0..tb-1, each later one the next g
// qubits, so every qubit's matrix is applied exactly once.
constexpr int LAYER_LT = 12, LAYER_R = 4;   // largest tile (2^12 amplitudes) and the low run (16)
struct LayerArgs {
    double m[LAYER_LT][8];   // per tile bit: m00, m01, m10, m11 (re, im), row-major
    int tb, r, lo, g;        // tile bits; low run bits; global bit of tile bit r; gate bits above the run
    int b0, b1;              // tile bits [b0, b1) carry a matrix
};

__device__ __forceinline__ int lpad(int i) { return i + (i >> 3); }   // one spare slot per 8 amplitudes

// the matrices m[0..rb) on bits 0..rb-1 of the index of v, ascending
__device__ __forceinline__ void mat3(cd (&v)[8], const double (*m)[8], int rb) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= rb) break;
        const cd m00 = cmk(m[k][0], m[k][1]), m01 = cmk(m[k][2], m[k][3]), m10 = cmk(m[k][4], m[k][5]),
                 m11 = cmk(m[k][6], m[k][7]);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (!((j >> k) & 1)) {
                cd x0 = cmul(m00, v[j]), x1 = cmul(m10, v[j]);
                cfma(x0, m01, v[j | (1 << k)]);
                cfma(x1, m11, v[j | (1 << k)]);
                v[j] = x0;
                v[j | (1 << k)] = x1;
            }
    }
}

template <typename S, int NT, int LT, int THREADS>
__global__ __launch_bounds__(THREADS) void sv_layer_kernel(S *__restrict__ psi, LayerArgs a) {
    __shared__ cd tile[(1 << LT) + ((1 << LT) >> 3)];
    const int T = 1 << a.tb;
    const u64 t = blockIdx.x;
    const int midb = a.lo - a.r;
    const u64 tmid = t & (((u64)1 << midb) - 1), thigh = t >> midb;
    const u64 base = (tmid << a.r) | (thigh << (a.lo + a.g));
#define LIDX(e) (base | (u64)((e) & ((1 << a.r) - 1)) | ((u64)((e) >> a.r) << a.lo))
    if (T == 8 * THREADS && a.b1 == a.tb && a.b1 - a.b0 >= 3) {
        // 8 amplitudes per thread, e = tid + u 2^UB: the top three tile bits are the register index
        // u, so their matrices apply as the tile is loaded; the last LDS pass stores to HBM directly
        const int UB = a.tb - 3;
        cd v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ldx<NT>(psi + LIDX(threadIdx.x + (u << UB)));
        mat3(v, a.m + UB, 3);
        if (a.b0 == UB) {
#pragma unroll
            for (int u = 0; u < 8; ++u) stx<NT>(psi + LIDX(threadIdx.x + (u << UB)), v[u]);
            return;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) tile[lpad(threadIdx.x + (u << UB))] = v[u];
        __syncthreads();
        for (int b = a.b0; b < UB;) {
            const int rb = (UB - b) < 3 ? (UB - b) : 3;
            const bool lastp = b + rb == UB;
            const int R = 1 << rb, ngroups = T >> rb, lowmask = (1 << b) - 1;
            for (int gi = threadIdx.x; gi < ngroups; gi += THREADS) {
                const int g0 = ((gi & ~lowmask) << rb) | (gi & lowmask);
                cd w[8];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (j < R) w[j] = tile[lpad(g0 | (j << b))];
                mat3(w, a.m + b, rb);
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (j < R) {
                        if (lastp) stx<NT>(psi + LIDX(g0 | (j << b)), w[j]);
                        else tile[lpad(g0 | (j << b))] = w[j];
                    }
            }
            if (!lastp) __syncthreads();
            b += rb;
        }
        return;
    }
    {
        // batches of up to 8 loads per thread, all issued before the first LDS write (element 0's address
        // stands in past the tile end, so that no load is conditional)
        constexpr int PER = (1 << LT) / THREADS, BT = PER < 8 ? PER : 8;
#pragma unroll 1
        for (int u0 = 0; u0 < PER; u0 += BT) {
            S raw[BT];
#pragma unroll
            for (int u = 0; u < BT; ++u) {
                const int e = threadIdx.x + (u0 + u) * THREADS;
                raw[u] = ldr<NT>(psi + LIDX(e < T ? e : 0));
            }
#pragma unroll
            for (int u = 0; u < BT; ++u) {
                const int e = threadIdx.x + (u0 + u) * THREADS;
                if (e < T) tile[lpad(e)] = to_cd(raw[u]);
            }
        }
    }
    __syncthreads();
    // up to three tile bits per pass: each thread holds the 8 amplitudes of one 3-bit group
    for (int b = a.b0; b < a.b1;) {
        const int rb = (a.b1 - b) < 3 ? (a.b1 - b) : 3;
        const int R = 1 << rb, ngroups = T >> rb, lowmask = (1 << b) - 1;
        for (int gi = threadIdx.x; gi < ngroups; gi += THREADS) {
            const int g0 = ((gi & ~lowmask) << rb) | (gi & lowmask);
            cd v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < R) v[j] = tile[lpad(g0 | (j << b))];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k >= rb) break;
                const double *m = a.m[b + k];
                const cd m00 = cmk(m[0], m[1]), m01 = cmk(m[2], m[3]), m10 = cmk(m[4], m[5]), m11 = cmk(m[6], m[7]);
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (j < R && !((j >> k) & 1)) {
                        cd x0 = cmul(m00, v[j]), x1 = cmul(m10, v[j]);
                        cfma(x0, m01, v[j | (1 << k)]);
                        cfma(x1, m11, v[j | (1 << k)]);
                        v[j] = x0;
                        v[j | (1 << k)] = x1;
                    }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < R) tile[lpad(g0 | (j << b))] = v[j];
        }
        __syncthreads();
        b += rb;
    }
    for (int e = threadIdx.x; e < T; e += THREADS) stx<NT>(psi + LIDX(e), tile[lpad(e)]);
#undef LIDX
}

// tile bits and workgroup size of the layer sweeps: 0 (12, 256), 1 (12, 512), 2 (11, 256), 3 (11, 512).
// Measured on 2^28 amplitudes: 8.2 / 4.9 / 10.0 / 15.3 ms per layer (complex128), 7.2 / 4.1 / 9.1 / 14.6
// (complex64): 512 threads hold the tile's 8 amplitudes each, the register-resident path.
constexpr int kLayerVariant = 1;
static int g_layer_variant = kLayerVariant;

template <typename S, int LT, int THREADS>
void launch_layer(hipStream_t st, void *psi, int n, bool nt, const LayerArgs &a) {
    const unsigned grid = (unsigned)((u64)1 << (n - a.tb));
    if (nt) sv_layer_kernel<S, 1, LT, THREADS><<<grid, THREADS, 0, st>>>((S *)psi, a);
    else sv_layer_kernel<S, 0, LT, THREADS><<<grid, THREADS, 0, st>>>((S *)psi, a);
}

template <typename S>
int sv_layer(hipStream_t st, void *psi, int n, const double *mats) {
    const bool nt = !(g_nt_off & 1) && ((u64)sizeof(S) << n) >= g_nt_min_bytes;
    const int v = g_layer_variant, LTv = v >= 2 ? 11 : 12;
    auto launch = [&](const LayerArgs &a) {
        switch (v) {
        case 1: launch_layer<S, 12, 512>(st, psi, n, nt, a); break;
        case 2: launch_layer<S, 11, 256>(st, psi, n, nt, a); break;
        case 3: launch_layer<S, 11, 512>(st, psi, n, nt, a); break;
        default: launch_layer<S, 12, 256>(st, psi, n, nt, a); break;
        }
        ORK_LAUNCH_CHECK();
        return 0;
    };
    LayerArgs a{};
    const int tb0 = n < LTv ? n : LTv;
    a.tb = a.r = a.lo = tb0;
    a.g = 0;
    a.b0 = 0;
    a.b1 = tb0;
    for (int q = 0; q < tb0; ++q)
        for (int i = 0; i < 8; ++i) a.m[q][i] = mats[8 * q + i];
    int e = launch(a);
    for (int lo = tb0; lo < n && !e; lo += LTv - LAYER_R) {
        const int g = (n - lo) < (LTv - LAYER_R) ? (n - lo) : (LTv - LAYER_R);
        LayerArgs b{};
        b.r = LAYER_R;
        b.g = g;
        b.tb = LAYER_R + g;
        b.lo = lo;
        b.b0 = LAYER_R;
        b.b1 = LAYER_R + g;
        for (int j = 0; j < g; ++j)
            for (int i = 0; i < 8; ++i) b.m[LAYER_R + j][i] = mats[8 * (lo + j) + i];
        e = launch(b);
    }
    return e;
}

}  // namespace

// ---- C ABI --------------------------------------------------------------------------------
extern "C" uint64_t ork_partials_len(void) { return 2 * RED_BLOCKS + 2; }

// mats: n x 8 doubles, the row-major 2x2 matrix of each qubit (re, im interleaved)
extern "C" int ork_sv_layer(void *stream, int prec, void *psi, int n, const double *mats) {
    if (n < 1 || n > 40 || !mats) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    return prec == ORK_F32 ? sv_layer<float2>(st, psi, n, mats) : sv_layer<double2>(st, psi, n, mats);
}

extern "C" int ork_sv_1q(void *stream, int prec, void *psi, int n, int q, int op, const double *coef) {
    Coef k;
    for (int i = 0; i < 8; ++i) k.v[i] = coef ? coef[i] : 0.0;
    hipStream_t st = (hipStream_t)stream;
    return prec == ORK_F32 ? sv_1q<float2>(st, psi, n, q, op, k) : sv_1q<double2>(st, psi, n, q, op, k);
}

// residency cap of the permutation / phase kernel (cx, cz, swap, ...) on complex64 states
constexpr int kMonoLds32 = 26624;
static int g_mono_lds32 = kMonoLds32;

// permutation kernels: lane-exchange form (every gate qubit below 6, one of them inside a 128 B line)
// and the paired complex64 form; set through orkan_set_launch_shape 21
struct PermShape {
    int lane_on, lane_all;     // lane kernel where a gate qubit lies inside a line / for every qubit set below 6
    int lane_u, lane_u32;      // amplitudes per thread (2, 4 or 8)
    int lane_lds, lane_lds32;  // residency caps
    int wl;                    // whole-line writes
    int x2_on, x2_lds;         // complex64: 16 B elements when no gate qubit is bit 0
};
// measured on 2^28 amplitudes (cx on every neighbour pair): complex128 lane kernel 4 amplitudes per thread
// under a 40 KB cap (low pairs 3.24 TB/s against 2.66 through the per-group kernel), complex64 4 per
// thread uncapped (low pairs 3.11-3.16, cx(4, 3) 5.2-5.3; 8 per thread 2.9 and 5.0 once the loads were
// batched, 2.9 and 3.8 before); whole-line writes (without them 2.0-2.2); the paired complex64 kernel
// under a 52 KB cap (5.9-6.2 TB/s on pairs from (4, 5) up, against 3.3 with 8 B accesses)
constexpr PermShape kPerm = {1, 0, 4, 4, 40960, 0, 1, 1, 53248};
static PermShape g_perm = kPerm;

template <typename S, int U>
void launch_perm_lane_u(hipStream_t st, void *psi, u64 dim, const PermLaneArgs &a, int xcd, int nt, int wl, int lds) {
    const u64 per = (u64)SV_THREADS * U;
    const unsigned grid = (unsigned)((dim + per - 1) / per);
    S *p = (S *)psi;
#define ORK_PERM_LAUNCH(X, N, W) sv_perm_lane_kernel<S, U, X, N, W><<<grid, SV_THREADS, lds, st>>>(p, dim, a)
    if (xcd) {
        if (nt) { if (wl) ORK_PERM_LAUNCH(1, 1, 1); else ORK_PERM_LAUNCH(1, 1, 0); }
        else { if (wl) ORK_PERM_LAUNCH(1, 0, 1); else ORK_PERM_LAUNCH(1, 0, 0); }
    } else {
        if (nt) { if (wl) ORK_PERM_LAUNCH(0, 1, 1); else ORK_PERM_LAUNCH(0, 1, 0); }
        else { if (wl) ORK_PERM_LAUNCH(0, 0, 1); else ORK_PERM_LAUNCH(0, 0, 0); }
    }
#undef ORK_PERM_LAUNCH
}

template <typename S>
void launch_perm_lane(hipStream_t st, void *psi, u64 dim, const PermLaneArgs &a, int u, int xcd, int nt, int wl,
                      int lds) {
    if (u >= 8) launch_perm_lane_u<S, 8>(st, psi, dim, a, xcd, nt, wl, lds);
    else if (u >= 4) launch_perm_lane_u<S, 4>(st, psi, dim, a, xcd, nt, wl, lds);
    else launch_perm_lane_u<S, 2>(st, psi, dim, a, xcd, nt, wl, lds);
}

// flags: bits 0-1 log2 of the lane kernel's amplitudes per thread (1..3), bit 2 whole-line writes,
// bit 3 lane kernel off, bit 4 lane kernel for every gate set below 6, bit 5 paired complex64 kernel
// off, bit 6 the settings are for complex64 states (u, lds), else complex128; lds: residency cap of the
// lane kernel (of the paired kernel when bit 7 is set)
extern "C" int ork_sv_set_perm_shape(int flags, int lds) {
    if (flags < 0 || flags > 255 || lds < 0 || lds > 160 * 1024 || (flags & 3) == 0) return (int)hipErrorInvalidValue;
    const int u = 1 << (flags & 3);
    g_perm.wl = (flags >> 2) & 1;
    g_perm.lane_on = !((flags >> 3) & 1);
    g_perm.lane_all = (flags >> 4) & 1;
    g_perm.x2_on = !((flags >> 5) & 1);
    if (flags & 128) {
        g_perm.x2_lds = lds;
    } else if (flags & 64) {
        g_perm.lane_u32 = u;
        g_perm.lane_lds32 = lds;
    } else {
        g_perm.lane_u = u;
        g_perm.lane_lds = lds;
    }
    return 0;
}

extern "C" int ork_sv_set_mono_lds32(int lds) {
    if (lds < 0 || lds > 160 * 1024) return (int)hipErrorInvalidValue;
    g_mono_lds32 = lds;
    return 0;
}

extern "C" int ork_sv_mono(void *stream, int prec, void *psi, int n, const ork_mono_t *mono) {
    MonoArgs m;
    m.k = mono->k;
    int srt[3] = {mono->qs[0], mono->qs[1], mono->qs[2]};
    for (int i = 0; i < m.k; ++i)
        for (int j = i + 1; j < m.k; ++j)
            if (srt[j] < srt[i]) { int t = srt[i]; srt[i] = srt[j]; srt[j] = t; }
    for (int i = 0; i < 3; ++i) m.pos[i] = i < m.k ? srt[i] : 0;
    auto spread = [&](int a) {
        u64 v = 0;
        for (int j = 0; j < mono->k; ++j)
            if (a >> j & 1) v |= (u64)1 << mono->qs[j];
        return v;
    };
    m.nt = 0;
    for (int a = 0; a < (1 << m.k); ++a) {
        if (mono->perm[a] == a && (mono->ph[a] & 3) == 0) continue;
        m.src[m.nt] = spread(a);
        m.dst[m.nt] = spread(mono->perm[a]);
        m.ph[m.nt] = mono->ph[a] & 3;
        ++m.nt;
    }
    if (m.nt == 0) return 0;
    const u64 nbase = ((u64)1 << n) >> m.k;
    hipStream_t st = (hipStream_t)stream;
    const int lds = prec == ORK_F32 ? g_mono_lds32 : 53248;   // residency cap (complex128: as the pair kernel)
    const u64 elem = prec == ORK_F32 ? 8 : 16;
    const bool big = !(g_nt_off & 2) && (elem << n) >= g_nt_min_bytes;
    const bool partial = (elem << m.pos[0]) < 128;   // a gate qubit inside a 128 B line
    if (g_perm.lane_on && m.pos[m.k - 1] < 6 && (partial || g_perm.lane_all)) {
        PermLaneArgs a;
        a.k = mono->k;
        for (int j = 0; j < 3; ++j) a.qs[j] = j < a.k ? mono->qs[j] : 0;
        for (int x = 0; x < 8; ++x) { a.pinv[x] = x; a.ph[x] = 0; }
        for (int x = 0; x < (1 << a.k); ++x) {
            a.pinv[mono->perm[x]] = x;
            a.ph[x] = mono->ph[x] & 3;
        }
        a.pinv_bits = a.ph_bits = 0;
        for (int x = 0; x < 8; ++x) {
            a.pinv_bits |= (unsigned)a.pinv[x] << (3 * x);
            a.ph_bits |= (unsigned)a.ph[x] << (2 * x);
        }
        const u64 dim = (u64)1 << n;
        const int u = prec == ORK_F32 ? g_perm.lane_u32 : g_perm.lane_u;
        const int xcd = dim >= (u64)SV_THREADS * u * 8;
        const int plds = prec == ORK_F32 ? g_perm.lane_lds32 : g_perm.lane_lds;
        // whole-line writes only where the moved amplitudes of a line come in runs shorter than 64 B: a gate
        // argument below the 64 B half of a line (bit 2 complex128, 3 complex64) that decides whether an
        // amplitude moves.  Aligned 64 B runs write as well partially (cx(2, 3): 3.8 against 3.2 TB/s)
        const int half_bit = prec == ORK_F32 ? 3 : 2;
        bool fine = false;
        for (int j = 0; j < a.k; ++j) {
            if (a.qs[j] >= half_bit) continue;
            for (int x = 0; x < (1 << a.k); ++x) {
                const bool mx = a.pinv[x] != x || a.ph[x], my = a.pinv[x ^ (1 << j)] != (x ^ (1 << j)) || a.ph[x ^ (1 << j)];
                if (mx != my) fine = true;
            }
        }
        const int wl = g_perm.wl && fine;
        if (prec == ORK_F32) launch_perm_lane<float2>(st, psi, dim, a, u, xcd, big, wl, plds);
        else launch_perm_lane<double2>(st, psi, dim, a, u, xcd, big, wl, plds);
        ORK_LAUNCH_CHECK();
        return 0;
    }
    // streaming accesses when no gate qubit lies inside a 128 B line (as for the pair kernel)
    const bool nt = big && !partial;
    if (prec == ORK_F32 && g_perm.x2_on && m.pos[0] >= 1) {
        // pairs of amplitudes as 16 B elements: gate qubits one lower, offsets halved
        MonoArgs m2 = m;
        for (int i = 0; i < m.k; ++i) m2.pos[i] = m.pos[i] - 1;
        for (int j = 0; j < m.nt; ++j) {
            m2.src[j] = m.src[j] >> 1;
            m2.dst[j] = m.dst[j] >> 1;
        }
        const u64 nb2 = nbase >> 1;
        const bool nt2 = big && (16ull << m2.pos[0]) >= 128;
        const u64 per = (u64)SV_THREADS * SV_UNROLL;
        const unsigned grid = (unsigned)((nb2 + per - 1) / per);
        if (nt2) sv_mono_x2_kernel<SV_UNROLL, 1><<<grid, SV_THREADS, g_perm.x2_lds, st>>>((f4v *)psi, nb2, m2);
        else sv_mono_x2_kernel<SV_UNROLL, 0><<<grid, SV_THREADS, g_perm.x2_lds, st>>>((f4v *)psi, nb2, m2);
        ORK_LAUNCH_CHECK();
        return 0;
    }
    if (prec == ORK_F32) {
        const u64 per = (u64)SV_THREADS * SV_UNROLL * 2;
        const unsigned grid = (unsigned)((nbase + per - 1) / per);
        if (nt) sv_mono_kernel<float2, SV_UNROLL * 2, 1><<<grid, SV_THREADS, lds, st>>>((float2 *)psi, nbase, m);
        else sv_mono_kernel<float2, SV_UNROLL * 2, 0><<<grid, SV_THREADS, lds, st>>>((float2 *)psi, nbase, m);
    } else {
        const u64 per = (u64)SV_THREADS * SV_UNROLL;
        const unsigned grid = (unsigned)((nbase + per - 1) / per);
        if (nt) sv_mono_kernel<double2, SV_UNROLL, 1><<<grid, SV_THREADS, lds, st>>>((double2 *)psi, nbase, m);
        else sv_mono_kernel<double2, SV_UNROLL, 0><<<grid, SV_THREADS, lds, st>>>((double2 *)psi, nbase, m);
    }
    ORK_LAUNCH_CHECK();
    return 0;
}

// exp_diag launch shape: amplitudes per thread (1, 2 or 4) and dynamic LDS bytes per workgroup
// (a residency cap).  Measured on 2^28 amplitudes: one call between other kernels (the pattern
// of a circuit) runs 6.45 TB/s with 1 per thread uncapped and 6.07 with 4 per thread under a 32 KB
// cap; back-to-back calls favour the latter slightly (6.26-6.29 against 6.11-6.14 TB/s).
constexpr int kExpU = 1, kExpLds = 0;
// complex64 (16 B per amplitude moved): one amplitude per thread leaves too few bytes in flight;
// 4 per thread, uncapped, measured 6.27 TB/s against 5.44 on single calls at 2^28
constexpr int kExpU32 = 4, kExpLds32 = 0;
static int g_exp_u[2] = {kExpU, kExpU32}, g_exp_lds[2] = {kExpLds, kExpLds32};

// flags bits 0-1: log2 of the amplitudes per thread (0..2); bit 2: the complex64 shape
extern "C" int ork_sv_set_exp_shape(int flags, int lds) {
    if (flags < 0 || flags > 7 || (flags & 3) > 2 || lds < 0 || lds > 160 * 1024) return (int)hipErrorInvalidValue;
    g_exp_u[flags >> 2] = 1 << (flags & 3);
    g_exp_lds[flags >> 2] = lds;
    return 0;
}

template <typename S, int NT>
static void launch_exp(hipStream_t st, void *psi, const double *diag, u64 dim, double t) {
    const int f = sizeof(S) == 8, u = g_exp_u[f], lds = g_exp_lds[f];
    const u64 per = (u64)SV_THREADS * u;
    const unsigned grid = (unsigned)((dim + per - 1) / per);
    if (u == 4) sv_exp_diag_kernel<S, NT, 4><<<grid, SV_THREADS, lds, st>>>((S *)psi, diag, dim, t);
    else if (u == 2) sv_exp_diag_kernel<S, NT, 2><<<grid, SV_THREADS, lds, st>>>((S *)psi, diag, dim, t);
    else sv_exp_diag_kernel<S, NT, 1><<<grid, SV_THREADS, lds, st>>>((S *)psi, diag, dim, t);
}

extern "C" int ork_sv_exp_diag(void *stream, int prec, void *psi, int n, const double *diag, double t) {
    const u64 dim = (u64)1 << n;
    hipStream_t st = (hipStream_t)stream;
    const bool nt = !(g_nt_off & 4) && dim * (prec == ORK_F32 ? 8 : 16) >= g_nt_min_bytes;
    if (prec == ORK_F32) {
        if (nt) launch_exp<float2, 1>(st, psi, diag, dim, t);
        else launch_exp<float2, 0>(st, psi, diag, dim, t);
    } else {
        if (nt) launch_exp<double2, 1>(st, psi, diag, dim, t);
        else launch_exp<double2, 0>(st, psi, diag, dim, t);
    }
    ORK_LAUNCH_CHECK();
    return 0;
}

// First-stage shape of the mean / matel reductions: dynamic LDS bytes per workgroup (a residency
// cap), loads in flight per thread (4 or 8) and the largest number of workgroups (RED_BLOCKS >> k).
// Measured on 2^28 amplitudes, host read included: 256 workgroups (one per CU: 96 KB of LDS each
// leaves room for no second one) with 4 loads in flight read 6.94 TB/s (mean) and 7.04 TB/s
// (matel_diag); 2048 workgroups grid-striding in four rounds of resident groups reached 5.85 / 5.94,
// 512 reached 6.74 / 6.17, and 128 ran out of loads in flight (4.13 / 5.60).
constexpr int kRedLds = 98304;
constexpr int kRedUnroll = 4, kRedShift = 3;
// complex64 storage: 8 B per amplitude leaves too few bytes in flight in one workgroup per CU
// (2.78 / 2.93 TB/s); 1024 uncapped workgroups, all resident at once, read 6.86 / 6.88 TB/s
constexpr int kRedLds32 = 0;
constexpr int kRedUnroll32 = 4, kRedShift32 = 1;
struct RedShape {
    int lds, unroll, shift;
};
static RedShape g_red[2] = {{kRedLds, kRedUnroll, kRedShift}, {kRedLds32, kRedUnroll32, kRedShift32}};
// flags bit 4 on: every target qubit uses (bit 0 xcd, bit 1 nontemporal, bits 2-3 log2 of the pairs
// per thread: 0..2) with LDS cap lds; bit 4 off: back to the per-qubit table
extern "C" int ork_sv_set_pair_shape(int flags, int lds) {
    if (lds < 0 || lds > 160 * 1024 || flags < 0 || flags > 31 || ((flags >> 2) & 3) == 3) return (int)hipErrorInvalidValue;
    if (!(flags & 16)) g_pair_override = SvShape{0, 0, 0, 0};
    else g_pair_override = SvShape{flags & 1, 1 << ((flags >> 2) & 3), lds, (flags >> 1) & 1};
    return 0;
}

// tuning hook: tile bits and workgroup size of the fused-layer sweeps (variants 0-3)
extern "C" int ork_sv_set_layer_variant(int v) {
    if (v < 0 || v > 3) return (int)hipErrorInvalidValue;
    g_layer_variant = v;
    return 0;
}

// tuning hook: nontemporal accesses off per kernel family (bit 0 pair kernels, bit 1 permutations,
// bit 2 exp_diag, bit 3 reductions); min_bytes: smallest state that uses them (0: 512 MiB)
extern "C" int ork_sv_set_nt_off(int mask, int min_bytes) {
    if (mask < 0 || mask > 15 || min_bytes < 0) return (int)hipErrorInvalidValue;
    g_nt_off = mask;
    g_nt_min_bytes = min_bytes ? (u64)min_bytes : kNtMinBytes;
    return 0;
}

extern "C" unsigned long long ork_nt_min_bytes(void) { return g_nt_min_bytes; }

extern "C" void ork_sv_reset_tuning(void) {
    g_perm = kPerm;
    g_pair_override = SvShape{0, 0, 0, 0};
    g_nt_off = 0;
    g_nt_min_bytes = kNtMinBytes;
    g_layer_variant = kLayerVariant;
    g_red[0] = RedShape{kRedLds, kRedUnroll, kRedShift};
    g_red[1] = RedShape{kRedLds32, kRedUnroll32, kRedShift32};
    g_mono_lds32 = kMonoLds32;
    g_exp_u[0] = kExpU;
    g_exp_u[1] = kExpU32;
    g_exp_lds[0] = kExpLds;
    g_exp_lds[1] = kExpLds32;
}

// flags bit 0: 8 loads in flight per thread (else 4); bits 1-3: k, at most RED_BLOCKS >> k workgroups;
// bit 4: the complex64 shape (else complex128)
extern "C" int ork_sv_set_red_lds(int flags, int lds) {
    if (lds < 0 || lds > 160 * 1024 || flags < 0 || flags > 31 || (flags & 15) > 9) return (int)hipErrorInvalidValue;
    g_red[flags >> 4] = RedShape{lds, (flags & 1) ? 8 : 4, (flags & 15) >> 1};
    return 0;
}

static int red_blocks(u64 dim, const RedShape &r) {
    const u64 cap = (u64)(RED_BLOCKS >> r.shift);
    u64 b = (dim + RED_THREADS * 8 - 1) / (RED_THREADS * 8);
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

template <typename S, int NT>
static void launch_mean(hipStream_t st, int nb, const void *psi, const double *obs, u64 dim, double *part) {
    const RedShape &r = g_red[sizeof(S) == 8];
    if (r.unroll == 8)
        sv_mean_kernel<S, NT, 8><<<nb, RED_THREADS, r.lds, st>>>((const S *)psi, obs, dim, part);
    else
        sv_mean_kernel<S, NT, 4><<<nb, RED_THREADS, r.lds, st>>>((const S *)psi, obs, dim, part);
}

template <typename S, int NT>
static void launch_matel(hipStream_t st, int nb, const void *bra, const void *ket, const double *obs, u64 dim,
                         double *part) {
    const RedShape &r = g_red[sizeof(S) == 8];
    if (r.unroll == 8)
        sv_matel_kernel<S, NT, 8><<<nb, RED_THREADS, r.lds, st>>>((const S *)bra, (const S *)ket, obs, dim, part);
    else
        sv_matel_kernel<S, NT, 4><<<nb, RED_THREADS, r.lds, st>>>((const S *)bra, (const S *)ket, obs, dim, part);
}

extern "C" int ork_sv_mean(void *stream, int prec, const void *psi, int n, const double *obs, double *part,
                           double *out) {
    const u64 dim = (u64)1 << n;
    const int nb = red_blocks(dim, g_red[prec == ORK_F32]);
    hipStream_t st = (hipStream_t)stream;
    const bool nt = !(g_nt_off & 8) && dim * (prec == ORK_F32 ? 8 : 16) >= g_nt_min_bytes;
    if (prec == ORK_F32) {
        if (nt) launch_mean<float2, 1>(st, nb, psi, obs, dim, part);
        else launch_mean<float2, 0>(st, nb, psi, obs, dim, part);
    } else {
        if (nt) launch_mean<double2, 1>(st, nb, psi, obs, dim, part);
        else launch_mean<double2, 0>(st, nb, psi, obs, dim, part);
    }
    ORK_LAUNCH_CHECK();
    finish_kernel<1><<<1, RED_THREADS, 0, st>>>(part, nb);
    ORK_LAUNCH_CHECK();
    hipError_t e = hipMemcpyAsync(out, part + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return (int)e;
    return (int)hipStreamSynchronize(st);
}

extern "C" int ork_sv_matel(void *stream, int prec, const void *bra, const void *ket, int n, const double *obs,
                            double *part, double *out) {
    const u64 dim = (u64)1 << n;
    const int nb = red_blocks(dim, g_red[prec == ORK_F32]);
    hipStream_t st = (hipStream_t)stream;
    const bool nt = !(g_nt_off & 8) && dim * (prec == ORK_F32 ? 8 : 16) >= g_nt_min_bytes;
    if (prec == ORK_F32) {
        if (nt) launch_matel<float2, 1>(st, nb, bra, ket, obs, dim, part);
        else launch_matel<float2, 0>(st, nb, bra, ket, obs, dim, part);
    } else {
        if (nt) launch_matel<double2, 1>(st, nb, bra, ket, obs, dim, part);
        else launch_matel<double2, 0>(st, nb, bra, ket, obs, dim, part);
    }
    ORK_LAUNCH_CHECK();
    finish_kernel<2><<<1, RED_THREADS, 0, st>>>(part, nb);
    ORK_LAUNCH_CHECK();
    hipError_t e = hipMemcpyAsync(out, part + 2 * RED_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return (int)e;
    return (int)hipStreamSynchronize(st);
}
