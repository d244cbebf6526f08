// k_qaoa.hip -- batched QAOA evaluation: thousands of (graph, angles) circuits
// of uneven qubit counts in one persistent launch.
//
// One evaluation is the reference's own QAOA call sequence:
//   state_plus(PURE, n); p x { exp_diag(C, gamma); rx(q, 2 beta) for q = 0..n-1 };
//   mean(C)
// (src/state/state_pure.c:30-43, src/circ/circ_pure.c:21-34,
//  src/gate/gate_pure.c:221-234, src/meas/meas_pure.c:20-30), with the same
// per-amplitude formulas for the phases and the rx with their common factor
// taken out (restored in the mean), so each instance matches a per-instance run
// of those calls to rounding (the rx of one layer act on distinct qubits and
// commute; states in HBM apply them in another order than 0..n-1).
//
// Execution: every workgroup of the persistent grid pulls evaluations from an
// atomic counter (largest first) and runs each one start to finish, so no
// workgroup ever waits for another.  A state of n <= LT qubits lives in LDS for
// the whole circuit.  A larger state (n <= 2 LT) lives in HBM and is swept in
// LDS tiles of 2^LT amplitudes, alternating contiguous tiles and tiles gathered
// at stride 2^LT; each sweep finishes the previous layer's rx on the qubits the
// previous sweep could not reach and runs the next layer's exp_diag and rx on
// every qubit it holds: p + 1 sweeps for p layers instead of p (n + 1) gate
// passes; the last sweep reduces sum_x C(x) |psi_x|^2.
#include "k_common.hpp"

namespace {

struct QaoaEval {
    int graph;       // index into the diagonal table
    int p;           // layers
    int angle_off;   // offset (in doubles) of this evaluation's per-layer (gamma, cos beta, sin beta)
    int pad;
};

struct QaoaGraph {
    u64 diag_off;    // offset (in doubles) of C in the diagonal arena
    u64 state_off;   // offset (in complex) of the state scratch in the state arena
    int n;
    int pad;
};

struct QaoaArgs {
    const QaoaEval *evals;
    const int *order;        // evaluation order (largest work first)
    const int *nevals;       // device scalar: evaluations in this launch
    const QaoaGraph *graphs;
    const double *angles;
    const double *diag;
    double2 *states;         // per-workgroup scratch slots for states larger than LDS
    u64 slot_elems;          // complex elements per scratch slot (2^nmax)
    double *out;
    unsigned *counter;
};

// rx(2 beta) on the pair (a, b): a' = c a - i s b, b' = -i s a + c b (gate_pure.c:221-234), with its
// scalar factor taken out: for |c| >= |s|, rx = c (I - i u X) with u = s / c; else rx = -i s (X + i u I)
// with u = c / s.  Either form costs 4 FMAs per pair instead of 4 products and 4 FMAs; the factor has
// modulus max(|c|, |s|) >= 1/sqrt 2 and, a global scale and phase of the whole layer, drops out of
// <C> once the sum is multiplied by the product of the squared moduli (max(c^2, s^2) per qubit).
struct RxScaled {
    double u;
    int swap;   // 1: the -i s (X + i u I) form
    __device__ RxScaled(double c, double s) {
        swap = fabs(c) < fabs(s);
        u = swap ? c / s : s / c;
    }
};

template <int SWAP>
__device__ __forceinline__ void rx_pair(cd &a, cd &b, double u) {
    if constexpr (!SWAP) {
        const cd na = cmk(fma(u, b.y, a.x), fma(-u, b.x, a.y));
        const cd nb = cmk(fma(u, a.y, b.x), fma(-u, a.x, b.y));
        a = na;
        b = nb;
    } else {
        const cd na = cmk(fma(-u, a.y, b.x), fma(u, a.x, b.y));
        const cd nb = cmk(fma(-u, b.y, a.x), fma(u, b.x, a.y));
        a = na;
        b = nb;
    }
}

// psi *= exp(-i C t) (circ_pure.c:27-32)
__device__ __forceinline__ cd phase(cd v, double d, double t) {
    double s, c;
    sincos(d * t, &s, &c);
    return cmk(c * v.x + s * v.y, c * v.y - s * v.x);
}

// LDS tiles are padded with one slot after every 8 amplitudes, so that a lane
// walking 8 consecutive amplitudes and its neighbours hit distinct banks
__device__ __forceinline__ int pad(int i) { return i + (i >> 3); }
constexpr int padded(int n) { return n + (n >> 3); }

// rx on LDS bits [b0, b1) of a tile of 2^lt amplitudes, qubits in ascending
// order, up to three qubits per pass: each thread holds the R = 2^r amplitudes of one
// r-bit group in registers and applies the r rx in qubit order, so every
// amplitude sees the same sequence of operations as gate-by-gate application.
// R is a template argument: full passes (R = 8) carry no run-time tests.
template <int THREADS, int SWAP, int R>
__device__ __forceinline__ void lds_rx_pass(cd *t, int lt, int b, double u) {
    constexpr int r = R == 8 ? 3 : (R == 4 ? 2 : 1);
    const int ngroups = 1 << (lt - r);
    const int lowmask = (1 << b) - 1;
    for (int gi = threadIdx.x; gi < ngroups; gi += THREADS) {
        const int base = ((gi & ~lowmask) << r) | (gi & lowmask);
        cd v[R];
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = t[pad(base | (j << b))];
#pragma unroll
        for (int k = 0; k < r; ++k) {
#pragma unroll
            for (int j = 0; j < R; ++j)
                if (!((j >> k) & 1)) rx_pair<SWAP>(v[j], v[j | (1 << k)], u);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) t[pad(base | (j << b))] = v[j];
    }
    __syncthreads();
}

template <int THREADS, int SWAP>
__device__ __forceinline__ void lds_rx_form(cd *t, int lt, int b0, int b1, double u) {
    int b = b0;
    for (; b + 3 <= b1; b += 3) lds_rx_pass<THREADS, SWAP, 8>(t, lt, b, u);
    if (b1 - b == 2) lds_rx_pass<THREADS, SWAP, 4>(t, lt, b, u);
    else if (b1 - b == 1) lds_rx_pass<THREADS, SWAP, 2>(t, lt, b, u);
}

// one form per call (the form is the same for every thread), so each instantiation holds one butterfly
template <int THREADS>
__device__ __forceinline__ void lds_rx(cd *t, int lt, int b0, int b1, const RxScaled &rx) {
    if (rx.swap) lds_rx_form<THREADS, 1>(t, lt, b0, b1, rx.u);
    else lds_rx_form<THREADS, 0>(t, lt, b0, b1, rx.u);
}

template <int THREADS>
__device__ __forceinline__ double block_sum(double v, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < THREADS / 64; ++w) s += red[w];
    __syncthreads();
    return s;
}

template <int LT, int THREADS>
__global__ __launch_bounds__(THREADS) void qaoa_persistent(QaoaArgs A) {
    __shared__ cd tile[padded(1 << LT)];
    __shared__ double red[THREADS / 64];
    __shared__ int item;
    constexpr int T = 1 << LT;
    // thread t holds tile elements t + THREADS u, u < PER; global loads go in batches of BT with every load of
    // a batch issued before the first is used (one load at a time left each thread waiting out the full
    // HBM latency per element)
    static_assert(T % THREADS == 0, "tile not a multiple of the workgroup");
    constexpr int PER = T / THREADS, BT = PER < 8 ? PER : 8;
    // loops that load C and evaluate a sincos per element: batches of 4, whose evaluations the compiler
    // interleaves (measured on the 4096-instance batch: 101k evaluations/s with 4, 98k with 2, 74k with 1)
    constexpr int BC = PER < 4 ? PER : 4;
    const int nev = *A.nevals;
    double2 *slot = A.states + (u64)blockIdx.x * A.slot_elems;

    for (;;) {
        if (threadIdx.x == 0) item = (int)atomicAdd(A.counter, 1u);
        __syncthreads();
        const int it = item;
        __syncthreads();
        if (it >= nev) break;
        const int e = A.order[it];
        const QaoaEval ev = A.evals[e];
        const QaoaGraph gr = A.graphs[ev.graph];
        const int n = gr.n;
        const u64 N = (u64)1 << n;
        const double *C = A.diag + gr.diag_off;
        const double *ang = A.angles + ev.angle_off;
        const double amp = 1.0 / sqrt((double)N);
        double acc = 0.0;
        // product over layers of the squared modulus the scaled rx leave out, once per qubit
        double scale = 1.0;
        for (int l = 0; l < ev.p; ++l) {
            const double c = ang[3 * l + 1], s = ang[3 * l + 2];
            const double f2 = fabs(c) < fabs(s) ? s * s : c * c;
            for (int q = 0; q < n; ++q) scale *= f2;
        }

        if (n <= LT) {
            // ---- whole state in LDS ----
            for (int x = threadIdx.x; x < (int)N; x += THREADS) tile[pad(x)] = cmk(amp, 0.0);
            __syncthreads();
            for (int l = 0; l < ev.p; ++l) {
                const double g = ang[3 * l];
                const RxScaled rx(ang[3 * l + 1], ang[3 * l + 2]);
                for (int u0 = 0; u0 < PER; u0 += BC) {
                    double cv[BC];
#pragma unroll
                    for (int u = 0; u < BC; ++u) {
                        const int x = threadIdx.x + (u0 + u) * THREADS;
                        cv[u] = x < (int)N ? C[x] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < BC; ++u) {
                        const int x = threadIdx.x + (u0 + u) * THREADS;
                        if (x < (int)N) tile[pad(x)] = phase(tile[pad(x)], cv[u], g);
                    }
                }
                __syncthreads();
                lds_rx<THREADS>(tile, n, 0, n, rx);
            }
            for (int u0 = 0; u0 < PER; u0 += BT) {
                double cv[BT];
#pragma unroll
                for (int u = 0; u < BT; ++u) {
                    const int x = threadIdx.x + (u0 + u) * THREADS;
                    cv[u] = x < (int)N ? C[x] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < BT; ++u) {
                    const int x = threadIdx.x + (u0 + u) * THREADS;
                    if (x < (int)N) {
                        const cd v = tile[pad(x)];
                        acc += cv[u] * (v.x * v.x + v.y * v.y);
                    }
                }
            }
        } else {
            // ---- state in HBM: p + 1 LDS-tiled passes ----
            // Pass k alternates between A tiles (contiguous: qubits 0..LT-1) and B tiles (all
            // 2^hb high indices x a run of 2^ls low indices: qubits 0..ls-1 and LT..n-1); both
            // hold tile bits ls..LT-1 that the other does not (A: qubits ls..LT-1, B: LT..n-1).
            // Pass k first finishes layer k-1 (rx on tile bits ls..LT-1, the qubits the previous
            // pass could not reach), then starts layer k (phase, rx on every tile bit), so one
            // pass per layer reads and writes the state; pass p reduces the mean instead.
            const int hb = n - LT;                 // qubits above the A tile
            const int ls = 2 * LT - n;             // log2 of the low run of a B tile
            const int S = 1 << ls;
            const u64 nchunk = (u64)1 << hb;
            for (int k = 0; k <= ev.p; ++k) {
                const bool btile = k & 1;
                const bool last = (k == ev.p);
                const double g = last ? 0.0 : ang[3 * k];
                const RxScaled rx(last ? 1.0 : ang[3 * k + 1], last ? 0.0 : ang[3 * k + 2]);
                const RxScaled rxp(k ? ang[3 * (k - 1) + 1] : 1.0, k ? ang[3 * (k - 1) + 2] : 0.0);
                for (u64 ch = 0; ch < nchunk; ++ch) {
                    const u64 lo0 = ch << ls;
                    // global index of tile element x
#define QIDX(x) (btile ? ((((u64)(x) >> ls) << LT) | (lo0 + (u64)((x) & (S - 1)))) : ((ch << LT) + (u64)(x)))
                    // C of the thread's elements of a batch, all loads in flight
#define LOAD_C(cv, u0)                                                                       \
    _Pragma("unroll") for (int u = 0; u < BC; ++u) cv[u] = C[QIDX(threadIdx.x + ((u0) + u) * THREADS)];
                    if (k == 0) {
                        for (int u0 = 0; u0 < PER; u0 += BC) {
                            double cv[BC];
                            LOAD_C(cv, u0)
#pragma unroll
                            for (int u = 0; u < BC; ++u)
                                tile[pad(threadIdx.x + (u0 + u) * THREADS)] = phase(cmk(amp, 0.0), cv[u], g);
                        }
                        __syncthreads();
                    } else {
                        for (int u0 = 0; u0 < PER; u0 += BT) {
                            cd v[BT];
#pragma unroll
                            for (int u = 0; u < BT; ++u) v[u] = slot[QIDX(threadIdx.x + (u0 + u) * THREADS)];
#pragma unroll
                            for (int u = 0; u < BT; ++u) tile[pad(threadIdx.x + (u0 + u) * THREADS)] = v[u];
                        }
                        __syncthreads();
                        lds_rx<THREADS>(tile, LT, ls, LT, rxp);
                        if (!last) {
                            for (int u0 = 0; u0 < PER; u0 += BC) {
                                double cv[BC];
                                LOAD_C(cv, u0)
#pragma unroll
                                for (int u = 0; u < BC; ++u) {
                                    const int x = threadIdx.x + (u0 + u) * THREADS;
                                    tile[pad(x)] = phase(tile[pad(x)], cv[u], g);
                                }
                            }
                            __syncthreads();
                        }
                    }
                    if (!last) {
                        lds_rx<THREADS>(tile, LT, 0, LT, rx);
                        for (int x = threadIdx.x; x < T; x += THREADS) slot[QIDX(x)] = tile[pad(x)];
                    } else {
                        for (int u0 = 0; u0 < PER; u0 += BT) {
                            double cv[BT];
#pragma unroll
                            for (int u = 0; u < BT; ++u) cv[u] = C[QIDX(threadIdx.x + (u0 + u) * THREADS)];
#pragma unroll
                            for (int u = 0; u < BT; ++u) {
                                const cd v = tile[pad(threadIdx.x + (u0 + u) * THREADS)];
                                acc += cv[u] * (v.x * v.x + v.y * v.y);
                            }
                        }
                    }
#undef LOAD_C
#undef QIDX
                    __syncthreads();
                }
            }
        }
        const double tot = block_sum<THREADS>(acc, red);
        if (threadIdx.x == 0) A.out[e] = tot * scale;
        __syncthreads();
    }
}

__global__ void reset_counter(unsigned *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}

// C(x) = sum_(i,j) w_ij [x_i != x_j] for one graph
__global__ __launch_bounds__(256) void maxcut_diag_kernel(double *__restrict__ out, u64 N, int ne,
                                                         const int *__restrict__ ei, const int *__restrict__ ej,
                                                         const double *__restrict__ w) {
    const u64 x = (u64)blockIdx.x * 256 + threadIdx.x;
    if (x >= N) return;
    double acc = 0.0;
    for (int k = 0; k < ne; ++k)
        if (((x >> ei[k]) ^ (x >> ej[k])) & 1) acc += w[k];
    out[x] = acc;
}

template <int LT, int THREADS>
int launch_qaoa(hipStream_t st, const QaoaArgs &a, int grid) {
    reset_counter<<<1, 64, 0, st>>>(a.counter);
    ORK_LAUNCH_CHECK();
    qaoa_persistent<LT, THREADS><<<grid, THREADS, 0, st>>>(a);
    ORK_LAUNCH_CHECK();
    return 0;
}

}  // namespace

// ---- C ABI (used by the host module ork_qaoa.c) ----------------------------------------------
extern "C" int ork_qaoa_launch(void *stream, int lt, int threads, int grid, const void *evals, const int *order,
                               const int *nevals, const void *graphs, const double *angles, const double *diag,
                               void *states, uint64_t slot_elems, double *out, unsigned *counter) {
    QaoaArgs a;
    a.evals = (const QaoaEval *)evals;
    a.order = order;
    a.nevals = nevals;
    a.graphs = (const QaoaGraph *)graphs;
    a.angles = angles;
    a.diag = diag;
    a.states = (double2 *)states;
    a.slot_elems = slot_elems;
    a.out = out;
    a.counter = counter;
    hipStream_t st = (hipStream_t)stream;
#define ORK_QCASE(L, TH) if (lt == L && threads == TH) return launch_qaoa<L, TH>(st, a, grid);
    ORK_QCASE(10, 256) ORK_QCASE(10, 512) ORK_QCASE(10, 1024)
    ORK_QCASE(11, 256) ORK_QCASE(11, 512) ORK_QCASE(11, 1024)
    ORK_QCASE(12, 256) ORK_QCASE(12, 512) ORK_QCASE(12, 1024)
    ORK_QCASE(13, 512) ORK_QCASE(13, 1024)
#undef ORK_QCASE
    return (int)hipErrorInvalidValue;
}

extern "C" int ork_qaoa_occupancy(int lt, int threads, int *blocks_per_cu) {
    hipError_t e = hipErrorInvalidValue;
#define ORK_QOCC(L, TH) \
    if (lt == L && threads == TH) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, qaoa_persistent<L, TH>, TH, 0);
    ORK_QOCC(10, 256) ORK_QOCC(10, 512) ORK_QOCC(10, 1024)
    ORK_QOCC(11, 256) ORK_QOCC(11, 512) ORK_QOCC(11, 1024)
    ORK_QOCC(12, 256) ORK_QOCC(12, 512) ORK_QOCC(12, 1024)
    ORK_QOCC(13, 512) ORK_QOCC(13, 1024)
#undef ORK_QOCC
    return (int)e;
}

extern "C" int ork_maxcut_diag(void *stream, double *out, int n, int ne, const int *ei, const int *ej,
                               const double *w) {
    const u64 N = (u64)1 << n;
    maxcut_diag_kernel<<<(unsigned)((N + 255) / 256), 256, 0, (hipStream_t)stream>>>(out, N, ne, ei, ej, w);
    ORK_LAUNCH_CHECK();
    return 0;
}
