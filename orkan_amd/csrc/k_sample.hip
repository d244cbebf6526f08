// k_sample.hip -- kernels of sample_matrix: the weighted sum of a linear combination of
// operations (acc += w * v) and the block product G = P^dagger diag(obs) Q of up to 64 x 64
// leaf vectors (no reference code: the operation is specified in docs/SAMPLE_MATRIX.md).
//
// Block product: every workgroup owns a fixed x-range of the 2^n amplitudes (the split depends
// on N only), stages 32 amplitudes of every P and Q vector through LDS per step, and each thread
// accumulates a 4 x 4 tile of (a, b) pairs in registers; a second kernel sums the workgroups'
// partials in workgroup order.  Each entry's summation order therefore depends on N alone, not
// on how many vectors a call holds, so results are bit-identical whatever the block size.
#include "k_common.hpp"

namespace {

constexpr int SM_THREADS = 256;
constexpr int SM_MAX = 64;          // vectors per side
constexpr int SM_STEP = 32;         // amplitudes per LDS step
constexpr int SM_STRIDE = SM_STEP + 1;

__global__ __launch_bounds__(SM_THREADS) void axpy_kernel(double2 *__restrict__ acc, const double2 *__restrict__ v,
                                                          double wr, double wi, u64 n, int first) {
    for (u64 i = (u64)blockIdx.x * SM_THREADS + threadIdx.x; i < n; i += (u64)gridDim.x * SM_THREADS) {
        const cd x = v[i];
        const cd t = cmk(wr * x.x - wi * x.y, wr * x.y + wi * x.x);
        acc[i] = first ? t : cadd(acc[i], t);
    }
}

__global__ __launch_bounds__(SM_THREADS) void block_partial_kernel(const double2 *__restrict__ P, int na,
                                                                   const double2 *__restrict__ Q, int nb,
                                                                   const double *__restrict__ obs, u64 N, u64 chunk,
                                                                   double2 *__restrict__ part) {
    __shared__ cd ps[SM_MAX * SM_STRIDE];
    __shared__ cd qs[SM_MAX * SM_STRIDE];
    const int tid = threadIdx.x;
    const int a0 = (tid >> 4) * 4, b0 = (tid & 15) * 4;   // this thread's 4 x 4 tile of (a, b)
    cd acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = cmk(0.0, 0.0);
    const u64 x0 = (u64)blockIdx.x * chunk, x1 = x0 + chunk < N ? x0 + chunk : N;
    for (u64 xs = x0; xs < x1; xs += SM_STEP) {
        for (int idx = tid; idx < SM_MAX * SM_STEP; idx += SM_THREADS) {
            const int v = idx / SM_STEP, xi = idx % SM_STEP;
            const u64 x = xs + (u64)xi;
            const bool in = x < x1;
            ps[v * SM_STRIDE + xi] = (in && v < na) ? P[(u64)v * N + x] : cmk(0.0, 0.0);
            qs[v * SM_STRIDE + xi] = (in && v < nb) ? cscale(Q[(u64)v * N + x], obs[x]) : cmk(0.0, 0.0);
        }
        __syncthreads();
        if (a0 < na && b0 < nb) {
            for (int xi = 0; xi < SM_STEP; ++xi) {
                cd pa[4], qb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) pa[i] = ps[(a0 + i) * SM_STRIDE + xi];
#pragma unroll
                for (int j = 0; j < 4; ++j) qb[j] = qs[(b0 + j) * SM_STRIDE + xi];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) cfma(acc[i][j], cconj(pa[i]), qb[j]);
            }
        }
        __syncthreads();
    }
    double2 *out = part + (u64)blockIdx.x * SM_MAX * SM_MAX;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) out[(a0 + i) * SM_MAX + b0 + j] = acc[i][j];
}

__global__ __launch_bounds__(SM_THREADS) void block_finish_kernel(const double2 *__restrict__ part, int nparts,
                                                                  int na, int nb, double2 *__restrict__ G) {
    const int p = blockIdx.x * SM_THREADS + threadIdx.x;
    if (p >= na * nb) return;
    const int a = p / nb, b = p % nb;
    cd s = cmk(0.0, 0.0);
    for (int w = 0; w < nparts; ++w) s = cadd(s, part[(u64)w * SM_MAX * SM_MAX + a * SM_MAX + b]);
    G[p] = s;
}

}  // namespace

extern "C" unsigned ork_sm_max_block(void) { return SM_MAX; }

// workgroups of the block product for 2^n amplitudes (a function of N only)
extern "C" unsigned ork_sm_parts(uint64_t N) {
    const uint64_t steps = (N + SM_STEP - 1) / SM_STEP;
    return (unsigned)(steps < 256 ? steps : 256);
}

extern "C" int ork_sm_axpy(void *stream, void *acc, const void *v, double wr, double wi, uint64_t n, int first) {
    uint64_t g = (n + SM_THREADS - 1) / SM_THREADS;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    axpy_kernel<<<(unsigned)g, SM_THREADS, 0, (hipStream_t)stream>>>((double2 *)acc, (const double2 *)v, wr, wi, n,
                                                                     first);
    ORK_LAUNCH_CHECK();
    return 0;
}

// G[a * nb + b] = sum_x conj(P[a][x]) obs[x] Q[b][x]; part: ork_sm_parts(N) * 64 * 64 complex
extern "C" int ork_sm_block(void *stream, const void *P, int na, const void *Q, int nb, const double *obs, uint64_t N,
                            void *part, void *G) {
    if (na < 1 || nb < 1 || na > SM_MAX || nb > SM_MAX) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    const unsigned parts = ork_sm_parts(N);
    const uint64_t steps = (N + SM_STEP - 1) / SM_STEP;
    const uint64_t chunk = (steps + parts - 1) / parts * SM_STEP;
    block_partial_kernel<<<parts, SM_THREADS, 0, st>>>((const double2 *)P, na, (const double2 *)Q, nb, obs, N, chunk,
                                                       (double2 *)part);
    ORK_LAUNCH_CHECK();
    block_finish_kernel<<<(na * nb + SM_THREADS - 1) / SM_THREADS, SM_THREADS, 0, st>>>((const double2 *)part, parts,
                                                                                     na, nb, (double2 *)G);
    ORK_LAUNCH_CHECK();
    return 0;
}
