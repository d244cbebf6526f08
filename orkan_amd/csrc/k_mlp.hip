// k_mlp.hip -- BF16 MFMA kernels of the QAOA energy surrogate (an MLP
// 128 -> 512 -> 512 -> 512 -> 1 with GELU): forward, backward and Adam.
//
// The GEMM is C[M x N] = A[M x K] * B[N x K]^T with both operands K-contiguous
// bf16 and fp32 accumulation on v_mfma_f32_32x32x16_bf16 (gfx950).  A 256-thread
// workgroup computes a 128 x 128 tile (2 x 2 waves, each 64 x 64 = 2 x 2 MFMA
// tiles) over K steps of 32, staged through double-buffered LDS.  The epilogue
// is fused: bias + GELU (forward), GELU' (backward), or plain fp32 (weight
// gradients).  Shapes must be multiples of the tile (M, N of 128, K of 32);
// the host pads the batch.
//
// No reference counterpart: this is the learned screening stage on top of the
// simulator.
#include "k_common.hpp"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int KP = BK + 8;            // padded LDS row (bf16 elements)
constexpr int GT = 256;               // threads per GEMM workgroup

enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_GELU_BWD = 2, EPI_F32_ACC = 3 };

// erf(x) by Abramowitz & Stegun, Handbook of Mathematical Functions, 7.1.26 (Hastings):
// 1 - (a1 t + ... + a5 t^5) exp(-x^2), t = 1 / (1 + p |x|), |error| <= 1.5e-7, far below the
// bf16 rounding of the activations it feeds; branch-free, unlike a piecewise erff, which
// dominated the GEMM epilogue.
__device__ __forceinline__ float erf_as(float x) {
    const float ax = fabsf(x);
    // hardware reciprocal (1 ulp): the approximation's own error is 1.5e-7
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.0f));
    const float poly =
        t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    return copysignf(1.0f - poly * __expf(-ax * ax), x);
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erf_as(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erf_as(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * __expf(-0.5f * x * x);
    return cdf + x * pdf;
}

struct EpiArgs {
    float *c32;          // EPI_F32 / EPI_F32_ACC: fp32 C, ldc
    int ldc;
    const float *bias;   // EPI_BIAS_GELU: bias[N]
    float *pre;          // EPI_BIAS_GELU: pre-activation out (fp32, ldc); EPI_GELU_BWD: pre-activation in
    bf16 *act;           // EPI_BIAS_GELU: gelu(pre) out (bf16, ldc); EPI_GELU_BWD: dPre out (bf16, ldc)
    long zstride;        // split K (gemm_nt_kernel, EPI_F32): c32 offset between the gridDim.z slices
};

template <int EPI>
__global__ __launch_bounds__(GT) void gemm_nt_kernel(const bf16 *__restrict__ A, int lda, const bf16 *__restrict__ B,
                                                     int ldb, int K, EpiArgs ep) {
    __shared__ __attribute__((aligned(16))) bf16 As[2][BM * KP];
    __shared__ __attribute__((aligned(16))) bf16 Bs[2][BN * KP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (gridDim.z > 1) {   // split K: slice z covers columns [z K, z K + K) of A and B
        A += (long)blockIdx.z * K;
        B += (long)blockIdx.z * K;
        ep.c32 += (long)blockIdx.z * ep.zstride;
    }
    const int wm = wave >> 1, wn = wave & 1;
    const long m0 = (long)blockIdx.y * BM, n0 = (long)blockIdx.x * BN;
    // global -> register staging: each thread moves 16 bf16 of A and 16 of B per K step
    const int lr = tid >> 1, lk = (tid & 1) * 16;
    const bf16 *ga = A + (m0 + lr) * (long)lda + lk;
    const bf16 *gb = B + (n0 + lr) * (long)ldb + lk;
    uint4 ra0, ra1, rb0, rb1;
#define GEMM_GLOAD(k0)                                  \
    do {                                                \
        ra0 = *(const uint4 *)(ga + (k0));              \
        ra1 = *(const uint4 *)(ga + (k0) + 8);          \
        rb0 = *(const uint4 *)(gb + (k0));              \
        rb1 = *(const uint4 *)(gb + (k0) + 8);          \
    } while (0)
#define GEMM_SSTORE(buf)                                         \
    do {                                                         \
        *(uint4 *)(&As[(buf)][lr * KP + lk]) = ra0;              \
        *(uint4 *)(&As[(buf)][lr * KP + lk + 8]) = ra1;          \
        *(uint4 *)(&Bs[(buf)][lr * KP + lk]) = rb0;              \
        *(uint4 *)(&Bs[(buf)][lr * KP + lk + 8]) = rb1;          \
    } while (0)
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int KT = K / BK;
    GEMM_GLOAD(0);
    GEMM_SSTORE(0);
    __syncthreads();
    const int l32 = lane & 31, lhi = lane >> 5;
    for (int kt = 0; kt < KT; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < KT) GEMM_GLOAD((kt + 1) * BK);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int kk = s * 16 + lhi * 8;
            bf16x8 av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = *(const bf16x8 *)(&As[buf][(wm * 64 + i * 32 + l32) * KP + kk]);
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = *(const bf16x8 *)(&Bs[buf][(wn * 64 + j * 32 + l32) * KP + kk]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < KT) GEMM_SSTORE(buf ^ 1);
        __syncthreads();
    }
#undef GEMM_GLOAD
#undef GEMM_SSTORE
    // epilogue: lane holds C[row][col] with col = l32, row = 8 (r / 4) + 4 lhi + (r % 4)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const long row = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * lhi + (r & 3);
                const long col = n0 + wn * 64 + j * 32 + l32;
                const long at = row * ep.ldc + col;
                const float v = acc[i][j][r];
                if constexpr (EPI == EPI_F32) ep.c32[at] = v;
                else if constexpr (EPI == EPI_F32_ACC) ep.c32[at] += v;
                else if constexpr (EPI == EPI_BIAS_GELU) {
                    const float z = v + ep.bias[col];
                    if (ep.pre) ep.pre[at] = z;
                    ep.act[at] = (bf16)gelu(z);
                } else {
                    ep.act[at] = (bf16)(v * gelu_grad(ep.pre[at]));
                }
            }
}

// ---- 256 x 256 tile variant: 8 waves of 128 x 64 (4 x 2 MFMA tiles each), K steps of 64 ------
// Each operand fragment read from LDS feeds two or four MFMAs; 128 accumulator registers per
// lane leave room for the staging registers without spills, and the 2 waves per SIMD plus the
// 64-deep K step (2048 MFMA cycles per SIMD between barriers) cover the global-load latency of
// the next step.  Used when M, N are multiples of 256 and K of 64.
constexpr int BM2 = 256, BN2 = 256, BK2 = 64, KP2 = BK2 + 8, GT2 = 512;

template <int EPI>
__global__ __launch_bounds__(GT2) void gemm_nt_big_kernel(const bf16 *__restrict__ A, int lda,
                                                          const bf16 *__restrict__ B, int ldb, int K, EpiArgs ep) {
    extern __shared__ __attribute__((aligned(16))) bf16 smem[];
    bf16 *As = smem;                          // [2][BM2 * KP2]
    bf16 *Bs = smem + 2 * BM2 * KP2;          // [2][BN2 * KP2]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;  // 2 x 4 waves: 128 rows x 64 columns each
    const long m0 = (long)blockIdx.y * BM2, n0 = (long)blockIdx.x * BN2;
    // staging: 8 lanes cover the 128 B of one row's K step, a wave 8 rows (fully used cache
    // lines); each thread moves 16 B of rows lr, lr + 64, lr + 128, lr + 192 of A and of B
    const int lr = tid >> 3, lk = (tid & 7) * 8;
    const bf16 *ga = A + (m0 + lr) * (long)lda + lk;
    const bf16 *gb = B + (n0 + lr) * (long)ldb + lk;
    const long sa = 64 * (long)lda, sb = 64 * (long)ldb;
    uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define BIG_GLOAD(k0)                                                                            \
    do {                                                                                         \
        ra0 = *(const uint4 *)(ga + (k0));          ra1 = *(const uint4 *)(ga + sa + (k0));      \
        ra2 = *(const uint4 *)(ga + 2 * sa + (k0)); ra3 = *(const uint4 *)(ga + 3 * sa + (k0));  \
        rb0 = *(const uint4 *)(gb + (k0));          rb1 = *(const uint4 *)(gb + sb + (k0));      \
        rb2 = *(const uint4 *)(gb + 2 * sb + (k0)); rb3 = *(const uint4 *)(gb + 3 * sb + (k0));  \
    } while (0)
#define BIG_SSTORE(buf)                                                                  \
    do {                                                                                 \
        bf16 *pa = As + (buf) * BM2 * KP2 + lr * KP2 + lk;                               \
        bf16 *pb = Bs + (buf) * BN2 * KP2 + lr * KP2 + lk;                               \
        *(uint4 *)(pa) = ra0; *(uint4 *)(pa + 64 * KP2) = ra1;                           \
        *(uint4 *)(pa + 128 * KP2) = ra2; *(uint4 *)(pa + 192 * KP2) = ra3;              \
        *(uint4 *)(pb) = rb0; *(uint4 *)(pb + 64 * KP2) = rb1;                           \
        *(uint4 *)(pb + 128 * KP2) = rb2; *(uint4 *)(pb + 192 * KP2) = rb3;              \
    } while (0)
    f32x16 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const int KT = K / BK2;
    BIG_GLOAD(0);
    BIG_SSTORE(0);
    __syncthreads();
    const int l32 = lane & 31, lhi = lane >> 5;
    for (int kt = 0; kt < KT; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < KT) BIG_GLOAD((kt + 1) * BK2);
        const bf16 *as = As + buf * BM2 * KP2 + (wm * 128 + l32) * KP2;
        const bf16 *bs = Bs + buf * BN2 * KP2 + (wn * 64 + l32) * KP2;
#pragma unroll
        for (int s = 0; s < BK2 / 16; ++s) {
            const int kk = s * 16 + lhi * 8;
            bf16x8 av[4], bv[2];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = *(const bf16x8 *)(as + i * 32 * KP2 + kk);
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = *(const bf16x8 *)(bs + j * 32 * KP2 + kk);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < KT) BIG_SSTORE(buf ^ 1);
        __syncthreads();
    }
#undef BIG_GLOAD
#undef BIG_SSTORE
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const long row = m0 + wm * 128 + i * 32 + 8 * (r >> 2) + 4 * lhi + (r & 3);
                const long col = n0 + wn * 64 + j * 32 + l32;
                const long at = row * ep.ldc + col;
                const float v = acc[i][j][r];
                if constexpr (EPI == EPI_F32) ep.c32[at] = v;
                else if constexpr (EPI == EPI_F32_ACC) ep.c32[at] += v;
                else if constexpr (EPI == EPI_BIAS_GELU) {
                    const float z = v + ep.bias[col];
                    if (ep.pre) ep.pre[at] = z;
                    ep.act[at] = (bf16)gelu(z);
                } else {
                    ep.act[at] = (bf16)(v * gelu_grad(ep.pre[at]));
                }
            }
}
constexpr int BIG_LDS = 2 * (BM2 + BN2) * KP2 * (int)sizeof(bf16);

// ---- 256 x 256 tile, operands streamed into LDS by DMA (global_load_lds), four stages ----------
// Same waves and MFMA tiles as gemm_nt_big_kernel, but K steps of 32 move from global memory
// straight into LDS (no staging registers), three steps ahead of the one being multiplied.  Each
// stage is its own __shared__ array, so the compiler's wait tracking keeps the DMA into one stage
// from blocking reads of another.  A stage row is 64 B (32 bf16); its four 16 B chunks are stored
// XOR-swizzled (position p holds K chunk p ^ ((row >> 2) & 3)), which keeps the 16 rows an MFMA
// operand read touches in distinct banks without padding (a DMA instruction writes 64 contiguous
// 16 B slots, one per lane).
constexpr int BK3 = 32, GT3 = 512;
typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ int swz(int row, int chunk) { return chunk ^ ((row >> 2) & 3); }

template <int EPI>
__global__ __launch_bounds__(GT3) void gemm_nt_dma_kernel(const bf16 *__restrict__ A, int lda,
                                                          const bf16 *__restrict__ B, int ldb, int K, EpiArgs ep) {
    __shared__ __attribute__((aligned(16))) bf16 sa0[BM2 * BK3], sa1[BM2 * BK3], sa2[BM2 * BK3], sa3[BM2 * BK3];
    __shared__ __attribute__((aligned(16))) bf16 sb0[BN2 * BK3], sb1[BN2 * BK3], sb2[BN2 * BK3], sb3[BN2 * BK3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    const long m0 = (long)blockIdx.y * BM2, n0 = (long)blockIdx.x * BN2;
    // DMA: wave w fills rows [32 w, 32 w + 32) of the A and B stage, 16 rows (1 KB) per instruction;
    // lane l lands in row 32 w + 16 i + l / 4 at chunk position l % 4
    const int drow = wave * 32 + (lane >> 2), dp = lane & 3;
    const bf16 *ga0 = A + (m0 + drow) * (long)lda + swz(drow, dp) * 8;
    const bf16 *ga1 = A + (m0 + drow + 16) * (long)lda + swz(drow + 16, dp) * 8;
    const bf16 *gb0 = B + (n0 + drow) * (long)ldb + swz(drow, dp) * 8;
    const bf16 *gb1 = B + (n0 + drow + 16) * (long)ldb + swz(drow + 16, dp) * 8;
    const int dst0 = wave * 32 * BK3, dst1 = (wave * 32 + 16) * BK3;
#define DMA(SA, SB, k0)                                                                         \
    do {                                                                                        \
        __builtin_amdgcn_global_load_lds((const void *)(ga0 + (k0)), (lds_void *)((SA) + dst0), 16, 0, 0); \
        __builtin_amdgcn_global_load_lds((const void *)(ga1 + (k0)), (lds_void *)((SA) + dst1), 16, 0, 0); \
        __builtin_amdgcn_global_load_lds((const void *)(gb0 + (k0)), (lds_void *)((SB) + dst0), 16, 0, 0); \
        __builtin_amdgcn_global_load_lds((const void *)(gb1 + (k0)), (lds_void *)((SB) + dst1), 16, 0, 0); \
    } while (0)
    f32x16 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const int KT = K / BK3;
    const int l32 = lane & 31, lhi = lane >> 5;
    int ra[4], rb[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) ra[i] = wm * 128 + i * 32 + l32;
#pragma unroll
    for (int j = 0; j < 2; ++j) rb[j] = wn * 64 + j * 32 + l32;
#define MMA(SA, SB)                                                                                    \
    do {                                                                                               \
        _Pragma("unroll") for (int s = 0; s < BK3 / 16; ++s) {                                         \
            const int c = 2 * s + lhi;                                                                 \
            bf16x8 av[4], bv[2];                                                                       \
            _Pragma("unroll") for (int i = 0; i < 4; ++i)                                              \
                av[i] = *(const bf16x8 *)((SA) + ra[i] * BK3 + swz(ra[i], c) * 8);                     \
            _Pragma("unroll") for (int j = 0; j < 2; ++j)                                              \
                bv[j] = *(const bf16x8 *)((SB) + rb[j] * BK3 + swz(rb[j], c) * 8);                     \
            _Pragma("unroll") for (int i = 0; i < 4; ++i)                                              \
                _Pragma("unroll") for (int j = 0; j < 2; ++j)                                          \
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i], bv[j], acc[i][j], 0, 0, 0); \
        }                                                                                              \
    } while (0)
    // stage kc sits in (CA, CB).  Every step issues one stage of DMA (past the end of K it reloads
    // the last stage into the slot nobody reads again), so at step kc the stages kc + 1 and kc + 2
    // are the only later ones in flight: wait until this wave has at most those 8 instructions
    // outstanding, barrier (every wave's DMA of stage kc landed and every wave is done with stage
    // kc - 1: its LDS reads fed MFMAs issued before the barrier, and the scheduling barrier keeps
    // them there), refill the slot of stage kc - 1 with stage kc + 3, multiply stage kc.
    // s_waitcnt immediate: vmcnt in bits 3:0, expcnt 6:4 = 7 and lgkmcnt 11:8 = 15 (no wait).  The
    // acquire fence (LDS only) keeps the stage reads below the barrier without the vmcnt(0) that a
    // release fence adds for DMA still in flight.
#define STEP(CA, CB, NA, NB, kc)                                                                       \
    do {                                                                                               \
        __builtin_amdgcn_s_waitcnt(0x0F70 | 8);                                                        \
        __builtin_amdgcn_s_barrier();                                                                  \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");                                \
        DMA(NA, NB, ((kc) + 3 < KT ? (kc) + 3 : KT - 1) * BK3);                                        \
        MMA(CA, CB);                                                                                   \
        __builtin_amdgcn_sched_barrier(0);                                                             \
    } while (0)
    // K is a multiple of 4 BK3 (the host checks), so the loop body runs whole
    DMA(sa0, sb0, 0);
    DMA(sa1, sb1, BK3);
    DMA(sa2, sb2, 2 * BK3);
    for (int kt = 0; kt < KT; kt += 4) {
        STEP(sa0, sb0, sa3, sb3, kt);
        STEP(sa1, sb1, sa0, sb0, kt + 1);
        STEP(sa2, sb2, sa1, sb1, kt + 2);
        STEP(sa3, sb3, sa2, sb2, kt + 3);
    }
    // no DMA may still be writing this workgroup's LDS when it ends
    __builtin_amdgcn_s_waitcnt(0x0F70);
#undef STEP
#undef MMA
#undef DMA
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const long row = m0 + wm * 128 + i * 32 + 8 * (r >> 2) + 4 * lhi + (r & 3);
                const long col = n0 + wn * 64 + j * 32 + l32;
                const long at = row * ep.ldc + col;
                const float v = acc[i][j][r];
                if constexpr (EPI == EPI_F32) ep.c32[at] = v;
                else if constexpr (EPI == EPI_F32_ACC) ep.c32[at] += v;
                else if constexpr (EPI == EPI_BIAS_GELU) {
                    const float z = v + ep.bias[col];
                    if (ep.pre) ep.pre[at] = z;
                    ep.act[at] = (bf16)gelu(z);
                } else {
                    ep.act[at] = (bf16)(v * gelu_grad(ep.pre[at]));
                }
            }
}


// ---- small kernels -----------------------------------------------------------------------------
__global__ void transpose_bf16_kernel(const bf16 *__restrict__ in, bf16 *__restrict__ out, int R, int Cc) {
    __shared__ bf16 t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int y = threadIdx.y; y < 32; y += 8)
        if (by + y < R && bx + threadIdx.x < Cc) t[y][threadIdx.x] = in[(long)(by + y) * Cc + bx + threadIdx.x];
    __syncthreads();
    for (int y = threadIdx.y; y < 32; y += 8)
        if (bx + y < Cc && by + threadIdx.x < R) out[(long)(bx + y) * R + by + threadIdx.x] = t[threadIdx.x][y];
}

__global__ void f32_to_bf16_kernel(const float *__restrict__ in, bf16 *__restrict__ out, long n) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) out[i] = (bf16)in[i];
}

// y[m] = sum_k h[m][k] w[k] + b   (one wave per row); optional loss gradient and
// per-row squared error (target given)
__global__ __launch_bounds__(256) void head_forward_kernel(const bf16 *__restrict__ h, const float *__restrict__ w, float b,
                                                          int M, int K, float *__restrict__ y,
                                                          const float *__restrict__ target, float *__restrict__ dy,
                                                          float *__restrict__ sqerr, int valid, float scale) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= M) return;
    float s = 0.0f;
    if (K % 8 == 0 && ((uintptr_t)h & 15) == 0 && ((uintptr_t)w & 15) == 0) {
        // 16 B per lane: one load instruction covers 512 activations of the row
        for (int k = lane * 8; k < K; k += 512) {
            const bf16x8 hv = *(const bf16x8 *)(h + (long)row * K + k);
            const float4 w0 = *(const float4 *)(w + k), w1 = *(const float4 *)(w + k + 4);
            s += (float)hv[0] * w0.x + (float)hv[1] * w0.y + (float)hv[2] * w0.z + (float)hv[3] * w0.w +
                 (float)hv[4] * w1.x + (float)hv[5] * w1.y + (float)hv[6] * w1.z + (float)hv[7] * w1.w;
        }
    } else {
        for (int k = lane; k < K; k += 64) s += (float)h[(long)row * K + k] * w[k];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0 && (target || row < valid)) {   // forward only: no writes past the valid rows
        const float v = s + b;
        y[row] = v;
        if (target) {
            const float e = (row < valid) ? (v - target[row]) : 0.0f;
            dy[row] = 2.0f * e * scale;
            sqerr[row] = e * e;
        }
    }
}

// dPre3[m][k] = dy[m] w[k] gelu'(pre3[m][k])   (bf16 out)
__global__ void head_backward_kernel(const float *__restrict__ dy, const float *__restrict__ w,
                                     const float *__restrict__ pre, bf16 *__restrict__ dpre, int M, int K) {
    const long n = (long)M * K;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const long m = i / K;
        const int k = (int)(i % K);
        dpre[i] = (bf16)(dy[m] * w[k] * gelu_grad(pre[i]));
    }
}

// out[k] = sum_m a[m][k] * s[m] (a bf16 [M][K]); bias grads use s = null (sum of rows).
// Stage 1: COLSUM_CHUNKS row chunks in parallel into part[chunk][K]; stage 2 adds the
// chunks in a fixed order (deterministic, no atomics).
constexpr int COLSUM_CHUNKS = 64;
__global__ __launch_bounds__(256) void colsum_part_kernel(const bf16 *__restrict__ a, const float *__restrict__ s, int M,
                                                         int K, float *__restrict__ part) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= K) return;
    const int rows = (M + COLSUM_CHUNKS - 1) / COLSUM_CHUNKS;
    const int m0 = blockIdx.y * rows, m1 = min(M, m0 + rows);
    float acc = 0.0f;
    for (int m = m0; m < m1; ++m) acc += (float)a[(long)m * K + k] * (s ? s[m] : 1.0f);
    part[(long)blockIdx.y * K + k] = acc;
}
// 8 columns per thread (16 B loads), 4 row subgroups of a chunk summed through LDS in a fixed order
__global__ __launch_bounds__(256) void colsum_part8_kernel(const bf16 *__restrict__ a, const float *__restrict__ s,
                                                          int M, int K, float *__restrict__ part) {
    __shared__ float red[4][512];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int k0 = blockIdx.x * 512 + tx * 8;
    const int rows = (M + COLSUM_CHUNKS - 1) / COLSUM_CHUNKS;
    const int m0 = blockIdx.y * rows, m1 = min(M, m0 + rows);
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (k0 < K)
        for (int m = m0 + ty; m < m1; m += 4) {
            const bf16x8 v = *(const bf16x8 *)(a + (long)m * K + k0);
            const float sc = s ? s[m] : 1.0f;
#pragma unroll
            for (int t = 0; t < 8; ++t) acc[t] += (float)v[t] * sc;
        }
#pragma unroll
    for (int t = 0; t < 8; ++t) red[ty][tx * 8 + t] = acc[t];
    __syncthreads();
    if (ty == 0 && k0 < K)
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int c = tx * 8 + t;
            part[(long)blockIdx.y * K + k0 + t] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
        }
}

__global__ __launch_bounds__(256) void colsum_finish_kernel(const float *__restrict__ part, int K,
                                                           float *__restrict__ out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= K) return;
    float acc = 0.0f;
    for (int c = 0; c < COLSUM_CHUNKS; ++c) acc += part[(long)c * K + k];
    out[k] = acc;
}

__global__ void sum_kernel(const float *__restrict__ v, int n, float *__restrict__ out) {
    __shared__ float red[4];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) acc += v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}

// Adam on fp32 master weights; refresh the bf16 copies W [K][N] and Wt [N][K]
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, long n, int ncols, float lr, float b1, float b2, float eps,
                            float c1, float c2, bf16 *__restrict__ wb, bf16 *__restrict__ wtb, int nrows) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = (bf16)wi;
        if (wtb) {
            const long r = i / ncols, c = i % ncols;
            wtb[c * nrows + r] = (bf16)wi;
        }
    }
}

// X[r] = [gf (ng floats), angles[r0 + r] (na floats)] as fp32 rows of width ng + na
// ---- first layer for screening ------------------------------------------------------------
// Every candidate row of one graph shares the ng graph features, so their share
// of layer 1 is computed once: u[j] = sum_{k<ng} bf16(gf[k]) W[k][j].  A row's
// pre-activation is then (u[j] + sum_k bf16(a[r][k]) W[ng+k][j]) + b[j]: the same
// bf16 products as the GEMM over the assembled row, summed in fp32 in another
// order.  W is the bf16 [K][N] copy of the layer.
__global__ void first_layer_graph_kernel(const float *__restrict__ gf, int ng, const bf16 *__restrict__ w, int N,
                                         float *__restrict__ u) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (int k = 0; k < ng; ++k) s += (float)(bf16)gf[k] * (float)w[(long)k * N + j];
    u[j] = s;
}

// act[r][j] = gelu(pre) for rows [0, rows) of the chunk starting at candidate r0,
// zero for the padding rows [rows, Mp); four consecutive columns per thread
// (one 8-byte weight load per angle, one 8-byte store).  NA > 0 fixes the angle
// count at compile time so that all loads of a row issue together.
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
template <int NA>
__global__ void first_layer_rows_kernel(const float *__restrict__ angles, int na_rt, long r0, int rows, int Mp,
                                        const float *__restrict__ u, const bf16 *__restrict__ wa,
                                        const float *__restrict__ bias, int N, bf16 *__restrict__ act) {
    const int na = NA > 0 ? NA : na_rt;
    // 8 outputs per thread: 16 B weight loads and one 16 B store
    const unsigned nq = (unsigned)(N >> 3);
    const unsigned total = (unsigned)Mp * nq;
    for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < total; i += gridDim.x * 256u) {
        const unsigned r = i / nq;
        const int j = (int)(i - r * nq) * 8;
        bf16x8 out;
        if ((int)r < rows) {
            const float4 u0 = *(const float4 *)(u + j), u1 = *(const float4 *)(u + j + 4);
            float z[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
            const float *a = angles + (r0 + (long)r) * na;
#pragma unroll 8
            for (int k = 0; k < na; ++k) {
                const float ak = (float)(bf16)a[k];
                const bf16x8 w = *(const bf16x8 *)(wa + (long)k * N + j);
#pragma unroll
                for (int t = 0; t < 8; ++t) z[t] += ak * (float)w[t];
            }
            const float4 b0 = *(const float4 *)(bias + j), b1 = *(const float4 *)(bias + j + 4);
            const float bb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
            for (int t = 0; t < 8; ++t) out[t] = (bf16)gelu(z[t] + bb[t]);
        } else {
#pragma unroll
            for (int t = 0; t < 8; ++t) out[t] = (bf16)0.0f;
        }
        *(bf16x8 *)(act + (long)r * N + j) = out;
    }
}

// ---- top-k selection (radix select on order-preserving float keys) ---------------------------
__device__ __forceinline__ unsigned fkey(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct TopkState {
    unsigned prefix, mask;   // bits of the k-th largest key fixed so far
    long need;               // elements still to take among those matching the prefix
    unsigned hist[256];
    unsigned n_gt, n_eq;     // compaction counters
};

__global__ __launch_bounds__(256) void topk_hist_kernel(const float *__restrict__ y, long n, TopkState *st, int shift) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const unsigned prefix = st->prefix, mask = st->mask;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const unsigned key = fkey(y[i]);
        if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], h[threadIdx.x]);
}

// walk the bins from the largest digit: the k-th largest key has the digit where the
// running count reaches `need`
__global__ void topk_select_kernel(TopkState *st, int shift) {
    if (threadIdx.x != 0) return;
    long cum = 0;
    int digit = 0;
    for (int b = 255; b >= 0; --b) {
        if (cum + (long)st->hist[b] >= st->need) {
            digit = b;
            break;
        }
        cum += st->hist[b];
    }
    st->need -= cum;
    st->prefix |= (unsigned)digit << shift;
    st->mask |= 255u << shift;
    for (int b = 0; b < 256; ++b) st->hist[b] = 0;
}

__global__ void topk_init_kernel(TopkState *st, long k) {
    if (threadIdx.x != 0) return;
    st->prefix = 0;
    st->mask = 0;
    st->need = k;
    st->n_gt = 0;
    st->n_eq = 0;
    for (int b = 0; b < 256; ++b) st->hist[b] = 0;
}

// keys above the threshold all go out; `need` of the keys equal to it (in the
// order the atomics hand out slots -- ties are re-ordered by index on the host)
__global__ __launch_bounds__(256) void topk_compact_kernel(const float *__restrict__ y, long n, TopkState *st,
                                                          int k, int *__restrict__ out) {
    const unsigned T = st->prefix;
    const long need = st->need;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const unsigned key = fkey(y[i]);
        if (key > T) {
            const unsigned s = atomicAdd(&st->n_gt, 1u);
            if (s < (unsigned)k) out[s] = (int)i;
        } else if (key == T) {
            const unsigned s = atomicAdd(&st->n_eq, 1u);
            if ((long)s < need) out[k - 1 - s] = (int)i;
        }
    }
}

unsigned g1(long n) {
    long g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

// ---- C ABI ------------------------------------------------------------------------------------
constexpr int kGemmVariant = 2;
static int g_gemm_big = kGemmVariant;   // 256 x 256 tiles where the shape allows: 1 register-staged, 2 DMA-staged
extern "C" void ork_mlp_gemm_reset(void) { g_gemm_big = kGemmVariant; }

extern "C" int ork_mlp_gemm_select(int variant) {
    if (variant < 0 || variant > 2) return (int)hipErrorInvalidValue;
    g_gemm_big = variant;
    return 0;
}

// epi: 0 fp32 C, 1 bias+GELU, 2 GELU backward, 3 fp32 C +=
extern "C" int ork_mlp_gemm(void *stream, int epi, const void *A, int lda, const void *B, int ldb, int M, int N, int K,
                            float *c32, int ldc, const float *bias, float *pre, void *act) {
    if (M % BM || N % BN || K % BK || M <= 0 || N <= 0 || K <= 0) return (int)hipErrorInvalidValue;
    EpiArgs ep{c32, ldc, bias, pre, (bf16 *)act};
    hipStream_t st = (hipStream_t)stream;
    const bf16 *a = (const bf16 *)A, *b = (const bf16 *)B;
    // the DMA moves 16 B per lane: rows of both operands must start on 16 B boundaries
    const bool aligned16 = lda % 8 == 0 && ldb % 8 == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;
    if (g_gemm_big == 2 && aligned16 && M % BM2 == 0 && N % BN2 == 0 && K % (4 * BK3) == 0 && K > 0) {
        dim3 grid3(N / BN2, M / BM2);
        switch (epi) {
        case EPI_F32: gemm_nt_dma_kernel<EPI_F32><<<grid3, GT3, 0, st>>>(a, lda, b, ldb, K, ep); break;
        case EPI_BIAS_GELU: gemm_nt_dma_kernel<EPI_BIAS_GELU><<<grid3, GT3, 0, st>>>(a, lda, b, ldb, K, ep); break;
        case EPI_GELU_BWD: gemm_nt_dma_kernel<EPI_GELU_BWD><<<grid3, GT3, 0, st>>>(a, lda, b, ldb, K, ep); break;
        case EPI_F32_ACC: gemm_nt_dma_kernel<EPI_F32_ACC><<<grid3, GT3, 0, st>>>(a, lda, b, ldb, K, ep); break;
        default: return (int)hipErrorInvalidValue;
        }
        ORK_LAUNCH_CHECK();
        return 0;
    }
    if (g_gemm_big && M % BM2 == 0 && N % BN2 == 0 && K % BK2 == 0) {
        dim3 grid2(N / BN2, M / BM2);
        switch (epi) {
        case EPI_F32: gemm_nt_big_kernel<EPI_F32><<<grid2, GT2, BIG_LDS, st>>>(a, lda, b, ldb, K, ep); break;
        case EPI_BIAS_GELU: gemm_nt_big_kernel<EPI_BIAS_GELU><<<grid2, GT2, BIG_LDS, st>>>(a, lda, b, ldb, K, ep); break;
        case EPI_GELU_BWD: gemm_nt_big_kernel<EPI_GELU_BWD><<<grid2, GT2, BIG_LDS, st>>>(a, lda, b, ldb, K, ep); break;
        case EPI_F32_ACC: gemm_nt_big_kernel<EPI_F32_ACC><<<grid2, GT2, BIG_LDS, st>>>(a, lda, b, ldb, K, ep); break;
        default: return (int)hipErrorInvalidValue;
        }
        ORK_LAUNCH_CHECK();
        return 0;
    }
    dim3 grid(N / BN, M / BM);
    switch (epi) {
    case EPI_F32: gemm_nt_kernel<EPI_F32><<<grid, GT, 0, st>>>(a, lda, b, ldb, K, ep); break;
    case EPI_BIAS_GELU: gemm_nt_kernel<EPI_BIAS_GELU><<<grid, GT, 0, st>>>(a, lda, b, ldb, K, ep); break;
    case EPI_GELU_BWD: gemm_nt_kernel<EPI_GELU_BWD><<<grid, GT, 0, st>>>(a, lda, b, ldb, K, ep); break;
    case EPI_F32_ACC: gemm_nt_kernel<EPI_F32_ACC><<<grid, GT, 0, st>>>(a, lda, b, ldb, K, ep); break;
    default: return (int)hipErrorInvalidValue;
    }
    ORK_LAUNCH_CHECK();
    return 0;
}

// c[r][col] = sum over the slices, in slice order (deterministic)
__global__ void splitk_sum_kernel(const float *__restrict__ work, int slices, long mn, int N, float *__restrict__ c,
                                  int ldc) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < mn; i += (long)gridDim.x * 256) {
        float s = 0.0f;
        for (int z = 0; z < slices; ++z) s += work[z * mn + i];
        c[(i / N) * ldc + i % N] = s;
    }
}

// C = A B^T with K split into `slices` equal slices on the 128 x 128 tile (fp32 partial products in
// work[slices][M][N], then summed in slice order): the batch-length weight-gradient products, whose
// M x N has only a few tiles
extern "C" int ork_mlp_gemm_splitk(void *stream, const void *A, int lda, const void *B, int ldb, int M, int N, int K,
                                   int slices, float *work, float *c32, int ldc) {
    if (slices < 1 || slices > 65535 || K % slices || (K / slices) % BK || K <= 0 || M % BM || N % BN || M <= 0 ||
        N <= 0)
        return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    EpiArgs ep{work, N, nullptr, nullptr, nullptr, (long)M * N};
    dim3 grid(N / BN, M / BM, slices);
    gemm_nt_kernel<EPI_F32><<<grid, GT, 0, st>>>((const bf16 *)A, lda, (const bf16 *)B, ldb, K / slices, ep);
    ORK_LAUNCH_CHECK();
    splitk_sum_kernel<<<g1((long)M * N), 256, 0, st>>>(work, slices, (long)M * N, N, c32, ldc);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_mlp_transpose(void *stream, const void *in, void *out, int R, int Cc) {
    dim3 grid((Cc + 31) / 32, (R + 31) / 32), blk(32, 8);
    transpose_bf16_kernel<<<grid, blk, 0, (hipStream_t)stream>>>((const bf16 *)in, (bf16 *)out, R, Cc);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_mlp_to_bf16(void *stream, const float *in, void *out, long n) {
    f32_to_bf16_kernel<<<g1(n), 256, 0, (hipStream_t)stream>>>(in, (bf16 *)out, n);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_mlp_head_forward(void *stream, const void *h, const float *w, float b, int M, int K, float *y,
                                    const float *target, float *dy, float *sqerr, int valid, float scale) {
    head_forward_kernel<<<(M + 3) / 4, 256, 0, (hipStream_t)stream>>>((const bf16 *)h, w, b, M, K, y, target, dy,
                                                                      sqerr, valid, scale);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_mlp_head_backward(void *stream, const float *dy, const float *w, const float *pre, void *dpre,
                                     int M, int K) {
    head_backward_kernel<<<g1((long)M * K), 256, 0, (hipStream_t)stream>>>(dy, w, pre, (bf16 *)dpre, M, K);
    ORK_LAUNCH_CHECK();
    return 0;
}

// part: scratch of COLSUM_CHUNKS * K floats
extern "C" int ork_mlp_colsum(void *stream, const void *a, const float *s, int M, int K, float *out, float *part) {
    if (K % 8 == 0 && ((uintptr_t)a & 15) == 0) {
        dim3 g8((K + 511) / 512, COLSUM_CHUNKS);
        colsum_part8_kernel<<<g8, 256, 0, (hipStream_t)stream>>>((const bf16 *)a, s, M, K, part);
    } else {
        dim3 g1d((K + 255) / 256, COLSUM_CHUNKS);
        colsum_part_kernel<<<g1d, 256, 0, (hipStream_t)stream>>>((const bf16 *)a, s, M, K, part);
    }
    ORK_LAUNCH_CHECK();
    colsum_finish_kernel<<<(K + 255) / 256, 256, 0, (hipStream_t)stream>>>(part, K, out);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_mlp_sum(void *stream, const float *v, int n, float *out) {
    sum_kernel<<<1, 256, 0, (hipStream_t)stream>>>(v, n, out);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_mlp_first_layer_graph(void *stream, const float *gf, int ng, const void *w, int N, float *u) {
    first_layer_graph_kernel<<<(N + 255) / 256, 256, 0, (hipStream_t)stream>>>(gf, ng, (const bf16 *)w, N, u);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_mlp_first_layer_rows(void *stream, const float *angles, int na, long r0, int rows, int Mp,
                                        const float *u, const void *wa, const float *bias, int N, void *act) {
    if (N % 8 || (long)Mp * (N / 8) >= (1L << 32)) return (int)hipErrorInvalidValue;
    const unsigned grid = g1((long)Mp * (N / 8));
    if (na == 8)
        first_layer_rows_kernel<8><<<grid, 256, 0, (hipStream_t)stream>>>(angles, na, r0, rows, Mp, u,
                                                                         (const bf16 *)wa, bias, N, (bf16 *)act);
    else
        first_layer_rows_kernel<0><<<grid, 256, 0, (hipStream_t)stream>>>(angles, na, r0, rows, Mp, u,
                                                                         (const bf16 *)wa, bias, N, (bf16 *)act);
    ORK_LAUNCH_CHECK();
    return 0;
}

// indices of the k largest of y_dev[0..n) into idx_dev[0..k) (unordered); st_dev: scratch of
// ork_topk_state_bytes() bytes
extern "C" unsigned long ork_topk_state_bytes(void) { return sizeof(TopkState); }

__global__ void gather_f32_kernel(const float *__restrict__ y, const int *__restrict__ idx, float *__restrict__ out,
                                  int k) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < k) out[i] = y[idx[i]];
}

extern "C" int ork_topk(void *stream, const float *y, long n, int k, int *idx, float *vals, void *st_dev) {
    if (k <= 0 || (long)k > n) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    TopkState *ts = (TopkState *)st_dev;
    long g = (n + 255) / 256;
    if (g > 2048) g = 2048;
    topk_init_kernel<<<1, 64, 0, st>>>(ts, k);
    ORK_LAUNCH_CHECK();
    for (int shift = 24; shift >= 0; shift -= 8) {
        topk_hist_kernel<<<(unsigned)g, 256, 0, st>>>(y, n, ts, shift);
        ORK_LAUNCH_CHECK();
        topk_select_kernel<<<1, 64, 0, st>>>(ts, shift);
        ORK_LAUNCH_CHECK();
    }
    topk_compact_kernel<<<(unsigned)g, 256, 0, st>>>(y, n, ts, k, idx);
    ORK_LAUNCH_CHECK();
    gather_f32_kernel<<<(k + 255) / 256, 256, 0, st>>>(y, idx, vals, k);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_mlp_adam(void *stream, float *w, const float *g, float *m, float *v, long n, int ncols, float lr,
                            float b1, float b2, float eps, float c1, float c2, void *wb, void *wtb, int nrows) {
    adam_kernel<<<g1(n), 256, 0, (hipStream_t)stream>>>(w, g, m, v, n, ncols, lr, b1, b2, eps, c1, c2, (bf16 *)wb,
                                                         (bf16 *)wtb, nrows);
    ORK_LAUNCH_CHECK();
    return 0;
}
