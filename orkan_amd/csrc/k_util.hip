// k_util.hip -- bulk fills (state_plus) and complex128 <-> complex64 storage
// conversion.
#include "k_common.hpp"

namespace {
constexpr int UT_THREADS = 256;

template <typename S>
__global__ __launch_bounds__(UT_THREADS) void fill_kernel(S *__restrict__ p, u64 len, double re, double im) {
    for (u64 i = (u64)blockIdx.x * UT_THREADS + threadIdx.x; i < len; i += (u64)gridDim.x * UT_THREADS)
        stc(p + i, cmk(re, im));
}

// the D x D corner of the single padded tile of a tiled state with n < LOG_TILE_DIM
template <typename S>
__global__ __launch_bounds__(UT_THREADS) void corner_kernel(S *__restrict__ p, u64 dim, int L, double v) {
    const u64 e = (u64)blockIdx.x * UT_THREADS + threadIdx.x;
    if (e >= dim * dim) return;
    stc(p + ((e / dim) << L) + e % dim, cmk(v, 0.0));
}

template <typename A, typename B>
__global__ __launch_bounds__(UT_THREADS) void convert_kernel(const A *__restrict__ a, B *__restrict__ b, u64 len) {
    for (u64 i = (u64)blockIdx.x * UT_THREADS + threadIdx.x; i < len; i += (u64)gridDim.x * UT_THREADS)
        stc(b + i, ldc(a + i));
}

unsigned grid_len(u64 len) {
    u64 g = (len + UT_THREADS - 1) / UT_THREADS;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}
}  // namespace

extern "C" int ork_fill(void *stream, int prec, void *p, uint64_t len, double re, double im) {
    hipStream_t st = (hipStream_t)stream;
    if (prec == ORK_F32) fill_kernel<float2><<<grid_len(len), UT_THREADS, 0, st>>>((float2 *)p, len, re, im);
    else fill_kernel<double2><<<grid_len(len), UT_THREADS, 0, st>>>((double2 *)p, len, re, im);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_fill_tiled_corner(void *stream, int prec, void *p, int n, int L, double v) {
    hipStream_t st = (hipStream_t)stream;
    const u64 dim = (u64)1 << n;
    const unsigned grid = (unsigned)((dim * dim + UT_THREADS - 1) / UT_THREADS);
    if (prec == ORK_F32) corner_kernel<float2><<<grid, UT_THREADS, 0, st>>>((float2 *)p, dim, L, v);
    else corner_kernel<double2><<<grid, UT_THREADS, 0, st>>>((double2 *)p, dim, L, v);
    ORK_LAUNCH_CHECK();
    return 0;
}

extern "C" int ork_convert(void *stream, const void *src, int sp, void *dst, int dp, uint64_t len) {
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_len(len);
    if (sp == ORK_F64 && dp == ORK_F32) convert_kernel<double2, float2><<<g, UT_THREADS, 0, st>>>((const double2 *)src, (float2 *)dst, len);
    else if (sp == ORK_F32 && dp == ORK_F64) convert_kernel<float2, double2><<<g, UT_THREADS, 0, st>>>((const float2 *)src, (double2 *)dst, len);
    else if (sp == ORK_F64) convert_kernel<double2, double2><<<g, UT_THREADS, 0, st>>>((const double2 *)src, (double2 *)dst, len);
    else convert_kernel<float2, float2><<<g, UT_THREADS, 0, st>>>((const float2 *)src, (float2 *)dst, len);
    ORK_LAUNCH_CHECK();
    return 0;
}
