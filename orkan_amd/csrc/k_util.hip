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

// ---- layout conversion: each destination entry gathers its logical (r, c) from the source ----
// logical entry (r, c) of a stored layout (Hermitian: the upper triangle is the conjugate of the
// stored lower one; a PURE source is the projector |psi><psi|)
template <typename S>
__device__ __forceinline__ cd logical_entry(const S *__restrict__ src, int layout, u64 dim, int L, u64 r, u64 c) {
    if (layout == ORK_PURE) {
        const cd a = ldc(src + r), b = ldc(src + c);
        return cmk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);   // a conj(b)
    }
    const bool up = r < c;
    const u64 rr = up ? c : r, cc = up ? r : c;
    cd v;
    if (layout == ORK_PACKED) {
        v = ldc(src + packed_off(dim, rr, cc));
    } else {
        // tiled: diagonal tiles hold both triangles, so read (r, c) itself when its tile is stored
        const bool stored = (r >> L) >= (c >> L);
        v = stored ? ldc(src + tiled_off(L, r, c)) : ldc(src + tiled_off(L, c, r));
        if (stored) return v;
        return cconj(v);
    }
    return up ? cconj(v) : v;
}

template <typename S>
__global__ __launch_bounds__(UT_THREADS) void relayout_kernel(const S *__restrict__ src, int src_layout,
                                                              S *__restrict__ dst, int dst_layout, int n, int L,
                                                              u64 len) {
    const u64 dim = (u64)1 << n;
    for (u64 e = (u64)blockIdx.x * UT_THREADS + threadIdx.x; e < len; e += (u64)gridDim.x * UT_THREADS) {
        u64 r, c;
        if (dst_layout == ORK_PACKED) {
            // column c holds dim - c entries starting at c (2 dim - c + 1) / 2
            const double b = 2.0 * (double)dim + 1.0;
            u64 cc = (u64)((b - sqrt(b * b - 8.0 * (double)e)) * 0.5);
            while (cc > 0 && cc * (2 * dim - cc + 1) / 2 > e) --cc;
            while ((cc + 1) * (2 * dim - cc) / 2 <= e) ++cc;
            c = cc;
            r = c + (e - c * (2 * dim - c + 1) / 2);
        } else {
            const u64 T = (u64)1 << L, t = e >> (2 * L), w = e & (T * T - 1);
            u64 tr, tc;
            tri_decode(t, tr, tc);
            r = tr * T + (w >> L);
            c = tc * T + (w & (T - 1));
            if (r >= dim || c >= dim) {   // padding of the single tile of n < L
                stc(dst + e, cmk(0.0, 0.0));
                continue;
            }
        }
        stc(dst + e, logical_entry(src, src_layout, dim, L, r, c));
    }
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

// dst (layout dst_layout: packed or tiled) <- the logical matrix of src (pure, packed or tiled),
// both of precision prec; len = stored entries of dst
extern "C" int ork_relayout(void *stream, int prec, const void *src, int src_layout, void *dst, int dst_layout, int n,
                            int log_tile, uint64_t len) {
    if (dst_layout != ORK_PACKED && dst_layout != ORK_TILED) return (int)hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_len(len);
    if (prec == ORK_F32)
        relayout_kernel<float2><<<g, UT_THREADS, 0, st>>>((const float2 *)src, src_layout, (float2 *)dst, dst_layout, n,
                                                         log_tile, len);
    else
        relayout_kernel<double2><<<g, UT_THREADS, 0, st>>>((const double2 *)src, src_layout, (double2 *)dst,
                                                          dst_layout, n, log_tile, len);
    ORK_LAUNCH_CHECK();
    return 0;
}
