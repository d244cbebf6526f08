// k_common.hpp -- shared device helpers of the gfx950 kernels: complex128
// arithmetic, storage conversion (complex128 or complex64 in HBM, complex128
// in registers), index enumeration and layout offsets.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ork_kernels.h"

typedef uint64_t u64;
typedef double2 cd;   // complex128 in registers: x = re, y = im

#define ORK_SQRT1_2 0.70710678118654752440

__device__ __forceinline__ cd cmk(double re, double im) { cd v; v.x = re; v.y = im; return v; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cscale(cd a, double s) { return cmk(a.x * s, a.y * s); }
// products are fused explicitly (fma of one product into the other), so every kernel that forms a
// given product rounds it the same way whatever the surrounding code lets the compiler contract
__device__ __forceinline__ cd cmul(cd a, cd b) { return cmk(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)); }
__device__ __forceinline__ cd cconj(cd a) { return cmk(a.x, -a.y); }
__device__ __forceinline__ cd cneg(cd a) { return cmk(-a.x, -a.y); }
// multiply by i^k, k in 0..3 (exact)
__device__ __forceinline__ cd cipow(cd a, int k) {
    switch (k & 3) {
    case 0: return a;
    case 1: return cmk(-a.y, a.x);
    case 2: return cmk(-a.x, -a.y);
    default: return cmk(a.y, -a.x);
    }
}
// acc += a * b
__device__ __forceinline__ void cfma(cd &acc, cd a, cd b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}

// ---- storage element types ---------------------------------------------------------------
template <typename S> struct Store;
template <> struct Store<double2> {
    static __device__ __forceinline__ cd ld(const double2 *p) { return *p; }
    static __device__ __forceinline__ void st(double2 *p, cd v) { *p = v; }
    // nontemporal (streaming) forms: the two halves merge into one 16 B access with the nt bit
    static __device__ __forceinline__ cd ldn(const double2 *p) {
        return cmk(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
    }
    static __device__ __forceinline__ void stn(double2 *p, cd v) {
        __builtin_nontemporal_store(v.x, &p->x);
        __builtin_nontemporal_store(v.y, &p->y);
    }
};
template <> struct Store<float2> {
    static __device__ __forceinline__ cd ld(const float2 *p) {
        float2 f = *p;
        return cmk((double)f.x, (double)f.y);
    }
    static __device__ __forceinline__ void st(float2 *p, cd v) {
        float2 f;
        f.x = (float)v.x;
        f.y = (float)v.y;
        *p = f;
    }
    static __device__ __forceinline__ cd ldn(const float2 *p) {
        return cmk((double)__builtin_nontemporal_load(&p->x), (double)__builtin_nontemporal_load(&p->y));
    }
    static __device__ __forceinline__ void stn(float2 *p, cd v) {
        __builtin_nontemporal_store((float)v.x, &p->x);
        __builtin_nontemporal_store((float)v.y, &p->y);
    }
};
// raw storage loads, converted later: a complex64 load whose float -> double conversion follows at once inside a
// condition makes the compiler wait for it there, one load in flight per thread; loading the raw values of a
// whole batch without conditions first and converting afterwards keeps every load of the batch in flight
template <int NT, typename S> __device__ __forceinline__ S ldr(const S *p) {
    if constexpr (NT) {
        S v;
        v.x = __builtin_nontemporal_load(&p->x);
        v.y = __builtin_nontemporal_load(&p->y);
        return v;
    } else {
        return *p;
    }
}
// an empty register barrier: a conversion of v cannot move above it, so one placed after a batch of
// conditional loads keeps the compiler from sinking each conversion into its load's condition (and waiting
// for that load there)
template <typename S> __device__ __forceinline__ void hold(S &v) { asm volatile("" : "+v"(v.x), "+v"(v.y)); }
__device__ __forceinline__ cd to_cd(double2 v) { return v; }
__device__ __forceinline__ cd to_cd(float2 v) { return cmk((double)v.x, (double)v.y); }
template <typename S> __device__ __forceinline__ cd ldc(const S *p) { return Store<S>::ld(p); }
template <typename S> __device__ __forceinline__ void stc(S *p, cd v) { Store<S>::st(p, v); }
// lane exchange (within a wavefront) of one complex value as four 32-bit shuffles
__device__ __forceinline__ double shfl_xor_f64(double v, int m) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __shfl_xor((int)(unsigned)b, m, 64), hi = __shfl_xor((int)(unsigned)(b >> 32), m, 64);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ cd shfl_xor_cd(cd v, int m) { return cmk(shfl_xor_f64(v.x, m), shfl_xor_f64(v.y, m)); }
// the value of lane `src` (any lane of the wavefront; every lane takes part)
__device__ __forceinline__ double shfl_f64(double v, int src) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __shfl((int)(unsigned)b, src, 64), hi = __shfl((int)(unsigned)(b >> 32), src, 64);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ cd shfl_cd(cd v, int src) { return cmk(shfl_f64(v.x, src), shfl_f64(v.y, src)); }
// component-wise select (a select of the vector type itself goes through memory)
__device__ __forceinline__ cd sel(bool c, cd a, cd b) { return cmk(c ? a.x : b.x, c ? a.y : b.y); }

// NT selects the nontemporal forms
template <int NT, typename S> __device__ __forceinline__ cd ldx(const S *p) {
    if constexpr (NT) return Store<S>::ldn(p);
    else return Store<S>::ld(p);
}
template <int NT, typename S> __device__ __forceinline__ void stx(S *p, cd v) {
    if constexpr (NT) Store<S>::stn(p, v);
    else Store<S>::st(p, v);
}

// ---- index enumeration -------------------------------------------------------------------
// insert a 0 bit at position `pos`
__device__ __forceinline__ u64 ins0(u64 v, int pos) {
    const u64 low = v & ((((u64)1) << pos) - 1);
    return low | ((v ^ low) << 1);
}
// insert 0 bits at the ascending positions p[0..k-1]
__device__ __forceinline__ u64 ins0k(u64 v, const int *p, int k) {
    for (int j = 0; j < k; ++j) v = ins0(v, p[j]);
    return v;
}

// ---- layout offsets ----------------------------------------------------------------------
// packed lower triangle, column-major, element (r, c) with r >= c
__device__ __forceinline__ u64 packed_off(u64 dim, u64 r, u64 c) {
    return c * (2 * dim - c + 1) / 2 + (r - c);
}
// tiled, element (r, c) whose tile (r >> L, c >> L) is a stored (lower or diagonal) tile
__device__ __forceinline__ u64 tiled_off(int L, u64 r, u64 c) {
    const u64 tr = r >> L, tc = c >> L, m = (((u64)1) << L) - 1;
    return ((tr * (tr + 1) / 2 + tc) << (2 * L)) + ((r & m) << L) + (c & m);
}

// triangular decode: w -> (R, C) with R >= C, w = R (R + 1) / 2 + C
__device__ __forceinline__ void tri_decode(u64 w, u64 &R, u64 &C) {
    u64 r = (u64)((sqrt(8.0 * (double)w + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > w) --r;
    while ((r + 1) * (r + 2) / 2 <= w) ++r;
    R = r;
    C = w - r * (r + 1) / 2;
}

#define ORK_LAUNCH_CHECK() \
    do { hipError_t e_ = hipGetLastError(); if (e_ != hipSuccess) return (int)e_; } while (0)
