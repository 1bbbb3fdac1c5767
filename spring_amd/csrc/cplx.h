// Small complex helpers over HIP vector types (double2 / float2), plus compensated sums.
#pragma once
#include <hip/hip_runtime.h>

namespace spr {

template <typename R> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };
template <typename R> using cplx_t = typename C2<R>::T;

template <typename R> __device__ __forceinline__ cplx_t<R> cmk(R re, R im) {
    cplx_t<R> z;
    z.x = re;
    z.y = im;
    return z;
}
template <typename Z> __device__ __forceinline__ Z cadd(Z a, Z b) {
    a.x += b.x;
    a.y += b.y;
    return a;
}
// acc + a * b
template <typename Z> __device__ __forceinline__ Z cfma(Z a, Z b, Z acc) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}
// acc + r * b  (r real)
template <typename R, typename Z> __device__ __forceinline__ Z crfma(R r, Z b, Z acc) {
    acc.x = fma(r, b.x, acc.x);
    acc.y = fma(r, b.y, acc.y);
    return acc;
}
template <typename Z> __device__ __forceinline__ auto cabs2(Z a) { return a.x * a.x + a.y * a.y; }

// Neumaier-compensated accumulator
template <typename R> struct KSum {
    R s = 0, c = 0;
    __device__ __forceinline__ void add(R x) {
        R t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};

}  // namespace spr
