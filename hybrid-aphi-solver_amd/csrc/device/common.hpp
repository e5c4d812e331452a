// Device-side helpers shared by the A-Phi kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aphi_dev {

struct cd {  // complex<double>, interleaved (re, im) like std::complex<double>
    double re, im;
};
struct cf {  // complex<float>
    float re, im;
};

template <typename T> struct cplx;
template <> struct cplx<double> { using type = cd; };
template <> struct cplx<float> { using type = cf; };

// ---------------------------------------------------------------------------
// Arithmetic that mirrors the host's std::complex<double> evaluation exactly
// (used by assembly so the device and the reference round identically). The
// callers build with floating-point contraction off.

__device__ __forceinline__ cd cmul(cd a, cd b) {  // (a*c - b*d, a*d + b*c)
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }

// Smith's complex division with the branch structure of libgcc's __divdc3
// (no scaling), which is what `std::complex<double>::operator/` compiles to.
__device__ __forceinline__ cd cdiv(cd x, cd y) {
    const double a = x.re, b = x.im, c = y.re, d = y.im;
    double ratio, denom;
    if (fabs(c) < fabs(d)) {
        ratio = c / d;
        denom = (c * ratio) + d;
        return {((a * ratio) + b) / denom, ((b * ratio) - a) / denom};
    }
    ratio = d / c;
    denom = (d * ratio) + c;
    return {((b * ratio) + a) / denom, (b - (a * ratio)) / denom};
}

// ---------------------------------------------------------------------------
// Generic complex helpers for the factorization (T = float or double). These
// are free to contract into FMAs.
template <typename C> __device__ __forceinline__ C cz() { return C{0, 0}; }

template <typename C> __device__ __forceinline__ C gmul(C a, C b) {
    return C{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename C> __device__ __forceinline__ C gsub(C a, C b) { return C{a.re - b.re, a.im - b.im}; }
template <typename C> __device__ __forceinline__ C gadd(C a, C b) { return C{a.re + b.re, a.im + b.im}; }
// a - b*c
template <typename C> __device__ __forceinline__ C gfms(C a, C b, C c) {
    return C{a.re - b.re * c.re + b.im * c.im, a.im - b.re * c.im - b.im * c.re};
}
// a + b*c
template <typename C> __device__ __forceinline__ C gfma(C a, C b, C c) {
    return C{a.re + b.re * c.re - b.im * c.im, a.im + b.re * c.im + b.im * c.re};
}
template <typename C> __device__ __forceinline__ C grecip(C a) {
    // 1/a with a scaling that keeps |a|^2 in range
    using T = decltype(a.re);
    const T s = fabs(a.re) > fabs(a.im) ? fabs(a.re) : fabs(a.im);
    const T ar = a.re / s, ai = a.im / s;
    const T den = (ar * ar + ai * ai) * s;
    return C{ar / den, -ai / den};
}
template <typename C> __device__ __forceinline__ auto gabs(C a) -> decltype(a.re) { return hypot(a.re, a.im); }

// Order-preserving bit patterns of non-negative doubles, for atomic max.
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(addr), static_cast<unsigned long long>(__double_as_longlong(v)));
}

// Opt a kernel into more than 64 KiB of dynamic LDS (gfx950 has 160 KiB per CU).
template <typename K>
inline void allow_large_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(bytes));
}

}  // namespace aphi_dev
