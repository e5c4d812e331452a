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
    // 1/a = conj(a)/|a|^2, with a scaled by the power of two 2^-e that brings
    // max(|re|, |im|) into [0.5, 1) (exact), so |a|^2 stays in range; one division
    using T = decltype(a.re);
    int e;
    frexp(fmax(fabs(a.re), fabs(a.im)), &e);
    const T ar = ldexp(a.re, -e), ai = ldexp(a.im, -e);
    const T inv = T(1) / (ar * ar + ai * ai);
    return C{ldexp(ar * inv, -e), ldexp(-ai * inv, -e)};
}
template <typename C> __device__ __forceinline__ auto gabs(C a) -> decltype(a.re) { return hypot(a.re, a.im); }

template <typename T> struct Mfma;

// Operands: lane l supplies A[l%16][l/16] and B[l/16][l%16] for both types. The
// accumulator layouts differ (measured on gfx950): register v of lane l holds
// D[l/16 + 4v][l%16] for f64 and D[4(l/16) + v][l%16] for f32.
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return lane / 16 + 4 * v; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane / 16) + v; }
};

// One wave: acc (a 16x16 complex tile in the MFMA accumulator layout, real and
// imaginary parts apart) += A (16 x K) B (K x 16), with A and B given as element
// functors a_at(p, k) and b_at(k, q). K runs in chunks of 64: the 16 A values a
// lane needs for a chunk are loaded together (one memory latency per chunk),
// then 64 MFMAs consume them. Lane l supplies A[l%16][k0 + l/16] and
// B[k0 + l/16][l%16]; a complex product is four real MFMAs. b_at must return
// zeros for K <= k < K rounded up to 64, and a_at finite values there.
template <typename T, typename FA, typename FB>
__device__ __forceinline__ void cmma16(typename Mfma<T>::v4& accr, typename Mfma<T>::v4& acci, int K, FA a_at,
                                       FB b_at) {
    const int lane = static_cast<int>(threadIdx.x % 64);
    const int p = lane % 16, kq = lane / 16;
    for (int k0 = 0; k0 < K; k0 += 64) {
        decltype(a_at(0, 0)) a[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) a[u] = a_at(p, k0 + 4 * u + kq);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const auto b = b_at(k0 + 4 * u + kq, p);
            accr = Mfma<T>::mac(a[u].re, b.re, accr);
            accr = Mfma<T>::mac(-a[u].im, b.im, accr);
            acci = Mfma<T>::mac(a[u].re, b.im, acci);
            acci = Mfma<T>::mac(a[u].im, b.re, acci);
        }
    }
}

// Scalar read of lane `src`'s value (src uniform): the broadcasts of the
// register-resident diagonal factorization and
// the ordered reductions.
__device__ __forceinline__ float lane_read(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double lane_read(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), src);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <typename Cx>
__device__ __forceinline__ Cx lane_read(Cx v, int src) {
    return Cx{lane_read(v.re, src), lane_read(v.im, src)};
}

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
