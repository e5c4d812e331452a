// Counter-based normal variates, the same code on the host and the device: the
// Philox-4x32-10 block cipher (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011), two 53-bit uniforms per block, and the
// Box-Muller transform with log, sqrt, sin and cos written out in IEEE
// additions, multiplications and divisions (no library transcendental, no
// contraction), so both sides produce the same bits.
//
// Normal number k of stream `seed` comes from counter block k / 2 (counter
// words: its low and high 32 bits, then 0, 0; key words: the seed's low and
// high 32 bits): u1 = ((w1:w0 >> 12) + 0.5) 2^-52 in (0, 1), u2 = (w3:w2 >> 11)
// 2^-53 in [0, 1), r = sqrt(-2 log u1); k even takes r cos(2 pi u2), k odd
// r sin(2 pi u2).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// (each function below turns floating-point contraction off in its own body)
namespace aphi_rng {

struct Block {
    uint32_t w[4];
};

__host__ __device__ inline Block philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    return Block{{c0, c1, c2, c3}};
}

// x = m 2^e with m in [0.5, 1) for x > 0 (exact)
__host__ __device__ inline double split_exp(double x, int& e) {
    uint64_t b;
    __builtin_memcpy(&b, &x, sizeof(b));
    const int be = static_cast<int>((b >> 52) & 0x7ff);
    e = be - 1022;
    b = (b & 0x800fffffffffffffull) | (static_cast<uint64_t>(1022) << 52);
    double m;
    __builtin_memcpy(&m, &b, sizeof(m));
    return m;
}

__host__ __device__ inline double times_pow2(double x, int e) {  // x 2^e for |e| < 1000, normal result
#pragma clang fp contract(off)
    const uint64_t b = static_cast<uint64_t>(e + 1023) << 52;
    double s;
    __builtin_memcpy(&s, &b, sizeof(s));
    return x * s;
}

// natural log of x in (0, 1]: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// log x = e log 2 + 2 atanh(s), s = (m - 1) / (m + 1), |s| < 0.172, series to s^23
__host__ __device__ inline double det_log(double x) {
#pragma clang fp contract(off)
    int e;
    double m = split_exp(x, e);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = 1.0 / 21.0 + s2 * p;
    p = 1.0 / 19.0 + s2 * p;
    p = 1.0 / 17.0 + s2 * p;
    p = 1.0 / 15.0 + s2 * p;
    p = 1.0 / 13.0 + s2 * p;
    p = 1.0 / 11.0 + s2 * p;
    p = 1.0 / 9.0 + s2 * p;
    p = 1.0 / 7.0 + s2 * p;
    p = 1.0 / 5.0 + s2 * p;
    p = 1.0 / 3.0 + s2 * p;
    const double atanh2 = 2.0 * (s + s * (s2 * p));
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    return (static_cast<double>(e) * ln2_hi + (static_cast<double>(e) * ln2_lo + atanh2));
}

// square root of y > 0 by six Newton steps from a linear start on [0.5, 2)
__host__ __device__ inline double det_sqrt(double y) {
#pragma clang fp contract(off)
    if (!(y > 0.0)) return 0.0;
    int e;
    double m = split_exp(y, e);
    if (e & 1) {
        m = m * 2.0;
        e = e - 1;
    }
    double g = 0.5 + 0.5 * m;
    for (int i = 0; i < 6; ++i) g = 0.5 * (g + m / g);
    return times_pow2(g, e / 2);
}

// sin and cos of 2 pi u for u in [0, 1): quadrant q = floor(4u), the remaining
// angle phi = pi/4 + x with x = (4u - q - 1/2) pi/2 in [-pi/4, pi/4), Taylor series
// of sin x and cos x to x^19 / x^20, then the rotations by pi/4 and q pi/2
__host__ __device__ inline void det_sincos_2pi(double u, double& sn, double& cs) {
#pragma clang fp contract(off)
    const double t = u * 4.0;
    const int q = static_cast<int>(t);
    const double x = ((t - static_cast<double>(q)) - 0.5) * 1.57079632679489661923;
    const double x2 = x * x;
    double ps = -1.0 / 121645100408832000.0;           // -1/19!
    ps = 1.0 / 355687428096000.0 + x2 * ps;             // 1/17!
    ps = -1.0 / 1307674368000.0 + x2 * ps;              // -1/15!
    ps = 1.0 / 6227020800.0 + x2 * ps;                  // 1/13!
    ps = -1.0 / 39916800.0 + x2 * ps;                   // -1/11!
    ps = 1.0 / 362880.0 + x2 * ps;                      // 1/9!
    ps = -1.0 / 5040.0 + x2 * ps;                       // -1/7!
    ps = 1.0 / 120.0 + x2 * ps;                         // 1/5!
    ps = -1.0 / 6.0 + x2 * ps;                          // -1/3!
    const double sx = x + x * (x2 * ps);
    double pc = 1.0 / 2432902008176640000.0;            // 1/20!
    pc = -1.0 / 6402373705728000.0 + x2 * pc;           // -1/18!
    pc = 1.0 / 20922789888000.0 + x2 * pc;              // 1/16!
    pc = -1.0 / 87178291200.0 + x2 * pc;                // -1/14!
    pc = 1.0 / 479001600.0 + x2 * pc;                   // 1/12!
    pc = -1.0 / 3628800.0 + x2 * pc;                    // -1/10!
    pc = 1.0 / 40320.0 + x2 * pc;                       // 1/8!
    pc = -1.0 / 720.0 + x2 * pc;                        // -1/6!
    pc = 1.0 / 24.0 + x2 * pc;                          // 1/4!
    pc = -0.5 + x2 * pc;                                // -1/2!
    const double cx = 1.0 + x2 * pc;
    const double h = 0.70710678118654752440;            // sin(pi/4) = cos(pi/4)
    const double sp = (sx + cx) * h, cp = (cx - sx) * h;  // sin, cos of pi/4 + x
    switch (q & 3) {
        case 0: sn = sp; cs = cp; break;
        case 1: sn = cp; cs = -sp; break;
        case 2: sn = -sp; cs = -cp; break;
        default: sn = -cp; cs = sp; break;
    }
}

// uniforms 2b and 2b + 1 of stream `seed`, in [0, 1): the 53-bit words (w1:w0 >> 11)
// and (w3:w2 >> 11) of counter block b, times 2^-53 (exact)
__host__ __device__ inline void uniform_pair(uint64_t seed, uint64_t b, double& u0, double& u1) {
    const uint32_t ctr[4] = {static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32), 0u, 0u};
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    const Block r = philox4x32_10(ctr, key);
    const double inv53 = 1.0 / 9007199254740992.0;
    u0 = static_cast<double>(((static_cast<uint64_t>(r.w[1]) << 32) | r.w[0]) >> 11) * inv53;
    u1 = static_cast<double>(((static_cast<uint64_t>(r.w[3]) << 32) | r.w[2]) >> 11) * inv53;
}

// normals 2b and 2b + 1 of stream `seed`
__host__ __device__ inline void normal_pair(uint64_t seed, uint64_t b, double& z0, double& z1) {
#pragma clang fp contract(off)
    const uint32_t ctr[4] = {static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32), 0u, 0u};
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    const Block r = philox4x32_10(ctr, key);
    const uint64_t a = ((static_cast<uint64_t>(r.w[1]) << 32) | r.w[0]) >> 12;
    const uint64_t c = ((static_cast<uint64_t>(r.w[3]) << 32) | r.w[2]) >> 11;
    const double inv52 = 1.0 / 4503599627370496.0;  // 2^-52
    const double inv53 = 1.0 / 9007199254740992.0;  // 2^-53
    const double u1 = (static_cast<double>(a) + 0.5) * inv52;  // exact: a + 0.5 has 53 bits
    const double u2 = static_cast<double>(c) * inv53;
    const double rad = det_sqrt(-2.0 * det_log(u1));
    double sn, cs;
    det_sincos_2pi(u2, sn, cs);
    z0 = rad * cs;
    z1 = rad * sn;
}

}  // namespace aphi_rng
