// Device-side building blocks of the exposure kernels (gfx950, FP64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eqhip {

constexpr double kDateTol = 1e-9;  // date matching tolerance, products/irs.rs:151-154,173

// B(tau) = (1 - exp(-a tau)) / a, models/hull_white.rs:51-55 (same expression order).
__device__ __forceinline__ double hw_b(double a, double tau) {
    double e = exp(a * (-tau));
    return ((-e) + 1.0) / a;
}

// A(tau), models/hull_white.rs:59-69.
__device__ __forceinline__ double hw_a(double a, double b, double sigma, double tau, double bf) {
    double a2 = a * a;
    double sigma2 = sigma * sigma;
    double term1 = ((bf + (-tau)) * (a2 * b - sigma2 * 0.5)) / a2;
    double term2 = ((sigma2 * bf) * bf) / (a * 4.0);
    return exp(term1 - term2);
}

// e^x for the per-path zero-coupon bonds of the valuation loops. Cody-Waite reduction x = n ln2 + r
// with ln2 split into a 32-bit head (n ln2_hi exact) and a tail, |r| <= ln2/2; then
// e^r = 1 + r(1 + r(1/2 + r q(r))), q the degree-9 Chebyshev fit of (e^r - 1 - r - r^2/2) / r^3 on
// |r| <= 1.0005 ln2/2 (exp_poly.py, 256-bit arithmetic: fit error 0.004 ulp of e^r); 2^n by ldexp.
// Evaluated in FP64 with one rounding per FMA the result is within 0.81 ulp of e^x over |x| <= 708.
// Overflow and underflow come from ldexp (inf, or 0 through the subnormals). For |x| > 1e9 and
// infinities the polynomial is replaced by 1, so the result is ldexp(1, n) with n saturated by the
// conversion: inf or 0. NaN propagates through r.
__device__ __forceinline__ double bond_exp(double x) {
    const double n = __builtin_rint(x * 0x1.71547652b82fep+0);  // log2(e)
    double r = __builtin_fma(-n, 0x1.62e42fee00000p-1, x);  // ln2, head
    r = __builtin_fma(-n, 0x1.a39ef35793c76p-33, r);  // ln2, tail
    double q = __builtin_fma(0x1.1f66f85dfa20ap-29, r, 0x1.af38d53857513p-26);
    q = __builtin_fma(q, r, 0x1.27e4e1ea0f53ap-22);
    q = __builtin_fma(q, r, 0x1.71de0d9c145d0p-19);
    q = __builtin_fma(q, r, 0x1.a01a01a47db6fp-16);
    q = __builtin_fma(q, r, 0x1.a01a01a7c8d47p-13);
    q = __builtin_fma(q, r, 0x1.6c16c16c167dep-10);
    q = __builtin_fma(q, r, 0x1.11111111109adp-7);
    q = __builtin_fma(q, r, 0x1.5555555555555p-5);
    q = __builtin_fma(q, r, 0x1.5555555555556p-3);
    double p = __builtin_fma(q, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    // beyond |x| = 1e9 the result is inf or 0 whatever p's low word holds: the high word of 1.0 is enough
    // (one select), and a NaN x fails the test and propagates
    {
        const uint64_t pb = __builtin_bit_cast(uint64_t, p);
        const uint32_t ph = (__builtin_fabs(x) > 1.0e9) ? 0x3ff00000u : (uint32_t)(pb >> 32);
        p = __builtin_bit_cast(double, ((uint64_t)ph << 32) | (uint32_t)pb);
    }
    return __builtin_amdgcn_ldexp(p, (int)n);  // the conversion saturates and maps NaN to 0
}

// bond_exp with the same operations in the same order, its polynomial steps written as three-operand
// fused multiply-adds: where the coefficients sit in vector registers, the compiler otherwise forms
// each step as a copy of the coefficient into the accumulator followed by a two-operand multiply-add.
__device__ __forceinline__ double fma_v(double a, double b, double c) {  // a b + c
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
__device__ __forceinline__ double bond_exp_v(double x) {
    const double n = __builtin_rint(x * 0x1.71547652b82fep+0);
    double r = __builtin_fma(-n, 0x1.62e42fee00000p-1, x);
    r = __builtin_fma(-n, 0x1.a39ef35793c76p-33, r);
    double q = fma_v(0x1.1f66f85dfa20ap-29, r, 0x1.af38d53857513p-26);
    q = fma_v(q, r, 0x1.27e4e1ea0f53ap-22);
    q = fma_v(q, r, 0x1.71de0d9c145d0p-19);
    q = fma_v(q, r, 0x1.a01a01a47db6fp-16);
    q = fma_v(q, r, 0x1.a01a01a7c8d47p-13);
    q = fma_v(q, r, 0x1.6c16c16c167dep-10);
    q = fma_v(q, r, 0x1.11111111109adp-7);
    q = fma_v(q, r, 0x1.5555555555555p-5);
    q = fma_v(q, r, 0x1.5555555555556p-3);
    double p = __builtin_fma(q, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    // beyond |x| = 1e9 the result is inf or 0 whatever p's low word holds: the high word of 1.0 is enough
    // (one select), and a NaN x fails the test and propagates
    {
        const uint64_t pb = __builtin_bit_cast(uint64_t, p);
        const uint32_t ph = (__builtin_fabs(x) > 1.0e9) ? 0x3ff00000u : (uint32_t)(pb >> 32);
        p = __builtin_bit_cast(double, ((uint64_t)ph << 32) | (uint32_t)pb);
    }
    return __builtin_amdgcn_ldexp(p, (int)n);
}

// G2++ (models/hull_white_2f.rs:58-113, Brigo-Mercurio eq. 4.11), same expression order.
// Own variance of a factor with reversion z and volatility vol.
__device__ __forceinline__ double hw2_var(double z, double vol, double tau) {
    const double inv_z = pow(z, -1.0);
    const double e1 = exp(z * (-tau));
    const double e2 = exp(z * (-2.0 * tau));
    const double bracket = (((inv_z * 2.0) * e1 - (inv_z * 0.5) * e2) - inv_z * 1.5) + tau;
    return ((vol * vol) / (z * z)) * bracket;
}
// Covariance term of the two factors.
__device__ __forceinline__ double hw2_cross(double a, double b, double sigma, double eta, double rho, double tau) {
    const double apb = a + b;
    const double ta = (exp(a * (-tau)) + (-1.0)) / a;
    const double tb = (exp(b * (-tau)) + (-1.0)) / b;
    const double tab = (exp(apb * (-tau)) + (-1.0)) / apb;
    const double bracket = ((ta + tau) + tb) - tab;
    const double coef = ((sigma * eta) * (2.0 * rho)) / (a * b);
    return coef * bracket;
}
// The state-free part of ln P(t, t + tau): -r0 tau + V(tau) / 2, so that
// P = exp((c - B(a,tau) x) - B(b,tau) y).
__device__ __forceinline__ double hw2_c(double a, double b, double sigma, double eta, double rho, double r0,
                                        double tau) {
    const double v = (hw2_var(a, sigma, tau) + hw2_var(b, eta, tau)) + hw2_cross(a, b, sigma, eta, rho, tau);
    return r0 * (-tau) + v * 0.5;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): 10 rounds of two 32x32->64 multiplies, Weyl key schedule.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * c0;
        uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Two standard normals for (path, step pair): counter (pair, path_lo, path_hi, stream), key = seed.
// u1 = ((w_a >> 11) + 1) 2^-53 in (0,1], u2 = (w_b >> 11) 2^-53 in [0,1); Box-Muller. Stream 0 is
// the one-factor model's (two consecutive steps per pair); stream 1 the two-factor model's (the two
// factors' shocks of one step per pair).
__device__ __forceinline__ void philox_normal_pair(uint64_t seed, uint64_t path, uint32_t pair,
                                                   double& z0, double& z1, uint32_t stream = 0u) {
    uint32_t o[4];
    philox4x32_10(pair, (uint32_t)path, (uint32_t)(path >> 32), stream, (uint32_t)seed,
                  (uint32_t)(seed >> 32), o);
    uint64_t wa = ((uint64_t)o[1] << 32) | o[0];
    uint64_t wb = ((uint64_t)o[3] << 32) | o[2];
    const double two_m53 = 1.0 / 9007199254740992.0;
    double u1 = (double)((wa >> 11) + 1u) * two_m53;
    double u2 = (double)(wb >> 11) * two_m53;
    double rad = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586476925286766559 * u2;
    z0 = rad * cos(th);
    z1 = rad * sin(th);
}

}  // namespace eqhip
