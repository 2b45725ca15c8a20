// Device Poisson sampling for the measurement simulation (no reference
// counterpart: the reference feeds recorded or host-computed intensities).
//
// Every value is a fixed sequence of IEEE-754 double operations (add, sub,
// mul, div, floor, frexp, ldexp: all exact or correctly rounded; no fused
// multiply-add, the library is built with -ffp-contract=off), so a CPU
// restatement evaluating the same sequence reproduces every sample bit for
// bit.  Library exp/log/sqrt are not used because their last bits differ
// between implementations; the ones below follow textbook methods:
//   exp  Cody-Waite reduction x = k ln2 + r (ln2 split into a 32-bit head and
//        a tail so k*head is exact), Taylor series of e^r to degree 13
//        (|r| <= ln2/2: truncation < 2^-57 relative), scaling by 2^k;
//   log  x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//        s = (m-1)/(m+1), odd series to s^23 (|s| <= 0.1716);
//   sqrt single-precision estimate refined by two Newton steps in double.
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), counter = (element index, draw, 0),
// key = seed; a double uniform in [0, 1) takes 53 bits from two words.
// Poisson(lambda): inversion by sequential search of the CDF for lambda < 10;
// for lambda >= 10 the transformed rejection method with squeeze (PTRS,
// W. Hoermann, "The transformed rejection method for generating Poisson
// random variables", Insurance: Mathematics and Economics 12 (1993) 39-45)
// with log k! from exact factorials (k < 20) or Stirling's series.  Both
// loops are bounded (1000 CDF steps, 256 rejection rounds), so every thread
// ends; the caps are never reached in practice (probability < 1e-60).
#pragma once

#include <cstdint>

namespace imr {
namespace pois {

constexpr double kLn2Hi = 0x1.62e42fee00000p-1;    // ln 2 to 32 significant bits
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;   // ln 2 - kLn2Hi
constexpr double kInvLn2 = 0x1.71547652b82fep+0;   // 1 / ln 2
constexpr double kHalfLog2Pi = 0.9189385332046728; // log(2 pi) / 2
constexpr double kSqrtHalf = 0.70710678118654752440;

__host__ __device__ inline double exp_d(double x)
{
    if (!(x > -745.0)) return 0.0;   // also NaN -> 0 (not reached: callers pass -lambda, lambda in (0, 10))
    if (x > 709.0) return __builtin_huge_val();
    const double k = floor(x * kInvLn2 + 0.5);
    const double r = (x - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0 / 6227020800.0;   // 1/13!
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int) k);
}

__host__ __device__ inline double log_d(double x)
{
    if (!(x > 0.0)) return x == 0.0 ? -__builtin_huge_val() : __builtin_nan("");
    int e;
    double m = frexp(x, &e);   // x = m 2^e, m in [0.5, 1)
    if (m < kSqrtHalf) {
        m = m * 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z + 1.0;
    const double ed = (double) e;
    return ed * kLn2Hi + (2.0 * s * p + ed * kLn2Lo);
}

__host__ __device__ inline double sqrt_d(double x)
{
    if (!(x > 0.0)) return 0.0;   // callers pass 0 or [1, 1e30] (float-representable estimate)
    double s = (double) sqrtf((float) x);
    s = 0.5 * (s + x / s);
    s = 0.5 * (s + x / s);
    return s;
}

// log k! for an integer-valued k >= 0
__host__ __device__ inline double log_factorial(double k)
{
    if (k < 20.0) {
        double f = 1.0;
        for (int i = 2; i <= (int) k; ++i) f = f * (double) i;   // exact up to 22!
        return log_d(f);
    }
    const double x = k + 1.0;
    const double r = 1.0 / x, r2 = r * r;
    double c = -1.0 / 1680.0;
    c = c * r2 + 1.0 / 1260.0;
    c = c * r2 - 1.0 / 360.0;
    c = c * r2 + 1.0 / 12.0;
    return (x - 0.5) * log_d(x) - x + kHalfLog2Pi + c * r;
}

struct U4 {
    uint32_t x[4];
};

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t) 0xD2511F53u * c.x[0];
        const uint64_t p1 = (uint64_t) 0xCD9E8D57u * c.x[2];
        const uint32_t hi0 = (uint32_t) (p0 >> 32), lo0 = (uint32_t) p0;
        const uint32_t hi1 = (uint32_t) (p1 >> 32), lo1 = (uint32_t) p1;
        c = U4{{hi1 ^ c.x[1] ^ k0, lo1, hi0 ^ c.x[3] ^ k1, lo0}};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// two uniforms in [0, 1) of element `idx`, draw `draw`
__host__ __device__ inline void uniforms(uint64_t seed, uint64_t idx, uint32_t draw, double *u1, double *u2)
{
    const U4 r = philox4x32_10(U4{{(uint32_t) idx, (uint32_t) (idx >> 32), draw, 0u}}, (uint32_t) seed,
                               (uint32_t) (seed >> 32));
    const double two53 = 9007199254740992.0;
    *u1 = ((double) (r.x[0] >> 5) * 67108864.0 + (double) (r.x[1] >> 6)) / two53;
    *u2 = ((double) (r.x[2] >> 5) * 67108864.0 + (double) (r.x[3] >> 6)) / two53;
}

__host__ __device__ inline double poisson(double lam, uint64_t seed, uint64_t idx)
{
    if (!(lam > 0.0)) return 0.0;
    if (lam < 10.0) {
        double u, unused;
        uniforms(seed, idx, 0u, &u, &unused);
        double p = exp_d(-lam), F = p, k = 0.0;
        for (int i = 0; i < 1000 && u > F; ++i) {
            k = k + 1.0;
            p = p * lam / k;
            F = F + p;
        }
        return k;
    }
    const double slam = sqrt_d(lam);
    const double loglam = log_d(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    const double log_invalpha = log_d(invalpha);
    for (uint32_t it = 0; it < 256u; ++it) {
        double u1, v;
        uniforms(seed, idx, it, &u1, &v);
        const double U = u1 - 0.5;
        const double us = 0.5 - fabs(U);
        if (!(us > 0.0)) continue;
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && v <= vr) return k;
        if (k < 0.0 || (us < 0.013 && v > us)) continue;
        const double lhs = log_d(v) + log_invalpha - log_d(a / (us * us) + b);
        const double rhs = -lam + k * loglam - log_factorial(k);
        if (lhs <= rhs) return k;
    }
    return floor(lam + 0.5);
}

}  // namespace pois
}  // namespace imr
