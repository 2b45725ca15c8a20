/*
 * CPU oracle for the shrink-wrap / HIO hot path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * See oracle.h for the reference functions restated here.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif

#define ORC_MAX_RADIX 61

/* FFT threads: 1 by default, as the reference plans FFTW single-threaded
 * (shrinkWrap.cpp:152-155, no fftw_plan_with_nthreads). */
static int g_fft_threads = 1;

void orc_set_fft_threads(int n) { g_fft_threads = n > 0 ? n : 1; }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_sigma_schedule(float sigma0, float sigma_change, unsigned n, float *out)
{
    float sigma = sigma0;
    for (unsigned i = 0; i < n; ++i) {
        out[i] = sigma;
        sigma = fmaxf(1.5f, (1 - sigma_change) * sigma);
    }
}

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(x) CAT(x, S)

#define T float
#define S f32
#define SQRT sqrtf
#include "oracle_impl.h"
#undef T
#undef S
#undef SQRT

#define T double
#define S f64
#define SQRT sqrt
#include "oracle_impl.h"
#undef T
#undef S
#undef SQRT

/* fftShift (src/imresh/libs/fftShift.cpp:42-61): for iy < Ny/2 and every ix,
 * swap element (iy, ix) with ((iy + Ny/2) % Ny, (ix + Nx/2) % Nx); elements
 * of elem_bytes bytes (4: float, 8: complex float) */
void orc_fft_shift(void *data, unsigned nx, unsigned ny, unsigned elem_bytes)
{
    unsigned char *d = (unsigned char *) data;
    unsigned char tmp[16];
    for (unsigned iy = 0; iy < ny / 2; ++iy)
        for (unsigned ix = 0; ix < nx; ++ix) {
            const size_t a = (size_t) iy * nx + ix;
            const size_t b = (size_t) ((iy + ny / 2) % ny) * nx + (ix + nx / 2) % nx;
            memcpy(tmp, d + a * elem_bytes, elem_bytes);
            memcpy(d + a * elem_bytes, d + b * elem_bytes, elem_bytes);
            memcpy(d + b * elem_bytes, tmp, elem_bytes);
        }
}

/* ------------------------------------------------------------------------ */
/* measurement simulation (no reference counterpart): the specification the
 * device sampler implements, restated here in plain C (double arithmetic,
 * -ffp-contract=off, SSE2 on x86-64: every operation IEEE-rounded).        */
/* ------------------------------------------------------------------------ */

/* Philox4x32-10 (Salmon et al., SC'11): 10 rounds of two 32x32->64
 * multiplications with M0 = 0xD2511F53 (word 0) and M1 = 0xCD9E8D57
 * (word 2); the key grows by the Weyl constants 0x9E3779B9 / 0xBB67AE85
 * after every round */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t a = (uint64_t) 0xD2511F53u * c0;
        const uint64_t b = (uint64_t) 0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t) (b >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t) b;
        const uint32_t n2 = (uint32_t) (a >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t) a;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* e^x: x = k ln2 + r with ln2 = head (32 significant bits) + tail, the
 * Taylor polynomial of e^r of degree 13 by Horner, times 2^k */
double orc_exp_d(double x)
{
    static const double head = 0x1.62e42fee00000p-1, tail = 0x1.a39ef35793c76p-33;
    if (!(x > -745.0)) return 0.0;
    if (x > 709.0) return HUGE_VAL;
    const double k = floor(x * 0x1.71547652b82fep+0 + 0.5);
    const double r = (x - k * head) - k * tail;
    double fact = 1.0;
    double inv[14];
    /* 1/n! for n = 0..13, each a single correctly rounded division */
    for (int n = 0; n < 14; ++n) {
        if (n > 1) fact *= n;
        inv[n] = 1.0 / fact;
    }
    double p = inv[13];
    for (int n = 12; n >= 0; --n) p = p * r + inv[n];
    return ldexp(p, (int) k);
}

/* log x: x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 * s = (m-1)/(m+1), series sum_{j=0}^{11} s^(2j+1)/(2j+1) by Horner in s^2 */
double orc_log_d(double x)
{
    static const double head = 0x1.62e42fee00000p-1, tail = 0x1.a39ef35793c76p-33;
    if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : NAN;
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    for (int d = 21; d >= 1; d -= 2) p = p * z + 1.0 / (double) d;
    return (double) e * head + (2.0 * s * p + (double) e * tail);
}

/* sqrt: single-precision estimate, two Newton steps in double */
double orc_sqrt_d(double x)
{
    if (!(x > 0.0)) return 0.0;
    double s = (double) sqrtf((float) x);
    s = 0.5 * (s + x / s);
    s = 0.5 * (s + x / s);
    return s;
}

/* log k!: exact factorial then log for k < 20, else Stirling's series
 * (x - 1/2) log x - x + log(2 pi)/2 + 1/(12x) - 1/(360x^3) + 1/(1260x^5)
 * - 1/(1680x^7), x = k + 1, the bracket evaluated as ((c7 r2 + c5) r2 + c3) r2 + c1) r */
double orc_log_factorial(double k)
{
    if (k < 20.0) {
        double f = 1.0;
        for (int i = 2; i <= (int) k; ++i) f *= (double) i;
        return orc_log_d(f);
    }
    const double x = k + 1.0, r = 1.0 / x, r2 = r * r;
    const double c = ((((-1.0 / 1680.0) * r2 + 1.0 / 1260.0) * r2 - 1.0 / 360.0) * r2 + 1.0 / 12.0);
    return (x - 0.5) * orc_log_d(x) - x + 0.9189385332046728 + c * r;
}

/* uniforms u1, u2 in [0,1) of element idx, draw d: 53 bits each from words
 * (0,1) and (2,3) of Philox(counter = (idx lo, idx hi, d, 0), key = seed) */
static void orc_uniforms(uint64_t seed, uint64_t idx, uint32_t d, double *u1, double *u2)
{
    const uint32_t ctr[4] = {(uint32_t) idx, (uint32_t) (idx >> 32), d, 0u};
    const uint32_t key[2] = {(uint32_t) seed, (uint32_t) (seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    *u1 = ((double) (o[0] >> 5) * 67108864.0 + (double) (o[1] >> 6)) / 9007199254740992.0;
    *u2 = ((double) (o[2] >> 5) * 67108864.0 + (double) (o[3] >> 6)) / 9007199254740992.0;
}

/* Poisson(lambda) of element idx: CDF inversion (lambda < 10, at most 1000
 * steps) or Hoermann's PTRS (lambda >= 10, at most 256 rounds; constants
 * b = 0.931 + 2.53 sqrt(l), a = -0.059 + 0.02483 b, 1/alpha = 1.1239 +
 * 1.1328/(b - 3.4), v_r = 0.9277 - 3.6224/(b - 2)) */
double orc_poisson(double lam, uint64_t seed, uint64_t idx)
{
    if (!(lam > 0.0)) return 0.0;
    if (lam < 10.0) {
        double u, unused;
        orc_uniforms(seed, idx, 0u, &u, &unused);
        double p = orc_exp_d(-lam), F = p, k = 0.0;
        for (int i = 0; i < 1000 && u > F; ++i) {
            k += 1.0;
            p = p * lam / k;
            F = F + p;
        }
        return k;
    }
    const double sl = orc_sqrt_d(lam), ll = orc_log_d(lam);
    const double b = 0.931 + 2.53 * sl;
    const double a = -0.059 + 0.02483 * b;
    const double ia = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    const double lia = orc_log_d(ia);
    for (uint32_t it = 0; it < 256u; ++it) {
        double u, v;
        orc_uniforms(seed, idx, it, &u, &v);
        const double U = u - 0.5;
        const double us = 0.5 - fabs(U);
        if (!(us > 0.0)) continue;
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && v <= vr) return k;
        if (k < 0.0 || (us < 0.013 && v > us)) continue;
        if (orc_log_d(v) + lia - orc_log_d(a / (us * us) + b) <= -lam + k * ll - orc_log_factorial(k))
            return k;
    }
    return floor(lam + 0.5);
}

/* sum of m_i^2 in double in the device's fixed order: lanes j < 65536 add
 * elements j, j + 65536, ... in sequence; groups of 256 lanes, then the 256
 * group sums, combine by the halving tree s[t] += s[t + h], h = 128..1 */
double orc_sq_sum_fixed(const float *m, size_t n)
{
    double grp[256];
    for (int g = 0; g < 256; ++g) {
        double s[256];
        for (int t = 0; t < 256; ++t) {
            double acc = 0.0;
            for (size_t i = (size_t) g * 256 + t; i < n; i += 65536) {
                const double v = (double) m[i];
                acc = acc + v * v;
            }
            s[t] = acc;
        }
        for (int h = 128; h > 0; h >>= 1)
            for (int t = 0; t < h; ++t) s[t] = s[t] + s[t + h];
        grp[g] = s[0];
    }
    for (int h = 128; h > 0; h >>= 1)
        for (int t = 0; t < h; ++t) grp[t] = grp[t] + grp[t + h];
    return grp[0];
}

/* m_i <- (float) sqrt(Poisson((photons * m_i^2) / total)), total = orc_sq_sum_fixed */
void orc_poisson_modulus(float *m, size_t n, double photons, uint64_t seed)
{
    const double t = orc_sq_sum_fixed(m, n);
    for (size_t i = 0; i < n; ++i) {
        const double v = (double) m[i];
        const double lam = t > 0.0 ? photons * (v * v) / t : 0.0;
        m[i] = (float) orc_sqrt_d(orc_poisson(lam, seed, (uint64_t) i));
    }
}

/* n samples: out[j] = orc_poisson(lam[j], seed, idx0 + j) (test helper) */
void orc_poisson_many(const double *lam, size_t n, uint64_t seed, uint64_t idx0, double *out)
{
    for (size_t j = 0; j < n; ++j) out[j] = orc_poisson(lam[j], seed, idx0 + j);
}
