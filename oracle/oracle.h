/*
 * CPU oracle for the shrink-wrap / HIO hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.
 * The product (libimresh_hip.so) never links, loads or falls back to it.
 *
 * It restates, in plain C, the reference CPU path and the arithmetic of the
 * reference CUDA kernels (file:line citations are relative to the reference
 * repository root):
 *   shrinkWrap            src/imresh/algorithms/shrinkWrap.cpp:111-361
 *   complexNorm/modulus   src/imresh/algorithms/vectorElementwise.cpp:41-81
 *   vectorMax/Min/Sum     src/imresh/algorithms/vectorReduce.cpp:73-119
 *   gaussian blur         src/imresh/libs/gaussian.cpp:51-197,559-749
 *   calcGaussianKernel    src/imresh/libs/calcGaussianKernel.cpp:42-91
 *   calculateHioError     src/imresh/libs/hybridInputOutput.cpp:67-103
 *   hybridInputOutput     src/imresh/libs/hybridInputOutput.cpp:126-306
 *   fftShift              src/imresh/libs/fftShift.cpp:42-61
 *   CUDA kernels          src/imresh/algorithms/cuda/cudaVectorElementwise.cu:42-219
 * The FFT (FFTW3f in the reference, absent here) is replaced by a
 * project-written mixed-radix + Bluestein CPU FFT, unnormalized, with
 * forward = exp(-2*pi*i*j*k/n) as FFTW_FORWARD.
 *
 * Every function exists for float (_f32) and double (_f64).  Complex arrays
 * are interleaved (re, im) pairs, row-major H rows x W columns, DC at [0,0].
 *
 * Parity pins (what each function is checked against) are listed in
 * tests/test_oracle.py; every function here has at least one.
 */
#ifndef IMRESH_ORACLE_H
#define IMRESH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sw_stats {
    unsigned cycles_run;    /* shrink-wrap cycles executed (early exit aware) */
    float    last_error;    /* HIO error of the last executed cycle           */
    float    last_total;    /* sum mask*|g'| of the last executed cycle        */
    float    last_masked;   /* sum mask of the last executed cycle             */
} orc_sw_stats;

#define ORC_DECLARE(T, S)                                                          \
    int  orc_gauss_kernel_##S(double sigma, T *w, unsigned nmax);                  \
    void orc_blur_h_##S(T *d, unsigned w, unsigned h, double sigma);               \
    void orc_blur_v_##S(T *d, unsigned w, unsigned h, double sigma);               \
    void orc_blur_##S(T *d, unsigned w, unsigned h, double sigma);                 \
    T    orc_vmax_##S(const T *d, unsigned n);                                     \
    T    orc_vmin_##S(const T *d, unsigned n);                                     \
    T    orc_vsum_##S(const T *d, unsigned n);                                     \
    void orc_complex_norm_##S(T *out, const T *c, unsigned n);                     \
    void orc_apply_modulus_##S(T *c, const T *modulus, unsigned n);                \
    void orc_hio_constraint_##S(T *gprev, const T *gprime, const T *mask,          \
                                unsigned n, T beta);                               \
    void orc_cutoff_##S(T *d, unsigned n, T threshold, T lower, T upper);          \
    float orc_hio_error_##S(const T *gprime, const T *mask, unsigned n,            \
                            int invert, float *total, float *nmasked);             \
    int  orc_fft_##S(T *out, const T *in, unsigned n, int sign);                   \
    int  orc_fft2d_##S(T *out, const T *in, unsigned w, unsigned h, int sign);     \
    void orc_copy_to_real_##S(T *c, const T *x, unsigned n);                       \
    void orc_copy_from_real_##S(T *x, const T *c, unsigned n);                     \
    int  orc_make_mask_##S(const T *gprime, T *mask, unsigned w, unsigned h,       \
                           float sigma, float cut);                                \
    int  orc_hio_iterations_##S(T *gprime, T *gprev, const T *mask,                \
                                const T *modulus, unsigned w, unsigned h,          \
                                T beta, unsigned n_iter);                          \
    int  orc_shrinkwrap_##S(T *io, unsigned w, unsigned h, unsigned n_cycles,      \
                            float target_error, float hio_beta,                    \
                            float cutoff_autocorel, float cutoff, float sigma0,    \
                            float sigma_change, unsigned n_hio,                    \
                            orc_sw_stats *stats);                                  \
    int  orc_diffraction_intensity_##S(T *io, unsigned w, unsigned h);             \
    void orc_random_phase_##S(const T *norm, T *out, unsigned w, unsigned h);      \
    int  orc_hio_standalone_##S(T *io, const unsigned char *mask, unsigned w,      \
                                unsigned h, unsigned n_cycles, float target,       \
                                float beta, orc_sw_stats *stats);

ORC_DECLARE(float, f32)
ORC_DECLARE(double, f64)

/* measurement simulation (the specification of the device sampler):
 * Philox4x32-10, exp/log/sqrt/log k! from IEEE operations, Poisson(lambda)
 * per element from the counter-based stream (seed, index), the fixed-order
 * sum of squares and the whole modulus transformation */
void   orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_exp_d(double x);
double orc_log_d(double x);
double orc_sqrt_d(double x);
double orc_log_factorial(double k);
double orc_poisson(double lam, uint64_t seed, uint64_t idx);
double orc_sq_sum_fixed(const float *m, size_t n);
void   orc_poisson_modulus(float *m, size_t n, double photons, uint64_t seed);
void   orc_poisson_many(const double *lam, size_t n, uint64_t seed, uint64_t idx0, double *out);

/* fftShift (fftShift.cpp:42-61) on elements of elem_bytes (4 or 8) */
void orc_fft_shift(void *data, unsigned nx, unsigned ny, unsigned elem_bytes);

/* sigma schedule of the reference: sigma <- max(1.5, (1 - dsigma) * sigma),
 * in float (shrinkWrap.cpp:245, cudaShrinkWrap.cu:294).  Writes n values. */
void orc_sigma_schedule(float sigma0, float sigma_change, unsigned n, float *out);

/* number of OpenMP threads the elementwise loops use (1 without OpenMP) */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
