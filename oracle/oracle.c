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
