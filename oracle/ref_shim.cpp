// C entry points over the reference's own CPU sources -- TEST INFRASTRUCTURE
// ONLY (built into oracle/_ref/ by `make -C oracle ref`, never shipped in the
// product).  The reference files are compiled where they lie under the
// mounted reference tree; this shim adds nothing to their arithmetic.
//
//   calcGaussianKernel   src/imresh/libs/calcGaussianKernel.cpp:42-91
//   gaussianBlur*        src/imresh/libs/gaussian.cpp:180-197,559-749
//   vectorMax/Min/Sum    src/imresh/algorithms/vectorReduce.cpp:73-119
//   complexNorm/modulus  src/imresh/algorithms/vectorElementwise.cpp:41-81
//   calculateHioError    src/imresh/libs/hybridInputOutput.cpp:67-103
//
// vectorElementwise.cpp and hybridInputOutput.cpp only instantiate their
// templates for FFTW's complex types (absent here), so they are included as
// translation-unit text and instantiated for float[2], the same memory layout
// as fftwf_complex.  The FFTW-dependent bodies stay compiled out, exactly as
// in a reference build without USE_FFTW.
#include "libs/calcGaussianKernel.hpp"
#include "libs/gaussian.hpp"
#include "algorithms/vectorReduce.hpp"
#include "algorithms/vectorElementwise.cpp"
#include "libs/hybridInputOutput.cpp"

typedef float fcomplex[2];

namespace imresh { namespace algorithms {
template void complexNormElementwise<float, fcomplex>(float *const __restrict__,
    fcomplex const *const __restrict__, unsigned int const);
template void applyComplexModulus<fcomplex, float>(fcomplex *const __restrict__,
    float const *const __restrict__, unsigned int const);
} }
namespace imresh { namespace libs {
template float calculateHioError<fcomplex, float>(fcomplex const *const __restrict__,
    float const *const __restrict__, unsigned int const, bool const,
    float *const __restrict__, float *const __restrict__);
} }

extern "C" {

int ref_gauss_kernel_f32(double sigma, float *w, unsigned nmax)
{
    return imresh::libs::calcGaussianKernel<float>(sigma, w, nmax);
}

void ref_blur_h_f32(float *d, unsigned w, unsigned h, double sigma)
{
    imresh::libs::gaussianBlurHorizontal<float>(d, w, h, sigma);
}

void ref_blur_v_f32(float *d, unsigned w, unsigned h, double sigma)
{
    imresh::libs::gaussianBlurVertical<float>(d, w, h, sigma);
}

void ref_blur_f32(float *d, unsigned w, unsigned h, double sigma)
{
    imresh::libs::gaussianBlur<float>(d, w, h, sigma);
}

float ref_vmax_f32(const float *d, unsigned n) { return imresh::algorithms::vectorMax<float>(d, n, 1); }
float ref_vmin_f32(const float *d, unsigned n) { return imresh::algorithms::vectorMin<float>(d, n, 1); }
float ref_vsum_f32(const float *d, unsigned n) { return imresh::algorithms::vectorSum<float>(d, n, 1); }

void ref_complex_norm_f32(float *out, const float *c, unsigned n)
{
    imresh::algorithms::complexNormElementwise<float, fcomplex>(out, (const fcomplex *) c, n);
}

void ref_apply_modulus_f32(float *c, const float *modulus, unsigned n)
{
    imresh::algorithms::applyComplexModulus<fcomplex, float>((fcomplex *) c, modulus, n);
}

float ref_hio_error_f32(const float *gprime, const float *mask, unsigned n, int invert,
                        float *total, float *nmasked)
{
    return imresh::libs::calculateHioError<fcomplex, float>((const fcomplex *) gprime, mask, n,
                                                          invert != 0, total, nmasked);
}

}
