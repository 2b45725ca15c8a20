// imresh::libs::fftShift -- drop-in for the reference declaration
// (src/imresh/libs/fftShift.hpp, fftShift.cpp:42-61): host data, shifted on
// the GPU through imresh_fft_shift.  Instantiated, as in the reference, for
// float and for interleaved complex float (float[2], the layout of
// fftwf_complex).
#pragma once

namespace imresh
{
namespace libs
{

    template< class T_COMPLEX >
    void fftShift
    (
        T_COMPLEX * const data,
        unsigned int const Nx,
        unsigned int const Ny
    );

} // namespace libs
} // namespace imresh
