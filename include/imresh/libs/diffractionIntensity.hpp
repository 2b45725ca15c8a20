// imresh::libs::diffractionIntensity -- drop-in for the reference
// declaration (src/imresh/libs/diffractionIntensity.hpp, CUDA version
// diffractionIntensity.cpp:84-135): the real object in rIoData (host memory)
// is replaced by |FFT2((x, 0))|, unnormalized, frequency 0 at [0, 0].  The
// transform runs on the GPU through imresh_diffraction_intensity; the stream
// is a HIP stream (0: the default stream).  With rAsync the call returns
// once the work is queued; the buffer is then valid after the stream has
// been synchronized.
#pragma once

#include <hip/hip_runtime_api.h>

namespace imresh
{
namespace libs
{

    void diffractionIntensity
    (
        float * rIoData,
        unsigned int rImageWidth,
        unsigned int rImageHeight,
        hipStream_t rStream = 0,
        bool rAsync = false
    );

} // namespace libs
} // namespace imresh
