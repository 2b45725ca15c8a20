// imresh::algorithms::shrinkWrap -- drop-in for the reference declaration
// (src/imresh/algorithms/shrinkWrap.hpp:40-53), executed on an MI355X through
// the C ABI in imresh_hip.h (imresh_shrinkwrap).  Same parameters, defaults
// and "<= 0 / == 0 selects the default" rule (shrinkWrap.cpp:133-140;
// rnCycles has no such default).  Returns 0 on success, 1 for a null io
// buffer (as the reference), otherwise an imresh_status code.
#pragma once

namespace imresh
{
namespace algorithms
{

    int shrinkWrap
    (
        float * const rIoData,
        unsigned int const rImageWidth,
        unsigned int const rImageHeight,
        unsigned int rnCycles = 20,
        float rTargetError = 1e-5,
        float rHioBeta = 0.9,
        float rIntensityCutOffAutoCorel = 0.04,
        float rIntensityCutOff = 0.20,
        float sigma0 = 3.0,
        float rSigmaChange = 0.01,
        unsigned int rnHioCycles = 20
    );

} // namespace algorithms
} // namespace imresh
