// imresh::libs::hybridInputOutput -- drop-in for the reference declaration
// (src/imresh/libs/hybridInputOutput.hpp, hybridInputOutput.cpp:200-306),
// executed on an MI355X through the C ABI in imresh_hip.h
// (imresh_hybrid_input_output).  rSize = {width, height}; rIsMasked uint8,
// 1 = outside the support; rnCycles 0 -> UINT_MAX; rBeta <= 0 -> 0.9;
// rnCores is accepted and ignored.  Returns 0 on success, 1 for null buffers
// or a size that is not 2-D (as the reference), otherwise an imresh_status
// code (also for rnCycles = UINT_MAX without a positive target, which the
// reference would never finish).
#pragma once

#include <climits>
#include <cstdint>
#include <vector>

namespace imresh
{
namespace libs
{

    int hybridInputOutput
    (
        float * const rIoData,
        uint8_t const * const rIsMasked,
        std::vector<unsigned int> const rSize,
        unsigned int rnCycles = UINT_MAX,
        float rTargetError = 1e-6,
        float rBeta = 0.9,
        unsigned int rnCores = 0
    );

} // namespace libs
} // namespace imresh
