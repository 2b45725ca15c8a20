// Shared host/device definitions for the MI355X (gfx950) shrink-wrap/HIO
// library: status codes, launch descriptors and the line-FFT plan layout.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "imresh_hip.h"

namespace imr {

// Largest number of Stockham passes of a line plan (plans.hpp).
constexpr int kMaxPasses = 8;

}  // namespace imr

#define IMR_HIP_CHECK(expr)                                              \
    do {                                                                 \
        hipError_t _e = (expr);                                          \
        if (_e != hipSuccess) return imr::from_hip(_e);                  \
    } while (0)

#define IMR_CHECK(expr)                                                  \
    do {                                                                 \
        int _s = (expr);                                                 \
        if (_s != IMRESH_OK) return _s;                                  \
    } while (0)

namespace imr {
inline int from_hip(hipError_t e)
{
    if (e == hipSuccess) return IMRESH_OK;
    if (e == hipErrorOutOfMemory) return IMRESH_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return IMRESH_ERR_NO_DEVICE;
    return IMRESH_ERR_HIP;
}
}  // namespace imr
