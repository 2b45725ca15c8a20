// Shared host/device definitions for the MI355X (gfx950) shrink-wrap/HIO
// library: status codes, launch descriptors and the line-FFT plan layout.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <new>

#include <hip/hip_ext.h>
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
        if (_e != hipSuccess) return imr::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

// a HIP call's status as an imresh status (reported like IMR_HIP_CHECK)
#define IMR_HIP_STATUS(expr) imr::hip_status((expr), #expr, __FILE__, __LINE__)

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

// a failed HIP call checked by IMR_HIP_CHECK: its status, and with
// IMRESH_VERBOSE_ERRORS=1 in the environment a line on stderr naming the
// call and the HIP error
inline int hip_fail(hipError_t e, const char *expr, const char *file, int line)
{
    static const bool verbose = [] {
        const char *v = std::getenv("IMRESH_VERBOSE_ERRORS");
        return v != nullptr && v[0] == '1';
    }();
    if (verbose)
        std::fprintf(stderr, "[imresh] %s:%d: %s -> %s (%s)\n", file, line, expr, hipGetErrorName(e),
                     hipGetErrorString(e));
    return from_hip(e);
}

inline int hip_status(hipError_t e, const char *expr, const char *file, int line)
{
    return e == hipSuccess ? IMRESH_OK : hip_fail(e, expr, file, line);
}

// status of the last kernel launch (hipGetLastError), reported as hip_fail
inline int launch_status(const char *file, int line)
{
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? IMRESH_OK : hip_fail(e, "kernel launch", file, line);
}
}  // namespace imr

namespace imr {
// C-ABI entry points report statuses: a C++ exception (a host allocation or
// thread creation that fails) must not cross the C boundary.
template <class F>
inline int guarded(F &&f) noexcept
{
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return IMRESH_ERR_OUT_OF_MEMORY;
    } catch (...) {
        return IMRESH_ERR_INTERNAL;
    }
}
}  // namespace imr

namespace imr {
// Kernel-exact timing for the profiler entry point: while `stop` is set on
// the calling thread, IMR_LAUNCH records `start` when the first kernel of the
// timed operation begins and `stop` when each kernel ends
// (hipExtLaunchKernelGGL), so the pair brackets exactly the kernels'
// execution, as rocprofv3's kernel trace measures it, also for an operation
// of several launches; otherwise it is a plain launch.
struct ProfEvents {
    hipEvent_t start = nullptr, stop = nullptr;
    bool started = false;
};
extern thread_local ProfEvents g_prof;
}  // namespace imr

#define IMR_LAUNCH(K, grid, block, lds, s, ...)                                                 \
    do {                                                                                        \
        if (imr::g_prof.stop != nullptr) {                                                      \
            hipExtLaunchKernelGGL(K, grid, block, lds, s,                                       \
                                  imr::g_prof.started ? nullptr : imr::g_prof.start,            \
                                  imr::g_prof.stop, 0, __VA_ARGS__);                            \
            imr::g_prof.started = true;                                                         \
        } else {                                                                                \
            hipLaunchKernelGGL(K, grid, block, lds, s, __VA_ARGS__);                            \
        }                                                                                       \
    } while (0)
