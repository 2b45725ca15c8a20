// Host-side pieces shared by the single-image driver (engine.cpp), the
// batched driver (batch.cpp) and the request server (serve.cpp).
#pragma once

#include "internal.hpp"

namespace imr {

// call parameters after the reference's defaulting rules
// (shrinkWrap.cpp:133-140 for the CPU path; taskQueue.cu / cudaShrinkWrap.cu
// for the queue path: nCycles 0 -> 20, cutOffAutoCorel -> 0.02)
struct Params {
    unsigned n_cycles, n_hio;
    float target, beta, cutA, cut, sigma0, dsigma;
};

int normalize(const imresh_params *p, Params &q);
int current_device(int *dev);
// picks (once per device and size) the row / column kernel variants
int tune_size(int device, int W, int H, hipStream_t s);

}  // namespace imr
