// The allocation-free shrink-wrap entry point (imresh_shrinkwrap_device_noalloc).
//
// Same launches as imresh_shrinkwrap_device (engine.cpp), but the
// reconstruction's device state -- the workspace planes, partial sums,
// scalars, flags and the per-cycle blur weights -- lives in `scratch`, a
// fixed-size device array of this translation unit, instead of a pooled
// hipMalloc'ed workspace.  The call allocates no device or pinned memory,
// records no memset (state is cleared by kernels) and launches the cycles
// one by one (no graph of its own), so it can run inside a caller's stream
// capture.  Weights and zeroed scalars reach the device through kernel
// arguments (launch_store_words), not through staged copies.
//
// Every call on a device uses that device's one copy of `scratch`: calls
// that overlap on one device (other streams, other threads) race.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "engine.hpp"

using namespace imr;

namespace {

// a 2048 x 2048 reconstruction needs 121.7 MB of the 128 MiB
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr unsigned kMaxCycles = IMRESH_NOALLOC_MAX_CYCLES;   // weights rows kept in the scratch
constexpr size_t kWeightsBytes = sizeof(float) * kBlurWeightRow * kMaxCycles;
static_assert(kScratchBytes - kWeightsBytes == IMRESH_NOALLOC_MAX_BYTES, "limit stated in imresh_hip.h");

// In host code `scratch` names a host-side stand-in; the device address of
// each device's copy comes from hipGetSymbolAddress (scratch_ptr).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mtx;
std::map<int, void *> g_addr;   // device -> address of its `scratch`

int scratch_ptr(int dev, void **out)
{
    std::lock_guard<std::mutex> lk(g_mtx);
    auto it = g_addr.find(dev);
    if (it != g_addr.end()) {
        *out = it->second;
        return IMRESH_OK;
    }
    void *p = nullptr;
    IMR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) return IMRESH_ERR_INTERNAL;
    g_addr[dev] = p;
    *out = p;
    return IMRESH_OK;
}

// the zero-padded per-cycle weights table of padded_weights (engine.cpp)
// into a fixed host array: row c holds cycle c's taps centred in 2R+1 slots
int weights_rows(const Params &q, float *rows, int *R)
{
    float sig[kMaxCycles];
    int taps[kMaxCycles];
    float sigma = q.sigma0;
    int rmax = 0;
    for (unsigned c = 0; c < q.n_cycles; ++c) {
        sig[c] = sigma;
        sigma = std::fmax(1.5f, (1.0f - q.dsigma) * sigma);
        taps[c] = gaussian_weights(sig[c], nullptr, 0);
        if (taps[c] > 2 * kMaxBlurRadius + 1) return IMRESH_ERR_KERNEL_TOO_LARGE;
        rmax = taps[c] / 2 > rmax ? taps[c] / 2 : rmax;
    }
    const unsigned nrows = q.n_cycles > 0 ? q.n_cycles : 1;
    std::memset(rows, 0, sizeof(float) * kBlurWeightRow * nrows);
    for (unsigned c = 0; c < q.n_cycles; ++c) {
        float w[2 * kMaxBlurRadius + 1];
        gaussian_weights(sig[c], w, 2 * kMaxBlurRadius + 1);
        const int r = taps[c] / 2;
        for (int k = 0; k < taps[c]; ++k) rows[(size_t) c * kBlurWeightRow + (rmax - r) + k] = w[k];
    }
    *R = rmax;
    return IMRESH_OK;
}

int shrinkwrap_noalloc(float *io_dev, uint32_t width, uint32_t height, const imresh_params *params,
                       void *stream, imresh_result *result)
{
    if (io_dev == nullptr) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    Params q;
    IMR_CHECK(normalize(params, q));
    if (q.n_cycles > kMaxCycles) return IMRESH_ERR_TOO_LARGE;
    const size_t need = ws_layout(nullptr, (int) width, (int) height, nullptr);
    if (need + kWeightsBytes > kScratchBytes) return IMRESH_ERR_TOO_LARGE;
    int dev;
    IMR_CHECK(current_device(&dev));
    // the size's FFT tables must exist already (built by an earlier call of
    // this size or by imresh_autotune): creating them would allocate
    Plan2D p;
    IMR_CHECK(find_plan2d(dev, (int) width, (int) height, &p));
    void *base;
    IMR_CHECK(scratch_ptr(dev, &base));
    hipStream_t s = (hipStream_t) stream;

    Workspace w;
    w.device = dev;
    ws_layout(&w, (int) width, (int) height, base);
    w.weights = (float *) ((char *) base + kScratchBytes - kWeightsBytes);
    int R = 0;
    float rows[kBlurWeightRow * kMaxCycles];   // 64 KiB host copy of the weights table
    IMR_CHECK(weights_rows(q, rows, &R));
    w.wR = R;
    const unsigned nrows = q.n_cycles > 0 ? q.n_cycles : 1;
    IMR_CHECK(launch_store_words(w.weights, rows, (size_t) kBlurWeightRow * nrows, s));
    // scalars and flags start at zero (kernels, not memsets: capturable)
    IMR_CHECK(launch_store_words(w.scal, nullptr, 64 / sizeof(float), s));
    IMR_CHECK(launch_store_words((float *) w.flags, nullptr, 64 / sizeof(float), s));

    // recon_begin: modulus into the workspace, initial inverse transform
    IMR_HIP_CHECK(hipMemcpyAsync(w.M, io_dev, sizeof(float) * w.n, hipMemcpyDeviceToDevice, s));
    ColArgs ca{};
    ca.X = w.X;
    ca.M = w.M;
    ca.done = w.flags;
    IMR_CHECK(launch_col(COL_INIT, p, ca, s));
    RowArgs ra{};
    ra.X = w.X;
    ra.G = w.G;
    ra.mask = w.mask;
    ra.Bn = w.Bn;
    ra.B2 = w.B2;
    ra.thr = w.scal;
    ra.beta = q.beta;
    ra.perr = w.perr;
    ra.pcnt = w.pcnt;
    ra.done = w.flags;
    IMR_CHECK(launch_row(ROW_INIT, p, ra, s));
    for (unsigned c = 0; c < q.n_cycles; ++c) IMR_CHECK(enqueue_cycle(&w, q, p, s));
    IMR_CHECK(launch_copy_from_real(io_dev, w.X, w.n, s));
    if (result != nullptr) {
        // scalars {threshold, error, total, masked} and flags {done, cycles}
        float sc[4];
        int fl[2];
        IMR_HIP_CHECK(hipMemcpyAsync(sc, w.scal, sizeof sc, hipMemcpyDeviceToHost, s));
        IMR_HIP_CHECK(hipMemcpyAsync(fl, w.flags, sizeof fl, hipMemcpyDeviceToHost, s));
        IMR_HIP_CHECK(hipStreamSynchronize(s));
        result->cycles_run = (uint32_t) fl[1];
        result->last_error = sc[1];
        result->last_total = sc[2];
        result->last_masked = sc[3];
    }
    return IMRESH_OK;
}

}  // namespace

extern "C" {

int imresh_shrinkwrap_device_noalloc(float *io_dev, uint32_t width, uint32_t height,
                                     const imresh_params *params, void *stream, imresh_result *result)
{
    return guarded([&] { return shrinkwrap_noalloc(io_dev, width, height, params, stream, result); });
}

}  // extern "C"
