// Host-side pieces shared by the single-image driver (engine.cpp), the
// batched driver (batch.cpp) and the request server (serve.cpp).
#pragma once

#include <atomic>

#include <vector>

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

// device state of one W x H reconstruction (one allocation: iterate X,
// gPrev G, |g'| Bn, blurred B2, modulus M, mask, partial sums, scalars)
struct Workspace {
    int device = 0;
    int W = 0, H = 0;
    size_t n = 0;
    size_t bytes = 0;           // device bytes of the main allocation
    void *base = nullptr;
    float2 *X = nullptr, *G = nullptr;
    float *Bn = nullptr, *B2 = nullptr, *M = nullptr;
    uint8_t *mask = nullptr;
    float *pmax = nullptr;
    float *perr = nullptr;
    unsigned *pcnt = nullptr;
    float *scal = nullptr;      // [0] threshold, [1..3] error, total, nmasked
    int *flags = nullptr;       // [0] done, [1] cycles
    float *weights = nullptr;   // [cycles][kBlurWeightRow], zero-padded to radius wR
    unsigned wcap = 0;
    int wR = 0;
    float w_sigma0 = -1.f, w_dsigma = -1.f;
    unsigned w_ncycles = 0;
    std::vector<float> host_weights;   // host copy of the weights table
    float *host_wpin = nullptr;        // pinned staging of that table (the upload's source)
    size_t host_wpin_n = 0;            // its capacity in floats
    float *host_stage = nullptr;  // pinned, 64 floats + readback
    // set when a user returns the workspace with work still queued on its
    // stream, cleared by a host function enqueued behind that work.  (An
    // event recorded on the user's stream would do, but HIP reads the
    // stream an event was last recorded on when the event is queried or
    // waited for, and the caller may have destroyed that stream by then.)
    std::atomic<bool> busy{false};
    uint64_t last_use = 0;        // pool clock of the last return (idle trimming)
    // the captured cycle graph and the parameters it was captured with
    hipGraphExec_t gexec = nullptr;
    unsigned g_nhio = 0;
    float g_beta = 0, g_cutA = 0, g_cut = 0, g_target = 0;
    int g_R = -1, g_rowVar = -1, g_colVar = -1;
    const float *g_weights = nullptr;
};

// carves the workspace planes of W x H out of `base` (w NULL: only sizes);
// returns the bytes the layout needs
size_t ws_layout(Workspace *w, int W, int H, void *base);
int ws_alloc(int device, int W, int H, Workspace **out);
// blocks the calling thread until the workspace's previous user's work has
// run (no-op for an idle workspace); false after timeout_s seconds
bool ws_wait_idle(Workspace *w, double timeout_s);
// batch class buffers: free idle ones whose work has finished, least
// recently used first, until the idle bytes fit `cap`; bytes held idle
int trim_idle_classes(size_t cap);
size_t idle_class_bytes();
// bytes that idle pooled buffers of one kind (workspaces, batch buffers) may keep
size_t pool_cap();
void ws_release(Workspace *w);
// device bytes of a workspace (for HBM budgets)
size_t ws_bytes(int W, int H);
int ensure_weights(Workspace *w, const Params &q, hipStream_t s);
// zero-padded per-cycle weights rows (kBlurWeightRow floats each) and their
// common radius
int padded_weights(const Params &q, std::vector<float> &rows, int *R);
// cycles run as captured graphs (default) or as eager launches
extern std::atomic<bool> g_use_graphs;

// one reconstruction in steps (all asynchronous on s): begin = weights,
// modulus M (device) staged into the workspace, state reset, initial
// inverse transform; cycle = blur, threshold, n_hio fused iterations,
// error / early exit (one graph replay; the cycle index is kept on the
// device);
// end = Re(g') -> io_dev.  stage_results copies the scalars to the pinned
// host stage, read_staged decodes them once that copy has completed.
// picks the size's kernel variants on first use; it overwrites the
// workspace's planes (M included), so it runs before the modulus is staged
int ws_tune(Workspace *w, hipStream_t s);
int recon_begin(Workspace *w, const float *M, const Params &q, hipStream_t s);
int recon_cycle(Workspace *w, const float *M, const Params &q, unsigned c, hipStream_t s);
// the launches of one cycle (what recon_cycle replays as a graph)
int enqueue_cycle(Workspace *w, const Params &q, const Plan2D &p, hipStream_t s);
int recon_end(Workspace *w, float *io_dev, hipStream_t s);
int recon_stage_results(Workspace *w, hipStream_t s);
void recon_read_staged(const Workspace *w, imresh_result *res, int *done);
int current_device(int *dev);
// picks (once per device and size) the row / column kernel variants
int tune_size(int device, int W, int H, hipStream_t s);

}  // namespace imr
