// C-ABI implementation: workspaces, the shrink-wrap driver and the
// per-operator entry points.  See include/imresh_hip.h for the interface.
//
// Driver (one reconstruction, all on one HIP stream, no host round trip
// until the end -- the reference synchronized twice per cycle for the max
// and the error, cudaShrinkWrap.cu:274,368):
//
//   COL_INIT  : M -> inverse column FFT                  (cudaShrinkWrap.cu:229-232)
//   ROW_INIT  : inverse row FFT -> g' , gPrev = g', |g'|  (:232, :304-309, :240)
//   per cycle c:
//     blur2d       |g'| -> blurred + tile max              (:242, :274)
//     max_finalize threshold = cut_c * max                 (:275-276)
//     HIO loop:    ROW_FIRST_CUT (mask cut-off, :277 + constraint :314) or
//                  ROW_MID (inverse rows, constraint) ; COL_PROJ (FFT, modulus, IFFT)
//     ROW_LAST     inverse rows -> g', |g'| for the next blur, error sums (:368)
//     err_finalize error, early-exit flag (:374-375)
//   copy_from_real  Re(g') -> io                           (:401)
//
// The early exit sets a device flag that turns every later kernel of the
// reconstruction into a no-op, so the host never waits inside the loop.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "engine.hpp"

using namespace imr;

namespace {

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }


}  // namespace

namespace imr {

size_t ws_layout(Workspace *w, int W, int H, void *base)
{
    const size_t n = (size_t) W * H;
    const int nblur = blur_grid(W, H);
    const int nrow = H;   // >= number of row workgroups
    size_t off = 0;
    const size_t oX = off; off = align256(off + 8 * n);
    const size_t oG = off; off = align256(off + 8 * n);
    const size_t oBn = off; off = align256(off + 4 * n);
    const size_t oB2 = off; off = align256(off + 4 * n);
    const size_t oM = off; off = align256(off + 4 * n);
    const size_t oMask = off; off = align256(off + n);
    const size_t oPmax = off; off = align256(off + 4 * (size_t) nblur);
    const size_t oPerr = off; off = align256(off + 4 * (size_t) nrow);
    const size_t oPcnt = off; off = align256(off + 4 * (size_t) nrow);
    const size_t oScal = off; off = align256(off + 64);
    const size_t oFlags = off; off = align256(off + 64);
    if (w == nullptr) return off;
    w->W = W;
    w->H = H;
    w->n = n;
    w->bytes = off;
    w->base = base;
    char *b = (char *) base;
    w->X = (float2 *) (b + oX);
    w->G = (float2 *) (b + oG);
    w->Bn = (float *) (b + oBn);
    w->B2 = (float *) (b + oB2);
    w->M = (float *) (b + oM);
    w->mask = (uint8_t *) (b + oMask);
    w->pmax = (float *) (b + oPmax);
    w->perr = (float *) (b + oPerr);
    w->pcnt = (unsigned *) (b + oPcnt);
    w->scal = (float *) (b + oScal);
    w->flags = (int *) (b + oFlags);
    return off;
}

int ws_alloc(int device, int W, int H, Workspace **out)
{
    Workspace *w = new Workspace();
    w->device = device;
    void *base = nullptr;
    hipError_t e = hipMalloc(&base, ws_layout(nullptr, W, H, nullptr));
    if (e != hipSuccess) {
        delete w;
        return from_hip(e);
    }
    ws_layout(w, W, H, base);
    e = hipHostMalloc((void **) &w->host_stage, 4096, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void) hipFree(w->base);
        if (w->host_stage) (void) hipHostFree(w->host_stage);
        delete w;
        return from_hip(e);
    }
    *out = w;
    return IMRESH_OK;
}

void ws_release(Workspace *w)
{
    if (w == nullptr) return;
    if (w->base) (void) hipFree(w->base);
    if (w->weights) (void) hipFree(w->weights);
    if (w->host_stage) (void) hipHostFree(w->host_stage);
    if (w->host_wpin) (void) hipHostFree(w->host_wpin);
    if (w->gexec) (void) hipGraphExecDestroy(w->gexec);
    delete w;
}

bool ws_wait_idle(Workspace *w, double timeout_s)
{
    const auto t0 = std::chrono::steady_clock::now();
    while (w->busy.load(std::memory_order_acquire)) {
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > timeout_s)
            return false;
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    return true;
}

size_t ws_bytes(int W, int H)
{
    const size_t n = (size_t) W * H;
    return 29 * n + 4 * (size_t) (blur_grid(W, H) + 2 * H) + 8 * 256;
}

}  // namespace imr

namespace {

std::mutex g_ws_mtx;
std::map<std::tuple<int, int, int>, std::vector<Workspace *>> g_ws_free;

// Workspaces are reused in stream order: a checked-out workspace makes the
// new stream wait for the previous user's work (no host synchronization).
// A size keeps up to kMaxWorkspaces of them so reconstructions of one size
// on different streams run concurrently: an idle one (its last work done) is
// preferred, a new one is allocated while under the cap, and only then does
// the new stream queue behind a busy one.
constexpr int kMaxWorkspaces = 4;
std::map<std::tuple<int, int, int>, int> g_ws_count;

// Idle workspaces stay pooled for the next reconstruction of their size, up
// to g_pool_cap bytes in all (IMRESH_OPT_POOL_MIB); beyond it the least
// recently returned idle ones whose work has finished are freed, so a stream
// of many distinct sizes does not keep one workspace per size forever.
size_t g_idle_bytes = 0;
std::atomic<size_t> g_pool_cap{(size_t) 32 << 30};   // read without g_ws_mtx (batch.cpp)
uint64_t g_use_clock = 0;

// under g_ws_mtx: detach idle workspaces (work finished) until the idle bytes
// fit `cap`, least recently returned first; the caller frees them unlocked
void trim_idle(size_t cap, std::vector<Workspace *> &out)
{
    while (g_idle_bytes > cap) {
        Workspace *lru = nullptr;
        std::vector<Workspace *> *list = nullptr;
        size_t at = 0;
        for (auto &kv : g_ws_free)
            for (size_t i = 0; i < kv.second.size(); ++i) {
                Workspace *w = kv.second[i];
                if (w->busy.load(std::memory_order_acquire)) continue;
                if (lru == nullptr || w->last_use < lru->last_use) {
                    lru = w;
                    list = &kv.second;
                    at = i;
                }
            }
        if (lru == nullptr) return;   // everything idle is still running
        list->erase(list->begin() + (long) at);
        --g_ws_count[{lru->device, lru->W, lru->H}];
        g_idle_bytes -= lru->bytes;
        out.push_back(lru);
    }
}

int ws_checkout(int device, int W, int H, hipStream_t s, Workspace **out)
{
    Workspace *w = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_ws_mtx);
        auto &v = g_ws_free[{device, W, H}];
        for (size_t i = 0; i < v.size() && w == nullptr; ++i)
            if (!v[i]->busy.load(std::memory_order_acquire)) {
                w = v[i];
                v.erase(v.begin() + (long) i);
            }
        if (w == nullptr) {
            int &count = g_ws_count[{device, W, H}];
            if (count >= kMaxWorkspaces && !v.empty()) {
                w = v.front();
                v.erase(v.begin());
            } else {
                ++count;
            }
        }
        if (w != nullptr) g_idle_bytes -= w->bytes;
    }
    if (w == nullptr) {
        const int st = ws_alloc(device, W, H, out);
        if (st != IMRESH_OK) {
            std::lock_guard<std::mutex> lk(g_ws_mtx);
            --g_ws_count[{device, W, H}];
        }
        return st;
    }
    // a busy workspace taken over (kMaxWorkspaces of this size in use): its
    // previous user's work runs first.  Should that work not finish in
    // reasonable time (a caller stream destroyed with its work dropped would
    // never run the host function), the workspace stays in the pool and a
    // new one is allocated instead of waiting on.
    (void) s;
    if (!ws_wait_idle(w, 10.0)) {
        {
            std::lock_guard<std::mutex> lk(g_ws_mtx);
            g_ws_free[{device, W, H}].push_back(w);
            g_idle_bytes += w->bytes;
            ++g_ws_count[{device, W, H}];
        }
        const int st = ws_alloc(device, W, H, out);
        if (st != IMRESH_OK) {
            std::lock_guard<std::mutex> lk(g_ws_mtx);
            --g_ws_count[{device, W, H}];
        }
        return st;
    }
    *out = w;
    return IMRESH_OK;
}

void mark_idle(void *p) { static_cast<Workspace *>(p)->busy.store(false, std::memory_order_release); }

// synced: the caller has synchronized the stream after its last use of
// the workspace (then it is idle at once; otherwise a host function behind
// the queued work marks it idle)
void ws_return(Workspace *w, hipStream_t s, bool synced = false)
{
    if (synced) {
        w->busy.store(false, std::memory_order_release);
    } else {
        w->busy.store(true, std::memory_order_relaxed);
        if (hipLaunchHostFunc(s, mark_idle, w) != hipSuccess) {
            (void) hipStreamSynchronize(s);   // no host function: wait here instead
            w->busy.store(false, std::memory_order_release);
        }
    }
    std::vector<Workspace *> drop;
    {
        std::lock_guard<std::mutex> lk(g_ws_mtx);
        w->last_use = ++g_use_clock;
        g_ws_free[{w->device, w->W, w->H}].push_back(w);
        g_idle_bytes += w->bytes;
        trim_idle(g_pool_cap, drop);
    }
    for (Workspace *x : drop) ws_release(x);
}

}  // namespace

namespace imr {

size_t pool_cap() { return g_pool_cap.load(); }

int normalize(const imresh_params *p, Params &q)
{
    imresh_params d;
    imresh_params_default(&d);
    if (p == nullptr) p = &d;
    const bool queue = p->defaults == IMRESH_DEFAULTS_QUEUE;
    q.n_cycles = p->n_cycles;
    if (queue && q.n_cycles == 0) q.n_cycles = 20;
    q.target = p->target_error <= 0 ? 1e-5f : p->target_error;
    q.beta = p->hio_beta <= 0 ? 0.9f : p->hio_beta;
    q.cutA = p->cutoff_autocorel <= 0 ? (queue ? 0.02f : 0.04f) : p->cutoff_autocorel;
    q.cut = p->cutoff <= 0 ? 0.2f : p->cutoff;
    q.sigma0 = p->sigma0 <= 0 ? 3.0f : p->sigma0;
    q.dsigma = p->sigma_change <= 0 ? 0.01f : p->sigma_change;
    q.n_hio = p->n_hio_cycles == 0 ? 20 : p->n_hio_cycles;
    return IMRESH_OK;
}

}  // namespace imr

namespace {

}  // namespace

namespace imr {

// Zero-padded per-cycle weights: row c (kBlurWeightRow floats) holds cycle
// c's 2r_c+1 taps centred in 2R+1 slots, R = widest radius over the cycles.
// A zero tap adds exactly +0 to the reference-order sum of non-negative
// inputs, so one blur kernel (radius R) serves every cycle bit-identically,
// and the cycle's row is picked on the device from the completed-cycle
// count -- the same work every cycle, which is what lets a cycle be replayed
// as one graph.
int padded_weights(const Params &q, std::vector<float> &rows, int *R)
{
    std::vector<float> sig;
    sigma_schedule(q.sigma0, q.dsigma, q.n_cycles, sig);
    int rmax = 0;
    std::vector<int> taps(q.n_cycles);
    for (unsigned c = 0; c < q.n_cycles; ++c) {
        taps[c] = gaussian_weights(sig[c], nullptr, 0);
        if (taps[c] > 2 * kMaxBlurRadius + 1) return IMRESH_ERR_KERNEL_TOO_LARGE;
        rmax = std::max(rmax, (taps[c] - 1) / 2);
    }
    rows.assign((size_t) std::max(1u, q.n_cycles) * kBlurWeightRow, 0.f);
    for (unsigned c = 0; c < q.n_cycles; ++c) {
        float w[2 * kMaxBlurRadius + 1];
        gaussian_weights(sig[c], w, 2 * kMaxBlurRadius + 1);
        const int r = (taps[c] - 1) / 2;
        for (int k = 0; k < taps[c]; ++k) rows[(size_t) c * kBlurWeightRow + (rmax - r) + k] = w[k];
    }
    *R = rmax;
    return IMRESH_OK;
}

// the workspace's weights table for q, uploaded once per parameter set from
// pinned memory kept with the workspace, so the copy is asynchronous (a copy
// from pageable memory would hold the calling thread until the stream's
// earlier work had finished).  The staging is rewritten only once the
// workspace's previous user has finished, whose copy may still read it.
int ensure_weights(Workspace *w, const Params &q, hipStream_t s)
{
    if (w->weights && w->w_sigma0 == q.sigma0 && w->w_dsigma == q.dsigma && w->w_ncycles == q.n_cycles)
        return IMRESH_OK;
    std::vector<float> &host = w->host_weights;
    int R = 0;
    IMR_CHECK(padded_weights(q, host, &R));
    // a previous user's copy may still read the staging (checked-out
    // workspaces are idle; this only guards)
    if (!ws_wait_idle(w, 10.0)) return IMRESH_ERR_INTERNAL;
    if (w->host_wpin_n < host.size()) {
        if (w->host_wpin) (void) hipHostFree(w->host_wpin);
        w->host_wpin = nullptr;
        w->host_wpin_n = 0;
        IMR_HIP_CHECK(hipHostMalloc((void **) &w->host_wpin, sizeof(float) * host.size(), hipHostMallocDefault));
        w->host_wpin_n = host.size();
    }
    std::memcpy(w->host_wpin, host.data(), sizeof(float) * host.size());
    const unsigned rows = std::max(1u, q.n_cycles);
    if (w->wcap < rows) {
        if (w->weights) {
            IMR_HIP_CHECK(hipStreamSynchronize(s));
            (void) hipFree(w->weights);
            w->weights = nullptr;
        }
        w->wcap = rows;
        IMR_HIP_CHECK(hipMalloc(&w->weights, sizeof(float) * kBlurWeightRow * w->wcap));
    }
    IMR_HIP_CHECK(hipMemcpyAsync(w->weights, w->host_wpin, sizeof(float) * host.size(),
                                 hipMemcpyHostToDevice, s));
    w->wR = R;
    w->w_sigma0 = q.sigma0;
    w->w_dsigma = q.dsigma;
    w->w_ncycles = q.n_cycles;
    return IMRESH_OK;
}

}  // namespace imr

namespace {

// ---------------------------------------------------------------- autotuner
// Picks, per device and image size, the fastest row-pass and column-pass
// kernel variant (element budget / rows or columns per workgroup / LDS
// exchange) by timing each on a synthetic resident state with HIP events.
// Runs once per (device, W, H) and process; IMRESH_AUTOTUNE=0 keeps the
// defaults, IMRESH_VARIANTS=r,c forces a choice.
int time_variant(bool row, int var, Plan2D p, Workspace *w, hipStream_t s, float *ms)
{
    RowArgs ra{};
    ra.X = w->X; ra.G = w->G; ra.mask = w->mask; ra.Bn = w->Bn; ra.B2 = w->B2;
    ra.thr = w->scal; ra.beta = 0.9f; ra.perr = w->perr; ra.pcnt = w->pcnt;
    ColArgs ca{};
    ca.X = w->X; ca.M = w->M;
    ca.in = w->X; ca.out = w->X;
    // a length whose projection pass is the four-step one uses the column
    // variant only for its plain transforms: time those
    const int col_mode = p.col.fs_h2 > 0 ? COL_FWD : COL_PROJ;
    if (row) p.rowVar = var; else p.colVar = var;
    hipEvent_t e0, e1;
    IMR_HIP_CHECK(hipEventCreate(&e0));
    IMR_HIP_CHECK(hipEventCreate(&e1));
    int st = IMRESH_OK;
    // one trial of 8 back-to-back launches on a refilled state (fill values
    // are below 1; 8 unnormalized row launches stay finite); the caller
    // repeats the trials in interleaved rounds
    constexpr int reps = 8, trials = 1;
    float best = 1e30f;
    for (int k = 0; k < trials && st == IMRESH_OK; ++k) {
        st = launch_fill_state(w->X, w->G, w->mask, w->M, w->Bn, w->n, w->W, s);
        for (int r = 0; r <= reps && st == IMRESH_OK; ++r) {
            if (r == 1) st = IMR_HIP_STATUS(hipEventRecord(e0, s));   // first launch warms up
            if (st == IMRESH_OK) st = row ? launch_row(ROW_MID, p, ra, s) : launch_col(col_mode, p, ca, s);
        }
        if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventRecord(e1, s));
        if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventSynchronize(e1));
        float t = 0.f;
        if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventElapsedTime(&t, e0, e1));
        best = std::min(best, t / reps);
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    *ms = best;
    if (st == IMRESH_OK && std::getenv("IMRESH_AUTOTUNE_VERBOSE") != nullptr)
        std::fprintf(stderr, "[imresh autotune] %dx%d %s variant %d: %.2f us\n", p.W, p.H,
                     row ? "row" : "column", var, best * 1e3f);
    return st;
}

int autotune(int device, int W, int H, hipStream_t s, Workspace *w)
{
    int rv = 0, cv = 0;
    if (get_tuned_variants(device, W, H, &rv, &cv)) return IMRESH_OK;
    Plan2D p;
    IMR_CHECK(get_plan2d(device, W, H, &p));
    const int nr = row_variants(p.row), nc = col_variants(p.col);
    const char *force = std::getenv("IMRESH_VARIANTS");
    const char *off = std::getenv("IMRESH_AUTOTUNE");
    if (force != nullptr && std::sscanf(force, "%d,%d", &rv, &cv) == 2) {
        rv = std::min(std::max(rv, 0), std::max(nr - 1, 0));
        cv = std::min(std::max(cv, 0), std::max(nc - 1, 0));
    } else if (!(off != nullptr && off[0] == '0') && (nr > 1 || nc > 1)) {
        // every variant once per round, best of kTuneRounds rounds: timing
        // the variants one after the other would bias against the first
        // ones, which can run while the GPU is still raising its clocks
        constexpr int kTuneRounds = 3;
        std::vector<float> br(nr > 1 ? nr : 0, 1e30f), bc(nc > 1 ? nc : 0, 1e30f);
        for (int round = 0; round < kTuneRounds; ++round) {
            for (int v = 0; v < (int) br.size(); ++v) {
                float ms;
                IMR_CHECK(time_variant(true, v, p, w, s, &ms));
                br[v] = std::min(br[v], ms);
            }
            for (int v = 0; v < (int) bc.size(); ++v) {
                float ms;
                IMR_CHECK(time_variant(false, v, p, w, s, &ms));
                bc[v] = std::min(bc[v], ms);
            }
        }
        if (!br.empty()) rv = (int) (std::min_element(br.begin(), br.end()) - br.begin());
        if (!bc.empty()) cv = (int) (std::min_element(bc.begin(), bc.end()) - bc.begin());
    }
    set_tuned_variants(device, W, H, rv, cv);
    return IMRESH_OK;
}

}  // namespace

namespace imr {

namespace {
ColArgs col_args(const Workspace *w, const float *M)
{
    ColArgs ca{};
    ca.X = w->X;
    ca.M = M;
    ca.done = w->flags;
    return ca;
}

RowArgs row_args(const Workspace *w, const Params &q)
{
    RowArgs ra{};
    ra.X = w->X;
    ra.G = w->G;
    ra.mask = w->mask;
    ra.Bn = w->Bn;
    ra.B2 = w->B2;
    ra.thr = w->scal;
    ra.beta = q.beta;
    ra.perr = w->perr;
    ra.pcnt = w->pcnt;
    ra.done = w->flags;
    return ra;
}
}  // namespace

int ws_tune(Workspace *w, hipStream_t s) { return autotune(w->device, w->W, w->H, s, w); }

int recon_begin(Workspace *w, const float *M, const Params &q, hipStream_t s)
{
    IMR_CHECK(ensure_weights(w, q, s));
    if (M != w->M)
        IMR_HIP_CHECK(hipMemcpyAsync(w->M, M, sizeof(float) * w->n, hipMemcpyDeviceToDevice, s));
    IMR_HIP_CHECK(hipMemsetAsync(w->flags, 0, 64, s));
    IMR_HIP_CHECK(hipMemsetAsync(w->scal, 0, 64, s));
    Plan2D p;
    IMR_CHECK(get_plan2d(w->device, w->W, w->H, &p));
    IMR_CHECK(launch_col(COL_INIT, p, col_args(w, w->M), s));
    return launch_row(ROW_INIT, p, row_args(w, q), s);
}

// one shrink-wrap cycle on the workspace (the batch kernels with one image):
// blur of |g'| with the cycle's weights row, threshold, n_hio fused HIO
// iterations, closing inverse rows, error and early-exit flag
int enqueue_cycle(Workspace *w, const Params &q, const Plan2D &p, hipStream_t s)
{
    BlurArgs ba{};
    ba.in = w->Bn;
    ba.out = w->B2;
    ba.wh = ba.wv = w->weights;
    ba.rh = ba.rv = w->wR;
    ba.W = w->W;
    ba.H = w->H;
    ba.pmax = w->pmax;
    ba.done = w->flags;
    ba.cycles = (const unsigned *) (w->flags + 1);
    IMR_CHECK(launch_blur(ba, s));
    IMR_CHECK(launch_max_finalize_batch(w->pmax, blur_grid(w->W, w->H), q.cutA, q.cut, w->scal,
                                        w->flags, 1, s));
    const ColArgs ca = col_args(w, w->M);
    const RowArgs ra = row_args(w, q);
    for (unsigned it = 0; it < q.n_hio; ++it) {
        IMR_CHECK(launch_row(it == 0 ? ROW_FIRST_CUT : ROW_MID, p, ra, s));
        IMR_CHECK(launch_col(COL_PROJ, p, ca, s));
    }
    IMR_CHECK(launch_row(ROW_LAST, p, ra, s));
    return launch_err_finalize_batch(w->perr, w->pcnt, row_grid(p), q.target, w->scal, w->flags, 1, s);
}

namespace {

std::mutex g_cap_mtx;
std::map<int, hipStream_t> g_cap_streams;

// capture streams (a capture records, it runs nothing; the legacy default
// stream cannot capture)
int capture_stream(int device, hipStream_t *out)
{
    std::lock_guard<std::mutex> lk(g_cap_mtx);
    auto it = g_cap_streams.find(device);
    if (it != g_cap_streams.end()) {
        *out = it->second;
        return IMRESH_OK;
    }
    hipStream_t cs;
    IMR_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    g_cap_streams[device] = cs;
    *out = cs;
    return IMRESH_OK;
}

bool same_key(const Workspace *w, const Params &q, const Plan2D &p)
{
    return w->gexec != nullptr && w->g_nhio == q.n_hio && w->g_beta == q.beta && w->g_cutA == q.cutA &&
           w->g_cut == q.cut && w->g_target == q.target && w->g_R == w->wR && w->g_weights == w->weights &&
           w->g_rowVar == p.rowVar && w->g_colVar == p.colVar;
}

// the cycle as a graph, captured once per workspace and parameter set
int cycle_graph(Workspace *w, const Params &q, const Plan2D &p)
{
    if (same_key(w, q, p)) return IMRESH_OK;
    if (w->gexec) {
        // replays of this graph by the workspace's previous user must have
        // run before it is destroyed (ws_checkout hands out idle workspaces)
        if (!ws_wait_idle(w, 10.0)) return IMRESH_ERR_INTERNAL;
        (void) hipGraphExecDestroy(w->gexec);
        w->gexec = nullptr;
    }
    IMR_CHECK(prepare_blur(w->wR));
    hipStream_t cs;
    IMR_CHECK(capture_stream(w->device, &cs));
    hipGraph_t g = nullptr;
    std::lock_guard<std::mutex> lk(g_cap_mtx);   // one capture per stream at a time
    IMR_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    const int st = enqueue_cycle(w, q, p, cs);
    const hipError_t e = hipStreamEndCapture(cs, &g);
    if (st != IMRESH_OK) {
        if (g) (void) hipGraphDestroy(g);
        return st;
    }
    IMR_HIP_CHECK(e);
    const hipError_t ei = hipGraphInstantiate(&w->gexec, g, nullptr, nullptr, 0);
    (void) hipGraphDestroy(g);
    IMR_HIP_CHECK(ei);
    w->g_nhio = q.n_hio;
    w->g_beta = q.beta;
    w->g_cutA = q.cutA;
    w->g_cut = q.cut;
    w->g_target = q.target;
    w->g_R = w->wR;
    w->g_weights = w->weights;
    w->g_rowVar = p.rowVar;
    w->g_colVar = p.colVar;
    return IMRESH_OK;
}

}  // namespace

std::atomic<bool> g_use_graphs{true};

int recon_cycle(Workspace *w, const float *M, const Params &q, unsigned c, hipStream_t s)
{
    (void) M;
    (void) c;   // the cycle's weights and cut-off are chosen on the device
    Plan2D p;
    IMR_CHECK(get_plan2d(w->device, w->W, w->H, &p));
    if (g_use_graphs) {
        IMR_CHECK(cycle_graph(w, q, p));
        return IMR_HIP_STATUS(hipGraphLaunch(w->gexec, s));
    }
    return enqueue_cycle(w, q, p, s);
}

int recon_end(Workspace *w, float *io_dev, hipStream_t s)
{
    return launch_copy_from_real(io_dev, w->X, w->n, s);
}

// the scalars and the flags lie kStageFlags floats apart in the workspace
// allocation (each region 256-byte aligned), so one copy stages both
constexpr size_t kStageFlags = 64;

int recon_stage_results(Workspace *w, hipStream_t s)
{
    static_assert(sizeof(float) * (kStageFlags + 2) <= 4096, "staging buffer");
    if ((const void *) w->flags != (const void *) (w->scal + kStageFlags)) return IMRESH_ERR_INTERNAL;
    IMR_HIP_CHECK(hipMemcpyAsync(w->host_stage, w->scal, sizeof(float) * (kStageFlags + 2),
                                 hipMemcpyDeviceToHost, s));
    return IMRESH_OK;
}

void recon_read_staged(const Workspace *w, imresh_result *res, int *done)
{
    const float *st = w->host_stage;
    int flags[2];
    std::memcpy(flags, st + kStageFlags, 8);
    if (res != nullptr) {
        res->cycles_run = (uint32_t) flags[1];
        res->last_error = st[1];
        res->last_total = st[2];
        res->last_masked = st[3];
    }
    if (done != nullptr) *done = flags[0];
}

}  // namespace imr

namespace {

int run_reconstruction(Workspace *w, float *io_dev, const Params &q, hipStream_t s)
{
    IMR_CHECK(recon_begin(w, io_dev, q, s));
    for (unsigned c = 0; c < q.n_cycles; ++c) IMR_CHECK(recon_cycle(w, io_dev, q, c, s));
    return recon_end(w, io_dev, s);
}

int read_result(Workspace *w, hipStream_t s, imresh_result *res)
{
    IMR_CHECK(recon_stage_results(w, s));
    IMR_HIP_CHECK(hipStreamSynchronize(s));
    recon_read_staged(w, res, nullptr);
    return IMRESH_OK;
}


hipStream_t S(void *s) { return (hipStream_t) s; }

// Device scratch of the asynchronous per-operator entry points (the blur's
// temporary plane, the Poisson simulation's sums), pooled per device: a
// buffer is held by its call, then stays busy until a host function behind
// the call's work on the caller's stream frees it.  (Stream-ordered
// allocation would leave the allocator tracking the caller's stream, which
// the caller may destroy.)
struct PoolBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int dev = 0;
    std::atomic<int> state{0};   // 0 free, 1 held by a call, 2 busy on a stream
};
std::mutex g_pb_mtx;
std::vector<PoolBuf *> g_pb;

void pool_buf_free(void *p) { static_cast<PoolBuf *>(p)->state.store(0, std::memory_order_release); }

struct Scratch {
    void *p = nullptr;
    PoolBuf *b = nullptr;
    hipStream_t s;
    explicit Scratch(hipStream_t st) : s(st) {}
    int alloc(size_t bytes)
    {
        int dev;
        IMR_CHECK(current_device(&dev));
        {
            std::lock_guard<std::mutex> lk(g_pb_mtx);
            for (PoolBuf *x : g_pb) {
                int fr = 0;
                if (x->dev == dev && x->bytes >= bytes && x->state.compare_exchange_strong(fr, 1)) {
                    b = x;
                    break;
                }
            }
        }
        if (b == nullptr) {
            size_t cap = 4096;
            while (cap < bytes) cap <<= 1;
            PoolBuf *x = new PoolBuf();
            x->bytes = cap;
            x->dev = dev;
            x->state.store(1);
            const hipError_t e = hipMalloc(&x->p, cap);
            if (e != hipSuccess) {
                delete x;
                return IMR_HIP_STATUS(e);
            }
            std::lock_guard<std::mutex> lk(g_pb_mtx);
            g_pb.push_back(x);
            b = x;
        }
        p = b->p;
        return IMRESH_OK;
    }
    ~Scratch()
    {
        if (b == nullptr) return;
        b->state.store(2, std::memory_order_relaxed);
        if (hipLaunchHostFunc(s, pool_buf_free, b) != hipSuccess) {
            (void) hipStreamSynchronize(s);   // no host function: wait here instead
            b->state.store(0, std::memory_order_release);
        }
    }
};

// free pooled scratch buffers (imresh_release_cached) / their bytes
void pool_bufs_release()
{
    std::vector<PoolBuf *> drop;
    {
        std::lock_guard<std::mutex> lk(g_pb_mtx);
        std::vector<PoolBuf *> keep;
        for (PoolBuf *x : g_pb) {
            int fr = 0;
            if (x->state.compare_exchange_strong(fr, 1)) drop.push_back(x);
            else keep.push_back(x);
        }
        g_pb.swap(keep);
    }
    for (PoolBuf *x : drop) {
        (void) hipFree(x->p);
        delete x;
    }
}

size_t pool_bufs_bytes()
{
    std::lock_guard<std::mutex> lk(g_pb_mtx);
    size_t t = 0;
    for (PoolBuf *x : g_pb)
        if (x->state.load() == 0) t += x->bytes;
    return t;
}

int check_n(const void *p, uint32_t n)
{
    if (p == nullptr) return IMRESH_ERR_NULL;
    (void) n;
    return IMRESH_OK;
}

// Scratch of the synchronous per-operator reductions, reused across calls:
// per device a free list of slots, each a device block (the partial sums)
// and a pinned host block (the result), so a call allocates nothing after
// its device's first calls and reads its result back without staging
// through pageable memory.  Concurrent calls take different slots.
struct ReduceSlot {
    void *dev = nullptr;
    float *host = nullptr;
};
constexpr size_t kReduceSlotBytes = sizeof(float) * (2 * kReduceBlocks + 64);
std::mutex g_rs_mtx;
std::map<int, std::vector<ReduceSlot>> g_rs_free;

int reduce_slot_get(int dev, ReduceSlot *out)
{
    {
        std::lock_guard<std::mutex> lk(g_rs_mtx);
        auto &v = g_rs_free[dev];
        if (!v.empty()) {
            *out = v.back();
            v.pop_back();
            return IMRESH_OK;
        }
    }
    ReduceSlot r;
    IMR_HIP_CHECK(hipMalloc(&r.dev, kReduceSlotBytes));
    const hipError_t e = hipHostMalloc((void **) &r.host, 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void) hipFree(r.dev);
        return from_hip(e);
    }
    *out = r;
    return IMRESH_OK;
}

void reduce_slot_put(int dev, const ReduceSlot &r)
{
    std::lock_guard<std::mutex> lk(g_rs_mtx);
    g_rs_free[dev].push_back(r);
}

// idle slots (their calls have synchronized, so no kernel writes them) are
// freed by imresh_release_cached; slots held by running calls stay
void reduce_slots_release()
{
    std::map<int, std::vector<ReduceSlot>> drop;
    {
        std::lock_guard<std::mutex> lk(g_rs_mtx);
        drop.swap(g_rs_free);
    }
    for (auto &kv : drop)
        for (const ReduceSlot &r : kv.second) {
            (void) hipFree(r.dev);
            (void) hipHostFree(r.host);
        }
}

size_t reduce_slots_bytes()
{
    std::lock_guard<std::mutex> lk(g_rs_mtx);
    size_t k = 0;
    for (auto &kv : g_rs_free) k += kv.second.size();
    return k * kReduceSlotBytes;
}

// a slot held for one call and returned on every exit path; a call that
// did not reach its own synchronization (an error after a launch) first
// drains its stream, so no queued kernel still writes the slot
struct SlotHold {
    int dev = 0;
    ReduceSlot r;
    bool held = false, done = false;
    hipStream_t s = nullptr;
    int take(hipStream_t st)
    {
        s = st;
        IMR_CHECK(current_device(&dev));
        IMR_CHECK(reduce_slot_get(dev, &r));
        held = true;
        return IMRESH_OK;
    }
    ~SlotHold()
    {
        if (!held) return;
        if (!done) (void) hipStreamSynchronize(s);   // release the slot only after its last use
        reduce_slot_put(dev, r);
    }
};

int vector_reduce(int op, const float *d, uint32_t n, float *out, void *stream)
{
    IMR_CHECK(check_n(d, n));
    if (out == nullptr) return IMRESH_ERR_NULL;
    hipStream_t s = S(stream);
    SlotHold sh;
    IMR_CHECK(sh.take(s));
    float *part = (float *) sh.r.dev, *res = part + kReduceBlocks;
    IMR_CHECK(launch_reduce(op, d, n, part, kReduceBlocks, res, s));
    IMR_HIP_CHECK(hipMemcpyAsync(sh.r.host, res, sizeof(float), hipMemcpyDeviceToHost, s));
    IMR_HIP_CHECK(hipStreamSynchronize(s));
    sh.done = true;
    *out = sh.r.host[0];
    return IMRESH_OK;
}

int hio_error(const float *g, const void *mask, int kind, uint32_t n, int invert, float *err,
              float *tot, float *nm, void *stream)
{
    if (g == nullptr || mask == nullptr || err == nullptr) return IMRESH_ERR_NULL;
    hipStream_t s = S(stream);
    SlotHold sh;
    IMR_CHECK(sh.take(s));
    float *perr = (float *) sh.r.dev;
    unsigned *pcnt = (unsigned *) (perr + kReduceBlocks);
    float *res = perr + 2 * kReduceBlocks;
    IMR_CHECK(launch_hio_error((const float2 *) g, mask, kind, n, invert, perr, pcnt, kReduceBlocks,
                               res, s));
    IMR_HIP_CHECK(hipMemcpyAsync(sh.r.host, res, 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    IMR_HIP_CHECK(hipStreamSynchronize(s));
    sh.done = true;
    const float *h = sh.r.host;
    *err = h[0];
    if (tot) *tot = h[1];
    if (nm) *nm = h[2];
    return IMRESH_OK;
}

// Per-device table of Gaussian weight vectors for the per-operator blur,
// keyed by sigma, as the reference keeps up to 128 per device keyed by sigma
// (GaussianKernelGpuBuffer, cudaGaussian.cu:64-298).  A slot holds the taps
// at [0, k), 1.0 at [64] (the weight of an identity pass) and zeros
// elsewhere; it is written once, on first use of its sigma (upload and
// stream synchronization, as the reference's first use), and never changed,
// so kernels on any stream read it while other sigmas are added.  With every
// slot taken the caller stages its weights per call instead.
constexpr int kWeightSlots = 128, kWeightSlotFloats = 128;
struct WeightTable {
    float *dev = nullptr;
    int used = 0;
    std::map<uint64_t, std::pair<int, bool>> slot;   // sigma bits -> {slot, uploaded}
};
std::mutex g_wt_mtx;
std::map<int, WeightTable> g_wt;

// *out = sigma's weights in the table, or nullptr when the caller must stage
// them itself (table full, or another thread's upload of sigma still
// running).  The upload and its stream synchronization run outside the lock.
int cached_weights(int dev, double sigma, const float *host, hipStream_t s, const float **out)
{
    uint64_t key;
    std::memcpy(&key, &sigma, sizeof key);
    *out = nullptr;
    float *p;
    {
        std::lock_guard<std::mutex> lk(g_wt_mtx);
        WeightTable &t = g_wt[dev];
        auto it = t.slot.find(key);
        if (it != t.slot.end()) {
            if (it->second.second) *out = t.dev + (size_t) it->second.first * kWeightSlotFloats;
            return IMRESH_OK;
        }
        if (t.used >= kWeightSlots) return IMRESH_OK;
        if (t.dev == nullptr)
            IMR_HIP_CHECK(hipMalloc(&t.dev, sizeof(float) * kWeightSlots * kWeightSlotFloats));
        const int idx = t.used++;
        t.slot.emplace(key, std::make_pair(idx, false));
        p = t.dev + (size_t) idx * kWeightSlotFloats;
    }
    hipError_t e = hipMemcpyAsync(p, host, sizeof(float) * kWeightSlotFloats, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lk(g_wt_mtx);
    WeightTable &t = g_wt[dev];
    if (e != hipSuccess) {
        t.slot.erase(key);   // the slot stays unused; a later call retries in another
        return from_hip(e);
    }
    t.slot[key].second = true;
    *out = p;
    return IMRESH_OK;
}

// The per-operator blur: asynchronous once sigma's weights are in the
// device table (synchronous on first use of a sigma, and on every call
// when the table is full, for the host weights must outlive their copy).
int blur_impl(float *d, uint32_t W, uint32_t H, double sigma, int which, void *stream)
{
    if (d == nullptr) return IMRESH_ERR_NULL;
    if (W == 0 || H == 0) return IMRESH_ERR_INVALID;
    const int k = gaussian_weights(sigma, nullptr, 0);
    if (k > 63) return IMRESH_ERR_KERNEL_TOO_LARGE;
    float w[kWeightSlotFloats] = {};
    gaussian_weights(sigma, w, 64);
    w[64] = 1.f;
    hipStream_t s = S(stream);
    int dev;
    IMR_CHECK(current_device(&dev));
    const float *dw = nullptr;
    IMR_CHECK(cached_weights(dev, sigma, w, s, &dw));
    const size_t n = (size_t) W * H;
    const bool h_inplace = which == 1 && (size_t) W * 4 + 256 <= 120 * 1024;
    const size_t stage = dw == nullptr ? kWeightSlotFloats : 0;
    Scratch sc(s);
    const size_t need = stage + (h_inplace ? 0 : n);
    if (need) IMR_CHECK(sc.alloc(sizeof(float) * need));
    float *tmp = (float *) sc.p + stage;
    if (dw == nullptr) {
        IMR_HIP_CHECK(hipMemcpyAsync(sc.p, w, sizeof(w), hipMemcpyHostToDevice, s));
        dw = (const float *) sc.p;
    }
    const int r = (k - 1) / 2;
    if (h_inplace) {
        IMR_CHECK(launch_blur_h_inplace(d, W, H, dw, r, s));
    } else {
        BlurArgs ba{};
        ba.in = d;
        ba.out = tmp;
        ba.wh = (which & 1) ? dw : dw + 64;
        ba.wv = (which & 2) ? dw : dw + 64;
        ba.rh = (which & 1) ? r : 0;
        ba.rv = (which & 2) ? r : 0;
        ba.h_identity = !(which & 1);
        ba.v_identity = !(which & 2);
        ba.W = W;
        ba.H = H;
        IMR_CHECK(launch_blur(ba, s));
        IMR_HIP_CHECK(hipMemcpyAsync(d, tmp, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    }
    // staged host weights must outlive their copy
    if (stage) IMR_HIP_CHECK(hipStreamSynchronize(s));
    return IMRESH_OK;
}

}  // namespace

namespace imr {

int current_device(int *dev)
{
    hipError_t e = hipGetDevice(dev);
    return from_hip(e);
}

int tune_size(int device, int W, int H, hipStream_t s)
{
    Workspace *w;
    IMR_CHECK(ws_checkout(device, W, H, s, &w));
    const int st = autotune(device, W, H, s, w);
    ws_return(w, s, st == IMRESH_OK);   // autotune synchronizes when it succeeds
    return st;
}

}  // namespace imr

// ====================================================================== ABI
extern "C" {

int imresh_abi_version(void) { return IMRESH_ABI_VERSION; }

const char *imresh_strerror(int status)
{
    switch (status) {
    case IMRESH_OK: return "ok";
    case IMRESH_ERR_NULL: return "null pointer";
    case IMRESH_ERR_INVALID: return "invalid argument";
    case IMRESH_ERR_UNSUPPORTED_SIZE: return "unsupported transform size";
    case IMRESH_ERR_OUT_OF_MEMORY: return "out of device memory";
    case IMRESH_ERR_NO_DEVICE: return "no HIP device";
    case IMRESH_ERR_HIP: return "HIP runtime error";
    case IMRESH_ERR_KERNEL_TOO_LARGE: return "Gaussian kernel wider than 63 taps";
    case IMRESH_ERR_NOT_INITIALIZED: return "task queue not initialized";
    case IMRESH_ERR_BUDGET: return "request exceeds the HBM budget";
    case IMRESH_ERR_IO: return "file unreadable, unwritable or of an unsupported format";
    case IMRESH_ERR_INTERNAL: return "internal host-side failure";
    case IMRESH_ERR_TOO_LARGE: return "exceeds the fixed scratch of the allocation-free entry point";
    default: return "unknown status";
    }
}

int imresh_params_default(imresh_params *p)
{
    if (p == nullptr) return IMRESH_ERR_NULL;
    p->n_cycles = 20;
    p->target_error = 1e-5f;
    p->hio_beta = 0.9f;
    p->cutoff_autocorel = 0.04f;
    p->cutoff = 0.20f;
    p->sigma0 = 3.0f;
    p->sigma_change = 0.01f;
    p->n_hio_cycles = 20;
    p->defaults = IMRESH_DEFAULTS_CPU;
    return IMRESH_OK;
}

int imresh_device_count(int *count)
{
    if (count == nullptr) return IMRESH_ERR_NULL;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice) c = 0;
    else if (e != hipSuccess) return from_hip(e);
    *count = c;
    return IMRESH_OK;
}

int imresh_dev_alloc(void **dev, uint64_t bytes)
{
    if (dev == nullptr) return IMRESH_ERR_NULL;
    return IMR_HIP_STATUS(hipMalloc(dev, bytes ? bytes : 1));
}

int imresh_dev_free(void *dev) { return IMR_HIP_STATUS(hipFree(dev)); }

int imresh_memcpy_h2d(void *dev, const void *host, uint64_t bytes, void *stream)
{
    if (!dev || !host) return IMRESH_ERR_NULL;
    IMR_HIP_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, S(stream)));
    return IMR_HIP_STATUS(hipStreamSynchronize(S(stream)));
}

int imresh_memcpy_d2h(void *host, const void *dev, uint64_t bytes, void *stream)
{
    if (!dev || !host) return IMRESH_ERR_NULL;
    IMR_HIP_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, S(stream)));
    return IMR_HIP_STATUS(hipStreamSynchronize(S(stream)));
}

int imresh_memcpy_d2d(void *dst, const void *src, uint64_t bytes, void *stream)
{
    if (!dst || !src) return IMRESH_ERR_NULL;
    return IMR_HIP_STATUS(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)));
}

int imresh_stream_sync(void *stream) { return IMR_HIP_STATUS(hipStreamSynchronize(S(stream))); }

int imresh_set_option(int option, int value)
{
    switch (option) {
    case IMRESH_OPT_CYCLE_GRAPHS: g_use_graphs = value != 0; return IMRESH_OK;
    case IMRESH_OPT_POOL_MIB: {
        if (value < 0) return IMRESH_ERR_INVALID;
        std::vector<Workspace *> drop;
        {
            std::lock_guard<std::mutex> lk(g_ws_mtx);
            g_pool_cap = (size_t) value << 20;
            trim_idle(g_pool_cap, drop);
        }
        for (Workspace *x : drop) ws_release(x);
        // the per-operator scratch and reduction slots are not kept under a
        // new cap either
        reduce_slots_release();
        pool_bufs_release();
        return guarded([&] { return trim_idle_classes((size_t) value << 20); });
    }
    default: return IMRESH_ERR_INVALID;
    }
}

int imresh_release_cached(void)
{
    std::vector<Workspace *> drop;
    {
        std::lock_guard<std::mutex> lk(g_ws_mtx);
        trim_idle(0, drop);
    }
    for (Workspace *x : drop) ws_release(x);
    reduce_slots_release();
    pool_bufs_release();
    return guarded([] { return trim_idle_classes(0); });
}

int imresh_cached_bytes(uint64_t *bytes)
{
    if (bytes == nullptr) return IMRESH_ERR_NULL;
    size_t ws = 0;
    {
        std::lock_guard<std::mutex> lk(g_ws_mtx);
        ws = g_idle_bytes;
    }
    // the pools' locks are never held together
    *bytes = ws + idle_class_bytes() + reduce_slots_bytes() + pool_bufs_bytes();
    return IMRESH_OK;
}

int imresh_stream_create(void **stream)
{
    if (stream == nullptr) return IMRESH_ERR_NULL;
    hipStream_t s = nullptr;
    IMR_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *) s;
    return IMRESH_OK;
}

int imresh_stream_destroy(void *stream)
{
    if (stream == nullptr) return IMRESH_ERR_NULL;
    return IMR_HIP_STATUS(hipStreamDestroy(S(stream)));
}

static int shrinkwrap_device_impl(float *io_dev, uint32_t width, uint32_t height,
                           const imresh_params *params, void *stream, imresh_result *result)
{
    if (io_dev == nullptr) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    Params q;
    IMR_CHECK(normalize(params, q));
    int dev;
    IMR_CHECK(current_device(&dev));
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, width, height, &p));
    hipStream_t s = S(stream);
    Workspace *w;
    IMR_CHECK(ws_checkout(dev, width, height, s, &w));
    int st = ws_tune(w, s);
    if (st == IMRESH_OK) st = run_reconstruction(w, io_dev, q, s);
    if (st == IMRESH_OK && result != nullptr) st = read_result(w, s, result);
    ws_return(w, s, st == IMRESH_OK && result != nullptr);
    return st;
}

static int shrinkwrap_host_impl(float *io_host, uint32_t width, uint32_t height,
                         const imresh_params *params, void *stream, imresh_result *result)
{
    if (io_host == nullptr) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    Params q;
    IMR_CHECK(normalize(params, q));
    int dev;
    IMR_CHECK(current_device(&dev));
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, width, height, &p));
    hipStream_t s = S(stream);
    Workspace *w;
    IMR_CHECK(ws_checkout(dev, width, height, s, &w));
    const size_t bytes = sizeof(float) * w->n;
    // tuning runs on the workspace's planes, so before the modulus lands in M
    int st = ws_tune(w, s);
    if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipMemcpyAsync(w->M, io_host, bytes, hipMemcpyHostToDevice, s));
    if (st == IMRESH_OK) st = run_reconstruction(w, w->M, q, s);
    if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipMemcpyAsync(io_host, w->M, bytes, hipMemcpyDeviceToHost, s));
    if (st == IMRESH_OK) {
        imresh_result r;
        st = read_result(w, s, &r);
        if (st == IMRESH_OK && result) *result = r;
    }
    ws_return(w, s, st == IMRESH_OK);
    return st;
}

int imresh_shrinkwrap_device(float *io_dev, uint32_t width, uint32_t height,
                             const imresh_params *params, void *stream, imresh_result *result)
{
    return guarded([&] { return shrinkwrap_device_impl(io_dev, width, height, params, stream, result); });
}

int imresh_shrinkwrap(float *io_host, uint32_t width, uint32_t height,
                      const imresh_params *params, void *stream, imresh_result *result)
{
    return guarded([&] { return shrinkwrap_host_impl(io_host, width, height, params, stream, result); });
}

// Test hook: the reconstruction's own launches up to cycle `cycle`, with the
// cycle's input g' copied out before the cycle and its mask, blurred |g'|
// and threshold after it (the cycle's first row pass writes the mask; no
// later pass of the cycle changes the mask or the blurred plane).
static int debug_cycle_mask_impl(const float *modulus_host, uint32_t width, uint32_t height,
                                 const imresh_params *params, uint32_t cycle, float *gprime_host,
                                 uint8_t *mask_host, float *blurred_host, float *threshold_host,
                                 void *stream)
{
    if (modulus_host == nullptr || gprime_host == nullptr || mask_host == nullptr) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    Params q;
    IMR_CHECK(normalize(params, q));
    if (cycle >= q.n_cycles) return IMRESH_ERR_INVALID;
    int dev;
    IMR_CHECK(current_device(&dev));
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, width, height, &p));
    hipStream_t s = S(stream);
    Workspace *w;
    IMR_CHECK(ws_checkout(dev, width, height, s, &w));
    const size_t n = w->n;
    int st = ws_tune(w, s);
    if (st == IMRESH_OK)
        st = IMR_HIP_STATUS(hipMemcpyAsync(w->M, modulus_host, sizeof(float) * n, hipMemcpyHostToDevice, s));
    if (st == IMRESH_OK) st = recon_begin(w, w->M, q, s);
    for (unsigned c = 0; c < cycle && st == IMRESH_OK; ++c) st = recon_cycle(w, w->M, q, c, s);
    if (st == IMRESH_OK)
        st = IMR_HIP_STATUS(hipMemcpyAsync(gprime_host, w->X, sizeof(float2) * n, hipMemcpyDeviceToHost, s));
    if (st == IMRESH_OK) st = recon_cycle(w, w->M, q, cycle, s);
    if (st == IMRESH_OK)
        st = IMR_HIP_STATUS(hipMemcpyAsync(mask_host, w->mask, n, hipMemcpyDeviceToHost, s));
    if (st == IMRESH_OK && blurred_host != nullptr)
        st = IMR_HIP_STATUS(hipMemcpyAsync(blurred_host, w->B2, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    if (st == IMRESH_OK) st = recon_stage_results(w, s);
    if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipStreamSynchronize(s));
    if (st == IMRESH_OK && threshold_host != nullptr) *threshold_host = w->host_stage[0];
    if (st != IMRESH_OK) (void) hipStreamSynchronize(s);   // drain; st is the status
    ws_return(w, s, true);
    return st;
}

int imresh_debug_cycle_mask(const float *modulus_host, uint32_t width, uint32_t height,
                            const imresh_params *params, uint32_t cycle, float *gprime_host,
                            uint8_t *mask_host, float *blurred_host, float *threshold_host,
                            void *stream)
{
    return guarded([&] {
        return debug_cycle_mask_impl(modulus_host, width, height, params, cycle, gprime_host, mask_host,
                                     blurred_host, threshold_host, stream);
    });
}

}  // extern "C"

namespace {

// libs::hybridInputOutput (hybridInputOutput.cpp:200-306) on a workspace whose
// M holds |F| and mask the caller's support (uint8, 1 = outside):
//   random phase (:126-197) -> inverse 2-D FFT, gPrev = g' -> per iteration
//   [constraint -> FFT -> modulus -> inverse FFT], with the error check after
//   every inverse transform when target > 0 (fused row passes otherwise).
// Leaves g' in X; cycles_run = HIO iterations executed.
int run_hio_standalone(Workspace *w, unsigned n_cycles, float target, float beta, hipStream_t s,
                       imresh_result *res)
{
    const size_t n = w->n;
    Plan2D p;
    IMR_CHECK(get_plan2d(w->device, w->W, w->H, &p));
    // the random real field: glibc rand() after srand(2623091912), as :179-184
    {
        std::vector<float> field(n);
        srand(2623091912u);
        // RAND_MAX converts to float (2^31) as in the reference's expression
        for (size_t i = 0; i < n; ++i) field[i] = (float) rand() / (float) RAND_MAX;
        IMR_HIP_CHECK(hipMemcpyAsync(w->B2, field.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
        IMR_HIP_CHECK(hipStreamSynchronize(s));
    }
    IMR_CHECK(launch_copy_to_real(w->X, w->B2, n, s));
    RowArgs rf{};
    rf.in = w->X;
    rf.out = w->X;
    IMR_CHECK(launch_row(ROW_FWD, p, rf, s));
    ColArgs cf{};
    cf.in = w->X;
    cf.out = w->X;
    IMR_CHECK(launch_col(COL_FWD, p, cf, s));
    IMR_CHECK(launch_random_phase(w->X, w->M, n, s));
    IMR_HIP_CHECK(hipMemsetAsync(w->flags, 0, 64, s));
    IMR_HIP_CHECK(hipMemsetAsync(w->scal, 0, 64, s));
    // first inverse transform; gPrevious = g' (:238-239)
    cf.in = w->X;
    IMR_CHECK(launch_col(COL_INV, p, cf, s));
    Params q;
    normalize(nullptr, q);
    q.beta = beta;
    RowArgs ra = row_args(w, q);
    const ColArgs ca = col_args(w, w->M);
    const bool check = target > 0;
    const int nrow = row_grid(p);
    auto finalize = [&] {
        return launch_err_finalize(w->perr, w->pcnt, nrow, target, w->scal + 1, w->flags,
                                   (unsigned *) (w->flags + 1), s);
    };
    IMR_CHECK(launch_row(check ? ROW_INIT_ERR : ROW_INIT, p, ra, s));
    if (check) IMR_CHECK(finalize());
    constexpr unsigned kChunk = 64;   // iterations queued between done-flag reads
    unsigned done_iters = 0;
    for (unsigned c = 0; c < n_cycles;) {
        const unsigned chunk = std::min(kChunk, n_cycles - c);
        for (unsigned k = 0; k < chunk; ++k, ++c) {
            if (check) {
                IMR_CHECK(launch_row(ROW_FIRST_MASK, p, ra, s));
                IMR_CHECK(launch_col(COL_PROJ, p, ca, s));
                IMR_CHECK(launch_row(ROW_LAST, p, ra, s));
                IMR_CHECK(finalize());
            } else {
                IMR_CHECK(launch_row(c == 0 ? ROW_FIRST_MASK : ROW_MID, p, ra, s));
                IMR_CHECK(launch_col(COL_PROJ, p, ca, s));
            }
        }
        done_iters = c;
        if (check && c < n_cycles) {   // stop queueing once the target is met
            IMR_CHECK(recon_stage_results(w, s));
            IMR_HIP_CHECK(hipStreamSynchronize(s));
            int d = 0;
            recon_read_staged(w, nullptr, &d);
            if (d) break;
        }
    }
    if (!check && done_iters > 0) IMR_CHECK(launch_row(ROW_LAST_PLAIN, p, ra, s));
    IMR_CHECK(recon_stage_results(w, s));
    IMR_HIP_CHECK(hipStreamSynchronize(s));
    imresh_result r{};
    recon_read_staged(w, &r, nullptr);
    if (check) {
        r.cycles_run = r.cycles_run > 0 ? r.cycles_run - 1 : 0;
    } else {
        r.cycles_run = done_iters;
        r.last_error = r.last_total = r.last_masked = 0.f;
    }
    if (res) *res = r;
    return IMRESH_OK;
}

int hio_standalone_entry(float *io, const uint8_t *mask, uint32_t W, uint32_t H, uint32_t n_cycles,
                         float target, float beta, void *stream, imresh_result *result, bool device_io)
{
    if (io == nullptr || mask == nullptr) return IMRESH_ERR_NULL;
    if (W == 0 || H == 0) return IMRESH_ERR_INVALID;
    if (n_cycles == 0) n_cycles = UINT_MAX;   // :218
    if (beta <= 0) beta = 0.9f;               // :219
    if (n_cycles == UINT_MAX && !(target > 0)) return IMRESH_ERR_INVALID;   // would never end
    int dev;
    IMR_CHECK(current_device(&dev));
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, W, H, &p));
    hipStream_t s = S(stream);
    Workspace *w;
    IMR_CHECK(ws_checkout(dev, W, H, s, &w));
    const size_t n = w->n;
    const hipMemcpyKind in_kind = device_io ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    int st = ws_tune(w, s);
    if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipMemcpyAsync(w->M, io, sizeof(float) * n, in_kind, s));
    if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipMemcpyAsync(w->mask, mask, n, in_kind, s));
    if (st == IMRESH_OK) st = run_hio_standalone(w, n_cycles, target, beta, s, result);
    if (st == IMRESH_OK) {
        if (device_io) {
            st = launch_copy_from_real(io, w->X, n, s);
        } else {
            st = launch_copy_from_real(w->B2, w->X, n, s);
            if (st == IMRESH_OK)
                st = IMR_HIP_STATUS(hipMemcpyAsync(io, w->B2, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        }
    }
    if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipStreamSynchronize(s));
    ws_return(w, s, st == IMRESH_OK);
    return st;
}

// Forward model |FFT2((x, 0))| (diffractionIntensity.cpp:49-80, CUDA
// :84-135): the column pass reads the real plane and writes the complex
// half-transform into the workspace, the row pass finishes the transform and
// stores only |z|, in place over the input (24 B/element of HBM traffic; the
// reference's path moves 4+8 (H2D), 8 (memset), cuFFT 2 x 16, 8+4 (norm), 4
// (D2H)).  Host io goes through the workspace's M plane.
int diffraction_entry(float *io, uint32_t W, uint32_t H, void *stream, bool device_io, bool async)
{
    if (io == nullptr) return IMRESH_ERR_NULL;
    if (W == 0 || H == 0) return IMRESH_ERR_INVALID;
    int dev;
    IMR_CHECK(current_device(&dev));
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, W, H, &p));
    hipStream_t s = S(stream);
    Workspace *w;
    IMR_CHECK(ws_checkout(dev, W, H, s, &w));
    const size_t n = w->n;
    float *plane = device_io ? io : w->M;
    int st = IMRESH_OK;
    if (!device_io) st = IMR_HIP_STATUS(hipMemcpyAsync(plane, io, sizeof(float) * n, hipMemcpyHostToDevice, s));
    if (st == IMRESH_OK) {
        ColArgs ca{};
        ca.M = plane;
        ca.out = w->X;
        ca.W = (int) W;
        st = launch_col(COL_REAL_FWD, p, ca, s);
    }
    if (st == IMRESH_OK) {
        RowArgs ra{};
        ra.in = w->X;
        ra.Bn = plane;
        ra.H = (int) H;
        st = launch_row(ROW_FWD_NORM, p, ra, s);
    }
    if (st == IMRESH_OK && !device_io)
        st = IMR_HIP_STATUS(hipMemcpyAsync(io, plane, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    if (st == IMRESH_OK && !async) st = IMR_HIP_STATUS(hipStreamSynchronize(s));
    ws_return(w, s, st == IMRESH_OK && !async);
    return st;
}

}  // namespace

extern "C" {

int imresh_hybrid_input_output(float *io, const uint8_t *mask, uint32_t width, uint32_t height,
                               uint32_t n_cycles, float target_error, float hio_beta, void *stream,
                               imresh_result *result)
{
    return guarded([&] {
        return hio_standalone_entry(io, mask, width, height, n_cycles, target_error, hio_beta, stream,
                                    result, false);
    });
}

int imresh_hybrid_input_output_device(float *io_dev, const uint8_t *mask_dev, uint32_t width,
                                      uint32_t height, uint32_t n_cycles, float target_error,
                                      float hio_beta, void *stream, imresh_result *result)
{
    return guarded([&] {
        return hio_standalone_entry(io_dev, mask_dev, width, height, n_cycles, target_error, hio_beta,
                                    stream, result, true);
    });
}

int imresh_hio_iterations(float *gprime_dev, float *gprev_dev, const uint8_t *mask_dev,
                          const float *modulus_dev, uint32_t width, uint32_t height,
                          float hio_beta, uint32_t n_iter, void *stream)
{
    if (!gprime_dev || !gprev_dev || !mask_dev || !modulus_dev) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    if (n_iter == 0) return IMRESH_OK;
    int dev;
    IMR_CHECK(current_device(&dev));
    hipStream_t s = S(stream);
    {
        int rv, cv;
        if (!get_tuned_variants(dev, width, height, &rv, &cv)) {
            Plan2D p0;
            IMR_CHECK(get_plan2d(dev, width, height, &p0));
            Workspace *w;
            IMR_CHECK(ws_checkout(dev, width, height, s, &w));
            const int st = autotune(dev, width, height, s, w);
            ws_return(w, s, st == IMRESH_OK);   // autotune synchronizes when it succeeds
            IMR_CHECK(st);
        }
    }
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, width, height, &p));
    RowArgs ra{};
    ra.X = (float2 *) gprime_dev;
    ra.G = (float2 *) gprev_dev;
    ra.mask = (uint8_t *) mask_dev;
    ra.B2 = modulus_dev;   // unused without the cut-off flag; never null in a kernel
    ra.beta = hio_beta;
    ColArgs ca{};
    ca.X = (float2 *) gprime_dev;
    ca.M = modulus_dev;
    for (uint32_t it = 0; it < n_iter; ++it) {
        IMR_CHECK(launch_row(it == 0 ? ROW_FIRST_MASK : ROW_MID, p, ra, s));
        IMR_CHECK(launch_col(COL_PROJ, p, ca, s));
    }
    IMR_CHECK(launch_row(ROW_LAST_PLAIN, p, ra, s));
    return IMRESH_OK;
}

int imresh_copy_to_real_part(float *cplx, const float *real, uint32_t n, void *stream)
{
    if (!cplx || !real) return IMRESH_ERR_NULL;
    return launch_copy_to_real((float2 *) cplx, real, n, S(stream));
}

int imresh_copy_from_real_part(float *real, const float *cplx, uint32_t n, void *stream)
{
    if (!cplx || !real) return IMRESH_ERR_NULL;
    return launch_copy_from_real(real, (const float2 *) cplx, n, S(stream));
}

int imresh_complex_norm(float *out, const float *cplx, uint32_t n, void *stream)
{
    if (!out || !cplx) return IMRESH_ERR_NULL;
    return launch_complex_norm(out, (const float2 *) cplx, n, S(stream));
}

int imresh_apply_complex_modulus(float *out, const float *in, const float *modulus, uint32_t n,
                                 void *stream)
{
    if (!out || !in || !modulus) return IMRESH_ERR_NULL;
    return launch_apply_modulus((float2 *) out, (const float2 *) in, modulus, n, S(stream));
}

int imresh_hio_domain_constraints(float *gprev, const float *gprime, const float *mask, uint32_t n,
                                  float hio_beta, void *stream)
{
    if (!gprev || !gprime || !mask) return IMRESH_ERR_NULL;
    return launch_hio_constraint((float2 *) gprev, (const float2 *) gprime, mask, n, hio_beta,
                                 S(stream));
}

int imresh_cutoff(float *data, uint32_t n, float threshold, float lower, float upper, void *stream)
{
    if (!data) return IMRESH_ERR_NULL;
    return launch_cutoff(data, n, threshold, lower, upper, S(stream));
}

int imresh_vector_max(const float *d, uint32_t n, float *r, void *s) { return vector_reduce(0, d, n, r, s); }
int imresh_vector_min(const float *d, uint32_t n, float *r, void *s) { return vector_reduce(1, d, n, r, s); }
int imresh_vector_sum(const float *d, uint32_t n, float *r, void *s) { return vector_reduce(2, d, n, r, s); }

int imresh_calculate_hio_error(const float *g, const float *mask, uint32_t n, int invert,
                               float *err, float *tot, float *nm, void *stream)
{
    return hio_error(g, mask, kMaskFloat, n, invert, err, tot, nm, stream);
}

int imresh_calculate_hio_error_u8(const float *g, const uint8_t *mask, uint32_t n, int invert,
                                  float *err, float *tot, float *nm, void *stream)
{
    return hio_error(g, mask, kMaskU8, n, invert, err, tot, nm, stream);
}

int imresh_calculate_hio_error_bits(const float *g, const uint32_t *mask_bits, uint32_t n, int invert,
                                    float *err, float *tot, float *nm, void *stream)
{
    return hio_error(g, mask_bits, kMaskBits, n, invert, err, tot, nm, stream);
}

int imresh_pack_mask(const uint8_t *mask_dev, uint32_t n, uint32_t *bits_dev, void *stream)
{
    if (mask_dev == nullptr || bits_dev == nullptr) return IMRESH_ERR_NULL;
    return launch_pack_mask(mask_dev, n, bits_dev, S(stream));
}

int imresh_gaussian_kernel(double sigma, float *w, uint32_t capacity)
{
    return gaussian_weights(sigma, w, capacity);
}

int imresh_gaussian_blur(float *d, uint32_t W, uint32_t H, double sigma, void *stream)
{
    return blur_impl(d, W, H, sigma, 3, stream);
}

int imresh_gaussian_blur_horizontal(float *d, uint32_t W, uint32_t H, double sigma, void *stream)
{
    return blur_impl(d, W, H, sigma, 1, stream);
}

int imresh_gaussian_blur_vertical(float *d, uint32_t W, uint32_t H, double sigma, void *stream)
{
    return blur_impl(d, W, H, sigma, 2, stream);
}

int imresh_fft2d(float *out, const float *in, uint32_t W, uint32_t H, int direction, void *stream)
{
    if (!out || !in) return IMRESH_ERR_NULL;
    if (W == 0 || H == 0 || (direction != -1 && direction != 1)) return IMRESH_ERR_INVALID;
    int dev;
    IMR_CHECK(current_device(&dev));
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, W, H, &p));
    hipStream_t s = S(stream);
    RowArgs ra{};
    ra.in = (const float2 *) in;
    ra.out = (float2 *) out;
    IMR_CHECK(launch_row(direction < 0 ? ROW_FWD : ROW_INV, p, ra, s));
    ColArgs ca{};
    ca.in = (const float2 *) out;
    ca.out = (float2 *) out;
    return launch_col(direction < 0 ? COL_FWD : COL_INV, p, ca, s);
}

int imresh_fft_supported(uint32_t n) { return line_supported((int) n) ? 1 : 0; }

int imresh_diffraction_intensity(float *io, uint32_t width, uint32_t height, void *stream, int async)
{
    return guarded([&] { return diffraction_entry(io, width, height, stream, false, async != 0); });
}

int imresh_diffraction_intensity_device(float *io_dev, uint32_t width, uint32_t height, void *stream)
{
    return guarded([&] { return diffraction_entry(io_dev, width, height, stream, true, true); });
}

int imresh_poisson_modulus(float *modulus_dev, uint32_t width, uint32_t height, double photons,
                           uint64_t seed, void *stream)
{
    if (modulus_dev == nullptr) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0 || !(photons > 0.0) || !(photons <= 1e30)) return IMRESH_ERR_INVALID;
    hipStream_t s = S(stream);
    Scratch sc(s);   // the sums: kAuxDoubles doubles, stream-ordered
    IMR_CHECK(sc.alloc(sizeof(double) * kAuxDoubles));
    return launch_poisson_modulus(modulus_dev, (size_t) width * height, photons, seed, (double *) sc.p, s);
}

int imresh_fft_shift(void *data_dev, uint32_t nx, uint32_t ny, uint32_t element_bytes, void *stream)
{
    if (data_dev == nullptr) return IMRESH_ERR_NULL;
    return launch_fft_shift(data_dev, nx, ny, element_bytes, S(stream));
}

int imresh_variant_counts(uint32_t width, uint32_t height, int *row_variants_out, int *col_variants_out)
{
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    int dev;
    IMR_CHECK(current_device(&dev));
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, width, height, &p));
    if (row_variants_out) *row_variants_out = std::max(row_variants(p.row), 1);
    if (col_variants_out) *col_variants_out = std::max(col_variants(p.col), 1);
    return IMRESH_OK;
}

int imresh_set_variants(uint32_t width, uint32_t height, int row_variant, int col_variant)
{
    int nr, nc;
    IMR_CHECK(imresh_variant_counts(width, height, &nr, &nc));
    if (row_variant < 0 || row_variant >= nr || col_variant < 0 || col_variant >= nc)
        return IMRESH_ERR_INVALID;
    int dev;
    IMR_CHECK(current_device(&dev));
    set_tuned_variants(dev, (int) width, (int) height, row_variant, col_variant);
    return IMRESH_OK;
}

int imresh_autotune(uint32_t width, uint32_t height, void *stream, int *row_variant,
                    int *col_variant)
{
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    int dev;
    IMR_CHECK(current_device(&dev));
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, width, height, &p));
    hipStream_t s = S(stream);
    Workspace *w;
    IMR_CHECK(ws_checkout(dev, width, height, s, &w));
    int st = autotune(dev, width, height, s, w);
    ws_return(w, s, st == IMRESH_OK);   // autotune synchronizes when it succeeds
    IMR_CHECK(st);
    IMR_CHECK(get_plan2d(dev, width, height, &p));
    if (row_variant) *row_variant = p.rowVar;
    if (col_variant) *col_variant = p.colVar;
    return IMRESH_OK;
}

int imresh_profile_kernels(uint32_t W, uint32_t H, uint32_t reps, float *out_ms, void *stream)
{
    if (out_ms == nullptr) return IMRESH_ERR_NULL;
    if (W == 0 || H == 0 || reps == 0) return IMRESH_ERR_INVALID;
    int dev;
    IMR_CHECK(current_device(&dev));
    hipStream_t s = S(stream);
    Workspace *w;
    IMR_CHECK(ws_checkout(dev, W, H, s, &w));
    Params q;
    normalize(nullptr, q);
    q.n_cycles = 1;
    double acc[4] = {0, 0, 0, 0};
    int cnt[4] = {0, 0, 0, 0};
    int st = autotune(dev, W, H, s, w);
    Plan2D p;
    if (st == IMRESH_OK) st = get_plan2d(dev, W, H, &p);
    if (st == IMRESH_OK) st = ensure_weights(w, q, s);
    std::vector<hipEvent_t> ev;
    const RowArgs ra = row_args(w, q);
    const ColArgs ca = col_args(w, w->M);
    BlurArgs ba{};
    ba.in = w->Bn; ba.out = w->B2; ba.wh = ba.wv = w->weights;
    ba.W = W; ba.H = H; ba.pmax = w->pmax; ba.rh = ba.rv = w->wR;
    // each launch timed on its own execution (kernel-exact events): the
    // fused HIO iteration as it runs in a reconstruction (row pass, column
    // pass alternating on a filled state), then the cycle-boundary kernels
    auto timed = [&](int k, auto launch) {
        hipEvent_t a = nullptr, b = nullptr;
        if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventCreate(&a));
        if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventCreate(&b));
        if (st != IMRESH_OK) return;
        g_prof.start = a;
        g_prof.stop = b;
        st = launch();
        g_prof = ProfEvents{};
        ev.push_back(a);
        ev.push_back(b);
        ev.push_back((hipEvent_t) (intptr_t) k);
    };
    const int iters = (int) std::max(4u, reps);
    for (int round = 0; round < 2 && st == IMRESH_OK; ++round) {   // round 0 warms up
        ev.clear();
        st = launch_fill_state(w->X, w->G, w->mask, w->M, w->Bn, w->n, W, s);
        for (int i = 0; i < iters && st == IMRESH_OK; ++i) {
            timed(0, [&] { return launch_row(ROW_MID, p, ra, s); });
            timed(1, [&] { return launch_col(COL_PROJ, p, ca, s); });
        }
        for (int i = 0; i < iters && st == IMRESH_OK; ++i) {
            timed(2, [&] { return launch_blur(ba, s); });
            timed(3, [&] { return launch_row(ROW_LAST, p, ra, s); });
        }
        if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipStreamSynchronize(s));
        for (size_t e = 0; e + 2 < ev.size(); e += 3) {
            const int k = (int) (intptr_t) ev[e + 2];
            float ms = 0.f;
            if (st == IMRESH_OK && round == 1 && hipEventElapsedTime(&ms, ev[e], ev[e + 1]) == hipSuccess) {
                acc[k] += ms;
                cnt[k] += 1;
            }
            (void) hipEventDestroy(ev[e]);
            (void) hipEventDestroy(ev[e + 1]);
        }
        ev.clear();
    }
    const int ss = IMR_HIP_STATUS(hipStreamSynchronize(s));
    ws_return(w, s, ss == IMRESH_OK);
    if (st == IMRESH_OK) st = ss;
    if (st == IMRESH_OK)
        for (int k = 0; k < 4; ++k) out_ms[k] = cnt[k] ? (float) (acc[k] / cnt[k]) : 0.f;
    return st;
}

int imresh_profile_forward_model(uint32_t W, uint32_t H, uint32_t reps, float *out_ms, void *stream)
{
    if (out_ms == nullptr) return IMRESH_ERR_NULL;
    if (W == 0 || H == 0 || reps == 0) return IMRESH_ERR_INVALID;
    int dev;
    IMR_CHECK(current_device(&dev));
    hipStream_t s = S(stream);
    Plan2D p;
    IMR_CHECK(get_plan2d(dev, W, H, &p));
    Workspace *w;
    IMR_CHECK(ws_checkout(dev, W, H, s, &w));
    int st = launch_fill_state(w->X, w->G, w->mask, w->M, w->Bn, w->n, W, s);
    ColArgs ca{};
    ca.M = w->Bn;   // a filled real plane
    ca.out = w->X;
    RowArgs ra{};
    ra.in = w->X;
    ra.Bn = w->B2;
    double acc[2] = {0, 0};
    int cnt[2] = {0, 0};
    for (int round = 0; round < 2 && st == IMRESH_OK; ++round) {   // round 0 warms up
        for (uint32_t i = 0; i < std::max(4u, reps) && st == IMRESH_OK; ++i) {
            for (int k = 0; k < 2 && st == IMRESH_OK; ++k) {
                hipEvent_t a = nullptr, b = nullptr;
                st = IMR_HIP_STATUS(hipEventCreate(&a));
                if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventCreate(&b));
                if (st == IMRESH_OK) {
                    g_prof.start = a;
                    g_prof.stop = b;
                    st = k == 0 ? launch_col(COL_REAL_FWD, p, ca, s) : launch_row(ROW_FWD_NORM, p, ra, s);
                    g_prof = ProfEvents{};
                }
                if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventSynchronize(b));
                float ms = 0.f;
                if (st == IMRESH_OK && round == 1 && hipEventElapsedTime(&ms, a, b) == hipSuccess) {
                    acc[k] += ms;
                    cnt[k] += 1;
                }
                if (a) (void) hipEventDestroy(a);
                if (b) (void) hipEventDestroy(b);
            }
        }
    }
    const int ss = IMR_HIP_STATUS(hipStreamSynchronize(s));
    ws_return(w, s, ss == IMRESH_OK);
    if (st == IMRESH_OK) st = ss;
    if (st == IMRESH_OK)
        for (int k = 0; k < 2; ++k) out_ms[k] = cnt[k] ? (float) (acc[k] / cnt[k]) : 0.f;
    return st;
}

}  // extern "C"
