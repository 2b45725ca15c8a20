// Batched shrink-wrap: many independent images of mixed sizes advance
// together.  The reference reconstructs one image per call and overlaps
// calls only through its 3-stream task queue (taskQueue.cu:91-193); here
//
//   * images of one size form a class whose planes are stacked in HBM
//     ([B][H][W]); every kernel of the single-image path runs once per
//     class with the image index in the grid (rows/columns/blur tiles x B),
//     per-image threshold, error, cycle count and early-exit flag on the
//     device;
//   * one shrink-wrap cycle of every class is captured once as a HIP graph,
//     each class on its own branch (the classes run concurrently), and the
//     graph is replayed once per cycle.  The cycle-dependent inputs live on
//     the device: the blur reads the weights row of the image's completed
//     cycle count and the threshold kernel picks cutA / cut from it.  The
//     Gaussian weights of every cycle are centred in a row of the widest
//     cycle's taps, padded with zero taps that add exactly +0 to the
//     reference-order sums, so one blur kernel serves all cycles;
//   * kernel variants are the autotuned choice for each size.
//
// A batch of one image computes exactly what imresh_shrinkwrap_device does
// with the same kernel variants.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "engine.hpp"

namespace imr {

namespace {

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// stacked device state of one size class
struct ClassBuf {
    int device = 0, W = 0, H = 0, cap = 0;
    size_t n = 0;
    void *base = nullptr;
    float2 *X = nullptr, *G = nullptr;
    float *M = nullptr, *Bn = nullptr, *B2 = nullptr;
    uint8_t *mask = nullptr;
    float *pmax = nullptr, *perr = nullptr;
    unsigned *pcnt = nullptr;
    float *scal = nullptr;   // [cap][kScalStride]
    int *flags = nullptr;    // [cap][kFlagStride]
    void **tab = nullptr;    // [cap] image pointers (gather / scatter)
    hipStream_t branch = nullptr;
    hipEvent_t join = nullptr, ready = nullptr;
    bool pending = false;
    size_t bytes = 0;        // device bytes of the allocation
    uint64_t last_use = 0;   // pool clock of the last return
};

std::mutex g_mtx;
std::map<std::tuple<int, int, int>, std::vector<ClassBuf *>> g_free;
size_t g_idle = 0;       // bytes of the idle buffers in g_free
uint64_t g_clock = 0;

// under g_mtx: detach idle buffers whose work has finished, least recently
// returned first, until the idle bytes fit `cap`
void trim_classes(size_t cap, std::vector<ClassBuf *> &out)
{
    while (g_idle > cap) {
        ClassBuf *lru = nullptr;
        std::vector<ClassBuf *> *list = nullptr;
        size_t at = 0;
        for (auto &kv : g_free)
            for (size_t i = 0; i < kv.second.size(); ++i) {
                ClassBuf *c = kv.second[i];
                if (c->pending && hipEventQuery(c->ready) != hipSuccess) continue;
                if (lru == nullptr || c->last_use < lru->last_use) {
                    lru = c;
                    list = &kv.second;
                    at = i;
                }
            }
        if (lru == nullptr) return;
        list->erase(list->begin() + (long) at);
        g_idle -= lru->bytes;
        out.push_back(lru);
    }
}

void class_free(ClassBuf *c)
{
    if (c->base) (void) hipFree(c->base);
    if (c->branch) (void) hipStreamDestroy(c->branch);
    if (c->join) (void) hipEventDestroy(c->join);
    if (c->ready) (void) hipEventDestroy(c->ready);
    delete c;
}

int class_alloc(int device, int W, int H, int cap, ClassBuf **out)
{
    ClassBuf *c = new ClassBuf();
    c->device = device;
    c->W = W;
    c->H = H;
    c->cap = cap;
    c->n = (size_t) W * H;
    const size_t N = c->n * cap;
    const size_t ntiles = (size_t) blur_grid(W, H) * cap;
    const size_t nrow = (size_t) H * cap;   // >= row workgroups of every variant
    size_t off = 0;
    const size_t oX = off; off = align256(off + 8 * N);
    const size_t oG = off; off = align256(off + 8 * N);
    const size_t oM = off; off = align256(off + 4 * N);
    const size_t oBn = off; off = align256(off + 4 * N);
    const size_t oB2 = off; off = align256(off + 4 * N);
    const size_t oMask = off; off = align256(off + N);
    const size_t oPmax = off; off = align256(off + 4 * ntiles);
    const size_t oPerr = off; off = align256(off + 4 * nrow);
    const size_t oPcnt = off; off = align256(off + 4 * nrow);
    const size_t oScal = off; off = align256(off + sizeof(float) * kScalStride * cap);
    const size_t oFlags = off; off = align256(off + sizeof(int) * kFlagStride * cap);
    const size_t oTab = off; off = align256(off + sizeof(void *) * cap);
    c->bytes = off;
    hipError_t e = hipMalloc(&c->base, off);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->branch, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ready, hipEventDisableTiming);
    if (e != hipSuccess) {
        class_free(c);
        return from_hip(e);
    }
    char *b = (char *) c->base;
    c->X = (float2 *) (b + oX);
    c->G = (float2 *) (b + oG);
    c->M = (float *) (b + oM);
    c->Bn = (float *) (b + oBn);
    c->B2 = (float *) (b + oB2);
    c->mask = (uint8_t *) (b + oMask);
    c->pmax = (float *) (b + oPmax);
    c->perr = (float *) (b + oPerr);
    c->pcnt = (unsigned *) (b + oPcnt);
    c->scal = (float *) (b + oScal);
    c->flags = (int *) (b + oFlags);
    c->tab = (void **) (b + oTab);
    *out = c;
    return IMRESH_OK;
}

// stream-ordered reuse, as the single-image workspaces
int class_checkout(int device, int W, int H, int count, hipStream_t s, ClassBuf **out)
{
    ClassBuf *c = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mtx);
        auto &v = g_free[{device, W, H}];
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i]->cap >= count) {
                c = v[i];
                v.erase(v.begin() + (long) i);
                g_idle -= c->bytes;
                break;
            }
    }
    if (c == nullptr) return class_alloc(device, W, H, count, out);
    if (c->pending) IMR_HIP_CHECK(hipStreamWaitEvent(s, c->ready, 0));
    *out = c;
    return IMRESH_OK;
}

void class_return(ClassBuf *c, hipStream_t s)
{
    c->pending = hipEventRecord(c->ready, s) == hipSuccess;
    if (!c->pending) (void) hipStreamSynchronize(s);   // no event: wait here instead
    std::vector<ClassBuf *> drop;
    {
        std::lock_guard<std::mutex> lk(g_mtx);
        c->last_use = ++g_clock;
        g_free[{c->device, c->W, c->H}].push_back(c);
        g_idle += c->bytes;
        trim_classes(pool_cap(), drop);
    }
    for (ClassBuf *x : drop) class_free(x);
}

// per-device internal stream (graph capture needs a non-legacy stream) and
// the lock that gives one batch call at a time the use of it: a batch fills
// the device, and two captures on one stream would interfere
std::map<int, hipStream_t> g_streams;
std::map<int, std::mutex> g_run_mtx;

std::mutex &run_lock(int device)
{
    std::lock_guard<std::mutex> lk(g_mtx);
    return g_run_mtx[device];
}

int work_stream(int device, hipStream_t *out)
{
    std::lock_guard<std::mutex> lk(g_mtx);
    auto it = g_streams.find(device);
    if (it != g_streams.end()) {
        *out = it->second;
        return IMRESH_OK;
    }
    hipStream_t s;
    IMR_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    g_streams[device] = s;
    *out = s;
    return IMRESH_OK;
}

struct Class {
    int W, H;
    std::vector<int> idx;   // positions in the caller's list
    ClassBuf *buf = nullptr;
    Plan2D p;
    int nrowwg = 0, ntiles = 0;
};

// one cycle of one class (captured into the cycle graph)
int enqueue_cycle(const Class &k, const Params &q, const float *weights, int R, hipStream_t s)
{
    ClassBuf *c = k.buf;
    const int B = (int) k.idx.size();
    BlurArgs ba{};
    ba.in = c->Bn;
    ba.out = c->B2;
    ba.wh = ba.wv = weights;
    ba.rh = ba.rv = R;
    ba.W = k.W;
    ba.H = k.H;
    ba.pmax = c->pmax;
    ba.done = c->flags;
    ba.batch = B;
    ba.istride = c->n;
    ba.cycles = (const unsigned *) (c->flags + 1);
    IMR_CHECK(launch_blur(ba, s));
    IMR_CHECK(launch_max_finalize_batch(c->pmax, k.ntiles, q.cutA, q.cut, c->scal, c->flags, B, s));
    RowArgs ra{};
    ra.X = c->X;
    ra.G = c->G;
    ra.mask = c->mask;
    ra.Bn = c->Bn;
    ra.B2 = c->B2;
    ra.thr = c->scal;
    ra.beta = q.beta;
    ra.perr = c->perr;
    ra.pcnt = c->pcnt;
    ra.done = c->flags;
    ra.batch = B;
    ra.istride = c->n;
    ColArgs ca{};
    ca.X = c->X;
    ca.M = c->M;
    ca.done = c->flags;
    ca.batch = B;
    ca.istride = c->n;
    for (unsigned it = 0; it < q.n_hio; ++it) {
        IMR_CHECK(launch_row(it == 0 ? ROW_FIRST_CUT : ROW_MID, k.p, ra, s));
        IMR_CHECK(launch_col(COL_PROJ, k.p, ca, s));
    }
    IMR_CHECK(launch_row(ROW_LAST, k.p, ra, s));
    return launch_err_finalize_batch(c->perr, c->pcnt, k.nrowwg, q.target, c->scal, c->flags, B, s);
}

int run_batch(const imresh_image *imgs, uint32_t count, const imresh_params *params, hipStream_t user,
              bool device_io, imresh_result *results)
{
    if (count == 0) return IMRESH_OK;
    if (imgs == nullptr) return IMRESH_ERR_NULL;
    Params q;
    IMR_CHECK(normalize(params, q));
    int dev;
    IMR_CHECK(current_device(&dev));
    std::map<std::pair<int, int>, Class> classes;
    for (uint32_t i = 0; i < count; ++i) {
        const imresh_image &im = imgs[i];
        if (im.data == nullptr) return IMRESH_ERR_NULL;
        if (im.width == 0 || im.height == 0) return IMRESH_ERR_INVALID;
        if (!line_supported((int) im.width) || !line_supported((int) im.height))
            return IMRESH_ERR_UNSUPPORTED_SIZE;
        Class &k = classes[{(int) im.width, (int) im.height}];
        k.W = (int) im.width;
        k.H = (int) im.height;
        k.idx.push_back((int) i);
    }
    std::vector<float> wrows;
    int R = 0;
    IMR_CHECK(padded_weights(q, wrows, &R));

    std::lock_guard<std::mutex> run(run_lock(dev));
    hipStream_t s;
    IMR_CHECK(work_stream(dev, &s));
    hipEvent_t ev = nullptr;
    IMR_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    float *weights = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int st = IMRESH_OK;
    auto hip = [&](hipError_t e) {
        if (st == IMRESH_OK && e != hipSuccess) st = from_hip(e);
        return st == IMRESH_OK;
    };
    // the work stream follows the caller's stream
    hip(hipEventRecord(ev, user));
    hip(hipStreamWaitEvent(s, ev, 0));
    for (auto &kv : classes) {
        Class &k = kv.second;
        if (st == IMRESH_OK) st = tune_size(dev, k.W, k.H, s);
        if (st == IMRESH_OK) st = get_plan2d(dev, k.W, k.H, &k.p);
        if (st == IMRESH_OK) st = class_checkout(dev, k.W, k.H, (int) k.idx.size(), s, &k.buf);
        if (st != IMRESH_OK) break;
        k.nrowwg = row_grid(k.p);
        k.ntiles = blur_grid(k.W, k.H);
    }
    hip(hipMallocAsync((void **) &weights, sizeof(float) * wrows.size(), s));
    hip(hipMemcpyAsync(weights, wrows.data(), sizeof(float) * wrows.size(), hipMemcpyHostToDevice, s));

    // inputs, state reset and the initial inverse transform of every class
    std::vector<void *> tab;
    for (auto &kv : classes) {
        if (st != IMRESH_OK) break;
        Class &k = kv.second;
        ClassBuf *c = k.buf;
        const int B = (int) k.idx.size();
        if (device_io) {
            tab.assign(B, nullptr);
            for (int b = 0; b < B; ++b) tab[b] = imgs[k.idx[b]].data;
            hip(hipMemcpyAsync(c->tab, tab.data(), sizeof(void *) * B, hipMemcpyHostToDevice, s));
            hip(hipStreamSynchronize(s));   // tab is reused for the next class
            if (st == IMRESH_OK) st = launch_gather_images(c->M, (const float *const *) c->tab, c->n, B, s);
        } else {
            for (int b = 0; b < B; ++b)
                hip(hipMemcpyAsync(c->M + (size_t) b * c->n, imgs[k.idx[b]].data, sizeof(float) * c->n,
                                   hipMemcpyHostToDevice, s));
        }
        hip(hipMemsetAsync(c->scal, 0, sizeof(float) * kScalStride * B, s));
        hip(hipMemsetAsync(c->flags, 0, sizeof(int) * kFlagStride * B, s));
        ColArgs ca{};
        ca.X = c->X;
        ca.M = c->M;
        ca.done = c->flags;
        ca.batch = B;
        ca.istride = c->n;
        if (st == IMRESH_OK) st = launch_col(COL_INIT, k.p, ca, s);
        RowArgs ra{};
        ra.X = c->X;
        ra.G = c->G;
        ra.mask = c->mask;
        ra.Bn = c->Bn;
        ra.done = c->flags;
        ra.batch = B;
        ra.istride = c->n;
        if (st == IMRESH_OK) st = launch_row(ROW_INIT, k.p, ra, s);
    }

    // one cycle of every class as a graph, each class on its own branch;
    // without graphs every class runs all its cycles on its own stream
    if (st == IMRESH_OK && q.n_cycles > 0) st = prepare_blur(R);
    if (st == IMRESH_OK && q.n_cycles > 0 && !g_use_graphs) {
        hip(hipEventRecord(ev, s));
        for (auto &kv : classes) {
            Class &k = kv.second;
            if (st != IMRESH_OK) break;
            hip(hipStreamWaitEvent(k.buf->branch, ev, 0));
            for (unsigned cyc = 0; cyc < q.n_cycles && st == IMRESH_OK; ++cyc)
                st = enqueue_cycle(k, q, weights, R, k.buf->branch);
            hip(hipEventRecord(k.buf->join, k.buf->branch));
            hip(hipStreamWaitEvent(s, k.buf->join, 0));
        }
    } else if (st == IMRESH_OK && q.n_cycles > 0) {
        hip(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        hip(hipEventRecord(ev, s));
        for (auto &kv : classes) {
            Class &k = kv.second;
            if (st != IMRESH_OK) break;
            hip(hipStreamWaitEvent(k.buf->branch, ev, 0));
            if (st == IMRESH_OK) st = enqueue_cycle(k, q, weights, R, k.buf->branch);
            hip(hipEventRecord(k.buf->join, k.buf->branch));
            hip(hipStreamWaitEvent(s, k.buf->join, 0));
        }
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        hip(ec);
        if (st == IMRESH_OK) hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        for (unsigned cyc = 0; cyc < q.n_cycles && st == IMRESH_OK; ++cyc) hip(hipGraphLaunch(exec, s));
    }

    // outputs Re(g') and the per-image results
    std::vector<float> sc;
    std::vector<int> fl;
    for (auto &kv : classes) {
        Class &k = kv.second;
        ClassBuf *c = k.buf;
        if (c == nullptr) continue;
        const int B = (int) k.idx.size();
        if (st == IMRESH_OK) {
            if (device_io) {
                tab.assign(B, nullptr);
                for (int b = 0; b < B; ++b) tab[b] = imgs[k.idx[b]].data;
                hip(hipMemcpyAsync(c->tab, tab.data(), sizeof(void *) * B, hipMemcpyHostToDevice, s));
                hip(hipStreamSynchronize(s));
                if (st == IMRESH_OK) st = launch_scatter_real((float *const *) c->tab, c->X, c->n, B, s);
            } else {
                if (st == IMRESH_OK) st = launch_copy_from_real(c->B2, c->X, c->n * B, s);
                for (int b = 0; b < B; ++b)
                    hip(hipMemcpyAsync(imgs[k.idx[b]].data, c->B2 + (size_t) b * c->n, sizeof(float) * c->n,
                                       hipMemcpyDeviceToHost, s));
            }
        }
        if (st == IMRESH_OK && results != nullptr) {
            sc.assign((size_t) kScalStride * B, 0.f);
            fl.assign((size_t) kFlagStride * B, 0);
            hip(hipMemcpyAsync(sc.data(), c->scal, sizeof(float) * sc.size(), hipMemcpyDeviceToHost, s));
            hip(hipMemcpyAsync(fl.data(), c->flags, sizeof(int) * fl.size(), hipMemcpyDeviceToHost, s));
            hip(hipStreamSynchronize(s));
            for (int b = 0; b < B && st == IMRESH_OK; ++b) {
                imresh_result &r = results[k.idx[b]];
                r.cycles_run = (uint32_t) fl[(size_t) b * kFlagStride + 1];
                r.last_error = sc[(size_t) b * kScalStride + 1];
                r.last_total = sc[(size_t) b * kScalStride + 2];
                r.last_masked = sc[(size_t) b * kScalStride + 3];
            }
        }
    }
    hipError_t es = hipStreamSynchronize(s);
    hip(es);
    for (auto &kv : classes)
        if (kv.second.buf) class_return(kv.second.buf, s);
    if (exec) (void) hipGraphExecDestroy(exec);
    if (graph) (void) hipGraphDestroy(graph);
    if (weights) (void) hipFreeAsync(weights, s);
    (void) hipEventDestroy(ev);
    return st;
}

}  // namespace

}  // namespace imr

namespace imr {

int trim_idle_classes(size_t cap)
{
    std::vector<ClassBuf *> drop;
    {
        std::lock_guard<std::mutex> lk(g_mtx);
        trim_classes(cap, drop);
    }
    for (ClassBuf *x : drop) class_free(x);
    return IMRESH_OK;
}

size_t idle_class_bytes()
{
    std::lock_guard<std::mutex> lk(g_mtx);
    return g_idle;
}

}  // namespace imr

extern "C" {

int imresh_shrinkwrap_batch(const imresh_image *images, uint32_t count, const imresh_params *params,
                            void *stream, imresh_result *results)
{
    return imr::guarded([&] { return imr::run_batch(images, count, params, (hipStream_t) stream, false, results); });
}

int imresh_shrinkwrap_batch_device(const imresh_image *images, uint32_t count,
                                   const imresh_params *params, void *stream, imresh_result *results)
{
    return imr::guarded([&] { return imr::run_batch(images, count, params, (hipStream_t) stream, true, results); });
}

}  // extern "C"
