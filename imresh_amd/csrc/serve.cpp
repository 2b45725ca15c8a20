// Reconstruction server: continuous batching of shrink-wrap requests of
// mixed sizes under an HBM budget.  The reference's task queue
// (taskQueue.cu:91-193) starts one host thread per request on one of 3
// streams per device and keeps no state between calls; here one scheduler
// thread per server
//
//   * admits queued requests in arrival order while the number of active
//     requests is below max_active and their device state (iterate, gPrev,
//     |g'|, blurred |g'|, modulus, mask: ws_bytes(W, H)) fits the budget;
//     idle cached states are released first when a new one does not fit;
//   * keeps every admitted request's state resident on the device and
//     advances all active requests cycle by cycle, at most kDepth cycles
//     queued per request, so requests sharing a stream interleave at cycle
//     granularity and a request admitted while others run starts at once;
//   * reads each cycle's early-exit flag back asynchronously (pinned stage)
//     and stops queueing cycles for a converged request;
//   * completes a request when its last queued work has finished: Re(g')
//     in the caller's buffer, the result record, the latency (submit to
//     completion), the callback.
//
// The scheduler thread never waits on a lane: host buffers go through
// pinned staging (a copy from or to pageable memory would hold the thread
// until the lane's earlier work, other requests' cycles included, had
// finished), and a new size's kernel variants are timed on a stream of
// their own.
//
// Results are those of imresh_shrinkwrap on the same input: the same
// kernels run on the request's own state, only interleaved with others.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace imr {

namespace {

constexpr unsigned kDepth = 2;   // cycles queued ahead per request

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    float *io = nullptr;
    uint32_t W = 0, H = 0;
    int device_io = 0;
    Params q{};
    imresh_server_callback cb = nullptr;
    void *user = nullptr;
    Clock::time_point t_submit;
    size_t bytes = 0;
    // active state
    Workspace *w = nullptr;
    hipStream_t lane = nullptr;
    unsigned next_cycle = 0;
    std::deque<std::pair<hipEvent_t, unsigned>> inflight;   // (event, cycle) in order
    bool converged = false;
    bool finishing = false;
    hipEvent_t fin = nullptr;
    int status = IMRESH_OK;
    void *pin = nullptr;   // pinned staging of the host buffer (host I/O only)
};

// Pinned host buffers by power-of-two capacity, reused across requests
// (scheduler thread only).
struct PinnedPool {
    std::map<size_t, std::vector<void *>> free;
    std::vector<void *> all;

    static size_t cap(size_t bytes)
    {
        size_t c = 64 << 10;
        while (c < bytes) c <<= 1;
        return c;
    }
    void *get(size_t bytes)
    {
        auto &v = free[cap(bytes)];
        if (!v.empty()) {
            void *p = v.back();
            v.pop_back();
            return p;
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, cap(bytes), hipHostMallocDefault) != hipSuccess) return nullptr;
        all.push_back(p);
        return p;
    }
    void put(void *p, size_t bytes)
    {
        if (p) free[cap(bytes)].push_back(p);
    }
    void release()
    {
        for (void *p : all) (void) hipHostFree(p);
        all.clear();
        free.clear();
    }
};

}  // namespace

struct Server;
// the server whose scheduler thread this is (completion callbacks run there):
// a drain or destroy of that server from a callback would wait for the very
// thread running it
thread_local const Server *t_scheduler_of = nullptr;

struct Server {
    int device = 0;
    uint64_t budget = 0;
    uint32_t max_active = 0;
    std::vector<hipStream_t> lanes;
    hipStream_t tune = nullptr;   // kernel-variant timing of new sizes
    unsigned next_lane = 0;
    PinnedPool pinned;

    std::mutex mtx;
    std::condition_variable cv;       // work arrived / stop
    std::condition_variable cv_idle;  // drained
    std::deque<Request *> queue;
    std::vector<Request *> active;
    std::deque<imresh_server_completion> done;
    bool stop = false;
    uint64_t next_id = 1;
    uint64_t submitted = 0, completed = 0;

    // device memory: workspaces in use + idle cached ones, per size
    uint64_t used = 0, peak = 0;
    std::map<std::pair<int, int>, std::vector<Workspace *>> idle;
    std::vector<hipEvent_t> event_pool;

    std::thread thr;

    hipEvent_t get_event()
    {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        return e;
    }
    void put_event(hipEvent_t e)
    {
        if (e) event_pool.push_back(e);
    }

    // a workspace for W x H within the budget (scheduler thread only)
    int acquire(int W, int H, Workspace **out)
    {
        auto &v = idle[{W, H}];
        if (!v.empty()) {
            *out = v.back();
            v.pop_back();
            return IMRESH_OK;
        }
        const uint64_t need = ws_bytes(W, H);
        for (auto it = idle.begin(); it != idle.end() && used + need > budget; ++it)
            while (!it->second.empty() && used + need > budget) {
                Workspace *w = it->second.back();
                it->second.pop_back();
                used -= ws_bytes(w->W, w->H);
                ws_release(w);
            }
        if (used + need > budget) return IMRESH_ERR_BUDGET;
        IMR_CHECK(ws_alloc(device, W, H, out));
        used += need;
        if (used > peak) peak = used;
        return IMRESH_OK;
    }

    void finish(Request *r, const imresh_result &res)
    {
        imresh_server_completion c{};
        c.id = r->id;
        c.status = r->status;
        c.result = res;
        c.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - r->t_submit).count();
        if (r->w) idle[{(int) r->W, (int) r->H}].push_back(r->w);
        pinned.put(r->pin, sizeof(float) * (size_t) r->W * r->H);
        r->pin = nullptr;
        for (auto &e : r->inflight) put_event(e.first);
        put_event(r->fin);
        if (r->cb) r->cb(r->user, &c, r->io);
        {
            std::lock_guard<std::mutex> lk(mtx);
            done.push_back(c);
            ++completed;
        }
        delete r;
        cv_idle.notify_all();
    }

    // admission in arrival order
    bool admit()
    {
        bool progress = false;
        for (;;) {
            Request *r = nullptr;
            {
                std::lock_guard<std::mutex> lk(mtx);
                if (queue.empty() || active.size() >= max_active) break;
                r = queue.front();
            }
            Workspace *w = nullptr;
            const int st = acquire((int) r->W, (int) r->H, &w);
            if (st == IMRESH_ERR_BUDGET && !active.empty()) break;   // wait for memory
            {
                std::lock_guard<std::mutex> lk(mtx);
                queue.pop_front();
            }
            progress = true;
            if (st != IMRESH_OK) {
                r->status = st;
                finish(r, imresh_result{});
                continue;
            }
            r->w = w;
            r->lane = lanes[next_lane++ % lanes.size()];
            const size_t bytes = sizeof(float) * w->n;
            const float *M = r->device_io ? r->io : w->M;
            int s2 = ws_tune(w, tune);
            if (s2 == IMRESH_OK && !r->device_io) {
                r->pin = pinned.get(bytes);
                if (r->pin == nullptr) {
                    s2 = IMRESH_ERR_OUT_OF_MEMORY;
                } else {
                    std::memcpy(r->pin, r->io, bytes);
                    s2 = IMR_HIP_STATUS(hipMemcpyAsync(w->M, r->pin, bytes, hipMemcpyHostToDevice, r->lane));
                }
            }
            if (s2 == IMRESH_OK) s2 = recon_begin(w, M, r->q, r->lane);
            r->status = s2;
            std::lock_guard<std::mutex> lk(mtx);
            active.push_back(r);
        }
        return progress;
    }

    // queue cycles / the closing step of every active request
    bool feed()
    {
        bool progress = false;
        for (Request *r : active) {
            if (r->finishing || r->status != IMRESH_OK) continue;
            const float *M = r->device_io ? r->io : r->w->M;
            while (!r->converged && r->next_cycle < r->q.n_cycles && r->inflight.size() < kDepth) {
                int st = recon_cycle(r->w, M, r->q, r->next_cycle, r->lane);
                if (st == IMRESH_OK) st = recon_stage_results(r->w, r->lane);
                hipEvent_t e = get_event();
                if (st == IMRESH_OK && e == nullptr) st = IMRESH_ERR_HIP;
                if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventRecord(e, r->lane));
                if (st != IMRESH_OK) {
                    put_event(e);
                    r->status = st;
                    break;
                }
                r->inflight.emplace_back(e, r->next_cycle++);
                progress = true;
            }
            if (r->status == IMRESH_OK && r->inflight.empty() &&
                (r->converged || r->next_cycle >= r->q.n_cycles)) {
                float *out = r->device_io ? r->io : r->w->M;
                int st = recon_end(r->w, out, r->lane);
                if (st == IMRESH_OK && !r->device_io)
                    st = IMR_HIP_STATUS(hipMemcpyAsync(r->pin, r->w->M, sizeof(float) * r->w->n,
                                                 hipMemcpyDeviceToHost, r->lane));
                if (st == IMRESH_OK) st = recon_stage_results(r->w, r->lane);
                r->fin = get_event();
                if (st == IMRESH_OK && r->fin == nullptr) st = IMRESH_ERR_HIP;
                if (st == IMRESH_OK) st = IMR_HIP_STATUS(hipEventRecord(r->fin, r->lane));
                r->status = st;
                r->finishing = true;
                progress = true;
            }
        }
        return progress;
    }

    // retire finished cycles and requests
    bool poll()
    {
        bool progress = false;
        for (size_t i = 0; i < active.size();) {
            Request *r = active[i];
            while (!r->inflight.empty() && hipEventQuery(r->inflight.front().first) == hipSuccess) {
                put_event(r->inflight.front().first);
                r->inflight.pop_front();
                int d = 0;
                recon_read_staged(r->w, nullptr, &d);
                if (d) r->converged = true;
                progress = true;
            }
            bool gone = false;
            if (r->status != IMRESH_OK) {
                (void) hipStreamSynchronize(r->lane);   // drain what was queued before the error
                gone = true;
            } else if (r->finishing && hipEventQuery(r->fin) == hipSuccess) {
                gone = true;
            }
            if (gone) {
                imresh_result res{};
                if (r->status == IMRESH_OK) {
                    recon_read_staged(r->w, &res, nullptr);
                    if (!r->device_io) std::memcpy(r->io, r->pin, sizeof(float) * r->w->n);
                }
                {
                    std::lock_guard<std::mutex> lk(mtx);
                    active.erase(active.begin() + (long) i);
                }
                finish(r, res);
                progress = true;
            } else {
                ++i;
            }
        }
        return progress;
    }

    void loop()
    {
        (void) hipSetDevice(device);   // the creating thread's current device, a valid index
        t_scheduler_of = this;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mtx);
                if (active.empty()) {
                    cv.wait(lk, [&] { return stop || !queue.empty(); });
                    if (stop && queue.empty()) break;
                }
            }
            const bool a = admit();
            const bool f = feed();
            const bool p = poll();
            if (!(a || f || p)) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
};

}  // namespace imr

using imr::Server;

extern "C" {

struct imresh_server {
    Server s;
};

static int server_create_impl(uint64_t hbm_budget_bytes, uint32_t max_active, uint32_t lanes, imresh_server **out)
{
    if (out == nullptr) return IMRESH_ERR_NULL;
    *out = nullptr;
    int dev = 0;
    IMR_CHECK(imr::current_device(&dev));
    imresh_server *h = new imresh_server();
    Server &s = h->s;
    s.device = dev;
    s.budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t) 200 << 30;
    // 16 requests on 16 streams: on the serving trace (sizes 128-2048, Poisson
    // arrivals) 8 on 8 reached 121 reconstructions/s, 16 on 16 126-127 with a
    // lower p99, 32 on 32 no more
    s.max_active = max_active ? max_active : 16;
    const uint32_t nl = lanes ? lanes : 16;
    for (uint32_t i = 0; i < nl; ++i) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : s.lanes) (void) hipStreamDestroy(x);
            delete h;
            return IMRESH_ERR_HIP;
        }
        s.lanes.push_back(st);
    }
    if (hipStreamCreateWithFlags(&s.tune, hipStreamNonBlocking) != hipSuccess) {
        for (hipStream_t x : s.lanes) (void) hipStreamDestroy(x);
        delete h;
        return IMRESH_ERR_HIP;
    }
    try {
        s.thr = std::thread([&s] { s.loop(); });
    } catch (...) {
        for (hipStream_t x : s.lanes) (void) hipStreamDestroy(x);
        (void) hipStreamDestroy(s.tune);
        delete h;
        return IMRESH_ERR_INTERNAL;
    }
    *out = h;
    return IMRESH_OK;
}

static int server_submit_impl(imresh_server *h, float *io, uint32_t width, uint32_t height,
                       const imresh_params *params, int device_io, imresh_server_callback cb, void *user,
                       uint64_t *id)
{
    if (h == nullptr || io == nullptr) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    if (!imr::line_supported((int) width) || !imr::line_supported((int) height))
        return IMRESH_ERR_UNSUPPORTED_SIZE;
    Server &s = h->s;
    imr::Params q{};
    IMR_CHECK(imr::normalize(params, q));
    auto *r = new imr::Request();
    r->q = q;
    r->io = io;
    r->W = width;
    r->H = height;
    r->device_io = device_io;
    r->cb = cb;
    r->user = user;
    r->bytes = imr::ws_bytes((int) width, (int) height);
    if (r->bytes > s.budget) {
        delete r;
        return IMRESH_ERR_BUDGET;
    }
    r->t_submit = imr::Clock::now();
    {
        std::lock_guard<std::mutex> lk(s.mtx);
        r->id = s.next_id++;
        ++s.submitted;
        s.queue.push_back(r);
    }
    if (id) *id = r->id;
    s.cv.notify_one();
    return IMRESH_OK;
}

int imresh_server_create(uint64_t hbm_budget_bytes, uint32_t max_active, uint32_t lanes,
                         imresh_server **out)
{
    return imr::guarded([&] { return server_create_impl(hbm_budget_bytes, max_active, lanes, out); });
}

int imresh_server_submit(imresh_server *h, float *io, uint32_t width, uint32_t height,
                         const imresh_params *params, int device_io, imresh_server_callback cb,
                         void *user, uint64_t *id)
{
    return imr::guarded(
        [&] { return server_submit_impl(h, io, width, height, params, device_io, cb, user, id); });
}

int imresh_server_poll(imresh_server *h, imresh_server_completion *out, uint32_t capacity,
                       uint32_t *count)
{
    if (h == nullptr || count == nullptr || (out == nullptr && capacity > 0)) return IMRESH_ERR_NULL;
    Server &s = h->s;
    std::lock_guard<std::mutex> lk(s.mtx);
    uint32_t n = 0;
    while (n < capacity && !s.done.empty()) {
        out[n++] = s.done.front();
        s.done.pop_front();
    }
    *count = n;
    return IMRESH_OK;
}

int imresh_server_drain(imresh_server *h)
{
    if (h == nullptr) return IMRESH_ERR_NULL;
    Server &s = h->s;
    if (imr::t_scheduler_of == &s) return IMRESH_ERR_INVALID;   // from a completion callback
    std::unique_lock<std::mutex> lk(s.mtx);
    s.cv_idle.wait(lk, [&] { return s.completed == s.submitted; });
    return IMRESH_OK;
}

int imresh_server_stats(imresh_server *h, imresh_server_info *info)
{
    if (h == nullptr || info == nullptr) return IMRESH_ERR_NULL;
    Server &s = h->s;
    std::lock_guard<std::mutex> lk(s.mtx);
    info->submitted = s.submitted;
    info->completed = s.completed;
    info->queued = s.queue.size();
    info->active = s.active.size();
    info->bytes_in_use = s.used;
    info->peak_bytes = s.peak;
    info->budget = s.budget;
    return IMRESH_OK;
}

int imresh_server_destroy(imresh_server *h)
{
    if (h == nullptr) return IMRESH_ERR_NULL;
    Server &s = h->s;
    if (imr::t_scheduler_of == &s) return IMRESH_ERR_INVALID;   // from a completion callback
    imresh_server_drain(h);
    {
        std::lock_guard<std::mutex> lk(s.mtx);
        s.stop = true;
    }
    s.cv.notify_all();
    if (s.thr.joinable()) s.thr.join();
    for (hipStream_t st : s.lanes) (void) hipStreamSynchronize(st);
    for (auto &kv : s.idle)
        for (imr::Workspace *w : kv.second) imr::ws_release(w);
    for (hipEvent_t e : s.event_pool) (void) hipEventDestroy(e);
    for (hipStream_t st : s.lanes) (void) hipStreamDestroy(st);
    (void) hipStreamDestroy(s.tune);
    s.pinned.release();
    delete h;
    return IMRESH_OK;
}

}  // extern "C"
