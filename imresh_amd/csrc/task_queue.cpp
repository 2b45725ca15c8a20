// Multi-stream task queue and the C++ drop-in API.
//
// Reference behaviour (src/imresh/io/taskQueue.cu:91-251): taskQueueInit
// creates 3 streams per device; addTask starts one std::thread per task and,
// when as many threads as streams exist, joins the OLDEST one first; each
// thread takes the least recently used stream, runs cudaShrinkWrap (with
// per-call cudaMalloc / cufftPlan2d) and then the write-out callback.
//
// Here: one persistent worker thread per (device, stream) slot, each with a
// reusable device workspace (the engine's workspace pool), fed from a
// bounded FIFO.  addTask blocks while the FIFO holds as many pending tasks as
// there are slots, so at most 2 x slots tasks are in flight (the reference
// allowed `slots`); callbacks run on the worker threads, as in the reference.
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "imresh/algorithms/shrinkWrap.hpp"
#include "imresh/libs/diffractionIntensity.hpp"
#include "imresh/libs/fftShift.hpp"
#include "imresh/libs/hybridInputOutput.hpp"
#include "imresh/io/taskQueue.hpp"
#include "internal.hpp"

namespace {

struct Task {
    float *io;
    uint32_t w, h;
    imresh_params params;
    std::function<void(float *, uint32_t, uint32_t, int)> done;
};

struct Slot {
    int device;
    hipStream_t stream;
    std::thread worker;
};

struct Queue {
    std::mutex mtx;
    std::condition_variable cv_work, cv_space;
    std::deque<Task> pending;
    std::vector<Slot> slots;
    size_t capacity = 0;
    bool running = false;
    bool closing = false;
} g_q;

std::mutex g_init_mtx;
// set on the queue's worker threads: a deinit from a task callback would
// have to join its own thread
thread_local bool t_queue_worker = false;

void worker_main(size_t idx)
{
    t_queue_worker = true;
    const Slot &slot = g_q.slots[idx];
    (void) hipSetDevice(slot.device);   // selected once already in queue_init
    for (;;) {
        Task t;
        {
            std::unique_lock<std::mutex> lk(g_q.mtx);
            g_q.cv_work.wait(lk, [] { return g_q.closing || !g_q.pending.empty(); });
            if (g_q.pending.empty()) return;   // closing and drained
            t = std::move(g_q.pending.front());
            g_q.pending.pop_front();
        }
        g_q.cv_space.notify_one();
        const int st = imresh_shrinkwrap(t.io, t.w, t.h, &t.params, (void *) slot.stream, nullptr);
        if (t.done) t.done(t.io, t.w, t.h, st);
    }
}

int queue_init(int per_device)
{
    std::lock_guard<std::mutex> lk(g_init_mtx);
    if (g_q.running) return IMRESH_OK;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return IMRESH_ERR_NO_DEVICE;
    if (per_device <= 0) per_device = 3;   // taskQueue.cu:217
    int caller_dev = 0;
    e = hipGetDevice(&caller_dev);
    if (e != hipSuccess) return imr::from_hip(e);
    g_q.closing = false;
    g_q.slots.clear();
    // streams on every device; on a failure those created so far are
    // destroyed and the queue stays uninitialized
    auto unwind = [&](hipError_t err) {
        for (auto &s : g_q.slots) (void) hipStreamDestroy(s.stream);
        g_q.slots.clear();
        (void) hipSetDevice(caller_dev);   // restore only; err is the status
        return imr::from_hip(err);
    };
    for (int d = 0; d < ndev; ++d) {
        e = hipSetDevice(d);
        if (e != hipSuccess) return unwind(e);
        for (int j = 0; j < per_device; ++j) {
            Slot s;
            s.device = d;
            e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
            if (e != hipSuccess) return unwind(e);
            g_q.slots.push_back(std::move(s));
        }
    }
    (void) hipSetDevice(caller_dev);   // the caller's current device is left as it was
    g_q.capacity = g_q.slots.size();
    g_q.running = true;
    for (size_t i = 0; i < g_q.slots.size(); ++i) g_q.slots[i].worker = std::thread(worker_main, i);
    return IMRESH_OK;
}

int queue_deinit()
{
    if (t_queue_worker) return IMRESH_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_init_mtx);
    if (!g_q.running) return IMRESH_OK;
    {
        std::lock_guard<std::mutex> l2(g_q.mtx);
        g_q.closing = true;
    }
    // producers blocked in queue_add wake up and fail; workers drain the FIFO
    g_q.cv_space.notify_all();
    g_q.cv_work.notify_all();
    for (auto &s : g_q.slots)
        if (s.worker.joinable()) s.worker.join();
    int st = IMRESH_OK;
    for (auto &s : g_q.slots) {
        hipError_t e = hipStreamDestroy(s.stream);
        if (e != hipSuccess) st = imr::from_hip(e);
    }
    g_q.slots.clear();
    g_q.running = false;
    return st;
}

int queue_add(Task &&t)
{
    std::unique_lock<std::mutex> lk(g_q.mtx);
    if (!g_q.running || g_q.closing) return IMRESH_ERR_NOT_INITIALIZED;
    g_q.cv_space.wait(lk, [] { return g_q.closing || g_q.pending.size() < g_q.capacity; });
    if (g_q.closing) return IMRESH_ERR_NOT_INITIALIZED;
    g_q.pending.push_back(std::move(t));
    lk.unlock();
    g_q.cv_work.notify_one();
    return IMRESH_OK;
}

}  // namespace

extern "C" {

int imresh_task_queue_init(int streams_per_device)
{
    return imr::guarded([&] { return queue_init(streams_per_device); });
}

int imresh_task_queue_deinit(void)
{
    return imr::guarded([] { return queue_deinit(); });
}

int imresh_add_task(float *io_host, uint32_t width, uint32_t height, const imresh_params *params,
                    imresh_task_callback callback, void *user)
{
    if (io_host == nullptr) return IMRESH_ERR_NULL;
    if (width == 0 || height == 0) return IMRESH_ERR_INVALID;
    Task t;
    t.io = io_host;
    t.w = width;
    t.h = height;
    if (params) t.params = *params;
    else {
        imresh_params_default(&t.params);
        t.params.defaults = IMRESH_DEFAULTS_QUEUE;
    }
    return imr::guarded([&] {
        if (callback)
            t.done = [callback, user](float *io, uint32_t w, uint32_t h, int st) { callback(user, io, w, h, st); };
        return queue_add(std::move(t));
    });
}

}  // extern "C"

// ------------------------------------------------------------ C++ drop-in API
namespace imresh {
namespace algorithms {

int shrinkWrap(float *const rIoData, unsigned int const rImageWidth, unsigned int const rImageHeight,
               unsigned int rnCycles, float rTargetError, float rHioBeta,
               float rIntensityCutOffAutoCorel, float rIntensityCutOff, float sigma0,
               float rSigmaChange, unsigned int rnHioCycles)
{
    if (rIoData == nullptr) return 1;   // shrinkWrap.cpp:133
    imresh_params p;
    p.n_cycles = rnCycles;
    p.target_error = rTargetError;
    p.hio_beta = rHioBeta;
    p.cutoff_autocorel = rIntensityCutOffAutoCorel;
    p.cutoff = rIntensityCutOff;
    p.sigma0 = sigma0;
    p.sigma_change = rSigmaChange;
    p.n_hio_cycles = rnHioCycles;
    p.defaults = IMRESH_DEFAULTS_CPU;
    return imresh_shrinkwrap(rIoData, rImageWidth, rImageHeight, &p, nullptr, nullptr);
}

}  // namespace algorithms

namespace io {

void taskQueueInit() { queue_init(0); }

void taskQueueDeinit() { queue_deinit(); }

void addTask(float *const ioData, unsigned int const imageWidth, unsigned int const imageHeight,
             WriteOutCallBack writeOutFunc, std::string fileName, unsigned int nCycles,
             unsigned int nHioCycles, float targetError, float hioBeta,
             float intensityCutOffAutoCorel, float intensityCutOff, float sigma0, float sigmaChange)
{
    Task t;
    t.io = ioData;
    t.w = imageWidth;
    t.h = imageHeight;
    t.params.n_cycles = nCycles;
    t.params.n_hio_cycles = nHioCycles;
    t.params.target_error = targetError;
    t.params.hio_beta = hioBeta;
    t.params.cutoff_autocorel = intensityCutOffAutoCorel;
    t.params.cutoff = intensityCutOff;
    t.params.sigma0 = sigma0;
    t.params.sigma_change = sigmaChange;
    t.params.defaults = IMRESH_DEFAULTS_QUEUE;
    t.done = [writeOutFunc, fileName](float *io, uint32_t w, uint32_t h, int) {
        if (writeOutFunc) writeOutFunc(io, w, h, fileName);
    };
    queue_add(std::move(t));
}

}  // namespace io
namespace libs {

int hybridInputOutput(float *const rIoData, uint8_t const *const rIsMasked,
                      std::vector<unsigned int> const rSize, unsigned int rnCycles, float rTargetError,
                      float rBeta, unsigned int /* rnCores */)
{
    if (rSize.size() != 2) return 1;   // hybridInputOutput.cpp:209
    if (rIoData == nullptr || rIsMasked == nullptr) return 1;
    const int st = imresh_hybrid_input_output(rIoData, rIsMasked, rSize[0], rSize[1], rnCycles,
                                              rTargetError, rBeta, nullptr, nullptr);
    return st == IMRESH_OK ? 0 : st;
}

void diffractionIntensity(float *rIoData, unsigned int rImageWidth, unsigned int rImageHeight,
                          hipStream_t rStream, bool rAsync)
{
    // the reference exits on a CUDA or cuFFT error; here the status is dropped
    // as the signature has no place for it (the C entry point reports it)
    imresh_diffraction_intensity(rIoData, rImageWidth, rImageHeight, (void *) rStream, rAsync ? 1 : 0);
}

template <class T_COMPLEX>
void fftShift(T_COMPLEX *const data, unsigned int const Nx, unsigned int const Ny)
{
    // the reference returns nothing and asserts on a null pointer
    if (data == nullptr || Nx == 0 || Ny == 0) return;
    const size_t bytes = sizeof(T_COMPLEX) * (size_t) Nx * Ny;
    void *d = nullptr;
    if (imresh_dev_alloc(&d, bytes) != IMRESH_OK) return;
    if (imresh_memcpy_h2d(d, data, bytes, nullptr) == IMRESH_OK &&
        imresh_fft_shift(d, Nx, Ny, (uint32_t) sizeof(T_COMPLEX), nullptr) == IMRESH_OK)
        imresh_memcpy_d2h(data, d, bytes, nullptr);
    imresh_stream_sync(nullptr);
    imresh_dev_free(d);
}

template void fftShift<float>(float *const, unsigned int const, unsigned int const);
template void fftShift<float[2]>(float (*const)[2], unsigned int const, unsigned int const);

}  // namespace libs
}  // namespace imresh
