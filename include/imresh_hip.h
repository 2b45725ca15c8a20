/*
 * imresh_hip.h -- C ABI of the MI355X (gfx950) shrink-wrap / HIO library.
 *
 * Plain C: pointers, sizes and an opaque HIP stream (hipStream_t passed as
 * void*; NULL = the default stream).  Every function returns an int status
 * (IMRESH_OK = 0); nothing calls exit()/abort() -- the reference printed and
 * exited on any CUDA/cuFFT error (src/imresh/libs/cudacommon.cu:39-56,
 * src/imresh/libs/checkCufftError.cpp:82-91).
 *
 * Memory convention: "dev" pointers are device (HBM) pointers, "host"
 * pointers are ordinary host memory.  Complex arrays are interleaved float
 * pairs (re, im), row-major H rows x W columns, DC at [0,0], matching
 * cufftComplex / fftwf_complex in the reference.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef IMRESH_HIP_H
#define IMRESH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMRESH_ABI_VERSION 1

enum imresh_status {
    IMRESH_OK = 0,
    IMRESH_ERR_NULL = 1,             /* null pointer (shrinkWrap.cpp:133 returns 1) */
    IMRESH_ERR_INVALID = 2,          /* invalid size / parameter                 */
    IMRESH_ERR_UNSUPPORTED_SIZE = 3, /* no FFT plan for this length               */
    IMRESH_ERR_OUT_OF_MEMORY = 4,
    IMRESH_ERR_NO_DEVICE = 5,        /* taskQueue.cu:206-211 exited instead       */
    IMRESH_ERR_HIP = 6,              /* any other HIP runtime error               */
    IMRESH_ERR_KERNEL_TOO_LARGE = 7, /* Gaussian kernel > 63 taps (gaussian.cpp:189-192) */
    IMRESH_ERR_NOT_INITIALIZED = 8,  /* addTask before taskQueueInit (taskQueue.cu:163) */
    IMRESH_ERR_BUDGET = 9,           /* request exceeds the configured HBM budget */
    IMRESH_ERR_IO = 10,              /* file unreadable / unwritable or format unsupported */
    IMRESH_ERR_INTERNAL = 11,        /* unexpected host-side failure (e.g. thread creation) */
    IMRESH_ERR_TOO_LARGE = 12        /* beyond the fixed scratch of the allocation-free entry point */
};

/* Default-substitution rule for parameters <= 0 (or == 0 for counts). */
enum imresh_defaults {
    /* shrinkWrap.cpp:133-140: n_cycles has no default (0 => no cycles),
     * cutoff_autocorel default 0.04 */
    IMRESH_DEFAULTS_CPU = 0,
    /* cudaShrinkWrap.cu:193-200 (the taskQueue path): n_cycles 0 => 20,
     * cutoff_autocorel default 0.02 */
    IMRESH_DEFAULTS_QUEUE = 1
};

/* Parameters of one reconstruction; names and meaning follow
 * shrinkWrap.hpp:40-53 / cudaShrinkWrap.hpp:114-130. */
typedef struct imresh_params {
    uint32_t n_cycles;        /* shrink-wrap cycles                        */
    float    target_error;    /* early exit when HIO error < this          */
    float    hio_beta;        /* HIO feedback beta                         */
    float    cutoff_autocorel;/* mask cut-off of cycle 0 (rel. to max)     */
    float    cutoff;          /* mask cut-off of later cycles              */
    float    sigma0;          /* first Gaussian blur sigma                 */
    float    sigma_change;    /* sigma <- max(1.5, (1-sigma_change)*sigma) */
    uint32_t n_hio_cycles;    /* HIO iterations per cycle                  */
    int32_t  defaults;        /* enum imresh_defaults                      */
} imresh_params;

/* Outcome of one reconstruction (read back from the device). */
typedef struct imresh_result {
    uint32_t cycles_run;      /* cycles executed (early exit aware)        */
    float    last_error;      /* sqrtf(sum mask*|g'|)/sum mask of last cycle */
    float    last_total;      /* sum mask*|g'| of the last executed cycle  */
    float    last_masked;     /* sum mask of the last executed cycle       */
} imresh_result;

/* ---- library ----------------------------------------------------------- */
int         imresh_abi_version(void);
const char *imresh_strerror(int status);
/* Fills the reference defaults (shrinkWrap.hpp:45-52) into *p. */
int imresh_params_default(imresh_params *p);
/* Number of visible GPUs (taskQueue.cu:203-211 used cudaGetDeviceCount). */
int imresh_device_count(int *count);

/* ---- device memory helpers (plumbing for callers without a runtime) ---- */
int imresh_dev_alloc(void **dev, uint64_t bytes);
int imresh_dev_free(void *dev);
int imresh_memcpy_h2d(void *dev, const void *host, uint64_t bytes, void *stream);
int imresh_memcpy_d2h(void *host, const void *dev, uint64_t bytes, void *stream);
/* asynchronous device-to-device copy on `stream` */
int imresh_memcpy_d2d(void *dst, const void *src, uint64_t bytes, void *stream);
int imresh_stream_sync(void *stream);
/* Library options (process-wide).  IMRESH_OPT_CYCLE_GRAPHS: 1 (default)
 * replays each shrink-wrap cycle as a captured HIP graph, 0 launches its
 * kernels one by one; both compute the same bits.  IMRESH_OPT_POOL_MIB: the
 * device memory (MiB, default 32768) that idle per-size workspaces may keep
 * between calls; beyond it the least recently used idle ones are freed
 * (setting it also frees the idle per-operator scratch). */
enum { IMRESH_OPT_CYCLE_GRAPHS = 1, IMRESH_OPT_POOL_MIB = 2 };
int imresh_set_option(int option, int value);
/* Frees every idle pooled workspace and batch buffer whose work has
 * finished (all of them after the caller's streams are synchronized), the
 * idle scratch slots of the synchronous reductions and the idle scratch
 * buffers of the asynchronous per-operator entry points. */
int imresh_release_cached(void);
/* Device bytes held by idle pooled workspaces, batch buffers and reduction
 * scratch slots. */
int imresh_cached_bytes(uint64_t *bytes);

/* A non-blocking HIP stream at the default priority (for callers without a
 * HIP runtime of their own, e.g. to overlap reconstructions), and its
 * destruction. */
int imresh_stream_create(void **stream);
int imresh_stream_destroy(void *stream);

/* ---- whole reconstruction ---------------------------------------------- */
/* Replaces imresh::algorithms::shrinkWrap (src/imresh/algorithms/shrinkWrap.cpp:111-361,
 * decl shrinkWrap.hpp:40-53) and cudaShrinkWrap (src/imresh/algorithms/cuda/
 * cudaShrinkWrap.cu:164-420).  io_host holds the Fourier modulus |F| (W*H
 * floats) on entry and Re(g') on return, unnormalized exactly as the
 * reference.  Synchronous.  result may be NULL. */
int imresh_shrinkwrap(float *io_host, uint32_t width, uint32_t height,
                      const imresh_params *params, void *stream, imresh_result *result);

/* Same algorithm with the modulus already resident in HBM: io_dev (W*H
 * floats, device) is read as |F| and overwritten with Re(g').  Asynchronous
 * on `stream` when result == NULL; with result != NULL it synchronizes the
 * stream and reads the cycle count / error back. */
int imresh_shrinkwrap_device(float *io_dev, uint32_t width, uint32_t height,
                             const imresh_params *params, void *stream, imresh_result *result);

/* Allocation-free variant of imresh_shrinkwrap_device.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The reconstruction's state lives in a
 * fixed device array of the library (one per GPU, held until the process
 * ends) instead of a pooled workspace: the call allocates nothing, records
 * no memset and launches each cycle's kernels directly.  Limits: the
 * workspace of W x H (29 W*H bytes plus partial sums) must fit in
 * IMRESH_NOALLOC_MAX_BYTES (2048 x 2048 fits) and n_cycles must not exceed
 * IMRESH_NOALLOC_MAX_CYCLES, else IMRESH_ERR_TOO_LARGE.  The FFT tables of
 * the size must already exist on the device (built by any earlier call of
 * the size, e.g. imresh_autotune), else IMRESH_ERR_NOT_INITIALIZED.  Every
 * call on one GPU uses that GPU's single array: calls that overlap on one
 * GPU (other streams or threads) race.  The result equals
 * imresh_shrinkwrap_device's bit for bit.  Asynchronous when result ==
 * NULL (the only form usable under capture); otherwise synchronizes. */
#define IMRESH_NOALLOC_MAX_CYCLES 256
#define IMRESH_NOALLOC_MAX_BYTES (134217728ull - 65536ull)
int imresh_shrinkwrap_device_noalloc(float *io_dev, uint32_t width, uint32_t height,
                                     const imresh_params *params, void *stream,
                                     imresh_result *result);

/* Test hook (no reference counterpart; it exposes the state of the
 * reference's mask update, shrinkWrap.cpp:191-242 / cudaShrinkWrap.cu:
 * 237-277, as the fused kernels hold it): runs the reconstruction of the host
 * modulus with `params` through cycle `cycle` (< n_cycles) and returns
 *   gprime_host   : g' the cycle's mask is built from (W*H complex, 8 B each),
 *   mask_host     : that cycle's support mask, uint8 (1 = outside),
 *   blurred_host  : the blurred |g'| (W*H floats; may be NULL),
 *   threshold_host: cut_c * max of the blurred plane (may be NULL).
 * Synchronous. */
int imresh_debug_cycle_mask(const float *modulus_host, uint32_t width, uint32_t height,
                            const imresh_params *params, uint32_t cycle, float *gprime_host,
                            uint8_t *mask_host, float *blurred_host, float *threshold_host,
                            void *stream);

/* ---- batched reconstruction (no reference counterpart: the reference runs
 * one image per cudaShrinkWrap call and overlaps calls only through its task
 * queue, taskQueue.cu:91-193) -------------------------------------------- */
typedef struct imresh_image {
    float   *data;            /* width*height floats: |F| in, Re(g') out     */
    uint32_t width;
    uint32_t height;
} imresh_image;

/* Reconstructs `count` independent images of any supported sizes together
 * with one parameter set (the reference's per-call defaults).  Images of one
 * size are stacked and run through batched kernels; each cycle of all sizes
 * is one HIP graph replay.  Each image's result equals
 * imresh_shrinkwrap(images[i], params) run with the same kernel variants.
 * images[i].data are host buffers; results (may be NULL) receives one
 * imresh_result per image.  Synchronous. */
int imresh_shrinkwrap_batch(const imresh_image *images, uint32_t count,
                            const imresh_params *params, void *stream, imresh_result *results);
/* Same with images[i].data device buffers (inputs resident in HBM). */
int imresh_shrinkwrap_batch_device(const imresh_image *images, uint32_t count,
                                   const imresh_params *params, void *stream,
                                   imresh_result *results);

/* ---- reconstruction server (no reference counterpart: the reference's
 * task queue, taskQueue.cu:91-193, runs each request on a fresh host thread
 * and stream slot with no device state kept between calls) -------------- */
typedef struct imresh_server imresh_server;

typedef struct imresh_server_completion {
    uint64_t      id;           /* from imresh_server_submit               */
    int32_t       status;       /* IMRESH_OK or an error code              */
    imresh_result result;
    double        latency_ms;   /* submit -> completion observed           */
} imresh_server_completion;

typedef struct imresh_server_info {
    uint64_t submitted, completed, queued, active;
    uint64_t bytes_in_use;      /* device bytes held (active + cached)      */
    uint64_t peak_bytes;
    uint64_t budget;
} imresh_server_info;

/* Called on the server's scheduler thread when a request completes; io is
 * the request's buffer (Re(g') written).  Must not block for long.  It may
 * submit, poll and read stats; imresh_server_drain and imresh_server_destroy
 * of the same server return IMRESH_ERR_INVALID there (they would wait for
 * the thread running the callback). */
typedef void (*imresh_server_callback)(void *user, const imresh_server_completion *c, float *io);

/* Continuous batching on the current device: requests of any supported
 * size are admitted in arrival order while fewer than max_active run and
 * their device state fits hbm_budget_bytes (0: 200 GiB; max_active 0: 16;
 * lanes = HIP streams, 0: 16); every admitted request advances cycle by cycle
 * with its state resident on the device, interleaved with the others.  Each
 * result equals imresh_shrinkwrap on the same input. */
int imresh_server_create(uint64_t hbm_budget_bytes, uint32_t max_active, uint32_t lanes,
                         imresh_server **server);
/* Queues one reconstruction of io (|F| in, Re(g') out; device memory when
 * device_io != 0, else host).  The buffer must stay valid until completion.
 * IMRESH_ERR_BUDGET when the request alone exceeds the budget. */
int imresh_server_submit(imresh_server *server, float *io, uint32_t width, uint32_t height,
                         const imresh_params *params, int device_io,
                         imresh_server_callback callback, void *user, uint64_t *id);
/* Moves up to capacity completion records (oldest first) into out. */
int imresh_server_poll(imresh_server *server, imresh_server_completion *out, uint32_t capacity,
                       uint32_t *count);
/* Blocks until every submitted request has completed. */
int imresh_server_drain(imresh_server *server);
int imresh_server_stats(imresh_server *server, imresh_server_info *info);
/* Drains, stops the scheduler and frees the device state. */
int imresh_server_destroy(imresh_server *server);

/* Standalone HIO with a given support (libs::hybridInputOutput,
 * src/imresh/libs/hybridInputOutput.cpp:200-306): io = |F| in, Re(g') out;
 * mask uint8 (1 = outside the support, :267); the first guess is |F| with
 * the phases of the FFT of glibc rand()/RAND_MAX after srand(2623091912)
 * (:126-197, like the reference this reseeds the process's rand());
 * n_cycles 0 -> unbounded (UINT_MAX), hio_beta <= 0 -> 0.9; with
 * target_error > 0 the HIO error is checked after every inverse transform
 * and the loop stops below it (n_cycles 0 with no target: IMRESH_ERR_INVALID
 * instead of the reference's endless loop).  result->cycles_run = HIO
 * iterations run.  The modulus projection maps |G'| = 0 to 0 (the reference
 * divides by zero there).  Synchronous. */
int imresh_hybrid_input_output(float *io, const uint8_t *mask, uint32_t width, uint32_t height,
                               uint32_t n_cycles, float target_error, float hio_beta, void *stream,
                               imresh_result *result);
int imresh_hybrid_input_output_device(float *io_dev, const uint8_t *mask_dev, uint32_t width,
                                      uint32_t height, uint32_t n_cycles, float target_error,
                                      float hio_beta, void *stream, imresh_result *result);

/* HIO inner loop from a given device state, n_iter iterations
 * (cudaShrinkWrap.cu:311-365 / shrinkWrap.cpp:260-319): gprime (complex, g'),
 * gprev (complex), mask (uint8, 1 = outside support), modulus (float).
 * Runs the fused row/column kernels the reconstruction uses. */
int imresh_hio_iterations(float *gprime_dev, float *gprev_dev, const uint8_t *mask_dev,
                          const float *modulus_dev, uint32_t width, uint32_t height,
                          float hio_beta, uint32_t n_iter, void *stream);

/* ---- per-operator entry points (device pointers, asynchronous) --------- */
/* cudaKernelCopyToRealPart, cudaVectorElementwise.cu:75-95 */
int imresh_copy_to_real_part(float *cplx_dev, const float *real_dev, uint32_t n, void *stream);
/* cudaKernelCopyFromRealPart, cudaVectorElementwise.cu:98-117 */
int imresh_copy_from_real_part(float *real_dev, const float *cplx_dev, uint32_t n, void *stream);
/* cudaKernelComplexNormElementwise, cudaVectorElementwise.cu:120-141;
 * complexNormElementwise, vectorElementwise.cpp:41-56 */
int imresh_complex_norm(float *out_dev, const float *cplx_dev, uint32_t n, void *stream);
/* cudaKernelApplyComplexModulus, cudaVectorElementwise.cu:168-195;
 * applyComplexModulus, vectorElementwise.cpp:59-81.  out may equal in. */
int imresh_apply_complex_modulus(float *out_dev, const float *in_dev, const float *modulus_dev,
                                 uint32_t n, void *stream);
/* cudaKernelApplyHioDomainConstraints, cudaVectorElementwise.cu:42-72 (float mask) */
int imresh_hio_domain_constraints(float *gprev_dev, const float *gprime_dev, const float *mask_dev,
                                  uint32_t n, float hio_beta, void *stream);
/* cudaKernelCutOff, cudaVectorElementwise.cu:198-219 */
int imresh_cutoff(float *data_dev, uint32_t n, float threshold, float lower, float upper,
                  void *stream);
/* cudaVectorMax/Min/Sum, cudaVectorReduce.cu:216-252 (host result, synchronous) */
int imresh_vector_max(const float *data_dev, uint32_t n, float *result_host, void *stream);
int imresh_vector_min(const float *data_dev, uint32_t n, float *result_host, void *stream);
int imresh_vector_sum(const float *data_dev, uint32_t n, float *result_host, void *stream);
/* cudaCalculateHioError, cudaVectorReduce.cu:357-403 (float mask) and the
 * unsigned char mask instantiation (:480-503).  Returns sqrtf(total)/nmasked
 * in *error_host; total / nmasked may be NULL.  Synchronous. */
int imresh_calculate_hio_error(const float *gprime_dev, const float *mask_dev, uint32_t n,
                               int invert_mask, float *error_host, float *total_host,
                               float *nmasked_host, void *stream);
int imresh_calculate_hio_error_u8(const float *gprime_dev, const uint8_t *mask_dev, uint32_t n,
                                  int invert_mask, float *error_host, float *total_host,
                                  float *nmasked_host, void *stream);
/* The same with a bit-packed mask (cudaCalculateHioErrorBitPacked,
 * benchmark/imresh/algorithms/cuda/cudaVectorReduce.cu:402-505): 32 elements
 * per uint32 word, element i at bit 31 - i%32 of word i/32. */
int imresh_calculate_hio_error_bits(const float *gprime_dev, const uint32_t *mask_bits_dev, uint32_t n,
                                    int invert_mask, float *error_host, float *total_host,
                                    float *nmasked_host, void *stream);
/* Byte mask (nonzero = 1) -> that bit-packed layout, ceil(n/32) words, a
 * partial last word zero-filled.  Asynchronous. */
int imresh_pack_mask(const uint8_t *mask_dev, uint32_t n, uint32_t *bits_dev, void *stream);
/* calcGaussianKernel, calcGaussianKernel.cpp:42-91 (host).  Returns the tap
 * count; writes the weights when weights != NULL and capacity suffices. */
int imresh_gaussian_kernel(double sigma, float *weights_host, uint32_t capacity);
/* cudaGaussianBlur / gaussianBlur (cudaGaussian.cu:1044-1057, gaussian.cpp:737-749),
 * the horizontal pass (cudaGaussian.cu:655-726) and the vertical pass
 * (cudaGaussian.cu:993-1041); in place, clamp-to-edge, up to 63 taps.
 * Each sigma's weights go into a per-device table of 128 slots on its first
 * use (as GaussianKernelGpuBuffer, cudaGaussian.cu:64-298); that first call
 * synchronizes the stream, later calls with the sigma are asynchronous.
 * With all slots taken, calls with a new sigma synchronize (no abort). */
int imresh_gaussian_blur(float *data_dev, uint32_t width, uint32_t height, double sigma,
                         void *stream);
int imresh_gaussian_blur_horizontal(float *data_dev, uint32_t width, uint32_t height,
                                    double sigma, void *stream);
int imresh_gaussian_blur_vertical(float *data_dev, uint32_t width, uint32_t height,
                                  double sigma, void *stream);
/* cufftExecC2C on a cufftPlan2d(H, W, C2C) plan (cudaShrinkWrap.cu:225-232,319,343):
 * unnormalized 2-D transform; direction -1 = FORWARD, +1 = INVERSE/BACKWARD.
 * out may equal in. */
int imresh_fft2d(float *out_dev, const float *in_dev, uint32_t width, uint32_t height,
                 int direction, void *stream);
/* Forward model (src/imresh/libs/diffractionIntensity.cpp:49-80, CUDA
 * :84-135): io (a width x height real object, row-major) is replaced in place
 * by |FFT2((x, 0))|, the unnormalized forward transform's modulus with
 * frequency 0 at [0, 0], i.e. the modulus input of imresh_shrinkwrap.
 * imresh_diffraction_intensity takes host memory and synchronizes the stream
 * unless async != 0 (then io is valid once the stream has finished);
 * the _device variant takes device memory and is asynchronous. */
int imresh_diffraction_intensity(float *io, uint32_t width, uint32_t height, void *stream, int async);
int imresh_diffraction_intensity_device(float *io_dev, uint32_t width, uint32_t height, void *stream);
/* Measurement simulation (no reference counterpart; the reference reads
 * recorded or host-computed intensities): a device modulus |F| (e.g. from
 * imresh_diffraction_intensity_device) is replaced in place by
 * sqrt(k), k ~ Poisson(photons |F|^2 / sum |F|^2) per element, i.e. the
 * modulus of a measurement that counted `photons` photons in expectation
 * (0 < photons <= 1e30).
 * Deterministic for a given seed (counter-based Philox4x32-10 stream per
 * element; the sum in a fixed order), independent of the launch shape.
 * Asynchronous. */
int imresh_poisson_modulus(float *modulus_dev, uint32_t width, uint32_t height, double photons,
                           uint64_t seed, void *stream);
/* 1 if this line length has a transform (1 ... 8192: a compiled plan or
 * Bluestein), else 0. */
int imresh_fft_supported(uint32_t n);
/* fftShift (src/imresh/libs/fftShift.cpp:42-61) in place on a device
 * nx x ny array of element_bytes = 4 (float) or 8 (complex float) elements:
 * (iy, ix) for iy < ny/2 swaps with ((iy+ny/2)%ny, (ix+nx/2)%nx). */
int imresh_fft_shift(void *data_dev, uint32_t nx, uint32_t ny, uint32_t element_bytes, void *stream);
/* Picks (once per device and size, by timing) the row- and column-pass
 * kernel variants used for W x H and reports them.  Called implicitly by the
 * first reconstruction of a size; no reference counterpart (cuFFT planned
 * internally, cudaShrinkWrap.cu:226). */
int imresh_autotune(uint32_t width, uint32_t height, void *stream, int *row_variant,
                    int *col_variant);
/* Number of row- and column-pass kernel variants of W x H (every variant
 * computes the same bits; they differ in tiling and scheduling). */
int imresh_variant_counts(uint32_t width, uint32_t height, int *row_variants, int *col_variants);
/* Pins the variants used for W x H on the current device instead of timing
 * them (a known tuning, or a test of one variant).  No reference
 * counterpart. */
int imresh_set_variants(uint32_t width, uint32_t height, int row_variant, int col_variant);

/* ---- file formats (readInFuncs.cpp, writeOutFuncs.cpp; host only) -------- */
/* Readers: with out == NULL only *width / *height are set (size query);
 * otherwise out receives width*height floats (capacity in floats, too small:
 * IMRESH_ERR_INVALID).  IMRESH_ERR_IO: unreadable file or unsupported layout.
 * PNG (readPNG, readInFuncs.cpp:108-157): red channel as 16 bits (8-bit
 * samples x 256) / 65535, buffer row iy = file row height-1-iy (pngwriter's
 * bottom-left origin); 1/2/4/8/16-bit gray, palette, RGB(A), gray+alpha,
 * non-interlaced.  TIFF (readTiff, :160-195 / ReadTiff.cpp): 32-bit IEEE
 * float, one sample per pixel, uncompressed strips, either byte order.
 * Text (readTxt, :56-104): whitespace-separated floats, width = values on
 * the first line, height = count / width. */
int imresh_read_png(const char *path, float *out, size_t capacity, uint32_t *width, uint32_t *height);
int imresh_read_tiff(const char *path, float *out, size_t capacity, uint32_t *width, uint32_t *height);
int imresh_read_txt(const char *path, float *out, size_t capacity, uint32_t *width, uint32_t *height);
/* writeOutPNG (writeOutFuncs.cpp:69-114): 16-bit RGB, R = G = B =
 * int(v/max*65535 + 0.5f) clamped to [0, 65535] (max 0 -> 1), buffer row iy
 * = file row height-1-iy. */
int imresh_write_png(const char *path, const float *data, uint32_t width, uint32_t height);

/* ---- task queue (taskQueue.hpp:38-84, taskQueue.cu:91-251) --------------- */
/* Completion callback: (user, io_host, width, height, status). */
typedef void (*imresh_task_callback)(void *user, float *io_host, uint32_t width,
                                     uint32_t height, int status);
/* taskQueueInit, taskQueue.cu:231-234 (streams per device: 0 = default) */
int imresh_task_queue_init(int streams_per_device);
/* taskQueueDeinit, taskQueue.cu:236-251: waits for all tasks; producers
 * still blocked in imresh_add_task return IMRESH_ERR_NOT_INITIALIZED.
 * Called from a task callback (a worker thread, which it would have to
 * join) it returns IMRESH_ERR_INVALID and leaves the queue running. */
int imresh_task_queue_deinit(void);
/* addTask, taskQueue.cu:147-193: blocks while all slots are busy */
int imresh_add_task(float *io_host, uint32_t width, uint32_t height,
                    const imresh_params *params, imresh_task_callback callback, void *user);

/* ---- measurement ------------------------------------------------------- */
/* Average device duration (ms, HIP events on `stream`) of each fused kernel
 * of one HIO iteration at W x H, over `reps` back-to-back launches on a
 * resident synthetic state.  out_ms[0] = row pass (inverse+constraint+
 * forward), out_ms[1] = column pass (forward+modulus+inverse),
 * out_ms[2] = blur+max, out_ms[3] = row pass closing a cycle (norm+error). */
int imresh_profile_kernels(uint32_t width, uint32_t height, uint32_t reps, float *out_ms,
                           void *stream);
/* Same for the two kernels of imresh_diffraction_intensity_device:
 * out_ms[0] = column pass (real plane -> forward columns), out_ms[1] = row
 * pass (forward rows -> |z|). */
int imresh_profile_forward_model(uint32_t width, uint32_t height, uint32_t reps, float *out_ms,
                                 void *stream);

#ifdef __cplusplus
}
#endif
#endif
