// Cycle-boundary and standalone kernels of the shrink-wrap / HIO path for
// MI355X (gfx950): the fused 2-D Gaussian blur with per-tile maximum, the
// max / error finalizers that keep the threshold and the early-exit flag on
// the device, wave64 reductions (max/min/sum, HIO error) and the reference's
// elementwise operators as standalone kernels.  Blur taps are summed from 0
// in the reference order w[0]*x[i-r] ... w[2r]*x[i+r] (gaussian.cpp:51-163,
// 559-734; cudaGaussian.cu:539-652, 776-990) with clamp-to-edge borders.
#include <algorithm>
#include <cfloat>
#include <utility>

#include "internal.hpp"
#include "ops_device.hpp"

namespace imr {

// ------------------------------------------------------------ Gaussian blur
// Out-of-place separable blur of a kBlurTX x kBlurTY output tile for a tap
// radius R known at compile time (every weight stays in registers):
//   1. the input tile with a clamped R-wide halo is staged in LDS;
//   2. horizontal pass in place in LDS: 16 lanes per row, each reading a
//      (4 + 2R)-wide window with 16-byte LDS reads and writing 4 outputs;
//      a row is owned by one wave, so its reads precede its writes;
//   3. vertical pass: each lane keeps a (16 + 2R)-long column window in
//      registers and writes 16 outputs, coalesced along the rows;
//   4. optional tile maximum of the outputs (wave64 DPP reduction).
// Each output is summed in the reference tap order from 0.  The identity
// flags turn a pass into a copy (the one-directional blur entry points).
template <int R>
__global__ void __launch_bounds__(256) k_blur2d(BlurArgs a)
{
    // XCD-aware tile order (a.xcd_tiles): workgroups are dealt round-robin
    // to the 8 XCDs (each with its own L2) in the order x + gx (y + gy z), so
    // row-major tiles put every horizontal neighbour on another XCD and fetch
    // the 128-byte lines their halos share twice; with xcd_tiles each XCD
    // takes a contiguous range of the grid's tiles instead (launch_blur sets
    // it where the planes stay in the Infinity Cache)
    int tx = blockIdx.x, ty = blockIdx.y, b = blockIdx.z;   // b: image of a batch
    if (a.xcd_tiles) {
        const int gx = gridDim.x, gy = gridDim.y, nt = gx * gy * (int) gridDim.z;
        const int lin = tx + gx * (ty + gy * b), x8 = lin & 7;
        const int k = x8 * (nt >> 3) + min(x8, nt & 7) + (lin >> 3);
        tx = k % gx;
        ty = (k / gx) % gy;
        b = k / (gx * gy);
    }
    if (a.done != nullptr && a.done[b * kFlagStride]) return;
    constexpr int TX = kBlurTX, TY = kBlurTY, NT = 2 * R + 1;
    constexpr int AW = TX + 2 * R, AH = TY + 2 * R, AWP = (AW + 3) & ~3;
    constexpr int NV = (4 + 2 * R + 3) / 4;   // 16-byte reads per row window
    constexpr int RPT = TY / 4;               // output rows per lane
    static_assert(TX == 64 && TY % 4 == 0, "tile shape");
    extern __shared__ float4 blur_sm[];
    float *A = reinterpret_cast<float *>(blur_sm);
    __shared__ float red[4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int x0 = tx * TX, y0 = ty * TY, W = a.W, H = a.H;
    const size_t wrow = a.cycles != nullptr ? (size_t) kBlurWeightRow * a.cycles[b * kFlagStride] : 0;
    const float *in = a.in + (size_t) b * a.istride;
    float *out = a.out + (size_t) b * a.istride;
    float wh[NT], wv[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        wh[k] = a.wh[wrow + k];
        wv[k] = a.wv[wrow + k];
    }
    {
        // all global loads of the staged tile are issued before the LDS
        // stores; rows/columns past the tile are clamped reads, not stored
        constexpr int NRW = (AH + 3) / 4, NCX = (AW + 63) / 64;
        float buf[NRW][NCX];
#pragma unroll
        for (int i = 0; i < NRW; ++i) {
            const int gy = min(max(y0 - R + wave + 4 * i, 0), H - 1);
            const float *src = in + (size_t) gy * W;
#pragma unroll
            for (int j = 0; j < NCX; ++j)
                buf[i][j] = src[min(max(x0 - R + lane + 64 * j, 0), W - 1)];
        }
#pragma unroll
        for (int i = 0; i < NRW; ++i) {
            const int ry = wave + 4 * i;
#pragma unroll
            for (int j = 0; j < NCX; ++j) {
                const int cx = lane + 64 * j;
                if (ry < AH && cx < AW) A[ry * AWP + cx] = buf[i][j];
            }
        }
    }
    __syncthreads();
    {
        const int q = 4 * (t & 15);
        for (int ry = t >> 4; ry < AH; ry += 16) {
            float *row = A + ry * AWP;
            float o[4];
            if (a.h_identity) {
#pragma unroll
                for (int k = 0; k < 4; ++k) o[k] = row[q + R + k];
            } else {
                float v[4 * NV];
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    const float4 f = *reinterpret_cast<const float4 *>(row + q + 4 * i);
                    v[4 * i] = f.x;
                    v[4 * i + 1] = f.y;
                    v[4 * i + 2] = f.z;
                    v[4 * i + 3] = f.w;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    float sum = 0.f;
#pragma unroll
                    for (int j = 0; j < NT; ++j) sum += v[k + j] * wh[j];
                    o[k] = sum;
                }
            }
            *reinterpret_cast<float4 *>(row + q) = make_float4(o[0], o[1], o[2], o[3]);
        }
    }
    __syncthreads();
    float lmax = -FLT_MAX;
    {
        const int r0 = wave * RPT, gx = x0 + lane;
        float v[RPT + 2 * R];
#pragma unroll
        for (int j = 0; j < RPT + 2 * R; ++j) v[j] = A[(r0 + j) * AWP + lane];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            float sum;
            if (a.v_identity) {
                sum = v[k + R];
            } else {
                sum = 0.f;
#pragma unroll
                for (int j = 0; j < NT; ++j) sum += v[k + j] * wv[j];
            }
            const int gy = y0 + r0 + k;
            if (gx < W && gy < H) {
                out[(size_t) gy * W + gx] = sum;
                lmax = fmax_ref(lmax, sum);
            }
        }
    }
    if (a.pmax != nullptr) {
        lmax = wave_reduce(lmax, [](float m, float x) { return fmax_ref(m, x); });
        if (lane == 0) red[wave] = lmax;
        __syncthreads();
        if (t == 0)
            a.pmax[((size_t) b * gridDim.y + ty) * gridDim.x + tx] =
                fmax_ref(fmax_ref(red[0], red[1]), fmax_ref(red[2], red[3]));
    }
}

// In-place horizontal blur, one row per workgroup, row staged in LDS
// (cudaGaussian.cu:539-652 processes a row per block as well).
__global__ void __launch_bounds__(256) k_blur_h_inplace(float *d, int W, const float *w, int r)
{
    extern __shared__ float sm[];
    float *ws = sm;
    float *row = sm + 64;
    float *p = d + (size_t) blockIdx.x * W;
    for (int i = threadIdx.x; i < 2 * r + 1; i += blockDim.x) ws[i] = w[i];
    for (int i = threadIdx.x; i < W; i += blockDim.x) row[i] = p[i];
    __syncthreads();
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        float sum = 0.f;
        for (int k = 0; k <= 2 * r; ++k) {
            const int j = min(max(i - r + k, 0), W - 1);
            sum += row[j] * ws[k];
        }
        p[i] = sum;
    }
}

// ------------------------------------------------------------ finalizers
// threshold = cut * max over the blur tiles (cudaShrinkWrap.cu:274-276)
__global__ void __launch_bounds__(1024) k_max_finalize(const float *pmax, int np, float cut,
                                                      float *thr, const int *done)
{
    if (done != nullptr && *done) return;
    __shared__ float red[1024];
    float m = -FLT_MAX;
    for (int i = threadIdx.x; i < np; i += blockDim.x) m = fmax_ref(m, pmax[i]);
    m = block_reduce(m, red, [](float x, float y) { return fmax_ref(x, y); });
    if (threadIdx.x == 0) {
        float t = m;
        t *= cut;
        *thr = t;
    }
}

// batched form: workgroup b finalizes image b; the cut-off factor is cutA
// in the image's first cycle and cut afterwards (cudaShrinkWrap.cu:275-276)
__global__ void __launch_bounds__(1024) k_max_finalize_batch(const float *pmax, int np, float cutA,
                                                           float cut, float *scal, const int *flags)
{
    const int b = blockIdx.x;
    if (flags[b * kFlagStride]) return;
    __shared__ float red[1024];
    float m = -FLT_MAX;
    for (int i = threadIdx.x; i < np; i += blockDim.x) m = fmax_ref(m, pmax[(size_t) b * np + i]);
    m = block_reduce(m, red, [](float x, float y) { return fmax_ref(x, y); });
    if (threadIdx.x == 0) {
        float t = m;
        t *= flags[b * kFlagStride + 1] == 0 ? cutA : cut;
        scal[b * kScalStride] = t;
    }
}

// error = sqrtf(sum mask*|g'|) / sum mask (hybridInputOutput.cpp:93-102,
// cudaVectorReduce.cu:402); early-exit flag as shrinkWrap.cpp:328-329.
// res[0] = error, res[1] = total, res[2] = nmasked.
__global__ void __launch_bounds__(1024) k_err_finalize(const float *perr, const unsigned *pcnt,
                                                      int np, float target, float *res, int *done,
                                                      unsigned *cycles)
{
    if (done != nullptr && *done) return;
    __shared__ float red_f[1024];
    __shared__ unsigned red_u[1024];
    float s = 0.f;
    unsigned c = 0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        s += perr[i];
        c += pcnt[i];
    }
    s = block_reduce(s, red_f, [](float x, float y) { return x + y; });
    c = block_reduce(c, red_u, [](unsigned x, unsigned y) { return x + y; });
    if (threadIdx.x == 0) {
        const float nm = (float) c;
        const float err = sqrtf(s) / nm;
        res[0] = err;
        res[1] = s;
        res[2] = nm;
        if (cycles != nullptr) *cycles += 1;
        if (done != nullptr && target > 0 && err < target) *done = 1;
    }
}

// ------------------------------------------------------------ elementwise
__global__ void k_copy_from_real(float *out, const float2 *in, size_t n)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        out[i] = in[i].x;
}

__global__ void k_copy_to_real(float2 *out, const float *in, size_t n)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        out[i] = make_float2(in[i], 0.f);
}

__global__ void k_complex_norm(float *out, const float2 *in, size_t n)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        out[i] = cnorm(((const cf *) in)[i]);
}

__global__ void k_apply_modulus(float2 *out, const float2 *in, const float *M, size_t n)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        ((cf *) out)[i] = modulus_project(((const cf *) in)[i], M[i]);
}

__global__ void k_hio_constraint(float2 *gp, const float2 *g, const float *mask, size_t n,
                                 float beta)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        ((cf *) gp)[i] = hio_constraint(((cf *) gp)[i], ((const cf *) g)[i], mask[i] == 1, beta);
}

__global__ void k_cutoff(float *d, size_t n, float t, float lo, float hi)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        d[i] = d[i] < t ? lo : hi;
}

// ------------------------------------------------------------ reductions
template <int OP>
__device__ __forceinline__ float red_op(float a, float b)
{
    if constexpr (OP == 0) return fmax_ref(a, b);
    else if constexpr (OP == 1) return fmin_ref(a, b);
    else return a + b;
}
template <int OP> __device__ __forceinline__ float red_init()
{
    if constexpr (OP == 0) return -FLT_MAX;   // numeric_limits<float>::lowest()
    else if constexpr (OP == 1) return FLT_MAX;
    else return 0.f;
}

template <int OP>
__global__ void __launch_bounds__(256) k_reduce(const float *d, size_t n, float *partials)
{
    __shared__ float red[1024];
    float acc = red_init<OP>();
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        acc = red_op<OP>(acc, d[i]);
    acc = block_reduce(acc, red, [](float x, float y) { return red_op<OP>(x, y); });
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

template <int OP>
__global__ void __launch_bounds__(1024) k_reduce_final(const float *partials, int np, float *result)
{
    __shared__ float red[1024];
    float acc = red_init<OP>();
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc = red_op<OP>(acc, partials[i]);
    acc = block_reduce(acc, red, [](float x, float y) { return red_op<OP>(x, y); });
    if (threadIdx.x == 0) *result = acc;
}

// mask element i: a float or unsigned char per element, or bit-packed
// (MaskBits: 32 elements per word, element i at bit 31 - i % 32 of word
// i / 32, the layout of cudaKernelCalculateHioErrorBitPacked,
// benchmark/imresh/algorithms/cuda/cudaVectorReduce.cu:402-458)
struct MaskBits {
    uint32_t w;
};
template <class MT>
__device__ __forceinline__ bool mask_at(const MT *m, size_t i) { return m[i] != 0; }
template <>
__device__ __forceinline__ bool mask_at<MaskBits>(const MaskBits *m, size_t i)
{
    return ((m[i >> 5].w >> (31u - (unsigned) (i & 31))) & 1u) != 0;
}

// HIO error partial sums with a float, unsigned char or bit-packed mask
// (cudaVectorReduce.cu:284-355: shouldBeZero = invert xor (bool) mask)
template <class MT>
__global__ void __launch_bounds__(256) k_hio_error(const float2 *g, const MT *mask, size_t n,
                                                   int invert, float *perr, unsigned *pcnt)
{
    __shared__ float red_f[1024];
    __shared__ unsigned red_u[1024];
    float s = 0.f;
    unsigned c = 0;
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x) {
        const bool z = (invert != 0) != mask_at(mask, i);
        if (z) {
            s += cnorm(((const cf *) g)[i]);
            c += 1;
        }
    }
    s = block_reduce(s, red_f, [](float x, float y) { return x + y; });
    c = block_reduce(c, red_u, [](unsigned x, unsigned y) { return x + y; });
    if (threadIdx.x == 0) {
        perr[blockIdx.x] = s;
        pcnt[blockIdx.x] = c;
    }
}

// deterministic synthetic state for kernel timing (no host round trip)
__global__ void k_fill_state(float2 *X, float2 *G, uint8_t *mask, float *M, float *Bn, size_t n,
                             int W)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x) {
        const unsigned h = (unsigned) (i * 2654435761u) ^ (unsigned) (i >> 7);
        const float u = (float) (h & 0xffff) * (1.f / 65536.f);
        const int x = (int) (i % (size_t) W), y = (int) (i / (size_t) W);
        X[i] = make_float2(u - 0.25f, 0.5f - u);
        G[i] = make_float2(0.5f * u, -0.25f * u);
        mask[i] = ((x ^ y) & 4) ? 1 : 0;
        M[i] = 1.f + u;
        Bn[i] = u;
    }
}

// up to kStoreWords host values carried in the kernel arguments (copied at
// launch, so the host array may be reused at once) stored to device memory:
// uploads that need no staging allocation and can be captured in a graph
struct StoreWords {
    float v[kStoreWords];
};

__global__ void k_store_words(float *dst, unsigned n, StoreWords a)
{
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x) dst[i] = a.v[i];
}

// ------------------------------------------------------------ launchers
static int grid_1d(size_t n)
{
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (int) g;
}

int blur_grid(int W, int H)
{
    return ((W + kBlurTX - 1) / kBlurTX) * ((H + kBlurTY - 1) / kBlurTY);
}

template <int R>
static int launch_blur_r(const BlurArgs &a, hipStream_t s)
{
    constexpr int AWP = (kBlurTX + 2 * R + 3) & ~3, AH = kBlurTY + 2 * R;
    constexpr size_t lds = sizeof(float) * (size_t) AWP * AH;
    if constexpr (lds > 64 * 1024) {
        static const hipError_t attr = hipFuncSetAttribute(
            (const void *) k_blur2d<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
        if (attr != hipSuccess) return from_hip(attr);
    }
    const dim3 grid((a.W + kBlurTX - 1) / kBlurTX, (a.H + kBlurTY - 1) / kBlurTY,
                    a.batch > 1 ? a.batch : 1);
    IMR_LAUNCH(k_blur2d<R>, grid, dim3(256), lds, s, a);
    return imr::launch_status(__FILE__, __LINE__);
}

template <int R>
static int prepare_blur_r()
{
    constexpr int AWP = (kBlurTX + 2 * R + 3) & ~3, AH = kBlurTY + 2 * R;
    constexpr size_t lds = sizeof(float) * (size_t) AWP * AH;
    if constexpr (lds > 64 * 1024) {
        static const hipError_t attr = hipFuncSetAttribute(
            (const void *) k_blur2d<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
        return from_hip(attr);
    }
    return IMRESH_OK;
}

template <int... Rs>
static int prepare_blur_dispatch(int r, std::integer_sequence<int, Rs...>)
{
    int st = IMRESH_ERR_INVALID;
    (void) ((r == Rs ? (st = prepare_blur_r<Rs>(), true) : false) || ...);
    return st;
}

int prepare_blur(int R)
{
    if (R < 0 || R > kMaxBlurRadius) return IMRESH_ERR_INVALID;
    return prepare_blur_dispatch(R, std::make_integer_sequence<int, kMaxBlurRadius + 1>{});
}

template <int... Rs>
static int launch_blur_dispatch(int r, const BlurArgs &a, hipStream_t s,
                                std::integer_sequence<int, Rs...>)
{
    int st = IMRESH_ERR_INVALID;
    (void) ((r == Rs ? (st = launch_blur_r<Rs>(a, s), true) : false) || ...);
    return st;
}

// one tap radius for both directions: rh == rv, or the identity direction's
// radius is 0 (the halo of the copied direction is then still staged)
int launch_blur(const BlurArgs &a, hipStream_t s)
{
    const int r = a.h_identity ? a.rv : a.v_identity ? a.rh : a.rh;
    if (!a.h_identity && !a.v_identity && a.rh != a.rv) return IMRESH_ERR_INVALID;
    if (r < 0 || r > kMaxBlurRadius || a.W < 1 || a.H < 1) return IMRESH_ERR_INVALID;
    // XCD-contiguous tiles while the launch's planes stay in the Infinity
    // Cache (measured: 1024^2 8.4 -> 7.8 us, 2048^2 17.2 -> 14.8 us, 4096^2
    // 44.8 -> 43.5 us); from DRAM the row-major order keeps the XCDs reading
    // neighbouring rows and is faster (8192^2: 155 against 166 us)
    BlurArgs b = a;
    const size_t elems = (size_t) a.W * a.H * (size_t) (a.batch > 1 ? a.batch : 1);
    b.xcd_tiles = elems <= ((size_t) 16 << 20);
    return launch_blur_dispatch(r, b, s, std::make_integer_sequence<int, kMaxBlurRadius + 1>{});
}

int launch_blur_h_inplace(float *d, int W, int H, const float *w, int r, hipStream_t s)
{
    const size_t lds = sizeof(float) * (64 + (size_t) W);
    static const hipError_t attr = hipFuncSetAttribute(
        (const void *) k_blur_h_inplace, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
    if (attr != hipSuccess) return from_hip(attr);
    hipLaunchKernelGGL(k_blur_h_inplace, dim3(H), dim3(256), lds, s, d, W, w, r);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_max_finalize(const float *pmax, int np, float cut, float *thr, const int *done,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_max_finalize, dim3(1), dim3(1024), 0, s, pmax, np, cut, thr, done);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_err_finalize(const float *perr, const unsigned *pcnt, int np, float target, float *res,
                        int *done, unsigned *cycles, hipStream_t s)
{
    hipLaunchKernelGGL(k_err_finalize, dim3(1), dim3(1024), 0, s, perr, pcnt, np, target, res,
                       done, cycles);
    return imr::launch_status(__FILE__, __LINE__);
}

// batched error finalizer: workgroup b sums the np row-pass partials of
// image b (in the order of k_err_finalize, so a batch reports the same
// error as the single-image path), writes {error, total, masked} after its
// threshold slot, counts the cycle and raises its done flag below the target
__global__ void __launch_bounds__(1024) k_err_finalize_batch(const float *perr, const unsigned *pcnt,
                                                           int np, float target, float *scal,
                                                           int *flags)
{
    const int b = blockIdx.x;
    if (flags[b * kFlagStride]) return;
    __shared__ float red_f[1024];
    __shared__ unsigned red_u[1024];
    float sum = 0.f;
    unsigned c = 0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        sum += perr[(size_t) b * np + i];
        c += pcnt[(size_t) b * np + i];
    }
    sum = block_reduce(sum, red_f, [](float x, float y) { return x + y; });
    c = block_reduce(c, red_u, [](unsigned x, unsigned y) { return x + y; });
    if (threadIdx.x == 0) {
        const float nm = (float) c;
        const float err = sqrtf(sum) / nm;
        float *res = scal + b * kScalStride + 1;
        res[0] = err;
        res[1] = sum;
        res[2] = nm;
        flags[b * kFlagStride + 1] += 1;
        if (target > 0 && err < target) flags[b * kFlagStride] = 1;
    }
}

int launch_max_finalize_batch(const float *pmax, int np, float cutA, float cut, float *scal,
                              int *flags, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(k_max_finalize_batch, dim3(batch), dim3(1024), 0, s, pmax, np, cutA, cut, scal,
                       flags);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_err_finalize_batch(const float *perr, const unsigned *pcnt, int np, float target,
                              float *scal, int *flags, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(k_err_finalize_batch, dim3(batch), dim3(1024), 0, s, perr, pcnt, np, target,
                       scal, flags);
    return imr::launch_status(__FILE__, __LINE__);
}

// addRandomPhase (hybridInputOutput.cpp:186-195): X holds the forward FFT of
// the random real field; X = |F| * (cos, sin)(arg X), the phase a float
// atan2 and cos/sin evaluated in double as the reference's unqualified cos
// of a float
__global__ void __launch_bounds__(256) k_random_phase(float2 *X, const float *M, size_t n)
{
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x) {
        const float2 z = X[i];
        const float phase = atan2f(z.y, z.x);
        const double m = (double) M[i];
        X[i] = make_float2((float) (m * cos((double) phase)), (float) (m * sin((double) phase)));
    }
}

int launch_random_phase(float2 *X, const float *M, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_random_phase, dim3(grid_1d(n)), dim3(256), 0, s, X, M, n);
    return imr::launch_status(__FILE__, __LINE__);
}

// fftShift (fftShift.cpp:42-61): element (iy, ix), iy < ny/2, swaps with
// ((iy + ny/2) % ny, (ix + nx/2) % nx); the pairs are disjoint, one thread each
template <class E>
__global__ void __launch_bounds__(256) k_fft_shift(E *d, unsigned nx, unsigned ny)
{
    const size_t half = (size_t) (ny / 2) * nx;
    for (size_t k = blockIdx.x * (size_t) blockDim.x + threadIdx.x; k < half;
         k += (size_t) gridDim.x * blockDim.x) {
        const unsigned iy = (unsigned) (k / nx), ix = (unsigned) (k - (size_t) iy * nx);
        const size_t b = (size_t) ((iy + ny / 2) % ny) * nx + (ix + nx / 2) % nx;
        const E t = d[k];
        d[k] = d[b];
        d[b] = t;
    }
}

int launch_fft_shift(void *d, unsigned nx, unsigned ny, unsigned elem_bytes, hipStream_t s)
{
    const size_t half = (size_t) (ny / 2) * nx;
    if (elem_bytes != 4 && elem_bytes != 8) return IMRESH_ERR_INVALID;
    if (half == 0) return IMRESH_OK;
    if (elem_bytes == 4)
        hipLaunchKernelGGL(k_fft_shift<float>, dim3(grid_1d(half)), dim3(256), 0, s, (float *) d, nx, ny);
    else if (elem_bytes == 8)
        hipLaunchKernelGGL(k_fft_shift<float2>, dim3(grid_1d(half)), dim3(256), 0, s, (float2 *) d, nx, ny);
    else
        return IMRESH_ERR_INVALID;
    return imr::launch_status(__FILE__, __LINE__);
}

// batch staging: image b of a pointer table <-> plane b of a batch buffer
__global__ void __launch_bounds__(256) k_gather_images(float *dst, const float *const *src, size_t n)
{
    const int b = blockIdx.y;
    const float *p = src[b];
    float *q = dst + (size_t) b * n;
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        q[i] = p[i];
}

__global__ void __launch_bounds__(256) k_scatter_real(float *const *dst, const float2 *X, size_t n)
{
    const int b = blockIdx.y;
    float *q = dst[b];
    const float2 *p = X + (size_t) b * n;
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x)
        q[i] = p[i].x;
}

int launch_gather_images(float *dst, const float *const *src, size_t n, int batch, hipStream_t s)
{
    const dim3 grid((unsigned) std::min<size_t>((n + 255) / 256, 1024), batch);
    hipLaunchKernelGGL(k_gather_images, grid, dim3(256), 0, s, dst, src, n);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_scatter_real(float *const *dst, const float2 *X, size_t n, int batch, hipStream_t s)
{
    const dim3 grid((unsigned) std::min<size_t>((n + 255) / 256, 1024), batch);
    hipLaunchKernelGGL(k_scatter_real, grid, dim3(256), 0, s, dst, X, n);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_copy_from_real(float *out, const float2 *in, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_from_real, dim3(grid_1d(n)), dim3(256), 0, s, out, in, n);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_copy_to_real(float2 *out, const float *in, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_to_real, dim3(grid_1d(n)), dim3(256), 0, s, out, in, n);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_complex_norm(float *out, const float2 *in, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_complex_norm, dim3(grid_1d(n)), dim3(256), 0, s, out, in, n);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_apply_modulus(float2 *out, const float2 *in, const float *M, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_apply_modulus, dim3(grid_1d(n)), dim3(256), 0, s, out, in, M, n);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_hio_constraint(float2 *gp, const float2 *g, const float *mask, size_t n, float beta,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_hio_constraint, dim3(grid_1d(n)), dim3(256), 0, s, gp, g, mask, n, beta);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_cutoff(float *d, size_t n, float t, float lo, float hi, hipStream_t s)
{
    hipLaunchKernelGGL(k_cutoff, dim3(grid_1d(n)), dim3(256), 0, s, d, n, t, lo, hi);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_reduce(int op, const float *d, size_t n, float *partials, int np, float *result,
                  hipStream_t s)
{
    const dim3 g(np), b(256), gf(1), bf(1024);
    switch (op) {
    case 0:
        hipLaunchKernelGGL(k_reduce<0>, g, b, 0, s, d, n, partials);
        hipLaunchKernelGGL(k_reduce_final<0>, gf, bf, 0, s, partials, np, result);
        break;
    case 1:
        hipLaunchKernelGGL(k_reduce<1>, g, b, 0, s, d, n, partials);
        hipLaunchKernelGGL(k_reduce_final<1>, gf, bf, 0, s, partials, np, result);
        break;
    case 2:
        hipLaunchKernelGGL(k_reduce<2>, g, b, 0, s, d, n, partials);
        hipLaunchKernelGGL(k_reduce_final<2>, gf, bf, 0, s, partials, np, result);
        break;
    default: return IMRESH_ERR_INVALID;
    }
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_hio_error(const float2 *g, const void *mask, int mask_kind, size_t n, int invert,
                     float *perr, unsigned *pcnt, int np, float *res, hipStream_t s)
{
    if (mask_kind == kMaskU8)
        hipLaunchKernelGGL(k_hio_error<uint8_t>, dim3(np), dim3(256), 0, s, g,
                           (const uint8_t *) mask, n, invert, perr, pcnt);
    else if (mask_kind == kMaskBits)
        hipLaunchKernelGGL(k_hio_error<MaskBits>, dim3(np), dim3(256), 0, s, g,
                           (const MaskBits *) mask, n, invert, perr, pcnt);
    else
        hipLaunchKernelGGL(k_hio_error<float>, dim3(np), dim3(256), 0, s, g, (const float *) mask,
                           n, invert, perr, pcnt);
    hipLaunchKernelGGL(k_err_finalize, dim3(1), dim3(1024), 0, s, perr, pcnt, np, 0.f, res,
                       (int *) nullptr, (unsigned *) nullptr);
    return imr::launch_status(__FILE__, __LINE__);
}

// word w of the bit-packed mask from elements 32w .. 32w+31 of a byte mask
// (nonzero = 1), element 32w + j at bit 31 - j; a partial last word is
// zero-filled
__global__ void __launch_bounds__(256) k_pack_mask(const uint8_t *mask, size_t n, uint32_t *bits)
{
    const size_t nw = (n + 31) / 32;
    for (size_t w = blockIdx.x * (size_t) blockDim.x + threadIdx.x; w < nw;
         w += (size_t) gridDim.x * blockDim.x) {
        uint32_t v = 0;
        const size_t i0 = w * 32;
        for (int j = 0; j < 32; ++j)
            if (i0 + j < n && mask[i0 + j] != 0) v |= 1u << (31 - j);
        bits[w] = v;
    }
}

int launch_pack_mask(const uint8_t *mask, size_t n, uint32_t *bits, hipStream_t s)
{
    const size_t nw = (n + 31) / 32;
    if (nw == 0) return IMRESH_OK;
    hipLaunchKernelGGL(k_pack_mask, dim3(grid_1d(nw)), dim3(256), 0, s, mask, n, bits);
    return imr::launch_status(__FILE__, __LINE__);
}

int launch_store_words(float *dst, const float *src, size_t n, hipStream_t s)
{
    for (size_t off = 0; off < n; off += kStoreWords) {
        StoreWords a;
        const unsigned k = (unsigned) std::min<size_t>(kStoreWords, n - off);
        for (unsigned i = 0; i < kStoreWords; ++i) a.v[i] = i < k ? (src ? src[off + i] : 0.f) : 0.f;
        hipLaunchKernelGGL(k_store_words, dim3(1), dim3(256), 0, s, dst + off, k, a);
        IMR_HIP_CHECK(hipGetLastError());
    }
    return IMRESH_OK;
}

int launch_fill_state(float2 *X, float2 *G, uint8_t *mask, float *M, float *Bn, size_t n, int W,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_state, dim3(grid_1d(n)), dim3(256), 0, s, X, G, mask, M, Bn, n, W);
    return imr::launch_status(__FILE__, __LINE__);
}

}  // namespace imr
