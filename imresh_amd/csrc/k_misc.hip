// Cycle-boundary and standalone kernels of the shrink-wrap / HIO path for
// MI355X (gfx950): the fused 2-D Gaussian blur with per-tile maximum, the
// max / error finalizers that keep the threshold and the early-exit flag on
// the device, wave64 reductions (max/min/sum, HIO error) and the reference's
// elementwise operators as standalone kernels.  Blur taps are summed from 0
// in the reference order w[0]*x[i-r] ... w[2r]*x[i+r] (gaussian.cpp:51-163,
// 559-734; cudaGaussian.cu:539-652, 776-990) with clamp-to-edge borders.
#include <cfloat>

#include "internal.hpp"
#include "ops_device.hpp"

namespace imr {

// ------------------------------------------------------------ Gaussian blur
// Out-of-place separable blur of a kBlurTX x kBlurTY output tile: the input
// tile with a clamped halo is staged in LDS, the horizontal pass runs over
// all halo rows, then the vertical pass; each output is summed in the
// reference tap order.  Optionally writes the tile maximum.
__global__ void __launch_bounds__(256) k_blur2d(BlurArgs a)
{
    if (a.done != nullptr && *a.done) return;
    extern __shared__ float sm[];
    __shared__ float red[1024];
    const int rh = a.rh, rv = a.rv;
    const int AW = kBlurTX + 2 * rh, AH = kBlurTY + 2 * rv;
    float *wh = sm;                  // 64
    float *wv = sm + 64;             // 64
    float *A = sm + 128;             // AH x AW
    float *Hs = A + AH * AW;         // AH x kBlurTX
    const int x0 = blockIdx.x * kBlurTX, y0 = blockIdx.y * kBlurTY;
    const int tid = threadIdx.x;
    for (int i = tid; i < 2 * rh + 1; i += blockDim.x) wh[i] = a.wh[i];
    for (int i = tid; i < 2 * rv + 1; i += blockDim.x) wv[i] = a.wv[i];
    for (int i = tid; i < AH * AW; i += blockDim.x) {
        const int ty = i / AW, tx = i - ty * AW;
        const int gy = min(max(y0 - rv + ty, 0), a.H - 1);
        const int gx = min(max(x0 - rh + tx, 0), a.W - 1);
        A[i] = a.in[(size_t) gy * a.W + gx];
    }
    __syncthreads();
    for (int i = tid; i < AH * kBlurTX; i += blockDim.x) {
        const int ty = i / kBlurTX, tx = i - ty * kBlurTX;
        const float *x = A + ty * AW + tx;
        float sum;
        if (a.h_identity) {
            sum = x[rh];
        } else {
            sum = 0.f;
            for (int k = 0; k <= 2 * rh; ++k) sum += x[k] * wh[k];
        }
        Hs[i] = sum;
    }
    __syncthreads();
    float lmax = -FLT_MAX;
    for (int i = tid; i < kBlurTY * kBlurTX; i += blockDim.x) {
        const int ty = i / kBlurTX, tx = i - ty * kBlurTX;
        const int gy = y0 + ty, gx = x0 + tx;
        if (gy < a.H && gx < a.W) {
            const float *x = Hs + ty * kBlurTX + tx;
            float sum;
            if (a.v_identity) {
                sum = x[rv * kBlurTX];
            } else {
                sum = 0.f;
                for (int k = 0; k <= 2 * rv; ++k) sum += x[k * kBlurTX] * wv[k];
            }
            a.out[(size_t) gy * a.W + gx] = sum;
            lmax = fmax_ref(lmax, sum);
        }
    }
    if (a.pmax != nullptr) {
        lmax = block_reduce(lmax, red, [](float m, float x) { return fmax_ref(m, x); });
        if (tid == 0) a.pmax[blockIdx.y * gridDim.x + blockIdx.x] = lmax;
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

// HIO error partial sums with a float or unsigned char mask
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
        const bool z = (invert != 0) != (mask[i] != 0);
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

int launch_blur(const BlurArgs &a, hipStream_t s)
{
    const dim3 grid((a.W + kBlurTX - 1) / kBlurTX, (a.H + kBlurTY - 1) / kBlurTY);
    const int AW = kBlurTX + 2 * a.rh, AH = kBlurTY + 2 * a.rv;
    const size_t lds = sizeof(float) * (128 + (size_t) AH * AW + (size_t) AH * kBlurTX);
    static const hipError_t attr =
        hipFuncSetAttribute((const void *) k_blur2d, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int) (sizeof(float) * (128 + (kBlurTY + 62) * (kBlurTX + 62) +
                                                    (kBlurTY + 62) * kBlurTX)));
    if (attr != hipSuccess) return from_hip(attr);
    hipLaunchKernelGGL(k_blur2d, grid, dim3(256), lds, s, a);
    return from_hip(hipGetLastError());
}

int launch_blur_h_inplace(float *d, int W, int H, const float *w, int r, hipStream_t s)
{
    const size_t lds = sizeof(float) * (64 + (size_t) W);
    static const hipError_t attr = hipFuncSetAttribute(
        (const void *) k_blur_h_inplace, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
    if (attr != hipSuccess) return from_hip(attr);
    hipLaunchKernelGGL(k_blur_h_inplace, dim3(H), dim3(256), lds, s, d, W, w, r);
    return from_hip(hipGetLastError());
}

int launch_max_finalize(const float *pmax, int np, float cut, float *thr, const int *done,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_max_finalize, dim3(1), dim3(1024), 0, s, pmax, np, cut, thr, done);
    return from_hip(hipGetLastError());
}

int launch_err_finalize(const float *perr, const unsigned *pcnt, int np, float target, float *res,
                        int *done, unsigned *cycles, hipStream_t s)
{
    hipLaunchKernelGGL(k_err_finalize, dim3(1), dim3(1024), 0, s, perr, pcnt, np, target, res,
                       done, cycles);
    return from_hip(hipGetLastError());
}

int launch_copy_from_real(float *out, const float2 *in, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_from_real, dim3(grid_1d(n)), dim3(256), 0, s, out, in, n);
    return from_hip(hipGetLastError());
}

int launch_copy_to_real(float2 *out, const float *in, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_to_real, dim3(grid_1d(n)), dim3(256), 0, s, out, in, n);
    return from_hip(hipGetLastError());
}

int launch_complex_norm(float *out, const float2 *in, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_complex_norm, dim3(grid_1d(n)), dim3(256), 0, s, out, in, n);
    return from_hip(hipGetLastError());
}

int launch_apply_modulus(float2 *out, const float2 *in, const float *M, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_apply_modulus, dim3(grid_1d(n)), dim3(256), 0, s, out, in, M, n);
    return from_hip(hipGetLastError());
}

int launch_hio_constraint(float2 *gp, const float2 *g, const float *mask, size_t n, float beta,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_hio_constraint, dim3(grid_1d(n)), dim3(256), 0, s, gp, g, mask, n, beta);
    return from_hip(hipGetLastError());
}

int launch_cutoff(float *d, size_t n, float t, float lo, float hi, hipStream_t s)
{
    hipLaunchKernelGGL(k_cutoff, dim3(grid_1d(n)), dim3(256), 0, s, d, n, t, lo, hi);
    return from_hip(hipGetLastError());
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
    return from_hip(hipGetLastError());
}

int launch_hio_error(const float2 *g, const void *mask, int mask_is_u8, size_t n, int invert,
                     float *perr, unsigned *pcnt, int np, float *res, hipStream_t s)
{
    if (mask_is_u8)
        hipLaunchKernelGGL(k_hio_error<uint8_t>, dim3(np), dim3(256), 0, s, g,
                           (const uint8_t *) mask, n, invert, perr, pcnt);
    else
        hipLaunchKernelGGL(k_hio_error<float>, dim3(np), dim3(256), 0, s, g, (const float *) mask,
                           n, invert, perr, pcnt);
    hipLaunchKernelGGL(k_err_finalize, dim3(1), dim3(1024), 0, s, perr, pcnt, np, 0.f, res,
                       (int *) nullptr, (unsigned *) nullptr);
    return from_hip(hipGetLastError());
}

int launch_fill_state(float2 *X, float2 *G, uint8_t *mask, float *M, float *Bn, size_t n, int W,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_state, dim3(grid_1d(n)), dim3(256), 0, s, X, G, mask, M, Bn, n, W);
    return from_hip(hipGetLastError());
}

}  // namespace imr
