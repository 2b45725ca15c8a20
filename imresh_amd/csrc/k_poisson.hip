// Measurement simulation on the device: a modulus |F| (the output of the
// forward model) becomes a photon-counting measurement's modulus
// sqrt(k), k ~ Poisson(photons |F|^2 / sum |F|^2), per element.  The sum is
// taken in double in a fixed order and the sampler is a fixed sequence of
// IEEE operations (poisson_device.hpp), so the result does not depend on the
// launch shape and a CPU restatement reproduces it bit for bit.
#include "internal.hpp"
#include "poisson_device.hpp"

namespace imr {

namespace {

constexpr int kSumThreads = 256;

// lane j = block * 256 + thread adds m[i]^2 for i = j, j + L, j + 2L, ...
// (L = 256 * 256 lanes) in sequence; a block's 256 lane sums combine by the
// halving tree s[t] += s[t + h], h = 128, 64, ..., 1
__global__ void __launch_bounds__(kSumThreads) k_sq_sum(const float *m, size_t n, double *partial)
{
    __shared__ double s[kSumThreads];
    const size_t lanes = (size_t) kPoissonSumBlocks * kSumThreads;
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t) kSumThreads + threadIdx.x; i < n; i += lanes) {
        const double v = (double) m[i];
        acc = acc + v * v;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kSumThreads / 2; h > 0; h >>= 1) {
        if ((int) threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// the kPoissonSumBlocks block sums by the same halving tree -> partial[kPoissonSumBlocks]
__global__ void __launch_bounds__(kPoissonSumBlocks) k_sum_blocks(double *partial)
{
    __shared__ double s[kPoissonSumBlocks];
    s[threadIdx.x] = partial[threadIdx.x];
    __syncthreads();
    for (int h = kPoissonSumBlocks / 2; h > 0; h >>= 1) {
        if ((int) threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[kPoissonSumBlocks] = s[0];
}

// element i: lambda = (photons * m_i^2) / total, k ~ Poisson(lambda) from the
// counter-based stream (seed, i), m_i <- (float) sqrt(k)
__global__ void __launch_bounds__(256) k_poisson_modulus(float *m, size_t n, const double *total,
                                                         double photons, uint64_t seed)
{
    const double t = *total;
    for (size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n;
         i += (size_t) gridDim.x * blockDim.x) {
        const double v = (double) m[i];
        const double lam = t > 0.0 ? photons * (v * v) / t : 0.0;
        m[i] = (float) pois::sqrt_d(pois::poisson(lam, seed, (uint64_t) i));
    }
}

}  // namespace

int launch_poisson_modulus(float *m, size_t n, double photons, uint64_t seed, double *aux, hipStream_t s)
{
    if (n == 0) return IMRESH_OK;
    hipLaunchKernelGGL(k_sq_sum, dim3(kPoissonSumBlocks), dim3(kSumThreads), 0, s, (const float *) m, n, aux);
    hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(kPoissonSumBlocks), 0, s, aux);
    const unsigned blocks = (unsigned) std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_poisson_modulus, dim3(blocks), dim3(256), 0, s, m, n,
                       (const double *) (aux + kPoissonSumBlocks), photons, seed);
    return imr::launch_status(__FILE__, __LINE__);
}

}  // namespace imr
