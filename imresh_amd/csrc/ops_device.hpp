// Elementwise device operators of the hot path and a deterministic
// workgroup reduction.  The library is compiled with -ffp-contract=off and
// IEEE sqrtf / division, so each operator rounds exactly like the
// reference's C++ code given identical inputs:
//   constraint   cudaVectorElementwise.cu:42-72, shrinkWrap.cpp:263-276
//   projection   cudaVectorElementwise.cu:168-195, vectorElementwise.cpp:59-81
//   norm         cudaVectorElementwise.cu:120-141, vectorElementwise.cpp:41-56
#pragma once

#include "fft_device.hpp"

namespace imr {

// if (masked or Re g' < 0): gPrev -= beta*g'  else gPrev = g'
__device__ __forceinline__ cf hio_constraint(cf gp, cf g, bool masked, float beta)
{
    if (masked || g.x < 0) {
        gp.x -= beta * g.x;
        gp.y -= beta * g.y;
        return gp;
    }
    return g;
}

__device__ __forceinline__ float cnorm(cf z) { return sqrtf(z.x * z.x + z.y * z.y); }

// |z| := m, keeping the phase; z == 0 stays 0
__device__ __forceinline__ cf modulus_project(cf z, float m)
{
    float norm = sqrtf(z.x * z.x + z.y * z.y);
    if (norm == 0) norm = 1;
    const float f = m / norm;
    return mkc(z.x * f, z.y * f);
}

// Global access at a workgroup-uniform base plus a 32-bit element index:
// base + zext(idx * sizeof(T)) maps onto the scalar-base + 32-bit vector
// offset addressing of global loads/stores (callers keep idx * sizeof(T)
// below 4 GiB: at most 8192 x 8192 complex values).
template <class T>
__device__ __forceinline__ T &at32(T *base, unsigned idx)
{
    return *(T *) ((char *) base + (unsigned) (idx * (unsigned) sizeof(T)));
}
template <class T>
__device__ __forceinline__ const T &at32(const T *base, unsigned idx)
{
    return *(const T *) ((const char *) base + (unsigned) (idx * (unsigned) sizeof(T)));
}

// Projection used inside the fused column pass: f = m * rsqrt(|z|^2) with
// the hardware reciprocal square root (1 ulp) instead of the correctly
// rounded sqrtf and division (~20 instructions); within 2 ulp of
// modulus_project, far below the FFT's own rounding.  Norms below 1e-15
// (|z|^2 < 1e-30, near the denormal range) take the exact path, which also
// keeps z == 0 at 0.
__device__ __forceinline__ cf modulus_project_fused(cf z, float m)
{
    const float n2 = z.x * z.x + z.y * z.y;
    if (n2 >= 1e-30f) {
        const float f = m * __builtin_amdgcn_rsqf(n2);
        return mkc(z.x * f, z.y * f);
    }
    return modulus_project(z, m);
}

// one wave64 lane move of a 32-bit value through DPP (CTRL: the DPP control
// word; every source lane used below lies inside the same row of 16 lanes)
template <int CTRL, class T>
__device__ __forceinline__ T dpp_move(T v)
{
    static_assert(sizeof(T) == 4, "32-bit lanes");
    return __builtin_bit_cast(
        T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

template <class T>
__device__ __forceinline__ T read_lane(T v, int lane)
{
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// wave64 reduction for a wave whose 64 lanes are all active: DPP butterflies
// inside each row of 16 lanes (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror), then the four row results combined in a fixed
// order.  Every lane returns the same, scheduling-independent value.
template <class T, class Op>
__device__ __forceinline__ T wave_reduce(T x, Op op)
{
    x = op(x, dpp_move<0xB1>(x));
    x = op(x, dpp_move<0x4E>(x));
    x = op(x, dpp_move<0x141>(x));
    x = op(x, dpp_move<0x140>(x));
    return op(op(read_lane(x, 0), read_lane(x, 16)), op(read_lane(x, 32), read_lane(x, 48)));
}

// deterministic workgroup reduction (any blockDim <= 1024; red holds at least
// blockDim entries): whole waves reduce with wave_reduce and combine their
// results in wave order; a block with a partial wave uses a fixed pairwise
// LDS tree.  Results do not depend on scheduling.
template <class T, class Op>
__device__ __forceinline__ T block_reduce(T x, T *red, Op op)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    if ((nt & 63) == 0) {
        x = wave_reduce(x, op);
        __syncthreads();
        if ((tid & 63) == 0) red[tid >> 6] = x;
        __syncthreads();
        T r = red[0];
        for (int i = 1; i < (nt >> 6); ++i) r = op(r, red[i]);
        return r;
    }
    __syncthreads();
    red[tid] = x;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (tid < s && tid + s < nt) red[tid] = op(red[tid], red[tid + s]);
        __syncthreads();
    }
    return red[0];
}

// outputs of the closing row pass / cut-off of the first (template FL of k_row)
constexpr int kRowWriteG = 1;     // write gPrev = g'
constexpr int kRowWriteBn = 2;    // write |g'| (blur input)
constexpr int kRowError = 4;      // per-workgroup HIO error sums
constexpr int kRowCut = 8;        // mask = blurred < threshold (else read mask)

}  // namespace imr
