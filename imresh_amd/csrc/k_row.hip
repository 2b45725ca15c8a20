// Row-pass kernels (one workgroup = P::L rows of length W = P::n, P::T
// threads per row).  Compiled once per plan group (-DIMR_GROUP=g).
//
//   RK_MID    inverse row FFT -> HIO constraint (mask, gPrev) -> forward row FFT
//             The steady-state half of an HIO iteration: the inverse 2-D FFT of
//             the previous iteration is finished here (its column half ran in
//             the column pass) and the forward 2-D FFT of this iteration starts
//             here, with the constraint (cudaVectorElementwise.cu:42-72)
//             applied in registers in between.
//   RK_FIRST  first iteration of a cycle: g' is already in real space; the
//             mask cut-off (cudaVectorElementwise.cu:198-219 with the device
//             threshold) is fused into the load, then constraint + forward FFT.
//   RK_INVOUT inverse row FFT closing a reconstruction step: writes g' and,
//             by flags, gPrev = g' (cudaShrinkWrap.cu:304-309), |g'| for the
//             blur (:240) and the HIO error sums (cudaVectorReduce.cu:284-355).
//   RK_FWD / RK_INV  plain row transforms of the 2-D FFT entry point.
#include "internal.hpp"
#include "ops_device.hpp"
#include "plans.hpp"

#ifndef IMR_GROUP
#error "compile with -DIMR_GROUP=<plan group>"
#endif

namespace imr {

namespace {

enum RowKernel { RK_FWD, RK_INV, RK_INVOUT, RK_FIRST, RK_MID };

template <class P, int K>
__global__ void __launch_bounds__(P::L *P::T) k_row(RowArgs a)
{
    if (a.done != nullptr && *a.done) return;
    extern __shared__ cf smem[];
    constexpr int n = P::n, T = P::T;
    const int l = threadIdx.x / T, t = threadIdx.x - l * T;
    const int row = blockIdx.x * P::L + l;
    const bool valid = row < a.H;
    // workgroup base (uniform) + 32-bit per-lane offsets; rows past the end
    // read row 0 of the workgroup and never store
    const size_t wb = (size_t) blockIdx.x * P::L * n;
    const int lo = valid ? l * n : 0;
    const LdsLine L{smem, l * n, 1};
    const cf *tw = (const cf *) a.tw;
    cf v[P::E];
    cf *X = (cf *) a.X + wb;
    auto ldX = [=](int e) { return X[lo + e]; };
    auto stX = [=](int e, cf x) { if (valid) X[lo + e] = x; };

    if constexpr (K == RK_FWD || K == RK_INV) {
        const cf *in = (const cf *) a.in + wb;
        cf *out = (cf *) a.out + wb;
        auto g0 = [=](int e) { return in[lo + e]; };
        auto sl = [=](int e, cf x) { if (valid) out[lo + e] = x; };
        if constexpr (K == RK_FWD) line_fft<P, -1>(v, tw, t, L, g0, sl);
        else line_fft<P, +1>(v, tw, t, L, g0, sl);
    } else if constexpr (K == RK_INVOUT) {
        __shared__ float red_f[P::L * P::T];
        __shared__ unsigned red_u[P::L * P::T];
        const int fl = a.flags;
        cf *G = (cf *) a.G + wb;
        float *Bn = a.Bn + wb;
        const uint8_t *mask = a.mask + wb;
        float err = 0.f;
        unsigned cnt = 0;
        auto sl = [&](int e, cf g) {
            if (!valid) return;
            X[lo + e] = g;
            if (fl & kRowWriteG) G[lo + e] = g;
            if (fl & (kRowWriteBn | kRowError)) {
                const float nrm = cnorm(g);
                if (fl & kRowWriteBn) Bn[lo + e] = nrm;
                if ((fl & kRowError) && mask[lo + e]) {
                    err += nrm;
                    cnt += 1;
                }
            }
        };
        line_fft<P, +1>(v, tw, t, L, ldX, sl);
        if (fl & kRowError) {
            err = block_reduce(err, red_f, [](float x, float y) { return x + y; });
            cnt = block_reduce(cnt, red_u, [](unsigned x, unsigned y) { return x + y; });
            if (threadIdx.x == 0) {
                a.perr[blockIdx.x] = err;
                a.pcnt[blockIdx.x] = cnt;
            }
        }
    } else if constexpr (K == RK_FIRST) {
        const bool cut = (a.flags & kRowCut) != 0;
        const float thr = cut ? *a.thr : 0.f;
        cf *G = (cf *) a.G + wb;
        uint8_t *mask = a.mask + wb;
        const float *B2 = a.B2 + wb;
        const float beta = a.beta;
        auto g0 = [=](int e) {
            const cf g = X[lo + e];
            bool m;
            if (cut) {
                m = B2[lo + e] < thr;
                if (valid) mask[lo + e] = m ? 1 : 0;
            } else {
                m = mask[lo + e] == 1;
            }
            const cf gp = hio_constraint(G[lo + e], g, m, beta);
            if (valid) G[lo + e] = gp;
            return gp;
        };
        line_fft<P, -1>(v, tw, t, L, g0, stX);
    } else if constexpr (K == RK_MID) {
        cf *G = (cf *) a.G + wb;
        const uint8_t *mask = a.mask + wb;
        const float beta = a.beta;
        auto mid = [=](int e, cf g) {
            const cf gp = hio_constraint(G[lo + e], g, mask[lo + e] == 1, beta);
            if (valid) G[lo + e] = gp;
            return gp;
        };
        line_pair<P, +1, -1>(v, tw, t, L, ldX, mid, stX);
    }
}

template <class P>
int launch_plan(int mode, RowArgs a, hipStream_t s)
{
    const dim3 grid((a.H + P::L - 1) / P::L), block(P::L * P::T);
    const size_t lds = sizeof(cf) * (size_t) lds_elems(P::L * P::n);
    if (lds > 64 * 1024) {
        static const hipError_t attr = [&] {
            hipError_t e = hipSuccess;
            for (const void *f : {(const void *) k_row<P, RK_FWD>, (const void *) k_row<P, RK_INV>,
                                  (const void *) k_row<P, RK_INVOUT>, (const void *) k_row<P, RK_FIRST>,
                                  (const void *) k_row<P, RK_MID>})
                if (e == hipSuccess)
                    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
            return e;
        }();
        if (attr != hipSuccess) return from_hip(attr);
    }
    switch (mode) {
    case ROW_FWD: a.flags = 0; hipLaunchKernelGGL((k_row<P, RK_FWD>), grid, block, lds, s, a); break;
    case ROW_INV: a.flags = 0; hipLaunchKernelGGL((k_row<P, RK_INV>), grid, block, lds, s, a); break;
    case ROW_INIT:
        a.flags = kRowWriteG | kRowWriteBn;
        hipLaunchKernelGGL((k_row<P, RK_INVOUT>), grid, block, lds, s, a);
        break;
    case ROW_LAST:
        a.flags = kRowWriteBn | kRowError;
        hipLaunchKernelGGL((k_row<P, RK_INVOUT>), grid, block, lds, s, a);
        break;
    case ROW_LAST_PLAIN:
        a.flags = 0;
        hipLaunchKernelGGL((k_row<P, RK_INVOUT>), grid, block, lds, s, a);
        break;
    case ROW_FIRST_CUT:
        a.flags = kRowCut;
        hipLaunchKernelGGL((k_row<P, RK_FIRST>), grid, block, lds, s, a);
        break;
    case ROW_FIRST_MASK:
        a.flags = 0;
        hipLaunchKernelGGL((k_row<P, RK_FIRST>), grid, block, lds, s, a);
        break;
    case ROW_MID: a.flags = 0; hipLaunchKernelGGL((k_row<P, RK_MID>), grid, block, lds, s, a); break;
    default: return IMRESH_ERR_INVALID;
    }
    return from_hip(hipGetLastError());
}

template <int G, int N, int... Rs>
int try_plan(int mode, int n, RowArgs a, hipStream_t s)
{
    if constexpr (G == IMR_GROUP) {
        if (n == N) return launch_plan<Plan<N, Rs...>>(mode, a, s);
    }
    return -1;
}

}  // namespace

#define IMR_CAT2(a, b) a##b
#define IMR_CAT(a, b) IMR_CAT2(a, b)
#define IMR_ROW_CASE(g, N, ...)                                                     \
    {                                                                               \
        const int r = try_plan<g, N, __VA_ARGS__>(mode, n, a, s);                   \
        if (r >= 0) return r;                                                       \
    }

int IMR_CAT(launch_row_g, IMR_GROUP)(int mode, int n, const LinePlan &lp, RowArgs a, hipStream_t s)
{
    a.tw = lp.tw;
    IMR_PLANS(IMR_ROW_CASE)
    return IMRESH_ERR_UNSUPPORTED_SIZE;
}

}  // namespace imr
