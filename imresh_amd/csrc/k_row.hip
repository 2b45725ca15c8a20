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
//             by its output set, gPrev = g' (cudaShrinkWrap.cu:304-309), |g'| for the
//             blur (:240) and the HIO error sums (cudaVectorReduce.cu:284-355).
//   RK_FWD / RK_INV  plain row transforms of the 2-D FFT entry point;
//            RK_FWD_NORM stores |z| of the forward one (the forward model).
#include "internal.hpp"
#include "blue_device.hpp"
#include "ops_device.hpp"
#include "plans.hpp"

#ifndef IMR_GROUP
#error "compile with -DIMR_GROUP=<plan group>"
#endif

namespace imr {

namespace {

enum RowKernel { RK_FWD, RK_INV, RK_INVOUT, RK_FIRST, RK_MID, RK_FWD_NORM };

// NL rows per workgroup; PF (RK_MID only): gPrev and the mask are read right
// after the row gather instead of at the junction; BL: Bluestein kernels of
// row length a.len with P the convolution plan; FL (RK_INVOUT, RK_FIRST): the
// kRow* outputs / cut-off, fixed at compile time (a uniform runtime flag
// tested per element slows the memory phases measurably)
template <class P, int NL, int K, bool PF = false, bool BL = false, int STG = 0, int FL = 0>
__global__ void __launch_bounds__(NL *P::T) k_row(RowArgs a)
{
    // image b of a batch (grid.y): its planes start b * istride elements in,
    // its scalars at b * kScalStride / kFlagStride
    const int b = blockIdx.y;
    if (a.done != nullptr && a.done[b * kFlagStride]) return;
    if constexpr (STG > 0) {
        // phase offset of the second half of the grid (see k_col)
        if (blockIdx.x >= (gridDim.x >> 1))
            for (int i = 0; i < STG; ++i) __builtin_amdgcn_s_sleep(127);
    }
    extern __shared__ cf smem[];
    constexpr int T = P::T;
    const int n = BL ? a.len : P::n;
    const Blue bl{n, (const cf *) a.bs};
    const int l = threadIdx.x / T, t = threadIdx.x - l * T;
    const int row = blockIdx.x * NL + l;
    const bool valid = row < a.H;
    // workgroup base (uniform) + 32-bit per-lane offsets; rows past the end
    // read row 0 of the workgroup and never store
    const size_t wb = (size_t) b * a.istride + (size_t) blockIdx.x * NL * n;
    const int lo = valid ? l * n : 0;
    const LdsC<1> L{smem, l * P::n};
    const cf *tw = (const cf *) a.tw;
    cf v[P::E];
    cf *X = (cf *) a.X + wb;
    auto ldX = [=](int e) { return at32(X, lo + e); };
    auto stX = [=](int e, cf x) { if (valid) at32(X, lo + e) = x; };

    if constexpr (K == RK_FWD || K == RK_INV || K == RK_FWD_NORM) {
        const cf *in = (const cf *) a.in + wb;
        auto g0 = [=](int e) { return at32(in, lo + e); };
        if constexpr (K == RK_FWD_NORM) {
            // the last half of the forward model |FFT2(x)|
            // (diffractionIntensity.cpp:72-77): only |z| leaves the registers
            float *Bn = a.Bn + wb;
            auto sl = [=](int e, cf x) { if (valid) at32(Bn, lo + e) = cnorm(x); };
            xfft<P, -1, BL, true>(v, tw, t, L, bl, g0, sl);
        } else {
            cf *out = (cf *) a.out + wb;
            auto sl = [=](int e, cf x) { if (valid) at32(out, lo + e) = x; };
            xfft<P, K == RK_FWD ? -1 : +1, BL, true>(v, tw, t, L, bl, g0, sl);
        }
    } else if constexpr (K == RK_INVOUT) {
        __shared__ float red_f[NL * P::T];
        __shared__ unsigned red_u[NL * P::T];
        constexpr int fl = FL;
        cf *G = (cf *) a.G + wb;
        float *Bn = a.Bn + wb;
        const uint8_t *mask = a.mask + wb;
        float err = 0.f;
        unsigned cnt = 0;
        auto sl = [&](int e, cf g) {
            if (!valid) return;
            at32(X, lo + e) = g;
            if constexpr ((fl & kRowWriteG) != 0) at32(G, lo + e) = g;
            if constexpr ((fl & (kRowWriteBn | kRowError)) != 0) {
                const float nrm = cnorm(g);
                if constexpr ((fl & kRowWriteBn) != 0) at32(Bn, lo + e) = nrm;
                if ((fl & kRowError) && at32(mask, lo + e)) {
                    err += nrm;
                    cnt += 1;
                }
            }
        };
        xfft<P, +1, BL, true>(v, tw, t, L, bl, ldX, sl);
        if constexpr ((fl & kRowError) != 0) {
            err = block_reduce(err, red_f, [](float x, float y) { return x + y; });
            cnt = block_reduce(cnt, red_u, [](unsigned x, unsigned y) { return x + y; });
            if (threadIdx.x == 0) {
                a.perr[b * gridDim.x + blockIdx.x] = err;
                a.pcnt[b * gridDim.x + blockIdx.x] = cnt;
            }
        }
    } else if constexpr (K == RK_FIRST) {
        constexpr bool cut = (FL & kRowCut) != 0;
        const float thr = cut ? a.thr[b * kScalStride] : 0.f;
        cf *G = (cf *) a.G + wb;
        uint8_t *mask = a.mask + wb;
        const float *B2 = a.B2 + wb;
        const float beta = a.beta;
        auto g0 = [=](int e) {
            const cf g = at32(X, lo + e);
            bool m;
            if constexpr (cut) {
                m = at32(B2, lo + e) < thr;
                if (valid) at32(mask, lo + e) = m ? 1 : 0;
            } else {
                m = at32(mask, lo + e) == 1;
            }
            const cf gp = hio_constraint(at32(G, lo + e), g, m, beta);
            if (valid) at32(G, lo + e) = gp;
            return gp;
        };
        xfft<P, -1, BL, true>(v, tw, t, L, bl, g0, stX);
    } else if constexpr (K == RK_MID) {
        cf *G = (cf *) a.G + wb;
        const uint8_t *mask = a.mask + wb;
        const float beta = a.beta;
        struct GM {
            cf g;
            bool m;
        };
        auto pre = [=](int e) {
            if constexpr (PF) return GM{at32(G, lo + e), at32(mask, lo + e) == 1};
            else return NoPre{};
        };
        auto mid = [=](int e, cf g, auto p) {
            cf gp;
            if constexpr (PF) gp = hio_constraint(p.g, g, p.m, beta);
            else gp = hio_constraint(at32(G, lo + e), g, at32(mask, lo + e) == 1, beta);
            if (valid) at32(G, lo + e) = gp;
            return gp;
        };
        xpair<P, +1, -1, BL, true>(v, tw, t, L, bl, ldX, pre, mid, stX);
    }
}

// the output sets of the closing row pass
constexpr int kInit = kRowWriteG | kRowWriteBn;
constexpr int kLast = kRowWriteBn | kRowError;
constexpr int kInitErr = kRowWriteG | kRowWriteBn | kRowError;

template <class P, int NL, bool PF = false, bool BL = false, int STG = 0>
int launch_variant(int mode, RowArgs a, hipStream_t s)
{
    static_assert(NL * P::T <= 1024, "row workgroup too large");
    const dim3 grid((a.H + NL - 1) / NL, a.batch > 1 ? a.batch : 1), block(NL * P::T);
    constexpr size_t lds = sizeof(cf) * (size_t) lds_elems(NL * P::n);
    static_assert(lds <= kLdsBudget, "row tile exceeds the LDS budget");
    if constexpr (lds > 64 * 1024) {
        static const hipError_t attr = [&] {
            hipError_t e = hipSuccess;
            for (const void *f : {(const void *) k_row<P, NL, RK_FWD, false, BL>,
                                  (const void *) k_row<P, NL, RK_FWD_NORM, false, BL>,
                                  (const void *) k_row<P, NL, RK_INV, false, BL>,
                                  (const void *) k_row<P, NL, RK_INVOUT, false, BL, 0, kInit>,
                                  (const void *) k_row<P, NL, RK_INVOUT, false, BL, 0, kLast>,
                                  (const void *) k_row<P, NL, RK_INVOUT, false, BL, 0, 0>,
                                  (const void *) k_row<P, NL, RK_INVOUT, false, BL, 0, kInitErr>,
                                  (const void *) k_row<P, NL, RK_FIRST, false, BL, 0, kRowCut>,
                                  (const void *) k_row<P, NL, RK_FIRST, false, BL, 0, 0>,
                                  (const void *) k_row<P, NL, RK_MID, PF, BL, STG>})
                if (e == hipSuccess)
                    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
            return e;
        }();
        if (attr != hipSuccess) return from_hip(attr);
    }
    switch (mode) {
    case ROW_FWD: IMR_LAUNCH((k_row<P, NL, RK_FWD, false, BL>), grid, block, lds, s, a); break;
    case ROW_FWD_NORM: IMR_LAUNCH((k_row<P, NL, RK_FWD_NORM, false, BL>), grid, block, lds, s, a); break;
    case ROW_INV: IMR_LAUNCH((k_row<P, NL, RK_INV, false, BL>), grid, block, lds, s, a); break;
    case ROW_INIT: IMR_LAUNCH((k_row<P, NL, RK_INVOUT, false, BL, 0, kInit>), grid, block, lds, s, a); break;
    case ROW_LAST: IMR_LAUNCH((k_row<P, NL, RK_INVOUT, false, BL, 0, kLast>), grid, block, lds, s, a); break;
    case ROW_LAST_PLAIN: IMR_LAUNCH((k_row<P, NL, RK_INVOUT, false, BL, 0, 0>), grid, block, lds, s, a); break;
    case ROW_INIT_ERR:
        IMR_LAUNCH((k_row<P, NL, RK_INVOUT, false, BL, 0, kInitErr>), grid, block, lds, s, a);
        break;
    case ROW_FIRST_CUT: IMR_LAUNCH((k_row<P, NL, RK_FIRST, false, BL, 0, kRowCut>), grid, block, lds, s, a); break;
    case ROW_FIRST_MASK: IMR_LAUNCH((k_row<P, NL, RK_FIRST, false, BL, 0, 0>), grid, block, lds, s, a); break;
    case ROW_MID: IMR_LAUNCH((k_row<P, NL, RK_MID, PF, BL, STG>), grid, block, lds, s, a); break;
    default: return IMRESH_ERR_INVALID;
    }
    return imr::launch_status(__FILE__, __LINE__);
}

// Row variants of a length (index: workgroup shape, then +count/2 for the
// prefetching RK_MID): shapes 0 = E 16 with ~256-thread groups; for long
// rows also 1 = E 32 (half the threads per row), 2 = E 16 with twice the
// rows, 3 = shape 0 with the second half of the grid started ~3.4 us late and,
// when a group holds 2 or more rows, a last shape with half the rows: fewer
// waves share each barrier, so the groups of a CU drift out of phase and
// overlap one's loads with another's arithmetic (2048^2: 23.8 -> 22.6 us);
// for short lines (<= 512 points, >= 8 rows per group) 1 = a quarter of the
// rows per group: 4x the workgroups, which spreads one small image over more
// CUs (a 256^2 reconstruction 199 -> 227 per second).
template <int N>
constexpr bool long_pow2() { return N >= 1024 && (N & (N - 1)) == 0; }

// shapes 0-3 of the long power-of-two lengths (2 when twice the rows do
// not fit a workgroup)
template <int N, int... Rs>
constexpr int long_row_shapes()
{
    using P16 = PlanE<16, N, Rs...>;
    return (2 * P16::L * P16::T <= 1024 &&
            sizeof(cf) * (size_t) lds_elems(2 * P16::L * N) <= kLdsBudget) ? 4 : 2;
}

template <int N, int... Rs>
constexpr int row_shape_count()
{
    using P16 = PlanE<16, N, Rs...>;
    if constexpr (long_pow2<N>()) {   // extra shapes for the long power-of-two lengths
        return long_row_shapes<N, Rs...>() + (P16::L >= 2 ? 1 : 0);   // + half the rows
    } else if constexpr (N <= 512 && P16::L >= 8) {
        return 2;   // short lines: also groups of a quarter of the rows (4x the workgroups)
    }
    return 1;
}

template <int N, int... Rs>
constexpr int row_variant_count() { return 2 * row_shape_count<N, Rs...>(); }

// Bluestein kernels exist for the power-of-two convolution lengths
template <int N>
constexpr bool blue_length() { return N >= 64 && N <= 2 * kMaxLine && (N & (N - 1)) == 0; }

template <class P, int NL, int STG = 0>
int launch_pf(bool pf, int mode, RowArgs a, hipStream_t s)
{
    return pf ? launch_variant<P, NL, true, false, STG>(mode, a, s)
              : launch_variant<P, NL, false, false, STG>(mode, a, s);
}

template <int N, int... Rs>
int launch_row_direct(int var, int mode, RowArgs a, hipStream_t s);

template <int N, int... Rs>
int launch_row_variant(int var, int mode, RowArgs a, hipStream_t s)
{
    using P16 = PlanE<16, N, Rs...>;
    if constexpr (blue_length<N>()) {
        if (var == kBlueVariant) return launch_variant<P16, P16::L, false, true>(mode, a, s);
        if (var == kBlueVariant + 1) return launch_variant<P16, P16::L, true, true>(mode, a, s);
        if (var == kBlueVariant + 2)
            return launch_variant<P16, (P16::L > 1 ? P16::L / 2 : 1), false, true>(mode, a, s);
    }
    if constexpr (N > kMaxLine) {
        return IMRESH_ERR_UNSUPPORTED_SIZE;   // a convolution length only
    } else {
        return launch_row_direct<N, Rs...>(var, mode, a, s);
    }
}

template <int N, int... Rs>
int launch_row_direct(int var, int mode, RowArgs a, hipStream_t s)
{
    using P16 = PlanE<16, N, Rs...>;
    constexpr int ns = row_shape_count<N, Rs...>();
    const bool pf = var >= ns;
    const int shape = pf ? var - ns : var;
    if constexpr (long_pow2<N>() && ns > 1) {
        using P32 = PlanE<32, N, Rs...>;
        if (shape == 1) return launch_pf<P32, P32::L>(pf, mode, a, s);
        constexpr int nb = long_row_shapes<N, Rs...>();
        if constexpr (nb == 4) {
            if (shape == 2) return launch_pf<P16, 2 * P16::L>(pf, mode, a, s);
            if (shape == 3) return launch_pf<P16, P16::L, 1>(pf, mode, a, s);
        }
        if constexpr (ns > nb) {
            if (shape == nb) return launch_pf<P16, P16::L / 2>(pf, mode, a, s);
        }
    } else if constexpr (ns > 1) {
        if (shape == 1) return launch_pf<P16, P16::L / 4>(pf, mode, a, s);
    }
    return launch_pf<P16, P16::L>(pf, mode, a, s);
}

// rows per workgroup of a variant (the host needs it for the partial sums;
// its buffers hold one slot per row, enough for any variant)
template <int N, int... Rs>
int row_variant_rows(int var)
{
    using P16 = PlanE<16, N, Rs...>;
    if (var == kBlueVariant + 2) return P16::L > 1 ? P16::L / 2 : 1;
    if (var >= kBlueVariant) return P16::L;
    constexpr int ns = row_shape_count<N, Rs...>();
    const int shape = var >= ns ? var - ns : var;
    if constexpr (long_pow2<N>() && ns > 1) {
        using P32 = PlanE<32, N, Rs...>;
        constexpr int nb = long_row_shapes<N, Rs...>();
        if (shape == 1) return P32::L;
        if (nb == 4 && shape == 2) return 2 * P16::L;
        if (ns > nb && shape == nb) return P16::L / 2;
    } else if constexpr (ns > 1) {
        if (shape == 1) return P16::L / 4;
    }
    return P16::L;
}

template <int G, int N, int... Rs>
int try_plan(int mode, int var, int n, RowArgs a, hipStream_t s)
{
    if constexpr (G == IMR_GROUP) {
        if (n == N) return launch_row_variant<N, Rs...>(var, mode, a, s);
    }
    return -1;
}

template <int G, int N, int... Rs>
int try_count(int n)
{
    if constexpr (G == IMR_GROUP) {
        if (n == N) return row_variant_count<N, Rs...>();
    }
    return -1;
}

template <int G, int N, int... Rs>
int try_rows(int n, int var)
{
    if constexpr (G == IMR_GROUP) {
        if (n == N) return row_variant_rows<N, Rs...>(var);
    }
    return -1;
}

}  // namespace

#define IMR_CAT2(a, b) a##b
#define IMR_CAT(a, b) IMR_CAT2(a, b)
#define IMR_ROW_CASE(g, N, ...)                                                     \
    {                                                                               \
        const int r = try_plan<g, N, __VA_ARGS__>(mode, var, n, a, s);              \
        if (r >= 0) return r;                                                       \
    }
#define IMR_ROW_COUNT(g, N, ...)                                                    \
    {                                                                               \
        const int r = try_count<g, N, __VA_ARGS__>(n);                              \
        if (r >= 0) return r;                                                       \
    }

int IMR_CAT(launch_row_g, IMR_GROUP)(int mode, int var, int n, const LinePlan &lp, RowArgs a,
                                     hipStream_t s)
{
    a.tw = lp.tw;
    IMR_PLANS(IMR_ROW_CASE)
    return IMRESH_ERR_UNSUPPORTED_SIZE;
}

int IMR_CAT(row_variants_g, IMR_GROUP)(int n)
{
    IMR_PLANS(IMR_ROW_COUNT)
    return 0;
}

#define IMR_ROW_ROWS(g, N, ...)                                                     \
    {                                                                               \
        const int r = try_rows<g, N, __VA_ARGS__>(n, var);                          \
        if (r >= 0) return r;                                                       \
    }

int IMR_CAT(row_rows_g, IMR_GROUP)(int n, int var)
{
    IMR_PLANS(IMR_ROW_ROWS)
    return 1;
}

}  // namespace imr
