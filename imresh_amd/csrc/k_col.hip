// Column-pass kernels (one workgroup = P::C adjacent columns of length
// H = P::n, the tile staged in LDS as [H][C], P::T threads per column).
// Compiled once per plan group (-DIMR_GROUP=g).
//
//   CK_PROJ  forward column FFT -> modulus projection -> inverse column FFT.
//            Finishes the forward 2-D FFT of an HIO iteration, replaces |G'|
//            by the measured modulus (cudaVectorElementwise.cu:168-195) in
//            registers, and starts the inverse 2-D FFT.
//   CK_INV   inverse column transform; CK_INV_REAL reads the real modulus M
//            instead, the first half of IFFT2((M, 0)) (cudaShrinkWrap.cu:229-232).
//   CK_FWD   forward column transform of the 2-D FFT entry point;
//            CK_FWD_REAL reads a real plane instead, the first half of the
//            forward model FFT2((x, 0)) (diffractionIntensity.cpp:59-70).
#include "internal.hpp"
#include "blue_device.hpp"
#include "ops_device.hpp"
#include "plans.hpp"

#include <type_traits>

#ifndef IMR_GROUP
#error "compile with -DIMR_GROUP=<plan group>"
#endif

namespace imr {

namespace {

enum ColKernel { CK_FWD, CK_INV, CK_PROJ, CK_FWD_REAL, CK_INV_REAL };

// C columns per workgroup; SPLIT: exchange real and imaginary parts through
// LDS one after the other (half the LDS of a complex tile); PF (CK_PROJ
// only): the measured modulus is read right after the column gather instead
// of at the junction.
template <class P, int C, bool SPLIT, int K, bool PF = false, bool BL = false, int STG = 0>
__global__ void __launch_bounds__(C *P::T) k_col(ColArgs a)
{
    // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs
    // (workgroup x + y * gridDim.x to XCD (x + y * gridDim.x) mod 8), each
    // with its own L2; give every XCD a contiguous range of tiles so
    // neighbouring tiles (which share 128-byte lines of X and M) meet in the
    // same L2.  A grid whose width is a multiple of 8 keeps its images in
    // place and remaps within each row of tiles; any other grid is remapped
    // as a whole (consecutive tiles, of one image and the next, per XCD).
    int bid = blockIdx.x, by = blockIdx.y;
    if ((gridDim.x & 7) == 0) {
        bid = (bid & 7) * (gridDim.x >> 3) + (bid >> 3);
    } else {
        const int gx = gridDim.x, nt = gx * gridDim.y, lin = bid + by * gx;
        const int x8 = lin & 7, k = x8 * (nt >> 3) + min(x8, nt & 7) + (lin >> 3);
        bid = k % gx;
        by = k / gx;
    }
    // image b of a batch and, for the middle step of a four-step column
    // pass, its sub-block sb (grid.y = images x sub)
    const int sub = a.sub > 1 ? a.sub : 1;
    const int b = by / sub, sb = by - b * sub;
    if (a.done != nullptr && a.done[b * kFlagStride]) return;
    if constexpr (STG > 0) {
        // phase offset: the second half of the grid shares CUs with the first
        // half; starting it STG x ~3.4 us later lets the loads of one overlap
        // the arithmetic of the other instead of every workgroup loading,
        // computing and storing in lockstep
        if (blockIdx.x >= (gridDim.x >> 1))
            for (int i = 0; i < STG; ++i) __builtin_amdgcn_s_sleep(127);
    }
    extern __shared__ cf smem[];
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const int col = bid * C + c;
    const bool valid = col < a.W;
    const int W = a.W;
    // workgroup base (uniform) + 32-bit per-lane offsets e*W + c; columns past
    // the edge read column 0 of the tile and never store
    const size_t wb = (size_t) b * a.istride + (size_t) sb * a.sub_xstride + (size_t) bid * C;
    const size_t mb = (size_t) b * a.istride + (size_t) sb * a.sub_mstride + (size_t) bid * C;
    const int MP = a.mpitch > 0 ? a.mpitch : W;   // row pitch of M (CK_PROJ)
    const int co = valid ? c : 0;
    using Lds = std::conditional_t<SPLIT, LdsS<C>, LdsC<C>>;
    const Lds L{(decltype(Lds::p)) smem, c};
    const cf *tw = (const cf *) a.tw;
    cf v[P::E];
    const Blue bl{BL ? a.len : P::n, (const cf *) a.bs};

    if constexpr (K == CK_FWD || K == CK_INV || K == CK_FWD_REAL || K == CK_INV_REAL) {
        constexpr bool real_in = K == CK_FWD_REAL || K == CK_INV_REAL;
        constexpr int DIR = (K == CK_FWD || K == CK_FWD_REAL) ? -1 : +1;
        cf *out = (cf *) a.out + wb;
        auto sl = [=](int e, cf x) { if (valid) at32(out, e * W + co) = x; };
        if constexpr (real_in) {
            const float *M = a.M + wb;
            auto g0 = [=](int e) { return mkc(at32(M, e * W + co), 0.f); };
            xfft<P, DIR, BL>(v, tw, t, L, bl, g0, sl);
        } else {
            const cf *in = (const cf *) a.in + wb;
            auto g0 = [=](int e) { return at32(in, e * W + co); };
            xfft<P, DIR, BL>(v, tw, t, L, bl, g0, sl);
        }
    } else if constexpr (K == CK_PROJ) {
        cf *X = (cf *) a.X + wb;
        const float *M = a.M + mb;
        auto g0 = [=](int e) { return at32(X, e * W + co); };
        auto pre = [=](int e) {
            if constexpr (PF) return at32(M, e * MP + co);
            else return NoPre{};
        };
        auto mid = [=](int e, cf z, auto m) {
            if constexpr (PF) return modulus_project_fused(z, m);
            else return modulus_project_fused(z, at32(M, e * MP + co));
        };
        auto sl = [=](int e, cf x) { if (valid) at32(X, e * W + co) = x; };
        xpair<P, -1, +1, BL>(v, tw, t, L, bl, g0, pre, mid, sl);
    }
}

// LDS bytes of a C-column tile
template <class P, int C, bool SPLIT>
constexpr size_t col_lds() { return (SPLIT ? sizeof(float) : sizeof(cf)) * (size_t) lds_elems(C * P::n); }

template <class P, int C, bool SPLIT, bool PF = false, bool BL = false, int STG = 0>
int launch_variant(int mode, ColArgs a, hipStream_t s)
{
    static_assert(C * P::T <= 1024, "column workgroup too large");
    static_assert(col_lds<P, C, SPLIT>() <= kLdsBudget, "column tile exceeds the LDS budget");
    const int images = a.batch > 1 ? a.batch : 1;
    const dim3 grid((a.W + C - 1) / C, images * (a.sub > 1 ? a.sub : 1)), block(C * P::T);
    constexpr size_t lds = col_lds<P, C, SPLIT>();
    if constexpr (lds > 64 * 1024) {
        static const hipError_t attr = [&] {
            hipError_t e = hipSuccess;
            for (const void *f : {(const void *) k_col<P, C, SPLIT, CK_FWD, false, BL>,
                                  (const void *) k_col<P, C, SPLIT, CK_INV, false, BL>,
                                  (const void *) k_col<P, C, SPLIT, CK_FWD_REAL, false, BL>,
                                  (const void *) k_col<P, C, SPLIT, CK_INV_REAL, false, BL>,
                                  (const void *) k_col<P, C, SPLIT, CK_PROJ, PF, BL, STG>})
                if (e == hipSuccess)
                    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
            return e;
        }();
        if (attr != hipSuccess) return from_hip(attr);
    }
    switch (mode) {
    case COL_FWD:
        IMR_LAUNCH((k_col<P, C, SPLIT, CK_FWD, false, BL>), grid, block, lds, s, a);
        break;
    case COL_REAL_FWD:
        IMR_LAUNCH((k_col<P, C, SPLIT, CK_FWD_REAL, false, BL>), grid, block, lds, s, a);
        break;
    case COL_INV:
        IMR_LAUNCH((k_col<P, C, SPLIT, CK_INV, false, BL>), grid, block, lds, s, a);
        break;
    case COL_INIT:
        a.out = a.X;
        IMR_LAUNCH((k_col<P, C, SPLIT, CK_INV_REAL, false, BL>), grid, block, lds, s, a);
        break;
    case COL_PROJ:
        IMR_LAUNCH((k_col<P, C, SPLIT, CK_PROJ, PF, BL, STG>), grid, block, lds, s, a);
        break;
    default: return IMRESH_ERR_INVALID;
    }
    return imr::launch_status(__FILE__, __LINE__);
}

// Column variants of a length (index: tile shape, then +count/2 for the
// prefetching CK_PROJ): shape 0 = E 16 with the plan's default tile; for
// long columns also 1 = E 32 (half the threads per column, 512-thread
// groups), 2 = E 16 with 4-column tiles (two workgroups per CU) and 3, 4 =
// the same with the second half of the grid started ~3.4 / ~6.8 us late.
template <int N>
constexpr bool blue_length() { return N >= 64 && N <= 2 * kMaxLine && (N & (N - 1)) == 0; }

template <int N, int... Rs>
constexpr int col_shape_count()
{
    using P16 = PlanE<16, N, Rs...>;
    constexpr bool pow2 = (N & (N - 1)) == 0;   // extra shapes for the power-of-two lengths
    if constexpr (N >= 1024 && pow2 && P16::C > 4) return 5;
    else if constexpr (N >= 1024 && pow2) return 2;
    return 1;
}

template <int N, int... Rs>
constexpr int col_variant_count() { return 2 * col_shape_count<N, Rs...>(); }

template <class P, int C, int STG = 0>
int launch_pf(bool pf, int mode, ColArgs a, hipStream_t s)
{
    return pf ? launch_variant<P, C, false, true, false, STG>(mode, a, s)
              : launch_variant<P, C, false, false, false, STG>(mode, a, s);
}

template <int N, int... Rs>
int launch_col_direct(int var, int mode, ColArgs a, hipStream_t s);

template <int N, int... Rs>
int launch_col_variant(int var, int mode, ColArgs a, hipStream_t s)
{
    using P16 = PlanE<16, N, Rs...>;
    if constexpr (blue_length<N>()) {
        // Bluestein variants: the plan's tile or half as many columns (twice
        // the workgroups, half the LDS), each plain or prefetching
        constexpr int C2 = P16::C >= 2 ? P16::C / 2 : 1;
        switch (var - kBlueVariant) {
        case 0: return launch_variant<P16, P16::C, false, false, true>(mode, a, s);
        case 1: return launch_variant<P16, C2, false, false, true>(mode, a, s);
        case 2: return launch_variant<P16, P16::C, false, true, true>(mode, a, s);
        case 3: return launch_variant<P16, C2, false, true, true>(mode, a, s);
        default: break;
        }
    }
    if constexpr (N > kMaxLine) {
        return IMRESH_ERR_UNSUPPORTED_SIZE;   // a convolution length only
    } else {
        return launch_col_direct<N, Rs...>(var, mode, a, s);
    }
}

template <int N, int... Rs>
int launch_col_direct(int var, int mode, ColArgs a, hipStream_t s)
{
    using P16 = PlanE<16, N, Rs...>;
    constexpr int ns = col_shape_count<N, Rs...>();
    const bool pf = var >= ns;
    const int shape = pf ? var - ns : var;
    if constexpr (ns > 1) {
        using P32 = PlanE<32, N, Rs...>;
        if (shape == 1) return launch_pf<P32, P32::C>(pf, mode, a, s);
        if constexpr (ns == 5) {
            if (shape == 2) return launch_pf<P16, 4>(pf, mode, a, s);
            if (shape == 3) return launch_pf<P16, 4, 1>(pf, mode, a, s);
            if (shape == 4) return launch_pf<P16, 4, 2>(pf, mode, a, s);
        }
    }
    return launch_pf<P16, P16::C>(pf, mode, a, s);
}

template <int G, int N, int... Rs>
int try_plan(int mode, int var, int n, ColArgs a, hipStream_t s)
{
    if constexpr (G == IMR_GROUP) {
        if (n == N) return launch_col_variant<N, Rs...>(var, mode, a, s);
    }
    return -1;
}

template <int G, int N, int... Rs>
int try_count(int n)
{
    if constexpr (G == IMR_GROUP) {
        if (n == N) return col_variant_count<N, Rs...>();
    }
    return -1;
}

}  // namespace

#define IMR_CAT2(a, b) a##b
#define IMR_CAT(a, b) IMR_CAT2(a, b)
#define IMR_COL_CASE(g, N, ...)                                                     \
    {                                                                               \
        const int r = try_plan<g, N, __VA_ARGS__>(mode, var, n, a, s);              \
        if (r >= 0) return r;                                                       \
    }
#define IMR_COL_COUNT(g, N, ...)                                                    \
    {                                                                               \
        const int r = try_count<g, N, __VA_ARGS__>(n);                              \
        if (r >= 0) return r;                                                       \
    }

int IMR_CAT(launch_col_g, IMR_GROUP)(int mode, int var, int n, const LinePlan &lp, ColArgs a,
                                     hipStream_t s)
{
    a.tw = lp.tw;
    IMR_PLANS(IMR_COL_CASE)
    return IMRESH_ERR_UNSUPPORTED_SIZE;
}

int IMR_CAT(col_variants_g, IMR_GROUP)(int n)
{
    IMR_PLANS(IMR_COL_COUNT)
    return 0;
}

}  // namespace imr
