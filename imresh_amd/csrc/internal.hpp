// Internal interfaces between the kernel translation unit, the planner and
// the engine (not part of the C ABI).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace imr {

// ---------------------------------------------------------------- kernels
enum RowMode {
    ROW_FWD = 0,        // in -> out, forward row transform
    ROW_INV,            // in -> out, inverse row transform
    ROW_INIT,           // X -> inverse -> X = g', G = g', Bn = |g'|
    ROW_FIRST_CUT,      // X = g'; mask = B2 < thr; G = constraint; forward -> X
    ROW_FIRST_MASK,     // X = g'; G = constraint(mask); forward -> X
    ROW_MID,            // X -> inverse -> constraint(mask, G) -> forward -> X
    ROW_LAST,           // X -> inverse -> X = g', Bn = |g'|, error partials
    ROW_LAST_PLAIN,     // X -> inverse -> X = g'
    ROW_INIT_ERR,       // ROW_INIT + error partials
    ROW_FWD_NORM,       // in -> forward row transform -> Bn = |z| (float)
    ROW_NMODES
};

enum ColMode {
    COL_FWD = 0,        // in -> out, forward column transform
    COL_INV,            // in -> out, inverse column transform
    COL_PROJ,           // X -> forward -> modulus projection -> inverse -> X
    COL_INIT,           // real M -> inverse -> X
    COL_REAL_FWD,       // real M -> forward -> out
    COL_NMODES
};

struct RowArgs {
    const float2 *in;
    float2 *out;
    float2 *X;
    float2 *G;
    uint8_t *mask;
    float *Bn;             // |g'| (RK_INVOUT) or |z| (ROW_FWD_NORM)
    const float *B2;
    const float *thr;      // device scalar: mask threshold
    float beta;
    float *perr;           // per-workgroup sum mask*|g'|
    unsigned *pcnt;        // per-workgroup sum mask
    const int *done;       // early-exit flag (NULL: never)
    int H;                 // number of rows (of one image)
    const float2 *tw;      // row-length twiddle table (set by launch_row)
    int batch;             // images (grid.y; 0 or 1: one image)
    size_t istride;        // elements between consecutive images
    int len;               // row length W (Bluestein kernels)
    const float2 *bs;      // Bluestein tables of W (set by launch_row)
};

struct ColArgs {
    const float2 *in;
    float2 *out;
    float2 *X;
    const float *M;        // modulus (COL_PROJ) or real input (COL_INIT)
    const int *done;
    int W;                 // number of columns = row pitch
    const float2 *tw;      // column-length twiddle table (set by launch_col)
    int batch;             // images (grid.y; 0 or 1: one image)
    size_t istride;        // elements between consecutive images
    int len;               // column length H (Bluestein kernels)
    const float2 *bs;      // Bluestein tables of H (set by launch_col)
    // sub-blocks (the middle step of a four-step column transform): grid.y
    // = images x sub; block s of image b starts at b * istride + s *
    // sub_xstride in X and at b * istride + s * sub_mstride in M, whose rows
    // are mpitch elements apart.  sub 0 / 1 and mpitch 0: one block, M rows W
    // apart.
    int sub;
    size_t sub_xstride, sub_mstride;
    int mpitch;
};

// the radix-8 outer steps of a four-step column transform of length H =
// 8 * H2 (k_step8.hip): X of an image viewed as [8][H2][W]
struct Step8Args {
    float2 *X;
    const int *done;      // early-exit flags of the images (NULL: never)
    const float2 *tw;     // exp(-2 pi i j / H), j < H
    int W, H2;
    size_t istride;       // elements between consecutive images
};
// forward step (DIR -1: 8-point DFTs over rows n2 + H2*n1, then twiddles
// w^(n2*k1)) or inverse step (DIR +1: conjugate twiddles, then 8-point DFTs)
int launch_step8(int dir, const Step8Args &a, int batch, hipStream_t s);

struct BlurArgs {
    const float *in;
    float *out;
    const float *wh;       // horizontal weights (2*rh+1)
    const float *wv;       // vertical weights (2*rv+1)
    int rh, rv;
    int h_identity, v_identity;
    int W, H;
    float *pmax;           // per-workgroup max (NULL: none)
    const int *done;
    int batch;             // images (grid.z; 0 or 1: one image)
    size_t istride;        // elements between consecutive images
    // non-NULL: the weights of image b are wh/wv + kBlurWeightRow * cycles[b],
    // cycles[b] = completed shrink-wrap cycles (one weights row per cycle)
    const unsigned *cycles;
    int xcd_tiles;         // set by launch_blur: tiles in contiguous ranges per XCD
};

// per-image device scalars of a batch: kScalStride floats {threshold, error,
// total, masked} and kFlagStride ints {done, cycles}
constexpr int kScalStride = 4;
constexpr int kFlagStride = 2;
constexpr int kBlurWeightRow = 64;
// double scratch of the Poisson measurement simulation: per-workgroup sums
// + total
constexpr int kPoissonSumBlocks = 256;
constexpr int kAuxDoubles = kPoissonSumBlocks + 1;

// A compile-time plan (plans.hpp) of one line length with its twiddles.
// A line length is transformed either by its own compiled plan or, when it
// has none, by the Bluestein kernels of the power-of-two convolution length
// m >= 2n-1 (group, T, L, C and tw then describe the length-m plan).
struct LinePlan {
    int n;
    int group;       // translation-unit group of the kernel instantiations
    int T;           // threads per line
    int L;           // rows per row workgroup
    int C;           // columns per column workgroup
    const float2 *tw;
    int blue;        // 1: Bluestein
    int m;           // convolution length (Bluestein)
    const float2 *bs;   // Bluestein tables (device)
    // four-step column pass (n = 8 * fs_h2, see launch_col): 0 = off
    int fs_h2;
    const float2 *fs_tw;   // exp(-2 pi i j / n), j < n (device)
};
// column lengths at or above this use the four-step column pass for the
// fused projection: their own tile holds only 2 columns (16-byte rows)
constexpr int kFourStepMinLine = 8192;

// variant index of the Bluestein kernels: kBlueVariant + {0: the plan's
// column tile, 1: half as many columns per workgroup, 2 / 3: the same with
// the prefetching junction}; rows use kBlueVariant + {0: the plan's rows
// per workgroup, 1: prefetching, 2: half as many rows}.  Sizes without a
// plan are not tuned one by one (a server meets hundreds of them):
// half-width column tiles measured fastest on every Bluestein size tried
// (300-2039: 1-1.4x), and half-height row groups spread a mid-size image
// over more CUs (700: row pass 19.5 -> 17.6 us), so both are fixed choices.
constexpr int kBlueVariant = 1000;
constexpr int kBlueColDefault = 1;
constexpr int kBlueRowDefault = 2;
// longest line length: lengths without a compiled plan up to it run the
// Bluestein kernels of a power-of-two convolution length up to 2 kMaxLine;
// plans longer than kMaxLine exist only as such convolution lengths
constexpr int kMaxLine = 8192;

struct Plan2D {
    int W, H;
    LinePlan row;    // n = W
    LinePlan col;    // n = H
    int rowVar;      // kernel variant of the row pass (autotuned)
    int colVar;      // kernel variant of the column pass (autotuned)
};

int row_variants(const LinePlan &lp);
int col_variants(const LinePlan &lp);
void set_tuned_variants(int device, int W, int H, int rowVar, int colVar);
bool get_tuned_variants(int device, int W, int H, int *rowVar, int *colVar);

int get_line_plan(int device, int n, LinePlan *out);
int get_plan2d(int device, int W, int H, Plan2D *out);
// the plan of W x H if both line plans already exist on the device (no
// allocation; IMRESH_ERR_NOT_INITIALIZED otherwise)
int find_plan2d(int device, int W, int H, Plan2D *out);
bool line_supported(int n);

int launch_row(int mode, const Plan2D &p, RowArgs a, hipStream_t s);
int launch_col(int mode, const Plan2D &p, ColArgs a, hipStream_t s);
// per translation-unit group (k_row.hip / k_col.hip compiled once per group)
#define IMR_DECLARE_GROUP(g)                                                                  \
    int launch_row_g##g(int mode, int var, int n, const LinePlan &lp, RowArgs a, hipStream_t s); \
    int launch_col_g##g(int mode, int var, int n, const LinePlan &lp, ColArgs a, hipStream_t s); \
    int row_variants_g##g(int n);                                                            \
    int row_rows_g##g(int n, int var);                                                       \
    int col_variants_g##g(int n);
IMR_DECLARE_GROUP(0)
IMR_DECLARE_GROUP(1)
IMR_DECLARE_GROUP(2)
IMR_DECLARE_GROUP(3)
IMR_DECLARE_GROUP(4)
IMR_DECLARE_GROUP(5)
IMR_DECLARE_GROUP(6)
IMR_DECLARE_GROUP(7)
// number of row workgroups of the plan's row variant
int row_grid(const Plan2D &p);
int blur_grid(int W, int H);
int launch_blur(const BlurArgs &a, hipStream_t s);
// sets the blur kernel's LDS attribute for radius R (before a graph capture)
int prepare_blur(int R);
int launch_blur_h_inplace(float *d, int W, int H, const float *w, int r, hipStream_t s);
int launch_max_finalize_batch(const float *pmax, int np, float cutA, float cut, float *scal,
                              int *flags, int batch, hipStream_t s);
int launch_err_finalize_batch(const float *perr, const unsigned *pcnt, int np, float target,
                              float *scal, int *flags, int batch, hipStream_t s);
int launch_max_finalize(const float *pmax, int np, float cut, float *thr, const int *done,
                        hipStream_t s);
int launch_err_finalize(const float *perr, const unsigned *pcnt, int np, float target,
                        float *res, int *done, unsigned *cycles, hipStream_t s);
int launch_copy_from_real(float *out, const float2 *in, size_t n, hipStream_t s);
int launch_random_phase(float2 *X, const float *M, size_t n, hipStream_t s);
// measurement simulation: modulus |F| -> sqrt(Poisson(photons |F|^2 / sum |F|^2)),
// aux = kAuxDoubles doubles of device scratch
int launch_poisson_modulus(float *m, size_t n, double photons, uint64_t seed, double *aux, hipStream_t s);
int launch_fft_shift(void *d, unsigned nx, unsigned ny, unsigned elem_bytes, hipStream_t s);
int launch_gather_images(float *dst, const float *const *src, size_t n, int batch, hipStream_t s);
int launch_scatter_real(float *const *dst, const float2 *X, size_t n, int batch, hipStream_t s);
int launch_copy_to_real(float2 *out, const float *in, size_t n, hipStream_t s);
int launch_complex_norm(float *out, const float2 *in, size_t n, hipStream_t s);
int launch_apply_modulus(float2 *out, const float2 *in, const float *M, size_t n, hipStream_t s);
int launch_hio_constraint(float2 *gp, const float2 *g, const float *mask, size_t n, float beta,
                          hipStream_t s);
int launch_cutoff(float *d, size_t n, float t, float lo, float hi, hipStream_t s);
// op: 0 = max, 1 = min, 2 = sum.  result: device float.
int launch_reduce(int op, const float *d, size_t n, float *partials, int np, float *result,
                  hipStream_t s);
// mask kinds of launch_hio_error
constexpr int kMaskFloat = 0, kMaskU8 = 1, kMaskBits = 2;
int launch_hio_error(const float2 *g, const void *mask, int mask_kind, size_t n, int invert,
                     float *perr, unsigned *pcnt, int np, float *res, hipStream_t s);
// byte mask -> bit-packed mask (32 elements per word, element 32w+j at bit 31-j)
int launch_pack_mask(const uint8_t *mask, size_t n, uint32_t *bits, hipStream_t s);
int launch_fill_state(float2 *X, float2 *G, uint8_t *mask, float *M, float *Bn, size_t n,
                      int W, hipStream_t s);
// n floats from host memory (src NULL: zeros) to device memory through
// kernel arguments, kStoreWords per launch: no staging buffer, capturable
constexpr int kStoreWords = 480;
int launch_store_words(float *dst, const float *src, size_t n, hipStream_t s);

constexpr int kReduceBlocks = 512;
constexpr int kBlurTX = 64;
constexpr int kBlurTY = 64;
constexpr int kMaxBlurRadius = 31;   // 63 taps

// host Gaussian weights: calcGaussianKernel semantics (calcGaussianKernel.cpp:42-91)
int gaussian_weights(double sigma, float *w, unsigned capacity);
// sigma_{c+1} = max(1.5, (1 - dsigma) * sigma_c) in float (shrinkWrap.cpp:245)
void sigma_schedule(float sigma0, float dsigma, unsigned n, std::vector<float> &out);

}  // namespace imr
