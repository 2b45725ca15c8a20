// Host-side planning: the compile-time plan registry (plans.hpp) viewed from
// the host, twiddle tables per device and length (cached for the process),
// kernel dispatch by plan group, and the Gaussian weights / sigma schedule
// of the reference (computed on the host exactly as the reference does, so
// the device blur uses bit-identical weights).
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "internal.hpp"
#include "plans.hpp"

namespace imr {

thread_local ProfEvents g_prof;

namespace {

std::recursive_mutex g_plan_mtx;   // get_line_plan recurses for Bluestein lengths
std::map<std::pair<int, int>, LinePlan> g_lines;   // (device, n) -> plan (twiddles on device)

template <int G, int N, int... Rs>
PlanInfo make_info()
{
    using P = Plan<N, Rs...>;
    PlanInfo pi{};
    pi.group = G;
    pi.n = N;
    pi.np = P::np;
    pi.T = P::T;
    pi.L = P::L;
    pi.C = P::C;
    for (int i = 0; i < P::np; ++i) pi.rad[i] = P::rf[i];
    return pi;
}

#define IMR_INFO_ENTRY(g, N, ...) make_info<g, N, __VA_ARGS__>(),

const std::vector<PlanInfo> &registry()
{
    static const std::vector<PlanInfo> r = {IMR_PLANS(IMR_INFO_ENTRY)};
    return r;
}

const PlanInfo *find_info(int n)
{
    for (const auto &pi : registry())
        if (pi.n == n) return &pi;
    return nullptr;
}

// exp(-2 pi i r k / (ns R)) for every pass of a radix sequence, evaluated in
// double; layout as Plan::twoff (k-major, r = 1..R-1)
void append_twiddles(const std::vector<int> &rad, std::vector<float2> &tw)
{
    int ns = 1;
    for (int R : rad) {
        for (int k = 0; k < ns; ++k)
            for (int r = 1; r < R; ++r) {
                const double a = -2.0 * M_PI * (double) r * (double) k / ((double) ns * R);
                tw.push_back(make_float2((float) std::cos(a), (float) std::sin(a)));
            }
        ns *= R;
    }
}

// forward DFT in double of a length whose prime factors are small (the
// compiled convolution lengths), recursive mixed radix; for the Bluestein
// tables
void dft_any(std::vector<std::complex<double>> &a)
{
    const int m = (int) a.size();
    if (m <= 1) return;
    int p = 2;
    while (p * p <= m && m % p != 0) ++p;
    if (m % p != 0) p = m;   // prime length: direct sum
    const int q = m / p;
    if (q == 1) {
        std::vector<std::complex<double>> out(m);
        for (int k = 0; k < m; ++k) {
            std::complex<double> s = 0;
            for (int j = 0; j < m; ++j) {
                const double ang = -2.0 * M_PI * (double) ((long long) j * k % m) / (double) m;
                s += a[j] * std::complex<double>(std::cos(ang), std::sin(ang));
            }
            out[k] = s;
        }
        a.swap(out);
        return;
    }
    std::vector<std::vector<std::complex<double>>> sub(p, std::vector<std::complex<double>>(q));
    for (int r = 0; r < p; ++r)
        for (int j = 0; j < q; ++j) sub[r][j] = a[(size_t) j * p + r];
    for (auto &v : sub) dft_any(v);
    for (int k = 0; k < q; ++k)
        for (int s2 = 0; s2 < p; ++s2) {
            const int kk = k + q * s2;
            std::complex<double> acc = 0;
            for (int r = 0; r < p; ++r) {
                const double ang = -2.0 * M_PI * (double) ((long long) r * kk % m) / (double) m;
                acc += sub[r][k] * std::complex<double>(std::cos(ang), std::sin(ang));
            }
            a[kk] = acc;
        }
}

// convolution length of a Bluestein line: the smallest power of two
// >= max(64, 2n-1).  (Smooth lengths 3*2^k pad less but measured slower per
// image on gfx950, with their plans ending in a radix-12 or radix-6 pass as
// well as with radix-3 passes.)
int blue_m(int n)
{
    int m = 64;
    while (m < 2 * n - 1) m <<= 1;
    return find_info(m) != nullptr ? m : 0;
}

// Bluestein tables of length n (layout in blue_device.hpp), float from double
std::vector<float2> blue_tables(int n, int m)
{
    std::vector<std::complex<double>> a(n), bf(m, 0.0), bi(m, 0.0);
    for (int j = 0; j < n; ++j) {
        // j^2 mod 2n keeps the angle exact for large j
        const long long q = ((long long) j * j) % (2LL * n);
        const double ang = -M_PI * (double) q / (double) n;
        a[j] = std::complex<double>(std::cos(ang), std::sin(ang));
    }
    for (int j = 0; j < n; ++j) {
        bf[j] = std::conj(a[j]);
        bi[j] = a[j];
        if (j > 0) {
            bf[m - j] = std::conj(a[j]);
            bi[m - j] = a[j];
        }
    }
    dft_any(bf);
    dft_any(bi);
    std::vector<float2> out;
    out.reserve((size_t) n + 2 * m);
    for (int j = 0; j < n; ++j) out.push_back(make_float2((float) a[j].real(), (float) a[j].imag()));
    for (int k = 0; k < m; ++k) out.push_back(make_float2((float) (bf[k].real() / m), (float) (bf[k].imag() / m)));
    for (int k = 0; k < m; ++k) out.push_back(make_float2((float) (bi[k].real() / m), (float) (bi[k].imag() / m)));
    return out;
}

// copies a host table into a new allocation on `device`, leaving the calling
// thread's current device as it was
int upload_table(int device, const std::vector<float2> &v, float2 **out)
{
    *out = nullptr;
    int prev = 0;
    IMR_HIP_CHECK(hipGetDevice(&prev));
    IMR_HIP_CHECK(hipSetDevice(device));
    float2 *p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(float2) * v.size());
    if (e == hipSuccess) {
        e = hipMemcpy(p, v.data(), sizeof(float2) * v.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) (void) hipFree(p);
    }
    const hipError_t r = hipSetDevice(prev);
    if (e != hipSuccess) return from_hip(e);
    if (r != hipSuccess) {
        (void) hipFree(p);
        return from_hip(r);
    }
    *out = p;
    return IMRESH_OK;
}

}  // namespace

bool line_supported(int n) { return n >= 1 && n <= kMaxLine; }

int get_line_plan(int device, int n, LinePlan *out)
{
    std::lock_guard<std::recursive_mutex> lk(g_plan_mtx);
    auto it = g_lines.find({device, n});
    if (it != g_lines.end()) {
        *out = it->second;
        return IMRESH_OK;
    }
    const PlanInfo *pi = find_info(n);
    if (pi == nullptr) {
        if (n < 1 || n > kMaxLine) return IMRESH_ERR_UNSUPPORTED_SIZE;
        // Bluestein: the convolution plan's kernels and twiddles + the tables
        const int m = blue_m(n);
        if (m == 0) return IMRESH_ERR_UNSUPPORTED_SIZE;
        LinePlan lm;
        IMR_CHECK(get_line_plan(device, m, &lm));
        float2 *dev = nullptr;
        IMR_CHECK(upload_table(device, blue_tables(n, m), &dev));
        LinePlan lp = lm;
        lp.n = n;
        lp.blue = 1;
        lp.m = m;
        lp.bs = dev;
        // the four-step column pass belongs to columns of the convolution
        // length itself: a Bluestein column of length n < m keeps the
        // one-pass kernels (the four-step outer steps would address m rows)
        lp.fs_h2 = 0;
        lp.fs_tw = nullptr;
        g_lines[{device, n}] = lp;
        *out = lp;
        return IMRESH_OK;
    }
    std::vector<int> fwd(pi->rad, pi->rad + pi->np), inv(fwd.rbegin(), fwd.rend());
    std::vector<float2> tw;
    append_twiddles(fwd, tw);
    append_twiddles(inv, tw);
    tw.push_back(make_float2(0.f, 0.f));   // never empty
    float2 *dev = nullptr;
    IMR_CHECK(upload_table(device, tw, &dev));
    LinePlan lp{};
    lp.n = n;
    lp.group = pi->group;
    lp.T = pi->T;
    lp.L = pi->L;
    lp.C = pi->C;
    lp.tw = dev;
    if (n >= kFourStepMinLine && n <= kMaxLine && n % 8 == 0 && find_info(n / 8) != nullptr) {
        // four-step column pass: its length-n/8 plan (created now: the pass
        // may first run inside a graph capture, where no allocation or copy
        // may happen) and the twiddles exp(-2 pi i j / n) of its outer steps
        LinePlan sub;
        IMR_CHECK(get_line_plan(device, n / 8, &sub));
        std::vector<float2> t(n);
        for (int j = 0; j < n; ++j) {
            const double ang = -2.0 * M_PI * (double) j / (double) n;
            t[j] = make_float2((float) std::cos(ang), (float) std::sin(ang));
        }
        float2 *fs = nullptr;
        IMR_CHECK(upload_table(device, t, &fs));
        lp.fs_h2 = n / 8;
        lp.fs_tw = fs;
    }
    g_lines[{device, n}] = lp;
    *out = lp;
    return IMRESH_OK;
}

namespace {
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_tuned;   // (device, W, H) -> variants
}

void set_tuned_variants(int device, int W, int H, int rowVar, int colVar)
{
    std::lock_guard<std::recursive_mutex> lk(g_plan_mtx);
    g_tuned[{device, W, H}] = {rowVar, colVar};
}

bool get_tuned_variants(int device, int W, int H, int *rowVar, int *colVar)
{
    std::lock_guard<std::recursive_mutex> lk(g_plan_mtx);
    auto it = g_tuned.find({device, W, H});
    if (it == g_tuned.end()) return false;
    *rowVar = it->second.first;
    *colVar = it->second.second;
    return true;
}

int get_plan2d(int device, int W, int H, Plan2D *out)
{
    if (W < 1 || H < 1) return IMRESH_ERR_INVALID;
    if (!line_supported(W) || !line_supported(H)) return IMRESH_ERR_UNSUPPORTED_SIZE;
    Plan2D p{};
    p.W = W;
    p.H = H;
    IMR_CHECK(get_line_plan(device, W, &p.row));
    IMR_CHECK(get_line_plan(device, H, &p.col));
    if (!get_tuned_variants(device, W, H, &p.rowVar, &p.colVar)) p.rowVar = p.colVar = 0;
    *out = p;
    return IMRESH_OK;
}

int find_plan2d(int device, int W, int H, Plan2D *out)
{
    if (W < 1 || H < 1) return IMRESH_ERR_INVALID;
    if (!line_supported(W) || !line_supported(H)) return IMRESH_ERR_UNSUPPORTED_SIZE;
    std::lock_guard<std::recursive_mutex> lk(g_plan_mtx);
    auto r = g_lines.find({device, W}), c = g_lines.find({device, H});
    if (r == g_lines.end() || c == g_lines.end()) return IMRESH_ERR_NOT_INITIALIZED;
    Plan2D p{};
    p.W = W;
    p.H = H;
    p.row = r->second;
    p.col = c->second;
    auto t = g_tuned.find({device, W, H});
    p.rowVar = t != g_tuned.end() ? t->second.first : 0;
    p.colVar = t != g_tuned.end() ? t->second.second : 0;
    *out = p;
    return IMRESH_OK;
}

int row_grid(const Plan2D &p)
{
    int L = 1;
    const int n = p.row.blue ? p.row.m : p.W;
    const int v = p.row.blue ? kBlueVariant + kBlueRowDefault : p.rowVar;
    switch (p.row.group) {
    case 0: L = row_rows_g0(n, v); break;
    case 1: L = row_rows_g1(n, v); break;
    case 2: L = row_rows_g2(n, v); break;
    case 3: L = row_rows_g3(n, v); break;
    case 4: L = row_rows_g4(n, v); break;
    case 5: L = row_rows_g5(n, v); break;
    case 6: L = row_rows_g6(n, v); break;
    case 7: L = row_rows_g7(n, v); break;
    default: break;
    }
    return (p.H + L - 1) / L;
}

int launch_row(int mode, const Plan2D &p, RowArgs a, hipStream_t s)
{
    a.H = p.H;
    a.len = p.W;
    a.bs = p.row.bs;
    const int n = p.row.blue ? p.row.m : p.W;
    const int v = p.row.blue ? kBlueVariant + kBlueRowDefault : p.rowVar;
    switch (p.row.group) {
    case 0: return launch_row_g0(mode, v, n, p.row, a, s);
    case 1: return launch_row_g1(mode, v, n, p.row, a, s);
    case 2: return launch_row_g2(mode, v, n, p.row, a, s);
    case 3: return launch_row_g3(mode, v, n, p.row, a, s);
    case 4: return launch_row_g4(mode, v, n, p.row, a, s);
    case 5: return launch_row_g5(mode, v, n, p.row, a, s);
    case 6: return launch_row_g6(mode, v, n, p.row, a, s);
    case 7: return launch_row_g7(mode, v, n, p.row, a, s);
    default: return IMRESH_ERR_UNSUPPORTED_SIZE;
    }
}

namespace {

// The fused column pass of a four-step length H = 8 * H2: the forward outer
// step, the existing fused forward/projection/inverse kernel of length H2
// on the 8 blocks of H2 rows of every image (block k1 holds frequencies
// k1 + 8 k2, so its modulus rows are k1 + 8 k2: M offset k1 rows, pitch 8
// rows), the inverse outer step.  52 bytes per element instead of 20, but
// every access covers whole 64-column row segments.
int launch_col_fourstep(const Plan2D &p, ColArgs a, hipStream_t s)
{
    const int h2 = p.col.fs_h2;
    const int images = a.batch > 1 ? a.batch : 1;
    const size_t istride = images > 1 ? a.istride : (size_t) p.W * p.H;
    Step8Args sa{};
    sa.X = a.X;
    sa.done = a.done;
    sa.tw = p.col.fs_tw;
    sa.W = p.W;
    sa.H2 = h2;
    sa.istride = istride;
    IMR_CHECK(launch_step8(-1, sa, images, s));
    int dev = 0;
    IMR_HIP_CHECK(hipGetDevice(&dev));
    Plan2D q = p;
    q.H = h2;
    IMR_CHECK(get_line_plan(dev, h2, &q.col));
    q.colVar = col_variants(q.col) / 2;   // the prefetching default tile
    ColArgs b = a;
    b.batch = images;
    b.sub = 8;
    b.istride = istride;
    b.sub_xstride = (size_t) h2 * p.W;
    b.sub_mstride = (size_t) p.W;
    b.mpitch = 8 * p.W;
    IMR_CHECK(launch_col(COL_PROJ, q, b, s));
    return launch_step8(+1, sa, images, s);
}

}  // namespace

int launch_col(int mode, const Plan2D &p, ColArgs a, hipStream_t s)
{
    if (mode == COL_PROJ && p.col.fs_h2 > 0 && a.sub <= 1) return launch_col_fourstep(p, a, s);
    a.W = p.W;
    a.len = p.H;
    a.bs = p.col.bs;
    const int n = p.col.blue ? p.col.m : p.H;
    const int v = p.col.blue ? kBlueVariant + kBlueColDefault : p.colVar;
    switch (p.col.group) {
    case 0: return launch_col_g0(mode, v, n, p.col, a, s);
    case 1: return launch_col_g1(mode, v, n, p.col, a, s);
    case 2: return launch_col_g2(mode, v, n, p.col, a, s);
    case 3: return launch_col_g3(mode, v, n, p.col, a, s);
    case 4: return launch_col_g4(mode, v, n, p.col, a, s);
    case 5: return launch_col_g5(mode, v, n, p.col, a, s);
    case 6: return launch_col_g6(mode, v, n, p.col, a, s);
    case 7: return launch_col_g7(mode, v, n, p.col, a, s);
    default: return IMRESH_ERR_UNSUPPORTED_SIZE;
    }
}

int row_variants(const LinePlan &lp)
{
    if (lp.blue) return 1;   // Bluestein: fixed choice (no per-size tuning, see kBlueColDefault)
    switch (lp.group) {
    case 0: return row_variants_g0(lp.n);
    case 1: return row_variants_g1(lp.n);
    case 2: return row_variants_g2(lp.n);
    case 3: return row_variants_g3(lp.n);
    case 4: return row_variants_g4(lp.n);
    case 5: return row_variants_g5(lp.n);
    case 6: return row_variants_g6(lp.n);
    case 7: return row_variants_g7(lp.n);
    default: return 0;
    }
}

int col_variants(const LinePlan &lp)
{
    if (lp.blue) return 1;
    switch (lp.group) {
    case 0: return col_variants_g0(lp.n);
    case 1: return col_variants_g1(lp.n);
    case 2: return col_variants_g2(lp.n);
    case 3: return col_variants_g3(lp.n);
    case 4: return col_variants_g4(lp.n);
    case 5: return col_variants_g5(lp.n);
    case 6: return col_variants_g6(lp.n);
    case 7: return col_variants_g7(lp.n);
    default: return 0;
    }
}

// calcGaussianKernel.cpp:42-91: radius ceil(2.884402748387961466*sigma - 0.5),
// float weights a*exp(i^2*b) summed in double, each divided by the sum.
int gaussian_weights(double sigma, float *w, unsigned capacity)
{
    const int nNeighbors = (int) std::ceil(2.884402748387961466 * sigma - 0.5);
    const int nWeights = 2 * nNeighbors + 1;
    if (w == nullptr || (unsigned) nWeights > capacity) return nWeights;
    double sum = 0;
    const double a = 1.0 / (std::sqrt(2.0 * M_PI) * sigma);
    const double b = -1.0 / (2.0 * sigma * sigma);
    for (int i = -nNeighbors; i <= nNeighbors; ++i) {
        const float weight = (float) (a * std::exp(i * i * b));
        w[nNeighbors + i] = weight;
        sum += weight;
    }
    for (int i = -nNeighbors; i <= nNeighbors; ++i)
        w[nNeighbors + i] = (float) ((double) w[nNeighbors + i] / sum);
    return nWeights;
}

void sigma_schedule(float sigma0, float dsigma, unsigned n, std::vector<float> &out)
{
    out.resize(n);
    float sigma = sigma0;
    for (unsigned i = 0; i < n; ++i) {
        out[i] = sigma;
        sigma = std::fmax(1.5f, (1.0f - dsigma) * sigma);
    }
}

}  // namespace imr
