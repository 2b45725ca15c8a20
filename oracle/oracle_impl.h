/*
 * Type-generic body of the CPU oracle (TEST INFRASTRUCTURE ONLY; see
 * oracle.h).  Included twice by oracle.c with
 *   T  = float  / double
 *   S  = f32    / f64
 * FN(x) expands to x_S.  Every loop keeps the reference's evaluation order so
 * that, for float, results are bit-identical to the reference CPU path
 * (compiled without FMA contraction, as g++ -O2 does on x86-64).
 */

typedef struct { T re, im; } FN(cplx);

/* ------------------------------------------------------------------------ */
/* calcGaussianKernel  (src/imresh/libs/calcGaussianKernel.cpp:42-91)       */
/* ------------------------------------------------------------------------ */
int FN(orc_gauss_kernel)(double sigma, T *w, unsigned nmax)
{
    const int nNeighbors = (int) ceil(2.884402748387961466 * sigma - 0.5);
    const int nWeights = 2 * nNeighbors + 1;
    if (w == NULL || (unsigned) nWeights > nmax)
        return nWeights;
    double sum = 0;
    const double a = 1.0 / (sqrt(2.0 * M_PI) * sigma);
    const double b = -1.0 / (2.0 * sigma * sigma);
    for (int i = -nNeighbors; i <= nNeighbors; ++i) {
        const T weight = (T) (a * exp(i * i * b));
        w[nNeighbors + i] = weight;
        sum += weight;
    }
    for (int i = -nNeighbors; i <= nNeighbors; ++i)
        w[nNeighbors + i] = (T) ((double) w[nNeighbors + i] / sum);
    return nWeights;
}

/* ------------------------------------------------------------------------ */
/* Gaussian blur, clamp-to-edge, taps summed w[0]*x[i-r] ... w[2r]*x[i+r]    */
/* from sum = 0 (gaussian.cpp:51-163 applyKernel, :559-734 vertical).        */
/* ------------------------------------------------------------------------ */
static void FN(blur_line)(T *d, unsigned n, size_t stride, const T *w, int r, T *tmp)
{
    for (unsigned i = 0; i < n; ++i)
        tmp[i] = d[(size_t) i * stride];
    for (unsigned i = 0; i < n; ++i) {
        T sum = 0;
        for (int t = 0; t <= 2 * r; ++t) {
            int j = (int) i - r + t;
            if (j < 0) j = 0;
            if (j > (int) n - 1) j = (int) n - 1;
            sum += tmp[j] * w[t];
        }
        d[(size_t) i * stride] = sum;
    }
}

void FN(orc_blur_h)(T *d, unsigned w, unsigned h, double sigma)
{
    T kern[64];
    const int k = FN(orc_gauss_kernel)(sigma, kern, 64);
    if (k > 64) abort();
    const int r = (k - 1) / 2;
    #pragma omp parallel
    {
        T *tmp = (T *) malloc(sizeof(T) * (w ? w : 1));
        #pragma omp for
        for (int y = 0; y < (int) h; ++y)
            FN(blur_line)(d + (size_t) y * w, w, 1, kern, r, tmp);
        free(tmp);
    }
}

void FN(orc_blur_v)(T *d, unsigned w, unsigned h, double sigma)
{
    T kern[64];
    const int k = FN(orc_gauss_kernel)(sigma, kern, 64);
    if (k > 64) abort();
    const int r = (k - 1) / 2;
    #pragma omp parallel
    {
        T *tmp = (T *) malloc(sizeof(T) * (h ? h : 1));
        #pragma omp for
        for (int x = 0; x < (int) w; ++x)
            FN(blur_line)(d + x, h, w, kern, r, tmp);
        free(tmp);
    }
}

/* gaussian.cpp:737-749: horizontal, then vertical, in place */
void FN(orc_blur)(T *d, unsigned w, unsigned h, double sigma)
{
    FN(orc_blur_h)(d, w, h, sigma);
    FN(orc_blur_v)(d, w, h, sigma);
}

/* ------------------------------------------------------------------------ */
/* reductions (vectorReduce.cpp:73-119); sequential order                    */
/* ------------------------------------------------------------------------ */
T FN(orc_vmax)(const T *d, unsigned n)
{
    T m = -(T) (sizeof(T) == 4 ? FLT_MAX : DBL_MAX);   /* numeric_limits::lowest */
    for (unsigned i = 0; i < n; ++i)
        m = (m < d[i]) ? d[i] : m;                     /* std::max(m, d[i]) */
    return m;
}

T FN(orc_vmin)(const T *d, unsigned n)
{
    T m = (T) (sizeof(T) == 4 ? FLT_MAX : DBL_MAX);
    for (unsigned i = 0; i < n; ++i)
        m = (d[i] < m) ? d[i] : m;                     /* std::min(m, d[i]) */
    return m;
}

T FN(orc_vsum)(const T *d, unsigned n)
{
    T s = 0;
    for (unsigned i = 0; i < n; ++i)
        s += d[i];
    return s;
}

/* ------------------------------------------------------------------------ */
/* elementwise (vectorElementwise.cpp:41-81; cudaVectorElementwise.cu)       */
/* ------------------------------------------------------------------------ */
void FN(orc_complex_norm)(T *out, const T *c, unsigned n)
{
    #pragma omp parallel for
    for (int i = 0; i < (int) n; ++i) {
        const T re = c[2 * (size_t) i], im = c[2 * (size_t) i + 1];
        out[i] = SQRT(re * re + im * im);
    }
}

void FN(orc_apply_modulus)(T *c, const T *modulus, unsigned n)
{
    #pragma omp parallel for
    for (int i = 0; i < (int) n; ++i) {
        const T re = c[2 * (size_t) i], im = c[2 * (size_t) i + 1];
        T norm = SQRT(re * re + im * im);
        if (norm == 0)
            norm = 1;
        const T factor = modulus[i] / norm;
        c[2 * (size_t) i]     = re * factor;
        c[2 * (size_t) i + 1] = im * factor;
    }
}

/* shrinkWrap.cpp:263-276 == cudaVectorElementwise.cu:42-72 */
void FN(orc_hio_constraint)(T *gprev, const T *gprime, const T *mask, unsigned n, T beta)
{
    #pragma omp parallel for
    for (int i = 0; i < (int) n; ++i) {
        const T re = gprime[2 * (size_t) i], im = gprime[2 * (size_t) i + 1];
        if (mask[i] == 1 || re < 0) {
            gprev[2 * (size_t) i]     -= beta * re;
            gprev[2 * (size_t) i + 1] -= beta * im;
        } else {
            gprev[2 * (size_t) i]     = re;
            gprev[2 * (size_t) i + 1] = im;
        }
    }
}

/* shrinkWrap.cpp:240-242 == cudaVectorElementwise.cu:198-219 */
void FN(orc_cutoff)(T *d, unsigned n, T threshold, T lower, T upper)
{
    #pragma omp parallel for
    for (int i = 0; i < (int) n; ++i)
        d[i] = d[i] < threshold ? lower : upper;
}

void FN(orc_copy_to_real)(T *c, const T *x, unsigned n)
{
    for (unsigned i = 0; i < n; ++i) {
        c[2 * (size_t) i] = x[i];
        c[2 * (size_t) i + 1] = 0;
    }
}

void FN(orc_copy_from_real)(T *x, const T *c, unsigned n)
{
    for (unsigned i = 0; i < n; ++i)
        x[i] = c[2 * (size_t) i];
}

/* hybridInputOutput.cpp:67-103.  The float build accumulates in float in
 * sequential order (the reference uses an OpenMP reduction whose order
 * depends on the thread count); the double build accumulates in double. */
float FN(orc_hio_error)(const T *gprime, const T *mask, unsigned n, int invert,
                        float *total, float *nmasked)
{
    T totalError = 0, nMasked = 0;
    for (unsigned i = 0; i < n; ++i) {
        const T re = gprime[2 * (size_t) i], im = gprime[2 * (size_t) i + 1];
        T shouldBeZero = mask[i];
        if (invert)
            shouldBeZero = 1 - shouldBeZero;
        totalError += shouldBeZero * SQRT(re * re + im * im);
        nMasked += shouldBeZero;
    }
    if (total) *total = (float) totalError;
    if (nmasked) *nmasked = (float) nMasked;
    return (float) (SQRT(totalError) / nMasked);
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time, Bluestein for lengths with */
/* a prime factor > ORC_MAX_RADIX.  Unnormalized; sign -1 = forward.          */
/* Twiddles are evaluated in double and rounded to T.                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    unsigned n;
    int nfac;
    unsigned fac[40];
    FN(cplx) *tw;      /* tw[k] = exp(sign*2*pi*i*k/n), k < n */
} FN(fftplan);

static int FN(fft_factor)(unsigned n, unsigned *fac)
{
    int nf = 0;
    while (n % 4 == 0) { fac[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { fac[nf++] = 2; n /= 2; }
    for (unsigned p = 3; p * p <= n; p += 2)
        while (n % p == 0) { fac[nf++] = p; n /= p; }
    if (n > 1) fac[nf++] = n;
    return nf;
}

static void FN(fft_rec)(const FN(cplx) *in, size_t is, FN(cplx) *out, unsigned n,
                        const unsigned *fac, const FN(fftplan) *pl, unsigned twstride)
{
    if (n == 1) { out[0] = in[0]; return; }
    const unsigned p = fac[0], m = n / p;
    for (unsigned q = 0; q < p; ++q)
        FN(fft_rec)(in + q * is, is * p, out + (size_t) q * m, m, fac + 1, pl, twstride * p);
    const unsigned N = pl->n;
    FN(cplx) t[ORC_MAX_RADIX + 1];
    for (unsigned k = 0; k < m; ++k) {
        for (unsigned q = 0; q < p; ++q) {
            const FN(cplx) x = out[(size_t) q * m + k];
            const FN(cplx) wq = pl->tw[(size_t) q * k * twstride];
            t[q].re = x.re * wq.re - x.im * wq.im;
            t[q].im = x.re * wq.im + x.im * wq.re;
        }
        for (unsigned s = 0; s < p; ++s) {
            T ar = 0, ai = 0;
            for (unsigned q = 0; q < p; ++q) {
                const FN(cplx) wp = pl->tw[(size_t) ((q * s) % p) * (N / p)];
                ar += t[q].re * wp.re - t[q].im * wp.im;
                ai += t[q].re * wp.im + t[q].im * wp.re;
            }
            out[k + (size_t) s * m].re = ar;
            out[k + (size_t) s * m].im = ai;
        }
    }
}

static FN(cplx) *FN(twiddles)(unsigned n, int sign)
{
    FN(cplx) *tw = (FN(cplx) *) malloc(sizeof(FN(cplx)) * n);
    for (unsigned k = 0; k < n; ++k) {
        const double a = sign * 2.0 * M_PI * (double) k / (double) n;
        tw[k].re = (T) cos(a);
        tw[k].im = (T) sin(a);
    }
    return tw;
}

/* Bluestein: X_k = c_k * sum_j (x_j c_j) conj(c_{k-j}), c_j = exp(sign*pi*i*j^2/n),
 * evaluated as a circular convolution of power-of-two length M >= 2n-1,
 * in double precision. */
static void FN(fft_bluestein)(const FN(cplx) *in, size_t is, FN(cplx) *out, unsigned n, int sign)
{
    unsigned M = 1;
    while (M < 2 * n - 1) M <<= 1;
    typedef struct { double re, im; } cd;
    cd *c = (cd *) malloc(sizeof(cd) * n);
    cd *a = (cd *) calloc(M, sizeof(cd));
    cd *f = (cd *) calloc(M, sizeof(cd));
    cd *tw = (cd *) malloc(sizeof(cd) * M);
    for (unsigned j = 0; j < n; ++j) {
        const unsigned long long jj = ((unsigned long long) j * j) % (2ull * n);
        const double ang = sign * M_PI * (double) jj / (double) n;
        c[j].re = cos(ang); c[j].im = sin(ang);
    }
    for (unsigned j = 0; j < n; ++j) {
        const double xr = in[j * is].re, xi = in[j * is].im;
        a[j].re = xr * c[j].re - xi * c[j].im;
        a[j].im = xr * c[j].im + xi * c[j].re;
        f[j].re = c[j].re; f[j].im = -c[j].im;
        if (j) { f[M - j].re = c[j].re; f[M - j].im = -c[j].im; }
    }
    /* radix-2 iterative FFT of length M in double */
    for (unsigned k = 0; k < M; ++k) {
        const double ang = -2.0 * M_PI * (double) k / (double) M;
        tw[k].re = cos(ang); tw[k].im = sin(ang);
    }
    cd *bufs[2] = { a, f };
    for (int b = 0; b < 3; ++b) {
        cd *x = b < 2 ? bufs[b] : a;
        const int inverse = (b == 2);
        /* bit reversal */
        for (unsigned i = 1, j = 0; i < M; ++i) {
            unsigned bit = M >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { cd tmp = x[i]; x[i] = x[j]; x[j] = tmp; }
        }
        for (unsigned len = 2; len <= M; len <<= 1) {
            const unsigned step = M / len;
            for (unsigned i = 0; i < M; i += len)
                for (unsigned k = 0; k < len / 2; ++k) {
                    cd w = tw[k * step];
                    if (inverse) w.im = -w.im;
                    const cd u = x[i + k], v0 = x[i + k + len / 2];
                    const cd v = { v0.re * w.re - v0.im * w.im, v0.re * w.im + v0.im * w.re };
                    x[i + k].re = u.re + v.re; x[i + k].im = u.im + v.im;
                    x[i + k + len / 2].re = u.re - v.re; x[i + k + len / 2].im = u.im - v.im;
                }
        }
        if (b == 1)
            for (unsigned k = 0; k < M; ++k) {
                const cd p = { a[k].re * f[k].re - a[k].im * f[k].im,
                               a[k].re * f[k].im + a[k].im * f[k].re };
                a[k] = p;
            }
    }
    for (unsigned k = 0; k < n; ++k) {
        const double yr = a[k].re / M, yi = a[k].im / M;
        out[k].re = (T) (yr * c[k].re - yi * c[k].im);
        out[k].im = (T) (yr * c[k].im + yi * c[k].re);
    }
    free(c); free(a); free(f); free(tw);
}

static void FN(fft_line)(const FN(cplx) *in, size_t is, FN(cplx) *out, unsigned n, int sign,
                         const FN(fftplan) *pl)
{
    if (pl->fac[pl->nfac - 1] > ORC_MAX_RADIX)
        FN(fft_bluestein)(in, is, out, n, sign);
    else
        FN(fft_rec)(in, is, out, n, pl->fac, pl, 1);
}

static void FN(plan_init)(FN(fftplan) *pl, unsigned n, int sign)
{
    pl->n = n;
    pl->nfac = FN(fft_factor)(n, pl->fac);
    if (pl->nfac == 0) { pl->fac[0] = 1; pl->nfac = 1; }
    pl->tw = FN(twiddles)(n, sign);
}

int FN(orc_fft)(T *out, const T *in, unsigned n, int sign)
{
    if (n == 0) return 1;
    FN(fftplan) pl;
    FN(plan_init)(&pl, n, sign);
    FN(cplx) *tmp = (FN(cplx) *) malloc(sizeof(FN(cplx)) * n);
    FN(fft_line)((const FN(cplx) *) in, 1, tmp, n, sign, &pl);
    memcpy(out, tmp, sizeof(FN(cplx)) * n);
    free(tmp);
    free(pl.tw);
    return 0;
}

/* 2-D transform (dims {H, W} as in shrinkWrap.cpp:151-155): rows, then columns */
int FN(orc_fft2d)(T *outp, const T *inp, unsigned w, unsigned h, int sign)
{
    if (w == 0 || h == 0) return 1;
    const FN(cplx) *in = (const FN(cplx) *) inp;
    FN(cplx) *out = (FN(cplx) *) outp;
    FN(fftplan) pw, ph;
    FN(plan_init)(&pw, w, sign);
    FN(plan_init)(&ph, h, sign);
    const unsigned nmax = w > h ? w : h;
    /* columns are copied out and back ORC_COL_BLOCK at a time, so every
     * cache line of a row serves several columns; each line's arithmetic is
     * the same as one column at a time */
    enum { ORC_COL_BLOCK = 16 };
    #pragma omp parallel num_threads(g_fft_threads)
    {
        FN(cplx) *tmp = (FN(cplx) *) malloc(sizeof(FN(cplx)) * nmax);
        FN(cplx) *col = (FN(cplx) *) malloc(sizeof(FN(cplx)) * (size_t) h * ORC_COL_BLOCK);
        #pragma omp for
        for (int y = 0; y < (int) h; ++y) {
            FN(fft_line)(in + (size_t) y * w, 1, tmp, w, sign, &pw);
            memcpy(out + (size_t) y * w, tmp, sizeof(FN(cplx)) * w);
        }
        #pragma omp for
        for (int x0 = 0; x0 < (int) w; x0 += ORC_COL_BLOCK) {
            const unsigned nb = w - (unsigned) x0 < ORC_COL_BLOCK ? w - (unsigned) x0 : ORC_COL_BLOCK;
            for (unsigned y = 0; y < h; ++y)
                for (unsigned b = 0; b < nb; ++b) col[(size_t) b * h + y] = out[(size_t) y * w + x0 + b];
            for (unsigned b = 0; b < nb; ++b) {
                FN(fft_line)(col + (size_t) b * h, 1, tmp, h, sign, &ph);
                memcpy(col + (size_t) b * h, tmp, sizeof(FN(cplx)) * h);
            }
            for (unsigned y = 0; y < h; ++y)
                for (unsigned b = 0; b < nb; ++b) out[(size_t) y * w + x0 + b] = col[(size_t) b * h + y];
        }
        free(tmp); free(col);
    }
    free(pw.tw); free(ph.tw);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* one mask update: shrinkWrap.cpp:191 (norm), :206 (blur), :234-242 (cut)   */
/* ------------------------------------------------------------------------ */
int FN(orc_make_mask)(const T *gprime, T *mask, unsigned w, unsigned h, float sigma, float cut)
{
    const unsigned n = w * h;
    FN(orc_complex_norm)(mask, gprime, n);
    FN(orc_blur)(mask, w, h, sigma);
    const T absMax = FN(orc_vmax)(mask, n);
    T threshold = absMax;
    threshold *= (T) cut;
    FN(orc_cutoff)(mask, n, threshold, 1, 0);
    return 0;
}

/* HIO inner loop (shrinkWrap.cpp:260-319) run n_iter times from a given state.
 * gprime = curData (g'), gprev = gPrevious, modulus = |F| */
int FN(orc_hio_iterations)(T *gprime, T *gprev, const T *mask, const T *modulus,
                           unsigned w, unsigned h, T beta, unsigned n_iter)
{
    const unsigned n = w * h;
    for (unsigned it = 0; it < n_iter; ++it) {
        FN(orc_hio_constraint)(gprev, gprime, mask, n, beta);
        FN(orc_fft2d)(gprime, gprev, w, h, -1);          /* toFreqSpace: FORWARD gPrev -> cur */
        FN(orc_apply_modulus)(gprime, modulus, n);
        FN(orc_fft2d)(gprime, gprime, w, h, +1);         /* toRealSpace: BACKWARD in place */
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* shrinkWrap (src/imresh/algorithms/shrinkWrap.cpp:111-361)                  */
/* ------------------------------------------------------------------------ */
int FN(orc_shrinkwrap)(T *io, unsigned w, unsigned h, unsigned n_cycles,
                       float target_error, float hio_beta, float cutoff_autocorel,
                       float cutoff, float sigma0, float sigma_change, unsigned n_hio,
                       orc_sw_stats *stats)
{
    if (io == NULL) return 1;
    if (target_error <= 0) target_error = 1e-5f;
    if (n_hio == 0) n_hio = 20;
    if (hio_beta <= 0) hio_beta = 0.9f;
    if (cutoff_autocorel <= 0) cutoff_autocorel = 0.04f;
    if (cutoff <= 0) cutoff = 0.2f;
    if (sigma0 <= 0) sigma0 = 3.0f;
    if (sigma_change <= 0) sigma_change = 0.01f;

    float sigma = sigma0;
    const unsigned n = w * h;
    T *cur = (T *) malloc(sizeof(T) * 2 * (size_t) n);
    T *gprev = (T *) malloc(sizeof(T) * 2 * (size_t) n);
    T *mask = (T *) malloc(sizeof(T) * (size_t) n);

    FN(orc_copy_to_real)(cur, io, n);
    FN(orc_fft2d)(cur, cur, w, h, +1);
    memcpy(gprev, cur, sizeof(T) * 2 * (size_t) n);

    unsigned cycles = 0;
    float err = 0, tot = 0, nm = 0;
    for (unsigned c = 0; c < n_cycles; ++c) {
        FN(orc_make_mask)(cur, mask, w, h, sigma, c == 0 ? cutoff_autocorel : cutoff);
        sigma = fmaxf(1.5f, (1 - sigma_change) * sigma);
        FN(orc_hio_iterations)(cur, gprev, mask, io, w, h, (T) hio_beta, n_hio);
        err = FN(orc_hio_error)(cur, mask, n, 0, &tot, &nm);
        ++cycles;
        if (target_error > 0 && err < target_error)
            break;
    }
    FN(orc_copy_from_real)(io, cur, n);
    if (stats) {
        stats->cycles_run = cycles;
        stats->last_error = err;
        stats->last_total = tot;
        stats->last_masked = nm;
    }
    free(cur); free(gprev); free(mask);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* forward model (diffractionIntensity.cpp:49-80, CUDA :84-135)              */
/* ------------------------------------------------------------------------ */
/* io (real object) <- |FFT2((x, 0))|: the real plane copied into the real
 * parts with zero imaginary parts (:59-63 / the memset + strided copy of
 * :98-113), the unnormalized FORWARD 2-D transform of H rows x W columns
 * (:64-68 / :114-118) and sqrt(re*re + im*im) per element (:72-77 /
 * cudaComplexNormElementwise), no fftshift */
int FN(orc_diffraction_intensity)(T *io, unsigned w, unsigned h)
{
    const size_t n = (size_t) w * h;
    T *tmp = (T *) malloc(sizeof(T) * 2 * n);
    if (tmp == NULL) return 1;
    for (size_t i = 0; i < n; ++i) {
        tmp[2 * i] = io[i];
        tmp[2 * i + 1] = 0;
    }
    const int rc = FN(orc_fft2d)(tmp, tmp, w, h, -1);
    if (rc == 0) FN(orc_complex_norm)(io, tmp, (unsigned) n);
    free(tmp);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* standalone HIO with a given support (hybridInputOutput.cpp:126-306)       */
/* ------------------------------------------------------------------------ */
/* addRandomPhase (:126-197): srand(2623091912); a real field rand()/RAND_MAX,
 * its forward FFT, phase = arg (float atan2), output = |F| * (cos, sin) of
 * the phase evaluated in double (unqualified cos on a float) */
void FN(orc_random_phase)(const T *norm, T *out, unsigned w, unsigned h)
{
    const unsigned n = w * h;
    T *tmp = (T *) malloc(sizeof(T) * 2 * (size_t) n);
    srand(2623091912u);
    for (unsigned i = 0; i < n; ++i) {
        tmp[2 * i] = (T) ((float) rand() / RAND_MAX);
        tmp[2 * i + 1] = 0;
    }
    FN(orc_fft2d)(tmp, tmp, w, h, -1);
    for (unsigned i = 0; i < n; ++i) {
        const float phase = atan2f((float) tmp[2 * i + 1], (float) tmp[2 * i]);
        out[2 * i] = (T) ((double) norm[i] * cos((double) phase));
        out[2 * i + 1] = (T) ((double) norm[i] * sin((double) phase));
    }
    free(tmp);
}

/* hybridInputOutput (:200-306): io = |F| in, Re(g') out; mask uint8 (1 =
 * outside the support); n_cycles 0 -> UINT_MAX, beta <= 0 -> 0.9; the error
 * check runs after every inverse transform when target > 0.  The modulus
 * replacement divides by |G'| without the zero guard of applyComplexModulus
 * (:292-297). */
int FN(orc_hio_standalone)(T *io, const unsigned char *mask8, unsigned w, unsigned h,
                           unsigned n_cycles, float target, float beta, orc_sw_stats *stats)
{
    if (io == NULL || mask8 == NULL) return 1;
    if (n_cycles == 0) n_cycles = UINT_MAX;
    if (beta <= 0) beta = 0.9f;
    const unsigned n = w * h;
    T *cur = (T *) malloc(sizeof(T) * 2 * (size_t) n);
    T *gprev = (T *) malloc(sizeof(T) * 2 * (size_t) n);
    T *mask = (T *) malloc(sizeof(T) * (size_t) n);
    for (unsigned i = 0; i < n; ++i) mask[i] = (T) (mask8[i] ? 1 : 0);
    FN(orc_random_phase)(io, cur, w, h);
    unsigned it = 0;
    float err = 0, tot = 0, nm = 0;
    for (;;) {
        FN(orc_fft2d)(cur, cur, w, h, +1);
        if (it == 0) memcpy(gprev, cur, sizeof(T) * 2 * (size_t) n);
        if (target > 0) {
            err = FN(orc_hio_error)(cur, mask, n, 0, &tot, &nm);
            if (err < target) break;
        }
        if (it >= n_cycles) break;
        for (unsigned i = 0; i < n; ++i) {
            if (mask8[i] == 1 || cur[2 * i] < 0) {
                cur[2 * i] = gprev[2 * i] - (T) beta * cur[2 * i];
                cur[2 * i + 1] = gprev[2 * i + 1] - (T) beta * cur[2 * i + 1];
            }
        }
        memcpy(gprev, cur, sizeof(T) * 2 * (size_t) n);
        FN(orc_fft2d)(cur, cur, w, h, -1);
        for (unsigned i = 0; i < n; ++i) {
            const T re = cur[2 * i], im = cur[2 * i + 1];
            const T factor = io[i] / (T) SQRT(re * re + im * im);
            cur[2 * i] *= factor;
            cur[2 * i + 1] *= factor;
        }
        ++it;
    }
    FN(orc_copy_from_real)(io, cur, n);
    if (stats) {
        stats->cycles_run = it;
        stats->last_error = err;
        stats->last_total = tot;
        stats->last_masked = nm;
    }
    free(cur); free(gprev); free(mask);
    return 0;
}
