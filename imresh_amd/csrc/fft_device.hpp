// Device building blocks of the LDS-staged Stockham line FFT for gfx950.
//
// A line of length n = prod(R_p) is transformed by T threads (compile-time
// plan, see plans.hpp).  Pass p (radix R, span ns = R_0*...*R_{p-1}) owns
// butterflies j = t + b*T; it reads elements j + r*(n/R), multiplies element r
// by exp(-+2*pi*i*r*(j%ns)/(ns*R)), runs a radix-R DFT in registers and
// writes element r to (j - j%ns)*R + j%ns + r*ns.  The first pass therefore
// reads and the last pass writes the line in natural order j + r*(n/R): the
// accesses are coalesced in global memory, and a fused prologue/epilogue
// (HIO constraint, modulus projection, norm, error sums) runs on registers at
// exactly those positions.  Between passes the data goes through LDS; each
// thread keeps at most E = 16 complex values in VGPRs.
//
// Forward = exp(-2*pi*i*j*k/n) (FFTW_FORWARD / CUFFT_FORWARD); both
// directions are unnormalized, as the reference's FFTs.  Forward transforms
// run the radices in plan order, inverse transforms in reverse order, so the
// last pass of one direction and the first pass of the other have the same
// radix and butterfly ownership: a forward->inverse (or inverse->forward)
// pair with an elementwise operator in between never leaves registers at the
// junction (line_pair).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace imr {

// native 2-wide vector: one 64-bit load / store per complex value
typedef float cf __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cf mkc(float x, float y)
{
    cf r;
    r.x = x;
    r.y = y;
    return r;
}

// ------------------------------------------------------------ complex math
// Products use one fma per component; the FFT is the only place that
// contracts -- the library is compiled with -ffp-contract=off so that all
// elementwise operators round exactly as the reference's C++ code.
__device__ __forceinline__ cf cadd(cf a, cf b) { return mkc(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cf csub(cf a, cf b) { return mkc(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cf cmul(cf a, cf b)
{
    return mkc(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ cf cmulc(cf a, cf b)
{
    return mkc(__builtin_fmaf(a.x, b.x, a.y * b.y), __builtin_fmaf(a.y, b.x, -(a.x * b.y)));
}
// a * (DIR < 0 ? -i : +i)
template <int DIR> __device__ __forceinline__ cf rot(cf a)
{
    return DIR < 0 ? mkc(a.y, -a.x) : mkc(-a.y, a.x);
}
// twiddle multiply: forward uses the table value, inverse its conjugate
template <int DIR> __device__ __forceinline__ cf twmul(cf a, cf w)
{
    return DIR < 0 ? cmul(a, w) : cmulc(a, w);
}

// cos/sin(2*pi*k/R) for the odd radices
template <int R> struct OddConst;
template <> struct OddConst<3> {
    static constexpr float c[3] = {1.f, -0.5f, -0.5f};
    static constexpr float s[3] = {0.f, 0.866025404f, -0.866025404f};
};
template <> struct OddConst<5> {
    static constexpr float c[5] = {1.f, 0.309016994f, -0.809016994f, -0.809016994f, 0.309016994f};
    static constexpr float s[5] = {0.f, 0.951056516f, 0.587785252f, -0.587785252f, -0.951056516f};
};

// ------------------------------------------------------------ butterflies
template <int DIR>
__device__ __forceinline__ void bf2(cf *v)
{
    const cf a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void bf4io(cf a, cf b, cf c, cf d, cf &o0, cf &o1, cf &o2, cf &o3)
{
    const cf t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = rot<DIR>(csub(b, d));
    o0 = cadd(t0, t2);
    o2 = csub(t0, t2);
    o1 = cadd(t1, t3);
    o3 = csub(t1, t3);
}

template <int DIR>
__device__ __forceinline__ void bf4(cf *v)
{
    cf o0, o1, o2, o3;
    bf4io<DIR>(v[0], v[1], v[2], v[3], o0, o1, o2, o3);
    v[0] = o0; v[1] = o1; v[2] = o2; v[3] = o3;
}

template <int DIR>
__device__ __forceinline__ void bf8(cf *v)
{
    constexpr float c = 0.707106781f;
    cf e0, e1, e2, e3, o0, o1, o2, o3;
    bf4io<DIR>(v[0], v[2], v[4], v[6], e0, e1, e2, e3);
    bf4io<DIR>(v[1], v[3], v[5], v[7], o0, o1, o2, o3);
    o1 = cmul(o1, mkc(c, DIR * c));
    o2 = rot<DIR>(o2);
    o3 = cmul(o3, mkc(-c, DIR * c));
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

template <int DIR>
__device__ __forceinline__ void bf16(cf *v)
{
    constexpr float c1 = 0.923879533f, s1 = 0.382683432f, c2 = 0.707106781f;
    cf y[16];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        bf4io<DIR>(v[q], v[q + 4], v[q + 8], v[q + 12], y[4 * q + 0], y[4 * q + 1], y[4 * q + 2],
                   y[4 * q + 3]);
    // y[4*q + s] *= W16^(q*s), W16^k = (cos(2 pi k/16), DIR sin(2 pi k/16))
    y[5] = cmul(y[5], mkc(c1, DIR * s1));     // k = 1
    y[6] = cmul(y[6], mkc(c2, DIR * c2));     // k = 2
    y[7] = cmul(y[7], mkc(s1, DIR * c1));     // k = 3
    y[9] = cmul(y[9], mkc(c2, DIR * c2));     // k = 2
    y[10] = rot<DIR>(y[10]);                  // k = 4
    y[11] = cmul(y[11], mkc(-c2, DIR * c2));  // k = 6
    y[13] = cmul(y[13], mkc(s1, DIR * c1));   // k = 3
    y[14] = cmul(y[14], mkc(-c2, DIR * c2));  // k = 6
    y[15] = cmul(y[15], mkc(-c1, -DIR * s1)); // k = 9
#pragma unroll
    for (int s = 0; s < 4; ++s)
        bf4io<DIR>(y[s], y[4 + s], y[8 + s], y[12 + s], v[s], v[s + 4], v[s + 8], v[s + 12]);
}

template <int R, int DIR>
__device__ __forceinline__ void bf_odd(cf *v)
{
    constexpr int H = (R - 1) / 2;
    cf t[H], u[H];
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        t[k - 1] = cadd(v[k], v[R - k]);
        u[k - 1] = csub(v[k], v[R - k]);
    }
    const cf x0 = v[0];
    cf sum = x0;
#pragma unroll
    for (int k = 0; k < H; ++k) sum = cadd(sum, t[k]);
#pragma unroll
    for (int s = 1; s <= H; ++s) {
        cf a = x0, b = mkc(0.f, 0.f);
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            const float cc = OddConst<R>::c[(k * s) % R];
            const float ss = OddConst<R>::s[(k * s) % R];
            a = mkc(__builtin_fmaf(t[k - 1].x, cc, a.x), __builtin_fmaf(t[k - 1].y, cc, a.y));
            b = mkc(__builtin_fmaf(u[k - 1].x, ss, b.x), __builtin_fmaf(u[k - 1].y, ss, b.y));
        }
        const cf rb = rot<DIR>(b);
        v[s] = cadd(a, rb);
        v[R - s] = csub(a, rb);
    }
    v[0] = sum;
}

template <int R, int DIR>
__device__ __forceinline__ void bfly(cf *v);

// inverse of a modulo m (a, m coprime, m small), for the index maps below
__host__ __device__ constexpr int inv_mod(int a, int m)
{
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return m == 1 ? 0 : -1;
}

// Radix N1*N2 with N1, N2 coprime (6, 10, 12, 15) by the prime-factor
// (Good-Thomas) algorithm: with input n = (N2 n1 + N1 n2) mod N and output
// k = (N2 (N2^-1 mod N1) k1 + N1 (N1^-1 mod N2) k2) mod N, the length-N DFT
// is N2 length-N1 DFTs over n1 followed by N1 length-N2 DFTs over n2, with
// no twiddle factors between them.  The index maps are compile-time
// register permutations.
template <int N1, int N2, int DIR>
__device__ __forceinline__ void bf_pfa(cf *v)
{
    constexpr int N = N1 * N2;
    constexpr int i1 = inv_mod(N2 % N1, N1), i2 = inv_mod(N1 % N2, N2);
    static_assert(i1 > 0 && i2 > 0, "prime-factor radices must be coprime");
    cf a[N1][N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
        cf c[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) c[n1] = v[(N2 * n1 + N1 * n2) % N];
        bfly<N1, DIR>(c);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) a[k1][n2] = c[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        bfly<N2, DIR>(a[k1]);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) v[(N2 * i1 * k1 + N1 * i2 * k2) % N] = a[k1][k2];
    }
}

template <int R, int DIR>
__device__ __forceinline__ void bfly(cf *v)
{
    if constexpr (R == 1) (void) v;   // length-1 line: identity
    else if constexpr (R == 2) bf2<DIR>(v);
    else if constexpr (R == 4) bf4<DIR>(v);
    else if constexpr (R == 8) bf8<DIR>(v);
    else if constexpr (R == 16) bf16<DIR>(v);
    else if constexpr (R == 6) bf_pfa<2, 3, DIR>(v);
    else if constexpr (R == 10) bf_pfa<2, 5, DIR>(v);
    else if constexpr (R == 12) bf_pfa<4, 3, DIR>(v);
    else if constexpr (R == 15) bf_pfa<3, 5, DIR>(v);
    else bf_odd<R, DIR>(v);
}

// ------------------------------------------------------------ LDS layout
// One padding slot every 32 values spreads the stride-R scatters of the
// early passes over the LDS banks.
__host__ __device__ constexpr int lpad(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int lds_elems(int n) { return n + (n >> 5) + 1; }

// One line in LDS: element e at p[lpad(base + e*STRIDE)].  LdsC holds
// complex values; LdsS holds one float component at a time (the exchange
// then runs once for the real and once for the imaginary parts, halving the
// LDS footprint of a tile).  at<STEP>(e0, r) addresses element e0 + r*STEP:
// when STEP*STRIDE is a multiple of 32 the padded index is affine in r, so
// the r-offsets become immediate offsets of the LDS instructions.
template <int STEP, int STRIDE>
__device__ __forceinline__ int lds_index(int base, int e0, int r)
{
    constexpr int d = STEP * STRIDE;
    const int x0 = base + e0 * STRIDE;
    if constexpr (d % 32 == 0) return lpad(x0) + r * (d / 32 * 33);
    else return lpad(x0 + r * d);
}

template <int STRIDE>
struct LdsC {
    cf *p;
    int base;
    template <int STEP>
    __device__ __forceinline__ cf &at(int e0, int r) const { return p[lds_index<STEP, STRIDE>(base, e0, r)]; }
};
template <int STRIDE>
struct LdsS {
    float *p;
    int base;
    template <int STEP>
    __device__ __forceinline__ float &at(int e0, int r) const { return p[lds_index<STEP, STRIDE>(base, e0, r)]; }
};

// ------------------------------------------------------------ passes
// Pass I of plan P in direction DIR (forward: plan order, inverse: reversed).
template <class P, int DIR, int I>
struct PassOf {
    static constexpr bool INV = DIR > 0;
    static constexpr int R = P::rad(INV, I), NS = P::ns(INV, I), TWOFF = P::twoff(INV, I);
    static constexpr int T = P::T, nb = P::n / R, B = (nb + T - 1) / T;
    static_assert(B * R <= P::E, "plan exceeds the per-thread element budget");
    __device__ static constexpr bool full() { return nb % T == 0; }
};

// gather the inputs of pass I (natural positions j + r*nb) from global memory
template <class P, int DIR, int I, class Src>
__device__ __forceinline__ void gather_global(cf (&v)[P::E], int t, Src &src)
{
    using Q = PassOf<P, DIR, I>;
#pragma unroll
    for (int b = 0; b < Q::B; ++b) {
        const int j = t + b * Q::T;
        if (Q::full() || j < Q::nb) {
#pragma unroll
            for (int r = 0; r < Q::R; ++r) v[b * Q::R + r] = src(j + r * Q::nb);
        }
    }
}

// twiddles of pass I for this thread's butterflies (none for a first pass)
template <class P, int DIR, int I>
struct TwRegs {
    using Q = PassOf<P, DIR, I>;
    static constexpr int count = Q::NS > 1 ? Q::B * (Q::R - 1) : 0;
    cf w[count > 0 ? count : 1];
};

// global (L1/L2-resident) twiddle loads of pass I; issued before the LDS
// exchange that precedes the pass so their latency overlaps the barrier
template <class P, int DIR, int I>
__device__ __forceinline__ void load_tw(TwRegs<P, DIR, I> &w, const cf *__restrict__ tw, int t)
{
    using Q = PassOf<P, DIR, I>;
    if constexpr (Q::NS > 1) {
#pragma unroll
        for (int b = 0; b < Q::B; ++b) {
            const int j = t + b * Q::T;
            if (Q::full() || j < Q::nb) {
                const cf *p = tw + Q::TWOFF + (j % Q::NS) * (Q::R - 1);
#pragma unroll
                for (int r = 1; r < Q::R; ++r) w.w[b * (Q::R - 1) + r - 1] = p[r - 1];
            }
        }
    }
}

// twiddle multiply + radix-R butterflies of pass I, in registers
template <class P, int DIR, int I>
__device__ __forceinline__ void compute_tw(cf (&v)[P::E], const TwRegs<P, DIR, I> &w, int t)
{
    using Q = PassOf<P, DIR, I>;
#pragma unroll
    for (int b = 0; b < Q::B; ++b) {
        const int j = t + b * Q::T;
        if (Q::full() || j < Q::nb) {
            if constexpr (Q::NS > 1) {
#pragma unroll
                for (int r = 1; r < Q::R; ++r)
                    v[b * Q::R + r] = twmul<DIR>(v[b * Q::R + r], w.w[b * (Q::R - 1) + r - 1]);
            }
            bfly<Q::R, DIR>(&v[b * Q::R]);
        }
    }
}

template <class P, int DIR, int I>
__device__ __forceinline__ void compute(cf (&v)[P::E], const cf *__restrict__ tw, int t)
{
    TwRegs<P, DIR, I> w;
    load_tw<P, DIR, I>(w, tw, t);
    compute_tw<P, DIR, I>(v, w, t);
}

// output positions of pass I: (j - j%ns)*R + j%ns + r*ns
template <class P, int DIR, int I, class F>
__device__ __forceinline__ void for_outputs(int t, F &&f)
{
    using Q = PassOf<P, DIR, I>;
#pragma unroll
    for (int b = 0; b < Q::B; ++b) {
        const int j = t + b * Q::T;
        if (Q::full() || j < Q::nb) {
            const int k = j % Q::NS;
            const int base = (j - k) * Q::R + k;
#pragma unroll
            for (int r = 0; r < Q::R; ++r) f(b * Q::R + r, base + r * Q::NS);
        }
    }
}

// input positions of pass I: j + r*nb
template <class P, int DIR, int I, class F>
__device__ __forceinline__ void for_inputs(int t, F &&f)
{
    using Q = PassOf<P, DIR, I>;
#pragma unroll
    for (int b = 0; b < Q::B; ++b) {
        const int j = t + b * Q::T;
        if (Q::full() || j < Q::nb) {
#pragma unroll
            for (int r = 0; r < Q::R; ++r) f(b * Q::R + r, j + r * Q::nb);
        }
    }
}

// scatter the outputs of the last pass to global memory (natural order)
template <class P, int DIR, int I, class Dst>
__device__ __forceinline__ void scatter_global(cf (&v)[P::E], int t, Dst &dst)
{
    for_outputs<P, DIR, I>(t, [&](int s, int e) { dst(e, v[s]); });
}

// elementwise operator on the natural-order outputs of a last pass
template <class P, int DIR, int I, class Mid>
__device__ __forceinline__ void apply_mid(cf (&v)[P::E], int t, Mid &mid)
{
    for_outputs<P, DIR, I>(t, [&](int s, int e) { v[s] = mid(e, v[s]); });
}

// LDS exchange between pass IA of direction DA (its outputs) and pass IB of
// direction DB (its inputs): barrier, write, barrier, read.  The leading
// barrier orders these writes after every read of the previous exchange, so
// a wave that has read its inputs goes on to compute without waiting for
// the others.  COMP selects the component moved (0 both as one complex,
// 1 real part, 2 imaginary).
template <class P, int DA, int IA, int DB, int IB, int COMP, class Lds>
__device__ __forceinline__ void exchange_part(cf (&v)[P::E], int t, const Lds &L)
{
    using A = PassOf<P, DA, IA>;
    using Bq = PassOf<P, DB, IB>;
    __syncthreads();
#pragma unroll
    for (int b = 0; b < A::B; ++b) {
        const int j = t + b * A::T;
        if (A::full() || j < A::nb) {
            const int k = j % A::NS;
            const int base = (j - k) * A::R + k;
#pragma unroll
            for (int r = 0; r < A::R; ++r) {
                if constexpr (COMP == 0) L.template at<A::NS>(base, r) = v[b * A::R + r];
                else if constexpr (COMP == 1) L.template at<A::NS>(base, r) = v[b * A::R + r].x;
                else L.template at<A::NS>(base, r) = v[b * A::R + r].y;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < Bq::B; ++b) {
        const int j = t + b * Bq::T;
        if (Bq::full() || j < Bq::nb) {
#pragma unroll
            for (int r = 0; r < Bq::R; ++r) {
                if constexpr (COMP == 0) v[b * Bq::R + r] = L.template at<Bq::nb>(j, r);
                else if constexpr (COMP == 1) v[b * Bq::R + r].x = L.template at<Bq::nb>(j, r);
                else v[b * Bq::R + r].y = L.template at<Bq::nb>(j, r);
            }
        }
    }
}

template <class P, int DA, int IA, int DB, int IB, int S>
__device__ __forceinline__ void exchange(cf (&v)[P::E], int t, const LdsC<S> &L)
{
    exchange_part<P, DA, IA, DB, IB, 0>(v, t, L);
}
template <class P, int DA, int IA, int DB, int IB, int S>
__device__ __forceinline__ void exchange(cf (&v)[P::E], int t, const LdsS<S> &L)
{
    exchange_part<P, DA, IA, DB, IB, 1>(v, t, L);
    exchange_part<P, DA, IA, DB, IB, 2>(v, t, L);
}

// passes [I, np) of direction DIR, each preceded by an LDS exchange
template <class P, int DIR, int I, class Lds>
__device__ __forceinline__ void passes_from(cf (&v)[P::E], const cf *tw, int t, const Lds &L)
{
    if constexpr (I < P::np) {
        TwRegs<P, DIR, I> w;
        load_tw<P, DIR, I>(w, tw, t);
        exchange<P, DIR, I - 1, DIR, I>(v, t, L);
        compute_tw<P, DIR, I>(v, w, t);
        passes_from<P, DIR, I + 1>(v, tw, t, L);
    }
}

// One full transform: global (src) -> registers/LDS -> global (dst).
template <class P, int DIR, class Lds, class Src, class Dst>
__device__ __forceinline__ void line_fft(cf (&v)[P::E], const cf *tw, int t, const Lds &L, Src &src,
                                         Dst &dst)
{
    gather_global<P, DIR, 0>(v, t, src);
    compute<P, DIR, 0>(v, tw, t);
    passes_from<P, DIR, 1>(v, tw, t, L);
    scatter_global<P, DIR, P::np - 1>(v, t, dst);
}

// Transform in direction DA, elementwise `mid` on its natural-order output,
// then the transform in direction DB.  A's last pass and B's first pass
// have the same radix (mirrored sequences), so the junction never leaves
// registers: B's first pass consumes exactly the positions A's last pass
// produced in the same thread.  `pre(e)` runs right after the input gather,
// so operands of `mid` read from global memory (pre returns them; `mid(e, z,
// p)` receives them) are in flight during the first transform instead of
// stalling the junction; a `pre` returning NoPre leaves the loads to `mid`.
struct NoPre {};

// the pair after its input gather: v holds the gathered line, pf the
// prefetched junction operands
template <class P, int DA, int DB, class Lds, class PT, class Mid, class Dst>
__device__ __forceinline__ void line_pair_loaded(cf (&v)[P::E], PT (&pf)[P::E], const cf *tw, int t,
                                                 const Lds &L, Mid &mid, Dst &dst)
{
    compute<P, DA, 0>(v, tw, t);
    passes_from<P, DA, 1>(v, tw, t, L);
    for_outputs<P, DA, P::np - 1>(t, [&](int s, int e) { v[s] = mid(e, v[s], pf[s]); });
    compute<P, DB, 0>(v, tw, t);
    passes_from<P, DB, 1>(v, tw, t, L);
    scatter_global<P, DB, P::np - 1>(v, t, dst);
}

template <class P, int DA, int DB, class Lds, class Src, class Pre, class Mid, class Dst>
__device__ __forceinline__ void line_pair(cf (&v)[P::E], const cf *tw, int t, const Lds &L, Src &src,
                                          Pre &pre, Mid &mid, Dst &dst)
{
    static_assert(DA == -DB, "a pair runs opposite directions");
    static_assert(P::rad(DA > 0, P::np - 1) == P::rad(DB > 0, 0), "mirrored junction radix");
    using PT = decltype(pre(0));
    PT pf[P::E];
    gather_global<P, DA, 0>(v, t, src);
    for_outputs<P, DA, P::np - 1>(t, [&](int s, int e) { pf[s] = pre(e); });
    line_pair_loaded<P, DA, DB>(v, pf, tw, t, L, mid, dst);
}

// ------------------------------------------------------------ reductions
// std::max / std::min semantics of the reference reductions
__device__ __forceinline__ float fmax_ref(float m, float x) { return (m < x) ? x : m; }
__device__ __forceinline__ float fmin_ref(float m, float x) { return (x < m) ? x : m; }

}  // namespace imr
