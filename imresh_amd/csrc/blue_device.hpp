// Bluestein (chirp-z) form of the line FFT for lengths without a compiled
// mixed-radix plan.  With a_j = exp(-i*pi*j^2/n),
//   X_k = sum_j x_j w^(jk) = a_k * sum_j (x_j a_j) conj(a_(k-j)),
// a circular convolution of length M >= 2n-1 (a power of two with a
// compiled Stockham plan PM).  One length-n transform is then
//   chirp -> FFT_M -> (x spectrum of the wrapped chirp / M) -> IFFT_M -> chirp
// entirely in registers/LDS of one kernel; the inverse direction uses
// conj(a).  PM's forward last pass and inverse first pass have the same
// radix (and the inverse last / forward first), so consecutive stages and
// consecutive transforms chain in registers exactly as line_pair does.
//
// Tables of one length (one device allocation, built on the host in double):
//   tab[0, n)       a_j
//   tab[n, n+M)     FFT_M(wrap(conj a)) / M   (forward transform)
//   tab[n+M, n+2M)  FFT_M(wrap(a)) / M        (inverse transform)
// wrap(b)[m] = b_m for m < n, b_(M-m) for m > M-n, 0 otherwise.
#pragma once

#include "fft_device.hpp"

namespace imr {

struct Blue {
    int n;
    const cf *tab;
};

template <int DIR>
__device__ __forceinline__ cf blue_chirp(const Blue &b, int e)
{
    const cf a = b.tab[e];
    return DIR < 0 ? a : mkc(a.x, -a.y);
}

// circular convolution with the chirp: v holds the chirped, zero-padded
// line at PM's forward first-pass input positions on entry and the result
// at PM's inverse last-pass output positions on exit.  HP: load the chirp's
// spectrum before the forward transform's exchanges, so its (L2) latency
// overlaps them (faster for the row kernels; the column kernels, with more
// registers live, were slower with it)
template <class PM, int DIR, bool HP, class Lds>
__device__ __forceinline__ void blue_conv(cf (&v)[PM::E], const cf *tw, int t, const Lds &L,
                                          const Blue &b)
{
    const cf *h = b.tab + b.n + (DIR < 0 ? 0 : PM::n);
    if constexpr (HP) {
        cf hv[PM::E];
        for_outputs<PM, -1, PM::np - 1>(t, [&](int s, int e) { hv[s] = h[e]; });
        compute<PM, -1, 0>(v, tw, t);
        passes_from<PM, -1, 1>(v, tw, t, L);
        for_outputs<PM, -1, PM::np - 1>(t, [&](int s, int e) { v[s] = cmul(v[s], hv[s]); (void) e; });
    } else {
        compute<PM, -1, 0>(v, tw, t);
        passes_from<PM, -1, 1>(v, tw, t, L);
        for_outputs<PM, -1, PM::np - 1>(t, [&](int s, int e) { v[s] = cmul(v[s], h[e]); });
    }
    compute<PM, +1, 0>(v, tw, t);
    passes_from<PM, +1, 1>(v, tw, t, L);
}

// one length-n transform of direction DIR: src(e) / dst(e, x) for e < n
template <class PM, int DIR, bool HP, class Lds, class Src, class Dst>
__device__ __forceinline__ void blue_fft(cf (&v)[PM::E], const cf *tw, int t, const Lds &L,
                                         const Blue &b, Src &src, Dst &dst)
{
    const int n = b.n;
    auto in = [&](int e) { return e < n ? cmul(src(e), blue_chirp<DIR>(b, e)) : mkc(0.f, 0.f); };
    gather_global<PM, -1, 0>(v, t, in);
    blue_conv<PM, DIR, HP>(v, tw, t, L, b);
    for_outputs<PM, +1, PM::np - 1>(t, [&](int s, int e) {
        if (e < n) dst(e, cmul(v[s], blue_chirp<DIR>(b, e)));
    });
}

// transform DA, mid(e, z, pre(e)) on its outputs, transform DB (as line_pair)
template <class PM, int DA, int DB, bool HP, class Lds, class Src, class Pre, class Mid, class Dst>
__device__ __forceinline__ void blue_pair(cf (&v)[PM::E], const cf *tw, int t, const Lds &L,
                                          const Blue &b, Src &src, Pre &pre, Mid &mid, Dst &dst)
{
    static_assert(DA == -DB, "a pair runs opposite directions");
    const int n = b.n;
    using PT = decltype(pre(0));
    PT pf[PM::E];
    auto in = [&](int e) { return e < n ? cmul(src(e), blue_chirp<DA>(b, e)) : mkc(0.f, 0.f); };
    gather_global<PM, -1, 0>(v, t, in);
    for_outputs<PM, +1, PM::np - 1>(t, [&](int s, int e) {
        if (e < n) pf[s] = pre(e);
    });
    blue_conv<PM, DA, HP>(v, tw, t, L, b);
    for_outputs<PM, +1, PM::np - 1>(t, [&](int s, int e) {
        if (e < n) {
            const cf z = mid(e, cmul(v[s], blue_chirp<DA>(b, e)), pf[s]);
            v[s] = cmul(z, blue_chirp<DB>(b, e));
        } else {
            v[s] = mkc(0.f, 0.f);
        }
    });
    blue_conv<PM, DB, HP>(v, tw, t, L, b);
    for_outputs<PM, +1, PM::np - 1>(t, [&](int s, int e) {
        if (e < n) dst(e, cmul(v[s], blue_chirp<DB>(b, e)));
    });
}

}  // namespace imr

namespace imr {

// direct (compiled plan of the line length) or Bluestein (BL: P is the
// convolution plan, the line length is b.n) line transforms
template <class P, int DIR, bool BL, bool HP = false, class Lds, class Src, class Dst>
__device__ __forceinline__ void xfft(cf (&v)[P::E], const cf *tw, int t, const Lds &L, const Blue &b,
                                     Src &src, Dst &dst)
{
    if constexpr (BL) blue_fft<P, DIR, HP>(v, tw, t, L, b, src, dst);
    else line_fft<P, DIR>(v, tw, t, L, src, dst);
}

template <class P, int DA, int DB, bool BL, bool HP = false, class Lds, class Src, class Pre, class Mid,
          class Dst>
__device__ __forceinline__ void xpair(cf (&v)[P::E], const cf *tw, int t, const Lds &L, const Blue &b,
                                      Src &src, Pre &pre, Mid &mid, Dst &dst)
{
    if constexpr (BL) blue_pair<P, DA, DB, HP>(v, tw, t, L, b, src, pre, mid, dst);
    else line_pair<P, DA, DB>(v, tw, t, L, src, pre, mid, dst);
}

}  // namespace imr
