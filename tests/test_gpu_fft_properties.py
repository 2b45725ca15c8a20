"""Size-independent properties of the library's 2-D FFT at the large
configurations' full sizes, where a float64 reference transform on the host
would dominate the test: unnormalised forward/inverse round trip, Parseval,
linearity, the shift theorem and two closed forms (a delta and a plane wave).
Each catches a different plausible mistake: a dropped twiddle or a wrong
pass order breaks the round trip and Parseval, a sign or index error in one
axis breaks the shift theorem, a non-linear slip (state carried between
lines) breaks linearity.  The sizes cover the compiled 4096- and 8192-point
plans, the four-step 8192-point column pass and Bluestein lines on both axes
(3000 points: convolution length 8192; 5000 points: 16384).

Tolerance: float32 transforms of N points carry a relative L2 error of about
2e-7 * log2(N) against exact arithmetic (the bound the numpy comparisons at
smaller sizes use); a property combining two transforms is allowed twice
that.
"""
import numpy as np
import pytest

import imresh_amd as im

pytestmark = pytest.mark.gpu
B = im.DeviceBuffer
L = im.lib

SIZES = [(4096, 4096), (3000, 3000), (5000, 8192), (8192, 8192)]   # (h, w)


def _fft(d, h, w, direction):
    im.check(L().imresh_fft2d(d.ptr, d.ptr, w, h, direction, None), f"fft {w}x{h}")


def _norm(a):
    """L2 norm accumulated in float64, chunked (bounded host memory)."""
    a = a.reshape(-1)
    s = 0.0
    for i in range(0, a.size, 1 << 22):
        c = a[i:i + (1 << 22)]
        s += float(np.sum(c.real.astype(np.float64) ** 2 + c.imag.astype(np.float64) ** 2))
    return np.sqrt(s)


def _rel(a, b):
    return _norm(a - b) / _norm(b)


def _random(shape, seed):
    rng = np.random.default_rng(seed)
    x = np.empty(shape, np.complex64)
    x.real = rng.standard_normal(shape, dtype=np.float32)
    x.imag = rng.standard_normal(shape, dtype=np.float32)
    return x


@pytest.mark.parametrize("h,w", SIZES)
def test_fft2d_full_size_properties(h, w):
    n = h * w
    tol = 2e-7 * np.log2(n)
    print(f"{w}x{h}: tol {tol:.2e}", flush=True)
    x = _random((h, w), seed=h * 31 + w)
    d = B.from_array(x)
    _fft(d, h, w, -1)
    F = d.download(np.complex64, (h, w))

    # Parseval: sum |F|^2 = N sum |x|^2
    px, pf = _norm(x) ** 2, _norm(F) ** 2
    assert abs(pf / (n * px) - 1) < tol, (pf / (n * px))

    # unnormalised inverse of the forward transform is N x
    _fft(d, h, w, +1)
    y = d.download(np.complex64, (h, w))
    e_rt = _rel(y / np.float32(n), x)
    print(f"  round trip rel L2 {e_rt:.2e}", flush=True)
    assert e_rt < 2 * tol

    # linearity: F(a x + b z) = a F(x) + b F(z)
    z = _random((h, w), seed=h * 37 + w + 1)
    a, b = np.float32(0.75), np.float32(-1.25)
    d = B.from_array(a * x + b * z)
    _fft(d, h, w, -1)
    lhs = d.download(np.complex64, (h, w))
    d = B.from_array(z)
    _fft(d, h, w, -1)
    rhs = a * F + b * d.download(np.complex64, (h, w))
    e_lin = _rel(lhs, rhs)
    print(f"  linearity rel L2 {e_lin:.2e}", flush=True)
    assert e_lin < 2 * tol
    del lhs, rhs, z

    # shift theorem: x shifted by (sy, sx) rows/columns (circularly) has the
    # spectrum F[ky, kx] * exp(-2 pi i (ky sy / h + kx sx / w))
    sy, sx = 3 % h, 5 % w
    d = B.from_array(np.roll(x, (sy, sx), axis=(0, 1)))
    _fft(d, h, w, -1)
    Fs = d.download(np.complex64, (h, w))
    py = np.exp(-2j * np.pi * sy * np.arange(h) / h).astype(np.complex64)
    pxs = np.exp(-2j * np.pi * sx * np.arange(w) / w).astype(np.complex64)
    exp = F * py[:, None] * pxs[None, :]
    e_sh = _rel(Fs, exp)
    print(f"  shift theorem rel L2 {e_sh:.2e}", flush=True)
    assert e_sh < 2 * tol


@pytest.mark.parametrize("h,w", [(8192, 8192), (3000, 3000)])
def test_fft2d_full_size_closed_forms(h, w):
    n = h * w
    # a delta at the origin transforms to all ones (every twiddle times 1)
    x = np.zeros((h, w), np.complex64)
    x[0, 0] = 1
    d = B.from_array(x)
    _fft(d, h, w, -1)
    F = d.download(np.complex64, (h, w))
    assert np.abs(F - 1).max() < 1e-4
    # a plane wave of frequency (ky, kx) transforms to N at (ky, kx), ~0 elsewhere
    ky, kx = 7, 11
    ph = (2 * np.pi * ky / h) * np.arange(h)[:, None] + (2 * np.pi * kx / w) * np.arange(w)[None, :]
    x = np.exp(1j * ph).astype(np.complex64)
    del ph
    d = B.from_array(x)
    _fft(d, h, w, -1)
    F = d.download(np.complex64, (h, w))
    peak = F[ky, kx]
    print(f"{w}x{h}: plane-wave peak {peak.real:.6e}{peak.imag:+.3e}i (N = {n})", flush=True)
    assert abs(peak - n) < 1e-4 * n
    F[ky, kx] = 0
    assert np.abs(F).max() < 1e-4 * n
