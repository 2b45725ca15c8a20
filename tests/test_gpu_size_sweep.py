"""A fixed sweep of arbitrary image sizes (GPU).

Lengths without a compiled plan run Bluestein transforms whose convolution
length, tile shape and (for 8192-point columns) four-step pass depend on the
length, so a mistake can hide in one range of sizes only.  40 (W, H) pairs,
log-uniform in [1, 8192] per axis with W*H <= 4 Mi elements, drawn once from
a fixed seed, each go through:
  * the 2-D FFT, forward and inverse, against numpy in float64 (relative L2
    below 1e-5);
  * one fused HIO iteration from an oracle state against the float64 oracle
    (relative L2 below 1e-5, the per-iteration bar), for W*H <= 1 Mi;
  * a 1-cycle x 2-iteration shrink-wrap: the masked count equal to the
    float64 oracle's and the error within 1e-3 relative (when the masked
    set is not empty: the error is 0/0 otherwise), for W*H <= 256 Ki.
The HIO inputs are the modulus of a sparse random object (20% of the
pixels, amplitudes U[0.2, 1]), which thin images also get (the benchmark's
shape generator leaves them empty).
"""
import numpy as np
import pytest

import imresh_amd as im
import oracle
from helpers import FLT_MIN
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu
B = im.DeviceBuffer
L = im.lib


def _sizes():
    rng = np.random.default_rng(20261203)
    out = []
    while len(out) < 40:
        w, h = (int(round(np.exp(rng.uniform(0.0, np.log(8192))))) for _ in range(2))
        if w * h <= 4 << 20:
            out.append((w, h))
    return out


SIZES = _sizes()


def _fft_gpu(x, direction):
    h, w = x.shape
    d = B.from_array(x.astype(np.complex64))
    im.check(L().imresh_fft2d(d.ptr, d.ptr, w, h, direction, None), f"fft {w}x{h}")
    return d.download(np.complex64, (h, w))


@pytest.mark.parametrize("w,h", SIZES)
def test_size_sweep(w, h):
    rng = np.random.default_rng(w * 8209 + h)
    x = (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)
    x64 = x.astype(np.complex128)
    e_f = syn.rel_l2(_fft_gpu(x, -1), np.fft.fft2(x64))
    e_i = syn.rel_l2(_fft_gpu(x, +1), np.fft.ifft2(x64) * (h * w))
    msg = f"{w}x{h}: FFT forward {e_f:.2e} inverse {e_i:.2e}"
    assert e_f < 1e-5 and e_i < 1e-5, msg
    if w * h <= 1 << 20:
        oracle.set_fft_threads(16)
        keep = rng.random((h, w)) < 0.2
        keep[0, 0] = True
        obj = np.where(keep, rng.uniform(0.2, 1.0, (h, w)), 0.0)
        M = np.abs(np.fft.fft2(obj)).astype(np.float32)
        g = oracle.fft2d(M.astype(np.complex64), +1, np.float32)
        mask = oracle.make_mask(g, 3.0, 0.04, np.float32)
        dg, dgp = B.from_array(g), B.from_array(g)
        dm, dM = B.from_array(mask.astype(np.uint8)), B.from_array(M)
        im.check(L().imresh_hio_iterations(dg.ptr, dgp.ptr, dm.ptr, dM.ptr, w, h, np.float32(0.9), 1, None))
        g64, _ = oracle.hio_iterations(g.astype(np.complex128), g.astype(np.complex128), mask.astype(np.float64),
                                       M.astype(np.float64), 0.9, 1, np.float64)
        e_h = syn.rel_l2(dg.download(np.complex64, (h, w)), g64)
        msg += f"; one HIO iteration {e_h:.2e}"
        assert e_h < 1e-5, msg
        if w * h <= 1 << 18:
            kw = dict(n_cycles=1, target_error=FLT_MIN, n_hio_cycles=2)
            y = M.copy()
            r = im.shrink_wrap(y, **kw)
            _, s64 = oracle.shrink_wrap(M.astype(np.float64), dtype=np.float64, **kw)
            msg += f"; 1x2 shrink-wrap masked {r.last_masked:.0f}/{s64.last_masked:.0f}"
            assert r.last_masked == s64.last_masked, msg
            if s64.last_masked > 0:
                rel = abs(r.last_error - s64.last_error) / abs(s64.last_error)
                msg += f", error rel {rel:.2e}"
                assert rel < 1e-3, msg
    print(msg)
