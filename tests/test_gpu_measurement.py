"""GPU parity of the measurement simulation: object -> forward model ->
Poisson-sampled modulus, all on the device.

imresh_poisson_modulus is a fixed sequence of IEEE double operations with a
counter-based random stream, so it must equal the oracle's restatement bit
for bit (the oracle itself is pinned in test_oracle.py: Philox known
answers, exp/log against libm, chi-square fits of the sampler); at full
size the photon total must be Poisson(photons).
"""
import math

import numpy as np
import pytest

import imresh_amd as im
import oracle
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu
B = im.DeviceBuffer
L = im.lib


def _modulus(w, h, index):
    obj, _ = syn.workload(w, h, n_shapes=6, central=min(w, h) // 2, index=index)
    d = B.from_array(obj.astype(np.float32))
    im.check(L().imresh_diffraction_intensity_device(d.ptr, w, h, None))
    return d


@pytest.mark.parametrize("w,h,photons,seed", [(64, 64, 1e6, 1), (131, 97, 1e9, 7), (600, 1000, 3e4, 11),
                                              (2048, 2048, 1e9, 0), (1, 5, 10.0, 3)])
def test_poisson_modulus_bit_exact_vs_oracle(w, h, photons, seed):
    d = _modulus(w, h, index=w + h)
    m = d.download(np.float32, (h, w))
    im.check(L().imresh_poisson_modulus(d.ptr, w, h, photons, seed, None))
    got = d.download(np.float32, (h, w))
    exp = oracle.poisson_modulus(m, photons, seed)
    assert np.array_equal(got, exp), (np.count_nonzero(got != exp), got.size)
    k = got.astype(np.float64) ** 2
    assert np.allclose(k, np.round(k), rtol=0, atol=1e-3 * np.maximum(1.0, k))
    if w * h >= 4096:
        assert abs(np.round(k).sum() - photons) < 6 * math.sqrt(photons)


def test_poisson_modulus_seeds_and_errors():
    d0 = _modulus(128, 128, 5)
    m = d0.download(np.float32, (128, 128))
    a, b = B.from_array(m), B.from_array(m)
    im.check(L().imresh_poisson_modulus(a.ptr, 128, 128, 1e7, 42, None))
    im.check(L().imresh_poisson_modulus(b.ptr, 128, 128, 1e7, 42, None))
    x, y = a.download(np.float32, (128, 128)), b.download(np.float32, (128, 128))
    assert np.array_equal(x, y)
    c = B.from_array(m)
    im.check(L().imresh_poisson_modulus(c.ptr, 128, 128, 1e7, 43, None))
    assert not np.array_equal(c.download(np.float32, (128, 128)), x)
    z = B.from_array(np.zeros((8, 8), np.float32))
    im.check(L().imresh_poisson_modulus(z.ptr, 8, 8, 1e6, 1, None))
    assert not z.download(np.float32, (8, 8)).any()
    assert L().imresh_poisson_modulus(None, 8, 8, 1e6, 1, None) == 1
    assert L().imresh_poisson_modulus(z.ptr, 8, 8, 0.0, 1, None) == 2
    assert L().imresh_poisson_modulus(z.ptr, 8, 8, 1e31, 1, None) == 2
    assert L().imresh_poisson_modulus(z.ptr, 0, 8, 1e6, 1, None) == 2


def test_simulated_measurement_reconstructs_on_device():
    """object -> |FFT2| -> Poisson modulus -> shrink-wrap without leaving the
    device equals shrink-wrap of the same modulus from host memory."""
    w, h = 256, 192
    d = _modulus(w, h, 9)
    im.check(L().imresh_poisson_modulus(d.ptr, w, h, 1e9, 77, None))
    M = d.download(np.float32, (h, w))
    p = im.default_params(n_cycles=3, n_hio_cycles=5)
    r = im.Result()
    im.check(L().imresh_shrinkwrap_device(d.ptr, w, h, p, None, r))
    host = M.copy()
    r2 = im.Result()
    im.check(L().imresh_shrinkwrap(host.ctypes.data, w, h, p, None, r2))
    assert np.array_equal(d.download(np.float32, (h, w)), host)
    assert r.last_error == r2.last_error and r.cycles_run == r2.cycles_run
