"""Randomized shapes (seeded, so every run checks the same cases): widths
and heights drawn from 1..700 and from the awkward corners (1, 2, primes,
plan lengths +-1, 4095/4096), mixing compiled-plan and Bluestein lengths in
one image.  For each shape:

* one fused HIO iteration from the same state against the float64 oracle
  (<= 1e-5 relative L2, the single-iteration bound) with gPrev exact against
  the float32 oracle;
* a short reconstruction run alone and inside a batch of mixed shapes gives
  the same bits.
"""
import numpy as np
import pytest

import imresh_amd as im
import oracle
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu
B = im.DeviceBuffer
L = im.lib


def _shapes():
    rng = np.random.default_rng(20261202)
    corners = [1, 2, 3, 7, 31, 63, 65, 127, 129, 255, 257, 509, 511, 513, 1023, 1025, 2039]
    out = [(int(rng.integers(1, 700)), int(rng.integers(1, 700))) for _ in range(14)]
    out += [(int(rng.choice(corners)), int(rng.integers(1, 300))) for _ in range(5)]
    out += [(int(rng.integers(1, 300)), int(rng.choice(corners))) for _ in range(5)]
    out += [(4095, 3), (5, 4096), (1, 1)]
    return out


SHAPES = _shapes()


def _state(w, h, index):
    _, M = syn.workload(w, h, n_shapes=4, central=max(1, min(w, h) // 2), index=index)
    g = oracle.fft2d(M.astype(np.complex64), +1, np.float32)
    mask = oracle.make_mask(g, 2.0, 0.04, np.float32)
    return M, g, mask


@pytest.mark.parametrize("w,h", SHAPES)
def test_one_iteration_random_shape(w, h):
    M, g, mask = _state(w, h, w * 1000 + h)
    dg, dgp = B.from_array(g), B.from_array(g.copy())
    dm, dM = B.from_array(mask.astype(np.uint8)), B.from_array(M)
    im.check(L().imresh_hio_iterations(dg.ptr, dgp.ptr, dm.ptr, dM.ptr, w, h, np.float32(0.9), 1, None))
    got = dg.download(np.complex64, (h, w))
    gp = dgp.download(np.complex64, (h, w))
    g64, _ = oracle.hio_iterations(g.astype(np.complex128), g.astype(np.complex128), mask.astype(np.float64),
                                   M.astype(np.float64), 0.9, 1, np.float64)
    _, gp32 = oracle.hio_iterations(g, g.copy(), mask, M, 0.9, 1, np.float32)
    if np.linalg.norm(g64) > 0:
        assert syn.rel_l2(got, g64) < 1e-5, (w, h)
    assert np.array_equal(gp.view(np.float32), gp32.view(np.float32)), (w, h)


def test_batch_of_random_shapes_equals_single():
    p = im.default_params(n_cycles=2, n_hio_cycles=3, target_error=float(np.finfo(np.float32).tiny))
    mods = [_state(w, h, 7 + i)[0] for i, (w, h) in enumerate(SHAPES)]
    alone = []
    for M in mods:
        x = M.copy()
        im.check(L().imresh_shrinkwrap(x.ctypes.data, x.shape[1], x.shape[0], p, None, None))
        alone.append(x)
    ios = [M.copy() for M in mods]
    imgs = (im.Image * len(ios))(*[im.Image(x.ctypes.data, x.shape[1], x.shape[0]) for x in ios])
    im.check(L().imresh_shrinkwrap_batch(imgs, len(ios), p, None, None))
    for x, y, (w, h) in zip(ios, alone, SHAPES):
        assert np.array_equal(x, y), (w, h)
