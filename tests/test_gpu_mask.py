"""The support mask of the fused path, element by element (GPU).

The reconstruction builds each cycle's mask from |g'| with the blur kernel
(zero-padded per-cycle weight rows), per-tile maxima reduced on the device
into the threshold cut_c * max, and the cut-off fused into the load of the
cycle's first row pass.  imresh_debug_cycle_mask runs those same launches
and returns g' (the cycle's input), the blurred plane, the threshold and the
byte mask, which must equal the reference's mask update applied to that g'
(shrinkWrap.cpp:191-242, cudaShrinkWrap.cu:237-277) bit for bit: the CPU
oracle's make_mask and, where oracle/_ref is built, the reference's own
compiled gaussian.cpp / vectorReduce.cpp.
"""
import numpy as np
import pytest

import imresh_amd as im
import oracle
from helpers import FLT_MIN
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu

REF = oracle.ref()

CASES = [
    ("256-config", lambda: syn.config_256(0), 3.0),
    ("2048-config", lambda: syn.config_2048(0), 3.0),
    ("bluestein-700x300", lambda: syn.workload(700, 300, n_shapes=8, central=150, index=5), 3.0),
    ("plan-rows-bluestein-cols-1024x333", lambda: syn.workload(1024, 333, n_shapes=8, central=160, index=6), 3.0),
]


def _check_cycle(M, cycle, sigma0, n_hio=20, label=""):
    p = im.default_params(n_cycles=cycle + 1, target_error=FLT_MIN, sigma0=sigma0, n_hio_cycles=n_hio)
    g, mask, blurred, thr = im.debug_cycle_mask(M, cycle, p)
    h, w = M.shape
    sigma = float(oracle.sigma_schedule(sigma0, 0.01, cycle + 1)[cycle])
    cut = np.float32(0.04 if cycle == 0 else 0.2)
    norm = oracle.complex_norm(g)
    b = oracle.blur(norm, sigma)
    mx = oracle.vmax(b)
    exp_thr = np.float32(np.float32(mx) * cut)
    exp_mask = oracle.make_mask(g, sigma, float(cut))
    n_out = int(mask.sum())
    print(f"{label} cycle {cycle}: sigma {sigma:.6g}, taps {len(oracle.gauss_kernel(sigma))}, "
          f"threshold {thr:.7g}, outside {n_out}/{mask.size}, inside {mask.size - n_out}")
    assert np.all(np.isfinite(g))
    assert np.array_equal(blurred, b), int(np.sum(blurred != b))
    assert thr == exp_thr, (thr, exp_thr)
    assert np.array_equal(mask, exp_mask.astype(np.uint8)), int(np.sum(mask != exp_mask))
    # a mask that is neither empty nor full, so the comparison is not vacuous
    assert 0 < n_out < mask.size
    if REF is not None and max(h, w) <= 2048:
        r = norm.copy()
        REF.ref_blur_f32(r.ctypes.data, w, h, sigma)
        assert np.array_equal(r, blurred)
        assert np.float32(REF.ref_vmax_f32(r.ctypes.data, r.size)) == np.float32(mx)
    return g, mask


@pytest.mark.parametrize("name,make,sigma0", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("cycle", [0, 1])
def test_fused_mask_bit_exact(name, make, sigma0, cycle):
    obj, M = make()
    _check_cycle(M, cycle, sigma0, label=name)


def test_fused_mask_later_cycle_sigma_schedule():
    """Cycle 5 of the 256^2 workload: sigma 3 * 0.99^5 picked from the
    device's completed-cycle count (one weights row per cycle)."""
    obj, M = syn.config_256(1)
    _check_cycle(M, 5, 3.0, n_hio=4, label="256 cycle 5")


def test_fused_mask_8192_sigma6():
    """The large-image workload (8192^2, sigma0 6.0: 35 taps, radius 17):
    cycle 0's mask from the initial inverse transform, and cycle 1's after
    two HIO iterations, against the oracle's mask update on the same g'."""
    obj, M = syn.config_8192(0)
    _check_cycle(M, 0, 6.0, n_hio=2, label="8192 sigma0 6")
    _check_cycle(M, 1, 6.0, n_hio=2, label="8192 sigma0 6")
