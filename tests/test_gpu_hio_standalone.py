"""GPU parity of the standalone HIO with a given support
(libs::hybridInputOutput, hybridInputOutput.cpp:200-306).

The first guess is |F| with the phases of the FFT of glibc's rand() field
after srand(2623091912), in both the product and the oracle; the GPU FFT of
that field and the device atan2/cos/sin round differently from the oracle's
FFT and libm, so states are compared with tolerances: the initial guess and
short runs (<= 3 iterations, where rounding has not been amplified) against
the fp64 oracle, early exits and iteration counts exactly.
"""
import numpy as np
import pytest

import imresh_amd as im
import oracle
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _case(w, h, index):
    obj, M = syn.workload(w, h, n_shapes=6, central=min(w, h) // 2, index=index)
    mask = (obj == 0).astype(np.uint8)
    return obj, M, mask


@pytest.mark.parametrize("w,h", [(64, 64), (96, 200), (256, 256), (97, 131)])
def test_first_guess_matches_oracle(w, h):
    """A target met at the first check: no iteration, output = Re(IFFT(first guess))."""
    obj, M, mask = _case(w, h, w + h)
    x = M.copy()
    r = im.hybrid_input_output(x, mask, n_cycles=0, target_error=1e30)
    ref, st = oracle.hio_standalone(M, mask, n_cycles=0, target_error=1e30, dtype=np.float64)
    e = syn.rel_l2(x, ref)
    print(f"{w}x{h}: first guess rel L2 vs fp64 oracle {e:.2e}; error {r.last_error:.6g} / {st.last_error:.6g}")
    assert r.cycles_run == st.cycles_run == 0
    assert e < 1e-5
    assert abs(r.last_error - st.last_error) <= 1e-5 * abs(st.last_error)
    assert r.last_masked == st.last_masked


@pytest.mark.parametrize("w,h,n", [(64, 64, 1), (128, 96, 3), (256, 256, 2), (97, 131, 2), (16, 8192, 2),
                                   (3000, 40, 2), (40, 3000, 2)])
def test_short_runs_match_oracle(w, h, n):
    obj, M, mask = _case(w, h, 7 * w + h)
    x = M.copy()
    r = im.hybrid_input_output(x, mask, n_cycles=n, target_error=0)
    ref, st = oracle.hio_standalone(M, mask, n_cycles=n, target_error=0, dtype=np.float64)
    e = syn.rel_l2(x, ref)
    print(f"{w}x{h} x{n}: rel L2 vs fp64 oracle {e:.2e}")
    assert r.cycles_run == st.cycles_run == n
    assert e < 1e-4


def test_checked_path_equals_fused_path():
    """With a target that is never met, the per-iteration checked sequence
    (separate inverse-row pass with error sums) must compute the same state
    as the fused row passes."""
    obj, M, mask = _case(128, 128, 3)
    a, b = M.copy(), M.copy()
    ra = im.hybrid_input_output(a, mask, n_cycles=5, target_error=0)
    rb = im.hybrid_input_output(b, mask, n_cycles=5, target_error=1e-30)
    assert ra.cycles_run == rb.cycles_run == 5
    assert syn.rel_l2(a, b) < 1e-6


def test_early_exit_count_matches_oracle():
    """A target between the errors of consecutive iterations: both stop after
    the same number of iterations (errors differ by < 1e-3 relative)."""
    obj, M, mask = _case(64, 64, 11)
    _, s1 = oracle.hio_standalone(M, mask, n_cycles=1, target_error=1e-30, dtype=np.float64)
    _, s2 = oracle.hio_standalone(M, mask, n_cycles=2, target_error=1e-30, dtype=np.float64)
    e1, e2 = s1.last_error, s2.last_error
    if not e2 < e1 * 0.99:
        pytest.skip(f"error not decreasing on this seed ({e1} -> {e2})")
    target = float(np.sqrt(e1 * e2))
    x = M.copy()
    r = im.hybrid_input_output(x, mask, n_cycles=0, target_error=target)
    _, st = oracle.hio_standalone(M, mask, n_cycles=0, target_error=target, dtype=np.float64)
    assert r.cycles_run == st.cycles_run == 2


def test_invalid_arguments():
    M = np.ones((16, 16), np.float32)
    mask = np.zeros((16, 16), np.uint8)
    L = im.lib()
    # unbounded iterations without a target would never end
    assert L.imresh_hybrid_input_output(M.ctypes.data, mask.ctypes.data, 16, 16, 0,
                                        np.float32(0), np.float32(0.9), None, None) == 2
    assert L.imresh_hybrid_input_output(None, mask.ctypes.data, 16, 16, 1, np.float32(0),
                                        np.float32(0.9), None, None) == 1
