"""Pins of the CPU oracle (runs on the CPU, no GPU).

Each oracle function is checked against something independent of its own
formula: the reference's own CPU sources compiled into oracle/_ref (bit-exact
where the reference is deterministic), the known-answer tests of the
reference's test suite (testGaussian.cpp, testVectorReduce.cu), closed forms,
float64 brute force, numpy's pocketfft, and committed golden fixtures
(tests/golden/).  A plausible mistake -- a dropped tap, a wrong sign, a
transposed index, a wrong clamp, '<=' for '<' -- fails at least one of them.
"""
import math
import os

import numpy as np
import pytest

import oracle
from imresh_amd import synthetic as syn
from helpers import FLT_EPS, KAT_SEED, RAND_MAX, compare_float_array, log_spaced_points, rand, srand

REF = oracle.ref()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (reference sources absent)")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_golden.npz")


# ------------------------------------------------------------ Gaussian kernel
@pytest.mark.parametrize("sigma,taps", [(0.1, 1), (0.5, 3), (1.0, 7), (3.0, 19), (4.0, 25),
                                        (6.0, 35), (6.4, 37), (10.9, 63), (11.0, 65)])
def test_gauss_kernel_size_closed_form(sigma, taps):
    """calcGaussianKernel.cpp:63: radius ceil(2.884402748387961466*sigma - 0.5)."""
    w = oracle.gauss_kernel(sigma)
    assert len(w) == taps == 2 * math.ceil(2.884402748387961466 * sigma - 0.5) + 1


@pytest.mark.parametrize("sigma", [0.1, 0.5, 1.0, 1.3, 1.5, 1.7, 2.0, 2.9403, 3.0, 4.0, 6.0, 6.4])
def test_gauss_kernel_properties(sigma):
    w = oracle.gauss_kernel(sigma).astype(np.float64)
    r = (len(w) - 1) // 2
    assert abs(w.sum() - 1.0) < 1e-6
    assert np.array_equal(w, w[::-1])
    assert np.argmax(w) == r
    # shape: ratio of neighbours follows exp(-(2i+1)/(2 sigma^2)) (to float rounding)
    if r >= 2:
        ratio = w[r + 1] / w[r]
        assert abs(ratio - math.exp(-1.0 / (2 * sigma * sigma))) < 1e-6


@needs_ref
@pytest.mark.parametrize("sigma", [0.1, 0.5, 1.0, 1.3, 1.7, 2.0, 3.0, 4.0, 6.0, 6.4, 10.0])
def test_gauss_kernel_bit_exact_vs_reference(sigma):
    w = oracle.gauss_kernel(sigma)
    buf = np.zeros(64, np.float32)
    n = REF.ref_gauss_kernel_f32(sigma, buf.ctypes.data, 64)
    assert n == len(w) and np.array_equal(w, buf[:n])


def test_sigma_schedule_closed_form():
    """sigma <- max(1.5, (1 - 0.01) * sigma) in float (shrinkWrap.cpp:245)."""
    s = oracle.sigma_schedule(3.0, 0.01, 100)
    exp = [np.float32(3.0)]
    for _ in range(99):
        exp.append(max(np.float32(1.5), np.float32(np.float32(0.99) * exp[-1])))
    assert np.array_equal(s, np.array(exp, np.float32))
    assert s[0] == 3.0 and s[-1] == 1.5 and np.all(np.diff(s) <= 0)
    # 3 * 0.99^n >= 1.5 until n = 69 (reference note cudaGaussian.cu:84-86)
    assert s[68] > 1.5 and s[70] == 1.5


# ------------------------------------------------------------ blur
@needs_ref
@pytest.mark.parametrize("h,w", [(1, 1), (1, 7), (9, 1), (7, 3), (37, 31), (64, 1021), (500, 16)])
def test_blur_bit_exact_vs_reference(h, w):
    rng = np.random.default_rng(h * 31 + w)
    d = (rng.random((h, w), dtype=np.float32) - 0.5).astype(np.float32)
    for s in (0.1, 0.5, 1.0, 1.7, 3.0, 4.0, 6.0, 10.0):
        for which, fn in (("h", REF.ref_blur_h_f32), ("v", REF.ref_blur_v_f32), ("hv", REF.ref_blur_f32)):
            a = oracle.blur(d, s, which)
            b = d.copy()
            fn(b.ctypes.data, w, h, s)
            assert np.array_equal(a, b), (h, w, s, which)


def _brute_blur_rows(d, w):
    """float64 direct convolution with clamp-to-edge along the last axis."""
    r = (len(w) - 1) // 2
    n = d.shape[-1]
    idx = np.clip(np.arange(n)[:, None] + np.arange(-r, r + 1)[None, :], 0, n - 1)
    return (d[..., idx] * w.astype(np.float64)).sum(-1)


@pytest.mark.parametrize("sigma", [0.5, 1.3, 3.0, 6.4])
def test_blur_vs_float64_brute_force(sigma):
    rng = np.random.default_rng(11)
    d = rng.random((40, 57), dtype=np.float32)
    w = oracle.gauss_kernel(sigma)
    h = _brute_blur_rows(d.astype(np.float64), w)
    hv = _brute_blur_rows(h.T, w).T
    for which, exp in (("h", h), ("v", _brute_blur_rows(d.astype(np.float64).T, w).T), ("hv", hv)):
        got = oracle.blur(d, sigma, which)
        assert compare_float_array(got, exp) <= 30 * FLT_EPS, which


def test_blur_dirac_deltas_kat():
    """testGaussian.cpp:217-337: spaced deltas reproduce the weight vector."""
    for s in (0.1, 0.5, 1, 1.3, 1.7, 2, 3, 4):
        w = oracle.gauss_kernel(s)
        k, r = len(w), (len(w) - 1) // 2
        for ncols in (1, 2, 3, 5, 10, 31, 37, 234):
            if ncols < k:
                continue
            d = np.zeros((3, ncols), np.float32)
            d[:, r:ncols - r:k] = 1
            sol = np.zeros_like(d)
            c = 0
            while c + k - 1 < ncols:
                sol[:, c:c + k] = w
                c += k
            assert compare_float_array(oracle.blur(d, s, "h"), sol) <= FLT_EPS * 30
            assert compare_float_array(oracle.blur(np.ascontiguousarray(d.T), s, "v"),
                                       np.ascontiguousarray(sol.T)) <= FLT_EPS * 30


def test_blur_single_element_and_constants():
    """testGaussian.cpp:340-544."""
    srand(KAT_SEED)
    for s in (0.1, 0.5, 1, 1.3, 1.7, 2, 3, 4):
        x = np.array([[np.float32(rand()) / np.float32(RAND_MAX) - np.float32(0.5)]], np.float32)
        assert compare_float_array(oracle.blur(x, s, "h"), x) <= FLT_EPS * 30
        c = np.full((31, 513), 7.37519, np.float32)
        for which in ("h", "v", "hv"):
            b = oracle.blur(c, s, which)
            assert np.all(b == b.flat[0])
            assert compare_float_array(b, c) <= FLT_EPS * 30


def test_blur_bound_invariants():
    """testGaussian.cpp:134-148: min does not decrease, max does not increase,
    max neighbour difference does not grow."""
    srand(KAT_SEED)
    n = 200 * 150
    d = ((np.array([rand() for _ in range(n)], np.float64) / RAND_MAX) - 0.5).astype(np.float32)
    d = d.reshape(200, 150)
    b = oracle.blur(d, 3.0, "h")
    for row_b, row_d in zip(b, d):
        assert row_d.min() <= row_b.min() and row_d.max() >= row_b.max()
        assert np.abs(np.diff(row_b)).max() <= np.abs(np.diff(row_d)).max()


# ------------------------------------------------------------ reductions
def test_reduce_obvious_value_kat():
    """testVectorReduce.cu:83-187 on the oracle (exact)."""
    srand(KAT_SEED)
    nmax = 1 << 16
    data = ((np.array([rand() for _ in range(nmax)], np.float64) / RAND_MAX) - 0.5).astype(np.float32)
    assert oracle.vmax(data[:1]) == data[0] == oracle.vmin(data[:1]) == oracle.vsum(data[:1])
    for n in log_spaced_points(2, nmax, 30):
        pos = rand() % n
        sub = data[:n].copy()
        sub[pos] = np.float32(7.37519)
        assert oracle.vmax(sub) == np.float32(7.37519)
        sub[pos] = np.float32(-7.37519)
        assert oracle.vmin(sub) == np.float32(-7.37519)


@needs_ref
def test_reduce_vs_reference():
    rng = np.random.default_rng(3)
    for n in (1, 17, 1000, 100003):
        d = (rng.standard_normal(n) * 3).astype(np.float32)
        assert oracle.vmax(d) == REF.ref_vmax_f32(d.ctypes.data, n)
        assert oracle.vmin(d) == REF.ref_vmin_f32(d.ctypes.data, n)
        s = float(np.sum(d.astype(np.float64)))
        assert abs(oracle.vsum(d) - s) <= n * FLT_EPS * float(np.sum(np.abs(d)))


# ------------------------------------------------------------ HIO error
def test_hio_error_pythagorean_kat():
    """testVectorReduce.cu:297-413: masked pixels = 3+4i => total = 5*nmasked."""
    srand(KAT_SEED)
    nmax = 1 << 15
    words = np.array([rand() % 0xFFFFFFFF for _ in range((nmax + 31) // 32)], np.uint64)
    bits = ((words[:, None] >> np.arange(31, -1, -1, dtype=np.uint64)[None, :]) & 1)
    mask = bits.reshape(-1)[:nmax].astype(np.float32)
    g = np.empty(nmax, np.complex64)
    for i in range(nmax):
        g[i] = complex(3, 4) if mask[i] else complex(np.float32(rand()) / np.float32(RAND_MAX),
                                                     np.float32(rand()) / np.float32(RAND_MAX))
    for n in log_spaced_points(2, nmax, 25):
        e, tot, nm = oracle.hio_error(g[:n], mask[:n])
        assert nm * 5 == tot and nm == mask[:n].sum()
        assert e == np.float32(np.sqrt(np.float32(tot))) / np.float32(nm) or (nm == 0 and np.isnan(e))
        e2, tot2, nm2 = oracle.hio_error(g[:n], mask[:n], invert=True)
        assert nm2 == n - nm


@needs_ref
def test_hio_error_vs_reference():
    import ctypes
    rng = np.random.default_rng(5)
    n = 100000
    g = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    m = (rng.random(n) < 0.3).astype(np.float32)
    gi = g.view(np.float32).copy()
    tot, nm = ctypes.c_float(), ctypes.c_float()
    REF.ref_hio_error_f32(gi.ctypes.data, m.ctypes.data, n, 0, ctypes.byref(tot), ctypes.byref(nm))
    e, t, c = oracle.hio_error(g, m)
    assert c == nm.value
    assert abs(t - tot.value) <= math.sqrt(n) * FLT_EPS * abs(tot.value) * 10


# ------------------------------------------------------------ elementwise
@needs_ref
def test_norm_and_modulus_bit_exact_vs_reference():
    rng = np.random.default_rng(6)
    n = 50000
    c = (rng.standard_normal(n) * 7 + 1j * rng.standard_normal(n)).astype(np.complex64)
    c[::9] = 0
    out = np.empty(n, np.float32)
    ci = c.view(np.float32).copy()
    REF.ref_complex_norm_f32(out.ctypes.data, ci.ctypes.data, n)
    assert np.array_equal(out, oracle.complex_norm(c))
    M = rng.random(n, dtype=np.float32)
    REF.ref_apply_modulus_f32(ci.ctypes.data, M.ctypes.data, n)
    assert np.array_equal(ci, oracle.apply_modulus(c, M).view(np.float32))


def test_modulus_projection_closed_form():
    c = np.array([3 + 4j, 0, -6 + 8j, 1e-3j], np.complex64)
    M = np.array([10, 5, 5, 2], np.float32)
    out = oracle.apply_modulus(c, M)
    assert out[0] == 6 + 8j and out[1] == 0 and out[2] == -3 + 4j and out[3] == 2j


def test_hio_constraint_closed_form():
    """cudaVectorElementwise.cu:61-70: feedback where mask == 1 or Re g' < 0."""
    g = np.array([1 + 2j, -1 + 2j, 1 + 2j, complex(-0.0, 5)], np.complex64)
    gp = np.array([10 + 10j] * 4, np.complex64)
    mask = np.array([0, 0, 1, 0], np.float32)
    out = oracle.hio_constraint(gp, g, mask, 0.5)
    assert out[0] == 1 + 2j                     # inside support, Re >= 0: copy
    assert out[1] == 10.5 + 9j                  # Re < 0: gPrev - beta*g'
    assert out[2] == 9.5 + 9j                   # masked: gPrev - beta*g'
    assert out[3] == complex(-0.0, 5)           # Re = -0.0 is not < 0: copy


def test_cutoff_is_strict():
    d = np.array([0.0, 0.5, 0.49999997, 1.0], np.float32)
    assert np.array_equal(oracle.cutoff(d, np.float32(0.5), 1, 0), np.array([1, 0, 1, 0], np.float32))


# ------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 12, 16, 60, 64, 96, 97, 128, 200, 243, 1021,
                               1031, 2048])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    s = np.abs(ref).max()
    assert np.abs(oracle.fft(x, -1, np.float64) - ref).max() <= 1e-12 * s
    assert np.abs(oracle.fft(x, +1, np.float64) - np.fft.ifft(x) * n).max() <= 1e-12 * s
    assert np.linalg.norm(oracle.fft(x, -1, np.float32) - ref) <= 2e-6 * np.linalg.norm(ref)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 13, 16, 24, 64])
def test_fft_vs_brute_force_dft(n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    assert np.abs(oracle.fft(x, -1, np.float64) - dft).max() <= 1e-12 * max(1, np.abs(dft).max())


def test_fft2d_closed_forms_and_identities():
    h, w = 24, 40
    d = np.zeros((h, w), np.complex128)
    d[0, 0] = 1
    assert np.allclose(oracle.fft2d(d, -1, np.float64), 1, atol=1e-15)
    assert np.allclose(oracle.fft2d(np.full((h, w), 2.0), -1, np.float64)[0, 0], 2 * h * w)
    ky, kx = 3, 7
    yy, xx = np.mgrid[0:h, 0:w]
    e = np.exp(2j * np.pi * (ky * yy / h + kx * xx / w))
    F = oracle.fft2d(e, -1, np.float64)
    assert abs(F[ky, kx] - h * w) < 1e-9 and np.abs(np.delete(F.ravel(), ky * w + kx)).max() < 1e-9
    rng = np.random.default_rng(1)
    x = rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))
    y = rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))
    X = oracle.fft2d(x, -1, np.float64)
    assert abs(np.sum(np.abs(X) ** 2) - h * w * np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(X) ** 2)
    assert np.allclose(oracle.fft2d(X, +1, np.float64), h * w * x, atol=1e-9)
    assert np.allclose(oracle.fft2d(2 * x + 3j * y, -1, np.float64),
                       2 * X + 3j * oracle.fft2d(y, -1, np.float64), atol=1e-9)
    assert np.allclose(X, np.fft.fft2(x), atol=1e-10)


# ------------------------------------------------------------ shrinkWrap
def test_shrinkwrap_zero_cycles_is_ifft_real_part():
    rng = np.random.default_rng(2)
    M = rng.random((32, 48), dtype=np.float32)
    out, st = oracle.shrink_wrap(M, n_cycles=0)
    assert st.cycles_run == 0
    exp = np.real(np.fft.ifft2(M.astype(np.float64))) * M.size
    assert np.linalg.norm(out - exp) <= 1e-6 * np.linalg.norm(exp)


def test_shrinkwrap_early_exit_and_defaults():
    from imresh_amd import synthetic as syn
    obj, M = syn.workload(32, 32, n_shapes=4, central=16, index=3)
    out, st = oracle.shrink_wrap(M, n_cycles=7, target_error=1e9, n_hio_cycles=2)
    assert st.cycles_run == 1                       # shrinkWrap.cpp:328-329
    a, sa = oracle.shrink_wrap(M, n_cycles=2, target_error=0, hio_beta=0, cutoff_autocorel=0,
                               cutoff=0, sigma0=0, sigma_change=0, n_hio_cycles=0)
    b, sb = oracle.shrink_wrap(M, n_cycles=2)
    assert np.array_equal(a, b)                     # <= 0 selects the defaults (:134-140)


def test_shrinkwrap_fp32_tracks_fp64_on_short_runs():
    from imresh_amd import synthetic as syn
    obj, M = syn.workload(64, 64, n_shapes=8, central=32, index=1)
    a, sa = oracle.shrink_wrap(M, n_cycles=1, n_hio_cycles=3, target_error=1e-38)
    b, sb = oracle.shrink_wrap(M, n_cycles=1, n_hio_cycles=3, target_error=1e-38, dtype=np.float64)
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
    assert abs(sa.last_error - sb.last_error) <= 1e-5 * sb.last_error


def test_make_mask_cycle_zero_is_autocorrelation_support():
    """The cycle-0 mask (1 = outside support) thresholds the blurred |IFFT(M)|
    at 0.04 of its maximum; the DC peak (max of the autocorrelation) is inside."""
    from imresh_amd import synthetic as syn
    obj, M = syn.workload(64, 64, n_shapes=8, central=32, index=2)
    g = oracle.fft2d(M.astype(np.complex64), +1, np.float32)
    mask = oracle.make_mask(g, 3.0, 0.04)
    assert mask[0, 0] == 0 and set(np.unique(mask)) <= {0.0, 1.0}
    b = oracle.blur(oracle.complex_norm(g), 3.0)
    assert np.array_equal(mask, (b < np.float32(b.max() * np.float32(0.04))).astype(np.float32))


# ------------------------------------------------------------ golden fixtures
def test_golden_fixtures():
    """tests/golden/oracle_golden.npz (make_golden.py): blur outputs of the
    reference's own CPU sources, FFTs from numpy's pocketfft, a short
    float64 shrink-wrap run and the KAT inputs."""
    z = np.load(GOLDEN, allow_pickle=False)
    for key in [k for k in z.files if k.startswith("blur_in_")]:
        tag = key[len("blur_in_"):]
        sigma = float(z[f"blur_sigma_{tag}"])
        assert np.array_equal(oracle.blur(z[key], sigma, "hv"), z[f"blur_out_{tag}"]), tag
    x = z["fft_in"]
    assert np.allclose(oracle.fft2d(x, -1, np.float64), z["fft_out"], rtol=0, atol=1e-10)
    for s, w in zip(z["kern_sigmas"], z["kern_weights"]):
        n = len(oracle.gauss_kernel(s))
        assert np.array_equal(oracle.gauss_kernel(s), w[:n])
    out, st = oracle.shrink_wrap(z["sw_modulus"], n_cycles=2, n_hio_cycles=3,
                                 target_error=1e-38, dtype=np.float64)
    assert np.allclose(out, z["sw_out_f64"], rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- standalone HIO
def test_random_phase_matches_glibc_field_and_numpy_fft():
    """addRandomPhase (hybridInputOutput.cpp:126-197): the real field is
    (float) rand() / RAND_MAX after srand(2623091912) (glibc's generator,
    drawn here through ctypes), its
    phases come from a forward FFT (numpy, float64 here) and the output keeps
    the input modulus."""
    import helpers
    h, w = 24, 40
    rng = np.random.default_rng(5)
    M = rng.uniform(0.5, 2.0, (h, w)).astype(np.float32)
    helpers.srand(2623091912)
    field = np.array([np.float32(helpers.rand()) / np.float32(helpers.RAND_MAX) for _ in range(h * w)],
                     dtype=np.float32).reshape(h, w)
    F = np.fft.fft2(field.astype(np.float64))
    expect = M.astype(np.float64) * np.exp(1j * np.angle(F))
    got = oracle.random_phase(M, np.float64)
    assert np.allclose(np.abs(got), M, rtol=1e-12)
    assert np.max(np.abs(got - expect)) < 1e-5
    got32 = oracle.random_phase(M, np.float32)
    assert np.max(np.abs(got32 - expect)) < 1e-4


def test_hio_standalone_is_random_phase_then_hio_iterations():
    """hybridInputOutput (:200-306) = random phase -> inverse FFT -> gPrev =
    g' -> n HIO iterations; with target <= 0 it must equal that composition of
    the pinned operators exactly (float64, no exact zeros in |G'|)."""
    obj, M = syn.workload(32, 48, n_shapes=4, central=16, index=4)
    mask = (obj == 0).astype(np.uint8)
    n = 4
    out, st = oracle.hio_standalone(M, mask, n_cycles=n, target_error=0, dtype=np.float64)
    assert st.cycles_run == n
    g = oracle.fft2d(oracle.random_phase(M, np.float64), +1, np.float64)
    # the reference's beta is a float (0.9f), widened to the working precision
    g, gp = oracle.hio_iterations(g, g.copy(), mask.astype(np.float64), M.astype(np.float64),
                                  float(np.float32(0.9)), n, np.float64)
    assert np.array_equal(out, g.real)


def test_hio_standalone_early_exit_and_defaults():
    obj, M = syn.workload(32, 32, n_shapes=4, central=16, index=9)
    mask = (obj == 0).astype(np.uint8)
    # a target met at the first check: no iteration, output = Re(IFFT(random phase))
    out, st = oracle.hio_standalone(M, mask, n_cycles=0, target_error=1e30, dtype=np.float64)
    assert st.cycles_run == 0
    g = oracle.fft2d(oracle.random_phase(M, np.float64), +1, np.float64)
    assert np.array_equal(out, g.real)
    e, tot, nm = oracle.hio_error(g, mask.astype(np.float64), 0, np.float64)
    assert st.last_error == np.float32(e) and st.last_masked == mask.sum()
    # beta <= 0 selects 0.9
    a, _ = oracle.hio_standalone(M, mask, n_cycles=2, target_error=0, beta=0, dtype=np.float64)
    b, _ = oracle.hio_standalone(M, mask, n_cycles=2, target_error=0, beta=0.9, dtype=np.float64)
    assert np.array_equal(a, b)


# ---------------------------------------------------------------- fftShift
@pytest.mark.parametrize("h,w", [(4, 6), (8, 8), (6, 10)])
def test_fft_shift_is_numpy_fftshift_for_even_sizes(h, w):
    """fftShift.cpp:42-61 swaps (iy, ix) with ((iy+Ny/2)%Ny, (ix+Nx/2)%Nx):
    for even sizes that is numpy's fftshift, and an involution."""
    rng = np.random.default_rng(h * w)
    a = rng.standard_normal((h, w)).astype(np.float32)
    assert np.array_equal(oracle.fft_shift(a), np.fft.fftshift(a))
    c = (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)
    assert np.array_equal(oracle.fft_shift(c), np.fft.fftshift(c))
    assert np.array_equal(oracle.fft_shift(oracle.fft_shift(c)), c)


def test_fft_shift_odd_sizes_brute_force():
    """Odd sizes: the reference loop (rows iy < Ny/2 only) is not numpy's
    fftshift; checked against a direct Python transcription on tiny inputs."""
    for h, w in [(5, 3), (3, 4), (1, 7), (7, 1)]:
        a = np.arange(h * w, dtype=np.float32).reshape(h, w)
        exp = a.copy().ravel()
        for iy in range(h // 2):
            for ix in range(w):
                i, j = iy * w + ix, ((iy + h // 2) % h) * w + (ix + w // 2) % w
                exp[i], exp[j] = exp[j], exp[i]
        assert np.array_equal(oracle.fft_shift(a).ravel(), exp), (h, w)


# ---------------------------------------------------------------- forward model
@pytest.mark.parametrize("h,w", [(8, 8), (6, 10), (5, 3), (1, 7), (12, 1), (64, 96), (97, 131)])
def test_diffraction_intensity_vs_numpy(h, w):
    """diffractionIntensity.cpp:49-80 is |FFT2((x, 0))| in H rows x W
    columns without a shift: numpy's |fft2| of the real array (non-square and
    odd shapes fail a transposed or shifted layout)."""
    rng = np.random.default_rng(h * 31 + w)
    x = rng.random((h, w))
    got = oracle.diffraction_intensity(x, dtype=np.float64)
    exp = np.abs(np.fft.fft2(x))
    assert np.max(np.abs(got - exp)) <= 1e-12 * max(1.0, exp.max())
    g32 = oracle.diffraction_intensity(x.astype(np.float32))
    assert g32.dtype == np.float32
    assert np.linalg.norm(g32 - exp) / np.linalg.norm(exp) < 2e-7 * max(1.0, np.log2(h * w))


def test_diffraction_intensity_closed_forms():
    h, w = 48, 40
    # a point object anywhere has a flat diffraction pattern |F| = value
    d = np.zeros((h, w))
    d[7, 13] = 2.5
    assert np.allclose(oracle.diffraction_intensity(d, dtype=np.float64), 2.5, rtol=0, atol=1e-12)
    # an a x b box at the origin: product of two Dirichlet kernels
    a, b = 5, 9
    box = np.zeros((h, w))
    box[:a, :b] = 1
    ky = np.arange(h)[:, None]
    kx = np.arange(w)[None, :]

    def dirichlet(k, n, m):
        s = np.sin(np.pi * k / n)
        with np.errstate(invalid="ignore", divide="ignore"):
            r = np.abs(np.sin(np.pi * m * k / n) / s)
        return np.where(k % n == 0, float(m), r)

    exp = dirichlet(ky, h, a) * dirichlet(kx, w, b)
    got = oracle.diffraction_intensity(box, dtype=np.float64)
    assert np.max(np.abs(got - exp)) < 1e-10
    # Parseval: sum |F|^2 = N sum x^2, and |F| is centro-symmetric for real x
    rng = np.random.default_rng(5)
    x = rng.standard_normal((h, w))
    F = oracle.diffraction_intensity(x, dtype=np.float64)
    assert abs((F ** 2).sum() - h * w * (x ** 2).sum()) < 1e-9 * h * w * (x ** 2).sum()
    assert np.allclose(F, np.roll(F[::-1, ::-1], (1, 1), axis=(0, 1)), rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- measurement simulation
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors published with Random123
    (kat_vectors: zero, all-ones and pi-digit counters/keys)."""
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, exp in kat:
        assert list(oracle.philox4x32_10(ctr, key)) == exp


def test_ieee_elementary_functions_vs_libm():
    """The sampler's exp / log / sqrt / log k! (IEEE operations only) agree
    with libm within 2 ulp (exp, log, sqrt) and 1e-14 relative (log k!)."""
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(-40, 0, 2000), rng.uniform(-744, 700, 2000), [0.0, -1e-300, 1e-12]])
    for x in xs:
        e = math.exp(x)
        assert abs(oracle.exp_d(x) - e) <= 2 * math.ulp(e), x
    ys = np.concatenate([10.0 ** rng.uniform(-300, 300, 3000), rng.uniform(0.5, 2.0, 2000), [1.0, 2.0, 0.5]])
    for y in ys:
        l = math.log(y)
        assert abs(oracle.log_d(y) - l) <= 2 * math.ulp(l) + 1e-300, y
        if 1e-30 <= y <= 1e30:   # the sampler's domain (photons <= 1e30)
            s = math.sqrt(y)
            assert abs(oracle.sqrt_d(y) - s) <= math.ulp(s), y
    assert oracle.log_d(0.0) == -math.inf and oracle.exp_d(-800.0) == 0.0
    for k in list(range(0, 40)) + [50, 99, 1000, 12345, 10 ** 6, 3 * 10 ** 9]:
        lf = math.lgamma(k + 1)
        assert abs(oracle.log_factorial(k) - lf) <= 1e-14 * max(1.0, lf), k


@pytest.mark.parametrize("lam", [0.05, 0.7, 3.0, 9.99, 10.0, 37.5, 1000.0, 2.5e6])
def test_poisson_sampler_distribution(lam):
    """Inversion (lambda < 10) and PTRS (lambda >= 10) draw Poisson(lambda):
    mean and variance within 5 sigma, chi-square goodness of fit against the
    exact pmf (bins with >= 20 expected counts), and integer outputs."""
    from scipy import stats
    n = 200000
    k = oracle.poisson_many(np.full(n, lam), seed=2024 + int(lam * 10), idx0=7)
    assert np.all(k == np.floor(k)) and k.min() >= 0
    m, v = k.mean(), k.var()
    assert abs(m - lam) < 5 * math.sqrt(lam / n)
    # variance of the sample variance of Poisson: (lam + 2 lam^2 (n/(n-1))) / n ~ (lam + 2 lam^2) / n
    assert abs(v - lam) < 5 * math.sqrt((lam + 2 * lam * lam) / n)
    lo, hi = int(stats.poisson.ppf(1e-7, lam)), int(stats.poisson.ppf(1 - 1e-7, lam)) + 1
    edges = np.arange(lo, hi + 1)
    pmf = stats.poisson.pmf(edges, lam)
    # merge into bins of >= 20 expected counts
    bins, cur, acc = [], [], 0.0
    for e, p in zip(edges, pmf):
        cur.append(e)
        acc += p
        if acc * n >= 20:
            bins.append((cur, acc))
            cur, acc = [], 0.0
    if cur:
        bins[-1] = (bins[-1][0] + cur, bins[-1][1] + acc)
    obs = np.array([np.isin(k, b).sum() if len(b) < 64 else ((k >= b[0]) & (k <= b[-1])).sum()
                    for b, _ in bins], dtype=np.float64)
    exp = np.array([p for _, p in bins]) * n
    exp *= obs.sum() / exp.sum()
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert stats.chi2.sf(chi2, len(bins) - 1) > 1e-6, (chi2, len(bins))


def test_poisson_modulus_sum_order_and_photon_budget():
    """The fixed-order sum agrees with an exactly rounded sum to 1e-13; the
    simulated measurement is sqrt of integer counts whose total is
    Poisson(photons); zero and constant moduli behave."""
    rng = np.random.default_rng(3)
    m = (rng.random(300001) * 100).astype(np.float32)
    exact = math.fsum(float(x) * float(x) for x in m.astype(np.float64))
    assert abs(oracle.sq_sum_fixed(m) - exact) <= 1e-13 * exact
    out = oracle.poisson_modulus(m, 1e8, seed=5)
    k = out.astype(np.float64) ** 2
    assert np.allclose(k, np.round(k), rtol=0, atol=1e-3 * np.maximum(1, k))
    assert abs(np.round(k).sum() - 1e8) < 6 * math.sqrt(1e8)
    assert np.array_equal(oracle.poisson_modulus(np.zeros(100, np.float32), 1e6, 1), np.zeros(100, np.float32))
    c = oracle.poisson_modulus(np.ones(10000, np.float32), 5e4, 9)   # lambda = 5 everywhere
    assert abs((c.astype(np.float64) ** 2).mean() - 5.0) < 5 * math.sqrt(5.0 / 10000)
    # the seed selects the stream; the same seed repeats it
    assert np.array_equal(oracle.poisson_modulus(m[:1000], 1e6, 5), oracle.poisson_modulus(m[:1000], 1e6, 5))
    assert not np.array_equal(oracle.poisson_modulus(m[:1000], 1e6, 5), oracle.poisson_modulus(m[:1000], 1e6, 6))
