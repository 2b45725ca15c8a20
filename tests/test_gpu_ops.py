"""GPU parity of the per-operator entry points against the CPU oracle.

Elementwise operators, blur, max/min and the cut-off are bit-exact given
identical inputs (the library rounds exactly as the reference's C++); sums
are checked against a float64 sum (1e-6 relative); the FFT against numpy's
float64 pocketfft.  Known-answer tests follow the reference's own
(testVectorReduce.cu, testGaussian.cpp).
"""
import ctypes

import numpy as np
import pytest

import imresh_amd as im
import oracle
from imresh_amd import synthetic as syn
from helpers import FLT_EPS, KAT_SEED, RAND_MAX, compare_float_array, log_spaced_points, rand, srand

pytestmark = pytest.mark.gpu
B = im.DeviceBuffer
L = im.lib


def _f(x):
    return ctypes.byref(x)


def run_norm(c):
    n = c.size
    dc, do = B.from_array(c), B(4 * n)
    im.check(L().imresh_complex_norm(do.ptr, dc.ptr, n, None))
    return do.download(np.float32, c.shape)


def test_complex_norm_bit_exact():
    rng = np.random.default_rng(1)
    n = 100003
    c = (rng.standard_normal(n) * 10 + 1j * rng.standard_normal(n)).astype(np.complex64)
    c[::17] = 0
    c[5] = complex(3, 4)
    got = run_norm(c)
    assert np.array_equal(got, oracle.complex_norm(c))
    assert got[5] == 5.0 and got[0] == 0.0


def test_apply_modulus_bit_exact_and_zero():
    rng = np.random.default_rng(2)
    n = 65536 + 7
    c = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    c[::13] = 0
    M = rng.random(n, dtype=np.float32) * 5
    dc, dM = B.from_array(c), B.from_array(M)
    im.check(L().imresh_apply_complex_modulus(dc.ptr, dc.ptr, dM.ptr, n, None))
    got = dc.download(np.complex64, n)
    exp = oracle.apply_modulus(c, M)
    assert np.array_equal(got.view(np.float32), exp.view(np.float32))
    assert np.all(got[::13] == 0)          # norm == 0 -> norm = 1 -> stays 0
    r = oracle.ref()
    if r is not None:
        ci = c.view(np.float32).copy()
        r.ref_apply_modulus_f32(ci.ctypes.data, M.ctypes.data, n)
        assert np.array_equal(got.view(np.float32), ci)


def test_hio_constraint_bit_exact():
    rng = np.random.default_rng(3)
    n = 50001
    g = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    gp = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    mask = (rng.random(n) < 0.4).astype(np.float32)
    gr = g.view(np.float32).reshape(-1, 2)
    gr[::11, 0] = -0.0                      # Re g' = -0.0: "< 0" is false
    dgp, dg, dm = B.from_array(gp), B.from_array(g), B.from_array(mask)
    im.check(L().imresh_hio_domain_constraints(dgp.ptr, dg.ptr, dm.ptr, n, np.float32(0.9), None))
    got = dgp.download(np.complex64, n)
    exp = oracle.hio_constraint(gp, g, mask, 0.9)
    assert np.array_equal(got.view(np.float32), exp.view(np.float32))


def test_cutoff_strict_less_bit_exact():
    rng = np.random.default_rng(4)
    n = 40000
    d = rng.random(n, dtype=np.float32)
    t = d[123]                              # values equal to the threshold map to upper
    dd = B.from_array(d)
    im.check(L().imresh_cutoff(dd.ptr, n, t, 1.0, 0.0, None))
    got = dd.download(np.float32, n)
    assert np.array_equal(got, oracle.cutoff(d, t, 1.0, 0.0))
    assert got[123] == 0.0


def test_copy_real_part_roundtrip():
    rng = np.random.default_rng(5)
    n = 1000
    x = rng.standard_normal(n).astype(np.float32)
    dx, dc = B.from_array(x), B(8 * n)
    im.check(L().imresh_copy_to_real_part(dc.ptr, dx.ptr, n, None))
    c = dc.download(np.complex64, n)
    assert np.array_equal(c.real, x) and np.all(c.imag == 0)
    dy = B(4 * n)
    im.check(L().imresh_copy_from_real_part(dy.ptr, dc.ptr, n, None))
    assert np.array_equal(dy.download(np.float32, n), x)


# ------------------------------------------------------------ reductions
def _reduce(name, d):
    dd = B.from_array(d)
    out = ctypes.c_float()
    im.check(getattr(L(), name)(dd.ptr, d.size, _f(out), None), name)
    return np.float32(out.value)


def test_reduce_length_one():
    """testVectorReduce.cu:93-99"""
    x = np.array([0.1234567], dtype=np.float32)
    for name in ("imresh_vector_max", "imresh_vector_min", "imresh_vector_sum"):
        assert _reduce(name, x) == x[0]


def test_reduce_obvious_value_kat():
    """testVectorReduce.cu:83-187: random U[-0.5, 0.5) data (srand 350471643),
    one slot set to +-7.37519 at rand() % n; max/min must be exact."""
    nmax = 1 << 22
    srand(KAT_SEED)
    data = np.empty(nmax, dtype=np.float32)
    # (float) rand() / RAND_MAX - 0.5f, vectorised over the glibc stream
    raw = np.array([rand() for _ in range(nmax)], dtype=np.float64)
    data[:] = (raw.astype(np.float32) / np.float32(RAND_MAX)).astype(np.float32) - np.float32(0.5)
    dd = B.from_array(data)
    for n in log_spaced_points(2, nmax, 30):
        pos = rand() % n
        for v, name in ((7.37519, "imresh_vector_max"), (-7.37519, "imresh_vector_min")):
            saved = data[pos]
            data[pos] = np.float32(v)
            dd.upload(data[:n])
            out = ctypes.c_float()
            im.check(getattr(L(), name)(dd.ptr, n, _f(out), None))
            assert np.float32(out.value) == np.float32(v), (n, name)
            assert np.float32(out.value) == (oracle.vmax(data[:n]) if v > 0 else oracle.vmin(data[:n]))
            data[pos] = saved


def test_reduce_max_min_exact_random():
    rng = np.random.default_rng(6)
    for n in (1, 63, 64, 65, 1000, 123457, 1 << 20):
        d = (rng.standard_normal(n) * 100).astype(np.float32)
        assert _reduce("imresh_vector_max", d) == oracle.vmax(d)
        assert _reduce("imresh_vector_min", d) == oracle.vmin(d)


def test_reduce_sum_within_1e6():
    rng = np.random.default_rng(7)
    for n in (1, 1000, 1 << 20, 3 << 20):
        d = rng.random(n, dtype=np.float32)
        exact = float(np.sum(d.astype(np.float64)))
        got = float(_reduce("imresh_vector_sum", d))
        assert abs(got - exact) <= 1e-6 * abs(exact), (n, got, exact)


# ------------------------------------------------------------ HIO error
def _kat_mask_and_data(nmax):
    """testVectorReduce.cu:297-338: random bit mask (MSB first per 32-bit
    word, srand(350471643)); masked pixels = 3 + 4i, others random."""
    srand(KAT_SEED)
    nwords = (nmax + 31) // 32
    words = np.array([rand() % 0xFFFFFFFF for _ in range(nwords)], dtype=np.uint64)
    bits = ((words[:, None] >> np.arange(31, -1, -1, dtype=np.uint64)[None, :]) & 1).astype(np.float32)
    mask = bits.reshape(-1)[:nmax].copy()
    data = np.empty(nmax, dtype=np.complex64)
    for i in range(nmax):
        if mask[i]:
            data[i] = complex(3.0, 4.0)
        else:
            re = np.float32(rand()) / np.float32(RAND_MAX)
            im_ = np.float32(rand()) / np.float32(RAND_MAX)
            data[i] = complex(re, im_)
    return mask, data


def test_hio_error_pythagorean_kat():
    nmax = 1 << 18
    mask, data = _kat_mask_and_data(nmax)
    dg, dmf, dmu = B.from_array(data), B.from_array(mask), B.from_array(mask.astype(np.uint8))
    last = 0.0
    for n in log_spaced_points(2, nmax, 30):
        for fn, dm in (("imresh_calculate_hio_error", dmf), ("imresh_calculate_hio_error_u8", dmu)):
            err, tot, nm = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
            im.check(getattr(L(), fn)(dg.ptr, dm.ptr, n, 0, _f(err), _f(tot), _f(nm), None))
            if tot.value < 16777216:
                assert nm.value * 5 == tot.value, (n, fn)
                assert nm.value == float(mask[:n].sum())
                assert nm.value >= last
            e_cpu, t_cpu, n_cpu = oracle.hio_error(data[:n], mask[:n])
            assert nm.value == n_cpu
            assert abs(tot.value - t_cpu) <= np.sqrt(n) * FLT_EPS * max(abs(t_cpu), 1e-30)
        last = nm.value


def test_hio_error_random_vs_float64():
    rng = np.random.default_rng(8)
    n = 1 << 20
    g = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    m = (rng.random(n) < 0.3).astype(np.float32)
    for invert in (0, 1):
        err, tot, nm = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        dg, dm = B.from_array(g), B.from_array(m)
        im.check(L().imresh_calculate_hio_error(dg.ptr, dm.ptr, n, invert, _f(err), _f(tot), _f(nm), None))
        w = (1 - m) if invert else m
        exact = float(np.sum(w.astype(np.float64) * np.abs(g.astype(np.complex128))))
        assert abs(tot.value - exact) <= 1e-6 * exact
        assert nm.value == float(w.sum())
        assert abs(err.value - np.float32(np.sqrt(np.float32(tot.value))) / nm.value) == 0


# ------------------------------------------------------------ blur
BLUR_SIZES = [(1, 1), (1, 5), (5, 1), (7, 3), (37, 31), (234, 512), (100, 1021), (333, 64)]
BLUR_SIGMAS = [0.1, 0.5, 1.0, 1.3, 1.7, 2.0, 3.0, 4.0, 6.0, 6.4, 10.0]


@pytest.mark.parametrize("h,w", BLUR_SIZES)
def test_blur_bit_exact_vs_oracle(h, w):
    rng = np.random.default_rng(h * 1000 + w)
    d = (rng.random((h, w), dtype=np.float32) - 0.5).astype(np.float32)
    dd = B(d.nbytes)
    for s in BLUR_SIGMAS:
        for which, fn in (("h", "imresh_gaussian_blur_horizontal"),
                          ("v", "imresh_gaussian_blur_vertical"), ("hv", "imresh_gaussian_blur")):
            dd.upload(d)
            im.check(getattr(L(), fn)(dd.ptr, w, h, s, None), fn)
            got = dd.download(np.float32, (h, w))
            exp = oracle.blur(d, s, which)
            assert np.array_equal(got, exp), (h, w, s, which, float(np.abs(got - exp).max()))


@pytest.mark.parametrize("h,w", [(4096, 4096), (4097, 4096), (8192, 2100)])
def test_blur_bit_exact_both_tile_orders(h, w):
    """Planes of up to 16 Mi elements take the blur's tiles in contiguous
    ranges per XCD, larger planes in row-major order: 4096^2 (the largest of
    the first kind), 4097 x 4096 and 8192 x 2100 (the second) are bit-exact
    against the oracle either way, for the 19- and 35-tap blurs."""
    rng = np.random.default_rng(h * 7 + w)
    d = (rng.random((h, w), dtype=np.float32) - 0.5).astype(np.float32)
    dd = B(d.nbytes)
    for s in (3.0, 6.0):
        for which, fn in (("v", "imresh_gaussian_blur_vertical"), ("hv", "imresh_gaussian_blur")):
            dd.upload(d)
            im.check(getattr(L(), fn)(dd.ptr, w, h, s, None), fn)
            got = dd.download(np.float32, (h, w))
            exp = oracle.blur(d, s, which)
            assert np.array_equal(got, exp), (h, w, s, which, float(np.abs(got - exp).max()))


def test_blur_concurrent_threads():
    """Four host threads, each on its own stream, blur with the same new
    sigmas at once (first uploads race for the same table slots); every
    result is bit-exact with the oracle."""
    import threading
    rng = np.random.default_rng(4)
    h, w = 64, 96
    d = (rng.random((h, w), dtype=np.float32) - 0.5).astype(np.float32)
    sigmas = [0.35 + 0.0137 * i for i in range(24)]
    exp = {s: oracle.blur(d, s, "hv") for s in sigmas}
    errors = []

    def worker(k):
        st = im.stream_create()
        dd = B(d.nbytes)
        try:
            for s in sigmas[k:] + sigmas[:k]:
                dd.upload(d)
                im.check(L().imresh_gaussian_blur(dd.ptr, w, h, s, st), "blur")
                im.sync(st)
                if not np.array_equal(dd.download(np.float32, (h, w)), exp[s]):
                    errors.append((k, s))
        except Exception as e:   # reported below, from the main thread
            errors.append((k, repr(e)))
        finally:
            dd.close()
            im.stream_destroy(st)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors[:4]


def test_blur_weight_table_and_overflow():
    """The per-operator blur keeps each sigma's weights in a per-device table
    of 128 slots (cudaGaussian.cu:64-298 keeps 128 per device); past that it
    stages the weights per call.  Both paths, a repeated sigma, and a call on
    a created stream are bit-exact with the oracle."""
    rng = np.random.default_rng(128)
    h, w = 37, 31
    d = (rng.random((h, w), dtype=np.float32) - 0.5).astype(np.float32)
    dd = B(d.nbytes)
    sigmas = [float(x) for x in np.linspace(0.2, 9.5, 140)] + [3.0, 3.0, 0.2]
    st = im.stream_create()
    try:
        for i, s in enumerate(sigmas):
            which, fn = (("h", "imresh_gaussian_blur_horizontal"), ("v", "imresh_gaussian_blur_vertical"),
                         ("hv", "imresh_gaussian_blur"))[i % 3]
            dd.upload(d)
            stream = st if i % 2 else None
            im.check(getattr(L(), fn)(dd.ptr, w, h, s, stream), fn)
            im.sync(stream)
            got = dd.download(np.float32, (h, w))
            assert np.array_equal(got, oracle.blur(d, s, which)), (i, s, which)
    finally:
        im.stream_destroy(st)


def test_blur_dirac_deltas_kat():
    """testGaussian.cpp:217-337: spaced deltas reproduce the weights."""
    for s in (0.1, 0.5, 1, 1.3, 1.7, 2, 3, 4):
        w = im.gaussian_kernel(s)
        k, r = len(w), (len(w) - 1) // 2
        for ncols, nrows in ((31, 5), (234, 10), (1021, 3)):
            if ncols < k:
                continue
            d = np.zeros((nrows, ncols), dtype=np.float32)
            d[:, r:ncols - r:k] = 1
            sol = np.zeros_like(d)
            c = 0
            while c + k - 1 < ncols:
                sol[:, c:c + k] = w
                c += k
            dd = B.from_array(d)
            im.check(L().imresh_gaussian_blur_horizontal(dd.ptr, ncols, nrows, s, None))
            got = dd.download(np.float32, d.shape)
            assert compare_float_array(got, sol) <= FLT_EPS * 30
            # vertical on the transpose
            dt = np.ascontiguousarray(d.T)
            dd = B.from_array(dt)
            im.check(L().imresh_gaussian_blur_vertical(dd.ptr, nrows, ncols, s, None))
            got = dd.download(np.float32, dt.shape)
            assert compare_float_array(got, np.ascontiguousarray(sol.T)) <= FLT_EPS * 30


def test_blur_constant_rows_identical():
    """testGaussian.cpp:400-544: a constant stays constant, all outputs identical."""
    for s in (0.5, 1.7, 3, 4):
        d = np.full((37, 513), 7.37519, dtype=np.float32)
        dd = B.from_array(d)
        im.check(L().imresh_gaussian_blur(dd.ptr, 513, 37, s, None))
        got = dd.download(np.float32, d.shape)
        assert np.all(got == got.flat[0])
        assert compare_float_array(got, d) <= FLT_EPS * 30


def test_blur_too_wide_kernel_is_an_error():
    d = B(4 * 64)
    st = L().imresh_gaussian_blur(d.ptr, 8, 8, 11.0, None)   # 65 taps
    assert st == 7  # IMRESH_ERR_KERNEL_TOO_LARGE


# ------------------------------------------------------------ FFT
FFT_SIZES = [(1, 1), (2, 4), (8, 8), (3, 5), (12, 60), (16, 32), (64, 64), (128, 256),
             (96, 200), (384, 768), (512, 512), (1024, 64), (2048, 16), (16, 2048),
             (4096, 8), (8, 8192), (160, 192), (320, 640), (1280, 1536), (2560, 8), (8, 3072),
             (5120, 4), (4, 6144), (8192, 8), (3000, 40)]


def _fft_gpu(x, direction):
    h, w = x.shape
    d = B.from_array(x.astype(np.complex64))
    im.check(L().imresh_fft2d(d.ptr, d.ptr, w, h, direction, None), f"fft {w}x{h}")
    return d.download(np.complex64, (h, w))


@pytest.mark.parametrize("h,w", FFT_SIZES)
def test_fft2d_vs_numpy(h, w):
    rng = np.random.default_rng(h * 7 + w)
    x = (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)
    x64 = x.astype(np.complex128)
    fwd = _fft_gpu(x, -1)
    exp = np.fft.fft2(x64)
    tol = 2e-7 * max(1.0, np.log2(h * w))
    assert np.linalg.norm(fwd - exp) / np.linalg.norm(exp) < tol
    inv = _fft_gpu(x, +1)
    exp = np.fft.ifft2(x64) * (h * w)
    assert np.linalg.norm(inv - exp) / np.linalg.norm(exp) < tol


@pytest.mark.parametrize("h,w", [(64, 64), (96, 200), (1000, 12), (2048, 2048), (8192, 8)])
def test_fft2d_and_modulus_out_of_place(h, w):
    """out != in gives the in-place result bit for bit and leaves the input
    untouched (imresh_fft2d and imresh_apply_complex_modulus accept both)."""
    rng = np.random.default_rng(h + w)
    x = (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)
    M = rng.random((h, w), dtype=np.float32)
    for direction in (-1, +1):
        src, dst, ip = B.from_array(x), B(x.nbytes), B.from_array(x)
        im.check(L().imresh_fft2d(dst.ptr, src.ptr, w, h, direction, None))
        im.check(L().imresh_fft2d(ip.ptr, ip.ptr, w, h, direction, None))
        assert np.array_equal(dst.download(np.complex64, (h, w)).view(np.uint32),
                              ip.download(np.complex64, (h, w)).view(np.uint32))
        assert np.array_equal(src.download(np.complex64, (h, w)).view(np.uint32), x.view(np.uint32))
    src, dst, dM = B.from_array(x), B(x.nbytes), B.from_array(M)
    im.check(L().imresh_apply_complex_modulus(dst.ptr, src.ptr, dM.ptr, h * w, None))
    exp = oracle.apply_modulus(x.reshape(-1), M.reshape(-1))
    assert np.array_equal(dst.download(np.complex64, h * w).view(np.uint32), exp.view(np.uint32))
    assert np.array_equal(src.download(np.complex64, (h, w)).view(np.uint32), x.view(np.uint32))


def test_fft2d_closed_forms():
    h, w = 64, 128
    d = np.zeros((h, w), np.complex64)
    d[0, 0] = 1
    assert np.allclose(_fft_gpu(d, -1), 1, atol=0)          # delta -> all ones, exactly
    c = np.full((h, w), 0.5, np.complex64)
    F = _fft_gpu(c, -1)
    assert F[0, 0] == 0.5 * h * w and np.abs(F).ravel()[1:].max() < 1e-3
    ky, kx = 5, 17
    yy, xx = np.mgrid[0:h, 0:w]
    e = np.exp(2j * np.pi * (ky * yy / h + kx * xx / w)).astype(np.complex64)
    F = _fft_gpu(e, -1)
    assert abs(F[ky, kx] - h * w) < 1e-2 * h * w
    F[ky, kx] = 0
    assert np.abs(F).max() < 1e-3 * h * w


def test_fft_unsupported_size_is_an_error():
    """Every length up to 8192 has a transform (own plan or Bluestein; the
    16384-point convolution plan is not a line length itself); anything
    else is an error."""
    d = B(8 * 8193 * 2)
    assert L().imresh_fft2d(d.ptr, d.ptr, 8193, 2, -1, None) == 3   # IMRESH_ERR_UNSUPPORTED_SIZE
    assert L().imresh_fft2d(d.ptr, d.ptr, 2, 8193, -1, None) == 3
    assert L().imresh_fft2d(d.ptr, d.ptr, 16384, 1, -1, None) == 3
    assert L().imresh_fft_supported(2048) == 1 and L().imresh_fft_supported(7) == 1
    assert L().imresh_fft_supported(4096) == 1 and L().imresh_fft_supported(8192) == 1
    assert L().imresh_fft_supported(4097) == 1 and L().imresh_fft_supported(8191) == 1
    assert L().imresh_fft_supported(8193) == 0 and L().imresh_fft_supported(0) == 0
    assert L().imresh_fft_supported(16384) == 0


@pytest.mark.parametrize("h,w", [(8, 8), (6, 10), (5, 3), (1, 7), (7, 1), (64, 96), (97, 131)])
def test_fft_shift_matches_oracle(h, w):
    """fftShift.cpp:42-61, float and complex elements, bit-exact."""
    rng = np.random.default_rng(h + 100 * w)
    for a in (rng.standard_normal((h, w)).astype(np.float32),
              (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)):
        d = B.from_array(a)
        im.check(L().imresh_fft_shift(d.ptr, w, h, a.itemsize, None))
        assert np.array_equal(d.download(a.dtype, a.shape), oracle.fft_shift(a))
    assert L().imresh_fft_shift(d.ptr, w, h, 3, None) == 2   # element size not 4 or 8


# ---------------------------------------------------------------- forward model
DIFF_SIZES = [(64, 64), (256, 256), (2048, 2048), (64, 96), (97, 131), (300, 200),
              (1280, 1536), (8, 8192), (1, 7), (8192, 16), (3000, 40)]


@pytest.mark.parametrize("h,w", DIFF_SIZES)
def test_diffraction_intensity_vs_oracle(h, w):
    """imresh_diffraction_intensity (diffractionIntensity.cpp:84-135) against
    the float64 oracle: relative L2 error within the fp32 FFT bound
    2e-7 * log2(N); the host and device entry points agree bit for bit."""
    rng = np.random.default_rng(h * 13 + w)
    x = rng.random((h, w)).astype(np.float32)
    exp = oracle.diffraction_intensity(x.astype(np.float64), dtype=np.float64)
    got = im.diffraction_intensity(x.copy())
    tol = 2e-7 * max(1.0, np.log2(h * w))
    assert np.linalg.norm(got - exp) / np.linalg.norm(exp) < tol
    d = B.from_array(x)
    im.check(L().imresh_diffraction_intensity_device(d.ptr, w, h, None))
    assert np.array_equal(d.download(np.float32, (h, w)), got)
    y = x.copy()
    im.check(L().imresh_diffraction_intensity(y.ctypes.data, w, h, None, 1))   # async
    im.sync()
    assert np.array_equal(y, got)


def test_diffraction_intensity_closed_forms_and_errors():
    h, w = 64, 128
    d = np.zeros((h, w), np.float32)
    d[0, 0] = 1   # every twiddle product is exact: all ones
    assert np.array_equal(im.diffraction_intensity(d.copy()), np.ones((h, w), np.float32))
    d[0, 0] = 0
    d[3, 5] = 1   # |exp(i phi)| rounded in fp32
    assert np.max(np.abs(im.diffraction_intensity(d.copy()) - 1)) <= 4 * FLT_EPS
    z = np.zeros((h, w), np.float32)
    assert np.array_equal(im.diffraction_intensity(z), z)
    assert L().imresh_diffraction_intensity(None, w, h, None, 0) == 1           # IMRESH_ERR_NULL
    assert L().imresh_diffraction_intensity(d.ctypes.data, 0, h, None, 0) == 2  # IMRESH_ERR_INVALID
    assert L().imresh_diffraction_intensity(d.ctypes.data, 8193, 1, None, 0) == 3


def test_device_forward_model_feeds_reconstruction():
    """Object -> forward model -> shrink-wrap entirely in device memory gives
    the same result as the host path (forward model on the host buffer, then
    imresh_shrinkwrap), bit for bit."""
    obj, _ = syn.workload(128, 96, n_shapes=5, central=48, index=4)
    p = im.default_params(n_cycles=3, n_hio_cycles=4)
    host = obj.astype(np.float32).copy()
    im.diffraction_intensity(host)
    r = im.Result()
    im.check(L().imresh_shrinkwrap(host.ctypes.data, 128, 96, p, None, r))
    d = B.from_array(obj.astype(np.float32))
    im.check(L().imresh_diffraction_intensity_device(d.ptr, 128, 96, None))
    r2 = im.Result()
    im.check(L().imresh_shrinkwrap_device(d.ptr, 128, 96, p, None, r2))
    assert np.array_equal(d.download(np.float32, (96, 128)), host)
    assert r.cycles_run == r2.cycles_run and r.last_error == r2.last_error


# ---------------------------------------------------------------- bit-packed mask
def _unpack_msb_first(words, n):
    """unpackBitMask (testVectorReduce.cu:199-228): element i = bit 31 - i%32 of word i/32."""
    b = np.unpackbits(np.asarray(words, ">u4").view(np.uint8), bitorder="big")
    return b[:n]


def test_pack_mask_layout():
    """0x33333333 unpacks to 0,0,1,1,... (testVectorReduce.cu:230-246); the
    packing kernel is its inverse, a partial last word zero-filled."""
    assert list(_unpack_msb_first([0x33333333], 30)) == [(i // 2) % 2 for i in range(30)]
    rng = np.random.default_rng(4)
    for n in (1, 31, 32, 33, 1000, 4097):
        m = (rng.random(n) < 0.5).astype(np.uint8) * rng.integers(1, 255, n).astype(np.uint8)
        dm, db = B.from_array(m), B(4 * ((n + 31) // 32))
        im.check(L().imresh_pack_mask(dm.ptr, n, db.ptr, None))
        words = db.download(np.uint32, ((n + 31) // 32,))
        assert np.array_equal(_unpack_msb_first(words, n), (m != 0).astype(np.uint8))
        tail = (n % 32)
        if tail:
            assert words[-1] & ((1 << (32 - tail)) - 1) == 0


def test_hio_error_bit_packed_pythagorean():
    """testVectorReduce.cu:248-410: random packed words (srand(350471643),
    rand() % UINT_MAX), masked elements 3+4i and unmasked random ones; for
    log-spaced prefix lengths the bit-packed error gives total = 5 * count
    while exact, the count of the unpacked mask, and equals the unsigned char
    and float mask versions."""
    n_max = 1 << 20
    nw = (n_max + 31) // 32
    srand(350471643)
    words = np.array([rand() % 0xFFFFFFFF for _ in range(nw)], np.uint32)
    mask = _unpack_msb_first(words, n_max)
    data = np.empty(n_max, np.complex64)
    idx = np.flatnonzero(mask == 0)
    data[mask == 1] = 3 + 4j
    r = np.array([rand() for _ in range(2 * idx.size)], np.float64) / RAND_MAX
    data[idx] = (r[0::2].astype(np.float32) + 1j * r[1::2].astype(np.float32)).astype(np.complex64)
    dd, db = B.from_array(data), B.from_array(words)
    du, df = B.from_array(mask.astype(np.uint8)), B.from_array(mask.astype(np.float32))
    err, tot, cnt = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    last = 0.0
    for n in log_spaced_points(2, n_max, 50):
        im.check(L().imresh_calculate_hio_error_bits(dd.ptr, db.ptr, n, 0, ctypes.byref(err),
                                                     ctypes.byref(tot), ctypes.byref(cnt), None))
        exp_cnt = float(mask[:n].sum())
        assert cnt.value == exp_cnt and cnt.value >= last
        last = cnt.value
        if tot.value < 16777216:
            assert tot.value == 5 * cnt.value
        got = (err.value, tot.value, cnt.value)
        for dm, fn in ((du, "imresh_calculate_hio_error_u8"), (df, "imresh_calculate_hio_error")):
            e2, t2, c2 = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
            im.check(getattr(L(), fn)(dd.ptr, dm.ptr, n, 0, ctypes.byref(e2), ctypes.byref(t2),
                                      ctypes.byref(c2), None))
            assert (e2.value, t2.value, c2.value) == got, (n, fn)
    # the inverted mask counts the complement
    im.check(L().imresh_calculate_hio_error_bits(dd.ptr, db.ptr, 1000, 1, ctypes.byref(err),
                                                 ctypes.byref(tot), ctypes.byref(cnt), None))
    assert cnt.value == 1000 - mask[:1000].sum()


# ---------------------------------------------------------------- pooled buffers
def test_pool_cap_and_release():
    """Idle per-size workspaces are kept up to IMRESH_OPT_POOL_MIB and freed
    least recently used first beyond it; imresh_release_cached frees every
    idle workspace and batch buffer.  Results do not depend on pooling."""
    cached = ctypes.c_uint64()
    p = im.default_params(n_cycles=2, n_hio_cycles=3)
    Ms = [syn.workload(n, n, n_shapes=4, central=n // 2, index=n)[1] for n in (192, 256, 320)]

    def run_all():
        outs = []
        for M in Ms:
            x = M.copy()
            im.check(L().imresh_shrinkwrap(x.ctypes.data, x.shape[1], x.shape[0], p, None, None))
            outs.append(x)
        return outs

    ref = run_all()
    try:
        im.check(L().imresh_set_option(2, 1))          # IMRESH_OPT_POOL_MIB: 1 MiB
        im.check(L().imresh_cached_bytes(ctypes.byref(cached)))
        assert cached.value <= 1 << 20
        again = run_all()
        # the workspace returned last may still count as busy (its completion
        # event was just recorded); the next trim, here re-applying the cap,
        # frees it
        im.check(L().imresh_stream_sync(None))
        im.check(L().imresh_set_option(2, 1))
        im.check(L().imresh_cached_bytes(ctypes.byref(cached)))
        assert cached.value <= 1 << 20
        for a, b in zip(ref, again):
            assert np.array_equal(a, b)
        im.check(L().imresh_set_option(2, 32768))
        run_all()
        im.shrink_wrap_batch([M.copy() for M in Ms], p)
        im.check(L().imresh_cached_bytes(ctypes.byref(cached)))
        assert cached.value > 1 << 20
        im.check(L().imresh_release_cached())
        im.check(L().imresh_cached_bytes(ctypes.byref(cached)))
        assert cached.value == 0
        for a, b in zip(ref, run_all()):
            assert np.array_equal(a, b)
        assert L().imresh_set_option(2, -1) == 2
    finally:
        im.check(L().imresh_set_option(2, 32768))


# ------------------------------------------------------------ against the reference's own code
@pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref (the reference's CPU sources) not built")
def test_ops_bit_exact_vs_reference_sources():
    """The GPU operators against the reference's own CPU sources compiled
    unchanged (gaussian.cpp, calcGaussianKernel.cpp, vectorReduce.cpp and
    the templates of vectorElementwise.cpp / hybridInputOutput.cpp, in
    oracle/_ref): blur (both passes and 2-D), weights, max, min, complex norm
    and modulus projection bit for bit; sum and HIO error within 1e-6 of
    float64 (the reference's OpenMP reductions are order-dependent)."""
    r = oracle.ref()
    rng = np.random.default_rng(20261202)
    for h, w in ((37, 31), (234, 512), (100, 1021)):
        d = (rng.random((h, w), dtype=np.float32) - 0.25).astype(np.float32)
        dd = B(d.nbytes)
        for s in (0.5, 1.7, 3.0, 6.0, 10.0):
            k = im.gaussian_kernel(s)
            kr = np.zeros(64, np.float32)
            assert r.ref_gauss_kernel_f32(s, kr.ctypes.data, 64) == len(k)
            assert np.array_equal(k, kr[:len(k)])
            for fn, rf in (("imresh_gaussian_blur_horizontal", r.ref_blur_h_f32),
                           ("imresh_gaussian_blur_vertical", r.ref_blur_v_f32),
                           ("imresh_gaussian_blur", r.ref_blur_f32)):
                dd.upload(d)
                im.check(getattr(L(), fn)(dd.ptr, w, h, s, None), fn)
                exp = d.copy()
                rf(exp.ctypes.data, w, h, s)
                assert np.array_equal(dd.download(np.float32, (h, w)), exp), (h, w, s, fn)
    n = 300007
    x = (rng.standard_normal(n) * 50).astype(np.float32)
    assert _reduce("imresh_vector_max", x) == np.float32(r.ref_vmax_f32(x.ctypes.data, n))
    assert _reduce("imresh_vector_min", x) == np.float32(r.ref_vmin_f32(x.ctypes.data, n))
    y = rng.random(n, dtype=np.float32)
    exact = float(np.sum(y.astype(np.float64)))
    assert abs(float(_reduce("imresh_vector_sum", y)) - exact) <= 1e-6 * exact
    assert abs(float(r.ref_vsum_f32(y.ctypes.data, n)) - exact) <= 1e-6 * exact
    c = (rng.standard_normal(n) * 3 + 1j * rng.standard_normal(n)).astype(np.complex64)
    c[::97] = 0
    got = run_norm(c)
    exp = np.empty(n, np.float32)
    r.ref_complex_norm_f32(exp.ctypes.data, c.ctypes.data, n)
    assert np.array_equal(got, exp)
    M = rng.random(n, dtype=np.float32) * 4
    dc, dM = B.from_array(c), B.from_array(M)
    im.check(L().imresh_apply_complex_modulus(dc.ptr, dc.ptr, dM.ptr, n, None))
    ci = c.copy()
    r.ref_apply_modulus_f32(ci.ctypes.data, M.ctypes.data, n)
    assert np.array_equal(dc.download(np.complex64, n).view(np.uint32), ci.view(np.uint32))
    mask = (rng.random(n) < 0.3).astype(np.float32)
    dg, dm = B.from_array(c), B.from_array(mask)
    err, tot, nm = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    im.check(L().imresh_calculate_hio_error(dg.ptr, dm.ptr, n, 0, _f(err), _f(tot), _f(nm), None))
    rt, rn = ctypes.c_float(), ctypes.c_float()
    r.ref_hio_error_f32(c.ctypes.data, mask.ctypes.data, n, 0, ctypes.byref(rt), ctypes.byref(rn))
    exact_tot = float(np.sum(mask.astype(np.float64) * np.abs(c.astype(np.complex128))))
    assert nm.value == rn.value == float(mask.sum())
    for t in (tot.value, rt.value):
        assert abs(t - exact_tot) <= 1e-6 * exact_tot


def test_reductions_concurrent_threads():
    """Four host threads on their own streams call the synchronous
    reductions and the HIO error at once, repeatedly (their scratch slots
    are pooled per device); every result equals the single-threaded one."""
    import threading
    rng = np.random.default_rng(44)
    n = 200003
    x = (rng.standard_normal(n) * 10).astype(np.float32)
    c = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    mask = (rng.random(n) < 0.5).astype(np.float32)
    dx, dc, dm = B.from_array(x), B.from_array(c), B.from_array(mask)
    exp = (_reduce("imresh_vector_max", x), _reduce("imresh_vector_min", x), _reduce("imresh_vector_sum", x))
    e0, t0, n0 = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    im.check(L().imresh_calculate_hio_error(dc.ptr, dm.ptr, n, 0, _f(e0), _f(t0), _f(n0), None))
    errors = []

    def worker(k):
        st = im.stream_create()
        try:
            for _ in range(25):
                got = []
                for name in ("imresh_vector_max", "imresh_vector_min", "imresh_vector_sum"):
                    out = ctypes.c_float()
                    im.check(getattr(L(), name)(dx.ptr, n, _f(out), st), name)
                    got.append(np.float32(out.value))
                e, t, m = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
                im.check(L().imresh_calculate_hio_error(dc.ptr, dm.ptr, n, 0, _f(e), _f(t), _f(m), st))
                if tuple(got) != exp or (e.value, t.value, m.value) != (e0.value, t0.value, n0.value):
                    errors.append((k, got, e.value))
        except Exception as ex:   # reported below, from the main thread
            errors.append((k, repr(ex)))
        finally:
            im.stream_destroy(st)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors[:4]
