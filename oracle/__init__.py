"""CPU oracle for the shrink-wrap/HIO hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed
CPU baseline).  The product library never routes through it.

``lib()`` loads ``oracle/liboracle.so`` (the C restatement in ``oracle.c``,
citing the reference file:line for every function); ``ref()`` loads
``oracle/_ref/libimresh_ref.so`` (the reference's own CPU sources compiled by
``make -C oracle ref``) when it was built, else returns None.

Parity pins per function are in ``tests/test_oracle.py``.  The FFT is the
project's own (FFTW3f, which the reference uses, is absent): it is pinned
against numpy's pocketfft, closed forms and a brute-force DFT, while its
bit-level agreement with FFTW is *parity unpinned*.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_uint, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


class SwStats(ctypes.Structure):
    _fields_ = [("cycles_run", c_uint), ("last_error", c_float),
                ("last_total", c_float), ("last_masked", c_float)]


def _declare(lib):
    for s, ct in (("f32", c_float), ("f64", c_double)):
        P = c_void_p
        sig = {
            "orc_gauss_kernel": (c_int, [c_double, P, c_uint]),
            "orc_blur_h": (None, [P, c_uint, c_uint, c_double]),
            "orc_blur_v": (None, [P, c_uint, c_uint, c_double]),
            "orc_blur": (None, [P, c_uint, c_uint, c_double]),
            "orc_vmax": (ct, [P, c_uint]),
            "orc_vmin": (ct, [P, c_uint]),
            "orc_vsum": (ct, [P, c_uint]),
            "orc_complex_norm": (None, [P, P, c_uint]),
            "orc_apply_modulus": (None, [P, P, c_uint]),
            "orc_hio_constraint": (None, [P, P, P, c_uint, ct]),
            "orc_cutoff": (None, [P, c_uint, ct, ct, ct]),
            "orc_hio_error": (c_float, [P, P, c_uint, c_int, POINTER(c_float), POINTER(c_float)]),
            "orc_fft": (c_int, [P, P, c_uint, c_int]),
            "orc_fft2d": (c_int, [P, P, c_uint, c_uint, c_int]),
            "orc_copy_to_real": (None, [P, P, c_uint]),
            "orc_copy_from_real": (None, [P, P, c_uint]),
            "orc_make_mask": (c_int, [P, P, c_uint, c_uint, c_float, c_float]),
            "orc_hio_iterations": (c_int, [P, P, P, P, c_uint, c_uint, ct, c_uint]),
            "orc_shrinkwrap": (c_int, [P, c_uint, c_uint, c_uint, c_float, c_float, c_float,
                                       c_float, c_float, c_float, c_uint, POINTER(SwStats)]),
            "orc_diffraction_intensity": (c_int, [P, c_uint, c_uint]),
            "orc_random_phase": (None, [P, P, c_uint, c_uint]),
            "orc_hio_standalone": (c_int, [P, P, c_uint, c_uint, c_uint, c_float, c_float,
                                           POINTER(SwStats)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, f"{name}_{s}")
            f.restype = res
            f.argtypes = args
    P = c_void_p
    for name, res, args in (("orc_philox4x32_10", None, [P, P, P]),
                            ("orc_exp_d", c_double, [c_double]),
                            ("orc_log_d", c_double, [c_double]),
                            ("orc_sqrt_d", c_double, [c_double]),
                            ("orc_log_factorial", c_double, [c_double]),
                            ("orc_poisson", c_double, [c_double, c_uint64, c_uint64]),
                            ("orc_sq_sum_fixed", c_double, [P, ctypes.c_size_t]),
                            ("orc_poisson_modulus", None, [P, ctypes.c_size_t, c_double, c_uint64]),
                            ("orc_poisson_many", None, [P, ctypes.c_size_t, c_uint64, c_uint64, P])):
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib.orc_fft_shift.restype = None
    lib.orc_fft_shift.argtypes = [c_void_p, c_uint, c_uint, c_uint]
    lib.orc_sigma_schedule.restype = None
    lib.orc_sigma_schedule.argtypes = [c_float, c_float, c_uint, c_void_p]
    lib.orc_num_threads.restype = c_int
    lib.orc_set_fft_threads.restype = None
    lib.orc_set_fft_threads.argtypes = [c_int]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


def ref():
    """The reference's own CPU sources compiled into oracle/_ref, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libimresh_ref.so")
        if not os.path.exists(path):
            return None
        r = ctypes.CDLL(path)
        P = c_void_p
        r.ref_gauss_kernel_f32.restype = c_int
        r.ref_gauss_kernel_f32.argtypes = [c_double, P, c_uint]
        for n in ("ref_blur_h_f32", "ref_blur_v_f32", "ref_blur_f32"):
            getattr(r, n).restype = None
            getattr(r, n).argtypes = [P, c_uint, c_uint, c_double]
        for n in ("ref_vmax_f32", "ref_vmin_f32", "ref_vsum_f32"):
            getattr(r, n).restype = c_float
            getattr(r, n).argtypes = [P, c_uint]
        r.ref_complex_norm_f32.restype = None
        r.ref_complex_norm_f32.argtypes = [P, P, c_uint]
        r.ref_apply_modulus_f32.restype = None
        r.ref_apply_modulus_f32.argtypes = [P, P, c_uint]
        r.ref_hio_error_f32.restype = c_float
        r.ref_hio_error_f32.argtypes = [P, P, c_uint, c_int, POINTER(c_float), POINTER(c_float)]
        _REF = r
    return _REF


# ---------------------------------------------------------------- helpers
def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _p(a):
    return a.ctypes.data_as(c_void_p)


def _real(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _cplx_view(a, dtype):
    """any ndarray -> interleaved (re, im) array of `dtype`, a fresh copy
    (real input gets a zero imaginary part)."""
    ct = np.complex64 if np.dtype(dtype) == np.float32 else np.complex128
    return np.array(np.asarray(a), dtype=ct, order="C", copy=True).view(dtype)


def _to_complex(a, dtype, shape):
    ct = np.complex64 if np.dtype(dtype) == np.float32 else np.complex128
    return a.view(ct).reshape(shape)


def set_fft_threads(n: int):
    lib().orc_set_fft_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def gauss_kernel(sigma, dtype=np.float32):
    L = lib()
    f = getattr(L, f"orc_gauss_kernel_{_sfx(dtype)}")
    n = f(float(sigma), None, 0)
    w = np.zeros(n, dtype=dtype)
    f(float(sigma), _p(w), n)
    return w


def blur(data, sigma, which="hv"):
    d = np.array(data, copy=True)
    dtype = d.dtype
    h, w = d.shape
    name = {"hv": "orc_blur", "h": "orc_blur_h", "v": "orc_blur_v"}[which]
    getattr(lib(), f"{name}_{_sfx(dtype)}")(_p(d), w, h, float(sigma))
    return d


def vmax(d):
    d = np.ascontiguousarray(d)
    return getattr(lib(), f"orc_vmax_{_sfx(d.dtype)}")(_p(d), d.size)


def vmin(d):
    d = np.ascontiguousarray(d)
    return getattr(lib(), f"orc_vmin_{_sfx(d.dtype)}")(_p(d), d.size)


def vsum(d):
    d = np.ascontiguousarray(d)
    return getattr(lib(), f"orc_vsum_{_sfx(d.dtype)}")(_p(d), d.size)


def complex_norm(c, dtype=np.float32):
    ci = _cplx_view(c, dtype)
    out = np.empty(ci.size // 2, dtype=dtype)
    getattr(lib(), f"orc_complex_norm_{_sfx(dtype)}")(_p(out), _p(ci), out.size)
    return out.reshape(np.shape(c))


def apply_modulus(c, modulus, dtype=np.float32):
    ci = _cplx_view(c, dtype)
    m = _real(modulus, dtype)
    getattr(lib(), f"orc_apply_modulus_{_sfx(dtype)}")(_p(ci), _p(m), m.size)
    return _to_complex(ci, dtype, np.shape(c))


def hio_constraint(gprev, gprime, mask, beta, dtype=np.float32):
    gp = _cplx_view(gprev, dtype)
    g = _cplx_view(gprime, dtype)
    m = _real(mask, dtype)
    getattr(lib(), f"orc_hio_constraint_{_sfx(dtype)}")(_p(gp), _p(g), _p(m), m.size, beta)
    return _to_complex(gp, dtype, np.shape(gprev))


def cutoff(d, threshold, lower=1.0, upper=0.0):
    d = np.array(d, copy=True)
    getattr(lib(), f"orc_cutoff_{_sfx(d.dtype)}")(_p(d), d.size, threshold, lower, upper)
    return d


def hio_error(gprime, mask, invert=False, dtype=np.float32):
    g = _cplx_view(gprime, dtype)
    m = _real(mask, dtype)
    tot, nm = c_float(), c_float()
    e = getattr(lib(), f"orc_hio_error_{_sfx(dtype)}")(_p(g), _p(m), m.size, int(invert),
                                                       ctypes.byref(tot), ctypes.byref(nm))
    return e, tot.value, nm.value


def fft(x, sign=-1, dtype=np.float32):
    xi = _cplx_view(x, dtype)
    out = np.empty_like(xi)
    getattr(lib(), f"orc_fft_{_sfx(dtype)}")(_p(out), _p(xi), xi.size // 2, sign)
    return _to_complex(out, dtype, np.shape(x))


def fft2d(x, sign=-1, dtype=np.float32):
    h, w = np.shape(x)
    xi = _cplx_view(x, dtype)
    out = np.empty_like(xi)
    getattr(lib(), f"orc_fft2d_{_sfx(dtype)}")(_p(out), _p(xi), w, h, sign)
    return _to_complex(out, dtype, (h, w))


def make_mask(gprime, sigma, cut, dtype=np.float32):
    h, w = np.shape(gprime)
    g = _cplx_view(gprime, dtype)
    m = np.empty(w * h, dtype=dtype)
    getattr(lib(), f"orc_make_mask_{_sfx(dtype)}")(_p(g), _p(m), w, h, sigma, cut)
    return m.reshape(h, w)


def hio_iterations(gprime, gprev, mask, modulus, beta, n_iter, dtype=np.float32):
    """n_iter HIO iterations from a given state; returns (gprime, gprev)."""
    h, w = np.shape(gprime)
    g = _cplx_view(gprime, dtype)
    gp = _cplx_view(gprev, dtype)
    m = _real(mask, dtype)
    M = _real(modulus, dtype)
    getattr(lib(), f"orc_hio_iterations_{_sfx(dtype)}")(_p(g), _p(gp), _p(m), _p(M), w, h,
                                                        beta, n_iter)
    return _to_complex(g, dtype, (h, w)), _to_complex(gp, dtype, (h, w))


def shrink_wrap(modulus, n_cycles=20, target_error=1e-5, hio_beta=0.9, cutoff_autocorel=0.04,
                cutoff=0.2, sigma0=3.0, sigma_change=0.01, n_hio_cycles=20, dtype=np.float32):
    """CPU shrinkWrap (shrinkWrap.cpp:111-361).  Returns (object, SwStats)."""
    io = np.array(modulus, dtype=dtype, copy=True)
    h, w = io.shape
    st = SwStats()
    rc = getattr(lib(), f"orc_shrinkwrap_{_sfx(dtype)}")(
        _p(io), w, h, n_cycles, target_error, hio_beta, cutoff_autocorel, cutoff, sigma0,
        sigma_change, n_hio_cycles, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_shrinkwrap returned {rc}")
    return io, st


def diffraction_intensity(x, dtype=np.float32):
    """libs::diffractionIntensity (diffractionIntensity.cpp:49-80): |FFT2(x)|
    of a real H x W object, unnormalized, frequency 0 at [0, 0] (a copy)."""
    io = np.array(x, dtype=dtype, copy=True, order="C")
    h, w = io.shape
    rc = getattr(lib(), f"orc_diffraction_intensity_{_sfx(dtype)}")(_p(io), w, h)
    if rc != 0:
        raise RuntimeError(f"orc_diffraction_intensity returned {rc}")
    return io


def random_phase(norm, dtype=np.float32):
    """addRandomPhase (hybridInputOutput.cpp:126-197): |F| * exp(i arg FFT(rand field))."""
    M = _real(norm, dtype)
    h, w = M.shape
    out = np.zeros(2 * w * h, dtype=dtype)
    getattr(lib(), f"orc_random_phase_{_sfx(dtype)}")(_p(M), _p(out), w, h)
    return _to_complex(out, dtype, (h, w))


def hio_standalone(modulus, mask, n_cycles=0, target_error=1e-6, beta=0.9, dtype=np.float32):
    """libs::hybridInputOutput (hybridInputOutput.cpp:200-306) with a given
    uint8 mask.  Returns (object, SwStats) with cycles_run = HIO iterations."""
    io = np.array(modulus, dtype=dtype, copy=True)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    h, w = io.shape
    st = SwStats()
    rc = getattr(lib(), f"orc_hio_standalone_{_sfx(dtype)}")(_p(io), _p(m), w, h, n_cycles,
                                                               target_error, beta, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_hio_standalone returned {rc}")
    return io, st


def fft_shift(a):
    """libs::fftShift (fftShift.cpp:42-61) of a float32 or complex64 H x W array (a copy)."""
    x = np.array(a, copy=True, order="C")
    if x.dtype not in (np.float32, np.complex64):
        raise ValueError("float32 or complex64")
    h, w = x.shape
    lib().orc_fft_shift(x.ctypes.data_as(c_void_p), w, h, x.itemsize)
    return x


def philox4x32_10(ctr, key):
    """Philox4x32-10 of a 4-word counter and a 2-word key."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def exp_d(x):
    return lib().orc_exp_d(float(x))


def log_d(x):
    return lib().orc_log_d(float(x))


def sqrt_d(x):
    return lib().orc_sqrt_d(float(x))


def log_factorial(k):
    return lib().orc_log_factorial(float(k))


def poisson_many(lam, seed, idx0=0):
    """Poisson samples of the counter-based stream: element idx0 + j draws
    with rate lam[j]."""
    l = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.empty_like(l)
    lib().orc_poisson_many(l.ctypes.data, l.size, seed, idx0, out.ctypes.data)
    return out


def sq_sum_fixed(m):
    a = np.ascontiguousarray(m, dtype=np.float32)
    return lib().orc_sq_sum_fixed(a.ctypes.data, a.size)


def poisson_modulus(m, photons, seed):
    """The measurement simulation: m <- sqrt(Poisson(photons m^2 / sum m^2)) (a copy)."""
    a = np.array(m, dtype=np.float32, copy=True, order="C")
    lib().orc_poisson_modulus(a.ctypes.data, a.size, float(photons), seed)
    return a


def sigma_schedule(sigma0, sigma_change, n):
    out = np.empty(n, dtype=np.float32)
    lib().orc_sigma_schedule(sigma0, sigma_change, n, _p(out))
    return out
