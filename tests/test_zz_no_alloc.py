"""The allocation-free entry point imresh_shrinkwrap_device_noalloc (GPU).

It runs the launches of imresh_shrinkwrap_device with the reconstruction's
state in the library's fixed device array, so its output and result must
equal imresh_shrinkwrap_device's bit for bit: checked on two calls in a row,
the second with a smaller (Bluestein) size, so leftovers of the first call
in the shared array would show.  Sizes beyond the array and sizes whose FFT
tables were never built are refused with a status.
"""
import ctypes

import numpy as np
import pytest

import imresh_amd as im
from helpers import FLT_MIN
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu
L = im.lib
TOO_LARGE = 12
NOT_INITIALIZED = 8


def _regular(M, p):
    d = im.DeviceBuffer.from_array(M)
    r = im.shrink_wrap_device(d.ptr, M.shape[1], M.shape[0], p)
    return d.download(np.float32, M.shape), r


def _noalloc(M, p, stream=None):
    d = im.DeviceBuffer.from_array(M)
    r = im.Result()
    im.check(L().imresh_shrinkwrap_device_noalloc(d.ptr, M.shape[1], M.shape[0], ctypes.byref(p), stream,
                                                   ctypes.byref(r)), "noalloc")
    return d.download(np.float32, M.shape), r


def _same(a, ra, b, rb):
    assert np.array_equal(a, b)
    assert (ra.cycles_run, ra.last_error, ra.last_total, ra.last_masked) == \
        (rb.cycles_run, rb.last_error, rb.last_total, rb.last_masked)


def test_noalloc_equals_device_entry_two_calls():
    p1 = im.default_params(n_cycles=6, n_hio_cycles=10, target_error=FLT_MIN)
    _, M1 = syn.workload(512, 512, n_shapes=8, central=256, index=41)
    _, M2 = syn.workload(300, 200, n_shapes=6, central=100, index=42)
    a1, r1 = _regular(M1, p1)          # builds the 512^2 tables
    p2 = im.default_params(n_cycles=4, n_hio_cycles=7, target_error=FLT_MIN, sigma0=2.0)
    a2, r2 = _regular(M2, p2)
    b1, s1 = _noalloc(M1, p1)
    _same(a1, r1, b1, s1)
    b2, s2 = _noalloc(M2, p2)          # smaller, after the larger one in the same array
    _same(a2, r2, b2, s2)
    print(f"512^2: {s1.cycles_run} cycles, error {s1.last_error:.6g}; "
          f"300x200: {s2.cycles_run} cycles, error {s2.last_error:.6g}")


def test_noalloc_four_step_and_long_bluestein():
    """8192-row images (four-step column pass) and 3000-long lines
    (8192-point Bluestein convolutions) through the fixed array."""
    p = im.default_params(n_cycles=2, n_hio_cycles=3, target_error=FLT_MIN)
    for i, (w, h) in enumerate([(16, 8192), (40, 3000), (3000, 40)]):
        _, M = syn.workload(w, h, n_shapes=4, central=min(w, h) // 2, index=45 + i)
        a, ra = _regular(M, p)
        b, rb = _noalloc(M, p)
        _same(a, ra, b, rb)


def test_noalloc_full_config_2048_and_early_exit():
    """The 2048^2 workload (20 x 20) fits the array; an early exit as well."""
    _, M = syn.config_2048(0)
    p = im.default_params(n_cycles=20, target_error=FLT_MIN)
    a, ra = _regular(M, p)
    b, rb = _noalloc(M, p)
    _same(a, ra, b, rb)
    _, S = syn.workload(128, 128, n_shapes=6, central=64, index=43)
    pe = im.default_params(n_cycles=20, n_hio_cycles=5, target_error=1e6)
    a, ra = _regular(S, pe)
    b, rb = _noalloc(S, pe)
    _same(a, ra, b, rb)
    assert rb.cycles_run == 1


def test_noalloc_asynchronous_on_a_stream():
    _, M = syn.workload(256, 256, n_shapes=8, central=128, index=44)
    p = im.default_params(n_cycles=3, n_hio_cycles=5, target_error=FLT_MIN)
    a, _ = _regular(M, p)
    s = im.stream_create()
    try:
        d = im.DeviceBuffer.from_array(M)
        im.check(L().imresh_shrinkwrap_device_noalloc(d.ptr, 256, 256, ctypes.byref(p), s, None))
        im.sync(s)
        assert np.array_equal(d.download(np.float32, M.shape), a)
    finally:
        im.stream_destroy(s)


def test_noalloc_refusals():
    p = im.default_params(n_cycles=2)
    d = im.DeviceBuffer(4 * 16)
    # 4096 x 4096 needs 486 MB: beyond the array (checked before anything runs)
    assert L().imresh_shrinkwrap_device_noalloc(d.ptr, 4096, 4096, ctypes.byref(p), None, None) == TOO_LARGE
    many = im.default_params(n_cycles=257)
    assert L().imresh_shrinkwrap_device_noalloc(d.ptr, 4, 4, ctypes.byref(many), None, None) == TOO_LARGE
    # a size whose FFT tables were never built in this process
    assert L().imresh_shrinkwrap_device_noalloc(d.ptr, 5, 4079, ctypes.byref(p), None, None) == NOT_INITIALIZED
    assert L().imresh_shrinkwrap_device_noalloc(None, 4, 4, ctypes.byref(p), None, None) == 1
