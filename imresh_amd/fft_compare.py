"""2-D FFT comparator: this library's LDS-staged Stockham row/column passes
(imresh_fft2d, two HBM passes) against rocFFT on the same device buffer.

    python -m imresh_amd.fft_compare --size 2048 8192

rocFFT (the ROCm system library librocfft.so, C API from rocfft/rocfft.h)
is only timed here, as a comparator; nothing in the library uses it.  Both
run a single-precision, unnormalized, in-place forward C2C transform of a
size x size complex64 buffer; times are wall-clock over `reps` back-to-back
transforms after a warm-up, with the stream synchronized on both sides.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import time

import numpy as np

import imresh_amd as im


class RocFFT:
    """Minimal ctypes binding of the rocFFT C API."""

    def __init__(self):
        self.lib = ctypes.CDLL("librocfft.so")
        L = self.lib
        for name in ("rocfft_setup", "rocfft_plan_create", "rocfft_plan_get_work_buffer_size",
                     "rocfft_execution_info_create", "rocfft_execution_info_set_work_buffer",
                     "rocfft_execution_info_set_stream", "rocfft_execute", "rocfft_plan_destroy",
                     "rocfft_execution_info_destroy"):
            getattr(L, name).restype = ctypes.c_int
        self._check(L.rocfft_setup(), "rocfft_setup")

    @staticmethod
    def _check(st, what):
        if st != 0:
            raise RuntimeError(f"{what} failed with rocfft_status {st}")

    def plan_2d(self, w: int, h: int):
        plan = ctypes.c_void_p()
        lengths = (ctypes.c_size_t * 2)(w, h)   # fastest dimension first
        # placement inplace = 0, transform complex_forward = 0, precision single = 0
        self._check(self.lib.rocfft_plan_create(ctypes.byref(plan), 0, 0, 0, ctypes.c_size_t(2), lengths,
                                                ctypes.c_size_t(1), None), "rocfft_plan_create")
        wsize = ctypes.c_size_t()
        self._check(self.lib.rocfft_plan_get_work_buffer_size(plan, ctypes.byref(wsize)),
                    "rocfft_plan_get_work_buffer_size")
        info = ctypes.c_void_p()
        self._check(self.lib.rocfft_execution_info_create(ctypes.byref(info)), "rocfft_execution_info_create")
        work = im.DeviceBuffer(wsize.value) if wsize.value else None
        if work is not None:
            self._check(self.lib.rocfft_execution_info_set_work_buffer(info, ctypes.c_void_p(work.ptr),
                                                                       ctypes.c_size_t(wsize.value)),
                        "rocfft_execution_info_set_work_buffer")
        return plan, info, work

    def execute(self, plan, info, ptr):
        bufs = (ctypes.c_void_p * 1)(ptr)
        self._check(self.lib.rocfft_execute(plan, bufs, None, info), "rocfft_execute")

    def destroy(self, plan, info):
        self.lib.rocfft_execution_info_destroy(info)
        self.lib.rocfft_plan_destroy(plan)


def compare(size: int, reps: int = 20) -> dict:
    n = size * size
    rng = np.random.default_rng(size)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    buf = im.DeviceBuffer.from_array(x)
    L = im.lib()

    def ours():
        im.check(L.imresh_fft2d(buf.ptr, buf.ptr, size, size, -1, None), "imresh_fft2d")

    def timeit(f):
        f()
        im.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        im.sync()
        return (time.perf_counter() - t0) / reps

    t_ours = timeit(ours)
    out = {"size": size, "imresh_fft2d_us": round(t_ours * 1e6, 2),
           "imresh_GBps_32B": round(32 * n / t_ours / 1e9, 1)}
    try:
        rf = RocFFT()
        plan, info, work = rf.plan_2d(size, size)
        t_roc = timeit(lambda: rf.execute(plan, info, buf.ptr))
        rf.destroy(plan, info)
        out.update({"rocfft_us": round(t_roc * 1e6, 2), "rocfft_GBps_32B": round(32 * n / t_roc / 1e9, 1),
                    "speedup_vs_rocfft": round(t_roc / t_ours, 3)})
    except (OSError, RuntimeError) as e:   # comparator only
        out["rocfft_error"] = str(e)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, nargs="+", default=[2048, 8192])
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    for s in a.size:
        print(json.dumps(compare(s, a.reps)), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
