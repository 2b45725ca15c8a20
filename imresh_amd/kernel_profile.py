"""Kernel profiling driver (run under rocprofv3 to collect counters).

    python -m imresh_amd.kernel_profile --size 2048 --reps 8

launches, in this order on the default stream:
  1. a plain 128 MiB device-to-device copy (hipMemcpyAsync, 4x the aggregate
     L2), the calibration kernel of known traffic (128 MiB read + 128 MiB
     written) for FETCH_SIZE / WRITE_SIZE;
  2. (after autotuning, or the variants given with --variants)
     imresh_profile_kernels: the fused HIO iteration as a reconstruction runs
     it (row pass and column pass alternating), then the cycle-boundary
     kernels (blur, closing row pass), each launch timed on its own execution
     with kernel-exact HIP events, and prints the mean durations.
"""
from __future__ import annotations

import argparse
import ctypes

import imresh_amd as im


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--reps", type=int, default=8)
    ap.add_argument("--variants", default="",
                    help="row,column kernel variants to pin instead of autotuning (e.g. 0,5)")
    a = ap.parse_args()
    L = im.lib()
    nbytes = 128 << 20
    src, dst = im.DeviceBuffer(nbytes), im.DeviceBuffer(nbytes)
    for _ in range(3):
        im.check(L.imresh_memcpy_d2d(dst.ptr, src.ptr, nbytes, None))
    im.sync()
    if a.variants:
        r, c = (int(x) for x in a.variants.split(","))
        im.check(L.imresh_set_variants(a.size, a.size, r, c), "imresh_set_variants")
    rv, cv = ctypes.c_int(), ctypes.c_int()
    im.check(L.imresh_autotune(a.size, a.size, None, ctypes.byref(rv), ctypes.byref(cv)))
    ms = (ctypes.c_float * 4)()
    im.check(L.imresh_profile_kernels(a.size, a.size, a.reps, ms, None))
    names = ["row pass (inverse rows, constraint, forward rows)",
             "column pass (forward cols, modulus, inverse cols)",
             "blur 2-D + tile max", "closing row pass (inverse rows, norm, error)"]
    print(f"{a.size}x{a.size}, variants row {rv.value} column {cv.value}")
    for n, t in zip(names, ms):
        print(f"  {n:50s} {t * 1e3:8.2f} us")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
