"""Benchmark: HIO iterations/s of the shrink-wrap reconstruction at 2048^2 fp32.

One step = one full shrink-wrap reconstruction (20 cycles x 20 HIO
iterations = 400 HIO iterations) of the 2048 x 2048 synthetic workload
(64 rectangles/discs in the central 1024^2, Poisson-sampled at 1e9 photons,
modulus = sqrt(intensity)), with the input resident in HBM: the measured
modulus is restored on the device (one 16 MiB device-to-device copy) and the
reconstruction runs in place on the device buffer, as the reference's
in-place io buffer.  Early exit is disabled with target error = FLT_MIN, as
the reference's own test does.

Prints ONE JSON line (the benchmark contract) with two extra objects:
  roofline      the dominant fused kernel's measured duration (HIP events on
                the launch stream, imresh_profile_kernels) against its
                compulsory HBM bytes per launch and the 8 TB/s HBM peak;
                `traffic` from rocprofv3 FETCH_SIZE/WRITE_SIZE summaries in
                profiles/ when present
  cpu_baseline  the CPU oracle (C restatement of the reference CPU path,
                single-threaded FFT as the reference's FFTW plans, OpenMP
                elementwise loops) timed on a bounded sample of HIO
                iterations of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import imresh_amd as im  # noqa: E402
from imresh_amd import synthetic as syn  # noqa: E402

METRIC = "HIO iters/s at 2048² fp32; batched recon/s + p50/p99 latency vs plain loop"
W = H = 2048
N_CYCLES, N_HIO = 20, 20
HBM_PEAK_GBS = 8000.0
FLT_MIN = float(np.finfo(np.float32).tiny)

# compulsory HBM bytes per element of each fused kernel:
#   row pass  (inverse rows, constraint, forward rows): read X 8 + G 8 + mask 1, write G 8 + X 8
#   column pass (forward cols, modulus, inverse cols) : read X 8 + M 4, write X 8
KERNEL_BYTES = {"row_mid": 33, "col_proj": 20, "blur": 8, "row_last": 21}
KERNEL_NAMES = ["row_mid", "col_proj", "blur", "row_last"]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` at 2048^2 measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE, scaled by a copy-kernel calibration
    (profiles/traffic_shrinkwrap_2048.md, lines `traffic <kernel> <bytes>`)."""
    p = os.path.join(ROOT, "profiles", "traffic_shrinkwrap_2048.md")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
                return float(parts[2])
    return None


def cpu_baseline(M: np.ndarray, budget_s: float = 12.0, max_iter: int = 20) -> dict:
    import oracle  # CPU checker; only the baseline leg uses it

    oracle.set_fft_threads(1)   # the reference plans FFTW single-threaded
    g = oracle.fft2d(M.astype(np.complex64), +1, np.float32)
    mask = oracle.make_mask(g, 3.0, 0.04, np.float32)
    gp = g.copy()
    done, t0 = 0, time.perf_counter()
    while done < max_iter and (time.perf_counter() - t0) < budget_s:
        g, gp = oracle.hio_iterations(g, gp, mask, M, 0.9, 1, np.float32)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": round(done / dt, 4), "unit": "HIO iters/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"{done} HIO iterations (constraint, FFT, modulus, IFFT) of the 2048x2048 "
                      f"workload on the CPU oracle; FFT single-threaded as the reference, "
                      f"elementwise loops on {oracle.num_threads()} OpenMP threads; "
                      f"{dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): independent images shard "
              f"across GPUs with no communication, see DESIGN.md", file=sys.stderr)
        return 2

    log("[bench] generating the 2048x2048 workload")
    obj, M = syn.config_2048(0)
    d_src = im.DeviceBuffer.from_array(M)
    d_io = im.DeviceBuffer(M.nbytes)
    L = im.lib()
    p = im.default_params(n_cycles=N_CYCLES, target_error=FLT_MIN, n_hio_cycles=N_HIO)

    def step():
        im.check(L.imresh_memcpy_d2d(d_io.ptr, d_src.ptr, M.nbytes, None), "d2d")
        im.check(L.imresh_shrinkwrap_device(d_io.ptr, W, H, p, None, None), "shrinkwrap")

    rv, cv = im.ctypes.c_int(), im.ctypes.c_int()
    im.check(L.imresh_autotune(W, H, None, im.ctypes.byref(rv), im.ctypes.byref(cv)), "autotune")
    log(f"[bench] autotuned kernel variants: row {rv.value}, column {cv.value}")
    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    im.sync()
    log(f"[bench] timing {args.steps} reconstructions")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    im.sync()
    dt = time.perf_counter() - t0
    iters = args.steps * N_CYCLES * N_HIO
    value = iters / dt
    ms_per_step = 1000.0 * dt / args.steps

    # quality check of the timed output (not part of the timing)
    out = d_io.download(np.float32, M.shape)
    if not np.all(np.isfinite(out)):
        log("[bench] non-finite reconstruction")
        return 1

    log("[bench] per-kernel durations (HIP events)")
    ms = (im.ctypes.c_float * 4)()
    im.check(L.imresh_profile_kernels(W, H, 50, ms, None), "profile")
    per = dict(zip(KERNEL_NAMES, [float(x) for x in ms]))
    # time share per reconstruction: 380 row_mid, 400 col_proj, 20 blur, 20 row_last
    counts = {"row_mid": N_CYCLES * (N_HIO - 1), "col_proj": N_CYCLES * N_HIO, "blur": N_CYCLES,
              "row_last": N_CYCLES}
    dom = max(per, key=lambda k: per[k] * counts[k])
    n = W * H
    kernels = {}
    for k in KERNEL_NAMES:
        gbs = KERNEL_BYTES[k] * n / (per[k] * 1e-3) / 1e9
        kernels[k] = {"ms": round(per[k], 5), "GB/s": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
    ach = KERNEL_BYTES[dom] * n / (per[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(dom),
                "kernel": dom, "bytes_per_launch": KERNEL_BYTES[dom] * n, "kernels": kernels}
    # whole-iteration view: the reference's operator sequence moves 80 B/element per
    # HIO iteration (76 + 4-byte mask); the fused path moves 53 B/element
    roofline["iteration_alg_GBps_80B"] = round(80 * n * value / 1e9, 1)

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(M)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "HIO iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded rectangles/discs, Poisson 1e9 photons)",
        "config": {"workload": "shrink-wrap 2048x2048 fp32, 20 cycles x 20 HIO, beta 0.9, "
                               "sigma0 3.0, cut 0.2/0.04, 64 shapes in central 1024^2",
                   "iterations_per_step": N_CYCLES * N_HIO, "width": W, "height": H,
                   "row_variant": rv.value, "col_variant": cv.value},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
