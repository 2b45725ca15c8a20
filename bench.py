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
  roofline      the dominant fused kernel's measured duration (kernel-exact
                HIP events recorded around each launch on the launch stream
                by imresh_profile_kernels, row and column passes alternating
                as in a reconstruction) against its compulsory HBM bytes per
                launch and the 8 TB/s HBM peak;
                `traffic` from rocprofv3 FETCH_SIZE/WRITE_SIZE summaries in
                profiles/ when present
  cpu_baseline  the CPU oracle (C restatement of the reference CPU path,
                single-threaded FFT as the reference's FFTW plans, OpenMP
                elementwise loops) timed on a bounded sample of HIO
                iterations of the same workload
Further legs (batched, serving, 8192^2, 2-D FFT vs rocFFT, forward model)
are reported beside them; --no-<leg> skips one.
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
# cycle-graph modes the batched and serving legs time (1 = per-cycle HIP graph
# replay, 0 = eager launches); --cycle-graphs narrows it to one
GRAPH_MODES = (1, 0)


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


def cpu_baseline(M: np.ndarray, M256: np.ndarray, M8=None, budget_s: float = 12.0,
                 max_iter: int = 20) -> dict:
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
    # the reference's CPU-runnable case: one full 20 x 20 reconstruction of the
    # 256 x 256 workload (shrinkWrap.cpp:111-361 restated)
    t1 = time.perf_counter()
    oracle.shrink_wrap(M256, n_cycles=N_CYCLES, target_error=FLT_MIN, n_hio_cycles=N_HIO)
    dt256 = time.perf_counter() - t1
    out = {"value": round(done / dt, 4), "unit": "HIO iters/s", "cores": oracle.num_threads(),
           "kind": "port",
           "sample": f"{done} HIO iterations (constraint, FFT, modulus, IFFT) of the 2048x2048 "
                     f"workload on the CPU oracle; FFT single-threaded as the reference's FFTW "
                     f"plans but the project's own CPU FFT (FFTW3f is not installed), "
                     f"elementwise loops on {oracle.num_threads()} OpenMP threads; "
                     f"{dt:.1f} s",
           "recon_256_s": round(dt256, 3),
           "recon_256_note": "one full 20x20 reconstruction of the 256x256 workload on the CPU oracle",
           "per_op_ms": cpu_per_op(g, gp, mask, M),
           "per_op_ms_reference": cpu_ref_per_op(g, mask, M),
           "fft2d_numpy_ms": _time_ms(lambda: np.fft.fft2(g)),
           "per_op_note": "per_op_ms: the oracle's restatement; per_op_ms_reference: the reference's "
                          "own CPU sources compiled unchanged (no FFT: FFTW is absent); "
                          "fft2d_numpy_ms: numpy's pocketfft (float64, one thread) on the same "
                          "plane, for scale against the oracle's FFT"}
    if M8 is not None:
        out["large_8192"] = cpu_large(M8)
    return out


def parity_check(L, M: np.ndarray) -> dict:
    """The stated errors of the 2048^2 workload's fused kernels against the
    CPU oracle (the checker), measured in this run:
      * the support mask the fused path builds in cycles 0 and 1 (blur, tile
        maxima, threshold, cut-off fused into the first row pass) against the
        oracle's mask update of the same g': elements equal (bit-exact bar);
      * one fused HIO iteration from that state against the float64 oracle:
        relative L2 (bar 1e-5), with the float32 oracle's own distance;
      * full 20 x 20 reconstructions of converging 256^2 and 300^2 workloads
        (300 has no compiled plan: Bluestein transforms) against the float64
        oracle's end results (tests/golden/converging.npz): NCC to the object
        (bar 1e-3), the object's relative L2 (bar 1e-3) and the HIO error."""
    import oracle  # CPU checker; only the baseline legs use it

    oracle.set_fft_threads(min(16, os.cpu_count() or 8))
    out = {"tolerances": {"mask": "bit-exact", "one_iteration_rel_l2": 1e-5, "ncc": 1e-3,
                          "object_rel_l2": 1e-3}}
    p = im.default_params(n_cycles=2, target_error=FLT_MIN, n_hio_cycles=N_HIO)
    masks = {}
    for c in (0, 1):
        g, mask, blurred, thr = im.debug_cycle_mask(M, c, p)
        sigma = float(oracle.sigma_schedule(3.0, 0.01, 2)[c])
        exp = oracle.make_mask(g, sigma, 0.04 if c == 0 else 0.2, np.float32).astype(np.uint8)
        masks[c] = (g, mask)
        out[f"cycle{c}_mask_elements_equal"] = int(np.sum(mask == exp))
        out[f"cycle{c}_mask_elements"] = int(mask.size)
        out[f"cycle{c}_masked"] = int(mask.sum())
    g, mask = masks[0]
    dg, dgp = im.DeviceBuffer.from_array(g), im.DeviceBuffer.from_array(g)
    dm, dM = im.DeviceBuffer.from_array(mask), im.DeviceBuffer.from_array(M)
    im.check(L.imresh_hio_iterations(dg.ptr, dgp.ptr, dm.ptr, dM.ptr, W, H, np.float32(0.9), 1, None))
    g_gpu = dg.download(np.complex64, M.shape)
    g64, _ = oracle.hio_iterations(g.astype(np.complex128), g.astype(np.complex128), mask.astype(np.float64),
                                   M.astype(np.float64), 0.9, 1, np.float64)
    g32, _ = oracle.hio_iterations(g, g.copy(), mask.astype(np.float32), M, 0.9, 1, np.float32)
    out["one_iteration_rel_l2_vs_fp64"] = float(f"{syn.rel_l2(g_gpu, g64):.3e}")
    out["one_iteration_rel_l2_fp32_oracle_vs_fp64"] = float(f"{syn.rel_l2(g32, g64):.3e}")
    z = np.load(os.path.join(ROOT, "tests", "golden", "converging.npz"), allow_pickle=False)
    for name, key in (("c256a", "converging_256"), ("b300", "converging_300_bluestein")):
        n, shapes, central, idx = (int(v) for v in z[f"{name}_spec"])
        obj, Mc = syn.workload(n, int(z[f"{name}_height"]), n_shapes=shapes, central=central, index=idx)
        x = Mc.copy()
        r = im.shrink_wrap(x, n_cycles=N_CYCLES, target_error=FLT_MIN, n_hio_cycles=N_HIO)
        y0, x0, h, w = (int(v) for v in z[f"{name}_box"])
        st = int(z[f"{name}_step"])
        ref = z[f"{name}_crop64"].astype(np.float64)
        out[key] = {
            "ncc_gpu": round(syn.ncc(obj, x), 9), "ncc_fp64": round(float(z[f"{name}_ncc64"]), 9),
            "object_rel_l2_vs_fp64":
                float(f"{np.linalg.norm(x[y0:y0 + h:st, x0:x0 + w:st] - ref) / np.linalg.norm(ref):.3e}"),
            "hio_error_gpu": float(r.last_error), "hio_error_fp64": float(z[f"{name}_err64"]),
            "hio_error_fp32_oracle": float(z[f"{name}_err32"]),
            "hio_error_first_cycle_fp64": float(z[f"{name}_err64_c1"])}
    return out


def _time_ms(fn) -> float:
    t = time.perf_counter()
    fn()
    return round((time.perf_counter() - t) * 1e3, 2)


def cpu_per_op(g, gp, mask, M) -> dict:
    """Each operator of one HIO iteration and of one cycle's mask update,
    timed once on the CPU oracle at the workload's size (the reference's CPU
    operators, shrinkWrap.cpp:235-301), called directly on preallocated
    buffers so only the operator is timed: the FFTs single-threaded, the
    rest on the OpenMP threads."""
    import ctypes

    import oracle
    L = oracle.lib()
    h, w = mask.shape
    n = w * h
    gf = np.ascontiguousarray(g, dtype=np.complex64).view(np.float32)
    gpf = np.ascontiguousarray(gp, dtype=np.complex64).view(np.float32).copy()
    mk = np.ascontiguousarray(mask, dtype=np.float32)
    Mf = np.ascontiguousarray(M, dtype=np.float32)
    out_c = np.empty_like(gf)
    absg = np.empty(n, np.float32)
    L.orc_complex_norm_f32(absg.ctypes.data, gf.ctypes.data, n)
    thr = np.float32(0.2) * np.float32(L.orc_vmax_f32(absg.ctypes.data, n))
    blur_buf, cut_buf, mod_buf = absg.copy(), absg.copy(), gf.copy()
    tot, nm = ctypes.c_float(), ctypes.c_float()
    return {
        "fft2d_forward": _time_ms(lambda: L.orc_fft2d_f32(out_c.ctypes.data, gpf.ctypes.data, w, h, -1)),
        "fft2d_inverse": _time_ms(lambda: L.orc_fft2d_f32(out_c.ctypes.data, gf.ctypes.data, w, h, +1)),
        "hio_constraint": _time_ms(lambda: L.orc_hio_constraint_f32(gpf.ctypes.data, gf.ctypes.data,
                                                                    mk.ctypes.data, n, 0.9)),
        "apply_modulus": _time_ms(lambda: L.orc_apply_modulus_f32(mod_buf.ctypes.data, Mf.ctypes.data, n)),
        "complex_norm": _time_ms(lambda: L.orc_complex_norm_f32(absg.ctypes.data, gf.ctypes.data, n)),
        "gaussian_blur_sigma3": _time_ms(lambda: L.orc_blur_f32(blur_buf.ctypes.data, w, h, 3.0)),
        "vector_max": _time_ms(lambda: L.orc_vmax_f32(absg.ctypes.data, n)),
        "cutoff": _time_ms(lambda: L.orc_cutoff_f32(cut_buf.ctypes.data, n, thr, 1.0, 0.0)),
        "hio_error": _time_ms(lambda: L.orc_hio_error_f32(gf.ctypes.data, mk.ctypes.data, n, 0,
                                                          ctypes.byref(tot), ctypes.byref(nm))),
    }


def cpu_ref_per_op(g, mask, M):
    """The same operators from the reference's own CPU sources
    (gaussian.cpp, vectorReduce.cpp, vectorElementwise.cpp,
    hybridInputOutput.cpp compiled unchanged into oracle/_ref), timed once;
    None where that library was not built (it needs the reference sources)."""
    import ctypes

    import oracle
    r = oracle.ref()
    if r is None:
        return None
    g = np.ascontiguousarray(g, dtype=np.complex64)
    gf = g.view(np.float32)
    mk = np.ascontiguousarray(mask, dtype=np.float32)
    Mf = np.ascontiguousarray(M, dtype=np.float32)
    n = mk.size
    h, w = mk.shape
    absg = np.empty(n, np.float32)
    r.ref_complex_norm_f32(absg.ctypes.data, gf.ctypes.data, n)
    blur_buf = absg.copy()
    mod_buf = gf.copy()
    tot, nm = ctypes.c_float(), ctypes.c_float()
    return {
        "complex_norm": _time_ms(lambda: r.ref_complex_norm_f32(absg.ctypes.data, gf.ctypes.data, n)),
        "apply_modulus": _time_ms(lambda: r.ref_apply_modulus_f32(mod_buf.ctypes.data, Mf.ctypes.data, n)),
        "gaussian_blur_sigma3": _time_ms(lambda: r.ref_blur_f32(blur_buf.ctypes.data, w, h, 3.0)),
        "vector_max": _time_ms(lambda: r.ref_vmax_f32(absg.ctypes.data, n)),
        "hio_error": _time_ms(lambda: r.ref_hio_error_f32(gf.ctypes.data, mk.ctypes.data, n, 0,
                                                          ctypes.byref(tot), ctypes.byref(nm))),
    }


def cpu_large(M8: np.ndarray) -> dict:
    """The 8192 x 8192 workload on the CPU oracle: one HIO iteration and one
    cycle's mask update timed, a full 20 x 20 reconstruction extrapolated
    (the initial inverse FFT + 20 mask updates + 400 iterations); one
    iteration is two single-threaded 8192^2 FFTs, tens of seconds."""
    import oracle
    oracle.set_fft_threads(1)
    t = time.perf_counter()
    g = oracle.fft2d(M8.astype(np.complex64), +1, np.float32)
    t_ifft = time.perf_counter() - t
    t = time.perf_counter()
    mask = oracle.make_mask(g, 6.0, 0.04, np.float32)
    t_mask = time.perf_counter() - t
    gp = g.copy()
    t = time.perf_counter()
    oracle.hio_iterations(g, gp, mask, M8, 0.9, 1, np.float32)
    t_iter = time.perf_counter() - t
    total = t_ifft + N_CYCLES * t_mask + N_CYCLES * N_HIO * t_iter
    return {"hio_iter_s": round(t_iter, 3), "mask_update_s": round(t_mask, 3),
            "initial_ifft_s": round(t_ifft, 3), "recon_s_extrapolated": round(total, 1),
            "hio_iters_per_s": round(1.0 / t_iter, 5),
            "note": "1 HIO iteration and 1 mask update (sigma 6, 35 taps) timed on the CPU oracle, "
                    "20x20 reconstruction extrapolated linearly"}


def _gen_batch_input(args):
    i, n = args
    shapes, central = syn.batch_shapes(n)
    return syn.workload(n, n, n_shapes=shapes, central=central, photons=1e7,
                        index=20_000_000 + i)[1]


def generate_inputs(batch_count: int, serve_count: int, light_count: int = 1000, light_rate: float = 100.0):
    """All synthetic inputs, generated in worker processes BEFORE this
    process touches the GPU (workers are forked; a process that has
    initialized the GPU must not fork helpers)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    sizes = syn.batch_sizes(batch_count) if batch_count else []
    trace = syn.serve_trace(serve_count) if serve_count else []
    light = syn.serve_trace(light_count, rate=light_rate, index=1) if serve_count else []
    log(f"[bench] generating {len(sizes)} batch and {len(trace)} + {len(light)} serving inputs")
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 4),
                             mp_context=mp.get_context("fork")) as ex:
        bmods = list(ex.map(_gen_batch_input, list(enumerate(sizes)), chunksize=8))
        smods = list(ex.map(_gen_serve_input, [(i, n) for i, (_, n, _) in enumerate(trace)],
                            chunksize=4))
        lmods = list(ex.map(_gen_serve_input, [(100_000 + i, n) for i, (_, n, _) in enumerate(light)],
                            chunksize=4))
    return (sizes, bmods), (trace, smods), (light, lmods)


def batch_bench(L, sizes, mods, reps: int = 2, plain_streams: int = 3) -> dict:
    """Batched workload: `count` images of uneven sizes (synthetic.config_batch),
    20 x 20 each, inputs resident in HBM.  Times one batched call
    (imresh_shrinkwrap_batch_device) against the plain loop: one
    imresh_shrinkwrap_device per image, round-robin over `plain_streams`
    streams (the reference task queue's 3 streams per device), no host
    synchronization until the end.  Both outputs are compared bitwise."""
    count = len(sizes)
    srcs = [im.DeviceBuffer.from_array(m) for m in mods]
    outs_b = [im.DeviceBuffer(m.nbytes) for m in mods]
    outs_p = [im.DeviceBuffer(m.nbytes) for m in mods]
    p = im.default_params(n_cycles=N_CYCLES, target_error=FLT_MIN, n_hio_cycles=N_HIO)

    def restore(outs):
        for a, b, m in zip(srcs, outs, mods):
            im.check(L.imresh_memcpy_d2d(b.ptr, a.ptr, m.nbytes, None), "d2d")
        im.sync()

    images = [(o.ptr, n, n) for o, n in zip(outs_b, sizes)]
    streams = [im.stream_create() for _ in range(plain_streams)]

    def plain():
        for i, (o, n) in enumerate(zip(outs_p, sizes)):
            im.check(L.imresh_shrinkwrap_device(o.ptr, n, n, p, streams[i % plain_streams], None),
                     "shrinkwrap")
        for st in streams:
            im.sync(st)

    # both paths timed with per-cycle graph replay and with eager launches;
    # each reports its faster mode
    modes = {}
    for graphs in GRAPH_MODES:
        im.set_option(im.OPT_CYCLE_GRAPHS, graphs)
        log(f"[bench] batch vs plain loop, cycle graphs {'on' if graphs else 'off'} (warmup first)")
        restore(outs_b)
        im.shrink_wrap_batch_device(images, p, want_results=False)
        restore(outs_p)
        plain()
        tb, tp = [], []
        for _ in range(reps):
            restore(outs_b)
            t0 = time.perf_counter()
            im.shrink_wrap_batch_device(images, p, want_results=False)
            tb.append(time.perf_counter() - t0)
            restore(outs_p)
            t0 = time.perf_counter()
            plain()
            tp.append(time.perf_counter() - t0)
        modes[graphs] = (min(tb), min(tp))
    im.set_option(im.OPT_CYCLE_GRAPHS, GRAPH_MODES[0])
    gb = min(modes, key=lambda g: modes[g][0])
    gp = min(modes, key=lambda g: modes[g][1])
    tb, tp = [modes[gb][0]], [modes[gp][1]]
    same = all(np.array_equal(a.download(np.float32, m.shape), b.download(np.float32, m.shape))
               for a, b, m in zip(outs_b, outs_p, mods))
    for st in streams:
        im.stream_destroy(st)
    elems = sum(n * n for n in sizes)
    rb, rp = count / min(tb), count / min(tp)
    log(f"[bench] batched {rb:.1f} recon/s, plain loop {rp:.1f} recon/s, bitwise equal {same}")
    return {"images": count, "sizes": "uniform over " + ",".join(map(str, syn.BATCH_SIZES)),
            "elements": elems, "photons": 1e7, "batched_recon_per_s": round(rb, 2),
            "plain_loop_recon_per_s": round(rp, 2), "plain_loop_streams": plain_streams,
            "speedup": round(rb / rp, 3), "batched_s": round(min(tb), 4), "plain_s": round(min(tp), 4),
            "batched_cycle_graphs": bool(gb), "plain_cycle_graphs": bool(gp),
            "seconds_by_mode": {("graphs" if g else "eager"): {"batched": round(v[0], 4), "plain": round(v[1], 4)}
                                for g, v in modes.items()},
            "batched_alg_GBps_80B": round(80 * elems * N_CYCLES * N_HIO / min(tb) / 1e9, 1),
            "outputs_bitwise_equal": bool(same)}


def _gen_serve_input(args):
    i, n = args
    return syn.fast_modulus(n, i)


def _percentiles(lat):
    a = np.sort(np.asarray(lat))
    return {"p50_ms": round(float(np.percentile(a, 50)), 2), "p99_ms": round(float(np.percentile(a, 99)), 2),
            "p99.9_ms": round(float(np.percentile(a, 99.9)), 2), "max_ms": round(float(a[-1]), 2)}


def serve_plain_loops(trace, mods, server_ios) -> dict:
    """The plain per-image GPU loop over the serving requests, every request
    available at t=0 (the counterpart of the server's capacity run): moduli
    resident on the device, one imresh_shrinkwrap_device call per request,
    either one image at a time on one stream or round-robin on 3 streams
    (the batched leg's plain loop).  Outputs are compared with the server's
    bit for bit."""
    count = len(trace)
    params = [im.default_params(n_cycles=c, n_hio_cycles=N_HIO, target_error=FLT_MIN) for _, _, c in trace]
    bufs = [im.DeviceBuffer.from_array(m) for m in mods]
    out = {"inputs": "device-resident moduli, all requests available at t=0"}
    same = True
    for name, nst in (("one_stream", 1), ("three_streams", 3)):
        for b, m in zip(bufs, mods):
            b.upload(m)
        streams = [im.stream_create() for _ in range(nst)]
        im.sync()
        t0 = time.perf_counter()
        for k, (b, m) in enumerate(zip(bufs, mods)):
            im.shrink_wrap_device(b.ptr, m.shape[1], m.shape[0], params[k], stream=streams[k % nst],
                                  want_result=False)
        for st in streams:
            im.sync(st)
        dt = time.perf_counter() - t0
        for st in streams:
            im.stream_destroy(st)
        out[name] = {"recon_per_s": round(count / dt, 2), "streams": nst}
        same = same and all(np.array_equal(b.download(np.float32, m.shape), x)
                            for b, m, x in zip(bufs, mods, server_ios))
    for b in bufs:
        b.close()
    out["outputs_bitwise_equal_to_server"] = bool(same)
    log(f"[bench] serving plain loops: {out}")
    return out


def serve_bench(L, trace, mods, rate: float = 200.0, budget_gib: int = 200, modes=None) -> dict:
    """Serving workload: `count` requests of synthetic.serve_trace (Poisson
    arrivals at `rate`/s, sizes log-uniform in [128, 2048], 5/10/20 cycles x
    20 HIO), host buffers as the reference API takes them.  The same trace is
    replayed open-loop (each request submitted at its arrival time) through
    the continuous-batching server and through the plain task queue (3
    streams, one reconstruction per slot, addTask blocking when all slots are
    busy).  Latency = completion - scheduled arrival."""
    import threading

    count = len(trace)
    params = [im.default_params(n_cycles=c, n_hio_cycles=N_HIO, target_error=FLT_MIN) for _, _, c in trace]
    iters = sum(c * N_HIO for _, _, c in trace)

    def replay(submit, wait):
        ios = [m.copy() for m in mods]
        t0 = time.perf_counter()
        for k, ((a, n, c), io) in enumerate(zip(trace, ios)):
            dt = t0 + a - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            submit(k, io)
        wait()
        return t0, ios

    def run_server():
        with im.Server(hbm_budget_bytes=budget_gib << 30) as srv:
            ids = {}
            depth = []

            def sub(k, io):
                ids[srv.submit(io, params[k])] = k
                if k % 10 == 0:
                    st = srv.stats()
                    depth.append(int(st.queued + st.active))

            _, ios = replay(sub, srv.drain)
            comp = srv.completions(count + 16)
            peak = srv.stats().peak_bytes
            run_server.depth = depth
        lat = {ids[c[0]]: c[3] for c in comp}   # from submit (= arrival: submit never blocks)
        errs = sum(1 for c in comp if c[1] != 0)
        span = max(trace[k][0] + lat[k] / 1e3 for k in range(count))
        return [lat[k] for k in range(count)], span, ios, peak, errs

    def run_queue():
        lock = threading.Lock()
        t_done = [0.0] * count
        im.task_queue_init(0)

        def sub_q(k, io):
            def cb(arr, w, h, f, k=k):
                with lock:
                    t_done[k] = time.perf_counter()
            _, n, c = trace[k]
            im.add_task(io, cb, "", n_cycles=c, n_hio_cycles=N_HIO, target_error=FLT_MIN,
                        cutoff_autocorel=0.04)

        t0, ios = replay(sub_q, lambda: None)
        im.task_queue_deinit()
        lat = [(t_done[k] - (t0 + trace[k][0])) * 1e3 for k in range(count)]
        return lat, max(t_done) - t0, ios

    res_s, res_q = {}, {}
    for graphs in (modes or GRAPH_MODES):
        im.set_option(im.OPT_CYCLE_GRAPHS, graphs)
        log(f"[bench] serving replay at {rate:g}/s, cycle graphs {'on' if graphs else 'off'}")
        res_s[graphs] = run_server()
        res_q[graphs] = run_queue()
    im.set_option(im.OPT_CYCLE_GRAPHS, GRAPH_MODES[0])
    gs = min(res_s, key=lambda g: res_s[g][1])
    gq = min(res_q, key=lambda g: res_q[g][1])
    lat_s, span_s, ios_s, peak, bad = res_s[gs]
    lat_q, span_q, ios_q = res_q[gq]
    same = all(np.array_equal(a, b) for a, b in zip(ios_s, ios_q))
    half = count // 2
    depth = getattr(run_server, "depth", [])
    out = {"requests": count, "arrival_rate_per_s": rate if np.isfinite(rate) else "all submitted at t=0", "sizes": "round(exp(U[ln128, ln2048]))",
           "cycles": "5/10/20 x 20 HIO", "hbm_budget_gib": budget_gib, "peak_state_gib": round(peak / 2**30, 3),
           "server": {"recon_per_s": round(count / span_s, 2),
                      "hio_iters_per_s": round(iters / span_s, 1), **_percentiles(lat_s), "errors": bad,
                      "p50_ms_first_half": _percentiles(lat_s[:half])["p50_ms"],
                      "p50_ms_second_half": _percentiles(lat_s[half:])["p50_ms"],
                      "requests_in_system_at_arrivals": {"max": max(depth) if depth else None,
                                                         "first_quarter_mean": round(float(np.mean(depth[:max(1, len(depth) // 4)])), 1) if depth else None,
                                                         "last_quarter_mean": round(float(np.mean(depth[-max(1, len(depth) // 4):])), 1) if depth else None}},
           "task_queue": {"recon_per_s": round(count / span_q, 2),
                          "hio_iters_per_s": round(iters / span_q, 1), **_percentiles(lat_q),
                          "p50_ms_first_half": _percentiles(lat_q[:half])["p50_ms"],
                          "p50_ms_second_half": _percentiles(lat_q[half:])["p50_ms"]},
           "server_cycle_graphs": bool(gs), "task_queue_cycle_graphs": bool(gq),
           "span_s_by_mode": {("graphs" if g else "eager"): {"server": round(res_s[g][1], 3),
                                                              "task_queue": round(res_q[g][1], 3)}
                              for g in res_s},
           "outputs_bitwise_equal": bool(same)}
    out["speedup_recon_per_s"] = round(out["server"]["recon_per_s"] / out["task_queue"]["recon_per_s"], 3)
    log(f"[bench] serving: server {out['server']}, task queue {out['task_queue']}, equal {same}")
    serve_bench.last_server_ios = ios_s
    return out


def large_bench(L, M8, steps: int = 2) -> dict:
    """The large single image: 8192 x 8192 (512 shapes in the central 4096^2,
    1e11 photons, sigma0 = 6.0: 35-tap blur), 20 x 20, resident in HBM."""
    n = M8.shape[0]
    d_src = im.DeviceBuffer.from_array(M8)
    d_io = im.DeviceBuffer(M8.nbytes)
    p = im.default_params(n_cycles=N_CYCLES, target_error=FLT_MIN, n_hio_cycles=N_HIO, sigma0=6.0)

    def step():
        im.check(L.imresh_memcpy_d2d(d_io.ptr, d_src.ptr, M8.nbytes, None), "d2d")
        im.check(L.imresh_shrinkwrap_device(d_io.ptr, n, n, p, None, None), "shrinkwrap 8192")

    step()   # warm-up: tuning, weights, graph capture
    im.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    im.sync()
    dt = (time.perf_counter() - t0) / steps
    out = d_io.download(np.float32, M8.shape)
    ms = (im.ctypes.c_float * 4)()
    im.check(L.imresh_profile_kernels(n, n, 8, ms, None), "profile 8192")
    per = dict(zip(KERNEL_NAMES, [float(x) for x in ms]))
    kernels = {k: {"ms": round(per[k], 4),
                   "GB/s": round(KERNEL_BYTES[k] * n * n / (per[k] * 1e-3) / 1e9, 1)} for k in KERNEL_NAMES}
    return {"workload": "shrink-wrap 8192x8192 fp32, 20 cycles x 20 HIO, sigma0 6.0 (35 taps), "
                        "512 shapes in central 4096^2, 1e11 photons",
            "hio_iters_per_s": round(N_CYCLES * N_HIO / dt, 1), "ms_per_reconstruction": round(dt * 1e3, 2),
            "finite": bool(np.all(np.isfinite(out))), "kernels": kernels,
            "iteration_alg_GBps_80B": round(80 * n * n * N_CYCLES * N_HIO / dt / 1e9, 1)}


def medium_bench(L, M4, steps: int = 3) -> dict:
    """4096 x 4096 (256 shapes in the central 2048^2, 1e10 photons), 20 x 20:
    the iteration's working set (iterate, gPrev, modulus, mask: 21
    B/element, 352 MB) exceeds the 256 MB Infinity Cache, so this rate is
    the HBM-resident counterpart of the 2048^2 figure."""
    n = M4.shape[0]
    d_src = im.DeviceBuffer.from_array(M4)
    d_io = im.DeviceBuffer(M4.nbytes)
    p = im.default_params(n_cycles=N_CYCLES, target_error=FLT_MIN, n_hio_cycles=N_HIO)

    def step():
        im.check(L.imresh_memcpy_d2d(d_io.ptr, d_src.ptr, M4.nbytes, None), "d2d")
        im.check(L.imresh_shrinkwrap_device(d_io.ptr, n, n, p, None, None), "shrinkwrap 4096")

    step()
    im.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    im.sync()
    dt = (time.perf_counter() - t0) / steps
    rate = N_CYCLES * N_HIO / dt
    ms = (im.ctypes.c_float * 4)()
    im.check(L.imresh_profile_kernels(n, n, 8, ms, None), "profile 4096")
    per = dict(zip(KERNEL_NAMES, [float(x) for x in ms]))
    kernels = {k: {"ms": round(per[k], 4),
                   "GB/s": round(KERNEL_BYTES[k] * n * n / (per[k] * 1e-3) / 1e9, 1),
                   "frac": round(KERNEL_BYTES[k] * n * n / (per[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
               for k in KERNEL_NAMES}
    return {"workload": "shrink-wrap 4096x4096 fp32, 20 cycles x 20 HIO, 256 shapes in central 2048^2, "
                        "1e10 photons",
            "hio_iters_per_s": round(rate, 1), "ms_per_reconstruction": round(dt * 1e3, 2),
            "kernels": kernels,
            "iteration_alg_GBps_77B": round(77 * n * n * rate / 1e9, 1),
            "iteration_frac_77B": round(77 * n * n * rate / 1e9 / HBM_PEAK_GBS, 4),
            "iteration_fused_GBps_53B": round(53 * n * n * rate / 1e9, 1)}


def gpu_per_op(L, M: np.ndarray, reps: int = 20) -> dict:
    """The per-operator entry points (the reference's unfused operator
    sequence, one launch per operator) at the workload's size, ms per call
    from host timing of `reps` back-to-back calls on the default stream;
    the reductions return their scalar to the host and synchronize per call
    as the reference's do (cudaVectorReduce.cu:198-211,379-395).  The state
    is the workload's: g' = IFFT2(M), the mask from one cycle's update."""
    h, w = M.shape
    n = w * h
    c = im.ctypes
    dM = im.DeviceBuffer.from_array(M)
    g, gp, tmp = im.DeviceBuffer(8 * n), im.DeviceBuffer(8 * n), im.DeviceBuffer(8 * n)
    a, mask = im.DeviceBuffer(4 * n), im.DeviceBuffer(4 * n)
    mx, err = c.c_float(), c.c_float()
    im.check(L.imresh_copy_to_real_part(g.ptr, dM.ptr, n, None), "copy")
    im.check(L.imresh_fft2d(g.ptr, g.ptr, w, h, +1, None), "ifft")
    im.check(L.imresh_memcpy_d2d(gp.ptr, g.ptr, 8 * n, None), "d2d")
    im.check(L.imresh_complex_norm(a.ptr, g.ptr, n, None), "norm")
    im.check(L.imresh_gaussian_blur(a.ptr, w, h, 3.0, None), "blur")
    im.check(L.imresh_vector_max(a.ptr, n, c.byref(mx), None), "max")
    im.check(L.imresh_cutoff(a.ptr, n, np.float32(0.04) * np.float32(mx.value), 1.0, 0.0, None), "cut")
    im.check(L.imresh_memcpy_d2d(mask.ptr, a.ptr, 4 * n, None), "d2d")
    im.sync()
    ops = {
        "fft2d_forward": lambda: L.imresh_fft2d(tmp.ptr, gp.ptr, w, h, -1, None),
        "fft2d_inverse": lambda: L.imresh_fft2d(tmp.ptr, g.ptr, w, h, +1, None),
        "hio_constraint": lambda: L.imresh_hio_domain_constraints(gp.ptr, g.ptr, mask.ptr, n,
                                                                  np.float32(0.9), None),
        "apply_modulus": lambda: L.imresh_apply_complex_modulus(tmp.ptr, g.ptr, dM.ptr, n, None),
        "complex_norm": lambda: L.imresh_complex_norm(a.ptr, g.ptr, n, None),
        "gaussian_blur_sigma3": lambda: L.imresh_gaussian_blur(a.ptr, w, h, 3.0, None),
        "vector_max": lambda: L.imresh_vector_max(a.ptr, n, c.byref(mx), None),
        "cutoff": lambda: L.imresh_cutoff(a.ptr, n, np.float32(mx.value), 1.0, 0.0, None),
        "hio_error": lambda: L.imresh_calculate_hio_error(g.ptr, mask.ptr, n, 0, c.byref(err),
                                                          None, None, None),
    }
    out = {}
    for name, fn in ops.items():
        im.check(fn(), name)
        im.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            im.check(fn(), name)
        im.sync()
        out[name] = round((time.perf_counter() - t0) * 1e3 / reps, 4)
    out["unfused_iteration"] = round(sum(out[k] for k in ("hio_constraint", "fft2d_forward",
                                                          "apply_modulus", "fft2d_inverse")), 4)
    for b in (dM, g, gp, tmp, a, mask):
        b.close()
    return out


def small_bench(L, M256: np.ndarray, reps: int = 20) -> dict:
    """The reference's CPU-runnable case (256 x 256, 8 shapes, no noise,
    20 x 20) on the GPU: reconstructions/s of back-to-back calls with the
    modulus resident (restored by a device copy before each)."""
    n = M256.shape[0]
    d_src = im.DeviceBuffer.from_array(M256)
    d_io = im.DeviceBuffer(M256.nbytes)
    p = im.default_params(n_cycles=N_CYCLES, target_error=FLT_MIN, n_hio_cycles=N_HIO)

    def step():
        im.check(L.imresh_memcpy_d2d(d_io.ptr, d_src.ptr, M256.nbytes, None), "d2d")
        im.check(L.imresh_shrinkwrap_device(d_io.ptr, n, n, p, None, None), "shrinkwrap 256")

    step()
    im.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    im.sync()
    dt = (time.perf_counter() - t0) / reps
    return {"workload": "shrink-wrap 256x256 fp32, 20 cycles x 20 HIO, 8 shapes in central 128^2",
            "recon_per_s": round(1.0 / dt, 1), "hio_iters_per_s": round(N_CYCLES * N_HIO / dt, 1),
            "ms_per_reconstruction": round(dt * 1e3, 3)}


def forward_model_bench(L, obj: np.ndarray, reps: int = 50) -> dict:
    """The device forward model |FFT2(object)| (imresh_diffraction_intensity_device)
    at the object's size: kernel-exact durations of its two fused passes
    (real-input column pass 12 B/element, modulus-output row pass 12 B/element),
    wall-clock rate of back-to-back calls (each preceded by a 16 MiB device
    copy restoring the object, timed with it), parity against numpy in
    float64, and the Poisson measurement simulation on its output
    (imresh_poisson_modulus, wall clock per call)."""
    h, w = obj.shape
    x = np.ascontiguousarray(obj, dtype=np.float32)
    d_src = im.DeviceBuffer.from_array(x)
    d_io = im.DeviceBuffer(x.nbytes)

    def step():
        im.check(L.imresh_memcpy_d2d(d_io.ptr, d_src.ptr, x.nbytes, None), "d2d")
        im.check(L.imresh_diffraction_intensity_device(d_io.ptr, w, h, None), "forward model")

    step()
    im.sync()
    got = d_io.download(np.float32, (h, w))
    exp = np.abs(np.fft.fft2(x.astype(np.float64)))
    rel = float(np.linalg.norm(got - exp) / np.linalg.norm(exp))
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    im.sync()
    dt = (time.perf_counter() - t0) / reps
    # measurement simulation on the forward model's output (in place, repeated)
    im.check(L.imresh_poisson_modulus(d_io.ptr, w, h, 1e9, 1, None), "poisson")
    im.sync()
    t0 = time.perf_counter()
    for r in range(reps):
        im.check(L.imresh_poisson_modulus(d_io.ptr, w, h, 1e9, r, None), "poisson")
    im.sync()
    dt_pois = (time.perf_counter() - t0) / reps
    ms = (im.ctypes.c_float * 2)()
    im.check(L.imresh_profile_forward_model(w, h, 20, ms, None), "profile forward model")
    n = w * h
    col_ms, row_ms = float(ms[0]), float(ms[1])
    return {"workload": f"|FFT2(object)| {w}x{h} fp32, object resident in HBM",
            "images_per_s_with_restore_copy": round(1.0 / dt, 1),
            "kernels": {"col_real_fwd": {"ms": round(col_ms, 5), "GB/s": round(12 * n / (col_ms * 1e-3) / 1e9, 1)},
                        "row_fwd_norm": {"ms": round(row_ms, 5), "GB/s": round(12 * n / (row_ms * 1e-3) / 1e9, 1)}},
            "alg_GBps_24B": round(24 * n / ((col_ms + row_ms) * 1e-3) / 1e9, 1),
            "rel_l2_vs_numpy_fp64": rel,
            "poisson_modulus_ms": round(dt_pois * 1e3, 4),
            "poisson_modulus_note": "sqrt(Poisson(1e9 |F|^2 / sum |F|^2)) per element, wall clock per call"}


def main() -> int:
    global GRAPH_MODES
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-count", type=int, default=300)
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-medium", action="store_true")
    ap.add_argument("--no-fft-compare", action="store_true")
    ap.add_argument("--no-forward-model", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--no-per-op", action="store_true")
    ap.add_argument("--cycle-graphs", choices=("both", "on", "off"), default="both",
                    help="cycle-graph modes of the batched and serving legs (default: time both); "
                         "'off' also runs the other legs with eager launches")
    ap.add_argument("--cpu-large", action="store_true",
                    help="also time the 8192x8192 workload on the CPU oracle (about a minute)")
    ap.add_argument("--batch-count", type=int, default=512)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): independent images shard "
              f"across GPUs with no communication (one engine per GPU)", file=sys.stderr)
        return 2

    log("[bench] generating the 2048x2048 workload")
    obj, M = syn.config_2048(0)
    M256 = syn.config_256(0)[1]
    M8 = None
    if not args.no_large or args.cpu_large:
        log("[bench] generating the 8192x8192 workload")
        M8 = syn.config_8192(0)[1]
    (bsizes, bmods), (trace, smods), (light, lmods) = generate_inputs(
        0 if args.no_batch else args.batch_count, 0 if args.no_serve else args.serve_count)
    d_src = im.DeviceBuffer.from_array(M)
    d_io = im.DeviceBuffer(M.nbytes)
    L = im.lib()
    GRAPH_MODES = {"both": (1, 0), "on": (1,), "off": (0,)}[args.cycle_graphs]
    im.set_option(im.OPT_CYCLE_GRAPHS, GRAPH_MODES[0])
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

    # quality check of the timed output (not part of the timing): HIO error
    # of one more identical reconstruction and NCC of the object estimate to
    # the known synthetic object (modulo translation and twin image)
    out = d_io.download(np.float32, M.shape)
    if not np.all(np.isfinite(out)):
        log("[bench] non-finite reconstruction")
        return 1
    res = im.Result()
    im.check(L.imresh_memcpy_d2d(d_io.ptr, d_src.ptr, M.nbytes, None), "d2d")
    im.check(L.imresh_shrinkwrap_device(d_io.ptr, W, H, p, None, im.ctypes.byref(res)), "shrinkwrap")
    out2 = d_io.download(np.float32, M.shape)
    quality = {"hio_error": float(res.last_error), "cycles_run": int(res.cycles_run),
               "ncc_to_object": round(syn.ncc(out, obj), 6),
               "bitwise_repeatable": bool(np.array_equal(out, out2))}

    log("[bench] per-kernel durations (kernel-exact HIP events)")
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
    # whole-iteration view: the reference's operator sequence moves 76 + m
    # B/element per HIO iteration, m = bytes per mask element: 77 with this
    # build's byte mask, 80 with the reference's float mask; the fused path
    # moves 53 B/element
    roofline["iteration_alg_GBps_77B"] = round(77 * n * value / 1e9, 1)
    roofline["iteration_frac_77B"] = round(77 * n * value / 1e9 / HBM_PEAK_GBS, 4)
    roofline["iteration_alg_GBps_80B"] = round(80 * n * value / 1e9, 1)

    batch = None
    if not args.no_batch:
        batch = batch_bench(L, bsizes, bmods)

    serve = None
    if not args.no_serve:
        serve = serve_bench(L, trace, smods)
        log("[bench] serving capacity: the same requests all at once")
        serve["capacity"] = serve_bench(L, [(0.0, n, c) for _, n, c in trace], smods, rate=float("inf"),
                                        modes=(0,))
        cap = serve["capacity"]
        cap["plain_loop"] = serve_plain_loops(trace, smods, serve_bench.last_server_ios)
        for name in ("one_stream", "three_streams"):
            cap[f"speedup_vs_plain_loop_{name}"] = round(
                cap["server"]["recon_per_s"] / cap["plain_loop"][name]["recon_per_s"], 3)
        log("[bench] serving steady state (1000 requests at 100/s)")
        serve["steady_state"] = serve_bench(L, light, lmods, rate=100.0, modes=(0,))

    medium = None
    if not args.no_medium:
        log("[bench] 4096x4096 reconstruction (working set beyond the Infinity Cache)")
        medium = medium_bench(L, syn.workload(4096, 4096, n_shapes=256, central=2048, photons=1e10)[1])

    large = None
    if M8 is not None and not args.no_large:
        log("[bench] 8192x8192 reconstruction")
        large = large_bench(L, M8)
        if not args.cpu_large:
            M8 = None

    per_op = None
    if not args.no_per_op:
        log("[bench] per-operator entry points (unfused sequence)")
        per_op = gpu_per_op(L, M)
        per_op["fused_iteration"] = round(1e3 / value, 4)

    small = None
    if not args.no_small:
        log("[bench] 256x256 reconstructions")
        small = small_bench(L, M256)

    fwd = None
    if not args.no_forward_model:
        log("[bench] forward model |FFT2(object)|")
        fwd = forward_model_bench(L, obj)

    fftc = None
    if not args.no_fft_compare:
        from imresh_amd import fft_compare
        log("[bench] 2-D FFT vs rocFFT")
        fftc = [fft_compare.compare(s_) for s_ in (2048, 8192)]

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(M, M256, M8 if args.cpu_large else None)
        log("[bench] parity against the oracle")
        parity = parity_check(L, M)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "HIO iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded rectangles/discs, Poisson 1e9 photons)",
        "config": {"workload": "shrink-wrap 2048x2048 fp32, 20 cycles x 20 HIO, beta 0.9, "
                               "sigma0 3.0, cut 0.2/0.04, 64 shapes in central 1024^2",
                   "iterations_per_step": N_CYCLES * N_HIO, "width": W, "height": H,
                   "row_variant": rv.value, "col_variant": cv.value},
        "roofline": roofline, "cpu_baseline": cpu, "parity": parity, "quality": quality, "small_256": small,
        "batched": batch,
        "serving": serve, "medium_4096": medium, "large_8192": large, "gpu_per_op_ms": per_op, "fft_vs_rocfft": fftc,
        "forward_model": fwd,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
