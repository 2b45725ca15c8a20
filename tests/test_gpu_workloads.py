"""The batched and serving workloads at their defined sizes (GPU).

* Batched: the 512-image workload (synthetic.config_batch: sizes uniform
  over 64 ... 1024, non-powers of two included, 1e7 photons per image),
  20 x 20 each, one imresh_shrinkwrap_batch_device call against the plain
  loop (imresh_shrinkwrap_device per image on 3 streams): every image
  bitwise equal, so each image's NCC to its object is the plain path's.
* Serving: a seeded 48-request sample of the serving trace
  (synthetic.serve_trace: sizes round(exp(U[ln 128, ln 2048])), 5 / 10 / 20
  cycles x 20), submitted back to back to the server so the requests
  overlap on the device: each result bitwise equal to the same request
  reconstructed alone.
"""
import numpy as np
import pytest

import imresh_amd as im
from helpers import FLT_MIN
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu
L = im.lib


def test_batch_512_workload_equals_plain_loop():
    sizes, objs, mods = syn.config_batch(512)
    p = im.default_params(n_cycles=20, target_error=FLT_MIN, n_hio_cycles=20)
    srcs = [im.DeviceBuffer.from_array(m) for m in mods]
    outs_b = [im.DeviceBuffer(m.nbytes) for m in mods]
    outs_p = [im.DeviceBuffer(m.nbytes) for m in mods]
    for a, b, c, m in zip(srcs, outs_b, outs_p, mods):
        im.check(L().imresh_memcpy_d2d(b.ptr, a.ptr, m.nbytes, None))
        im.check(L().imresh_memcpy_d2d(c.ptr, a.ptr, m.nbytes, None))
    im.sync()
    res = im.shrink_wrap_batch_device([(o.ptr, n, n) for o, n in zip(outs_b, sizes)], p)
    streams = [im.stream_create() for _ in range(3)]
    try:
        for i, (o, n) in enumerate(zip(outs_p, sizes)):
            im.check(L().imresh_shrinkwrap_device(o.ptr, n, n, p, streams[i % 3], None))
        for s in streams:
            im.sync(s)
    finally:
        for s in streams:
            im.stream_destroy(s)
    by_size = {}
    for i, (n, o, m) in enumerate(zip(sizes, objs, mods)):
        b = outs_b[i].download(np.float32, m.shape)
        q = outs_p[i].download(np.float32, m.shape)
        assert np.array_equal(b, q), (i, n)
        assert res[i].cycles_run == 20
        by_size.setdefault(n, []).append(syn.ncc(o, b))
    print("per-size NCC to the object (batched = plain loop, bitwise):")
    for n in sorted(by_size):
        v = np.array(by_size[n])
        print(f"  {n:5d}^2: {len(v):3d} images, NCC mean {v.mean():.4f} min {v.min():.4f} max {v.max():.4f}")
    assert len(sizes) == 512 and set(sizes) == set(syn.BATCH_SIZES)


def test_serving_trace_sample_equals_alone():
    trace = syn.serve_trace(48, index=7)
    mods = [syn.fast_modulus(n, 500 + i) for i, (_, n, _) in enumerate(trace)]
    params = [im.default_params(n_cycles=c, n_hio_cycles=20, target_error=FLT_MIN) for _, _, c in trace]
    alone = []
    for M, p in zip(mods, params):
        x = M.copy()
        im.check(L().imresh_shrinkwrap(x.ctypes.data, x.shape[1], x.shape[0], p, None, None))
        alone.append(x)
    ios = [m.copy() for m in mods]
    with im.Server() as srv:
        ids = {srv.submit(io, p): k for k, (io, p) in enumerate(zip(ios, params))}
        srv.drain()
        comp = srv.completions(64)
        peak_active = srv.stats()
    assert len(comp) == 48 and all(c[1] == 0 for c in comp)
    for rid, st, r, lat in comp:
        k = ids[rid]
        assert r.cycles_run == trace[k][2]
        assert np.array_equal(ios[k], alone[k]), (k, trace[k])
    sizes = [n for _, n, _ in trace]
    odd = sum(1 for n in sizes if n & (n - 1))
    print(f"48 requests, sizes {min(sizes)}..{max(sizes)} ({odd} not a power of two), "
          f"peak device bytes {peak_active.peak_bytes / 2**30:.2f} GiB")
