"""Reconstructions of one size on many streams with parameters that change
from call to call (GPU).

A size keeps at most four pooled workspaces, so a fifth and sixth stream
take over busy ones and queue behind their previous users' work.  When the
HIO parameters change, the workspace's captured cycle graph is rebuilt; the
old graph may still have replays queued on the previous user's stream and
must outlive them.  Every output must equal the same reconstruction run
alone, bit for bit, in graph and in eager mode.
"""
import numpy as np
import pytest

import imresh_amd as im
from helpers import FLT_MIN
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _alone(M, beta, cut):
    x = M.copy()
    im.shrink_wrap(x, n_cycles=3, n_hio_cycles=6, target_error=FLT_MIN, hio_beta=beta, cutoff=cut)
    return x


@pytest.mark.parametrize("graphs", [1, 0])
def test_alternating_parameters_on_six_streams(graphs):
    obj, M = syn.workload(192, 192, n_shapes=8, central=96, index=31)
    combos = [(0.9, 0.2), (0.7, 0.2), (0.9, 0.25), (0.7, 0.25), (0.8, 0.2), (0.6, 0.3)]
    refs = {c: _alone(M, *c) for c in combos}
    n = M.size
    src = im.DeviceBuffer.from_array(M)
    bufs = [im.DeviceBuffer(4 * n) for _ in range(6)]
    streams = [im.stream_create() for _ in range(6)]
    im.set_option(im.OPT_CYCLE_GRAPHS, graphs)
    try:
        for rnd in range(4):
            order = combos[rnd % 6:] + combos[:rnd % 6]
            for i, (s, b) in enumerate(zip(streams, bufs)):
                beta, cut = order[i]
                im.check(im.lib().imresh_memcpy_d2d(b.ptr, src.ptr, 4 * n, s))
                p = im.default_params(n_cycles=3, n_hio_cycles=6, target_error=FLT_MIN, hio_beta=beta,
                                      cutoff=cut)
                im.shrink_wrap_device(b.ptr, 192, 192, p, stream=s, want_result=False)
            for s in streams:
                im.sync(s)
            for i, b in enumerate(bufs):
                out = b.download(np.float32, M.shape)
                assert np.array_equal(out, refs[order[i]]), (rnd, i, order[i])
    finally:
        im.set_option(im.OPT_CYCLE_GRAPHS, 1)
        for s in streams:
            im.stream_destroy(s)


def test_pooled_state_outlives_destroyed_streams():
    """Reconstructions on caller streams that are destroyed afterwards, with
    servers (which create and destroy streams of their own) in between:
    the pooled workspaces must not keep anything that refers to the
    destroyed streams (HIP reads the stream an event was last recorded on
    when the event is queried, so a pool that tracked its workspaces with
    events recorded on caller streams failed here with
    hipErrorCapturedEvent)."""
    trace = syn.serve_trace(40, index=3)
    mods = [syn.fast_modulus(n, i) for i, (_, n, _) in enumerate(trace)]
    p = im.default_params(n_cycles=2, n_hio_cycles=5, target_error=FLT_MIN)
    ref = []
    for m in mods:
        x = m.copy()
        im.check(im.lib().imresh_shrinkwrap(x.ctypes.data, x.shape[1], x.shape[0], p, None, None))
        ref.append(x)
    im.set_option(im.OPT_CYCLE_GRAPHS, 0)
    try:
        for k in (4, 8, 16):
            bufs = [im.DeviceBuffer.from_array(m) for m in mods]
            with im.Server() as srv:
                for b, m in zip(bufs, mods):
                    srv.submit((b.ptr, m.shape[1], m.shape[0]), p)
                srv.drain()
            for b, m in zip(bufs, mods):
                b.upload(m)
            streams = [im.stream_create() for _ in range(k)]
            for i, (b, m) in enumerate(zip(bufs, mods)):
                im.check(im.lib().imresh_shrinkwrap_device(b.ptr, m.shape[1], m.shape[0], p, streams[i % k], None),
                         f"{k} streams, request {i}")
            for s in streams:
                im.sync(s)
            for s in streams:
                im.stream_destroy(s)
            for b, m, r in zip(bufs, mods, ref):
                assert np.array_equal(b.download(np.float32, m.shape), r)
    finally:
        im.set_option(im.OPT_CYCLE_GRAPHS, 1)


def test_concurrent_reconstructions_from_host_threads():
    """Eight host threads reconstruct 32 images of eight sizes no other test
    uses, half through the host entry point (default stream) and half
    through the device entry point on a stream of their own, so plans,
    Bluestein tables, tuned variants and pooled workspaces of a size are
    first created while other threads ask for them.  Every result must
    equal the same reconstruction run alone afterwards, bit for bit."""
    import ctypes
    import threading

    sizes = [(77, 91), (131, 263), (389, 157), (509, 97), (1031, 61), (61, 1031), (200, 333), (96, 97)]
    mods = [syn.workload(w, h, n_shapes=4, central=min(w, h) // 2, index=300 + i)[1]
            for i, (w, h) in enumerate(sizes)]
    p = im.default_params(n_cycles=2, n_hio_cycles=5, target_error=FLT_MIN)
    jobs = [j % len(sizes) for j in range(32)]
    out, errors = [None] * len(jobs), []

    def work(tid):
        s = im.stream_create()
        try:
            for j in range(tid, len(jobs), 8):
                m = mods[jobs[j]]
                h, w = m.shape
                if j % 2 == 0:
                    x = m.copy()
                    st = im.lib().imresh_shrinkwrap(x.ctypes.data, w, h, ctypes.byref(p), None, None)
                    out[j] = x
                else:
                    d = im.DeviceBuffer.from_array(m)
                    st = im.lib().imresh_shrinkwrap_device(d.ptr, w, h, ctypes.byref(p), s, None)
                    if st == 0:
                        im.sync(s)
                        out[j] = d.download(np.float32, m.shape)
                if st != 0:
                    errors.append((j, sizes[jobs[j]], st))
        finally:
            im.stream_destroy(s)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for i, m in enumerate(mods):
        x = m.copy()
        h, w = m.shape
        im.check(im.lib().imresh_shrinkwrap(x.ctypes.data, w, h, ctypes.byref(p), None, None))
        for j, k in enumerate(jobs):
            if k == i:
                assert np.array_equal(out[j], x), (j, sizes[i])
