"""GPU tests of the continuous-batching server (imresh_server_*).

Requests of mixed sizes (compiled-plan and Bluestein lengths) and iteration
budgets are submitted while others run; every result must equal the same
reconstruction run alone through imresh_shrinkwrap, bit for bit (the server
runs the same kernels on each request's own resident state, only
interleaved), under tight and generous HBM budgets and with host and device
buffers.
"""
import time

import numpy as np
import pytest

import imresh_amd as im
from helpers import FLT_MIN
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _requests(specs, seed):
    reqs = []
    for i, (w, h, cycles) in enumerate(specs):
        _, M = syn.workload(w, h, n_shapes=6, central=min(w, h) // 2, index=seed * 100 + i)
        reqs.append((M, im.default_params(n_cycles=cycles, n_hio_cycles=4, target_error=FLT_MIN)))
    return reqs


def _alone(M, p):
    x = M.copy()
    r = im.Result()
    im.check(im.lib().imresh_shrinkwrap(x.ctypes.data, x.shape[1], x.shape[0], p, None, r))
    return x, r


SPECS = [(128, 128, 3), (97, 97, 2), (200, 200, 1), (131, 160, 3), (128, 128, 2), (256, 256, 2),
         (64, 96, 3), (97, 97, 1)]


def _run(server, reqs, gap_s=0.0):
    ios = [M.copy() for M, _ in reqs]
    ids = []
    for io, (_, p) in zip(ios, reqs):
        ids.append(server.submit(io, p))
        if gap_s:
            time.sleep(gap_s)
    server.drain()
    comp = {c[0]: c for c in server.completions()}
    return ios, ids, comp


@pytest.mark.parametrize("budget_mib,max_active", [(0, 0), (8, 4), (64, 2)])
def test_server_results_equal_alone(budget_mib, max_active):
    reqs = _requests(SPECS, seed=budget_mib + max_active)
    with im.Server(hbm_budget_bytes=budget_mib << 20, max_active=max_active, lanes=4) as srv:
        ios, ids, comp = _run(srv, reqs, gap_s=0.002)
        info = srv.stats()
        assert info.completed == len(reqs) and info.queued == 0 and info.active == 0
        if budget_mib:
            assert info.peak_bytes <= budget_mib << 20
    for (M, p), io, rid in zip(reqs, ios, ids):
        x, r = _alone(M, p)
        _, status, res, lat = comp[rid]
        assert status == 0 and lat > 0
        assert res.cycles_run == r.cycles_run == p.n_cycles
        assert res.last_error == r.last_error
        assert np.array_equal(io, x), M.shape


def test_server_device_buffers_and_early_exit():
    specs = [(128, 128, 4), (97, 97, 4), (128, 128, 4)]
    reqs = [(M, im.default_params(n_cycles=4, n_hio_cycles=3, target_error=1e30))
            for M, _ in _requests(specs, seed=9)]
    bufs = [im.DeviceBuffer.from_array(M) for M, _ in reqs]
    with im.Server(lanes=2) as srv:
        ids = [srv.submit((b.ptr, M.shape[1], M.shape[0]), p) for b, (M, p) in zip(bufs, reqs)]
        srv.drain()
        comp = {c[0]: c for c in srv.completions()}
    for b, (M, p), rid in zip(bufs, reqs, ids):
        x, r = _alone(M, p)
        assert comp[rid][2].cycles_run == r.cycles_run == 1
        assert np.array_equal(b.download(np.float32, M.shape), x)


def test_server_rejects_over_budget_and_bad_sizes():
    with im.Server(hbm_budget_bytes=1 << 20) as srv:
        a = np.zeros((512, 512), np.float32)
        with pytest.raises(im.ImreshError):
            srv.submit(a)
        b = np.zeros((2, 8193), np.float32)
        with pytest.raises(im.ImreshError):
            srv.submit(b)


def test_server_many_small_requests_complete():
    specs = [(64, 64, 1)] * 40 + [(96, 96, 1)] * 20
    reqs = _requests(specs, seed=3)
    with im.Server(max_active=16, lanes=8) as srv:
        ios, ids, comp = _run(srv, reqs)
    assert sorted(comp) == sorted(ids)
    for (M, p), io in list(zip(reqs, ios))[::7]:
        assert np.array_equal(io, _alone(M, p)[0])


def test_drain_and_destroy_from_a_callback_are_refused():
    """Completion callbacks run on the server's scheduler thread: a drain or
    destroy of the same server there would wait for that very thread, so both
    return IMRESH_ERR_INVALID (2) and the server keeps running; submit, poll
    and stats from the callback work."""
    import ctypes
    L = im.lib()
    got = []
    M = syn.workload(64, 64, n_shapes=4, central=32, index=77)[1]
    p = im.default_params(n_cycles=1, n_hio_cycles=2, target_error=FLT_MIN)
    with im.Server(max_active=4, lanes=2) as srv:
        h = srv._h
        info = im.ServerInfo()

        @im.SERVER_CALLBACK
        def cb(user, comp, io):
            got.append((L.imresh_server_drain(h), L.imresh_server_destroy(h),
                        L.imresh_server_stats(h, ctypes.byref(info))))

        x = M.copy()
        rid = ctypes.c_uint64()
        im.check(L.imresh_server_submit(h, x.ctypes.data, 64, 64, ctypes.byref(p), 0, cb, None,
                                        ctypes.byref(rid)))
        srv.drain()
        assert got == [(2, 2, 0)]
        assert np.array_equal(x, _alone(M, p)[0])
        y = M.copy()                           # still serving after the refused calls
        srv.submit(y, p)
        srv.drain()
        assert np.array_equal(y, x)
