"""GPU tests of the batched reconstruction (imresh_shrinkwrap_batch[_device]).

A batch runs the single-image kernels with the image index in the grid and
replays one HIP graph per shrink-wrap cycle, so every image of a batch must
come out bit-identical to its own imresh_shrinkwrap call (same kernel
variants: the autotuner's choice is cached per size).  The single-image
path is checked against the oracle in test_gpu_hio.py; here the batch is
checked against the single-image path, image by image, on mixed sizes
(including non-powers of two and sizes whose column tiles use more than
64 KiB of LDS), with early exits, and through both the host and the device
entry points.  The graph is replayed once per cycle, so every multi-cycle
case here also checks replays after the first against the eager path.
"""
import ctypes

import numpy as np
import pytest

import imresh_amd as im
from helpers import FLT_MIN
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu

KW = dict(n_cycles=3, n_hio_cycles=5, target_error=FLT_MIN)


def _moduli(sizes, seed=0):
    out = []
    for i, (w, h) in enumerate(sizes):
        _, M = syn.workload(w, h, n_shapes=6, central=min(w, h) // 2, index=1000 * seed + i)
        out.append(M)
    return out


def _single(M, **kw):
    x = M.copy()
    r = im.shrink_wrap(x, **kw)
    return x, r


def _params(**kw):
    return im.default_params(**kw)


@pytest.mark.parametrize("sizes", [
    [(64, 64)],
    [(64, 64), (96, 96), (64, 64), (128, 128), (96, 96), (200, 200)],
    [(256, 256), (96, 200), (384, 384), (256, 256)],
])
def test_batch_equals_single_images(sizes):
    Ms = _moduli(sizes, seed=len(sizes))
    ios = [M.copy() for M in Ms]
    res = im.shrink_wrap_batch(ios, _params(**KW))
    for M, io, r, (w, h) in zip(Ms, ios, res, sizes):
        x, rs = _single(M, **KW)
        print(f"{w}x{h}: batch error {r.last_error:.6g} single {rs.last_error:.6g}, "
              f"max |diff| {np.max(np.abs(io - x)):.3g}")
        assert r.cycles_run == rs.cycles_run == KW["n_cycles"]
        assert np.array_equal(io, x), (w, h)
        assert r.last_error == rs.last_error and r.last_masked == rs.last_masked


def test_batch_large_columns():
    """1024^2 column tiles exceed 64 KiB of LDS (attribute set before capture)."""
    sizes = [(1024, 1024), (768, 768), (1024, 1024)]
    Ms = _moduli(sizes, seed=7)
    kw = dict(n_cycles=2, n_hio_cycles=3, target_error=FLT_MIN)
    ios = [M.copy() for M in Ms]
    im.shrink_wrap_batch(ios, _params(**kw))
    for M, io in zip(Ms, ios):
        x, _ = _single(M, **kw)
        assert np.array_equal(io, x)


def test_batch_device_entry_point():
    sizes = [(128, 128), (96, 96), (128, 128)]
    Ms = _moduli(sizes, seed=3)
    bufs = [im.DeviceBuffer.from_array(M) for M in Ms]
    res = im.shrink_wrap_batch_device([(b.ptr, w, h) for b, (w, h) in zip(bufs, sizes)],
                                      _params(**KW))
    for M, b, r, (w, h) in zip(Ms, bufs, res, sizes):
        x, rs = _single(M, **KW)
        assert np.array_equal(b.download(np.float32, (h, w)), x)
        assert r.cycles_run == rs.cycles_run and r.last_error == rs.last_error


def test_batch_early_exit_per_image():
    """A target every image reaches after its first cycle: each image stops
    after one cycle (cudaShrinkWrap.cu:374-375) exactly as alone."""
    sizes = [(64, 64), (96, 96), (64, 64)]
    Ms = _moduli(sizes, seed=5)
    kw = dict(n_cycles=4, n_hio_cycles=3, target_error=1e30)
    ios = [M.copy() for M in Ms]
    res = im.shrink_wrap_batch(ios, _params(**kw))
    for M, io, r in zip(Ms, ios, res):
        x, rs = _single(M, **kw)
        assert r.cycles_run == rs.cycles_run == 1
        assert np.array_equal(io, x)


def test_batch_mixed_early_exit():
    """Images with different convergence: a target between their first-cycle
    errors stops some images and not others; each matches its single run."""
    sizes = [(64, 64)] * 4
    Ms = _moduli(sizes, seed=11)
    singles = [_single(M, n_cycles=1, n_hio_cycles=3, target_error=FLT_MIN)[1].last_error
               for M in Ms]
    target = float(np.median(singles))
    kw = dict(n_cycles=3, n_hio_cycles=3, target_error=target)
    ios = [M.copy() for M in Ms]
    res = im.shrink_wrap_batch(ios, _params(**kw))
    runs = []
    for M, io, r in zip(Ms, ios, res):
        x, rs = _single(M, **kw)
        runs.append(r.cycles_run)
        assert r.cycles_run == rs.cycles_run
        assert np.array_equal(io, x)
    print("cycles run per image:", runs, "first-cycle errors", singles, "target", target)


def test_batch_zero_cycles_and_empty():
    Ms = _moduli([(64, 64), (96, 96)], seed=2)
    ios = [M.copy() for M in Ms]
    res = im.shrink_wrap_batch(ios, _params(n_cycles=0, n_hio_cycles=5))
    for M, io, r in zip(Ms, ios, res):
        x, rs = _single(M, n_cycles=0, n_hio_cycles=5)
        assert r.cycles_run == 0
        assert np.array_equal(io, x)
    assert im.shrink_wrap_batch([], _params(**KW)) == []


def test_batch_rejects_bad_input():
    L = im.lib()
    imgs = (im.Image * 1)()
    imgs[0] = im.Image(None, 64, 64)
    assert L.imresh_shrinkwrap_batch(imgs, 1, None, None, None) == 1   # IMRESH_ERR_NULL
    a = np.zeros((2, 8193), np.float32)
    imgs[0] = im.Image(a.ctypes.data, 8193, 2)
    assert L.imresh_shrinkwrap_batch(imgs, 1, None, None, None) == 3   # unsupported length


def test_batch_repeated_calls_reuse_buffers():
    """Two calls in a row (pooled class buffers, the second smaller) agree
    with single runs."""
    for sizes in ([(96, 96)] * 5, [(96, 96)] * 2):
        Ms = _moduli(sizes, seed=len(sizes))
        ios = [M.copy() for M in Ms]
        im.shrink_wrap_batch(ios, _params(**KW))
        for M, io in zip(Ms, ios):
            assert np.array_equal(io, _single(M, **KW)[0])


def test_cycle_graphs_equal_eager_launches():
    """Each cycle replays a captured graph (default); with graphs off the
    same kernels are launched one by one.  Several reconstructions in a row
    (replays after the first, eager work in between) must agree bitwise."""
    Ms = _moduli([(128, 128), (97, 97), (128, 128)], seed=21)
    kw = dict(n_cycles=4, n_hio_cycles=3, target_error=FLT_MIN)
    try:
        im.set_option(im.OPT_CYCLE_GRAPHS, 0)
        eager = [_single(M, **kw) for M in Ms]
    finally:
        im.set_option(im.OPT_CYCLE_GRAPHS, 1)
    for _ in range(2):
        for M, (x, r) in zip(Ms, eager):
            y, s = _single(M, **kw)
            assert np.array_equal(x, y)
            assert r.last_error == s.last_error and r.cycles_run == s.cycles_run


def test_concurrent_batch_calls_from_two_threads():
    """Batch calls from several host threads on one device run one at a time
    on the library's capture stream; each call's images equal their single
    reconstructions."""
    import threading
    sets = [[(64, 64), (96, 96), (128, 128)], [(200, 200), (64, 64), (97, 131)],
            [(128, 128), (256, 256)]]
    moduli = [_moduli(sz, seed=40 + i) for i, sz in enumerate(sets)]
    ios = [[M.copy() for M in Ms] for Ms in moduli]
    errors = []

    def work(k):
        try:
            for _ in range(3):
                for io, M in zip(ios[k], moduli[k]):
                    io[...] = M
                im.shrink_wrap_batch(ios[k], _params(**KW))
        except Exception as e:   # reported below, in the test's thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(sets))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not errors, errors
    assert not any(t.is_alive() for t in threads)
    for Ms, outs in zip(moduli, ios):
        for M, io in zip(Ms, outs):
            assert np.array_equal(io, _single(M, **KW)[0]), M.shape
