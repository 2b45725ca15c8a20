"""GPU tests of the task queue (imresh::io::taskQueueInit / addTask /
taskQueueDeinit, taskQueue.cu:91-251) and of the C++ drop-in flow.

* every queued reconstruction equals imresh_shrinkwrap of the same input
  with the queue's parameter rule (cudaShrinkWrap.cu:193-200), bit for bit,
  with mixed sizes and more tasks than stream slots; callbacks receive the
  buffer, its size and the file name;
* addTask outside init/deinit is an error;
* the C++ example built next to the library (examples/queue_png.cpp:
  readPNG -> diffractionIntensity -> addTask(writeOutAndFreePNG) ->
  taskQueueDeinit) writes PNGs equal to the same steps done through the C ABI.
"""
import os
import subprocess
import threading

import numpy as np
import pytest

import imresh_amd as im
from imresh_amd import synthetic as syn

pytestmark = pytest.mark.gpu
L = im.lib
EXAMPLE = os.path.join(os.path.dirname(im.LIB_PATH), "imresh_queue_example")


def _alone(M, n_cycles, n_hio):
    x = M.copy()
    p = im.default_params(im.DEFAULTS_QUEUE, n_cycles=n_cycles, n_hio_cycles=n_hio)
    im.check(L().imresh_shrinkwrap(x.ctypes.data, x.shape[1], x.shape[0], p, None, None))
    return x


def test_queue_results_equal_shrinkwrap():
    sizes = [(64, 64), (96, 80), (128, 128), (97, 131), (64, 64), (200, 120), (128, 128), (80, 96)]
    inputs = [syn.workload(w, h, n_shapes=5, central=min(w, h) // 2, index=i)[1]
              for i, (w, h) in enumerate(sizes)]
    got, lock = {}, threading.Lock()

    def write_out(io, w, h, name):
        with lock:
            got[name] = (io.copy(), w, h)

    bufs = [M.copy() for M in inputs]
    im.task_queue_init()
    try:
        for i, b in enumerate(bufs):
            im.add_task(b, write_out, f"task{i}", n_cycles=3, n_hio_cycles=4)
    finally:
        im.task_queue_deinit()
    assert sorted(got) == sorted(f"task{i}" for i in range(len(bufs)))
    for i, M in enumerate(inputs):
        io, w, h = got[f"task{i}"]
        assert (h, w) == M.shape
        assert np.array_equal(io, _alone(M, 3, 4)), sizes[i]


def test_add_task_requires_init():
    M = np.ones((16, 16), np.float32)
    assert L().imresh_add_task(M.ctypes.data, 16, 16, None, im.TASK_CALLBACK(), None) == 8


def test_cpp_dropin_example(tmp_path):
    assert os.path.exists(EXAMPLE), "build first: __graft_entry__.build()"
    obj, _ = syn.workload(128, 96, n_shapes=6, central=48, index=21)
    src = str(tmp_path / "object.png")
    im.write_png(src, obj.astype(np.float32))
    prefix = str(tmp_path / "out_")
    r = subprocess.run([EXAMPLE, src, prefix, "3", "4"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "done 3"
    # the same steps through the C ABI
    x = im.read_image(src)
    im.diffraction_intensity(x)
    x = _alone(x, 4, 20)
    ref = str(tmp_path / "ref.png")
    im.write_png(ref, x)
    exp = im.read_image(ref)
    for i in range(3):
        assert np.array_equal(im.read_image(f"{prefix}{i}.png"), exp), i


MINI = os.path.join(os.path.dirname(im.LIB_PATH), "imresh_shrinkwrap_example")


def test_cpp_shrinkwrap_mini_example(tmp_path):
    """examples/shrinkwrap_mini.cpp calls imresh::algorithms::shrinkWrap
    through the C++ header as the reference's miniExample.cpp:93-116 does:
    a null buffer returns 1, shrinkWrap(io, W, H, 64, 1e-6) equals
    imresh_shrinkwrap with the reference defaults bit for bit, and passing 0
    for every later parameter selects those defaults (shrinkWrap.cpp:133-140)."""
    assert os.path.exists(MINI), "build first: __graft_entry__.build()"
    obj, M = syn.workload(96, 80, n_shapes=6, central=40, index=23)
    src = str(tmp_path / "modulus.f32")
    M.astype(np.float32).tofile(src)
    prefix = str(tmp_path / "out_")
    r = subprocess.run([MINI, src, "96", "80", prefix, "6"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split()
    assert lines[:6] == ["null", "1", "a", "0", "b", "0"], r.stdout
    a = np.fromfile(prefix + "a.f32", np.float32).reshape(80, 96)
    b = np.fromfile(prefix + "b.f32", np.float32).reshape(80, 96)
    x = M.copy()
    p = im.default_params(n_cycles=6, target_error=1e-6)
    res = im.Result()
    im.check(L().imresh_shrinkwrap(x.ctypes.data, 96, 80, p, None, res))
    assert np.array_equal(a, x)
    assert np.array_equal(b, x)
    assert res.cycles_run >= 1


def test_deinit_from_callback_is_refused():
    """A task callback runs on a queue worker; a deinit there would join its
    own thread, so it returns IMRESH_ERR_INVALID and the queue keeps
    running."""
    got = []

    def write_out(io, w, h, name):
        got.append(L().imresh_task_queue_deinit())

    M = syn.workload(64, 64, n_shapes=4, central=32, index=24)[1]
    im.task_queue_init()
    try:
        im.add_task(M.copy(), write_out, "t", n_cycles=1, n_hio_cycles=2)
    finally:
        im.task_queue_deinit()
    assert got == [2]


def test_deinit_wakes_every_blocked_producer():
    """With one slot and more producers blocked in addTask than the FIFO
    holds, deinit must wake all of them (each returns OK or
    IMRESH_ERR_NOT_INITIALIZED) instead of leaving some asleep."""
    M = syn.workload(256, 256, n_shapes=8, central=128, index=25)[1]
    p = im.default_params(im.DEFAULTS_QUEUE, n_cycles=4, n_hio_cycles=20)
    im.check(L().imresh_task_queue_init(1))
    status, lock = [], threading.Lock()
    bufs = [M.copy() for _ in range(8)]

    def producer(b):
        st = L().imresh_add_task(b.ctypes.data, 256, 256, p, im.TASK_CALLBACK(), None)
        with lock:
            status.append(st)

    threads = [threading.Thread(target=producer, args=(b,), daemon=True) for b in bufs]
    for t in threads:
        t.start()
    import time
    time.sleep(0.05)
    im.check(L().imresh_task_queue_deinit())
    for t in threads:
        t.join(timeout=60)
    assert not any(t.is_alive() for t in threads), "a producer stayed blocked after deinit"
    assert len(status) == 8 and set(status) <= {0, 8}, status
