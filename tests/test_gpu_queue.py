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
