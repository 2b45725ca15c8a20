"""imresh_amd -- MI355X (gfx950) shrink-wrap / HIO phase retrieval.

Python view of the C ABI in ``include/imresh_hip.h`` (the product is the
shared library ``imresh_amd/lib/libimresh_hip.so``; the C++ drop-in API
``imresh::algorithms::shrinkWrap`` and ``imresh::io::taskQueue*`` lives in the
same library).  This module only binds it with ctypes: there is no Python or
CPU fallback -- every call fails loudly when the library or a GPU is missing.

Python mirror of the reference API:
  shrink_wrap(io, ...)        imresh::algorithms::shrinkWrap (shrinkWrap.hpp:40-53)
  task_queue_init/add_task/task_queue_deinit   imresh::io (taskQueue.hpp:38-84)
"""
from __future__ import annotations

import ctypes
import os
import threading
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libimresh_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "imresh_hip.h")

IMRESH_OK = 0
DEFAULTS_CPU = 0
DEFAULTS_QUEUE = 1


class ImreshError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = _lib().imresh_strerror(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


class Params(ctypes.Structure):
    """imresh_params (shrinkWrap.hpp:40-53 parameter set)."""
    _fields_ = [("n_cycles", c_uint32), ("target_error", c_float), ("hio_beta", c_float),
                ("cutoff_autocorel", c_float), ("cutoff", c_float), ("sigma0", c_float),
                ("sigma_change", c_float), ("n_hio_cycles", c_uint32), ("defaults", c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("cycles_run", c_uint32), ("last_error", c_float), ("last_total", c_float),
                ("last_masked", c_float)]


class Image(ctypes.Structure):
    """imresh_image: one image of a batch (data: host or device pointer)."""
    _fields_ = [("data", c_void_p), ("width", c_uint32), ("height", c_uint32)]


class ServerCompletion(ctypes.Structure):
    """imresh_server_completion."""
    _fields_ = [("id", c_uint64), ("status", ctypes.c_int32), ("result", Result),
                ("latency_ms", c_double)]


class ServerInfo(ctypes.Structure):
    _fields_ = [("submitted", c_uint64), ("completed", c_uint64), ("queued", c_uint64),
                ("active", c_uint64), ("bytes_in_use", c_uint64), ("peak_bytes", c_uint64),
                ("budget", c_uint64)]


SERVER_CALLBACK = ctypes.CFUNCTYPE(None, c_void_p, POINTER(ServerCompletion), POINTER(c_float))

TASK_CALLBACK = ctypes.CFUNCTYPE(None, c_void_p, POINTER(c_float), c_uint32, c_uint32, c_int)

_LIB = None
_LIB_LOCK = threading.Lock()

_P = c_void_p
_SIGS = {
    "imresh_abi_version": (c_int, []),
    "imresh_strerror": (ctypes.c_char_p, [c_int]),
    "imresh_params_default": (c_int, [POINTER(Params)]),
    "imresh_device_count": (c_int, [POINTER(c_int)]),
    "imresh_dev_alloc": (c_int, [POINTER(c_void_p), c_uint64]),
    "imresh_dev_free": (c_int, [_P]),
    "imresh_memcpy_h2d": (c_int, [_P, _P, c_uint64, _P]),
    "imresh_memcpy_d2h": (c_int, [_P, _P, c_uint64, _P]),
    "imresh_memcpy_d2d": (c_int, [_P, _P, c_uint64, _P]),
    "imresh_stream_sync": (c_int, [_P]),
    "imresh_stream_create": (c_int, [POINTER(c_void_p)]),
    "imresh_set_option": (c_int, [c_int, c_int]),
    "imresh_stream_destroy": (c_int, [_P]),
    "imresh_shrinkwrap": (c_int, [_P, c_uint32, c_uint32, POINTER(Params), _P, POINTER(Result)]),
    "imresh_shrinkwrap_device": (c_int, [_P, c_uint32, c_uint32, POINTER(Params), _P,
                                         POINTER(Result)]),
    "imresh_shrinkwrap_batch": (c_int, [POINTER(Image), c_uint32, POINTER(Params), _P,
                                        POINTER(Result)]),
    "imresh_shrinkwrap_batch_device": (c_int, [POINTER(Image), c_uint32, POINTER(Params), _P,
                                               POINTER(Result)]),
    "imresh_server_create": (c_int, [c_uint64, c_uint32, c_uint32, POINTER(c_void_p)]),
    "imresh_server_submit": (c_int, [_P, _P, c_uint32, c_uint32, POINTER(Params), c_int,
                                     SERVER_CALLBACK, _P, POINTER(c_uint64)]),
    "imresh_server_poll": (c_int, [_P, POINTER(ServerCompletion), c_uint32, POINTER(c_uint32)]),
    "imresh_server_drain": (c_int, [_P]),
    "imresh_server_stats": (c_int, [_P, POINTER(ServerInfo)]),
    "imresh_server_destroy": (c_int, [_P]),
    "imresh_hybrid_input_output": (c_int, [_P, _P, c_uint32, c_uint32, c_uint32, c_float, c_float,
                                           _P, POINTER(Result)]),
    "imresh_hybrid_input_output_device": (c_int, [_P, _P, c_uint32, c_uint32, c_uint32, c_float,
                                                  c_float, _P, POINTER(Result)]),
    "imresh_hio_iterations": (c_int, [_P, _P, _P, _P, c_uint32, c_uint32, c_float, c_uint32, _P]),
    "imresh_copy_to_real_part": (c_int, [_P, _P, c_uint32, _P]),
    "imresh_copy_from_real_part": (c_int, [_P, _P, c_uint32, _P]),
    "imresh_complex_norm": (c_int, [_P, _P, c_uint32, _P]),
    "imresh_apply_complex_modulus": (c_int, [_P, _P, _P, c_uint32, _P]),
    "imresh_hio_domain_constraints": (c_int, [_P, _P, _P, c_uint32, c_float, _P]),
    "imresh_cutoff": (c_int, [_P, c_uint32, c_float, c_float, c_float, _P]),
    "imresh_vector_max": (c_int, [_P, c_uint32, POINTER(c_float), _P]),
    "imresh_vector_min": (c_int, [_P, c_uint32, POINTER(c_float), _P]),
    "imresh_vector_sum": (c_int, [_P, c_uint32, POINTER(c_float), _P]),
    "imresh_calculate_hio_error": (c_int, [_P, _P, c_uint32, c_int, POINTER(c_float),
                                           POINTER(c_float), POINTER(c_float), _P]),
    "imresh_calculate_hio_error_u8": (c_int, [_P, _P, c_uint32, c_int, POINTER(c_float),
                                              POINTER(c_float), POINTER(c_float), _P]),
    "imresh_calculate_hio_error_bits": (c_int, [_P, _P, c_uint32, c_int, POINTER(c_float),
                                                POINTER(c_float), POINTER(c_float), _P]),
    "imresh_pack_mask": (c_int, [_P, c_uint32, _P, _P]),
    "imresh_gaussian_kernel": (c_int, [c_double, _P, c_uint32]),
    "imresh_gaussian_blur": (c_int, [_P, c_uint32, c_uint32, c_double, _P]),
    "imresh_gaussian_blur_horizontal": (c_int, [_P, c_uint32, c_uint32, c_double, _P]),
    "imresh_gaussian_blur_vertical": (c_int, [_P, c_uint32, c_uint32, c_double, _P]),
    "imresh_fft2d": (c_int, [_P, _P, c_uint32, c_uint32, c_int, _P]),
    "imresh_fft_supported": (c_int, [c_uint32]),
    "imresh_diffraction_intensity": (c_int, [_P, c_uint32, c_uint32, _P, c_int]),
    "imresh_diffraction_intensity_device": (c_int, [_P, c_uint32, c_uint32, _P]),
    "imresh_poisson_modulus": (c_int, [_P, c_uint32, c_uint32, c_double, c_uint64, _P]),
    "imresh_fft_shift": (c_int, [_P, c_uint32, c_uint32, c_uint32, _P]),
    "imresh_autotune": (c_int, [c_uint32, c_uint32, _P, POINTER(c_int), POINTER(c_int)]),
    "imresh_release_cached": (c_int, []),
    "imresh_cached_bytes": (c_int, [POINTER(c_uint64)]),
    "imresh_variant_counts": (c_int, [c_uint32, c_uint32, POINTER(c_int), POINTER(c_int)]),
    "imresh_set_variants": (c_int, [c_uint32, c_uint32, c_int, c_int]),
    "imresh_task_queue_init": (c_int, [c_int]),
    "imresh_task_queue_deinit": (c_int, []),
    "imresh_add_task": (c_int, [_P, c_uint32, c_uint32, POINTER(Params), TASK_CALLBACK, _P]),
    "imresh_profile_kernels": (c_int, [c_uint32, c_uint32, c_uint32, POINTER(c_float), _P]),
    "imresh_read_png": (c_int, [ctypes.c_char_p, _P, ctypes.c_size_t, POINTER(c_uint32), POINTER(c_uint32)]),
    "imresh_read_tiff": (c_int, [ctypes.c_char_p, _P, ctypes.c_size_t, POINTER(c_uint32), POINTER(c_uint32)]),
    "imresh_read_txt": (c_int, [ctypes.c_char_p, _P, ctypes.c_size_t, POINTER(c_uint32), POINTER(c_uint32)]),
    "imresh_write_png": (c_int, [ctypes.c_char_p, _P, c_uint32, c_uint32]),
    "imresh_profile_forward_model": (c_int, [c_uint32, c_uint32, c_uint32, POINTER(c_float), _P]),
    "imresh_shrinkwrap_device_noalloc": (c_int, [_P, c_uint32, c_uint32, POINTER(Params), _P,
                                                 POINTER(Result)]),
    "imresh_debug_cycle_mask": (c_int, [_P, c_uint32, c_uint32, POINTER(Params), c_uint32, _P, _P, _P,
                                        POINTER(c_float), _P]),
}


def _lib():
    """Load the built library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        with _LIB_LOCK:
            if _LIB is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: build it first (python -c "
                        "'import __graft_entry__ as g; g.build()')")
                lib = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in _SIGS.items():
                    f = getattr(lib, name)
                    f.restype = res
                    f.argtypes = args
                _LIB = lib
    return _LIB


lib = _lib


def check(status: int, what: str = "") -> None:
    if status != IMRESH_OK:
        raise ImreshError(status, what)


def device_count() -> int:
    n = c_int(0)
    check(_lib().imresh_device_count(ctypes.byref(n)), "imresh_device_count")
    return n.value


def default_params(defaults: int = DEFAULTS_CPU, **kw) -> Params:
    p = Params()
    check(_lib().imresh_params_default(ctypes.byref(p)))
    p.defaults = defaults
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gaussian_kernel(sigma: float) -> np.ndarray:
    n = _lib().imresh_gaussian_kernel(float(sigma), None, 0)
    w = np.zeros(n, dtype=np.float32)
    _lib().imresh_gaussian_kernel(float(sigma), w.ctypes.data, n)
    return w


# ------------------------------------------------------------ device memory
class DeviceBuffer:
    """A device allocation owned by this object (freed on close / GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = c_void_p()
        check(_lib().imresh_dev_alloc(ctypes.byref(p), self.nbytes), "imresh_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(_lib().imresh_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, None), "h2d")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(_lib().imresh_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, None), "d2h")
        return out

    def close(self) -> None:
        if getattr(self, "ptr", None):
            _lib().imresh_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sync(stream=None) -> None:
    check(_lib().imresh_stream_sync(stream), "imresh_stream_sync")


OPT_CYCLE_GRAPHS = 1


def set_option(option: int, value: int) -> None:
    check(_lib().imresh_set_option(option, value), "imresh_set_option")


def stream_create() -> int:
    s = c_void_p()
    check(_lib().imresh_stream_create(ctypes.byref(s)), "imresh_stream_create")
    return s.value


def stream_destroy(stream: int) -> None:
    check(_lib().imresh_stream_destroy(stream), "imresh_stream_destroy")


# ------------------------------------------------------------ reference API
def shrink_wrap(io: np.ndarray, n_cycles: int = 20, target_error: float = 1e-5,
                hio_beta: float = 0.9, cutoff_autocorel: float = 0.04, cutoff: float = 0.20,
                sigma0: float = 3.0, sigma_change: float = 0.01, n_hio_cycles: int = 20,
                stream=None, defaults: int = DEFAULTS_CPU) -> Result:
    """imresh::algorithms::shrinkWrap on the GPU.  `io` (float32, H x W,
    C-contiguous) holds |F| and is overwritten with Re(g') in place."""
    if io.dtype != np.float32 or not io.flags.c_contiguous or io.ndim != 2:
        raise ValueError("io must be a C-contiguous float32 H x W array")
    h, w = io.shape
    p = Params(n_cycles, target_error, hio_beta, cutoff_autocorel, cutoff, sigma0, sigma_change,
               n_hio_cycles, defaults)
    r = Result()
    check(_lib().imresh_shrinkwrap(io.ctypes.data, w, h, ctypes.byref(p), stream, ctypes.byref(r)),
          "imresh_shrinkwrap")
    return r


def shrink_wrap_device(io_dev: int, width: int, height: int, params: Params | None = None,
                       stream=None, want_result: bool = True):
    p = params if params is not None else default_params()
    r = Result()
    check(_lib().imresh_shrinkwrap_device(io_dev, width, height, ctypes.byref(p), stream,
                                          ctypes.byref(r) if want_result else None),
          "imresh_shrinkwrap_device")
    return r if want_result else None


def shrink_wrap_batch(ios: list, params: Params | None = None, stream=None) -> list:
    """Reconstructs a list of float32 H x W arrays (each |F| in, Re(g') out,
    in place) together; returns one Result per image."""
    for io in ios:
        if io.dtype != np.float32 or not io.flags.c_contiguous or io.ndim != 2:
            raise ValueError("every image must be a C-contiguous float32 H x W array")
    p = params if params is not None else default_params()
    imgs = (Image * max(1, len(ios)))()
    for i, io in enumerate(ios):
        imgs[i] = Image(io.ctypes.data, io.shape[1], io.shape[0])
    res = (Result * max(1, len(ios)))()
    check(_lib().imresh_shrinkwrap_batch(imgs, len(ios), ctypes.byref(p), stream, res),
          "imresh_shrinkwrap_batch")
    return [res[i] for i in range(len(ios))]


def shrink_wrap_batch_device(images: list, params: Params | None = None, stream=None,
                             want_results: bool = True):
    """Batch over device buffers: images = [(device_ptr, width, height), ...]."""
    p = params if params is not None else default_params()
    imgs = (Image * max(1, len(images)))()
    for i, (ptr, w, h) in enumerate(images):
        imgs[i] = Image(ptr, w, h)
    res = (Result * max(1, len(images)))()
    check(_lib().imresh_shrinkwrap_batch_device(imgs, len(images), ctypes.byref(p), stream,
                                                res if want_results else None),
          "imresh_shrinkwrap_batch_device")
    return [res[i] for i in range(len(images))] if want_results else None


def debug_cycle_mask(modulus: np.ndarray, cycle: int, params: Params | None = None, stream=None):
    """The state of shrink-wrap cycle `cycle` of a reconstruction of
    `modulus` as the fused kernels hold it: (g' the mask is built from,
    complex64 H x W; mask, uint8 H x W, 1 = outside; blurred |g'|, float32;
    threshold cut_c * max)."""
    m = np.ascontiguousarray(modulus, dtype=np.float32)
    h, w = m.shape
    p = params if params is not None else default_params()
    g = np.empty((h, w), np.complex64)
    mask = np.empty((h, w), np.uint8)
    blurred = np.empty((h, w), np.float32)
    thr = c_float()
    check(_lib().imresh_debug_cycle_mask(m.ctypes.data, w, h, ctypes.byref(p), cycle, g.ctypes.data,
                                         mask.ctypes.data, blurred.ctypes.data, ctypes.byref(thr),
                                         stream), "imresh_debug_cycle_mask")
    return g, mask, blurred, thr.value


def hybrid_input_output(io: np.ndarray, mask: np.ndarray, n_cycles: int = 0,
                        target_error: float = 1e-6, hio_beta: float = 0.9, stream=None) -> Result:
    """imresh::libs::hybridInputOutput on the GPU: `io` (float32 H x W) holds
    |F| and is overwritten with Re(g'); `mask` uint8 H x W, 1 = outside the
    support; n_cycles 0 = until target_error is met."""
    if io.dtype != np.float32 or not io.flags.c_contiguous or io.ndim != 2:
        raise ValueError("io must be a C-contiguous float32 H x W array")
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    if m.shape != io.shape:
        raise ValueError("mask must have the shape of io")
    h, w = io.shape
    r = Result()
    check(_lib().imresh_hybrid_input_output(io.ctypes.data, m.ctypes.data, w, h, n_cycles,
                                            target_error, hio_beta, stream, ctypes.byref(r)),
          "imresh_hybrid_input_output")
    return r


def read_image(path: str) -> np.ndarray:
    """readPNG / readTiff / readTxt by file extension (.png, .tif/.tiff,
    anything else as text): a float32 H x W array."""
    ext = os.path.splitext(path)[1].lower()
    fn = {".png": "imresh_read_png", ".tif": "imresh_read_tiff", ".tiff": "imresh_read_tiff"}.get(
        ext, "imresh_read_txt")
    L = _lib()
    w, h = c_uint32(), c_uint32()
    p = os.fsencode(path)
    check(getattr(L, fn)(p, None, 0, ctypes.byref(w), ctypes.byref(h)), fn)
    out = np.empty((h.value, w.value), np.float32)
    check(getattr(L, fn)(p, out.ctypes.data, out.size, ctypes.byref(w), ctypes.byref(h)), fn)
    return out


def write_png(path: str, data: np.ndarray) -> None:
    """writeOutPNG: 16-bit RGB scaled by the maximum."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim != 2:
        raise ValueError("H x W array expected")
    check(_lib().imresh_write_png(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0]),
          "imresh_write_png")


def diffraction_intensity(io: np.ndarray, stream=None) -> np.ndarray:
    """imresh::libs::diffractionIntensity on the GPU: `io` (float32 H x W,
    the real object) is overwritten with |FFT2(io)| (unnormalized, frequency 0
    at [0, 0]) and returned."""
    if io.dtype != np.float32 or not io.flags.c_contiguous or io.ndim != 2:
        raise ValueError("io must be a C-contiguous float32 H x W array")
    h, w = io.shape
    check(_lib().imresh_diffraction_intensity(io.ctypes.data, w, h, stream, 0),
          "imresh_diffraction_intensity")
    return io


class Server:
    """Continuous-batching reconstruction server (imresh_server_*).  submit()
    takes a float32 H x W host array (kept alive until completion) or a
    (device_ptr, width, height) tuple; completions() returns the finished
    requests as (id, status, Result, latency_ms)."""

    def __init__(self, hbm_budget_bytes: int = 0, max_active: int = 0, lanes: int = 0):
        h = c_void_p()
        check(_lib().imresh_server_create(hbm_budget_bytes, max_active, lanes, ctypes.byref(h)),
              "imresh_server_create")
        self._h = h.value
        self._keep: dict = {}

    def submit(self, io, params: Params | None = None) -> int:
        p = params if params is not None else default_params()
        rid = c_uint64()
        if isinstance(io, tuple):
            ptr, w, h = io
            dev = 1
        else:
            if io.dtype != np.float32 or not io.flags.c_contiguous or io.ndim != 2:
                raise ValueError("io must be a C-contiguous float32 H x W array")
            ptr, (h, w), dev = io.ctypes.data, io.shape, 0
        check(_lib().imresh_server_submit(self._h, ptr, w, h, ctypes.byref(p), dev,
                                          SERVER_CALLBACK(), None, ctypes.byref(rid)),
              "imresh_server_submit")
        self._keep[rid.value] = io
        return rid.value

    def completions(self, capacity: int = 4096) -> list:
        buf = (ServerCompletion * capacity)()
        n = c_uint32()
        check(_lib().imresh_server_poll(self._h, buf, capacity, ctypes.byref(n)), "imresh_server_poll")
        out = []
        for i in range(n.value):
            c = buf[i]
            self._keep.pop(c.id, None)
            out.append((c.id, c.status, c.result, c.latency_ms))
        return out

    def drain(self) -> None:
        check(_lib().imresh_server_drain(self._h), "imresh_server_drain")

    def stats(self) -> ServerInfo:
        info = ServerInfo()
        check(_lib().imresh_server_stats(self._h, ctypes.byref(info)), "imresh_server_stats")
        return info

    def close(self) -> None:
        if self._h:
            check(_lib().imresh_server_destroy(self._h), "imresh_server_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


_TASK_REFS: dict = {}
_TASK_LOCK = threading.Lock()


def task_queue_init(streams_per_device: int = 0) -> None:
    check(_lib().imresh_task_queue_init(streams_per_device), "imresh_task_queue_init")


def task_queue_deinit() -> None:
    check(_lib().imresh_task_queue_deinit(), "imresh_task_queue_deinit")
    with _TASK_LOCK:
        _TASK_REFS.clear()


def add_task(io: np.ndarray, write_out, file_name: str = "", n_cycles: int = 20,
             n_hio_cycles: int = 20, target_error: float = 1e-5, hio_beta: float = 0.9,
             cutoff_autocorel: float = 0.04, cutoff: float = 0.2, sigma0: float = 3.0,
             sigma_change: float = 0.01) -> None:
    """imresh::io::addTask: runs the reconstruction of `io` (float32 H x W,
    modified in place) on a queue slot, then write_out(io, W, H, file_name)
    on a worker thread.  Parameter order as taskQueue.hpp:55-69."""
    if io.dtype != np.float32 or not io.flags.c_contiguous or io.ndim != 2:
        raise ValueError("io must be a C-contiguous float32 H x W array")
    h, w = io.shape
    p = Params(n_cycles, target_error, hio_beta, cutoff_autocorel, cutoff, sigma0, sigma_change,
               n_hio_cycles, DEFAULTS_QUEUE)
    key = id(io) ^ (threading.get_ident() << 1)

    def _cb(user, ptr, ww, hh, status, _io=io, _key=key):
        try:
            write_out(_io, ww, hh, file_name)
        finally:
            with _TASK_LOCK:
                _TASK_REFS.pop(_key, None)

    cb = TASK_CALLBACK(_cb)
    with _TASK_LOCK:
        _TASK_REFS[key] = (cb, io)
    check(_lib().imresh_add_task(io.ctypes.data, w, h, ctypes.byref(p), cb, None), "imresh_add_task")
