"""Shared test helpers: glibc rand() (the reference's KATs seed it with
srand(350471643)), float32 epsilon, device round trips."""
from __future__ import annotations

import ctypes
import ctypes.util
import math

import numpy as np

FLT_EPS = float(np.finfo(np.float32).eps)
FLT_MIN = float(np.finfo(np.float32).tiny)
KAT_SEED = 350471643          # testVectorReduce.cu:58, testGaussian.cpp:660
RAND_MAX = 2147483647

_libc = ctypes.CDLL(ctypes.util.find_library("c"))
_libc.rand.restype = ctypes.c_int
_libc.srand.argtypes = [ctypes.c_uint]


def srand(seed: int) -> None:
    _libc.srand(seed)


def rand() -> int:
    return _libc.rand()


def rand_uniform_m05(n: int) -> np.ndarray:
    """(float) rand() / RAND_MAX - 0.5f for n draws (testVectorReduce.cu:60)."""
    out = np.empty(n, dtype=np.float32)
    for i in range(n):
        out[i] = np.float32(np.float32(rand()) / np.float32(RAND_MAX)) - np.float32(0.5)
    return out


def log_spaced_points(start: int, end: int, n: int) -> list[int]:
    """getLogSpacedSamplingPoints (tests/benchmarkHelper.cpp:64-143), float32 arithmetic."""
    dx = np.float32((np.float32(math.log(end)) - np.float32(math.log(start))) / np.float32(n))
    factor = np.float32(math.exp(float(dx)))
    naive = [np.float32(start)]
    for _ in range(1, n):
        naive.append(np.float32(naive[-1] * factor))
    naive[-1] = np.float32(end)
    tmp = [int(naive[0])]
    for i in range(1, n):
        if tmp[-1] != int(naive[i]):
            tmp.append(int(naive[i]))
    to_insert = n - len(tmp)
    pts = [tmp[0]]
    for src in tmp[1:]:
        while to_insert > 0 and pts[-1] < src - 1:
            pts.append(pts[-1] + 1)
            to_insert -= 1
        if len(pts) < n:
            pts.append(src)
    return pts


def compare_float_array(a: np.ndarray, b: np.ndarray, max_kernel_width: int = 30) -> float:
    """testGaussian.cpp:53-85: max|a-b| / max(|a|,|b|) <= FLT_EPSILON * maxKernelWidth."""
    err = float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))) if a.size else 0.0
    mx = max(float(np.max(np.abs(a))) if a.size else 0.0, float(np.max(np.abs(b))) if b.size else 0.0)
    if mx == 0:
        mx = 1.0
    return err / mx
