"""Writes tests/golden/converging.npz: end results of full 20 x 20
shrink-wrap reconstructions on workloads where the reconstruction
converges, computed by the float64 oracle (and the float32 one, for scale).

On these workloads every precision ends at the known object (NCC to it
1.000000 in float32 and float64), so the end result no longer depends on
rounding and can be compared across implementations: the GPU test
(tests/test_gpu_converging.py) regenerates the modulus with
imresh_amd.synthetic and compares its reconstruction with these numbers.

Per workload <w> (plain arrays, allow_pickle=False):
  <w>_spec      [width, n_shapes, central, index] of synthetic.workload (no Poisson noise)
  <w>_height    the image height (the width is spec[0])
  <w>_sigma0    the run's sigma0 (the other parameters are the defaults)
  <w>_err64     final HIO error of the float64 oracle
  <w>_err32     final HIO error of the float32 oracle
  <w>_err64_c1  HIO error after the first cycle (float64): the error scale
  <w>_masked    masked count of the last cycle (float64)
  <w>_ncc64, <w>_ncc32   NCC of the end result to the known object
  <w>_box       [y0, x0, h, w] of the region holding the reconstruction
  <w>_step      pixel stride of the stored crop (1: every pixel)
  <w>_crop64    the float64 end result inside the box, every step-th row
                and column (stored as float32)
  <w>_norm64    L2 norm of the whole float64 end result
  <w>_out64     L2 norm of the float64 end result outside the box
Runs for ~10 minutes (the 2048^2 runs take most of it).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from imresh_amd import synthetic as syn  # noqa: E402

FLT_MIN = float(np.finfo(np.float32).tiny)
# name -> (width, height, n_shapes, central, index, step, sigma0): found by
# running the oracle on seeds of the benchmark's object generator with 1-2
# shapes (objects with more shapes stagnate, differently in every
# precision); b300 and b1000 are lengths without a compiled plan (Bluestein
# transforms of 1024 and 2048 points on the GPU); c512s6 runs the 35-tap
# blur of sigma0 = 6; w8192 has 8192-point rows and h8192 8192-point columns
# (the four-step column pass); b3000 and b5000 have lines convolved over 8192
# and 16384 points.  Reconstructions that span the whole image
# (the object lands across its edges) keep every step-th row and column of
# the crop.
WORKLOADS = {
    "c256a": (256, 256, 1, 128, 100, 1, 3.0),
    "c256b": (256, 256, 2, 64, 101, 1, 3.0),
    "c2048": (2048, 2048, 2, 512, 101, 1, 3.0),
    "b300": (300, 300, 1, 150, 102, 1, 3.0),
    "b1000": (1000, 1000, 1, 500, 101, 4, 3.0),
    "c512s6": (512, 512, 1, 256, 102, 2, 6.0),
    "w8192": (8192, 64, 1, 32, 100, 4, 3.0),
    "h8192": (64, 8192, 1, 32, 100, 4, 3.0),
    "b3000": (3000, 64, 1, 32, 100, 4, 3.0),
    "b5000": (64, 5000, 1, 32, 100, 4, 3.0),
}


def box_of(x, margin=4):
    """Bounding box of the reconstructed object (|x| above 1% of the
    maximum), widened by `margin` pixels; what lies outside is the small
    residual the HIO error measures."""
    m = np.abs(x) > 1e-2 * np.abs(x).max()
    ys, xs = np.nonzero(m)
    y0, y1 = max(0, ys.min() - margin), min(x.shape[0], ys.max() + 1 + margin)
    x0, x1 = max(0, xs.min() - margin), min(x.shape[1], xs.max() + 1 + margin)
    return np.array([y0, x0, y1 - y0, x1 - x0], np.int64)


def main():
    oracle.set_fft_threads(int(os.environ.get("OMP_NUM_THREADS", "8")))
    out = {}
    for name, (n, hgt, shapes, central, idx, step, sigma0) in WORKLOADS.items():
        obj, M = syn.workload(n, hgt, n_shapes=shapes, central=central, index=idx)
        kw = dict(n_cycles=20, target_error=FLT_MIN, n_hio_cycles=20, sigma0=sigma0)
        x64, s64 = oracle.shrink_wrap(M, dtype=np.float64, **kw)
        x32, s32 = oracle.shrink_wrap(M, dtype=np.float32, **kw)
        _, c1 = oracle.shrink_wrap(M, dtype=np.float64, n_cycles=1, target_error=FLT_MIN, n_hio_cycles=20,
                                   sigma0=sigma0)
        b = box_of(x64)
        y0, x0, h, w = b
        out[f"{name}_spec"] = np.array([n, shapes, central, idx], np.int64)
        out[f"{name}_sigma0"] = np.float64(sigma0)
        out[f"{name}_height"] = np.int64(hgt)
        out[f"{name}_err64"] = np.float64(s64.last_error)
        out[f"{name}_err32"] = np.float64(s32.last_error)
        out[f"{name}_err64_c1"] = np.float64(c1.last_error)
        out[f"{name}_masked"] = np.float64(s64.last_masked)
        out[f"{name}_ncc64"] = np.float64(syn.ncc(obj, x64))
        out[f"{name}_ncc32"] = np.float64(syn.ncc(obj, x32))
        out[f"{name}_box"] = b
        out[f"{name}_step"] = np.int64(step)
        out[f"{name}_crop64"] = x64[y0:y0 + h:step, x0:x0 + w:step].astype(np.float32)
        out[f"{name}_norm64"] = np.float64(np.linalg.norm(x64))
        inside = np.zeros(x64.shape, bool)
        inside[y0:y0 + h, x0:x0 + w] = True
        out[f"{name}_out64"] = np.float64(np.linalg.norm(x64[~inside]))
        print(f"{name}: err64 {s64.last_error:.7g} err32 {s32.last_error:.7g} (first cycle {c1.last_error:.7g}); "
              f"NCC {out[f'{name}_ncc64']:.6f} / {out[f'{name}_ncc32']:.6f}; box {b.tolist()}; "
              f"fp32 vs fp64 rel L2 {np.linalg.norm(x32 - x64) / np.linalg.norm(x64):.2e}", flush=True)
    np.savez_compressed(os.path.join(HERE, "converging.npz"), **out)
    print("wrote", os.path.join(HERE, "converging.npz"))


if __name__ == "__main__":
    main()
