"""Writes tests/golden/oracle_golden.npz (run here, where the reference
sources are mounted and oracle/_ref is built; the .npz travels with the repo).

Contents (all plain arrays, loaded with allow_pickle=False):
  blur_in_<tag>, blur_sigma_<tag>, blur_out_<tag>
      random inputs and the 2-D Gaussian blur computed by the reference's
      own CPU sources (libs/gaussian.cpp via oracle/_ref)
  kern_sigmas, kern_weights
      calcGaussianKernel (libs/calcGaussianKernel.cpp via oracle/_ref)
  fft_in, fft_out
      a complex 24 x 40 input and numpy's float64 pocketfft fft2
  sw_modulus, sw_out_f64
      a 32 x 32 synthetic modulus and a 2-cycle x 3-iteration float64
      shrink-wrap run of the oracle (regression fixture)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from imresh_amd import synthetic as syn  # noqa: E402


def main():
    ref = oracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    rng = np.random.default_rng(20261202)
    out = {}
    for tag, (h, w, s) in {"a": (13, 17, 1.3), "b": (40, 33, 3.0), "c": (9, 70, 6.0)}.items():
        d = (rng.random((h, w), dtype=np.float32) - 0.5).astype(np.float32)
        b = d.copy()
        ref.ref_blur_f32(b.ctypes.data, w, h, s)
        out[f"blur_in_{tag}"], out[f"blur_sigma_{tag}"], out[f"blur_out_{tag}"] = d, np.float64(s), b
    sig = np.array([0.1, 0.5, 1.0, 3.0, 6.0, 6.4, 10.0])
    kw = np.zeros((len(sig), 64), np.float32)
    for i, s in enumerate(sig):
        ref.ref_gauss_kernel_f32(float(s), kw[i].ctypes.data, 64)
    out["kern_sigmas"], out["kern_weights"] = sig, kw
    x = rng.standard_normal((24, 40)) + 1j * rng.standard_normal((24, 40))
    out["fft_in"], out["fft_out"] = x, np.fft.fft2(x)
    obj, M = syn.workload(32, 32, n_shapes=4, central=16, index=5)
    o, st = oracle.shrink_wrap(M, n_cycles=2, n_hio_cycles=3, target_error=1e-38, dtype=np.float64)
    out["sw_modulus"], out["sw_out_f64"] = M, o
    np.savez_compressed(os.path.join(HERE, "oracle_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "oracle_golden.npz"))


if __name__ == "__main__":
    main()
