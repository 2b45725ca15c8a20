"""Synthetic reconstruction workloads and image metrics (host side, numpy).

Object recipe (fixed here because the workload definition leaves it open):
  * zero background; `n_shapes` axis-aligned rectangles and discs (alternating,
    first a rectangle) whose centres are uniform in the central region
    (size `central` x `central`, centred in the frame) and whose half-extent
    (rectangles, per axis) or radius (discs) is uniform in
    [0.02, 0.12] * central; shapes are clipped to the central region;
  * amplitude uniform in [0.2, 1]; overlapping shapes keep the maximum;
  * intensity I = |FFT2(object)|^2 evaluated in float64 (unshifted, DC at
    [0, 0], as the reference's diffractionIntensity, libs/diffractionIntensity.cpp:49-80);
  * optional Poisson sampling at `photons` total: lambda_i = photons*I_i/sum(I),
    numpy's Generator.poisson, then cast to float32;
  * the reconstruction input is the modulus sqrt(I) as float32: the reference
    projects |G| := io[i] (vectorElementwise.cpp:73-79), i.e. it consumes |F|.
Seeds: numpy.random.default_rng(splitmix64(base_seed ^ index)), base seed 20261202.
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 20261202
_MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & _MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return z ^ (z >> 31)


def rng_for(index: int = 0, base_seed: int = BASE_SEED) -> np.random.Generator:
    return np.random.default_rng(splitmix64((base_seed ^ index) & _MASK64))


def make_object(width: int, height: int, n_shapes: int, central: int,
                rng: np.random.Generator) -> np.ndarray:
    """Rectangles and discs in the central `central` x `central` region (float64)."""
    obj = np.zeros((height, width), dtype=np.float64)
    c = int(min(central, width, height))
    x0, y0 = (width - c) // 2, (height - c) // 2
    region = obj[y0:y0 + c, x0:x0 + c]

    def box(center, half):
        # pixel range that can satisfy the shape test (one pixel of margin)
        return max(0, int(np.floor(center - half)) - 1), min(c, int(np.ceil(center + half)) + 2)

    for s in range(n_shapes):
        cx, cy = rng.uniform(0, c, size=2)
        amp = rng.uniform(0.2, 1.0)
        if s % 2 == 0:
            hx, hy = rng.uniform(0.02, 0.12, size=2) * c
            (ax, bx), (ay, by) = box(cx, hx), box(cy, hy)
            yy, xx = np.mgrid[ay:by, ax:bx]
            m = (np.abs(xx - cx) <= hx) & (np.abs(yy - cy) <= hy)
        else:
            r = rng.uniform(0.02, 0.12) * c
            (ax, bx), (ay, by) = box(cx, r), box(cy, r)
            yy, xx = np.mgrid[ay:by, ax:bx]
            m = (xx - cx) ** 2 + (yy - cy) ** 2 <= r * r
        sub = region[ay:by, ax:bx]
        np.maximum(sub, np.where(m, amp, 0.0), out=sub)
    return obj


def intensity(obj: np.ndarray) -> np.ndarray:
    f = np.fft.fft2(obj)
    return (f.real ** 2 + f.imag ** 2)


def poisson_sample(I: np.ndarray, photons: float, rng: np.random.Generator) -> np.ndarray:
    lam = photons * I / I.sum()
    return rng.poisson(lam).astype(np.float64)


def workload(width: int, height: int | None = None, n_shapes: int = 8, central: int | None = None,
             photons: float | None = None, index: int = 0, base_seed: int = BASE_SEED):
    """Returns (object float64, modulus float32) for one synthetic image."""
    height = width if height is None else height
    central = min(width, height) // 2 if central is None else central
    rng = rng_for(index, base_seed)
    obj = make_object(width, height, n_shapes, central, rng)
    I = intensity(obj)
    if photons is not None:
        I = poisson_sample(I, photons, rng)
    return obj, np.sqrt(I).astype(np.float32)


# the configurations of the benchmark definition
def config_256(index: int = 0):
    """256 x 256, 8 shapes in the central 128 x 128, no Poisson noise."""
    return workload(256, 256, n_shapes=8, central=128, photons=None, index=index)


def config_2048(index: int = 0):
    """2048 x 2048, 64 shapes in the central 1024 x 1024, 1e9 photons."""
    return workload(2048, 2048, n_shapes=64, central=1024, photons=1e9, index=index)


def config_8192(index: int = 0):
    """8192 x 8192, 512 shapes in the central 4096 x 4096, 1e11 photons."""
    return workload(8192, 8192, n_shapes=512, central=4096, photons=1e11, index=index)


BATCH_SIZES = (64, 96, 128, 200, 256, 384, 512, 768, 1024)


def batch_shapes(n: int) -> tuple[int, int]:
    """Central region n/2 and 8 shapes per 128^2 of it (at least 2)."""
    central = n // 2
    return max(2, int(round(8 * (central / 128.0) ** 2))), central


def batch_sizes(count: int = 512, index: int = 0) -> list[int]:
    """Image sizes of the batched workload (uniform over BATCH_SIZES)."""
    rng = rng_for(10_000_000 + index)
    return [int(BATCH_SIZES[k]) for k in rng.integers(0, len(BATCH_SIZES), count)]


def config_batch(count: int = 512, index: int = 0):
    """The batched workload: `count` square images, sizes uniform over
    BATCH_SIZES (a log-spaced set, so uniform over it is log-uniform), each
    with batch_shapes(n) shapes and 1e7 photons.  Returns (sizes, objects,
    moduli)."""
    sizes = batch_sizes(count, index)
    objs, mods = [], []
    for i, n in enumerate(sizes):
        shapes, central = batch_shapes(n)
        o, m = workload(n, n, n_shapes=shapes, central=central, photons=1e7,
                        index=20_000_000 + 1000 * index + i)
        objs.append(o)
        mods.append(m)
    return sizes, objs, mods


def fast_modulus(n: int, index: int, photons: float = 1e9) -> np.ndarray:
    """A serving request's input: the object of batch_shapes(n), intensity
    from a complex64 FFT (scipy, all host threads: input generation only),
    Poisson-sampled at `photons`, modulus = sqrt."""
    import scipy.fft
    rng = rng_for(30_000_000 + index)
    shapes, central = batch_shapes(n)
    obj = make_object(n, n, shapes, central, rng)
    F = scipy.fft.fft2(obj.astype(np.complex64), workers=-1)
    I = (F.real.astype(np.float64) ** 2 + F.imag.astype(np.float64) ** 2)
    I = poisson_sample(I, photons, rng)
    return np.sqrt(I).astype(np.float32)


def serve_trace(count: int, rate: float = 200.0, lo: int = 128, hi: int = 2048, index: int = 0):
    """The serving workload: Poisson arrivals at `rate` per second, square
    sizes round(exp(U[ln lo, ln hi])), 5 / 10 / 20 cycles of 20 HIO
    iterations with equal probability.  Returns [(arrival_s, n, cycles)]."""
    rng = rng_for(40_000_000 + index)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    t -= t[0]
    sizes = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    cycles = rng.choice([5, 10, 20], count)
    return [(float(a), int(n), int(c)) for a, n, c in zip(t, sizes, cycles)]


# ------------------------------------------------------------------ metrics
def ncc(a: np.ndarray, b: np.ndarray) -> float:
    """Normalized cross-correlation maximised over circular translations and
    the twin image b(-x) (phase retrieval is blind to both)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    a = a - a.mean()
    b = b - b.mean()
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    if na == 0 or nb == 0:
        return 0.0
    Fa = np.fft.fft2(a)
    best = 0.0
    for bb in (b, np.roll(b[::-1, ::-1], 1, axis=(0, 1))):
        c = np.fft.ifft2(Fa * np.conj(np.fft.fft2(bb))).real
        best = max(best, float(c.max()))
    return best / (na * nb)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a.astype(np.complex128) if np.iscomplexobj(a) else a.astype(np.float64))
                       - (b.astype(np.complex128) if np.iscomplexobj(b) else b.astype(np.float64)))
    n = np.linalg.norm(b.astype(np.complex128) if np.iscomplexobj(b) else b.astype(np.float64))
    return float(d / n) if n else float(d)
