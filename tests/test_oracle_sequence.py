"""Pins of the oracle's shrink-wrap cycle sequence (CPU, no GPU).

The oracle's per-operator functions are pinned one by one in
test_oracle.py.  These tests pin how orc_shrinkwrap strings them together,
following src/imresh/algorithms/shrinkWrap.cpp:111-361:

  cur = IFFT2_unnorm((M, 0))                                   :157-164
  gPrev = cur                                                  :173-178
  per cycle c:
    mask = |cur| -> blur(sigma_c) -> < cut_c * max  (1 = out)  :191, :206, :234-242
           cut_0 = cutOffAutoCorel, cut_c = cutOff for c >= 1  :236-239
    sigma_{c+1} = max(1.5, (1 - dSigma) * sigma_c)  (float)    :245
    n_hio x { gPrev = constraint(cur, gPrev, mask, beta)       :263-276
              cur = FFT2(gPrev); cur = modulus(cur, M);        :279, :282
              cur = IFFT2(cur) }                               :301
    error = sqrtf(sum mask |cur|) / sum mask; exit if < target :331-335
  out = Re(cur)                                                :349-351

The sequence is rebuilt here in numpy: the elementwise steps (norm, max,
cut-off, constraint, modulus projection, copy) are written out directly in
float64 from the formulas above, and only the 2-D FFT and the blur are
taken from the oracle (pinned against numpy / brute force and the
reference's own compiled gaussian.cpp).  One and two cycles of
orc_shrinkwrap must equal this bit for bit in float64.  Each plausible
sequencing mistake (sigma updated before the blur, cutOff used in cycle 0,
cutOffAutoCorel kept after it, the mask built from gPrev, the modulus
projection dropped) is built as well and must give different bits, so the
test cannot pass for a sequence that makes one of them.

A closed form pins the HIO iteration on its own: for a real object
obj >= 0 and M = |FFT2(obj)|, one iteration from g' = s * obj (s > 0) with
every pixel inside the support returns N * obj whatever s is.
"""
import numpy as np
import pytest

import oracle

F32 = np.float32


def _fft2(x, sign):
    return oracle.fft2d(np.asarray(x, np.complex128), sign, np.float64)


def _mask_from(g, sigma, cut):
    """mask = (blur(|g|) < cut * max(blur(|g|))) ? 1 : 0, float64."""
    re, im = g.real.copy(), g.imag.copy()
    norm = np.sqrt(re * re + im * im)
    blurred = oracle.blur(norm, float(sigma))
    thr = blurred.max() * float(F32(cut))
    return np.where(blurred < thr, 1.0, 0.0)


def _hio(g, gp, mask, M, beta, n_iter, project=True):
    re, im = g.real.copy(), g.imag.copy()
    pre, pim = gp.real.copy(), gp.imag.copy()
    b = float(F32(beta))
    for _ in range(n_iter):
        out = (mask == 1) | (re < 0)
        pre = np.where(out, pre - b * re, re)
        pim = np.where(out, pim - b * im, im)
        G = _fft2(pre + 1j * pim, -1)
        gre, gim = G.real.copy(), G.imag.copy()
        if project:
            n = np.sqrt(gre * gre + gim * gim)
            n = np.where(n == 0, 1.0, n)
            f = M / n
            gre, gim = gre * f, gim * f
        g2 = _fft2(gre + 1j * gim, +1)
        re, im = g2.real.copy(), g2.imag.copy()
    return re + 1j * im, pre + 1j * pim


def _sequence(M, n_cycles, n_hio, beta, cutA, cut, sigma0, dsigma, mistake=None):
    """The shrink-wrap sequence composed from the steps above (float64)."""
    M = M.astype(np.float64)
    g = _fft2(M + 0j, +1)
    gp = g.copy()
    sigma = F32(sigma0)
    for c in range(n_cycles):
        s_blur = sigma
        if mistake == "sigma_first":
            s_blur = max(F32(1.5), F32((F32(1) - F32(dsigma)) * sigma))
        c_cut = cutA if c == 0 else cut
        if mistake == "cut_at_0" and c == 0:
            c_cut = cut
        if mistake == "cutA_kept":
            c_cut = cutA
        src = gp if (mistake == "mask_from_gprev" and c > 0) else g
        mask = _mask_from(src, s_blur, c_cut)
        sigma = max(F32(1.5), F32((F32(1) - F32(dsigma)) * sigma))
        g, gp = _hio(g, gp, mask, M, beta, n_hio, project=(mistake != "no_projection"))
    return g.real


PARAMS = dict(beta=0.9, cutA=0.04, cut=0.2, sigma0=3.0, dsigma=0.3)
MISTAKES = ["sigma_first", "cut_at_0", "cutA_kept", "mask_from_gprev", "no_projection"]


def _workload(w, h, seed):
    """A small non-square object (so a transposed index shows) and its
    float64 Fourier modulus."""
    rng = np.random.default_rng(seed)
    obj = np.zeros((h, w))
    for _ in range(5):
        x0, y0 = rng.integers(w // 4, 3 * w // 4 - 4), rng.integers(h // 4, 3 * h // 4 - 4)
        obj[y0:y0 + rng.integers(2, 6), x0:x0 + rng.integers(2, 6)] = rng.uniform(0.2, 1.0)
    return obj, np.abs(np.fft.fft2(obj))


@pytest.mark.parametrize("n_cycles", [1, 2, 3])
def test_shrinkwrap_sequence_bit_exact_float64(n_cycles):
    obj, M = _workload(40, 28, 7)
    p = PARAMS
    out, st = oracle.shrink_wrap(M, n_cycles=n_cycles, target_error=np.finfo(np.float32).tiny,
                                 hio_beta=p["beta"], cutoff_autocorel=p["cutA"], cutoff=p["cut"],
                                 sigma0=p["sigma0"], sigma_change=p["dsigma"], n_hio_cycles=3,
                                 dtype=np.float64)
    assert st.cycles_run == n_cycles
    exp = _sequence(M, n_cycles, 3, **p)
    assert np.array_equal(out, exp), float(np.max(np.abs(out - exp)))
    # every planted sequencing mistake changes the bits (so the equality
    # above discriminates between them); mistakes that act only from cycle 1
    # on are checked where a second cycle runs
    for m in MISTAKES:
        if n_cycles == 1 and m in ("cutA_kept", "mask_from_gprev"):
            continue
        alt = _sequence(M, n_cycles, 3, **p, mistake=m)
        assert not np.array_equal(alt, exp), m


def test_shrinkwrap_sequence_default_parameters():
    """The default parameters (sigma 3.0 -> 2.97, cut 0.04 -> 0.2) over two
    cycles of 20 iterations."""
    obj, M = _workload(48, 36, 11)
    out, st = oracle.shrink_wrap(M, n_cycles=2, target_error=np.finfo(np.float32).tiny,
                                 n_hio_cycles=20, dtype=np.float64)
    exp = _sequence(M, 2, 20, beta=0.9, cutA=0.04, cut=0.2, sigma0=3.0, dsigma=0.01)
    assert np.array_equal(out, exp)


def test_shrinkwrap_error_metric_is_last_cycles_masked_norm():
    """The error returned after the last cycle is sqrtf(sum mask |g'|) /
    sum mask over that cycle's mask (hybridInputOutput.cpp:67-103),
    recomputed here from the composed sequence."""
    obj, M = _workload(40, 28, 5)
    kw = dict(n_cycles=2, target_error=np.finfo(np.float32).tiny, hio_beta=0.9, cutoff_autocorel=0.04,
              cutoff=0.2, sigma0=3.0, sigma_change=0.3, n_hio_cycles=3)
    _, st = oracle.shrink_wrap(M, dtype=np.float64, **kw)
    # rebuild the last cycle's mask and g'
    M64 = M.astype(np.float64)
    g = _fft2(M64 + 0j, +1)
    gp = g.copy()
    sigma = F32(3.0)
    for c in range(2):
        mask = _mask_from(g, sigma, 0.04 if c == 0 else 0.2)
        sigma = max(F32(1.5), F32((F32(1) - F32(0.3)) * sigma))
        g, gp = _hio(g, gp, mask, M64, 0.9, 3)
    tot = float(np.sum(mask * np.abs(g)))
    nm = float(mask.sum())
    assert st.last_masked == nm
    assert abs(st.last_total - tot) <= 1e-6 * tot
    assert abs(st.last_error - np.sqrt(tot) / nm) <= 1e-6 * st.last_error


@pytest.mark.parametrize("scale", [1.0, 2.5, 1e-3])
def test_hio_iteration_fixed_point_closed_form(scale):
    """obj >= 0 inside the support (mask 0 everywhere), M = |FFT2(obj)|: one
    iteration from g' = scale * obj, gPrev = 0 gives gPrev = scale * obj and
    g' = N * obj exactly up to FFT rounding (the projection restores |F|)."""
    obj, _ = _workload(36, 20, 3)
    M = np.sqrt(np.abs(_fft2(obj + 0j, -1)) ** 2)
    mask = np.zeros_like(obj)
    g = (scale * obj).astype(np.complex128)
    gp = np.zeros_like(g)
    g1, gp1 = oracle.hio_iterations(g, gp, mask, M, 0.9, 1, np.float64)
    N = obj.size
    assert np.array_equal(gp1, g)
    assert np.max(np.abs(g1 - N * obj)) <= 1e-12 * N * obj.max()
    # a fixed point: a second iteration stays at N * obj
    g2, _ = oracle.hio_iterations(g1, gp1, mask, M, 0.9, 1, np.float64)
    assert np.max(np.abs(g2 - N * obj)) <= 1e-12 * N * obj.max()


def test_hio_iteration_outside_support_closed_form():
    """Pixels outside the support (mask 1) take gPrev - beta g' and pixels
    inside take g', before the transforms: with M = |FFT2(expected gPrev)|
    the iteration returns N * gPrev when that gPrev is real and >= 0."""
    rng = np.random.default_rng(4)
    h, w = 24, 32
    g = rng.uniform(0.1, 1.0, (h, w)) + 0j
    mask = (rng.random((h, w)) < 0.3).astype(np.float64)
    gp = g * (1 + 0.9) * (mask == 1) + rng.uniform(0, 1, (h, w)) * (mask == 0)
    exp_gp = np.where(mask == 1, gp - 0.9 * g, g)   # = g everywhere here
    M = np.abs(_fft2(exp_gp, -1))
    g1, gp1 = oracle.hio_iterations(g, gp, mask, M, 0.9, 1, np.float64)
    assert np.max(np.abs(gp1 - exp_gp)) <= 1e-15 * np.abs(exp_gp).max() * 4
    assert np.max(np.abs(g1 - g.size * exp_gp)) <= 1e-12 * g.size


@pytest.mark.parametrize("name", ["c256a", "b300"])
def test_converging_fixture_float32_within_the_gpu_bounds(name):
    """tests/golden/converging.npz (make_converging.py): the float64 oracle
    reproduces its end result (256^2, and 300^2 whose length has no compiled
    GPU plan), and the float32 oracle -- the reference's own precision --
    meets the bounds the GPU test applies (NCC within 1e-3, the
    reconstruction within 1e-3 relative L2)."""
    import os
    from imresh_amd import synthetic as syn
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "converging.npz"), allow_pickle=False)
    oracle.set_fft_threads(8)
    n, shapes, central, idx = (int(v) for v in z[f"{name}_spec"])
    obj, M = syn.workload(n, int(z[f"{name}_height"]), n_shapes=shapes, central=central, index=idx)
    kw = dict(n_cycles=20, target_error=np.finfo(np.float32).tiny, n_hio_cycles=20,
              sigma0=float(z[f"{name}_sigma0"]))
    x64, s64 = oracle.shrink_wrap(M, dtype=np.float64, **kw)
    y0, x0, h, w = (int(v) for v in z[f"{name}_box"])
    st = int(z[f"{name}_step"])
    assert np.array_equal(x64[y0:y0 + h:st, x0:x0 + w:st].astype(np.float32), z[f"{name}_crop64"])
    assert s64.last_error == np.float32(z[f"{name}_err64"])
    x32, s32 = oracle.shrink_wrap(M, dtype=np.float32, **kw)
    ref = z[f"{name}_crop64"].astype(np.float64)
    rel = np.linalg.norm(x32[y0:y0 + h:st, x0:x0 + w:st] - ref) / np.linalg.norm(ref)
    assert rel <= 1e-3
    assert abs(syn.ncc(obj, x32) - float(z[f"{name}_ncc64"])) <= 1e-3
    assert abs(s32.last_error - float(z[f"{name}_err64"])) <= 1e-3 * float(z[f"{name}_err64_c1"])
