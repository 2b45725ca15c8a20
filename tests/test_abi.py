"""The C ABI library on the CPU: it loads without a GPU, exports every
function include/imresh_hip.h declares plus the C++ drop-in API, and its
host-only functions (defaults, status strings, Gaussian weights) agree with
the reference semantics.  No device work is started here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import imresh_amd as im
import oracle

HEADER = im.HEADER_PATH


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(imresh_[a-z0-9_]+)\s*\(", text)))


def test_library_built_and_loads():
    assert os.path.exists(im.LIB_PATH), "build first: __graft_entry__.build()"
    im.lib()


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    L = ctypes.CDLL(im.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(names) <= set(im._SIGS), sorted(set(names) - set(im._SIGS))


def test_exports_cpp_dropin_api():
    out = subprocess.run(["nm", "-DC", "--defined-only", im.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("imresh::algorithms::shrinkWrap(float*, unsigned int, unsigned int, unsigned int, "
                "float, float, float, float, float, float, unsigned int)",
                "imresh::io::taskQueueInit()", "imresh::io::taskQueueDeinit()",
                "imresh::io::addTask(",
                "imresh::libs::hybridInputOutput(float*, unsigned char const*, std::vector<unsigned int",
                "imresh::libs::diffractionIntensity(float*, unsigned int, unsigned int, ihipStream_t*, bool)",
                "imresh::libs::fftShift<float>(float*, unsigned int, unsigned int)",
                "imresh::libs::fftShift<float [2]>(float (*) [2], unsigned int, unsigned int)",
                "imresh::io::readInFuncs::readPNG(",
                "imresh::io::readInFuncs::readTiff(",
                "imresh::io::readInFuncs::readTxt(",
                "imresh::io::writeOutFuncs::writeOutPNG(float*, unsigned int, unsigned int, std::",
                "imresh::io::writeOutFuncs::writeOutAndFreePNG(",
                "imresh::io::writeOutFuncs::justFree("):
        assert sym in out, sym


def test_status_strings_and_version():
    L = im.lib()
    assert L.imresh_abi_version() == 1
    assert L.imresh_strerror(0) == b"ok"
    assert b"null" in L.imresh_strerror(1)
    assert L.imresh_strerror(12345) == b"unknown status"


def test_params_defaults_match_reference():
    """shrinkWrap.hpp:45-52 defaults."""
    p = im.default_params()
    assert (p.n_cycles, p.n_hio_cycles) == (20, 20)
    assert np.float32(p.target_error) == np.float32(1e-5)
    assert np.float32(p.hio_beta) == np.float32(0.9)
    assert np.float32(p.cutoff_autocorel) == np.float32(0.04)
    assert np.float32(p.cutoff) == np.float32(0.2)
    assert (p.sigma0, np.float32(p.sigma_change)) == (3.0, np.float32(0.01))


@pytest.mark.parametrize("sigma", [0.1, 0.5, 1.0, 1.5, 2.9403, 3.0, 4.0, 6.0, 6.4, 10.9])
def test_host_gaussian_weights_bit_exact(sigma):
    """The library's host weights (uploaded to the device blur) equal the
    oracle's (and, through test_oracle, the reference's calcGaussianKernel)."""
    assert np.array_equal(im.gaussian_kernel(sigma), oracle.gauss_kernel(sigma))


def test_null_pointer_checks_need_no_device():
    L = im.lib()
    assert L.imresh_complex_norm(None, None, 4, None) == 1
    assert L.imresh_fft2d(None, None, 4, 4, -1, None) == 1
    assert L.imresh_shrinkwrap(None, 4, 4, None, None, None) == 1
    assert L.imresh_gaussian_kernel(3.0, None, 0) == 19
