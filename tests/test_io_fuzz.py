"""Fuzz pass over the file readers (imresh_read_png / _tiff / _txt, host code
in imresh_amd/csrc/io.cpp) under AddressSanitizer and UndefinedBehaviorSanitizer.

The seeds are small valid files written by this suite's own encoders (every
PNG colour type and bit depth, little- and big-endian float TIFF, text);
tests/fuzz/io_fuzz.cpp mutates them (structure-aware for PNG, byte-level for
all) and reads every mutant.  Each must be read or rejected with a status;
a memory error or undefined behaviour aborts the fuzzer and fails the test.
Built with g++ from io.cpp alone (no GPU code), so it runs on the CPU.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from test_io import encode_png, write_tiff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write_seeds(d):
    rng = np.random.default_rng(1)
    k = 0
    for ctype, depths, nch in ((0, (1, 2, 4, 8, 16), 1), (2, (8, 16), 3), (3, (1, 2, 4, 8), 1),
                               (4, (8, 16), 2), (6, (8, 16), 4)):
        for depth in depths:
            hi = (1 << depth) - 1 if ctype != 3 else min(15, (1 << depth) - 1)
            s = rng.integers(0, hi + 1, size=(5, 9, nch))
            plte = bytes(rng.integers(0, 256, size=48, dtype=np.uint8)) if ctype == 3 else None
            encode_png(os.path.join(d, f"s{k}.png"), s, depth, ctype, plte=plte)
            k += 1
    a = rng.standard_normal((6, 5)).astype(np.float32)
    write_tiff(os.path.join(d, "le.tif"), a, False, 2)
    write_tiff(os.path.join(d, "be.tif"), a, True, 4)
    with open(os.path.join(d, "a.txt"), "w") as f:
        for r in a:
            f.write(" ".join(f"{x:.6g}" for x in r) + "\n")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_readers_survive_mutated_files(tmp_path):
    seeds = tmp_path / "seeds"
    seeds.mkdir()
    _write_seeds(str(seeds))
    exe = str(tmp_path / "io_fuzz")
    build = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"),
             os.path.join(ROOT, "imresh_amd", "csrc", "io.cpp"), os.path.join(ROOT, "tests", "fuzz", "io_fuzz.cpp"),
             "-lz", "-o", exe]
    r = subprocess.run(build, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(seeds), "400", "20261202", str(tmp_path / "cur")], capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "fuzz done" in r.stdout
    read_ok = int(r.stdout.split("fuzz done:")[1].split()[0])
    assert read_ok > 500   # the mutants reach the decoders, not only the header checks
