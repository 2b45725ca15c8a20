"""Build the gfx950 library ``imresh_amd/lib/libimresh_hip.so`` (hipcc).

Every call recompiles all sources (the build check must exercise hipcc).
Compile flags:
  --offload-arch=gfx950              MI355X only
  -ffp-contract=off                  elementwise operators round exactly as the
                                     reference's C++ (the FFT uses explicit fma)
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE sqrtf and '/' (reference parity)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "imresh_amd", "csrc")
LIBDIR = os.path.join(ROOT, "imresh_amd", "lib")
LIBNAME = "libimresh_hip.so"
NUM_PLAN_GROUPS = 8   # IMR_NUM_PLAN_GROUPS in csrc/plans.hpp
# (source, object stem, extra defines): the row / column kernels are compiled
# once per plan group so the instantiations build in parallel
UNITS = ([("k_row.hip", f"k_row_g{g}", [f"-DIMR_GROUP={g}"]) for g in range(NUM_PLAN_GROUPS)]
         + [("k_col.hip", f"k_col_g{g}", [f"-DIMR_GROUP={g}"]) for g in range(NUM_PLAN_GROUPS)]
         + [(s, s.rsplit(".", 1)[0], []) for s in ("k_misc.hip", "k_poisson.hip", "k_step8.hip", "plan.cpp", "engine.cpp",
                                                    "batch.cpp", "serve.cpp", "task_queue.cpp", "io.cpp",
                                                    "noalloc.cpp")])


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build imresh_amd)")


def common_flags() -> list[str]:
    return [
        "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    ]


def _compile(src: str, obj: str, defines: list[str], verbose: bool) -> None:
    cmd = [hipcc(), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + common_flags() + defines
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr).strip()
    # hipcc's host pass does not know the packed-fp32 feature; that warning is expected
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n" + "\n".join(lines))
    if verbose and lines:
        print("\n".join(lines), flush=True)


def build(verbose: bool = False) -> str:
    objdir = os.path.join(ROOT, "imresh_amd", "_build")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = [os.path.join(objdir, stem + ".o") for _, stem, _ in UNITS]
    jobs = max(1, min(len(UNITS), os.cpu_count() or 4, 16))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda u: _compile(u[0][0], u[1], u[0][2], verbose), zip(UNITS, objs)))
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    _build_example(verbose)
    return out


EXAMPLE = os.path.join(LIBDIR, "imresh_queue_example")
MINI_EXAMPLE = os.path.join(LIBDIR, "imresh_shrinkwrap_example")


def _build_example(verbose: bool) -> None:
    """The C++ drop-in examples (examples/queue_png.cpp, examples/shrinkwrap_mini.cpp)
    against the built library."""
    for name, out in (("queue_png.cpp", EXAMPLE), ("shrinkwrap_mini.cpp", MINI_EXAMPLE)):
        src = os.path.join(ROOT, "imresh_amd", "examples", name)
        cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", out + ".tmp",
               "-L", LIBDIR, "-limresh_hip", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed ({name}):\n" + r.stdout + r.stderr)
        os.replace(out + ".tmp", out)


def build_oracle(verbose: bool = False) -> None:
    """CPU checker: oracle/liboracle.so, and oracle/_ref when the reference is mounted."""
    odir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-C", odir, "-B", "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if os.path.isdir("/root/reference/src/imresh"):
        r = subprocess.run(["make", "-C", odir, "-B", "ref"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference checker build failed:\n" + r.stdout + r.stderr)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    build_oracle()
