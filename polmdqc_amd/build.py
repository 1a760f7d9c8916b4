"""Builds polmdqc_amd/libpmq.so (HIP, gfx950) in-tree with hipcc.

    python -m polmdqc_amd.build [--jobs N]

Every translation unit under polmdqc_amd/csrc is compiled as HIP for gfx950
and linked into one shared library whose exported symbols are exactly the
C ABI of include/pmq.h.  The C++ host layer (polmdqc_amd/host: the
reference's energy/empole/epolar<CalcMode>() over global arrays, calling the
C ABI) is built with g++ into libpmq_host.so, and its test driver
tests/host/host_check against both.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpmq.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-x", "hip",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]
LDFLAGS = [f"--offload-arch={ARCH}", "-shared", "-fPIC", "-L/opt/rocm/lib", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib"]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


# Kernels whose pair algebra is evaluated with some multipole components or
# pair-vector components known to be zero (induced dipoles carry no charge
# or quadrupole; the quasi-internal frame's pair vector lies on z): these
# flags let the compiler drop the x*0 terms (no signed zeros, no NaN/Inf
# inputs).  They do not reassociate arithmetic.
EXTRA = {"k_polar.hip": ["-fno-signed-zeros", "-ffinite-math-only"],
         # the quasi-internal-frame pairs evaluate the algebra at r = (0, 0, |r|)
         "k_mpole.hip": ["-fno-signed-zeros", "-ffinite-math-only"],
         # parameter derivatives evaluate the pair terms with unit multipoles
         "k_params.hip": ["-fno-signed-zeros", "-ffinite-math-only"],
         "k_batch.hip": ["-fno-signed-zeros", "-ffinite-math-only"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, *EXTRA.get(os.path.basename(src), []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, *LDFLAGS, *objs, "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    build_host(verbose)
    return LIB


HOST_SRC = os.path.join(HERE, "host", "pmq_host.cpp")
HOST_LIB = os.path.join(HERE, "libpmq_host.so")
HOST_CHECK_SRC = os.path.join(ROOT, "tests", "host", "host_check.cpp")
HOST_CHECK = os.path.join(ROOT, "tests", "host", "host_check")


def build_host(verbose: bool = True) -> str:
    """g++: libpmq_host.so (links libpmq.so) and the host_check driver."""
    gxx = shutil.which("g++") or "g++"
    rpath = "-Wl,-rpath,$ORIGIN"
    cmds = [
        [gxx, "-std=c++17", "-O2", "-fPIC", "-shared", HOST_SRC, "-o", HOST_LIB, f"-L{HERE}", "-lpmq", rpath],
        [gxx, "-std=c++17", "-O2", HOST_CHECK_SRC, "-o", HOST_CHECK, f"-L{HERE}", "-lpmq_host", "-lpmq",
         f"-Wl,-rpath,{os.path.relpath(HERE, os.path.dirname(HOST_CHECK)).join(['$ORIGIN/', ''])}"],
    ]
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed:\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {HOST_LIB} and {HOST_CHECK}")
    return HOST_LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
