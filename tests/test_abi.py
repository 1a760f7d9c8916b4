"""The C-ABI library loads and exports every entry point include/pmq.h declares
(CPU only: no function that touches a GPU is called)."""
import ctypes
import os
import re

from conftest import ROOT


def declared():
    text = open(os.path.join(ROOT, "include", "pmq.h")).read()
    return sorted(set(re.findall(r"^\s*PMQ_API\s+(?:int|const char \*)\s*(pmq_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for need in ("pmq_create", "pmq_destroy", "pmq_set_topology", "pmq_set_params", "pmq_eval",
                 "pmq_set_coords", "pmq_compute", "pmq_fetch"):
        assert need in names


def test_library_exports_every_declared_symbol():
    from polmdqc_amd import _lib

    L = _lib.lib()
    for name in declared():
        assert hasattr(L, name), name
    assert L.pmq_abi_version() == 6
    assert set(_lib.EXPORTS) <= set(declared())


def test_library_exports_nothing_else():
    """Only the C ABI is visible (everything else is built hidden)."""
    import subprocess

    from polmdqc_amd import _lib

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if " T " in line}
    pmq = {s for s in syms if s.startswith("pmq_")}
    assert pmq == set(declared())


def test_header_and_bindings_agree_on_fft_flags():
    """The PME transform constants of the header match the Python bindings,
    and bench.py's --pin-fft spellings map onto them (no GPU needed)."""
    import sys

    sys.path.insert(0, ROOT)
    import bench
    from polmdqc_amd import engine as E

    hdr = open(os.path.join(ROOT, "include", "pmq.h")).read()
    consts = dict(re.findall(r"#define (PMQ_FFT_\w+) (\d+)", hdr))
    assert int(consts["PMQ_FFT_AUTO"]) == E.FFT_AUTO
    assert int(consts["PMQ_FFT_LIBRARY"]) == E.FFT_LIBRARY
    assert int(consts["PMQ_FFT_DIRECT"]) == E.FFT_DIRECT
    assert int(consts["PMQ_FFT_GRIDS_FP32"]) == E.FFT_GRIDS_FP32
    assert int(consts["PMQ_FFT_GRIDS_FP64"]) == E.FFT_GRIDS_FP64
    assert bench.pin_fft_values("library/fp32-direct") == (E.FFT_LIBRARY, E.FFT_DIRECT | E.FFT_GRIDS_FP32)
    assert bench.pin_fft_values("direct/fp64-library") == (E.FFT_DIRECT, E.FFT_LIBRARY | E.FFT_GRIDS_FP64)
