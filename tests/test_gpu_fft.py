"""GPU parity of the PME transform engines and gather layouts: the direct
LDS transforms (k_fft.hip: forward planes, z columns with the reciprocal
kernel, inverse planes) against rocFFT and against the oracle, and the two
gather layouts (a lane per stencil row, a lane per stencil plane).

Both engines compute the same unnormalized DFTs in different summation
orders, so the plain FP64 results agree to round-off (1e-12 relative) and
each meets the oracle bar (1e-9).  Grids: the cubic 24^3 of the 216-water
box and a non-cubic one with an odd first axis (27 x 20 x 32), which
exercises the half-complex bookkeeping (no m = K0/2 term).
"""
import numpy as np
import pytest

import oracle
from polmdqc_amd import water

pytestmark = pytest.mark.gpu
REL = 1e-9


@pytest.fixture(scope="module")
def engine():
    from polmdqc_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def box216(grid=(24, 24, 24)):
    xyz = water.lattice_box(6, 18.643, 216)
    s = water.water_system(xyz, "water03", box=18.643,
                           ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})
    s["nfft1"], s["nfft2"], s["nfft3"] = grid
    s["poleps"] = 1e-10
    return s


def run(engine, s, fft, precision):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import set_fft

    engine.set_system(s)
    set_fft(engine, fft)
    try:
        return engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR, precision)
    finally:
        set_fft(engine)


@pytest.mark.parametrize("grid", [(24, 24, 24), (27, 20, 32)])
def test_direct_transforms_match_library_and_oracle(engine, grid):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import FFT_DIRECT, FFT_LIBRARY

    print(f"[gpu] fft engines grid={grid}", flush=True)
    s = box216(grid)
    d = run(engine, s, FFT_DIRECT, L.PLAIN_FP64)
    b = run(engine, s, FFT_LIBRARY, L.PLAIN_FP64)
    for k in ("em", "ep"):
        assert abs(d[k] - b[k]) <= 1e-12 * abs(b[k]), (k, d[k], b[k])
    for k in ("dem", "dep", "uind"):
        assert rel(d[k], b[k]) <= 1e-11, k
    o = oracle.elec(s, want_grad=True, poleps=1e-10)
    for k in ("em", "ep"):
        assert abs(d[k] - o[k]) <= REL * abs(o[k])
    assert rel(d["dem"], o["dem"]) <= REL
    assert rel(d["dep"], o["dep"]) <= REL


def test_direct_transforms_mixed_form(engine):
    """FP32 grids through the direct transforms: the mixed result stays
    within its error bar of the FP64 one."""
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import FFT_DIRECT, FFT_LIBRARY

    print("[gpu] fft engines mixed", flush=True)
    s = box216()
    ref = run(engine, s, FFT_LIBRARY, L.PLAIN_FP64)
    for fft in (FFT_DIRECT, FFT_LIBRARY):
        m = run(engine, s, fft, L.MIXED)
        for k in ("em", "ep"):
            assert abs(m[k] - ref[k]) <= 1e-5 * abs(ref[k]), (fft, k, m[k], ref[k])
        rms = float(np.sqrt(np.mean(np.sum((m["desum"] - ref["desum"]) ** 2, 1))))
        assert rms < 1e-4, (fft, rms)


def test_direct_transforms_dispersion(engine):
    """HIPPO dispersion PME (its own kernel table, the m = 0 term kept)
    through the direct transforms against the oracle."""
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import FFT_DIRECT, set_fft

    print("[gpu] fft engines dispersion", flush=True)
    xyz = water.lattice_box(6, 18.643, 216)
    s = water.water_system(xyz, "hippo19", box=18.643,
                           ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0, "dewald": 0.45})
    engine.set_system(s)
    set_fft(engine, FFT_DIRECT)
    try:
        g = engine.eval(L.GRADIENT, L.TERM_DISP)
    finally:
        set_fft(engine)
    o = oracle.hippo_terms(s, want_grad=True)
    assert abs(g["edsp"] - o["edsp"]) <= REL * abs(o["edsp"])
    assert rel(g["dedsp"], o["dedsp"]) <= REL


@pytest.mark.parametrize("mode", ["rows", "planes"])
def test_gather_layouts_match_oracle(engine, mode):
    """Both gather layouts (one lane per stencil row, one per stencil plane)
    on multipoles, polarization and dispersion: each within the oracle bar,
    and the two within round-off of each other."""
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import GATHER_AUTO, GATHER_PLANES, GATHER_ROWS, set_gather_lanes

    print(f"[gpu] gather layout {mode}", flush=True)
    s = box216()
    engine.set_system(s)
    set_gather_lanes(engine, GATHER_PLANES if mode == "planes" else GATHER_ROWS)
    try:
        g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
        set_gather_lanes(engine, GATHER_ROWS if mode == "planes" else GATHER_PLANES)
        h = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    finally:
        set_gather_lanes(engine, GATHER_AUTO)
    o = oracle.elec(s, want_grad=True, poleps=1e-10)
    for k in ("em", "ep"):
        assert abs(g[k] - o[k]) <= REL * abs(o[k])
        assert abs(g[k] - h[k]) <= 1e-12 * abs(h[k])
    assert rel(g["dem"], o["dem"]) <= REL
    assert rel(g["dep"], o["dep"]) <= REL

    xyz = water.lattice_box(6, 18.643, 216)
    sd = water.water_system(xyz, "hippo19", box=18.643,
                            ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0, "dewald": 0.45})
    engine.set_system(sd)
    set_gather_lanes(engine, GATHER_PLANES if mode == "planes" else GATHER_ROWS)
    try:
        d = engine.eval(L.GRADIENT, L.TERM_DISP)
    finally:
        set_gather_lanes(engine, GATHER_AUTO)
    od = oracle.hippo_terms(sd, want_grad=True)
    assert abs(d["edsp"] - od["edsp"]) <= REL * abs(od["edsp"])
    assert rel(d["dedsp"], od["dedsp"]) <= REL


@pytest.mark.parametrize("grids", ["fp32", "fp64"])
def test_pinned_grid_precision(engine, grids):
    """FFT_GRIDS_FP32 / FFT_GRIDS_FP64 or'ed into the engine fix the mixed
    form's grid precision; either stays within the mixed form's bar."""
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import FFT_DIRECT, FFT_GRIDS_FP32, FFT_GRIDS_FP64, PmqError, fft_choice, set_fft

    print(f"[gpu] pinned grids {grids}", flush=True)
    s = box216()
    ref = run(engine, s, FFT_DIRECT, L.PLAIN_FP64)
    engine.set_system(s)
    set_fft(engine, FFT_DIRECT | (FFT_GRIDS_FP32 if grids == "fp32" else FFT_GRIDS_FP64))
    try:
        m = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR, L.MIXED)
        assert fft_choice(engine)["mixed_grids"] == grids
    finally:
        set_fft(engine)
    for k in ("em", "ep"):
        assert abs(m[k] - ref[k]) <= 1e-5 * abs(ref[k]), (k, m[k], ref[k])
    with pytest.raises(PmqError):
        set_fft(engine, FFT_DIRECT | FFT_GRIDS_FP32 | FFT_GRIDS_FP64)
