"""The mixed-precision form against the plain FP64 form on the GPU.

Mixed: FP32 pair terms (displacements still taken in FP64), the dipole-field
tensor stored in FP32, 64-bit fixed-point accumulation of fields, gradients
and torques, FP64 PCG reductions and energy totals.  Bars: RMS force error
below 1e-3 kcal/mol/A against the FP64 result on the same input, energy
error below 1e-5 kcal/mol per atom (the FP32 rounding of each pair energy
adds up like a random walk over the pairs, largest for the unscreened
Coulomb sums of a cluster), and gradients bit-identical from one run to
the next (the fixed-point sums do not depend on the order of the atomics).
"""
import numpy as np
import pytest

from polmdqc_amd import water

pytestmark = pytest.mark.gpu
RMS_F = 1e-3
E_PER_ATOM = 1e-5


@pytest.fixture(scope="module")
def engine():
    from polmdqc_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def box216():
    xyz = water.lattice_box(6, 18.643, 216)
    return water.water_system(xyz, "water03", box=18.643,
                              ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})


def rms(a, b):
    return float(np.sqrt(np.mean(np.sum((np.asarray(a) - np.asarray(b)) ** 2, axis=1))))


@pytest.mark.parametrize("periodic", [True, False])
def test_mixed_matches_fp64(engine, periodic):
    from polmdqc_amd import _lib as L

    print(f"[gpu] mixed vs fp64 periodic={periodic}", flush=True)
    s = box216() if periodic else water.water_system(water.cluster(40, np.random.default_rng(4)), "water03")
    engine.set_system(s)
    terms = L.TERM_MPOLE | L.TERM_POLAR
    d = engine.eval(L.GRADIENT, terms, L.PLAIN_FP64)
    m = engine.eval(L.GRADIENT, terms, L.MIXED)
    n = s["n"]
    assert abs(m["em"] - d["em"]) <= E_PER_ATOM * n
    assert abs(m["ep"] - d["ep"]) <= E_PER_ATOM * n
    assert rms(m["desum"], d["desum"]) < RMS_F
    assert rms(m["dem"], d["dem"]) < RMS_F
    assert abs(m["pcg_iters"] - d["pcg_iters"]) <= 1
    e = engine.eval(L.ENERGY, terms, L.MIXED)
    assert abs(e["em"] - d["em"]) <= E_PER_ATOM * n


def test_mixed_gradient_is_bit_reproducible(engine):
    from polmdqc_amd import _lib as L

    engine.set_system(box216())
    terms = L.TERM_MPOLE | L.TERM_POLAR
    a = engine.eval(L.GRADIENT, terms, L.MIXED)
    b = engine.eval(L.GRADIENT, terms, L.MIXED)
    assert np.array_equal(a["desum"], b["desum"])
    assert np.array_equal(a["uind"], b["uind"])


def test_mixed_analysis_equals_fp64_analysis(engine):
    """ANALYSIS in the mixed form: the same interaction counts as FP64 (an
    FP32 pair energy is zero only where the FP64 one is), the per-atom
    partitions and einter to 1e-5 of |em| and |ep|, and the partition adds up
    to em to 1e-7 relative."""
    from polmdqc_amd import _lib as L

    engine.set_system(box216())
    terms = L.TERM_MPOLE | L.TERM_POLAR
    d = engine.eval(L.ANALYSIS, terms, L.PLAIN_FP64)
    m = engine.eval(L.ANALYSIS, terms, L.MIXED)
    assert m["nem"] == d["nem"] and m["nep"] == d["nep"]
    assert np.max(np.abs(m["aem"] - d["aem"])) <= 1e-5 * abs(d["em"])
    assert np.max(np.abs(m["aep"] - d["aep"])) <= 1e-5 * abs(d["ep"])
    assert abs(m["einter"] - d["einter"]) <= 1e-5 * abs(d["em"])
    # the partition sums FP32 pair halves in fixed point, em the FP32 pair energies in FP64
    assert abs(m["aem"].sum() - m["em"]) <= 1e-7 * abs(m["em"])
