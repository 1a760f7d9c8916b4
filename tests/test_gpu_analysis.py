"""Analysis (nem, aem, einter; nep, aep) on the Ewald and polarization
paths on the GPU against the oracle (tests/test_oracle_analysis.py pins the
oracle's definitions).  Bars: nem and nep exact; energies, einter and the
per-atom partitions to 1e-9 relative (FP64), to 1e-5 of |em| in the mixed
form.

Cases: the 648-atom AMOEBA water box (PME 24^3); the alanine tetrapeptide
in a periodic box (intramolecular pairs beyond 1-4 are unscaled, so einter
differs from em); the HIPPO water box (charge penetration, Ewald); the
induced-dipole energy partition with and without Ewald.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from polmdqc_amd import tinker, water

pytestmark = pytest.mark.gpu
REL = 1e-9


@pytest.fixture(scope="module")
def engine():
    from polmdqc_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def amoeba_box():
    return water.water_system(water.lattice_box(6, 18.643, 216), "water03", box=18.643,
                              ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})


def hippo_box():
    return water.water_system(water.lattice_box(6, 18.643, 216), "hippo19", box=18.643,
                              ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})


def alatet_box():
    from test_oracle_analysis import in_box

    s = tinker.system_from_files(os.path.join(GOLDEN, "tinker/chkpole/alatet.xyz"), None,
                                 [os.path.join(GOLDEN, "tinker/params")])
    return in_box(s, 30.0, 9.0, aewald=0.4, grid=32)


def close(a, b, scale):
    assert abs(a - b) <= REL * max(1.0, abs(scale)), (a, b)


@pytest.mark.parametrize("name", ["amoeba_box", "alatet_box", "hippo_box"])
def test_ewald_multipole_analysis(engine, name):
    from polmdqc_amd import _lib as L

    print(f"[gpu] analysis {name}", flush=True)
    s = {"amoeba_box": amoeba_box, "alatet_box": alatet_box, "hippo_box": hippo_box}[name]()
    engine.set_system(s)
    g = engine.eval(L.ANALYSIS, L.TERM_MPOLE, L.PLAIN_FP64)
    o = oracle.elec(s, want_grad=False, polar=False)
    assert g["nem"] == o["nem"]
    close(g["em"], o["em"], o["em"])
    close(g["einter"], o["einter"], o["em"])
    assert np.max(np.abs(g["aem"] - o["aem"])) <= REL * max(1.0, abs(o["em"]))
    assert abs(g["aem"].sum() - g["em"]) <= 1e-10 * abs(g["em"])
    if name == "alatet_box":
        assert abs(g["em"] - g["einter"]) > 1.0  # the intramolecular part is not empty
    # the mixed form: the same counts, partitions and einter to 1e-5 of |em|
    m = engine.eval(L.ANALYSIS, L.TERM_MPOLE, L.MIXED)
    assert m["nem"] == o["nem"]
    assert np.max(np.abs(m["aem"] - o["aem"])) <= 1e-5 * abs(o["em"])
    assert abs(m["einter"] - o["einter"]) <= 1e-5 * abs(o["em"])


@pytest.mark.parametrize("periodic", [True, False])
def test_polarization_analysis(engine, periodic):
    from polmdqc_amd import _lib as L

    print(f"[gpu] polarization analysis periodic={periodic}", flush=True)
    s = amoeba_box() if periodic else water.water_system(water.cluster(30, np.random.default_rng(3)), "water03")
    s["poleps"] = 1e-10
    engine.set_system(s)
    g = engine.eval(L.ANALYSIS, L.TERM_MPOLE | L.TERM_POLAR, L.PLAIN_FP64)
    o = oracle.elec(s, want_grad=False, poleps=1e-10)
    assert g["nep"] == o["nep"] == int(s["n"])
    close(g["ep"], o["ep"], o["ep"])
    assert np.max(np.abs(g["aep"] - o["aep"])) <= 1e-8 * max(1.0, np.max(np.abs(o["aep"])))
    assert abs(g["aep"].sum() - g["ep"]) <= 1e-10 * abs(g["ep"])
    assert g["nem"] == o["nem"]
    assert np.max(np.abs(g["aem"] - o["aem"])) <= REL * max(1.0, abs(o["em"]))
