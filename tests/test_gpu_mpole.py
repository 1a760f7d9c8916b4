"""GPU parity of the permanent-multipole path (plain FP64) through the C ABI.

Each reference test system (and the two synthetic water boxes, 648 and
12,288 atoms, in non-Ewald form) is run on the GPU and compared with the reference's own
dumped outputs and with the oracle.  Bar: relative error <= 1e-9 on energy,
gradient and virial; nem exact; einter and the per-atom partition aem too.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_fixture

pytestmark = pytest.mark.gpu

SYSTEMS = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))
REL = 1e-9


@pytest.fixture(scope="module")
def engine():
    from polmdqc_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def rel_err(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("system", SYSTEMS)
def test_mpole_matches_reference(engine, system):
    print(f"[gpu] mpole parity {system}", flush=True)
    fx = load_fixture(system)
    engine.set_system(fx)
    from polmdqc_amd import _lib as L

    a = engine.eval(L.ANALYSIS, L.TERM_MPOLE, L.PLAIN_FP64, want_rpole=True)
    v = engine.eval(L.VIRIAL, L.TERM_MPOLE, L.PLAIN_FP64)
    em = float(fx["em"])
    assert a["nem"] == int(fx["nem"])
    assert abs(a["em"] - float(fx["em_analysis"])) <= REL * max(1.0, abs(em))
    assert abs(v["em"] - em) <= REL * max(1.0, abs(em))
    assert abs(a["einter"] - float(fx["einter"])) <= REL * max(1.0, abs(em))
    sa, ga = fx["sample_atoms"], fx["grad_atoms"]
    assert np.max(np.abs(a["aem"][sa] - fx["aem"])) <= REL * max(1.0, np.abs(fx["aem"]).max())
    assert np.max(np.abs(a["rpole"][sa].ravel() - fx["rpole"])) <= 1e-14
    assert rel_err(v["dem"][ga].ravel(), fx["dem"]) <= REL
    assert rel_err(v["vir"].ravel(), fx["vir"]) <= REL


@pytest.mark.parametrize("system", ["amoeba_water09", "hippo_water21", "amoebap_waterap", "chkpole_alatet"])
def test_mpole_matches_oracle_at_moved_coordinates(engine, system):
    """Same system, coordinates displaced at random: GPU vs oracle (no golden)."""
    fx = load_fixture(system)
    engine.set_system(fx)
    from polmdqc_amd import _lib as L

    rng = np.random.default_rng(11)
    x0 = np.stack([fx["x"], fx["y"], fx["z"]], 1)
    for trial in range(3):
        x = x0 + rng.normal(scale=0.05, size=x0.shape)
        g = engine.eval(L.VIRIAL, L.TERM_MPOLE, L.PLAIN_FP64, xyz=x)
        o = oracle.empole(fx, oracle.VIRIAL, x=x)
        assert abs(g["em"] - o["em"]) <= REL * max(1.0, abs(o["em"]))
        assert rel_err(g["dem"], o["dem"]) <= REL
        assert rel_err(g["vir"], o["vir"]) <= REL


def test_energy_mode_equals_gradient_mode_energy(engine):
    fx = load_fixture("synth_water216")
    engine.set_system(fx)
    from polmdqc_amd import _lib as L

    e = engine.eval(L.ENERGY)["em"]
    g = engine.eval(L.GRADIENT)["em"]
    assert abs(e - g) <= 1e-12 * abs(g)


def test_repeat_is_stable(engine):
    """Two evaluations of one configuration agree to FP64 atomics round-off."""
    fx = load_fixture("bounds_orthogonal")
    engine.set_system(fx)
    from polmdqc_amd import _lib as L

    a = engine.eval(L.VIRIAL)
    b = engine.eval(L.VIRIAL)
    assert abs(a["em"] - b["em"]) <= 1e-12 * abs(a["em"])
    assert rel_err(a["dem"], b["dem"]) <= 1e-12
