"""Pins the FP64 oracle's multipole path (CPU only, no GPU).

* against the golden vectors of the reference's own tests (test/*.h), with the
  reference's own tolerances (absolute, as include/testrt.h:8-19 compares);
* against the reference's energy<Analysis>/energy<Virial> outputs dumped by
  the compiled reference (virial included, which no reference test covers);
* against what the mathematics fixes: central finite differences of the
  energy, translational and rotational invariance, and the pair energy's
  symmetry under exchanging the two sites.
"""
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_fixture

HDR = json.load(open(os.path.join(GOLDEN, "ref_headers.json")))

# reference test namespace -> fixture system (test/*.cpp argv)
CASES = {
    "amoeba1": "amoeba_water09", "amoeba2": "amoeba_water09",
    "amoebap1": "amoebap_waterap", "amoebap2": "amoebap_waterap",
    "hippo1": "hippo_water21", "hippo2": "hippo_water21",
    "rotpole1": "rotpole_water09_Na", "rotpole4": "rotpole_water09_Na",
    "rotpole2": "rotpole_benzene_ethyne_water", "rotpole5": "rotpole_benzene_ethyne_water",
    "rotpole3": "rotpole_ammonia", "rotpole6": "rotpole_ammonia",
    "chkpole1": "chkpole_alatet", "chkpole2": "chkpole_alatet",
    # amoebaMisc-2/3 read a two-frame archive (alatet, then the water dimer);
    # the goldens hold the last frame's results
    "amoebaMisc1": "misc_water09_2", "amoebaMisc2": "amoeba_water09", "amoebaMisc3": "amoeba_water09",
    "amoebaMisc5": "misc_water09NSeq",
    "active1": "active_water09_Na_Cls_active", "active2": "active_water09_Na_Cls_inactive",
    "active3": "active_waterbox30",
    "bounds1": "bounds_orthogonal", "bounds2": "bounds_monoclinic", "bounds3": "bounds_triclinic",
    "bounds4": "bounds_octahedron", "bounds5": "bounds_dodecahedron",
}

_cache = {}


def run(system, mode):
    key = (system, mode)
    if key not in _cache:
        _cache[key] = oracle.empole(load_fixture(system), mode)
    return _cache[key]


@pytest.mark.parametrize("ns", sorted(CASES))
def test_reference_golden_vectors(ns):
    g = HDR[ns]
    system = CASES[ns]
    if "em" in g:
        a = run(system, oracle.ANALYSIS)
        eps = g["eps"]
        assert a["nem"] == int(g["nem"])
        assert abs(a["em"] - g["em"]) <= eps
        assert abs(a["einter"] - g["einter"]) <= eps
        if "aem" in g:
            np.testing.assert_allclose(a["aem"], g["aem"], rtol=0, atol=eps)
        if "rpole" in g:
            np.testing.assert_allclose(a["rpole"], np.array(g["rpole"]), rtol=0, atol=g["epsR"])
    if "dem" in g:
        v = run(system, oracle.GRADIENT)
        np.testing.assert_allclose(v["dem"].ravel(), g["dem"], rtol=0, atol=g["eps1"])


SYSTEMS = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("system", SYSTEMS)
def test_reference_dump(system):
    """Every output of the reference's energy<Analysis>/<Virial> on its own inputs."""
    fx = load_fixture(system)
    a = run(system, oracle.ANALYSIS)
    v = run(system, oracle.VIRIAL)
    # the reference is compiled with -ffast-math: its sums are reassociated,
    # which moves energies with strong charge-penetration cancellation by a
    # few 1e-12 relative
    scale = max(1.0, abs(float(fx["em"])))
    assert a["nem"] == int(fx["nem"])
    assert abs(a["em"] - fx["em_analysis"]) <= 1e-11 * scale
    assert abs(v["em"] - fx["em"]) <= 1e-11 * scale
    assert abs(a["einter"] - fx["einter"]) <= 1e-11 * scale
    sa, ga = fx["sample_atoms"], fx["grad_atoms"]
    np.testing.assert_allclose(a["aem"][sa], fx["aem"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(a["rpole"][sa].ravel(), fx["rpole"], rtol=0, atol=1e-15)
    gscale = max(1.0, np.abs(fx["dem"]).max())
    np.testing.assert_allclose(v["dem"][ga].ravel(), fx["dem"], rtol=0, atol=1e-11 * gscale)
    vscale = max(1.0, np.abs(fx["vir"]).max())
    np.testing.assert_allclose(v["vir"].ravel(), fx["vir"], rtol=0, atol=1e-11 * vscale)


# absolute tolerances of the reference's own numerical-gradient checks (eps2)
FD_TOL = {"amoeba_water09": 1e-6, "amoebap_waterap": 1e-6, "hippo_water21": 1e-6,
          "rotpole_benzene_ethyne_water": 1e-6, "rotpole_ammonia": 1e-4, "chkpole_alatet": 1e-7}


@pytest.mark.parametrize("system", sorted(FD_TOL))
def test_gradient_matches_finite_difference(system):
    """Central differences of the energy (step 1e-5 A, as testgrad.cpp uses),
    with the reference's own absolute tolerances."""
    fx = load_fixture(system)
    n = int(fx["n"])
    x0 = np.stack([fx["x"], fx["y"], fx["z"]], 1)
    g = oracle.empole(fx, oracle.GRADIENT, x=x0)["dem"]
    h = 1e-5
    rng = np.random.default_rng(7)
    for i in rng.choice(n, size=min(n, 6), replace=False):
        for a in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[i, a] += h
            xm[i, a] -= h
            num = (oracle.empole(fx, oracle.ENERGY, x=xp)["em"] - oracle.empole(fx, oracle.ENERGY, x=xm)["em"]) / (2 * h)
            assert abs(num - g[i, a]) < FD_TOL[system]


@pytest.mark.parametrize("system", ["amoeba_water09", "rotpole_benzene_ethyne_water", "chkpole_alatet"])
def test_translation_and_rotation_invariance(system):
    """A cluster without a box: forces sum to zero; energy is invariant under a
    rigid rotation, and the rotated gradient is the rotated gradient."""
    fx = load_fixture(system)
    x0 = np.stack([fx["x"], fx["y"], fx["z"]], 1)
    r0 = oracle.empole(fx, oracle.GRADIENT, x=x0)
    np.testing.assert_allclose(r0["dem"].sum(0), 0.0, atol=1e-10)
    th = 0.7
    rot = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    rot = rot @ np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    r1 = oracle.empole(fx, oracle.GRADIENT, x=x0 @ rot.T + np.array([1.5, -2.0, 0.25]))
    assert abs(r1["em"] - r0["em"]) < 1e-11 * max(1, abs(r0["em"]))
    np.testing.assert_allclose(r1["dem"], r0["dem"] @ rot.T, atol=1e-10)


def test_pair_symmetry_and_torque_balance():
    """Exchanging the sites flips the force; total torque about the origin
    (site torques plus r x F) vanishes for any isolated pair."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        mi, mk = rng.normal(size=10), rng.normal(size=10)
        for m in (mi, mk):  # traceless quadrupoles
            tr = (m[4] + m[7] + m[9]) / 3
            m[4] -= tr; m[7] -= tr; m[9] -= tr
        r = rng.normal(size=3) * 2.0
        a = oracle.pair_mpole(r, 1.0, 332.0, mi, mk)
        b = oracle.pair_mpole(-r, 1.0, 332.0, mk, mi)
        assert abs(a[0] - b[0]) < 1e-12 * max(1, abs(a[0]))
        np.testing.assert_allclose(a[1:4], -b[1:4], atol=1e-11)
        # site k sits at r, site i at the origin: total torque = ti + tk + r x F_k
        fk = -a[1:4]  # gradient on k is -frc; force is minus gradient
        total = a[4:7] + a[7:10] + np.cross(r, -fk)
        np.testing.assert_allclose(total, 0.0, atol=1e-10)
