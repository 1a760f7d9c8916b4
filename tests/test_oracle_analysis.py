"""Pins the oracle's analysis of the Ewald and polarization paths (nem,
aem, einter, aep, nep), which no reference test covers (the reference has
no Ewald energy and no polarization energy).

* without Ewald, orc_elec's analysis equals the established multipole
  analysis (orc_empole, pinned by the reference's goldens) on the alanine
  tetrapeptide: nem, einter, aem;
* with Ewald, aem sums to em (the per-site reciprocal and self terms are
  not dropped), nem is the number of real-space pairs within the cutoff
  plus one self term per site (counted here by brute force), and
  em - einter, the molecule's own unscreened scaled energy, equals the
  energy of the same molecule computed without Ewald and without periodic
  images once the cutoff covers the molecule;
* the per-site polarization energy of two polarizable sites equals the
  closed form -(f/2) mu_i . E_i of the 2x2 solve.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_fixture
from polmdqc_amd import tinker
from test_oracle_damping import thole_model, two_sites

TD = os.path.join(GOLDEN, "tinker")


def alatet():
    return tinker.system_from_files(os.path.join(TD, "chkpole/alatet.xyz"), None, [os.path.join(TD, "params")])


def in_box(s, L, cut, aewald=0.35, grid=40):
    t = dict(s)
    c = np.stack([s["x"], s["y"], s["z"]], 1)
    c = c - c.mean(0) + L / 2
    t["x"], t["y"], t["z"] = (np.ascontiguousarray(c[:, q]) for q in range(3))
    t.update({"use_bounds": 1, "xbox": L, "ybox": L, "zbox": L, "alpha": 90.0, "beta": 90.0, "gamma": 90.0,
              "orthogonal": 1, "monoclinic": 0, "triclinic": 0, "octahedron": 0, "dodecadron": 0,
              "use_ewald": 1, "aewald": aewald, "ewaldcut": cut, "nfft1": grid, "nfft2": grid, "nfft3": grid,
              "bsorder": 5})
    return t


def test_without_ewald_equals_the_multipole_analysis():
    fx = load_fixture("misc_alatet_water09")
    a = oracle.elec(fx, want_grad=False, polar=False)
    b = oracle.empole(fx, oracle.ANALYSIS)
    assert a["nem"] == b["nem"] == int(fx["nem"])
    assert abs(a["einter"] - b["einter"]) <= 1e-12 * max(1.0, abs(b["em"]))
    assert np.max(np.abs(a["aem"] - b["aem"])) <= 1e-12 * max(1.0, abs(b["em"]))


def test_ewald_partition_counts_and_intramolecular_energy():
    s = alatet()
    iso = oracle.empole(s, oracle.ENERGY)["em"]  # no box, no cutoff
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    cut = np.sqrt(np.max(np.sum((xyz[:, None] - xyz[None]) ** 2, axis=2))) + 1.0  # the molecule's diameter + 1
    t = in_box(s, 3.0 * cut, cut)
    o = oracle.elec(t, want_grad=False, polar=False)
    assert abs(o["aem"].sum() - o["em"]) <= 1e-10 * abs(o["em"])
    # every pair lies within the cutoff (one image each), plus 42 self terms
    n = int(s["n"])
    assert o["nem"] == n * (n - 1) // 2 + n
    # em - einter is the molecule's own energy (unscreened, scaled, all pairs)
    assert abs((o["em"] - o["einter"]) - iso) <= 1e-10 * abs(iso)
    # the intermolecular remainder is the small image interaction
    assert abs(o["einter"]) < 0.05 * abs(iso)
    # a mistake that dropped the self or reciprocal terms from aem would not pass
    assert abs(o["em_self"]) > 1.0 and abs(o["em_recip"]) > 1e-3


def test_ewald_pair_count_in_a_water_box():
    from polmdqc_amd import water

    s = water.water_system(water.lattice_box(6, 18.643, 216), "water03", box=18.643,
                           ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})
    o = oracle.elec(s, want_grad=False, polar=False)
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    d = xyz[:, None, :] - xyz[None, :, :]
    d -= 18.643 * np.round(d / 18.643)
    r2 = np.sum(d * d, axis=2)
    iu = np.triu_indices(len(xyz), 1)
    pairs = int(np.sum(r2[iu] <= 49.0))
    assert o["nem"] == pairs + len(xyz)
    assert abs(o["aem"].sum() - o["em"]) <= 1e-10 * abs(o["em"])
    assert o["einter"] == o["em"]  # water: every intramolecular pair is scaled to zero


@pytest.mark.parametrize("r", [1.6, 2.4])
def test_polarization_site_energies_closed_form(r):
    q, alpha, thole = np.array([0.7, -0.4]), np.array([0.9, 1.3]), np.array([0.39, 0.2])
    s = two_sites(r, q, alpha, thole)
    l3, l5 = thole_model(r, (alpha[0] * alpha[1]) ** (1 / 6), min(thole))[:2]
    EA, EB, tt = -q[1] * l3 / r ** 2, q[0] * l3 / r ** 2, (3 * l5 - l3) / r ** 3
    mu = np.linalg.solve(np.array([[1.0 / alpha[0], -tt], [-tt, 1.0 / alpha[1]]]), np.array([EA, EB]))
    f = float(s["electric"])
    o = oracle.elec(s, want_grad=False)
    np.testing.assert_allclose(o["aep"], [-0.5 * f * mu[0] * EA, -0.5 * f * mu[1] * EB], rtol=1e-10)
    assert o["nep"] == 2
