"""The oracle's charge flux beside Ewald summation and polarization
(oracle/pmq_oracle_elec.c orc_elec), pinned by derivatives of its own
energy.

With charge flux the monopoles depend on the bond lengths and angles
(alterchg.cpp:36-169), so the gradient of the multipole and polarization
energies carries the chain term sum_i (dE/dq_i) dq_i/dx (dcflux.cpp:31-225);
dE/dq_i collects the real-space pairs (charge-penetration parts linear in
the valence charge), the reciprocal potential and the self energy with
Ewald, and for polarization the potential of the induced dipoles at i
(the variational energy, so no response term).  Pinned by central
differences of the energy (step 1e-5 A, PCG to 1e-12 D) on the HIPPO water
dimer (water21: Gordon1 penetration and charge flux, no Ewald) and a
64-water water21 box with Ewald (7 A cutoff), and the box's virial by the
strain derivative of the energy.  Dropping any one of the pot
contributions (real, reciprocal, self, polarization) moves the gradient by
far more than the bars; the last test shows the size of the whole chain
term by holding the altered charges fixed.
"""
import os

import numpy as np

import oracle
from conftest import GOLDEN
from polmdqc_amd import tinker
from test_oracle_virial import strain_derivative

TD = os.path.join(GOLDEN, "tinker")
PRM = [os.path.join(TD, "params")]


def dimer():
    s = tinker.system_from_files(os.path.join(TD, "hippo/water21.xyz"), os.path.join(TD, "hippo/water21.key"), PRM)
    s["use_polar"] = 1
    s["poleps"] = 1e-12
    return s


def water21_box(tmp_path, nmol=64):
    """nmol waters of the 30 A box around its first one in a 24 A cell,
    water21 with Ewald (7 A)."""
    src = open(os.path.join(TD, "bounds/waterbox30.xyz")).read().splitlines()[1:]
    xyz = np.array([[float(v) for v in l.split()[2:5]] for l in src])
    d = xyz - xyz[0]
    mol = np.arange(len(xyz)) // 3
    keep = sorted({int(m) for m in mol[np.max(np.abs(d), axis=1) < 12.0]})[:nmol]
    rows = []
    for q, m in enumerate(keep):
        for a in range(3):
            w = src[3 * m + a].split()
            bonds = [3 * q + 2, 3 * q + 3] if a == 0 else [3 * q + 1]
            p = xyz[3 * m + a] - xyz[0]
            rows.append(f"{3 * q + a + 1:6d}  {w[1]:<3}{p[0]:12.6f}{p[1]:12.6f}{p[2]:12.6f}{int(w[5]):6d}"
                        + "".join(f"{b:6d}" for b in bonds))
    (tmp_path / "cf.xyz").write_text(f"{len(rows):6d}  box\n" + "\n".join(rows) + "\n")
    (tmp_path / "cf.key").write_text("parameters water21\na-axis 24.0\newald\newald-cutoff 7.0\n")
    s = tinker.system_from_files(str(tmp_path / "cf.xyz"), str(tmp_path / "cf.key"), PRM)
    s["use_polar"] = 1
    s["poleps"] = 1e-12
    return s


def energies(s, xyz=None):
    t = dict(s)
    if xyz is not None:
        t["x"], t["y"], t["z"] = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    o = oracle.elec(t, want_grad=False)
    return o["em"], o["ep"]


def fd_check(s, picks, h=1e-5):
    o = oracle.elec(s, want_grad=True)
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    for i, a in picks:
        ep, em = xyz.copy(), xyz.copy()
        ep[i, a] += h
        em[i, a] -= h
        (mp_, pp_), (mm_, pm_) = energies(s, ep), energies(s, em)
        for name, fd, g, scale in (("em", (mp_ - mm_) / (2 * h), o["dem"][i, a], np.abs(o["dem"]).max()),
                                   ("ep", (pp_ - pm_) / (2 * h), o["dep"][i, a], np.abs(o["dep"]).max())):
            assert abs(fd - g) <= 1e-6 * scale, (name, i, a, fd, g)
    return o


def test_dimer_gradient_with_charge_flux_and_polarization():
    s = dimer()
    assert s["use_chgflx"] and s["pentype"] == 1 and not s.get("use_ewald", 0)
    fd_check(s, [(i, a) for i in range(s["n"]) for a in range(3)])


def test_ewald_box_gradient_and_virial_with_charge_flux(tmp_path):
    s = water21_box(tmp_path)
    assert s["use_chgflx"] and s["use_ewald"]
    rng = np.random.default_rng(21)
    picks = [(int(rng.integers(s["n"])), int(rng.integers(3))) for _ in range(10)]
    o = fd_check(s, picks)
    for key, k in (("vir_m", 0), ("vir_p", 1)):
        W = strain_derivative(lambda t: energies(t)[k], s)
        assert np.max(np.abs(o[key] - W)) <= 1e-6 * np.max(np.abs(W)), (key, o[key], W)


def test_the_flux_term_is_large_against_the_bars(tmp_path):
    """The gradient with the charges' dependence on the geometry left out
    (the same altered charges, held fixed) misses the full gradient by far
    more than the bars above, in both terms."""
    s = water21_box(tmp_path, 16)
    full = oracle.elec(s, want_grad=True)
    pdelta = oracle.charge_flux_delta(s)
    frozen = dict(s)
    frozen["use_chgflx"] = 0
    pole = np.asarray(s["pole"], dtype=float).reshape(-1, 13).copy()
    pole[:, 0] = np.asarray(s["mono0"]) + pdelta
    frozen["pole"] = pole.ravel()
    frozen["pval0"] = np.asarray(s["pval0"]) + pdelta
    g0 = oracle.elec(frozen, want_grad=True)
    assert abs(g0["em"] - full["em"]) <= 1e-12 * abs(full["em"])
    assert abs(g0["ep"] - full["ep"]) <= 1e-10 * abs(full["ep"])
    for k in ("dem", "dep"):
        assert np.max(np.abs(full[k] - g0[k])) > 1e-3 * np.max(np.abs(full[k])), k
