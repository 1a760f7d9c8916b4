"""The oracle's virial, pinned by the derivative of its own energy under
homogeneous strain.

W_ab = dE/d(eps_ab) with every position and the periodic box strained,
x -> (1 + eps) x (the reference's convention, empole.cpp:310-320 and
torque.cpp:361-395, W = sum_i x_i (x) dE/dx_i for a cluster).  A symmetric
strain with eps_ab = eps_ba = d changes the energy by d (W_ab + W_ba); the
strained cell is rotated back to Tinker's triclinic convention (a along x,
b in the xy plane), which leaves the energy unchanged.  Central differences
with d = 1e-5; the real-space cutoffs are 9 A so that no pair energy jumps
at the cutoff within the step (erfc(0.5446 x 9) = 4e-12), the polarization is
converged to 1e-12 D, and dispersion and repulsion are smooth at theirs
(Ewald damping, quintic taper).  A wrong sign of the k-space term, a
dropped dipole/quadrupole strain term or a missing frame (torque) virial
fails these by far more than the 1e-6 relative bar.  For a cluster the
virial also equals sum_i x_i (x) g_i.
"""
import numpy as np
import pytest

import oracle
from polmdqc_amd import water

STEP = 1e-5
PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


def lattice(s):
    """Columns a, b, c of Tinker's cell (lattice.cpp) from lengths and angles."""
    A, B, C = (float(s[k]) for k in ("xbox", "ybox", "zbox"))
    al, be, ga = (np.deg2rad(float(s[k])) for k in ("alpha", "beta", "gamma"))
    cx = np.cos(be)
    cy = (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    cz = np.sqrt(1.0 - cx * cx - cy * cy)
    return np.array([[A, B * np.cos(ga), C * cx], [0.0, B * np.sin(ga), C * cy], [0.0, 0.0, C * cz]])


def strained(s, eps):
    """The system under strain eps (3x3), cell rotated to Tinker's convention."""
    t = dict(s)
    F = np.eye(3) + eps
    xyz = np.stack([s["x"], s["y"], s["z"]], 1) @ F.T
    if int(s.get("use_bounds", 0)):
        M = F @ lattice(s)
        a, b, c = M[:, 0], M[:, 1], M[:, 2]
        ln = [np.linalg.norm(v) for v in (a, b, c)]
        ang = lambda u, v: np.rad2deg(np.arccos(u @ v / (np.linalg.norm(u) * np.linalg.norm(v))))
        t.update({"xbox": ln[0], "ybox": ln[1], "zbox": ln[2], "alpha": ang(b, c), "beta": ang(a, c),
                  "gamma": ang(a, b), "triclinic": 1, "monoclinic": 0, "orthogonal": 0})
        R = lattice(t) @ np.linalg.inv(M)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and np.linalg.det(R) > 0
        xyz = xyz @ R.T
    t["x"], t["y"], t["z"] = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    return t


def strain_derivative(energy, s):
    """W from central differences of energy(system) under six symmetric strains."""
    W = np.zeros((3, 3))
    for a, b in PAIRS:
        e = np.zeros((3, 3))
        e[a, b] = e[b, a] = 1.0
        d = (energy(strained(s, STEP * e)) - energy(strained(s, -STEP * e))) / (2 * STEP)
        if a == b:
            W[a, a] = d
        else:
            W[a, b] = W[b, a] = 0.5 * d
    return W


def amoeba_box(cut=9.0):
    s = water.water_system(water.lattice_box(6, 18.643, 216), "water03", box=18.643,
                           ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": cut})
    s["poleps"] = 1e-12
    return s


def hippo_box(cut=9.0):
    s = water.water_system(water.lattice_box(6, 18.643, 216), "hippo19", box=18.643,
                           ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": cut, "dewald": 0.45})
    s["poleps"] = 1e-12
    s["dispcut"] = cut
    return s


def cluster(ff):
    s = water.water_system(water.cluster(12, np.random.default_rng(12)), ff)
    s["poleps"] = 1e-12
    if ff == "hippo19":
        s["dispcut"], s["repcut"] = 8.0, 5.0
    return s


def close(W, ref, tol=1e-6):
    assert np.max(np.abs(W - ref)) <= tol * np.max(np.abs(ref)), (W, ref)


@pytest.mark.parametrize("name", ["amoeba_box", "amoeba_cluster", "hippo_box", "hippo_cluster"])
def test_electrostatic_virial_matches_strain_derivative(name):
    s = {"amoeba_box": amoeba_box, "hippo_box": hippo_box}.get(name, lambda: None)()
    if s is None:
        s = cluster("water03" if name.startswith("amoeba") else "hippo19")
    o = oracle.elec(s, want_grad=True)
    Wm = strain_derivative(lambda t: oracle.elec(t, want_grad=False, polar=False)["em"], s)
    close(o["vir_m"], Wm)
    Wp = strain_derivative(lambda t: oracle.elec(t, want_grad=False)["ep"], s)
    close(o["vir_p"], Wp)
    if not int(s["use_bounds"]):
        # a cluster: W = sum_i x_i (x) g_i, symmetric
        xyz = np.stack([s["x"], s["y"], s["z"]], 1)
        for g, W in ((o["dem"], o["vir_m"]), (o["dep"], o["vir_p"])):
            X = xyz.T @ g
            close(W, 0.5 * (X + X.T), 1e-10)


@pytest.mark.parametrize("periodic", [True, False])
def test_dispersion_and_repulsion_virial_match_strain_derivative(periodic):
    s = hippo_box() if periodic else cluster("hippo19")
    o = oracle.hippo_terms(s, want_grad=True)
    close(o["vir_d"], strain_derivative(lambda t: oracle.hippo_terms(t, want_grad=False)["edsp"], s))
    close(o["vir_r"], strain_derivative(lambda t: oracle.hippo_terms(t, want_grad=False)["er"], s))
