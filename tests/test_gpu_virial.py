"""The virial of every term on the GPU against the oracle (FP64 bar 1e-9
relative to the largest component; the oracle's virial is pinned by strain
derivatives of its energy in test_oracle_virial.py).

Covered: Ewald multipoles alone, polarization alone and both (AMOEBA, the
216-water box of the 12,288-atom workload's geometry at 24^3 and the
12,288-atom box itself at 48^3), a triclinic cell, a non-periodic cluster
(where W = sum_i x_i (x) g_i also holds), the HIPPO terms (charge
penetration, dispersion PME, repulsion), and the mixed form against FP64
(the mixed form's virial is accumulated in FP64 from FP32 pair terms; bar
1e-5 relative for AMOEBA, measured 1.1e-7 at 12,288 atoms; 1e-4 for HIPPO,
measured 1.6e-5, whose charge-penetration pairs carry the larger FP32 force
error, RMS 2e-4 kcal/mol/A against RMS forces near 10).
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


def vrel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def box216(ff="water03"):
    xyz = water.lattice_box(6, 18.643, 216)
    ew = {"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0}
    if ff == "hippo19":
        ew["dewald"] = 0.45
    s = dict(water.water_system(xyz, ff, box=18.643, ewald=ew))
    s["poleps"] = 1e-10
    return s


@pytest.mark.parametrize("terms", ["mpole", "polar", "both"])
def test_amoeba_box_virial_matches_oracle(engine, terms):
    from polmdqc_amd import _lib as L

    print(f"[gpu] virial amoeba 648 {terms}", flush=True)
    s = box216()
    t = {"mpole": L.TERM_MPOLE, "polar": L.TERM_POLAR, "both": L.TERM_MPOLE | L.TERM_POLAR}[terms]
    engine.set_system(s)
    g = engine.eval(L.VIRIAL, t)
    o = oracle.elec(s, want_grad=True)
    ref = (o["vir_m"] if terms != "polar" else 0) + (o["vir_p"] if terms != "mpole" else 0)
    print(f"  rel {vrel(g['vir'], ref):.2e}", flush=True)
    assert vrel(g["vir"], ref) <= REL
    # the gradient of a VIRIAL evaluation is the GRADIENT one's
    d = engine.eval(L.GRADIENT, t)
    assert np.max(np.abs(g["desum"] - d["desum"])) <= 1e-12 * np.max(np.abs(d["desum"]))


def test_triclinic_and_cluster_virial_match_oracle(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] virial triclinic box and cluster", flush=True)
    s = box216()
    s["triclinic"], s["orthogonal"] = 1, 0
    s["alpha"], s["beta"], s["gamma"] = 88.0, 92.0, 89.0
    c = water.water_system(water.cluster(40, np.random.default_rng(4)), "water03")
    c["poleps"] = 1e-10
    for sysd in (s, c):
        engine.set_system(sysd)
        g = engine.eval(L.VIRIAL, L.TERM_MPOLE | L.TERM_POLAR)
        o = oracle.elec(sysd, want_grad=True)
        assert vrel(g["vir"], o["vir_m"] + o["vir_p"]) <= REL
    xyz = np.stack([c["x"], c["y"], c["z"]], 1)
    X = xyz.T @ g["desum"]
    assert vrel(g["vir"], 0.5 * (X + X.T)) <= 1e-10


def test_hippo_virial_matches_oracle(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] virial hippo 648 all terms", flush=True)
    s = box216("hippo19")
    engine.set_system(s)
    o = oracle.elec(s, want_grad=True)
    h = oracle.hippo_terms(s, want_grad=True)
    for t, ref in ((L.TERM_MPOLE | L.TERM_POLAR, o["vir_m"] + o["vir_p"]), (L.TERM_REPEL, h["vir_r"]),
                   (L.TERM_DISP, h["vir_d"])):
        g = engine.eval(L.VIRIAL, t)
        print(f"  terms {t}: rel {vrel(g['vir'], ref):.2e}", flush=True)
        assert vrel(g["vir"], ref) <= REL, t
    g = engine.eval(L.VIRIAL, L.TERM_MPOLE | L.TERM_POLAR | L.TERM_REPEL | L.TERM_DISP)
    assert vrel(g["vir"], o["vir_m"] + o["vir_p"] + h["vir_r"] + h["vir_d"]) <= REL


def test_amoeba12k_virial_fp64_and_mixed(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] virial amoeba 12288 fp64 vs oracle, mixed vs fp64", flush=True)
    xyz = water.lattice_box(16, 49.68, 4096)
    s = dict(water.water_system(xyz, "water03", box=49.68,
                                ewald={"aewald": 0.5446, "nfft": 48, "order": 5, "cutoff": 7.0}))
    s["poleps"] = 1e-10
    engine.set_system(s)
    terms = L.TERM_MPOLE | L.TERM_POLAR
    d = engine.eval(L.VIRIAL, terms, L.PLAIN_FP64)
    o = oracle.elec(s, want_grad=True)
    print(f"  fp64 rel {vrel(d['vir'], o['vir_m'] + o['vir_p']):.2e}", flush=True)
    assert vrel(d["vir"], o["vir_m"] + o["vir_p"]) <= REL
    m = engine.eval(L.VIRIAL, terms, L.MIXED)
    print(f"  mixed rel {vrel(m['vir'], d['vir']):.2e}", flush=True)
    assert vrel(m["vir"], d["vir"]) <= 1e-5


def test_hippo_mixed_virial_against_fp64(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] virial hippo mixed vs fp64", flush=True)
    s = box216("hippo19")
    engine.set_system(s)
    terms = L.TERM_MPOLE | L.TERM_POLAR | L.TERM_REPEL | L.TERM_DISP
    d = engine.eval(L.VIRIAL, terms, L.PLAIN_FP64)
    m = engine.eval(L.VIRIAL, terms, L.MIXED)
    print(f"  mixed rel {vrel(m['vir'], d['vir']):.2e}", flush=True)
    assert vrel(m["vir"], d["vir"]) <= 1e-4
