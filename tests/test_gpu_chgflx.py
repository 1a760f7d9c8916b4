"""Charge flux on the GPU beside Ewald summation and HIPPO polarization, in
both forms, against the oracle (oracle/pmq_oracle_elec.c orc_elec, whose
chain term is pinned by central differences in tests/test_oracle_chgflx.py).

Systems: the HIPPO water dimer (water21: Gordon1 penetration and charge
flux, no Ewald) and a 64-water water21 box with Ewald (7 A), polarization
on, PCG to 1e-10 D so both solvers sit on the same dipoles.  FP64: energies
to 1e-9 relative, gradients and virials to 1e-9 of their largest entry.
Mixed: 1e-5 of the largest gradient entry and 1e-6 relative on energies,
and two evaluations bit-identical (the monopole shifts and the scaled pairs'
dE/dq are summed in fixed point).
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from polmdqc_amd import tinker

pytestmark = pytest.mark.gpu

TD = os.path.join(GOLDEN, "tinker")
PRM = [os.path.join(TD, "params")]


@pytest.fixture(scope="module")
def engine():
    from polmdqc_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def dimer(tmp_path):
    s = tinker.system_from_files(os.path.join(TD, "hippo/water21.xyz"), os.path.join(TD, "hippo/water21.key"), PRM)
    s["use_polar"] = 1
    return s


def water21_box(tmp_path, nmol=64):
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
    return s


def worst(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", ["dimer", "box"])
def test_charge_flux_matches_oracle(engine, case, tmp_path):
    from polmdqc_amd import _lib as L

    print(f"[gpu] charge flux {case}", flush=True)
    s = (dimer if case == "dimer" else water21_box)(tmp_path)
    s["poleps"] = 1e-10
    assert s["use_chgflx"] and s["pentype"] == 1
    o = oracle.elec(s, want_grad=True)
    engine.set_system(s)
    terms = L.TERM_MPOLE | L.TERM_POLAR
    g = engine.eval(L.VIRIAL, terms, L.PLAIN_FP64)
    for k in ("em", "ep"):
        assert abs(g[k] - o[k]) <= 1e-9 * abs(o[k]), (k, g[k], o[k])
    assert worst(g["dem"], o["dem"]) <= 1e-9
    assert worst(g["dep"], o["dep"]) <= 1e-9
    assert worst(g["vir"], o["vir_m"] + o["vir_p"]) <= 1e-9
    # each term alone as in the sum
    m = engine.eval(L.GRADIENT, L.TERM_MPOLE, L.PLAIN_FP64)
    assert worst(m["dem"], o["dem"]) <= 1e-9
    mx = engine.eval(L.VIRIAL, terms, L.MIXED)
    for k in ("em", "ep"):
        assert abs(mx[k] - o[k]) <= 1e-6 * abs(o[k]), (k, mx[k], o[k])
    assert worst(mx["dem"], o["dem"]) <= 1e-5
    assert worst(mx["dep"], o["dep"]) <= 1e-5
    assert worst(mx["vir"], o["vir_m"] + o["vir_p"]) <= 1e-5
    mx2 = engine.eval(L.GRADIENT, terms, L.MIXED)
    assert np.array_equal(mx["desum"], mx2["desum"])
