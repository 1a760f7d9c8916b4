"""The C++ host layer (polmdqc_amd/host) used the way the reference is used.

tests/host/host_check fills pmq_host::g (the reference's global names),
calls pmq_host::energy<CalcMode>() and writes em/esum/nem/einter, the total
gradient and the virial.  It must reproduce the reference's own outputs
(the dumped goldens) to 1e-9 in Virial and Analysis mode, and the Python
front end's polarization result exactly on a periodic box.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_fixture
from polmdqc_amd import water

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "host", "host_check")
REL = 1e-9


def write_system(path, s):
    with open(path, "w") as f:
        for k, v in s.items():
            a = np.atleast_1d(np.asarray(v))
            if a.dtype.kind not in "biuf":
                continue
            a = a.astype(np.float64).ravel()
            f.write(f"{k} {a.size} " + " ".join(repr(float(x)) for x in a) + "\n")


def read_out(path):
    out = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            out[p[0]] = np.array([float(x) for x in p[2:]])
    return out


def run(tmp_path, s, mode, terms, precision=0):
    sysf, outf = tmp_path / "sys.txt", tmp_path / "out.txt"
    write_system(sysf, s)
    print(f"[gpu] host_check mode={mode} terms={terms}", flush=True)
    r = subprocess.run([EXE, str(sysf), str(outf), str(mode), str(terms), str(precision)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return read_out(outf)


@pytest.mark.parametrize("system", ["amoeba_water09", "hippo_water21", "bounds_triclinic", "misc_alatet_water09"])
def test_host_energy_matches_reference(tmp_path, system):
    fx = load_fixture(system)
    s = dict(fx)
    v = run(tmp_path, s, 3, 1)
    a = run(tmp_path, s, 1, 1)
    em = float(fx["em"])
    assert abs(v["em"][0] - em) <= REL * max(1.0, abs(em))
    assert abs(a["em"][0] - float(fx["em_analysis"])) <= REL * max(1.0, abs(em))
    assert int(a["nem"][0]) == int(fx["nem"])
    assert abs(a["einter"][0] - float(fx["einter"])) <= REL * max(1.0, abs(em))
    g = v["desum"].reshape(-1, 3)[fx["grad_atoms"]].ravel()
    assert np.max(np.abs(g - fx["dem"])) <= REL * np.abs(fx["dem"]).max()
    assert np.max(np.abs(v["vir"] - fx["vir"])) <= REL * np.abs(fx["vir"]).max()


def test_host_polarization_matches_python_front_end(tmp_path):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import Engine

    xyz = water.lattice_box(6, 18.643, 216)
    s = water.water_system(xyz, "water03", box=18.643,
                           ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})
    h = run(tmp_path, s, 2, 3)
    e = Engine(0)
    try:
        e.set_system(s)
        p = e.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    finally:
        e.close()
    assert int(h["pcg_iters"][0]) == p["pcg_iters"]
    assert abs(h["esum"][0] - (p["em"] + p["ep"])) <= 1e-12 * abs(p["em"] + p["ep"])
    assert np.max(np.abs(h["desum"].reshape(-1, 3) - p["desum"])) <= 1e-12 * np.abs(p["desum"]).max()
