"""The multipole tiles in the quasi-internal frame (pmq_set_pair_form
PMQ_PAIR_QI; the formulation of the reference's pairMpoleQI.h:22-278 with
the frame of qi.h:11-64) against the reference's own dumped outputs and
against the Cartesian form on the same input.

Each pair's multipoles are rotated so that the pair vector lies on z, the
algebra is evaluated there and the force and torques are rotated back: an
error in the rotation of the quadrupole, the frame's handedness or the
back-rotation moves the energy, gradient or virial by far more than the
bars.  FP64: the reference's outputs to 1e-9 relative (as the Cartesian
form), the Cartesian form to 1e-11.  Mixed precision: both forms round
each pair in FP32 differently, so the bar is that of mixed against FP64.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_fixture
from polmdqc_amd import water

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


def both_forms(engine, *args, **kw):
    from polmdqc_amd.engine import PAIR_CARTESIAN, PAIR_QI, set_pair_form

    set_pair_form(engine, PAIR_QI)
    try:
        q = engine.eval(*args, **kw)
    finally:
        set_pair_form(engine, PAIR_CARTESIAN)
    return q, engine.eval(*args, **kw)


@pytest.mark.parametrize("system", SYSTEMS)
def test_qi_matches_reference_and_cartesian(engine, system):
    from polmdqc_amd import _lib as L

    print(f"[gpu] qi parity {system}", flush=True)
    fx = load_fixture(system)
    engine.set_system(fx)
    q, c = both_forms(engine, L.VIRIAL, L.TERM_MPOLE, L.PLAIN_FP64)
    em = float(fx["em"])
    assert abs(q["em"] - em) <= REL * max(1.0, abs(em))
    ga = fx["grad_atoms"]
    assert rel_err(q["dem"][ga].ravel(), fx["dem"]) <= REL
    assert rel_err(q["vir"].ravel(), fx["vir"]) <= REL
    assert abs(q["em"] - c["em"]) <= 1e-11 * max(1.0, abs(c["em"]))
    assert rel_err(q["dem"], c["dem"]) <= 1e-11
    assert rel_err(q["vir"], c["vir"]) <= 1e-11


def test_qi_analysis_counts_and_partition(engine):
    from polmdqc_amd import _lib as L

    fx = load_fixture("misc_alatet_water09")
    engine.set_system(fx)
    q, c = both_forms(engine, L.ANALYSIS, L.TERM_MPOLE, L.PLAIN_FP64)
    assert q["nem"] == c["nem"] == int(fx["nem"])
    assert rel_err(q["aem"], c["aem"]) <= 1e-11
    assert abs(q["einter"] - c["einter"]) <= 1e-11 * max(1.0, abs(c["einter"]))


@pytest.mark.parametrize("ff", ["water03", "hippo19"])
def test_qi_ewald_box_fp64_and_mixed(engine, ff):
    """Ewald real space, and for HIPPO the charge-penetration pair form."""
    from polmdqc_amd import _lib as L

    print(f"[gpu] qi ewald {ff}", flush=True)
    ew = {"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0}
    if ff == "hippo19":
        ew["dewald"] = 0.45
    s = water.water_system(water.lattice_box(6, 18.643, 216), ff, box=18.643, ewald=ew)
    engine.set_system(s)
    q, c = both_forms(engine, L.VIRIAL, L.TERM_MPOLE, L.PLAIN_FP64)
    assert abs(q["em"] - c["em"]) <= 1e-11 * abs(c["em"])
    assert rel_err(q["dem"], c["dem"]) <= 1e-11
    assert rel_err(q["vir"], c["vir"]) <= 1e-11
    qm, cm = both_forms(engine, L.VIRIAL, L.TERM_MPOLE, L.MIXED)
    n = s["n"]
    for m in (qm, cm):
        assert abs(m["em"] - c["em"]) <= 1e-5 * n
        assert float(np.sqrt(np.mean(np.sum((m["dem"] - c["dem"]) ** 2, axis=1)))) < 1e-3
