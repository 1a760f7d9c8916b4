"""The Tinker readers and parameter assignment (polmdqc_amd/tinker.py)
against the reference's own setup, system by system.

Each reference test system is rebuilt from its input files (the .xyz and
.key files of the reference's test/testFiles and the parameter files of
params/, kept under tests/golden/tinker as data) and every array the
electrostatics path reads is compared with what the compiled reference's
initial/getcart/mechanic left in its modules (the fixtures, written by
tests/golden/make_golden.py).  Integer arrays (types, 1-2 .. 1-5 and
polarization-group lists, frames, molecules, active atoms, bonds and
angles) must be equal; parameters equal to the last bit (the reference
reads the same decimal text and applies the same unit factors in the same
order).  This covers AMOEBA (amoebabio09: the alanine tetrapeptide with its
chiral frames and 1-5 lists, ions, non-sequential atom numbering, active
and inactive atoms), AMOEBA+ (charge penetration Gordon2, charge flux),
HIPPO (Gordon1, charge flux, 3-fold and bisector frames of ammonia and
benzene) and all five box shapes.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_fixture
from polmdqc_amd import tinker

TD = os.path.join(GOLDEN, "tinker")
PRM = [os.path.join(TD, "params")]

SYSTEMS = {
    "amoeba_water09": ("amoeba/water09.xyz", None),
    "amoebap_waterap": ("amoebap/waterap.xyz", None),
    "hippo_water21": ("hippo/water21.xyz", None),
    "rotpole_water09_Na": ("rotpole/water09_Na.xyz", None),
    "rotpole_benzene_ethyne_water": ("rotpole/benzene_ethyne_water.xyz", None),
    "rotpole_ammonia": ("rotpole/ammonia.xyz", None),
    "chkpole_alatet": ("chkpole/alatet.xyz", None),
    "misc_alatet_water09": ("amoebaMisc/alatet_water09.xyz", None),
    "misc_water09NSeq": ("amoebaMisc/water09NSeq.xyz", None),
    "misc_water09_2": ("amoebaMisc/water09.xyz_2", "amoebaMisc/water09.key"),
    "active_water09_Na_Cls_active": ("active/water09_Na_Cls.xyz", "active/active.key"),
    "active_water09_Na_Cls_inactive": ("active/water09_Na_Cls.xyz", "active/inactive.key"),
    "active_waterbox30": ("active/waterbox30.xyz", None),
    "bounds_orthogonal": ("bounds/waterbox30.xyz", "bounds/orthogonal.key"),
    "bounds_monoclinic": ("bounds/waterbox30.xyz", "bounds/monoclinic.key"),
    "bounds_triclinic": ("bounds/waterbox30.xyz", "bounds/triclinic.key"),
    "bounds_octahedron": ("bounds/waterbox30.xyz", "bounds/octahedron.key"),
    "bounds_dodecahedron": ("bounds/waterbox30.xyz", "bounds/dodecahedron.key"),
}

INTS = ["type", "atomic", "n12", "i12", "n13", "i13", "n14", "i14", "n15", "i15", "molcule", "use", "polaxe",
        "zaxis", "xaxis", "yaxis", "np11", "ip11", "np12", "ip12", "np13", "ip13", "np14", "ip14", "nbond", "ibnd",
        "nangle", "iang", "balist", "pentype", "use_chgflx", "use_polar", "use_mpole", "use_bounds"]
REALS = ["x", "y", "z", "pole", "mono0", "polarity", "thole", "pdamp", "pcore", "pval0", "palpha", "bl", "bflx",
         "anat", "aflx", "abflx", "sizpr", "dmppr", "elepr", "electric", "dielec", "mpolecut", "ewaldcut",
         "usolvcut", "poleps"]
SCALES = sorted(tinker.SCALE_DEFAULTS)
BOX = ["xbox", "ybox", "zbox", "alpha", "beta", "gamma", "orthogonal", "monoclinic", "triclinic", "octahedron",
       "dodecadron"]


def build(name):
    xyz, key = SYSTEMS[name]
    return tinker.system_from_files(os.path.join(TD, xyz), os.path.join(TD, key) if key else None, PRM)


@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_setup_equals_reference(name):
    fx = load_fixture(name)
    s = build(name)
    assert int(s["n"]) == int(fx["n"])
    names = INTS + REALS + SCALES + (BOX if int(fx["use_bounds"]) else [])
    checked = 0
    for k in names:
        if k not in fx:
            continue
        a, b = np.asarray(s[k]).ravel(), np.asarray(fx[k]).ravel()
        assert a.shape == b.shape, (k, a.shape, b.shape)
        if k in INTS:
            assert np.array_equal(a.astype(np.int64), b.astype(np.int64)), k
        else:
            assert np.array_equal(a.astype(float), b.astype(float)), (k, np.max(np.abs(a - b)))
        checked += 1
    assert checked >= 50
    assert str(s["poltyp"]) == str(fx["poltyp"])


def test_system_runs_through_the_oracle_like_the_fixture():
    """The rebuilt alanine tetrapeptide gives the reference's energy."""
    import oracle

    fx = load_fixture("chkpole_alatet")
    s = build("chkpole_alatet")
    o = oracle.empole(s, oracle.ENERGY)
    assert abs(o["em"] - float(fx["em"])) <= 1e-10 * abs(float(fx["em"]))


def test_reader_details():
    # POLARIZE field rule: an integer-led field is a group type
    f = tinker._polarize_fields(["1.334", "0.390", "2", "5"])
    assert f == {"alpha": 1.334, "thole": 0.39, "tholed": 0.0, "group": [2, 5]}
    f = tinker._polarize_fields(["0.837", "0.390", "0.700", "7"])
    assert f["thole"] == 0.39 and f["tholed"] == 0.7 and f["group"] == [7]
    # frame kinds from the signs of the frame types
    assert tinker.frame_kind(2, 3, 0) == tinker.FRAME["Z-then-X"]
    assert tinker.frame_kind(-2, -2, 0) == tinker.FRAME["Bisector"]
    assert tinker.frame_kind(2, -3, -3) == tinker.FRAME["Z-Bisect"]
    assert tinker.frame_kind(-2, -2, -2) == tinker.FRAME["3-Fold"]
    assert tinker.frame_kind(2, 0, 0) == tinker.FRAME["Z-Only"]
    assert tinker.frame_kind(0, 0, 0) == tinker.FRAME["None"]
    # ACTIVE ranges
    assert tinker._atom_list([-14, 16, 5, 6, 7], 100) == [13, 14, 15, 4, 5, 6]


def test_archive_loop_reaches_the_reference_goldens():
    """amoebaMisc-2: analyze over the two-structure archive (the alanine
    tetrapeptide, then a water dimer: a new setup for the second); the
    reference's goldens hold the last structure's results."""
    import json

    import oracle

    g = json.load(open(os.path.join(GOLDEN, "ref_headers.json")))["amoebaMisc2"]
    frames = list(tinker.archive_systems(os.path.join(TD, "amoebaMisc/alatet_water09.xyz"), None, PRM))
    assert [f["n"] for f in frames] == [42, 6]
    res = [oracle.empole(s, oracle.ANALYSIS) for s in frames]
    a = res[-1]
    assert a["nem"] == int(g["nem"])
    assert abs(a["em"] - g["em"]) <= g["eps"] and abs(a["einter"] - g["einter"]) <= g["eps"]
    np.testing.assert_allclose(a["aem"], g["aem"], rtol=0, atol=g["eps"])
    # the first structure against the reference's own evaluation of it
    fx = load_fixture("misc_alatet_water09")
    assert res[0]["nem"] == int(fx["nem"])
    assert abs(res[0]["em"] - float(fx["em_analysis"])) <= 1e-9 * abs(float(fx["em"]))


def test_unchanged_topology_keeps_the_setup(tmp_path):
    """A second structure with the same atoms takes only new coordinates
    (analyze.cpp:202-212): no new chirality check at the new positions."""
    src = open(os.path.join(TD, "chkpole/alatet.xyz")).read().splitlines()
    n = int(src[0].split()[0])
    body = src[1:1 + n]
    moved = []
    for line in body:
        w = line.split()
        x, y, z = (float(v) + 0.01 for v in w[2:5])
        moved.append(f"{w[0]:>6}  {w[1]:<3}{x:12.6f}{y:12.6f}{z:12.6f}  " + "  ".join(w[5:]))
    p = tmp_path / "two.xyz"
    p.write_text("\n".join([src[0]] + body + [src[0]] + moved) + "\n")
    (tmp_path / "two.key").write_text(open(os.path.join(TD, "chkpole/alatet.key")).read())
    a, b = list(tinker.archive_systems(str(p), None, PRM))
    assert np.array_equal(a["pole"], b["pole"]) and np.array_equal(a["yaxis"], b["yaxis"])
    assert np.allclose(b["x"] - a["x"], 0.01)
