"""Pins the oracle's buffered 14-7 van der Waals and charge-transfer terms
(oracle/pmq_oracle_elec.c orc_vdwct) and the reader's vdW tables.  The
reference has the parameters of both (kvdw.cpp, kchgtrn.cpp) and neither
energy (energy.cpp:268,276), so these checks fix the forms from what the
mathematics and the parameter files determine:

* at r = R_ik the buffered 14-7 form gives exactly -eps_ik; R_ik and eps_ik
  are recomputed here from the VDW lines of the parameter file (diameters
  halved, cubic-mean radius, HHG depth) and the reader's tables must equal
  them;
* the hydrogen reduction equals moving each hydrogen to its reduced site
  kred x_H + (1 - kred) x_X by hand;
* gradients equal central differences of the energy and the virial equals
  the strain derivative (both terms, periodic and not);
* scaled and excluded pairs: a lone water molecule has no energy in either
  term; a wrong unit factor, a dropped scale or a missing reduction would
  fail one of these.
"""
import os
import re

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from polmdqc_amd import tinker
from test_oracle_virial import close, strain_derivative

TD = os.path.join(GOLDEN, "tinker")
PRM = [os.path.join(TD, "params")]


def build(xyz, key=None):
    return tinker.system_from_files(os.path.join(TD, xyz), os.path.join(TD, key) if key else None, PRM)


def vdw_lines(prm):
    out = {}
    for line in open(os.path.join(TD, "params", prm)):
        m = re.match(r"vdw\s+(\d+)\s+(\S+)\s+(\S+)", line)
        if m:
            out[int(m.group(1))] = (float(m.group(2)), float(m.group(3)))
    return out


def test_tables_and_the_well_depth_at_r_min():
    s = build("chkpole/alatet.xyz")
    v = vdw_lines("amoebabio09.prm")
    cls = sorted({int(c) for c in np.unique(s["vdwclass"])})
    nv = int(s["nvdw"])
    # the classes in the order the reader indexes them
    P = tinker.parse_prm(os.path.join(TD, "params", "amoebabio09.prm"))
    types = s["type"] + 1
    present = sorted({P["atom"][int(t)]["class"] for t in types})
    assert len(present) == nv
    R = np.array([0.5 * v[c][0] for c in present])           # diameters halved
    E = np.array([v[c][1] for c in present])
    for a in range(nv):
        for b in range(nv):
            r0 = 2.0 * (R[a] ** 3 + R[b] ** 3) / (R[a] ** 2 + R[b] ** 2)
            e0 = 4.0 * E[a] * E[b] / (np.sqrt(E[a]) + np.sqrt(E[b])) ** 2
            assert abs(s["radmin"][a * nv + b] - r0) < 1e-14 and abs(s["epsilon"][a * nv + b] - e0) < 1e-15
    # two unbonded sites of classes a, b at r = R_ab: E = -eps_ab
    a, b = 0, nv - 1
    t = two_atoms(s, a, b, s["radmin"][a * nv + b])
    o = oracle.vdwct(t, want_grad=False)
    assert abs(o["ev"] + s["epsilon"][a * nv + b]) < 1e-14


def two_atoms(s, a, b, r):
    from test_oracle_damping import two_sites

    t = two_sites(r, [0.0, 0.0], [0.0, 0.0], [0.0, 0.0])
    t.update({"use_vdw": 1, "nvdw": s["nvdw"], "vdwclass": np.array([a, b], np.int32), "radmin": s["radmin"],
              "epsilon": s["epsilon"], "radmin4": s["radmin4"], "epsilon4": s["epsilon4"], "kred": np.zeros(2),
              "ired": np.array([0, 1], np.int32), "v2scale": 0.0, "v3scale": 0.0, "v4scale": 1.0, "v5scale": 1.0,
              "vdwcut": 50.0, "vdwtaper": 45.0, "use_ctrn": 0})
    return t


def test_reduction_equals_moved_hydrogens():
    s = build("amoeba/water09.xyz")
    s["use_vdw"] = 1
    full = oracle.vdwct(s, want_grad=False)["ev"]
    t = dict(s)
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    red = xyz.copy()
    for i in range(len(xyz)):
        j = int(s["ired"][i])
        if j != i:
            red[i] = s["kred"][i] * (xyz[i] - xyz[j]) + xyz[j]
    t["x"], t["y"], t["z"] = (np.ascontiguousarray(red[:, q]) for q in range(3))
    t["kred"] = np.zeros(len(xyz))
    t["ired"] = np.arange(len(xyz), dtype=np.int32)
    assert abs(oracle.vdwct(t, want_grad=False)["ev"] - full) < 1e-13
    assert np.any(s["ired"] != np.arange(len(xyz)))


def fd_grad(s, key, h=1e-5):
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    g = np.zeros_like(xyz)
    for i in range(len(xyz)):
        for a in range(3):
            for sg in (1, -1):
                x = xyz.copy()
                x[i, a] += sg * h
                t = dict(s)
                t["x"], t["y"], t["z"] = (np.ascontiguousarray(x[:, q]) for q in range(3))
                g[i, a] += sg * oracle.vdwct(t, want_grad=False)[key] / (2 * h)
    return g


@pytest.mark.parametrize("case", ["vdw_water", "vdw_alatet", "ctrn_hippo"])
def test_gradient_matches_central_differences(case):
    if case == "ctrn_hippo":
        s = build("hippo/water21.xyz")
        s["use_ctrn"], key, gk = 1, "ect", "dect"
    else:
        s = build("amoeba/water09.xyz" if case == "vdw_water" else "chkpole/alatet.xyz")
        s["use_vdw"], key, gk = 1, "ev", "dev"
    o = oracle.vdwct(s)
    g = fd_grad(s, key)
    assert np.max(np.abs(o[gk] - g)) <= 1e-6 * max(1.0, np.max(np.abs(g)))
    assert np.max(np.abs(o[gk].sum(0))) < 1e-10  # no net force
    # the cluster's virial is sum x (x) g
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    X = xyz.T @ o[gk]
    close(o["vir_v" if key == "ev" else "vir_c"], 0.5 * (X + X.T), 1e-10)


def small_box(ff_key_text, tmp_path):
    src = open(os.path.join(TD, "bounds/waterbox30.xyz")).read().splitlines()
    n = 3 * 64
    # 64 waters of the 30 A box, in a 12.5 A box around the first one
    lines = src[1:1 + 900 * 3]
    xyz = np.array([[float(v) for v in l.split()[2:5]] for l in lines])
    c = xyz[0]
    d = xyz - c
    mol = np.arange(len(xyz)) // 3
    keep = sorted({int(m) for m in mol[np.max(np.abs(d), axis=1) < 12.0]})[:64]
    rows = []
    for q, m in enumerate(keep):
        for a in range(3):
            w = lines[3 * m + a].split()
            idx = 3 * q + a + 1
            bonds = [3 * q + 2, 3 * q + 3] if a == 0 else [3 * q + 1]
            p = xyz[3 * m + a] - c
            rows.append(f"{idx:6d}  {w[1]:<3}{p[0]:12.6f}{p[1]:12.6f}{p[2]:12.6f}{int(w[5]):6d}" +
                        "".join(f"{b:6d}" for b in bonds))
    (tmp_path / "box.xyz").write_text(f"{n:6d}  box\n" + "\n".join(rows) + "\n")
    (tmp_path / "box.key").write_text(ff_key_text)
    return tinker.system_from_files(str(tmp_path / "box.xyz"), None, PRM)


@pytest.mark.parametrize("term", ["vdw", "ctrn"])
def test_periodic_virial_matches_strain_derivative(term, tmp_path):
    if term == "vdw":
        s = small_box("parameters water03\na-axis 24.0\nvdw-cutoff 9.0\nvdw-taper 8.0\n", tmp_path)
        s["use_vdw"], key, vk = 1, "ev", "vir_v"
    else:
        s = small_box("parameters water21\na-axis 24.0\nchgtrn-cutoff 6.0\nchgtrn-taper 5.4\n", tmp_path)
        s["use_ctrn"], key, vk = 1, "ect", "vir_c"
    o = oracle.vdwct(s)
    W = strain_derivative(lambda t: oracle.vdwct(t, want_grad=False)[key], s)
    close(o[vk], W, 1e-6)
    assert abs(o[key]) > 1e-3


def test_lone_molecule_has_no_energy():
    s = build("amoeba/water09.xyz")
    one = dict(s)
    for k in ("x", "y", "z"):
        one[k] = s[k][:3].copy()
    one["n"] = 3
    for k in ("type", "atomic", "molcule", "vdwclass", "kred", "ired", "chgct", "dmpct", "polaxe", "zaxis", "xaxis",
              "yaxis", "pcore", "pval0", "palpha", "mono0", "polarity", "thole", "pdamp"):
        one[k] = np.asarray(s[k])[:3].copy()
    one["pole"] = s["pole"][:39].copy()
    one["use"] = s["use"][:4].copy()
    for a in ("12", "13", "14", "15"):
        cnt = s["n" + a][:3]
        one["n" + a], one["i" + a] = cnt.copy(), s["i" + a][:int(cnt.sum())].copy()
    one["use_vdw"], one["use_ctrn"] = 1, 1
    one["chgct"] = np.ones(3)
    one["dmpct"] = np.full(3, 3.0)
    o = oracle.vdwct(one, want_grad=False)
    assert o["ev"] == 0.0 and o["ect"] == 0.0


def subset_xyz(src, atoms, dst):
    """Atoms (1-based, whole molecules) of a Tinker .xyz file, renumbered,
    written as their own file."""
    lines = open(src).read().splitlines()
    n = int(lines[0].split()[0])
    body = lines[1:1 + n] if len(lines[1].split()) > 1 and not _is_cell(lines[1]) else lines[2:2 + n]
    new = {a: q + 1 for q, a in enumerate(sorted(atoms))}
    rows = []
    for a in sorted(atoms):
        w = body[a - 1].split()
        bonds = [new[int(b)] for b in w[6:]]
        rows.append(f"{new[a]:6d}  {w[1]:<3}{float(w[2]):12.6f}{float(w[3]):12.6f}{float(w[4]):12.6f}{int(w[5]):6d}"
                    + "".join(f"{b:6d}" for b in bonds))
    open(dst, "w").write(f"{len(rows):6d}  subset\n" + "\n".join(rows) + "\n")


def _is_cell(line):
    try:
        return len([float(v) for v in line.split()]) == 6
    except ValueError:
        return False


@pytest.mark.parametrize("term", ["vdw", "ctrn"])
def test_active_pairs_are_total_minus_the_inactive_subsystem(term, tmp_path):
    """With a set A of active atoms, the pairs that count are those with an
    atom of A: the energy is E(all) - E(the rest alone), the rest's energy
    computed from a file holding only those atoms."""
    if term == "vdw":
        xyz, prm, active = os.path.join(TD, "active/water09_Na_Cls.xyz"), "amoebabio09", [5, 6, 7, 14, 15, 16]
        key, flag = "ev", "use_vdw"
    else:
        # the 8 waters of the 30 A box nearest its first one, as a cluster
        box = os.path.join(TD, "bounds/waterbox30.xyz")
        lines = open(box).read().splitlines()[1:]
        ox = np.array([[float(v) for v in lines[3 * m].split()[2:5]] for m in range(len(lines) // 3)])
        near = np.argsort(np.linalg.norm(ox - ox[0], axis=1))[:8]
        xyz = str(tmp_path / "cluster.xyz")
        subset_xyz(box, [3 * int(m) + a for m in near for a in (1, 2, 3)], xyz)
        prm, active, key, flag = "water21", [1, 2, 3], "ect", "use_ctrn"
    n = int(open(xyz).readline().split()[0])
    rest = [a for a in range(1, n + 1) if a not in active]
    (tmp_path / "a.key").write_text(f"parameters {prm}\nactive " + " ".join(map(str, active)) + "\n")
    (tmp_path / "all.key").write_text(f"parameters {prm}\n")
    subset_xyz(xyz, rest, str(tmp_path / "rest.xyz"))
    (tmp_path / "rest.key").write_text(f"parameters {prm}\n")

    def energy(x, k):
        s = tinker.system_from_files(x, str(k), PRM)
        s[flag] = 1
        return oracle.vdwct(s, want_grad=False)[key], s

    e_act, s_act = energy(xyz, tmp_path / "a.key")
    assert int(np.sum(s_act["use"][1:])) == len(active)
    e_all, _ = energy(xyz, tmp_path / "all.key")
    e_rest, _ = energy(str(tmp_path / "rest.xyz"), tmp_path / "rest.key")
    assert abs(e_rest) > 1e-3 and abs(e_act - e_all) > 1e-3
    assert abs(e_act - (e_all - e_rest)) < 1e-12 * max(1.0, abs(e_all))
