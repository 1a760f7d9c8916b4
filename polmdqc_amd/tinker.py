"""Tinker input for the electrostatics path: .xyz coordinates and
connectivity, .key keywords and .prm parameter files, and the parameter
assignment the reference's mechanic() performs over them, for any molecule.

The result is the system dict the engine takes (the reference's global
arrays after setup, the layout of the fixtures).  Each step follows the
reference's own:

* coordinates and bonds: readxyz.cpp:38-213 (optional unit-cell line, atom
  tags renumbered in file order, bond lists sorted)
* keywords: getkey.cpp:29 (the .key file), the .prm file's keyword lines
  first (prmkey.cpp:38); a later line wins
* parameters: readprm.cpp:64 (ATOM :186, BOND :220, ANGLE :320, REPULSION
  :1032, DISPERSION :1047, MULTIPOLE :1157, CHGPEN :1237, POLARIZE :1250,
  BNDCFLUX :1371, ANGCFLUX :1394)
* 1-3, 1-4, 1-5 lists: attach.cpp:12-129; molecules: molecule.cpp:15-60;
  bonds and angles in the order bonds()/angles() list them
* multipole frames: kmpole.cpp:290-441 (four passes over the MULTIPOLE
  records in order: z and x among the 1-2 atoms, x among the 1-3 atoms
  bonded to z, z only, no frame atom), units kmpole.cpp:538-552 (dipole x
  bohr, quadrupole x bohr^2/3), chirality check chkpole.cpp:38-78
* polarizabilities and groups: kpolar.cpp:290-370 (POLARIZE fields),
  :461-476 (pdamp = alpha^(1/6)), polargrp :592-853 (group closure over
  bonded group members, then groups 1-2, 1-3, 1-4 apart)
* charge penetration by class: kmpole.cpp:660-676; charge flux by class
  pairs and triples: kchgflx.cpp:150-230
* active atoms: active.cpp (ACTIVE / INACTIVE lists and ranges,
  ACTIVE-SPHERE)

Parity: tests/test_tinker_inputs.py rebuilds every reference test system
from its own input files and compares each array with what the compiled
reference's setup produced (the fixtures).
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

BOHR = 0.529177210903          # units.h:75
COULOMB = 332.0637133          # units.h:79
FRAME = {"None": 0, "Z-then-X": 1, "Z-Only": 2, "Bisector": 3, "Z-Bisect": 4, "3-Fold": 5}  # kmulti.h:23-31

_INT = re.compile(r"[+-]?\d+")


def int_prefix(tok: str) -> int:
    """The integer an istream would read from the start of tok (0 if none):
    how the reference tells group types from reals (getnumb.cpp:9-22)."""
    m = _INT.match(tok)
    return int(m.group(0)) if m else 0


def frame_kind(kz: int, kx: int, ky: int) -> int:
    """Frame type of a MULTIPOLE record from the signs of its frame atom
    types (readprm.cpp:1206-1211)."""
    axt = "Z-then-X"
    if kz == 0:
        axt = "None"
    if kz != 0 and kx == 0:
        axt = "Z-Only"
    if kz < 0 or kx < 0:
        axt = "Bisector"
    if kx < 0 and ky < 0:
        axt = "Z-Bisect"
    if max(kz, kx, ky) < 0:
        axt = "3-Fold"
    return FRAME[axt]


# ---------------------------------------------------------------------------
# files

def _read_frame(lines: list[str], k: int, path: str):
    """One structure starting at line k (readxyz.cpp:66-213); returns it
    and the line after it, or (None, k) at the end of the file."""
    while k < len(lines) and not lines[k].split():
        k += 1
    if k >= len(lines):
        return None, k
    head = lines[k].split()
    n = int(head[0])
    title = lines[k][lines[k].index(head[0]) + len(head[0]):].strip()
    k += 1
    box = None
    tags, names, xyz, types, bonded = [], [], [], [], []
    while len(tags) < n:
        if k >= len(lines):
            raise ValueError(f"{path}: {len(tags)} of {n} atoms")
        w = lines[k].split()
        k += 1
        if not w:
            continue
        if not tags and not any(ch.isalpha() for ch in lines[k - 1]):
            box = [float(v) for v in w[:6]]  # a b c alpha beta gamma
            continue
        tags.append(int(w[0]))
        names.append(w[1])
        xyz.append([float(w[2]), float(w[3]), float(w[4])])
        types.append(int(w[5]))
        bonded.append([int(v) for v in w[6:]])
    index = {t: i for i, t in enumerate(tags)}
    i12 = [sorted(index[b] for b in bl) for bl in bonded]
    for i, bl in enumerate(i12):
        for j in bl:
            if i not in i12[j]:
                raise ValueError(f"{path}: bond {i + 1}-{j + 1} listed on one side only")
    return {"n": n, "title": title, "name": names, "xyz": np.array(xyz, dtype=float),
            "type": np.array(types, dtype=np.int64), "i12": i12, "box": box}, k


def read_xyz(path: str) -> dict:
    """Atoms of a Tinker .xyz file: names, coordinates, 1-based types and
    0-based sorted bond lists, and the unit-cell line when there is one."""
    with open(path) as f:
        lines = f.read().splitlines()
    frame, _ = _read_frame(lines, 0, path)
    if frame is None:
        raise ValueError(f"{path}: no atoms")
    return frame


def read_frames(path: str) -> list[dict]:
    """Every structure of a coordinate archive (structures one after
    another, as analyze reads them, analyze.cpp:200-269)."""
    with open(path) as f:
        lines = f.read().splitlines()
    out, k = [], 0
    while True:
        frame, k = _read_frame(lines, k, path)
        if frame is None:
            return out
        out.append(frame)


def _records(lines: list[str]) -> list[str]:
    out = []
    for raw in lines:
        line = raw.split("!!")[0].rstrip()
        w = line.split()
        if not w or w[0].startswith("#"):
            continue
        out.append(line.strip())
    return out


def read_key(path: str) -> list[str]:
    with open(path) as f:
        return _records(f.read().splitlines())


def parse_prm(path: str) -> dict:
    """Records of a .prm file in file order, and its keyword lines."""
    with open(path) as f:
        lines = f.read().splitlines()
    P = {"keywords": [], "atom": {}, "multipole": [], "polarize": {}, "chgpen": {}, "repulsion": {},
         "dispersion": {}, "bond": [], "angle": [], "anglep": [], "bndcflux": [], "angcflux": [], "vdw": {},
         "vdw14": {}, "vdwpair": {}, "chgtrn": {}}
    i = 0
    while i < len(lines):
        line = lines[i].split("!!")[0].strip()
        i += 1
        w = line.split()
        if not w or w[0].startswith("#"):
            continue
        key = w[0].lower()
        rec = key in P and key != "keywords"
        if rec and (len(w) < 2 or not _INT.fullmatch(w[1])):
            continue  # prose in the file's literature notes
        if key == "atom":
            m = re.match(r'\s*atom\s+(\d+)\s+(\d+)\s+(\S+)\s+"([^"]*)"\s+(\d+)\s+(\S+)\s+(\d+)', line, re.I)
            t, cls = int(m.group(1)), int(m.group(2))
            P["atom"][t] = {"class": cls if cls else t, "symbol": m.group(3), "atomic": int(m.group(5)),
                            "mass": float(m.group(6)), "valence": int(m.group(7))}
        elif key == "multipole":
            ids = [int(v) for v in w[1:-1]]
            c = float(w[-1])
            more = []
            for _ in range(4):
                more.append([float(v) for v in lines[i].split()])
                i += 1
            P["multipole"].append(_mp_record(ids, c, more))
        elif key == "polarize":
            P["polarize"][int(w[1])] = _polarize_fields(w[2:])
        elif key == "chgpen":
            P["chgpen"][int(w[1])] = (abs(float(w[2])), float(w[3]))
        elif key == "repulsion":
            P["repulsion"][int(w[1])] = (float(w[2]), float(w[3]), -abs(float(w[4])))
        elif key == "dispersion":
            P["dispersion"][int(w[1])] = (float(w[2]), float(w[3]))
        elif key == "vdw":
            v = [float(x) for x in w[2:5]]
            P["vdw"][int(w[1])] = tuple(v + [0.0] * (3 - len(v)))
        elif key == "vdw14":
            P["vdw14"][int(w[1])] = (float(w[2]), float(w[3]))
        elif key in ("vdwpair", "vdwpr"):
            P["vdwpair"][tuple(sorted((int(w[1]), int(w[2]))))] = (float(w[3]), float(w[4]))
        elif key == "chgtrn":
            P["chgtrn"][int(w[1])] = (float(w[2]), float(w[3]))
        elif key == "bond":
            P["bond"].append(((int(w[1]), int(w[2])), float(w[3]), float(w[4])))
        elif key == "angle":
            ang = [float(v) for v in w[5:8]] + [0.0] * (3 - len(w[5:8]))
            P["angle"].append(((int(w[1]), int(w[2]), int(w[3])), float(w[4]), ang))
        elif key == "anglep":
            ang = [float(v) for v in w[5:7]] + [0.0] * (2 - len(w[5:7]))
            P["anglep"].append(((int(w[1]), int(w[2]), int(w[3])), float(w[4]), ang))
        elif key == "bndcflux":
            P["bndcflux"].append(((int(w[1]), int(w[2])), float(w[3])))
        elif key == "angcflux":
            P["angcflux"].append(((int(w[1]), int(w[2]), int(w[3])), [float(v) for v in w[4:8]]))
        elif not rec:
            P["keywords"].append(line)
    return P


def _mp_record(ids, c, more) -> dict:
    ids = ids + [0] * (4 - len(ids))
    t, kz, kx, ky = ids[:4]
    (dx, dy, dz), (qxx,), (qyx, qyy), (qzx, qzy, qzz) = [tuple(r) for r in more]
    # c, d, Q row by row (readprm.cpp:1219-1233)
    pole = [c, dx, dy, dz, qxx, qyx, qzx, qyx, qyy, qzy, qzx, qzy, qzz]
    return {"type": t, "z": abs(kz), "x": abs(kx), "y": abs(ky), "frame": frame_kind(kz, kx, ky), "pole": pole}


def _polarize_fields(toks: list[str]) -> dict:
    """polarizability, then Thole a and the direct-field damping when the
    next fields are not group types, then the group types
    (kpolar.cpp:292-345)."""
    pol = float(toks[0])
    thl, thd = -1.0, -1.0
    rest = toks[1:]
    if rest and int_prefix(rest[0]) == 0:
        thl = float(rest[0])
        rest = rest[1:]
        if rest and int_prefix(rest[0]) == 0:
            thd = float(rest[0])
            rest = rest[1:]
    grp = []
    for tok in rest:
        g = int_prefix(tok)
        if g == 0:
            break
        grp.append(g)
    return {"alpha": pol, "thole": max(0.0, thl), "tholed": max(0.0, thd), "group": grp}


def find_prm(name: str, key_dir: str, search: list[str]) -> str:
    """The parameter file a PARAMETERS line names: as given (relative to the
    key file), else by base name in the search directories."""
    cands = [name, name + ".prm"]
    for c in cands:
        p = c if os.path.isabs(c) else os.path.join(key_dir, c)
        if os.path.isfile(p):
            return p
    base = os.path.basename(name)
    for d in search:
        for c in (base, base + ".prm"):
            p = os.path.join(d, c)
            if os.path.isfile(p):
                return p
    raise FileNotFoundError(f"parameter file {name!r} not found (searched {search})")


# ---------------------------------------------------------------------------
# keywords

SCALE_KEYS = {
    "mpole-12-scale": "m2scale", "mpole-13-scale": "m3scale", "mpole-14-scale": "m4scale", "mpole-15-scale": "m5scale",
    "polar-12-scale": "p2scale", "polar-13-scale": "p3scale", "polar-14-scale": "p4scale", "polar-15-scale": "p5scale",
    "polar-12-intra": "p2iscale", "polar-13-intra": "p3iscale", "polar-14-intra": "p4iscale",
    "polar-15-intra": "p5iscale",
    "direct-11-scale": "d1scale", "direct-12-scale": "d2scale", "direct-13-scale": "d3scale",
    "direct-14-scale": "d4scale",
    "mutual-11-scale": "u1scale", "mutual-12-scale": "u2scale", "mutual-13-scale": "u3scale",
    "mutual-14-scale": "u4scale",
    "induce-12-scale": "w2scale", "induce-13-scale": "w3scale", "induce-14-scale": "w4scale",
    "induce-15-scale": "w5scale",
    "rep-12-scale": "r2scale", "rep-13-scale": "r3scale", "rep-14-scale": "r4scale", "rep-15-scale": "r5scale",
    "disp-12-scale": "dsp2scale", "disp-13-scale": "dsp3scale", "disp-14-scale": "dsp4scale",
    "disp-15-scale": "dsp5scale",
}

# initprm.cpp defaults
SCALE_DEFAULTS = {
    "m2scale": 0.0, "m3scale": 0.0, "m4scale": 1.0, "m5scale": 1.0,
    "p2scale": 0.0, "p3scale": 0.0, "p4scale": 1.0, "p5scale": 1.0,
    "p2iscale": 0.0, "p3iscale": 0.0, "p4iscale": 0.5, "p5iscale": 1.0,
    "d1scale": 0.0, "d2scale": 1.0, "d3scale": 1.0, "d4scale": 1.0,
    "u1scale": 1.0, "u2scale": 1.0, "u3scale": 1.0, "u4scale": 1.0,
    "w2scale": 1.0, "w3scale": 1.0, "w4scale": 1.0, "w5scale": 1.0,
    "r2scale": 0.0, "r3scale": 0.0, "r4scale": 1.0, "r5scale": 1.0,
    "dsp2scale": 0.0, "dsp3scale": 0.0, "dsp4scale": 1.0, "dsp5scale": 1.0,
}

TERMS = ("multipoleterm", "polarizeterm", "repulsionterm", "dispersionterm", "chgflxterm")


def settings(lines: list[str]) -> dict:
    """The keywords the electrostatics setup reads (prm lines first, then
    the key file's: a later line wins)."""
    s = dict(SCALE_DEFAULTS)
    s.update({"terms": {t: True for t in TERMS}, "cutoff": None, "mpole-cutoff": None, "ewald-cutoff": None,
              "ewald": False, "ewald-alpha": None, "pme-grid": None, "pme-order": 5, "polarization": "MUTUAL",
              "polar-eps": 1e-6, "polar-iter": 100, "usolve-cutoff": 4.5, "usolve-diagonal": 2.0,
              "pcg-precond": True, "dielectric": 1.0, "d-equals-p": False, "penetration": "GORDON1",
              "box": {}, "octahedron": False, "dodecahedron": False, "active": [], "inactive": [],
              "active-sphere": [], "repuls-cutoff": None, "disp-cutoff": None, "dewald": False,
              "dewald-alpha": None, "vdw-cutoff": None, "vdw-taper": 0.9, "chgtrn-cutoff": None,
              "chgtrn-taper": 0.9, "vdwtype": "LENNARD-JONES", "radiusrule": "ARITHMETIC", "radiustype": "R-MIN",
              "radiussize": "RADIUS", "epsilonrule": "GEOMETRIC", "vdwindex": "CLASS",
              "v2scale": 0.0, "v3scale": 0.0, "v4scale": 1.0, "v5scale": 1.0})
    s["terms"].update({"vdwterm": True, "chgtrnterm": True})
    for line in lines:
        try:
            _keyword(s, line)
        except (ValueError, IndexError):
            pass  # a value that does not read leaves the setting as it was (prmkey.cpp)
    return s


def _keyword(s: dict, line: str) -> None:
    w = line.split()
    k = w[0].lower()
    arg = w[1:]
    if True:
        if k in SCALE_KEYS:
            v = float(arg[0])
            s[SCALE_KEYS[k]] = 1.0 / v if v > 1.0 else v  # prmkey.cpp: "1-n scale" given as a divisor
        elif k in TERMS:
            v = arg[0].upper() if arg else ""
            if v == "ONLY":
                # potoff (prmkey.cpp:635-665) clears every energy term but
                # leaves charge flux as it is
                for t in list(s["terms"]):
                    if t != "chgflxterm":
                        s["terms"][t] = False
                s["terms"][k] = True
            elif v == "NONE":
                s["terms"][k] = False
            else:
                s["terms"][k] = True
        elif k in ("cutoff", "mpole-cutoff", "ewald-cutoff", "repuls-cutoff", "disp-cutoff", "vdw-cutoff",
                   "chgtrn-cutoff"):
            s[k] = float(arg[0])
        elif k in ("vdw-taper", "chgtrn-taper"):
            s[k] = float(arg[0])
        elif k in ("vdwtype", "radiusrule", "radiustype", "radiussize", "epsilonrule", "vdwindex"):
            s[k] = arg[0].upper()
        elif k in ("vdw-12-scale", "vdw-13-scale", "vdw-14-scale", "vdw-15-scale"):
            v = float(arg[0])
            s["v" + k[4] + "scale"] = 1.0 / v if v > 1.0 else v  # prmkey.cpp:391-404
        elif k in ("vdwterm", "chgtrnterm"):
            v = arg[0].upper() if arg else ""
            s["terms"][k] = v != "NONE"
        elif k == "ewald":
            s["ewald"] = True
        elif k == "dewald":
            s["dewald"] = True
        elif k in ("ewald-alpha", "dewald-alpha"):
            s[k] = float(arg[0])
        elif k == "pme-grid":
            g = [int(v) for v in arg[:3]]
            s["pme-grid"] = (g + [g[0]] * 3)[:3] if len(g) < 3 else g
        elif k == "pme-order":
            s["pme-order"] = int(arg[0])
        elif k == "polarization":
            s["polarization"] = arg[0].upper()
        elif k == "polar-eps":
            s["polar-eps"] = float(arg[0])
        elif k == "polar-iter":
            s["polar-iter"] = int(arg[0])
        elif k == "usolve-cutoff":
            s["usolve-cutoff"] = float(arg[0])
        elif k == "usolve-diagonal":
            s["usolve-diagonal"] = float(arg[0])
        elif k == "pcg-noprecond":
            s["pcg-precond"] = False
        elif k == "pcg-precond":
            s["pcg-precond"] = True
        elif k == "dielectric":
            s["dielectric"] = float(arg[0])
        elif k == "d-equals-p":
            s["d-equals-p"] = True
        elif k == "penetration":
            s["penetration"] = arg[0].upper()
        elif k in ("a-axis", "b-axis", "c-axis", "alpha", "beta", "gamma"):
            s["box"][k] = float(arg[0])
        elif k == "octahedron":
            s["octahedron"] = True
        elif k == "dodecahedron":
            s["dodecahedron"] = True
        elif k == "active":
            s["active"] += [int(v) for v in arg]
        elif k == "inactive":
            s["inactive"] += [int(v) for v in arg]
        elif k == "active-sphere":
            s["active-sphere"].append([float(v) for v in arg])


def _atom_list(vals: list[int], n: int) -> list[int]:
    """ACTIVE/INACTIVE entries: atom numbers, or -a b for the range a..b
    (active.cpp); entries beyond n are dropped."""
    vals = [0 if abs(v) > n else v for v in vals]
    out, i = [], 0
    while i < len(vals) and vals[i] != 0:
        if vals[i] > 0:
            out.append(vals[i] - 1)
            i += 1
        else:
            out += list(range(abs(vals[i]) - 1, abs(vals[i + 1])))
            i += 2
    return out


def use_flags(S: dict, xyz: np.ndarray) -> np.ndarray:
    """use[0] = 0, use[1..n] the active atoms (active.cpp)."""
    n = len(xyz)
    use = np.ones(n + 1, dtype=np.int32)
    use[0] = 0
    for k, sph in enumerate(S["active-sphere"]):
        if k == 0:
            use[1:] = 0
        if len(sph) >= 4 and sph[3] != 0.0:
            c, r = np.array(sph[:3]), sph[3]
        else:
            c, r = xyz[int(sph[0]) - 1], sph[1]
        d2 = np.sum((xyz - c) ** 2, axis=1)
        use[1:][d2 <= r * r] = 1
    for i in _atom_list(S["inactive"], n):
        use[i + 1] = 0
    mob = _atom_list(S["active"], n)
    if mob:
        use[1:] = 0
        for i in mob:
            use[i + 1] = 1
    return use


# ---------------------------------------------------------------------------
# topology

def attach(i12: list[list[int]]):
    """1-3, 1-4 and 1-5 lists (attach.cpp:12-129): atoms one bond further
    out that are not already closer, sorted and unique."""
    n = len(i12)
    i13, i14, i15 = [], [], []
    for i in range(n):
        s12 = set(i12[i])
        c = {k for j in i12[i] for k in i12[j] if k != i and k not in s12}
        i13.append(sorted(c))
    for i in range(n):
        s = set(i12[i]) | set(i13[i])
        c = {k for j in i13[i] for k in i12[j] if k != i and k not in s}
        i14.append(sorted(c))
    for i in range(n):
        s = set(i12[i]) | set(i13[i]) | set(i14[i])
        c = {k for j in i14[i] for k in i12[j] if k != i and k not in s}
        i15.append(sorted(c))
    return i13, i14, i15


def molecules(i12: list[list[int]]) -> np.ndarray:
    """Connected components numbered in the order of their first atom
    (molecule.cpp:15-60)."""
    n = len(i12)
    mol = -np.ones(n, dtype=np.int32)
    nm = 0
    for i in range(n):
        if mol[i] >= 0:
            continue
        stack = [i]
        mol[i] = nm
        while stack:
            a = stack.pop()
            for b in i12[a]:
                if mol[b] < 0:
                    mol[b] = nm
                    stack.append(b)
        nm += 1
    return mol


def bonds_angles(i12: list[list[int]]):
    """Bonds (i < j, in atom order) and angles (j-i-k, j < k, by central
    atom i) as bonds() and angles() list them, and each angle's two bonds."""
    bnd, where = [], {}
    for i, bl in enumerate(i12):
        for j in bl:
            if j > i:
                where[(i, j)] = len(bnd)
                bnd.append((i, j))
    ang, bal = [], []
    for i, bl in enumerate(i12):
        for a in range(len(bl)):
            for b in range(a + 1, len(bl)):
                j, k = bl[a], bl[b]
                ang.append((j, i, k))
                bal.append((where[tuple(sorted((j, i)))], where[tuple(sorted((i, k)))]))
    return bnd, ang, bal


def flat(lists):
    cnt = np.array([len(v) for v in lists], dtype=np.int32)
    arr = np.array([x for v in lists for x in v], dtype=np.int32)
    return cnt, arr


# ---------------------------------------------------------------------------
# parameter assignment

def assign_multipoles(types: np.ndarray, i12, i13, records: list[dict]):
    """pole (n x 13, prm units), frame type and frame atoms (1-based, 0 =
    none) by the four passes of kmpole.cpp:298-441 over the records in
    order; the first record that fits wins."""
    n = len(types)
    pole = np.zeros((n, 13))
    polaxe = np.zeros(n, dtype=np.int32)
    zax = np.zeros(n, dtype=np.int32)
    xax = np.zeros(n, dtype=np.int32)
    yax = np.zeros(n, dtype=np.int32)
    done = np.zeros(n, dtype=bool)

    def take(i, r, z=-1, x=-1, y=-1):
        pole[i] = r["pole"]
        polaxe[i] = r["frame"]
        zax[i], xax[i], yax[i] = z + 1, x + 1, y + 1
        done[i] = True

    for i in range(n):
        it = types[i]
        mine = [r for r in records if r["type"] == it]
        # z and x (and y) among the 1-2 atoms
        for r in mine:
            if done[i]:
                break
            for j in i12[i]:
                if done[i] or types[j] != r["z"]:
                    continue
                for k in i12[i]:
                    if done[i] or types[k] != r["x"] or k == j:
                        continue
                    if r["y"] == 0:
                        take(i, r, j, k)
                        break
                    for m in i12[i]:
                        if types[m] == r["y"] and m != j and m != k:
                            take(i, r, j, k, m)
                            break
        # z among the 1-2 atoms, x (and y) among the 1-3 atoms bonded to z
        for r in mine:
            if done[i]:
                break
            for j in i12[i]:
                if done[i] or types[j] != r["z"]:
                    continue
                for k in i13[i]:
                    if done[i] or types[k] != r["x"] or j not in i12[k]:
                        continue
                    if r["y"] == 0:
                        take(i, r, j, k)
                        break
                    for m in i13[i]:
                        if types[m] == r["y"] and m != k and j in i12[m]:
                            take(i, r, j, k, m)
                            break
        # only a z atom
        for r in mine:
            if done[i]:
                break
            if r["x"] != 0:
                continue
            for j in i12[i]:
                if types[j] == r["z"]:
                    take(i, r, j)
                    break
        # no frame atom
        for r in mine:
            if done[i]:
                break
            if r["z"] == 0:
                take(i, r)
    return pole, polaxe, zax, xax, yax


def chkpole(xyz, polaxe, zax, xax, yax, pole):
    """Invert the chiral sites whose y atom lies on the wrong side
    (chkpole.cpp:38-78)."""
    for i in range(len(polaxe)):
        k = int(yax[i])
        if polaxe[i] != FRAME["Z-then-X"] or k == 0:
            continue
        a, b, c, d = xyz[i], xyz[zax[i] - 1], xyz[xax[i] - 1], xyz[abs(k) - 1]
        ad, bd, cd = a - d, b - d, c - d
        c1 = bd[1] * cd[2] - bd[2] * cd[1]
        c2 = cd[1] * ad[2] - cd[2] * ad[1]
        c3 = ad[1] * bd[2] - ad[2] * bd[1]
        vol = ad[0] * c1 + bd[0] * c2 + cd[0] * c3
        if (k < 0 and vol > 0.0) or (k > 0 and vol < 0.0):
            yax[i] = -k
            for q in (2, 5, 7, 9, 11):
                pole[i, q] = -pole[i, q]


def polar_groups(types: np.ndarray, i12, polrec: dict):
    """ip11 .. ip14 (polargrp, kpolar.cpp:592-853): group members are atoms
    bonded through type pairs that each list the other as a group type,
    closed over bonded members; then the groups one, two and three group
    bonds away."""
    n = len(types)
    ip11 = [[i] for i in range(n)]
    for i in range(n):
        grp = polrec.get(int(types[i]), {}).get("group", [])
        for j in i12[i]:
            if int(types[j]) in grp:
                ip11[i].append(j)
    for i in range(n):
        for j in ip11[i]:
            if i not in ip11[j]:
                raise ValueError(f"polarization group of atoms {i + 1} and {j + 1} is not symmetric")
    for i in range(n):
        have = set(ip11[i])
        start = 0
        while start < len(ip11[i]):
            stop = len(ip11[i])
            for j in ip11[i][start:stop]:
                for k in ip11[j]:
                    if k not in have:
                        have.add(k)
                        ip11[i].append(k)
            start = stop
        ip11[i].sort()
    ip12 = []
    for i in range(n):
        mine = set(ip11[i])
        keep = [k for j in ip11[i] for k in i12[j] if k not in mine]
        ip12.append(sorted({m for k in keep for m in ip11[k]}))
    ip13 = []
    for i in range(n):
        seen = set(ip11[i]) | set(ip12[i])
        ip13.append(sorted({k for j in ip12[i] for k in ip12[j] if k not in seen}))
    ip14 = []
    for i in range(n):
        seen = set(ip11[i]) | set(ip12[i]) | set(ip13[i])
        ip14.append(sorted({k for j in ip13[i] for k in ip12[j] if k not in seen}))
    return ip11, ip12, ip13, ip14


def atomic_of(types, atom):
    return np.array([atom[int(t)]["atomic"] for t in types], dtype=np.int64)


def ideal_angles(ang, cls, atomic, i12, P) -> np.ndarray:
    """Ideal angle of each bond angle (kangle.cpp:296-430; no small-ring
    variants): the first ANGLE record of the class triple, else for a
    trigonal centre the first ANGLEP record, with a nonzero value at index
    nh = 1 + the hydrogens on the centre outside the angle.  The reference
    indexes each record's values (ANGLE: three, the second and third equal
    to the first when both are zero; ANGLEP: two) as one flat array, so an
    index past a record's own values reads the next record's first."""
    def flat(records, width):
        v = []
        for _, _, a in records:
            a = list(a)
            if all(x == 0.0 for x in a[1:]):
                a[1:] = [a[0]] * (width - 1)
            v += a
        return v

    fa, fp = flat(P["angle"], 3), flat(P["anglep"], 2)
    out = np.zeros(len(ang))
    for q, (a, b, c) in enumerate(ang):
        ca, cb, cc = int(cls[a]), int(cls[b]), int(cls[c])
        key3 = (ca, cb, cc) if ca <= cc else (cc, cb, ca)
        nh = 1 + sum(1 for h in i12[b] if h != a and h != c and atomic[h] == 1)
        done = False
        for recs, fl, width in ((P["angle"], fa, 3), (P["anglep"], fp, 2) if len(i12[b]) == 3 else (None, None, 0)):
            if recs is None or done:
                continue
            for j, ((c1, c2, c3), _, _) in enumerate(recs):
                k3 = (c1, c2, c3) if c1 <= c3 else (c3, c2, c1)
                idx = width * j + nh
                if k3 == key3 and idx < len(fl) and fl[idx] != 0.0:
                    out[q] = fl[idx]
                    done = True
                    break
    return out


def _unit_cell(S: dict, xyzbox):
    """Cell lengths and angles: the .xyz cell line, else A-AXIS .. GAMMA
    (b and c default to a, angles to 90)."""
    if xyzbox is not None:
        a, b, c, al, be, ga = xyzbox
    elif "a-axis" in S["box"]:
        B = S["box"]
        a = B["a-axis"]
        b = B.get("b-axis", a)
        c = B.get("c-axis", a)
        al, be, ga = B.get("alpha", 90.0), B.get("beta", 90.0), B.get("gamma", 90.0)
    else:
        return None
    return a, b, c, al, be, ga


def system_from_files(xyz_path: str, key_path: str | None = None, prm_search: list[str] | None = None) -> dict:
    """The system dict of a Tinker input: the .xyz file, its .key file (by
    default the .key beside it with the same base name) and the parameter
    file that key names."""
    X = read_xyz(xyz_path)
    if key_path is None:
        cand = os.path.splitext(xyz_path)[0] + ".key"
        key_path = cand if os.path.isfile(cand) else None
    key = read_key(key_path) if key_path else []
    kdir = os.path.dirname(os.path.abspath(key_path)) if key_path else os.path.dirname(os.path.abspath(xyz_path))
    pl = [l for l in key if l.split()[0].lower() == "parameters"]
    if not pl:
        raise ValueError("no PARAMETERS keyword")
    P = parse_prm(find_prm(pl[-1].split()[1], kdir, prm_search or []))
    return build_system(X, P, key)


def archive_systems(xyz_path: str, key_path: str | None = None, prm_search: list[str] | None = None):
    """The system of each structure of an archive, as analyze.cpp:200-212
    steps through it: the full setup again when the atom count or the atom
    types change, otherwise the last setup with the new coordinates."""
    if key_path is None:
        cand = os.path.splitext(xyz_path)[0] + ".key"
        key_path = cand if os.path.isfile(cand) else None
    key = read_key(key_path) if key_path else []
    kdir = os.path.dirname(os.path.abspath(key_path)) if key_path else os.path.dirname(os.path.abspath(xyz_path))
    pl = [l for l in key if l.split()[0].lower() == "parameters"]
    if not pl:
        raise ValueError("no PARAMETERS keyword")
    P = parse_prm(find_prm(pl[-1].split()[1], kdir, prm_search or []))
    last = None
    for X in read_frames(xyz_path):
        if last is None or X["n"] != last["n"] or not np.array_equal(X["type"] - 1, last["type"]):
            last = build_system(X, P, key)
        else:
            last = dict(last)
            last["x"], last["y"], last["z"] = (X["xyz"][:, q].copy() for q in range(3))
        yield last


def build_system(X: dict, P: dict, key: list[str]) -> dict:
    n = X["n"]
    xyz = X["xyz"]
    types = X["type"]                     # 1-based Tinker types
    S = settings(P["keywords"] + key)
    i12 = X["i12"]
    i13, i14, i15 = attach(i12)
    atom = P["atom"]
    cls = np.array([atom[int(t)]["class"] for t in types], dtype=np.int64)

    # multipoles: key-file records with positive types first (kmpole.cpp:72-243)
    recs = []
    lines = key
    for q, line in enumerate(lines):
        w = line.split()
        if w[0].lower() == "multipole" and int(w[1]) > 0:
            ids = [int(v) for v in w[1:-1]]
            more = [[float(v) for v in lines[q + r].split()] for r in (1, 2, 3, 4)]
            recs.append(_mp_record(ids, float(w[-1]), more))
    recs += P["multipole"]
    pole, polaxe, zax, xax, yax = assign_multipoles(types, i12, i13, recs)
    # the quadrupole factor as one constant: the reference is built with
    # -ffast-math, which folds bohr*bohr/3 (its fixtures hold those bits)
    pole[:, 1:4] *= BOHR
    pole[:, 4:13] = pole[:, 4:13] * (BOHR * BOHR / 3.0)

    # polarizabilities by type (kpolar.cpp:461-476), key records override
    polrec = dict(P["polarize"])
    for line in key:
        w = line.split()
        if w[0].lower() == "polarize" and int(w[1]) > 0:
            polrec[int(w[1])] = _polarize_fields(w[2:])
    pr = [polrec.get(int(t)) for t in types]
    polarity = np.array([r["alpha"] if r else 0.0 for r in pr])
    thole = np.array([r["thole"] if r else 0.0 for r in pr])
    # alpha^(1/6) as sqrt(cbrt(alpha)): the form -ffast-math gives the
    # reference's pow(alpha, 1/6) (equal bits on every fixture)
    pdamp = np.sqrt(np.cbrt(polarity))

    terms = S["terms"]
    use_polar = bool(terms["polarizeterm"]) and bool(np.any(polarity != 0.0))
    use_mpole = bool(terms["multipoleterm"])
    chkpole(xyz, polaxe, zax, xax, yax, pole)
    ip11, ip12, ip13, ip14 = polar_groups(types, i12, polrec)

    # charge penetration by class (kmpole.cpp:660-676)
    cp = [P["chgpen"].get(int(c)) for c in cls]
    pcore = np.array([v[0] if v else 0.0 for v in cp])
    palpha = np.array([v[1] if v else 0.0 for v in cp])
    pval0 = pole[:, 0] - pcore
    use_chgpen = bool(np.any(palpha != 0.0))
    pentype = 0
    if use_chgpen:
        pentype = {"GORDON1": 1, "GORDON2": 2}[S["penetration"]]

    # repulsion and dispersion by class
    rp = [P["repulsion"].get(int(c)) for c in cls]
    dp = [P["dispersion"].get(int(c)) for c in cls]
    sizpr = np.array([v[0] if v else 0.0 for v in rp])
    dmppr = np.array([v[1] if v else 0.0 for v in rp])
    elepr = np.array([v[2] if v else 0.0 for v in rp])
    csix = np.array([v[0] if v else 0.0 for v in dp])
    adisp = np.array([v[1] if v else 0.0 for v in dp])

    # bonds, angles and charge flux by class (kchgflx.cpp:150-230)
    bnd, ang, bal = bonds_angles(i12)
    bl = np.zeros(len(bnd))
    for q, (a, b) in enumerate(bnd):
        ca, cb = sorted((int(cls[a]), int(cls[b])))
        for (c1, c2), fc, length in P["bond"]:
            if tuple(sorted((c1, c2))) == (ca, cb):
                bl[q] = length
                break
    anat = ideal_angles(ang, cls, atomic_of(types, atom), i12, P)
    cfb = {}
    for (ia, ib), v in P["bndcflux"]:
        k2 = (ia, ib) if ia <= ib else (ib, ia)
        cfb[k2] = v if ia < ib else (-v if ib < ia else 0.0)
    bflx = np.zeros(len(bnd))
    nflx = 0
    for q, (a, b) in enumerate(bnd):
        ca, cb = int(cls[a]), int(cls[b])
        k2 = (ca, cb) if ca <= cb else (cb, ca)
        if k2 in cfb:
            nflx += 1
            bflx[q] = cfb[k2] if ca <= cb else -cfb[k2]
    cfa = {}
    for (ia, ib, ic), v in P["angcflux"]:
        if ia <= ic:
            cfa[(ia, ib, ic)] = v
        else:
            cfa[(ic, ib, ia)] = [v[1], v[0], v[3], v[2]]
    aflx = np.zeros((len(ang), 2))
    abflx = np.zeros((len(ang), 2))
    for q, (a, b, c) in enumerate(ang):
        ca, cb, cc = int(cls[a]), int(cls[b]), int(cls[c])
        k3 = (ca, cb, cc) if ca <= cc else (cc, cb, ca)
        if k3 in cfa:
            nflx += 1
            v = cfa[k3]
            if ca <= cc:
                aflx[q], abflx[q] = v[0:2], v[2:4]
            else:
                aflx[q], abflx[q] = v[1::-1], v[3:1:-1]
    # kchgflx.cpp:362-363
    use_chgflx = 1 if (nflx > 0 and terms["chgflxterm"] and (use_mpole or use_polar)) else 0

    s_vdw = vdw_setup(S, P, types, cls, i12)
    ct = [P["chgtrn"].get(int(c)) for c in cls]
    chgct = np.array([v[0] if v else 0.0 for v in ct])
    dmpct = np.array([v[1] if v else 0.0 for v in ct])

    nb, ib12 = flat(i12)
    n13, ib13 = flat(i13)
    n14, ib14 = flat(i14)
    n15, ib15 = flat(i15)
    np11, a11 = flat(ip11)
    np12, a12 = flat(ip12)
    np13, a13 = flat(ip13)
    np14, a14 = flat(ip14)
    s = {
        "n": n, "x": xyz[:, 0].copy(), "y": xyz[:, 1].copy(), "z": xyz[:, 2].copy(),
        "type": (types - 1).astype(np.int32), "atomic": np.array([atom[int(t)]["atomic"] for t in types], np.int32),
        "pole": pole.ravel(), "polaxe": polaxe, "zaxis": zax, "xaxis": xax, "yaxis": yax,
        "use": use_flags(S, xyz), "molcule": molecules(i12),
        "n12": nb, "i12": ib12, "n13": n13, "i13": ib13, "n14": n14, "i14": ib14, "n15": n15, "i15": ib15,
        "np11": np11, "ip11": a11, "np12": np12, "ip12": a12, "np13": np13, "ip13": a13, "np14": np14, "ip14": a14,
        "electric": COULOMB, "dielec": S["dielectric"],
        "pentype": pentype, "use_chgflx": use_chgflx,
        "nbond": len(bnd), "ibnd": np.array(bnd, np.int32).ravel(), "bl": bl,
        "nangle": len(ang), "iang": np.array(ang, np.int32).ravel(), "anat": anat,
        # the flux parameters and the bonds of each angle exist with charge
        # flux only (kchgflx)
        "bflx": bflx if use_chgflx else np.zeros(0), "aflx": aflx.ravel() if use_chgflx else np.zeros(0),
        "abflx": abflx.ravel() if use_chgflx else np.zeros(0),
        "balist": np.array(bal if use_chgflx else [], np.int32).ravel(),
        "mono0": pole[:, 0].copy(), "pcore": pcore, "pval0": pval0, "palpha": palpha,
        "polarity": polarity, "thole": thole, "pdamp": pdamp,
        "sizpr": sizpr, "dmppr": dmppr, "elepr": elepr, "csix": csix, "adisp": adisp,
        "poleps": S["polar-eps"], "politer": S["polar-iter"], "poltyp": S["polarization"],
        "usolvcut": S["usolve-cutoff"], "udiag": S["usolve-diagonal"], "pcgprec": 1 if S["pcg-precond"] else 0,
        "use_polar": 1 if use_polar else 0, "use_mpole": 1 if use_mpole else 0,
        "use_repel": 1 if (terms["repulsionterm"] and np.any(sizpr != 0.0)) else 0,
        "dpequal": 1 if S["d-equals-p"] else 0,
    }
    s.update({k: S[k] for k in SCALE_DEFAULTS})
    cell = _unit_cell(S, X["box"])
    if cell is None:
        s.update({"use_bounds": 0, "xbox": 0.0, "ybox": 0.0, "zbox": 0.0, "alpha": 0.0, "beta": 0.0,
                  "gamma": 0.0, "orthogonal": 0, "monoclinic": 0, "triclinic": 0, "octahedron": 0,
                  "dodecadron": 0})
    else:
        a, b, c, al, be, ga = cell
        octa, dode = bool(S["octahedron"]), bool(S["dodecahedron"])
        if octa or dode:
            # non-prism cells (unitcell.cpp:141-151)
            b, c = a, (a if octa else a * math.sqrt(2.0))
            al = be = ga = 90.0
        ortho = not (octa or dode) and al == 90.0 and be == 90.0 and ga == 90.0
        mono = not (octa or dode) and al == 90.0 and ga == 90.0 and be != 90.0
        s.update({"use_bounds": 1, "xbox": a, "ybox": b, "zbox": c, "alpha": al, "beta": be, "gamma": ga,
                  "orthogonal": int(ortho), "monoclinic": int(mono), "triclinic": int(not (ortho or mono or octa or dode)),
                  "octahedron": int(octa), "dodecadron": int(dode)})
    periodic = cell is not None
    big = 1e12
    cut = S["cutoff"]
    s["mpolecut"] = S["mpole-cutoff"] or cut or (9.0 if periodic else big)
    s["ewaldcut"] = S["ewald-cutoff"] or cut or 7.0
    s["use_ewald"] = 1 if (S["ewald"] and periodic) else 0
    s["repcut"] = S["repuls-cutoff"] or cut or 6.0
    s["dispcut"] = S["disp-cutoff"] or cut or (9.0 if periodic else big)
    s["use_dewald"] = 1 if (S["dewald"] and periodic) else 0
    s["bsorder"] = S["pme-order"] if s["use_ewald"] else 5
    if s["use_ewald"]:
        s["aewald"] = S["ewald-alpha"] if S["ewald-alpha"] is not None else ewald_coefficient(s["ewaldcut"])
        grid = S["pme-grid"] or [fft_size(1.2 * v) for v in (s["xbox"], s["ybox"], s["zbox"])]
        s["nfft1"], s["nfft2"], s["nfft3"] = grid
    else:
        s.update({"aewald": 0.0, "nfft1": 0, "nfft2": 0, "nfft3": 0})
    s["dewald"] = (S["dewald-alpha"] if S["dewald-alpha"] is not None else ewald_coefficient(s["dispcut"])) \
        if s["use_dewald"] else 0.0
    # van der Waals and charge transfer (cutoffs.cpp:44-70: vdW 9 A with a
    # box, else none; charge transfer 6 A; tapers given as a fraction of the
    # cutoff when below 1)
    s.update(s_vdw)
    s["vdwcut"] = S["vdw-cutoff"] or cut or (9.0 if periodic else big)
    s["vdwtaper"] = S["vdw-taper"] * s["vdwcut"] if S["vdw-taper"] < 1.0 else S["vdw-taper"]
    s["use_vdw"] = 1 if (terms["vdwterm"] and s_vdw["nvdw"] and np.any(s_vdw["vdw_eps"] != 0.0)) else 0
    s["chgct"], s["dmpct"] = chgct, dmpct
    s["ctrncut"] = S["chgtrn-cutoff"] or 6.0
    s["ctrntaper"] = S["chgtrn-taper"] * s["ctrncut"] if S["chgtrn-taper"] < 1.0 else S["chgtrn-taper"]
    s["use_ctrn"] = 1 if (terms["chgtrnterm"] and np.any(chgct != 0.0)) else 0
    return s


def vdw_setup(S: dict, P: dict, types, cls, i12) -> dict:
    """Buffered 14-7 parameters (kvdw.cpp): per class (or type) radius and
    well depth, halved when the file gives diameters, 2^(1/6) x when it gives
    sigma; pair tables by the radius and epsilon rules (VDWPAIR records
    override), the 1-4 tables from VDW14; hydrogen reduction factors and
    their bonded partner."""
    key = types if S["vdwindex"] == "TYPE" else cls
    present = sorted({int(k) for k in key})
    nv = len(present)
    idx = {k: q for q, k in enumerate(present)}
    rad, eps, red, rad4, eps4 = (np.zeros(nv) for _ in range(5))
    for q, k in enumerate(present):
        r, e, rd = P["vdw"].get(k, (0.0, 0.0, 0.0))
        r4, e4 = P["vdw14"].get(k, (0.0, 0.0))
        r4 = r4 or r
        e4 = e4 or e
        if S["radiustype"] == "SIGMA":
            r, r4 = r * 2.0 ** (1.0 / 6.0), r4 * 2.0 ** (1.0 / 6.0)
        if S["radiussize"] == "DIAMETER":
            r, r4 = 0.5 * r, 0.5 * r4
        rad[q], eps[q], red[q], rad4[q], eps4[q] = r, abs(e), rd, r4, abs(e4)

    def rule(ri, rk, ei, ek):
        if ri == 0.0 and rk == 0.0:
            rd = 0.0
        elif S["radiusrule"] == "ARITHMETIC":
            rd = ri + rk
        elif S["radiusrule"] == "GEOMETRIC":
            rd = 2.0 * math.sqrt(ri) * math.sqrt(rk)
        elif S["radiusrule"] == "CUBIC-MEAN":
            rd = 2.0 * (ri ** 3 + rk ** 3) / (ri * ri + rk * rk)
        else:
            rd = ri + rk
        if ei == 0.0 and ek == 0.0:
            ep = 0.0
        elif S["epsilonrule"] == "ARITHMETIC":
            ep = 0.5 * (ei + ek)
        elif S["epsilonrule"] == "HARMONIC":
            ep = 2.0 * ei * ek / (ei + ek)
        elif S["epsilonrule"] == "HHG":
            ep = 4.0 * ei * ek / (math.sqrt(ei) + math.sqrt(ek)) ** 2
        else:
            ep = math.sqrt(ei) * math.sqrt(ek)
        return rd, ep

    radmin, epsilon, radmin4, epsilon4 = (np.zeros((nv, nv)) for _ in range(4))
    for a in range(nv):
        for b in range(nv):
            radmin[a, b], epsilon[a, b] = rule(rad[a], rad[b], eps[a], eps[b])
            radmin4[a, b], epsilon4[a, b] = rule(rad4[a], rad4[b], eps4[a], eps4[b])
    for (ka, kb), (rp, ep) in P["vdwpair"].items():
        if ka in idx and kb in idx:
            a, b = idx[ka], idx[kb]
            rp = rp * 2.0 ** (1.0 / 6.0) if S["radiustype"] == "SIGMA" else rp
            for A, B in ((a, b), (b, a)):
                radmin[A, B] = radmin4[A, B] = rp
                epsilon[A, B] = epsilon4[A, B] = abs(ep)
    jv = np.array([idx[int(k)] for k in key], dtype=np.int32)
    kred = red[jv].copy()
    ired = np.array([i12[i][0] if (len(i12[i]) == 1 and kred[i] != 0.0) else i for i in range(len(types))],
                    dtype=np.int32)
    return {"nvdw": nv, "vdwclass": jv, "radmin": radmin.ravel(), "epsilon": epsilon.ravel(),
            "radmin4": radmin4.ravel(), "epsilon4": epsilon4.ravel(), "kred": kred, "ired": ired,
            "vdw_eps": eps, "v2scale": S["v2scale"], "v3scale": S["v3scale"], "v4scale": S["v4scale"],
            "v5scale": S["v5scale"]}


def ewald_coefficient(cut: float, eps: float = 1e-8) -> float:
    """The Gaussian width whose erfc(a cut)/cut falls to eps, by bisection
    (the criterion of kewald.cpp:286-318)."""
    x = 0.5
    while math.erfc(x * cut) / cut >= eps:
        x *= 2.0
    lo, hi = 0.0, x
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if math.erfc(mid * cut) / cut >= eps:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def fft_size(minimum: float) -> int:
    """The smallest grid size of the form 2^a 3^b 5^c at or above minimum."""
    k = max(8, int(math.ceil(minimum)))
    while True:
        m = k
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return k
        k += 1
