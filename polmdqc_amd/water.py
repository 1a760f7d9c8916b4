"""Seeded synthetic water systems (rigid waters on a lattice or in clusters).

The workloads this package is measured on are water boxes built the same way
every time from a seed:

* a cubic periodic box of side ``L`` holding ``k**3`` rigid waters
  (O-H 0.9572 A, H-O-H 104.52 deg) whose oxygens sit on the lattice sites
  ``(i + 1/2) L / k``, each with a uniformly random orientation (Shoemake's
  quaternion method) and a per-molecule translation jitter uniform in
  [-0.1, 0.1] A per axis;
* non-periodic clusters of n waters placed at random in a sphere whose
  volume gives 1 g/cm^3 (29.9 A^3 per molecule), rejecting O-O < 2.5 A.

Atom order inside a molecule is O, H, H; the Tinker atom types are those of
the chosen parameter file (water03.prm: O=1, H=2; hippo19.prm: O=16, H=17),
which is how ``readxyz.cpp:38`` would see the same system.
"""
from __future__ import annotations

import numpy as np

ROH = 0.9572
THETA_HOH = np.deg2rad(104.52)
WATER_VOLUME = 29.9  # A^3 per molecule at 1 g/cm^3

TYPES = {"water03": (1, 2), "hippo19": (16, 17)}


def _molecule_frame():
    """O at the origin, the two H in the xz plane, bisector along +z."""
    s, c = np.sin(THETA_HOH / 2), np.cos(THETA_HOH / 2)
    return np.array([[0.0, 0.0, 0.0], [ROH * s, 0.0, ROH * c], [-ROH * s, 0.0, ROH * c]])


def random_rotations(rng: np.random.Generator, count: int) -> np.ndarray:
    """Uniform random rotation matrices from uniform unit quaternions (Shoemake 1992)."""
    u1, u2, u3 = rng.random((3, count))
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    qx, qy = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    qz, qw = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    r = np.empty((count, 3, 3))
    r[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    r[:, 0, 1] = 2 * (qx * qy - qz * qw)
    r[:, 0, 2] = 2 * (qx * qz + qy * qw)
    r[:, 1, 0] = 2 * (qx * qy + qz * qw)
    r[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz)
    r[:, 1, 2] = 2 * (qy * qz - qx * qw)
    r[:, 2, 0] = 2 * (qx * qz - qy * qw)
    r[:, 2, 1] = 2 * (qy * qz + qx * qw)
    r[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    return r


def lattice_box(k: int, L: float, seed: int, nmol: int | None = None) -> np.ndarray:
    """Coordinates (3*nmol, 3) of the seeded lattice water box, O,H,H per
    molecule.  nmol defaults to k^3 (every site); a smaller count leaves
    k^3 - nmol sites empty, chosen by the seed (the 32,000-molecule HIPPO box
    on a 32^3 lattice at 1 g/cm^3)."""
    rng = np.random.default_rng(seed)
    g = (np.arange(k) + 0.5) * (L / k)
    ox = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(k ** 3, 3)
    if nmol is not None and nmol < k ** 3:
        keep = np.sort(rng.choice(k ** 3, size=nmol, replace=False))
        ox = ox[keep]
    nmol = len(ox)
    rot = random_rotations(rng, nmol)
    jitter = rng.uniform(-0.1, 0.1, size=(nmol, 3))
    mol = np.einsum("mij,aj->mai", rot, _molecule_frame())
    xyz = mol + (ox + jitter)[:, None, :]
    return xyz.reshape(3 * nmol, 3)


def cluster(nmol: int, rng: np.random.Generator, min_oo: float = 2.5) -> np.ndarray:
    """One non-periodic cluster of nmol waters inside a 1 g/cm^3 sphere."""
    radius = (3.0 * nmol * WATER_VOLUME / (4.0 * np.pi)) ** (1.0 / 3.0)
    oxy = []
    while len(oxy) < nmol:
        p = rng.uniform(-radius, radius, size=3)
        if p @ p > radius * radius:
            continue
        if oxy and np.min(np.sum((np.array(oxy) - p) ** 2, axis=1)) < min_oo * min_oo:
            continue
        oxy.append(p)
    rot = random_rotations(rng, nmol)
    mol = np.einsum("mij,aj->mai", rot, _molecule_frame())
    return (mol + np.array(oxy)[:, None, :]).reshape(3 * nmol, 3)


def cluster_batch(nclusters: int, seed: int, min_mol: int = 2, max_mol: int = 64):
    """Independent clusters of U{min_mol..max_mol} waters (seeded): the
    concatenated coordinates and the atom offsets of each cluster.  Each
    cluster sits at its own origin; clusters never interact."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(min_mol, max_mol + 1, size=nclusters)
    xyz = [cluster(int(m), rng) for m in sizes]
    off = np.concatenate([[0], np.cumsum(3 * sizes)]).astype(np.int32)
    return np.concatenate(xyz), off


def topology(nmol: int, ff: str = "water03"):
    """Atom types and 0-based bond list (O-H1, O-H2) for nmol waters."""
    to, th = TYPES[ff]
    types = np.tile(np.array([to, th, th], dtype=np.int32), nmol)
    o = 3 * np.arange(nmol)
    bonds = np.concatenate([np.stack([o, o + 1], 1), np.stack([o, o + 2], 1)])
    return types, bonds


def write_tinker_xyz(path: str, xyz: np.ndarray, ff: str = "water03", title: str = "water") -> None:
    """Tinker .xyz file (readxyz.cpp:38 format) for a water system."""
    n = len(xyz)
    nmol = n // 3
    types, _ = topology(nmol, ff)
    lines = [f"{n:6d}  {title}"]
    for m in range(nmol):
        i = 3 * m
        o, h1, h2 = i + 1, i + 2, i + 3
        lines.append(f"{o:6d}  O  {xyz[i,0]:14.8f}{xyz[i,1]:14.8f}{xyz[i,2]:14.8f}{types[i]:6d}{h1:6d}{h2:6d}")
        lines.append(f"{h1:6d}  H  {xyz[i+1,0]:14.8f}{xyz[i+1,1]:14.8f}{xyz[i+1,2]:14.8f}{types[i+1]:6d}{o:6d}")
        lines.append(f"{h2:6d}  H  {xyz[i+2,0]:14.8f}{xyz[i+2,1]:14.8f}{xyz[i+2,2]:14.8f}{types[i+2]:6d}{o:6d}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")



# ---------------------------------------------------------------------------
# parameter assignment for water systems (what mechanic() leaves in the
# reference's modules for these molecules)

BOHR = 0.529177210903          # units.h:75
COULOMB = 332.0637133          # units.h:79 (electric)
FRAME_ZTHENX, FRAME_BISECTOR = 1, 3
DATA = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "data")


def _params(ff: str) -> dict:
    from . import prm

    return prm.load(__import__("os").path.join(DATA, f"{ff}.json"))


def water_system(xyz: np.ndarray, ff: str = "water03", box: float | None = None, *, ewald: dict | None = None,
                 polar: bool = True, cutoff: float | None = None) -> dict:
    """System dict (the reference's global arrays) for rigid waters O,H,H.

    Frames as kmpole assigns them for water (kmpole.cpp:290-441): O is a
    bisector site with z = first H and x = second H, each H is Z-then-X with
    z = O and x = the other H.  Multipoles in kmpole's units (dipole x bohr,
    quadrupole x bohr^2/3, kmpole.cpp:543-552); one polarization group per
    molecule (kpolar.cpp:592-853); Thole a and pdamp = alpha^(1/6)
    (kpolar.cpp:461-476).  `ewald` = {aewald, nfft, order, cutoff} switches on
    Ewald summation for a cubic box of side `box`.
    """
    P = _params(ff)
    xyz = np.asarray(xyz, dtype=np.float64)
    n = len(xyz)
    nmol = n // 3
    to, th = TYPES[ff]
    mp = {m["ids"][0]: m for m in P["multipole"]}
    pole = np.zeros((n, 13))
    for t, rows in ((to, slice(0, None, 3)), (th, None)):
        m = mp[t]
        v = np.concatenate([[m["c"]], np.array(m["d"]) * BOHR, np.array(m["q"]) * BOHR * BOHR / 3.0])
        if rows is not None:
            pole[rows] = v
        else:
            pole[1::3] = v
            pole[2::3] = v
    idx = np.arange(n)
    o = 3 * np.arange(nmol)
    polaxe = np.full(n, FRAME_ZTHENX, dtype=np.int32)
    polaxe[o] = FRAME_BISECTOR
    zaxis = np.empty(n, dtype=np.int32)
    xaxis = np.empty(n, dtype=np.int32)
    zaxis[o], xaxis[o] = o + 2, o + 3            # 1-based
    zaxis[o + 1], xaxis[o + 1] = o + 1, o + 3
    zaxis[o + 2], xaxis[o + 2] = o + 1, o + 2
    yaxis = np.zeros(n, dtype=np.int32)
    n12 = np.where(idx % 3 == 0, 2, 1).astype(np.int32)
    i12 = np.empty(2 * n - 2 * nmol + 2 * nmol, dtype=np.int32)
    i12 = np.concatenate([np.array([[a + 1, a + 2], [a, -1], [a, -1]]) for a in o]).ravel()
    i12 = i12[i12 >= 0].astype(np.int32)
    n13 = np.where(idx % 3 == 0, 0, 1).astype(np.int32)
    i13 = np.stack([o + 2, o + 1], 1).ravel().astype(np.int32)
    zero = np.zeros(n, dtype=np.int32)
    grp = np.repeat(o, 3)
    ip11 = (grp[:, None] + np.arange(3)[None, :]).ravel().astype(np.int32)
    types = np.where(idx % 3 == 0, to, th).astype(np.int32)
    cls = np.array([P["atom"][int(t)]["class"] for t in types])
    s = {
        "n": n, "x": xyz[:, 0].copy(), "y": xyz[:, 1].copy(), "z": xyz[:, 2].copy(),
        "type": types, "pole": pole.ravel(), "polaxe": polaxe, "zaxis": zaxis, "xaxis": xaxis, "yaxis": yaxis,
        "use": np.concatenate([[0], np.ones(n, dtype=np.int32)]).astype(np.int32),
        "molcule": (idx // 3).astype(np.int32),
        "n12": n12, "i12": i12, "n13": n13, "i13": i13, "n14": zero, "i14": np.zeros(0, np.int32),
        "n15": zero, "i15": np.zeros(0, np.int32),
        "np11": np.full(n, 3, dtype=np.int32), "ip11": ip11,
        "np12": zero, "ip12": np.zeros(0, np.int32), "np13": zero, "ip13": np.zeros(0, np.int32),
        "np14": zero, "ip14": np.zeros(0, np.int32),
        "electric": COULOMB, "dielec": float(P["keywords"].get("dielectric", 1.0)),
        "pentype": 0, "use_chgflx": 0, "nbond": 0, "nangle": 0,
        "ibnd": np.zeros(0, np.int32), "iang": np.zeros(0, np.int32), "balist": np.zeros(0, np.int32),
        "bl": np.zeros(0), "bflx": np.zeros(0), "anat": np.zeros(0), "aflx": np.zeros(0), "abflx": np.zeros(0),
        "mono0": pole[:, 0].copy(), "pcore": np.zeros(n), "pval0": pole[:, 0].copy(), "palpha": np.zeros(n),
    }
    sc = {"m2scale": 0.0, "m3scale": 0.0, "m4scale": 1.0, "m5scale": 1.0, "p2scale": 0.0, "p3scale": 0.0,
          "p4scale": 1.0, "p5scale": 1.0, "p2iscale": 0.0, "p3iscale": 0.0, "p4iscale": 0.5, "p5iscale": 1.0,
          "d1scale": 0.0, "d2scale": 1.0, "d3scale": 1.0, "d4scale": 1.0,
          "u1scale": 1.0, "u2scale": 1.0, "u3scale": 1.0, "u4scale": 1.0}
    sc.update({k: v for k, v in P["scales"].items() if k in sc})
    s.update(sc)
    # polarizability (by type, or by class for HIPPO files)
    # POLARIZE: polarizability, then Thole a when the next field is not an
    # integer group type (readprm.cpp:1250-1275)
    def thole_of(v):
        return v[1] if len(v) > 1 and float(v[1]) != int(v[1]) else 0.0
    pol = np.array([P["polarize"][int(t)][0] for t in types])
    thl = np.array([thole_of(P["polarize"][int(t)]) for t in types])
    s["polarity"], s["thole"], s["pdamp"] = pol, thl, pol ** (1.0 / 6.0)
    s["poleps"], s["politer"] = 1e-6, 100
    s["use_polar"] = 1 if polar else 0
    if P["chgpen"]:
        cp = np.array([P["chgpen"][int(c)] for c in cls])
        s["pentype"] = 1
        s["pcore"], s["palpha"] = cp[:, 0].copy(), cp[:, 1].copy()
        s["pval0"] = pole[:, 0] - cp[:, 0]
    # HIPPO dispersion (csix, adisp by class: kdisp) and Pauli repulsion (size,
    # damping exponent, valence electrons by class: krepel)
    n0 = np.zeros(n)
    s.update({"csix": n0.copy(), "adisp": n0.copy(), "sizpr": n0.copy(), "dmppr": n0.copy(), "elepr": n0.copy()})
    # defaults of initprm.cpp:332-341,390-393
    for k in ("dsp2scale", "dsp3scale", "dsp4scale", "dsp5scale", "r2scale", "r3scale", "r4scale", "r5scale"):
        s[k] = float(P["scales"].get(k, 1.0 if k[-6] in "45" else 0.0))
    for k in ("w2scale", "w3scale", "w4scale", "w5scale"):
        s[k] = float(P["scales"].get(k, 1.0))
    # d-equals-p: the direct-field scales follow the p scales (prmkey.cpp:531)
    s["dpequal"] = 1 if P["keywords"].get("d-equals-p") else 0
    if P.get("dispersion"):
        dp = np.array([P["dispersion"][int(c)] for c in cls])
        s["csix"], s["adisp"] = dp[:, 0].copy(), dp[:, 1].copy()
    if P.get("repulsion"):
        rp = np.array([P["repulsion"][int(c)] for c in cls])
        # valence electron counts are stored negative (readprm.cpp:1042)
        s["sizpr"], s["dmppr"], s["elepr"] = rp[:, 0].copy(), rp[:, 1].copy(), -np.abs(rp[:, 2])
    # box and method
    if box is None:
        s.update({"use_bounds": 0, "xbox": 0.0, "ybox": 0.0, "zbox": 0.0, "alpha": 90.0, "beta": 90.0,
                  "gamma": 90.0, "orthogonal": 0, "monoclinic": 0, "triclinic": 0, "octahedron": 0,
                  "dodecadron": 0})
    else:
        s.update({"use_bounds": 1, "xbox": float(box), "ybox": float(box), "zbox": float(box), "alpha": 90.0,
                  "beta": 90.0, "gamma": 90.0, "orthogonal": 1, "monoclinic": 0, "triclinic": 0,
                  "octahedron": 0, "dodecadron": 0})
    s["mpolecut"] = float(cutoff) if cutoff is not None else (9.0 if box is not None else 1e12)
    if ewald:
        g = ewald["nfft"]
        s.update({"use_ewald": 1, "aewald": float(ewald["aewald"]), "ewaldcut": float(ewald.get("cutoff", 7.0)),
                  "nfft1": int(g), "nfft2": int(g), "nfft3": int(g), "bsorder": int(ewald.get("order", 5))})
        # dispersion PME on the same grid (dewald: its Gaussian coefficient)
        s.update({"use_dewald": 1 if "dewald" in ewald else 0, "dewald": float(ewald.get("dewald", 0.0))})
    else:
        s.update({"use_ewald": 0, "aewald": 0.0, "ewaldcut": 7.0, "nfft1": 0, "nfft2": 0, "nfft3": 0, "bsorder": 5,
                  "use_dewald": 0, "dewald": 0.0})
    s["dispcut"] = float(ewald.get("cutoff", 7.0)) if ewald else s["mpolecut"]
    # repulsion cutoff: 6 A by default (cutoffs.cpp:57), kept inside the
    # electrostatic cutoff
    s["repcut"] = min(6.0, float(ewald.get("cutoff", 7.0)) if ewald else s["mpolecut"])
    return s
