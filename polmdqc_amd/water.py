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


def lattice_box(k: int, L: float, seed: int) -> np.ndarray:
    """Coordinates (3*k^3, 3) of the seeded lattice water box, O,H,H per molecule."""
    rng = np.random.default_rng(seed)
    nmol = k ** 3
    g = (np.arange(k) + 0.5) * (L / k)
    ox = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(nmol, 3)
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

