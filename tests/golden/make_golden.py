"""Regenerate the golden fixtures in tests/golden/ from the reference.

Needs /root/reference and the compiled reference checker (``make -f
oracle/ref.mk``); neither exists on the GPU test machines, which only read
the fixtures written here.

Two kinds of fixtures:

* ``ref_headers.json`` -- the golden vectors the reference's own Catch2 tests
  hold (test/amoeba.h, amoebap.h, hippo.h, rotpole.h, chkpole.h, active.h,
  bounds.h, amoebaMisc.h), parsed from those headers as data.
* ``<system>.npz`` -- for each reference test system (and for two synthetic
  water boxes in non-Ewald form), every array the multipole path reads as the
  reference's setup produced it, plus what the reference's energy<Analysis>()
  and energy<Virial>() computed (em, nem, einter, aem, rpole, dem, vir).
  Written by oracle/ref_dump.cpp, run pinned to one core.

    python tests/golden/make_golden.py [--only NAME ...]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
DUMP = os.path.join(REPO, "oracle", "_ref", "ref_dump")
RUNDIR = os.path.join(REPO, "oracle", "_ref", "run", "a", "b")
sys.path.insert(0, REPO)

HEADERS = ["amoeba", "amoebap", "hippo", "rotpole", "chkpole", "active", "bounds", "amoebaMisc"]

T = "../../test/testFiles/"
# name -> (xyz, key or None, gradient kept in full?)
SYSTEMS = {
    "amoeba_water09": (T + "amoeba/water09.xyz", None, True),
    "amoebap_waterap": (T + "amoebap/waterap.xyz", None, True),
    "hippo_water21": (T + "hippo/water21.xyz", None, True),
    "rotpole_water09_Na": (T + "rotpole/water09_Na.xyz", None, True),
    "rotpole_benzene_ethyne_water": (T + "rotpole/benzene_ethyne_water.xyz", None, True),
    "rotpole_ammonia": (T + "rotpole/ammonia.xyz", None, True),
    "chkpole_alatet": (T + "chkpole/alatet.xyz", None, True),
    "misc_alatet_water09": (T + "amoebaMisc/alatet_water09.xyz", None, True),
    "misc_water09NSeq": (T + "amoebaMisc/water09NSeq.xyz", None, True),
    "misc_water09_2": (T + "amoebaMisc/water09.xyz_2", T + "amoebaMisc/water09.key", True),
    "active_water09_Na_Cls_active": (T + "active/water09_Na_Cls.xyz", T + "active/active.key", True),
    "active_water09_Na_Cls_inactive": (T + "active/water09_Na_Cls.xyz", T + "active/inactive.key", True),
    "active_waterbox30": (T + "active/waterbox30.xyz", None, False),
    "bounds_orthogonal": (T + "bounds/waterbox30.xyz", T + "bounds/orthogonal.key", True),
    "bounds_monoclinic": (T + "bounds/waterbox30.xyz", T + "bounds/monoclinic.key", False),
    "bounds_triclinic": (T + "bounds/waterbox30.xyz", T + "bounds/triclinic.key", True),
    "bounds_octahedron": (T + "bounds/waterbox30.xyz", T + "bounds/octahedron.key", True),
    "bounds_dodecahedron": (T + "bounds/waterbox30.xyz", T + "bounds/dodecahedron.key", False),
}
# synthetic water boxes (non-Ewald multipoles, 7 A cutoff):
# name -> (k, L, seed, full gradient, parameter file)
SYNTH = {
    "synth_water216": (6, 18.643, 216, True, "water03"),
    "synth_water4096": (16, 49.68, 4096, False, "water03"),
    "synth_hippo27": (3, 9.32, 27, True, "hippo19"),
}
SAMPLE = 300  # atoms whose gradient is kept when the full gradient is not


def parse_header(path: str) -> dict:
    """namespace -> {name: value} for `real/int x = v;` and `std::vector<real> x = {..};`."""
    text = re.sub(r"namespace\s+polmdqc\s*\{", "", open(path).read(), count=1)
    out = {}
    for m in re.finditer(r"namespace\s+(\w+)\s*\{(.*?)\n\}", text, re.S):
        ns, body = m.group(1), m.group(2)
        if ns == "polmdqc":
            continue
        d = {}
        tok = {"true": 1.0, "false": 0.0}
        for v in re.finditer(r"std::vector<\s*(?:real|int|bool)\s*>\s+(\w+)\s*=\s*\{(.*?)\};", body, re.S):
            d[v.group(1)] = [tok.get(t.strip(), None) if t.strip() in tok else float(t)
                             for t in v.group(2).replace("\n", " ").split(",") if t.strip()]
        for v in re.finditer(r"std::vector<std::vector<real>>\s+(\w+)\s*=\s*\{(.*?)\n\s*\};", body, re.S):
            rows = re.findall(r"\{([^{}]*)\}", v.group(2))
            d[v.group(1)] = [[float(t) for t in r.split(",") if t.strip()] for r in rows]
        for v in re.finditer(r"^\s*(?:real|int|bool)\s+(\w+)\s*=\s*([^;{]+);", body, re.M):
            t = v.group(2).strip()
            d[v.group(1)] = tok[t] if t in tok else float(t)
        out[ns] = d
    return out


def write_headers() -> None:
    gold = {}
    for h in HEADERS:
        gold.update(parse_header(os.path.join(REF, "test", h + ".h")))
    with open(os.path.join(HERE, "ref_headers.json"), "w") as f:
        json.dump(gold, f, indent=1, sort_keys=True)
    print("ref_headers.json:", len(gold), "namespaces")


def run_dump(xyz: str, key: str | None) -> dict:
    with tempfile.NamedTemporaryFile(suffix=".json", delete=False) as t:
        out = t.name
    cmd = ["taskset", "-c", "0", DUMP, out, xyz] + (["-k", key] if key else [])
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run(cmd, cwd=RUNDIR, env=env, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    d = json.load(open(out))
    os.unlink(out)
    return d


def save(name: str, d: dict, full_grad: bool) -> None:
    arrays = {}
    for k, v in d.items():
        if isinstance(v, str):
            arrays[k] = np.array(v)
        elif isinstance(v, list):
            a = np.array(v)
            arrays[k] = a.astype(np.int32) if a.dtype.kind == "i" else a.astype(np.float64)
        else:
            arrays[k] = np.array(v, dtype=np.float64 if isinstance(v, float) else np.int64)
    n = int(d["n"])
    if not full_grad:
        idx = np.sort(np.random.default_rng(12345).choice(n, size=min(SAMPLE, n), replace=False))
        for k in ("dem", "desum"):
            arrays[k] = arrays[k].reshape(n, 3)[idx].ravel()
        arrays["grad_atoms"] = idx.astype(np.int32)
    else:
        arrays["grad_atoms"] = np.arange(n, dtype=np.int32)
    if n > 1000:
        # large boxes: keep the rotated multipoles and the per-atom energy
        # partition on a seeded sample of atoms only
        rows = np.sort(np.random.default_rng(54321).choice(n, size=SAMPLE, replace=False))
        arrays["rpole"] = arrays["rpole"].reshape(n, 13)[rows].ravel()
        arrays["aem"] = arrays["aem"][rows]
        arrays["sample_atoms"] = rows.astype(np.int32)
    else:
        arrays["sample_atoms"] = np.arange(n, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print(f"{name}: n={n} em={d['em']:.10f} nem={d['nem']}")


def synth(name: str, k: int, L: float, seed: int, full_grad: bool, ff: str) -> None:
    from polmdqc_amd import water

    tmp = tempfile.mkdtemp(prefix="pmq_golden_")
    xyzf = os.path.join(tmp, name + ".xyz")
    keyf = os.path.join(tmp, name + ".key")
    water.write_tinker_xyz(xyzf, water.lattice_box(k, L, seed), ff, name)
    # multipoles only: with polarization on, kmpole.cpp:704-726 decides the
    # penetration type before kpolar counts the charge-penetration sites, so
    # a HIPPO file would run without charge penetration
    terms = "multipoleterm only\n"
    with open(keyf, "w") as f:
        f.write(f"parameters {REF}/params/{ff}.prm\ndigits 8\n{terms}"
                f"a-axis {L}\nmpole-cutoff 7.0\n")
    save(name, run_dump(xyzf, keyf), full_grad)


def write_water_params() -> None:
    """Water parameter tables the package's water builder reads (data
    extracted from the reference's params/*.prm)."""
    from polmdqc_amd import prm

    for ff, types in (("water03", {1, 2}), ("hippo19", {16, 17})):
        d = prm.parse_prm(open(os.path.join(REF, "params", ff + ".prm")).read(), types)
        prm.save(d, os.path.join(REPO, "polmdqc_amd", "data", ff + ".json"))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    args = ap.parse_args()
    if not os.path.exists(DUMP):
        sys.exit("build the reference checker first: make -f oracle/ref.mk")
    write_headers()
    for name, (xyz, key, full) in SYSTEMS.items():
        if args.only and name not in args.only:
            continue
        save(name, run_dump(xyz, key), full)
    for name, (k, L, seed, full, ff) in SYNTH.items():
        if args.only and name not in args.only:
            continue
        synth(name, k, L, seed, full, ff)
    write_water_params()


if __name__ == "__main__":
    main()
