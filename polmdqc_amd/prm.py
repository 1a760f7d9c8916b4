"""Reader for the subset of Tinker .prm/.key keywords the electrostatics path reads.

Follows the keyword semantics of the reference's parameter reader
(src/md/readprm.cpp:64; MULTIPOLE :1157, POLARIZE :1250, CHGPEN :1237,
REPULSION :1032, DISPERSION :1047) and of prmkey.cpp:38 for the scale
factors.  Only the fields the electrostatics terms use are kept.
"""
from __future__ import annotations

import json
import re

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


def parse_prm(text: str, types: set[int] | None = None) -> dict:
    """Parse a .prm text.  `types`, when given, keeps only those atom types
    (and the classes they map to)."""
    lines = text.splitlines()
    out = {"atom": {}, "multipole": [], "polarize": {}, "chgpen": {}, "repulsion": {}, "dispersion": {},
           "scales": {}, "keywords": {}}
    i = 0
    while i < len(lines):
        raw = lines[i]
        line = raw.split("!!")[0].strip()
        i += 1
        if not line or line.startswith("#"):
            continue
        w = line.split()
        key = w[0].lower()
        if key in ("atom", "multipole", "polarize", "chgpen", "repulsion", "dispersion") and (
                len(w) < 2 or not re.fullmatch(r"-?\d+", w[1])):
            continue  # prose in the file's literature section
        if key == "atom":
            t, cls = int(w[1]), int(w[2])
            m = re.search(r'"[^"]*"\s+(\d+)\s+([-\d.]+)\s+(\d+)', line)
            out["atom"][t] = {"class": cls, "atomic": int(m.group(1)), "mass": float(m.group(2)),
                              "valence": int(m.group(3))}
        elif key == "multipole":
            # header: type, up to three frame types, charge; then 4 continuation lines
            nums = w[1:]
            ids = [int(v) for v in nums[:-1]]
            charge = float(nums[-1])
            more = []
            while len(more) < 9 and i < len(lines):
                more += [float(v) for v in lines[i].split()]
                i += 1
            d = more[0:3]
            qxx, qxy, qyy, qxz, qyz, qzz = more[3], more[4], more[5], more[6], more[7], more[8]
            out["multipole"].append({"ids": ids, "c": charge, "d": d,
                                     "q": [qxx, qxy, qxz, qxy, qyy, qyz, qxz, qyz, qzz]})
        elif key == "polarize":
            t = int(w[1])
            vals = [float(v) for v in w[2:]]
            out["polarize"][t] = vals
        elif key in ("chgpen", "repulsion", "dispersion"):
            out[key][int(w[1])] = [float(v) for v in w[2:]]
        elif key in SCALE_KEYS:
            out["scales"][SCALE_KEYS[key]] = float(w[1])
        elif key in ("polarization", "dielectric", "d-equals-p", "forcefield"):
            out["keywords"][key] = w[1] if len(w) > 1 else True
    if types is not None:
        classes = {out["atom"][t]["class"] for t in types}
        out["atom"] = {t: v for t, v in out["atom"].items() if t in types}
        out["multipole"] = [m for m in out["multipole"] if m["ids"][0] in types]
        # POLARIZE is keyed by atom type (kpolar.cpp:469-471); CHGPEN,
        # REPULSION and DISPERSION by atom class (kmpole.cpp:645-649,
        # krepel.cpp:98-103)
        out["polarize"] = {t: v for t, v in out["polarize"].items() if t in types}
        for k in ("chgpen", "repulsion", "dispersion"):
            out[k] = {t: v for t, v in out[k].items() if t in classes}
    return out


def save(d: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)


def load(path: str) -> dict:
    with open(path) as f:
        d = json.load(f)
    for k in ("atom", "polarize", "chgpen", "repulsion", "dispersion"):
        d[k] = {int(t): v for t, v in d[k].items()}
    return d
