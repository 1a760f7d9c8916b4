"""Evaluations of one system replayed from a HIP graph (pmq_set_graphs, off
by default) against the same evaluations enqueued launch by launch.

After the safe first evaluation, the second is enqueued plainly, the third
is captured and later ones replay the graph.  Every replay is checked, with
other GPU work (a second context's evaluation) and host synchronizations
between replays, at new coordinates (the graph reads them from the device)
and after a change of box (a new key: enqueued plainly, then captured
again).  Two cases add the van der Waals term (its own pair structure
built and replaced inside the graph) and charge transfer.  The mixed form's gradients are fixed-point sums, so they must match
exactly; energy totals and FP64 gradients are summed with FP64 atomics and
match to 1e-12.  Covered: AMOEBA with PME and the
sparse preconditioner (fused SpMV + spreading), HIPPO with all four terms,
a cluster without PME (the row product ends the operator), the 30 A AMOEBA
water box read from its input files with van der Waals, and a 64-water
HIPPO box with charge transfer beside its other terms.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from polmdqc_amd import tinker, water

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engines():
    from polmdqc_amd.engine import Engine

    a, b = Engine(0), Engine(0)
    yield a, b
    a.close()
    b.close()


def amoeba_box():
    return water.water_system(water.lattice_box(6, 18.643, 216), "water03", box=18.643,
                              ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})


def hippo_box():
    return water.water_system(water.lattice_box(6, 18.643, 216), "hippo19", box=18.643,
                              ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0, "dewald": 0.45})


def cluster():
    return water.water_system(water.cluster(40, np.random.default_rng(4)), "water03")


def tinker_vdw_box(tmp):
    td = os.path.join(GOLDEN, "tinker")
    key = os.path.join(tmp, "vdw.key")
    open(key, "w").write("parameters water03\na-axis 30\newald\newald-cutoff 7.0\nvdw-cutoff 9.0\n")
    s = tinker.system_from_files(os.path.join(td, "bounds/waterbox30.xyz"), key, [os.path.join(td, "params")])
    s["use_vdw"] = 1
    return s


def hippo_ct_box(tmp):
    from test_gpu_vdwct import hippo_box
    import pathlib

    s = hippo_box(pathlib.Path(tmp))
    s["use_chgflx"] = 0
    return s


def keys(terms):
    from polmdqc_amd import _lib as L

    k = ["em", "dem"]
    if terms & L.TERM_POLAR:
        k += ["ep", "dep", "pcg_iters"]
    if terms & L.TERM_REPEL:
        k += ["er"]
    if terms & L.TERM_DISP:
        k += ["edsp"]
    if terms & L.TERM_VDW:
        k += ["ev", "dev"]
    if terms & L.TERM_CTRN:
        k += ["ect", "dect"]
    return k + ["desum"]


def agree(g, e, prec, terms):
    from polmdqc_amd import _lib as L

    for k in keys(terms):
        a, b = np.asarray(g[k], dtype=float), np.asarray(e[k], dtype=float)
        if k == "pcg_iters" or (prec == L.MIXED and a.ndim > 0):
            # the mixed form's fixed-point gradient sums do not depend on order
            assert np.array_equal(a, b), k
        else:
            # energy totals (and FP64 gradients) are summed with FP64 atomics
            assert np.max(np.abs(a - b)) <= 1e-12 * max(np.max(np.abs(b)), 1e-300), k


def eager(b, s, mode, terms, prec, xyz, box=None):
    from polmdqc_amd.engine import set_graphs

    set_graphs(b, False)
    b.set_system(s)
    for _ in range(3):  # safe, then speculative
        b.set_coords(xyz, box)
        b.compute(mode, terms, prec)
        r = b.fetch()
    return r


@pytest.mark.parametrize("name", ["amoeba", "hippo", "cluster", "amoeba_vdw", "hippo_ct"])
def test_graph_replays_match_plain_enqueue(engines, name, tmp_path):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import box_from, set_graphs

    print(f"[gpu] graphs {name}", flush=True)
    a, b = engines
    if name == "amoeba_vdw":
        s = tinker_vdw_box(str(tmp_path))
    elif name == "hippo_ct":
        s = hippo_ct_box(str(tmp_path))
    else:
        s = {"amoeba": amoeba_box, "hippo": hippo_box, "cluster": cluster}[name]()
    terms = L.TERM_MPOLE | L.TERM_POLAR
    if name in ("hippo", "hippo_ct"):
        terms |= L.TERM_REPEL | L.TERM_DISP
    if name == "amoeba_vdw":
        terms |= L.TERM_VDW
    if name == "hippo_ct":
        terms |= L.TERM_CTRN
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    rng = np.random.default_rng(7)
    moved = xyz + rng.normal(scale=0.02, size=xyz.shape)
    for prec in (L.PLAIN_FP64, L.MIXED):
        ref = eager(b, s, L.GRADIENT, terms, prec, xyz)
        ref_moved = eager(b, s, L.GRADIENT, terms, prec, moved)
        set_graphs(a, True)
        a.set_system(s)
        for k in range(6):
            a.set_coords(xyz)
            a.compute(L.GRADIENT, terms, prec)
            g = a.fetch()
            agree(g, ref, prec, terms)
            # other GPU work and a host sync between replays
            b.set_coords(moved)
            b.compute(L.GRADIENT, terms, prec)
            b.sync()
        for k in range(3):
            a.set_coords(moved)
            a.compute(L.GRADIENT, terms, prec)
            agree(a.fetch(), ref_moved, prec, terms)
        if s["use_bounds"]:
            # a new box: a new key (enqueued plainly, then captured again)
            bx = box_from(s)
            bx.xbox *= 1.01
            ref_box = eager(b, s, L.GRADIENT, terms, prec, xyz, bx)
            for k in range(4):
                a.set_coords(xyz, bx)
                a.compute(L.GRADIENT, terms, prec)
                agree(a.fetch(), ref_box, prec, terms)
            a.set_coords(xyz, box_from(s))
