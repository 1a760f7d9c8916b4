"""Parameter gradients and shapes of the batched clusters (one persistent
launch per size bucket, graph-captured).

- Every cluster's parameter gradients from the batch equal those of the
  one-system path on that cluster alone (FP64, 1e-9 of the largest), and the
  mixed form's are within 1e-4 of them.
- Central finite differences of each sampled cluster's own energy on the
  512-cluster U{2..64} batch (32 samples over per-atom multipole components,
  polarizabilities with pdamp = alpha^(1/6), and Thole values made distinct
  per atom so no pair sits at the min(a_i, a_k) kink), bar 1e-5 relative plus
  the difference quotient's resolution 1e-11 |E| / h, PCG to 1e-12 D.
- The energy is quadratic in the multipoles: sum(pole * dE/dpole) = 2 E.
- Every shape the autotuner may choose (lanes per row, workgroups per CU, per
  size bucket) gives the same energies, forces and parameter gradients, and
  the tuner records one shape per bucket that holds clusters.
"""
import numpy as np
import pytest

from polmdqc_amd import water

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from polmdqc_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def batch(nclusters, seed, min_mol=2, max_mol=12):
    xyz, off = water.cluster_batch(nclusters, seed, min_mol, max_mol)
    s = water.water_system(xyz, "water03")
    s["pcgprec"] = 0  # the polarizability preconditioner (the sparse one: test_gpu_batch.py)
    return s, off


def sub(s, a, b):
    """Atoms a:b of a water system as a system of their own (same parameters)."""
    t = water.water_system(np.stack([s["x"][a:b], s["y"][a:b], s["z"][a:b]], 1), "water03")
    for k in ("pole", "polarity", "pdamp", "thole"):
        v = np.asarray(s[k], dtype=float)
        t[k] = (v.reshape(-1, 13)[a:b].ravel() if k == "pole" else v[a:b]).copy()
    t["pcgprec"] = 0
    t["poleps"] = s["poleps"]
    return t


def test_batch_param_grad_equals_one_system_path(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] batch parameter gradients vs one-system path, 8 clusters", flush=True)
    s, off = batch(8, 17, 2, 20)
    s["poleps"] = 1e-10
    engine.set_system(s)
    engine.set_batch(off)
    engine.set_batch_param_grad(True)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR, L.PLAIN_FP64)
    pb = engine.param_grad()
    engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR, L.MIXED)
    pm = engine.param_grad()
    # sum(pole * dE/dpole) = 2 E over the whole batch
    lhs = float(np.sum(np.asarray(s["pole"]).reshape(-1, 13) * pb["dpole"]))
    assert abs(lhs - 2.0 * (g["em"] + g["ep"])) <= 1e-9 * abs(g["em"] + g["ep"])
    for c in range(len(off) - 1):
        lo, hi = off[c], off[c + 1]
        engine.set_system(sub(s, lo, hi))
        engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR, L.PLAIN_FP64)
        one = engine.param_grad()
        for k in ("dpole", "dpolarity", "dpdamp", "dthole"):
            ref = one[k]
            sc = max(np.abs(ref).max(), 1e-300)
            assert np.max(np.abs(pb[k][lo:hi] - ref)) <= 1e-9 * sc, (c, k)
            assert np.max(np.abs(pm[k][lo:hi] - ref)) <= 1e-4 * sc, (c, k)


def test_batch_param_grad_finite_differences_512(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] batch parameter FD on 512 clusters", flush=True)
    s, off = batch(512, 512, 2, 64)
    s["poleps"] = 1e-12
    n = s["n"]
    rng = np.random.default_rng(5)
    s["thole"] = np.asarray(s["thole"], dtype=float) * (1.0 + 0.2 * rng.random(n))
    # fixed shapes: every energy of the differences comes from the same kernels
    for b in range(3):
        engine.set_batch_shape(L.PLAIN_FP64, b, 1, 2)

    def run(t):
        engine.set_system(t)
        engine.set_batch(off)
        engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR, L.PLAIN_FP64)
        return engine.fetch_batch()

    engine.set_batch_param_grad(True)
    b0 = run(s)
    assert np.all(b0["pcg_iters"] > 0)
    pg = engine.param_grad()
    pole = np.asarray(s["pole"], dtype=float).reshape(-1, 13)
    alpha = np.asarray(s["polarity"], dtype=float)
    owner = np.repeat(np.arange(len(off) - 1), np.diff(off))
    picks = []
    for k in range(32):
        i = int(rng.integers(n))
        kind = ("pole", "polarity", "thole")[k % 3]
        comp = int(rng.choice([0, 1, 3, 4, 6, 8, 12])) if kind == "pole" else 0
        picks.append((kind, i, comp))
    for kind, i, comp in picks:
        if kind == "pole":
            v0 = pole[i, comp]
            h = 1e-6 * abs(v0) if v0 != 0 else 1e-6
            analytic = pg["dpole"][i, comp]
        elif kind == "polarity":
            v0 = alpha[i]
            h = 1e-3 * v0
            analytic = pg["dpolarity"][i] + pg["dpdamp"][i] * v0 ** (-5.0 / 6.0) / 6.0
        else:
            v0 = s["thole"][i]
            h = 1e-3 * v0
            analytic = pg["dthole"][i]
        es = []
        for sgn in (1, -1):
            t = dict(s)
            if kind == "pole":
                p = pole.copy()
                p[i, comp] += sgn * h
                t["pole"] = p.ravel()
            elif kind == "polarity":
                a = alpha.copy()
                a[i] += sgn * h
                t["polarity"] = a
                pd = np.asarray(s["pdamp"], dtype=float).copy()
                pd[i] = a[i] ** (1.0 / 6.0)
                t["pdamp"] = pd
            else:
                th = np.asarray(s["thole"], dtype=float).copy()
                th[i] += sgn * h
                t["thole"] = th
            bb = run(t)
            cl = owner[i]
            es.append(bb["em"][cl] + bb["ep"][cl])
        fd = (es[0] - es[1]) / (2 * h)
        e0 = abs(es[0]) + abs(es[1])
        assert abs(fd - analytic) <= 1e-5 * abs(fd) + 1e-11 * e0 / h, (kind, i, comp, fd, analytic)
    for b in range(3):
        engine.set_batch_shape(L.PLAIN_FP64, b, 0, 0)


def test_every_shape_gives_the_same_result(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] batch shapes", flush=True)
    # clusters in all three size buckets (2..80 waters: 6..240 atoms)
    s, off = batch(40, 23, 2, 80)
    engine.set_system(s)
    engine.set_batch(off)
    engine.set_batch_param_grad(True)
    terms = L.TERM_MPOLE | L.TERM_POLAR
    for prec in (L.PLAIN_FP64, L.MIXED):
        ref = engine.eval(L.GRADIENT, terms, prec)
        refp = engine.param_grad()
        tun = engine.batch_tuning(prec)
        sizes = np.diff(off)
        split = [int(np.sum(sizes > 128)), int(np.sum((sizes > 64) & (sizes <= 128))), int(np.sum(sizes <= 64))]
        # the buckets side by side, or all clusters in one merged launch
        assert [t["clusters"] for t in tun] in (split, [len(sizes), 0, 0])
        for t in tun:
            if t["clusters"]:
                assert t["lanes_per_row"] > 0 and t["wgs_per_cu"] > 0 and t["ms"] > 0
        lanes = {0: (1, 2), 1: (1, 2, 4), 2: (1, 2, 4, 8)}
        for b in range(3):
            for ln in lanes[b]:
                for w in (1, 2, 4):
                    engine.set_batch_shape(prec, b, ln, w)
                    g = engine.eval(L.GRADIENT, terms, prec)
                    p = engine.param_grad()
                    tol = 1e-11 if prec == L.PLAIN_FP64 else 1e-5
                    assert abs(g["em"] - ref["em"]) <= tol * abs(ref["em"]), (prec, b, ln, w)
                    assert abs(g["ep"] - ref["ep"]) <= tol * abs(ref["ep"]), (prec, b, ln, w)
                    sc = np.abs(ref["desum"]).max()
                    assert np.max(np.abs(g["desum"] - ref["desum"])) <= tol * sc, (prec, b, ln, w)
                    sc = np.abs(refp["dpole"]).max()
                    assert np.max(np.abs(p["dpole"] - refp["dpole"])) <= 10 * tol * sc, (prec, b, ln, w)
            engine.set_batch_shape(prec, b, 0, 0)
    engine.set_batch_param_grad(False)
