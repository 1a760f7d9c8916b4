"""The measured workloads at their own sizes on the GPU.

* 4,096 AMOEBA waters (12,288 atoms, L = 49.68 A, PME 48^3 order 5, alpha
  0.5446, 7 A cutoff): plain FP64 against the oracle (PCG to 1e-10 D on
  both sides), equal PCG iteration counts at the production 1e-6 D, and the
  mixed form's RMS force and per-term energy errors against FP64.
* 32,000 HIPPO waters (96,000 atoms, L = 98.58 A, multipole and dispersion
  PME on 100^3 grids): the mixed form against FP64 per term, the net force
  of each term, and invariance of every term under a translation by one
  grid spacing on each axis (which maps the PME grid onto itself); plus a
  5,184-atom HIPPO box against the oracle term by term.
* Differentiation on the 4,096-water box: 32 coordinates and 32 per-type
  parameters (multipole components, polarizability, Thole a) by central
  finite differences of the GPU energy.
* 512 clusters of U{2..64} waters (seed 512): a seeded sample of 32 clusters
  against the oracle, iteration counts at 1e-6 D, and mixed against FP64 on
  every cluster.

Oracle bars as in the smaller parity tests: 1e-9 relative on energies and
the multipole gradient, 1e-8 on induced dipoles and the polarization
gradient (the two solvers stop at different points inside the 1e-10 D
tolerance).  The mixed form: RMS force error < 1e-3 kcal/mol/A and per-term
energy errors below 1e-5 relative.
"""
import numpy as np
import pytest

import oracle
from polmdqc_amd import water

pytestmark = pytest.mark.gpu
REL = 1e-9
RMS_F = 1e-3


@pytest.fixture(scope="module")
def engine():
    from polmdqc_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rms(a, b):
    return float(np.sqrt(np.mean(np.sum((np.asarray(a) - np.asarray(b)) ** 2, axis=1))))


def amoeba12k():
    xyz = water.lattice_box(16, 49.68, 4096)
    return water.water_system(xyz, "water03", box=49.68,
                              ewald={"aewald": 0.5446, "nfft": 48, "order": 5, "cutoff": 7.0})


def hippo96k():
    xyz = water.lattice_box(32, 98.58, 32000, 32000)
    return water.water_system(xyz, "hippo19", box=98.58,
                              ewald={"aewald": 0.5446, "nfft": 100, "order": 5, "cutoff": 7.0, "dewald": 0.5446})


def hippo5k():
    xyz = water.lattice_box(12, 37.26, 1728)
    return water.water_system(xyz, "hippo19", box=37.26,
                              ewald={"aewald": 0.5446, "nfft": 40, "order": 5, "cutoff": 7.0, "dewald": 0.45})


# ---------------------------------------------------------------------------
# 12,288-atom AMOEBA box

def test_amoeba12k_fp64_matches_oracle(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] amoeba 12288 fp64 vs oracle (poleps 1e-10)", flush=True)
    s = dict(amoeba12k())
    s["poleps"] = 1e-10
    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    o = oracle.elec(s, want_grad=True)
    print(f"  em {g['em']:.10f} / {o['em']:.10f}  ep {g['ep']:.10f} / {o['ep']:.10f}  "
          f"iters {g['pcg_iters']} / {o['iters']}  dem {rel(g['dem'], o['dem']):.2e}  "
          f"dep {rel(g['dep'], o['dep']):.2e}  uind {rel(g['uind'], o['uind']):.2e}", flush=True)
    assert abs(g["em"] - o["em"]) <= REL * abs(o["em"])
    assert abs(g["ep"] - o["ep"]) <= REL * abs(o["ep"])
    assert rel(g["dem"], o["dem"]) <= REL
    assert rel(g["uind"], o["uind"]) <= 1e-8
    assert rel(g["uinp"], o["uinp"]) <= 1e-8
    assert rel(g["dep"], o["dep"]) <= 1e-8


def test_amoeba12k_iterations_at_production_tolerance(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] amoeba 12288 pcg iterations at 1e-6 D", flush=True)
    s = amoeba12k()
    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    g2 = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)  # the speculative path
    o = oracle.elec(s, want_grad=False)
    assert g["pcg_iters"] == o["iters"] == g2["pcg_iters"]
    assert abs(g["ep"] - o["ep"]) <= 1e-6 * abs(o["ep"])
    # FP64 atomics add in a run-dependent order: equal to rounding
    assert abs(g2["ep"] - g["ep"]) <= 1e-12 * abs(g["ep"])


def test_amoeba12k_mixed_against_fp64(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] amoeba 12288 mixed vs fp64", flush=True)
    engine.set_system(amoeba12k())
    terms = L.TERM_MPOLE | L.TERM_POLAR
    d = engine.eval(L.GRADIENT, terms, L.PLAIN_FP64)
    m = engine.eval(L.GRADIENT, terms, L.MIXED)
    e_em, e_ep = abs(m["em"] - d["em"]) / abs(d["em"]), abs(m["ep"] - d["ep"]) / abs(d["ep"])
    f = rms(m["desum"], d["desum"])
    print(f"  rms force {f:.3e}  em rel {e_em:.3e}  ep rel {e_ep:.3e}  iters {m['pcg_iters']}/{d['pcg_iters']}",
          flush=True)
    assert f < RMS_F
    assert rms(m["dem"], d["dem"]) < RMS_F and rms(m["dep"], d["dep"]) < RMS_F
    assert e_em < 1e-5 and e_ep < 1e-5
    assert abs(m["pcg_iters"] - d["pcg_iters"]) <= 1


# ---------------------------------------------------------------------------
# 96,000-atom HIPPO box

HIPPO_TERMS = (("em", "dem"), ("ep", "dep"), ("er", "der"), ("edsp", "dedsp"))


def test_hippo96k_mixed_against_fp64_per_term(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] hippo 96000 mixed vs fp64 per term", flush=True)
    engine.set_system(hippo96k())
    terms = L.TERM_MPOLE | L.TERM_POLAR | L.TERM_REPEL | L.TERM_DISP
    d = engine.eval(L.GRADIENT, terms, L.PLAIN_FP64)
    m = engine.eval(L.GRADIENT, terms, L.MIXED)
    for e, g in HIPPO_TERMS:
        er = abs(m[e] - d[e]) / abs(d[e])
        fr = rms(m[g], d[g])
        print(f"  {e}: fp64 {d[e]:.6f}  rel err {er:.2e}  rms force err {fr:.2e}", flush=True)
        assert er < 1e-5, e
        assert fr < RMS_F, g
    assert rms(m["desum"], d["desum"]) < RMS_F
    assert abs(m["pcg_iters"] - d["pcg_iters"]) <= 1


def test_hippo96k_net_force_and_grid_translation(engine):
    """Real-space terms conserve momentum pair by pair (repulsion exactly);
    PME's B-spline interpolation leaves a small net force, bounded here per
    term against the term's RMS force (the oracle on a 5,184-atom HIPPO box
    with a 0.93 A grid gives |sum F| / (sqrt(N) F_rms) = 2.5e-3 for the
    multipoles, 3.8e-3 for polarization, 9e-5 for dispersion).  A rigid translation by one grid
    spacing on every axis maps the grid onto itself, so every term's energy
    and forces are unchanged up to rounding."""
    from polmdqc_amd import _lib as L

    print("[gpu] hippo 96000 net force and grid translation (fp64)", flush=True)
    s = hippo96k()
    engine.set_system(s)
    terms = L.TERM_MPOLE | L.TERM_POLAR | L.TERM_REPEL | L.TERM_DISP
    a = engine.eval(L.GRADIENT, terms)
    n = s["n"]
    for e, g in HIPPO_TERMS:
        net = np.abs(a[g].sum(0)).max()
        frms = float(np.sqrt(np.mean(np.sum(a[g] ** 2, 1))))
        print(f"  {g}: net {net:.3e}  rms {frms:.3e}", flush=True)
        bound = np.sqrt(n) * frms * {"dem": 2e-2, "dep": 2e-2, "der": 1e-12, "dedsp": 1e-3}[g]
        assert net <= bound, (g, net, bound)
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    h = 98.58 / 100.0
    b = engine.eval(L.GRADIENT, terms, xyz=xyz + np.array([h, -2 * h, 3 * h]))
    for e, g in HIPPO_TERMS:
        print(f"  {e}: translated rel {abs(b[e] - a[e]) / abs(a[e]):.2e}  grad {rel(b[g], a[g]):.2e}", flush=True)
        assert abs(b[e] - a[e]) <= 1e-10 * abs(a[e]), e
        assert rel(b[g], a[g]) <= 1e-8, g


def test_hippo5k_matches_oracle(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] hippo 5184 fp64 vs oracle", flush=True)
    s = dict(hippo5k())
    s["poleps"] = 1e-10
    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR | L.TERM_REPEL | L.TERM_DISP)
    o = oracle.elec(s, want_grad=True)
    h = oracle.hippo_terms(s, want_grad=True)
    print(f"  em {rel(g['em'], o['em']):.2e} ep {rel(g['ep'], o['ep']):.2e} er {rel(g['er'], h['er']):.2e} "
          f"edsp {rel(g['edsp'], h['edsp']):.2e}", flush=True)
    assert abs(g["em"] - o["em"]) <= REL * abs(o["em"])
    assert abs(g["ep"] - o["ep"]) <= REL * abs(o["ep"])
    assert abs(g["er"] - h["er"]) <= REL * abs(h["er"])
    assert abs(g["edsp"] - h["edsp"]) <= REL * abs(h["edsp"])
    assert rel(g["dem"], o["dem"]) <= REL
    assert rel(g["dep"], o["dep"]) <= 1e-8
    assert rel(g["der"], h["der"]) <= REL
    assert rel(g["dedsp"], h["dedsp"]) <= REL
    assert rel(g["uind"], o["uind"]) <= 1e-8


# ---------------------------------------------------------------------------
# differentiation on the 4,096-water box

def energy(engine, s):
    from polmdqc_amd import _lib as L

    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    return g["em"] + g["ep"]


def per_type(x, types):
    return {t: np.asarray(x)[types == t].sum(axis=0) for t in np.unique(types)}


def test_amoeba12k_positions_by_finite_differences(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] amoeba 12288 FD of 32 coordinates", flush=True)
    s = dict(amoeba12k())
    s["poleps"] = 1e-12
    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    rng = np.random.default_rng(32)
    worst = 0.0
    for _ in range(32):
        i, a = int(rng.integers(len(xyz))), int(rng.integers(3))
        es = []
        for sgn in (1, -1):
            x = xyz.copy()
            x[i, a] += sgn * 1e-5
            e = engine.eval(L.ENERGY, L.TERM_MPOLE | L.TERM_POLAR, xyz=x)
            es.append(e["em"] + e["ep"])
        fd = (es[0] - es[1]) / 2e-5
        an = g["desum"][i, a]
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)) + 1e-11 * abs(es[0]) / 1e-5, (i, a, fd, an)
    print(f"  worst relative FD difference {worst:.2e}", flush=True)


def test_amoeba12k_parameters_by_finite_differences(engine):
    print("[gpu] amoeba 12288 FD of 32 parameters", flush=True)
    base = dict(amoeba12k())
    base["poleps"] = 1e-12
    types = np.asarray(base["type"])
    th = np.asarray(base["thole"], dtype=float).copy()
    th[types == types.max()] *= 0.85  # distinct per-type Thole values (no tie in min(a_i, a_k))
    base["thole"] = th
    energy(engine, base)
    pg = engine.param_grad()
    pole = np.asarray(base["pole"], dtype=float).reshape(-1, 13)
    alpha = np.asarray(base["polarity"], dtype=float)
    dpole_t = per_type(pg["dpole"], types)
    dal = pg["dpolarity"] + pg["dpdamp"] * np.where(alpha > 0, alpha ** (-5.0 / 6.0) / 6.0, 0.0)
    dal_t, dth_t = per_type(dal, types), per_type(pg["dthole"], types)
    checks = [("pole", t, c) for t in np.unique(types) for c in range(13)]
    checks += [("polarity", t, 0) for t in np.unique(types)] + [("thole", t, 0) for t in np.unique(types)]
    assert len(checks) == 30
    # every per-type parameter, and two of them twice with a second step size
    picks = [(k, t, c, 1.0) for k, t, c in checks] + [("polarity", types.min(), 0, 2.0),
                                                       ("thole", types.min(), 0, 2.0)]
    for kind, t, c, hs in picks:
        sel = types == t
        if kind == "pole":
            v0 = pole[sel, c][0]
            h = hs * (1e-6 * abs(v0) if v0 != 0 else 1e-6)
            analytic = dpole_t[t][c]
        elif kind == "polarity":
            h = hs * 1e-3 * alpha[sel][0]
            analytic = dal_t[t]
        else:
            h = hs * 1e-3 * th[sel][0]
            analytic = dth_t[t]
        es = []
        for sgn in (1, -1):
            s = dict(base)
            if kind == "pole":
                p = pole.copy()
                p[sel, c] += sgn * h
                s["pole"] = p.ravel()
            elif kind == "polarity":
                a = alpha.copy()
                a[sel] += sgn * h
                s["polarity"] = a
                s["pdamp"] = a ** (1.0 / 6.0)
            else:
                tt = th.copy()
                tt[sel] += sgn * h
                s["thole"] = tt
            es.append(energy(engine, s))
        fd = (es[0] - es[1]) / (2 * h)
        e0 = abs(es[0]) + abs(es[1])
        assert abs(fd - analytic) <= 1e-5 * abs(fd) + 1e-11 * e0 / h, (kind, t, c, fd, analytic)


# ---------------------------------------------------------------------------
# 512 clusters of U{2..64} waters

def cluster512():
    xyz, off = water.cluster_batch(512, 512, 2, 64)
    s = water.water_system(xyz, "water03")
    s["pcgprec"] = 0  # as benchmarked: the polarizability preconditioner (the sparse one: test_gpu_batch.py)
    return s, off


def sub(s, a, b, poleps):
    t = water.water_system(np.stack([s["x"][a:b], s["y"][a:b], s["z"][a:b]], 1), "water03")
    t["pcgprec"] = 0
    t["poleps"] = poleps
    return t


def test_clusters512_sample_matches_oracle(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] 512 clusters fp64 vs oracle on 32 sampled clusters", flush=True)
    s, off = cluster512()
    s["poleps"] = 1e-10
    engine.set_system(s)
    engine.set_batch(off)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    b = engine.fetch_batch()
    assert np.all(b["pcg_iters"] > 0)
    pick = np.random.default_rng(512).choice(len(off) - 1, size=32, replace=False)
    # the largest cluster is always in the sample
    pick[0] = int(np.argmax(np.diff(off)))
    for c in pick:
        lo, hi = off[c], off[c + 1]
        o = oracle.elec(sub(s, lo, hi, 1e-10), want_grad=True)
        assert abs(b["em"][c] - o["em"]) <= REL * max(1.0, abs(o["em"])), c
        assert abs(b["ep"][c] - o["ep"]) <= REL * max(1.0, abs(o["ep"])), c
        scale = max(1.0, np.abs(o["dem"] + o["dep"]).max())
        assert np.max(np.abs(g["desum"][lo:hi] - (o["dem"] + o["dep"]))) <= 1e-8 * scale, c
    assert abs(g["em"] - b["em"].sum()) <= 1e-12 * abs(g["em"])
    engine.set_batch(None)


def test_clusters512_iterations_and_mixed(engine):
    from polmdqc_amd import _lib as L

    print("[gpu] 512 clusters iterations at 1e-6 D and mixed vs fp64", flush=True)
    s, off = cluster512()
    engine.set_system(s)
    engine.set_batch(off)
    terms = L.TERM_MPOLE | L.TERM_POLAR
    d = engine.eval(L.GRADIENT, terms, L.PLAIN_FP64)
    bd = engine.fetch_batch()
    m = engine.eval(L.GRADIENT, terms, L.MIXED)
    bm = engine.fetch_batch()
    assert np.all(bd["pcg_iters"] > 0) and np.all(bm["pcg_iters"] > 0)
    for c in np.random.default_rng(7).choice(len(off) - 1, size=32, replace=False):
        o = oracle.elec(sub(s, off[c], off[c + 1], 1e-6), want_grad=False)
        assert bd["pcg_iters"][c] == o["iters"], c
    f = rms(m["desum"], d["desum"])
    nat = np.diff(off)
    e_em = np.abs(bm["em"] - bd["em"]) / nat
    e_ep = np.abs(bm["ep"] - bd["ep"]) / nat
    print(f"  rms force {f:.3e}  max per-atom energy err em {e_em.max():.2e} ep {e_ep.max():.2e}", flush=True)
    assert f < RMS_F
    assert e_em.max() < 1e-5 and e_ep.max() < 1e-5
    assert np.max(np.abs(bm["pcg_iters"] - bd["pcg_iters"])) <= 1
    engine.set_batch(None)
