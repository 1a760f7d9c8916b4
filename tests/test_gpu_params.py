"""Parameter gradients (FP64) against central finite differences of the GPU
energy, per atom type, on the 216-water box (Ewald) and a 40-water cluster.

Parameters: the 13 local multipole components of each type, the
polarizability (with pdamp = alpha^(1/6) as the parameter files tie them) and
the Thole parameter.  The PCG is converged to 1e-12 D so the energy is a
smooth function of the parameters; the two types get distinct Thole values.  Steps: 1e-6 relative for multipole components (1e-6
absolute for components that are zero); 1e-3 relative for the polarizability
and the Thole parameter, whose derivatives are small against the total (a
Thole derivative of 0.04 kcal/mol per unit beside |E| = 300 kcal/mol would
keep only ~4 digits from a 1e-6 step).  Bar: 1e-5 relative, plus the
difference quotient's own resolution 1e-11 |E| / h.  The energy is quadratic
in the multipoles, so sum(pole * dE/dpole) = 2 E is checked too.
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


def systems():
    box = water.water_system(water.lattice_box(6, 18.643, 216), "water03", box=18.643,
                             ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})
    clu = water.water_system(water.cluster(40, np.random.default_rng(4)), "water03")
    return {"box216": box, "cluster40": clu}


def energy(engine, s):
    from polmdqc_amd import _lib as L

    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    return g["em"] + g["ep"]


def per_type(x, types):
    return {t: np.asarray(x)[types == t].sum(axis=0) for t in np.unique(types)}


@pytest.mark.parametrize("name", ["box216", "cluster40"])
def test_multipole_homogeneity(engine, name):
    from polmdqc_amd import _lib as L

    s = dict(systems()[name])
    s["poleps"] = 1e-12
    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    p = engine.param_grad()
    lhs = float(np.sum(np.asarray(s["pole"]).reshape(-1, 13) * p["dpole"]))
    assert abs(lhs - 2.0 * (g["em"] + g["ep"])) <= 1e-9 * abs(g["em"] + g["ep"])
    engine.eval(L.GRADIENT, L.TERM_MPOLE)
    pm = engine.param_grad()
    e = engine.eval(L.GRADIENT, L.TERM_MPOLE)
    assert abs(float(np.sum(np.asarray(s["pole"]).reshape(-1, 13) * pm["dpole"])) - 2.0 * e["em"]) <= 1e-9 * abs(e["em"])


@pytest.mark.parametrize("name", ["box216", "cluster40"])
def test_parameter_gradients_match_finite_differences(engine, name):
    print(f"[gpu] parameter FD {name}", flush=True)
    base = dict(systems()[name])
    base["poleps"] = 1e-12
    types = np.asarray(base["type"])
    # distinct Thole values per type: the pair value min(a_i, a_k) is smooth
    # in each type's parameter away from a tie (at a tie a central difference
    # sees the mean of the one-sided slopes only to first order in the step)
    th = np.asarray(base["thole"], dtype=float).copy()
    th[types == types.max()] *= 0.85
    base["thole"] = th
    energy(engine, base)
    pg = engine.param_grad()
    pole = np.asarray(base["pole"], dtype=float).reshape(-1, 13)
    alpha = np.asarray(base["polarity"], dtype=float)
    dpole_t = per_type(pg["dpole"], types)
    # polarizability with pdamp = alpha^(1/6)
    dal = pg["dpolarity"] + pg["dpdamp"] * np.where(alpha > 0, alpha ** (-5.0 / 6.0) / 6.0, 0.0)
    dal_t, dth_t = per_type(dal, types), per_type(pg["dthole"], types)
    rng = np.random.default_rng(32)
    checks = [("pole", t, c) for t in np.unique(types) for c in range(13)]
    checks += [("polarity", t, 0) for t in np.unique(types)] + [("thole", t, 0) for t in np.unique(types)]
    picks = [checks[k] for k in rng.choice(len(checks), size=min(32, len(checks)), replace=False)]
    for kind, t, c in picks:
        sel = types == t
        if kind == "pole":
            v0 = pole[sel, c][0]
            h = 1e-6 * abs(v0) if v0 != 0 else 1e-6
            analytic = dpole_t[t][c]
        elif kind == "polarity":
            v0 = alpha[sel][0]
            h = 1e-3 * v0
            analytic = dal_t[t]
        else:
            v0 = np.asarray(base["thole"], dtype=float)[sel][0]
            h = 1e-3 * v0
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
                th = np.asarray(base["thole"], dtype=float).copy()
                th[sel] += sgn * h
                s["thole"] = th
            es.append(energy(engine, s))
        fd = (es[0] - es[1]) / (2 * h)
        e0 = abs(es[0]) + abs(es[1])
        assert abs(fd - analytic) <= 1e-5 * abs(fd) + 1e-11 * e0 / h, (kind, t, c, fd, analytic)


def test_positions_match_finite_differences(engine):
    """dE/dx of 32 random coordinates per box, central differences 1e-5 A."""
    from polmdqc_amd import _lib as L

    for name, s in systems().items():
        s = dict(s)
        s["poleps"] = 1e-12
        engine.set_system(s)
        g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
        xyz = np.stack([s["x"], s["y"], s["z"]], 1)
        rng = np.random.default_rng(4)
        for k in range(32):
            i, a = int(rng.integers(len(xyz))), int(rng.integers(3))
            es = []
            for sgn in (1, -1):
                x = xyz.copy()
                x[i, a] += sgn * 1e-5
                e = engine.eval(L.ENERGY, L.TERM_MPOLE | L.TERM_POLAR, xyz=x)
                es.append(e["em"] + e["ep"])
            fd = (es[0] - es[1]) / 2e-5
            an = g["desum"][i, a]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)) + 1e-11 * abs(es[0]) / 1e-5, (name, i, a, fd, an)


def test_autograd_function(engine):
    import torch

    from polmdqc_amd.autograd import ElecEnergy, ElecModel

    s = dict(systems()["cluster40"])
    s["poleps"] = 1e-12
    types = np.asarray(s["type"])
    ut, tidx = np.unique(types, return_inverse=True)
    pole = np.asarray(s["pole"], dtype=float).reshape(-1, 13)
    pole_t = torch.tensor(np.stack([pole[tidx == k][0] for k in range(len(ut))]), requires_grad=True)
    pol_t = torch.tensor([float(np.asarray(s["polarity"])[tidx == k][0]) for k in range(len(ut))], requires_grad=True)
    th = np.asarray(s["thole"], dtype=float)
    th_t = torch.tensor([float(th[tidx == k][0]) * (1.0 if k == 0 else 0.85) for k in range(len(ut))],
                        requires_grad=True)
    xyz = torch.tensor(np.stack([s["x"], s["y"], s["z"]], 1), requires_grad=True)
    model = ElecModel(s, tidx, engine)
    e = ElecEnergy.apply(xyz, pole_t, pol_t, th_t, model)
    e.backward()
    # one parameter of each kind against a central difference
    for tensor, idx, h in ((pole_t, (0, 0), 1e-6), (pol_t, (1,), 1e-3 * pol_t.detach()[1].item()),
                           (th_t, (0,), 1e-3 * th_t.detach()[0].item())):
        vals = []
        for sgn in (1, -1):
            with torch.no_grad():
                t2 = tensor.detach().clone()
                t2[idx] += sgn * h
                args = {id(pole_t): pole_t.detach(), id(pol_t): pol_t.detach(), id(th_t): th_t.detach()}
                args[id(tensor)] = t2
                vals.append(float(ElecEnergy.apply(xyz.detach(), args[id(pole_t)], args[id(pol_t)], args[id(th_t)],
                                                   model)))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - float(tensor.grad[idx])) <= 1e-5 * abs(fd) + 1e-11 * abs(vals[0]) / h
    assert xyz.grad.shape == xyz.shape
