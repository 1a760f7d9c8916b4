"""Context state rules of the C ABI on the GPU: a rejected setter leaves no
usable system behind, a setter first completes the evaluation still pending,
and a batched solve that runs out of iterations is reported.

The reference stops the program on bad input (fatal(), fatal.cpp:21-28);
the library returns an error code instead and must then refuse to compute
with half-updated state.
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


def box216():
    xyz = water.lattice_box(6, 18.643, 216)
    return water.water_system(xyz, "water03", box=18.643,
                              ewald={"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0})


def test_failed_set_topology_blocks_compute(engine):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import PmqError

    print("[gpu] state: failed set_topology", flush=True)
    s = box216()
    engine.set_system(s)
    engine.eval(L.GRADIENT, L.TERM_MPOLE)
    bad = dict(water.water_system(water.cluster(8, np.random.default_rng(1)), "water03"))
    pol = np.array(bad["polaxe"], dtype=np.int32).copy()
    pol[3] = 99  # no such frame type
    bad["polaxe"] = pol
    with pytest.raises(PmqError, match="PMQ_ERR_ARG"):
        engine.set_system(bad)
    with pytest.raises(PmqError, match="PMQ_ERR_STATE"):
        engine.compute(L.GRADIENT, L.TERM_MPOLE)
    # a good system afterwards works again
    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE)
    assert np.isfinite(g["em"])


def test_rejected_set_params_blocks_compute(engine):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import PmqError

    print("[gpu] state: rejected set_params", flush=True)
    s = box216()
    engine.set_system(s)
    ref = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    bad = dict(s)
    a = np.array(s["polarity"], dtype=float).copy()
    a[5] = -1.0
    bad["polarity"] = a
    with pytest.raises(PmqError, match="PMQ_ERR_ARG"):
        engine.set_system(bad)
    with pytest.raises(PmqError, match="PMQ_ERR_STATE"):
        engine.compute(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    engine.set_system(s)
    g = engine.eval(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    assert abs(g["ep"] - ref["ep"]) <= 1e-12 * abs(ref["ep"])  # FP64 atomics: equal to rounding


def test_set_coords_completes_the_pending_evaluation(engine):
    """compute -> set_coords(next) -> fetch returns the first geometry's
    result, also when the speculative run falls short (a denser geometry
    needs more pairs and iterations than the last one) and is redone."""
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import Engine

    print("[gpu] state: set_coords while pending", flush=True)
    s = box216()
    terms = L.TERM_MPOLE | L.TERM_POLAR
    engine.set_system(s)
    engine.eval(L.GRADIENT, terms)
    engine.eval(L.GRADIENT, terms)  # speculative from here on
    xyz = np.stack([s["x"], s["y"], s["z"]], 1)
    dense = xyz * np.array([0.6, 1.0, 1.0])
    engine.set_coords(dense)
    engine.compute(L.GRADIENT, terms)
    engine.set_coords(xyz)  # the dense evaluation is still pending here
    g = engine.fetch()
    fresh = Engine(0)
    try:
        fresh.set_system(s)
        f = fresh.eval(L.GRADIENT, terms, xyz=dense)
    finally:
        fresh.close()
    assert g["pcg_iters"] == f["pcg_iters"]
    assert abs(g["ep"] - f["ep"]) <= 1e-12 * abs(f["ep"])
    assert np.max(np.abs(g["desum"] - f["desum"])) <= 1e-12 * np.max(np.abs(f["desum"]))


def test_batch_nonconvergence_is_reported(engine):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import PmqError

    print("[gpu] state: batch without convergence", flush=True)
    xyz, off = water.cluster_batch(6, 9, 4, 10)
    s = water.water_system(xyz, "water03")
    s["pcgprec"] = 0
    s["politer"] = 1
    s["poleps"] = 1e-12
    engine.set_system(s)
    engine.set_batch(off)
    engine.compute(L.GRADIENT, L.TERM_MPOLE | L.TERM_POLAR)
    with pytest.raises(PmqError, match="PMQ_ERR_NOCONV"):
        engine.fetch()
    assert np.all(engine.fetch_batch()["pcg_iters"] < 0)
    engine.set_batch(None)
