"""The PME work forked onto the context's second stream against the same
evaluation on one stream.

With PME, an evaluation forks the permanent multipoles' reciprocal
potential (and, with HIPPO dispersion, the dispersion potential after it)
onto a second stream beside the real-space pair kernels and joins it before
the kernels that read it.  With the library's HIP-event timing on
(pmq_profile_enable) everything runs on the evaluation's own stream.  The
mixed form's gradients are fixed-point sums, so the two must agree bit for
bit there; FP64 gradients and every energy total are summed with FP64
atomics and agree to 1e-12.  Each run is also checked against a repeat of
itself, so a missing join would show as a difference between evaluations.
Systems: 216 AMOEBA waters (water03) and 216 HIPPO waters (hippo19, with
dispersion PME) in an 18.643 A box, 24^3 grid of order 5, 7 A cutoff.
"""
import numpy as np
import pytest

from polmdqc_amd import water

pytestmark = pytest.mark.gpu


def box(kind):
    ew = {"aewald": 0.5446, "nfft": 24, "order": 5, "cutoff": 7.0}
    if kind == "hippo":
        ew["dewald"] = 0.45
    return water.water_system(water.lattice_box(6, 18.643, 216), "water03" if kind == "amoeba" else "hippo19",
                              box=18.643, ewald=ew)


def agree(a, b, mixed):
    for k in ("em", "ep", "edsp", "er", "esum"):
        if k in a and k in b:
            assert abs(a[k] - b[k]) <= 1e-12 * max(abs(b[k]), 1e-300), (k, a[k], b[k])
    for k in ("desum", "dem", "dep"):
        if k in a and k in b:
            if mixed:
                assert np.array_equal(a[k], b[k]), k
            else:
                assert np.max(np.abs(a[k] - b[k])) <= 1e-12 * max(np.max(np.abs(b[k])), 1e-300), k


@pytest.mark.parametrize("kind", ["amoeba", "hippo"])
def test_side_stream_matches_one_stream(kind):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import Engine, profile_enable

    print(f"[gpu] side stream {kind}", flush=True)
    s = box(kind)
    terms = L.TERM_MPOLE | L.TERM_POLAR
    if kind == "hippo":
        terms |= L.TERM_REPEL | L.TERM_DISP
    e = Engine(0)
    try:
        e.set_system(s)
        for prec in (L.PLAIN_FP64, L.MIXED):
            mixed = prec == L.MIXED
            forked = []
            for _ in range(3):  # safe, then speculative twice
                forked.append(e.eval(L.GRADIENT, terms, prec))
            profile_enable(e, True)
            one = [e.eval(L.GRADIENT, terms, prec) for _ in range(2)]
            profile_enable(e, False)
            agree(forked[2], forked[1], mixed)
            agree(one[1], one[0], mixed)
            agree(forked[2], one[1], mixed)
    finally:
        e.close()
