"""The allocation-free batched evaluation (pmq_compute_batch_noalloc)
against the batched evaluation it is a variant of (pmq_compute on a batch,
then pmq_fetch / pmq_fetch_batch), through the C ABI.

Both run the same kernels with the same lanes-per-row variant.  The mixed
form sums in 64-bit fixed point, so its results must be bit-identical; the
FP64 form adds its scaled-pair corrections with FP64 atomics, whose order
may differ from run to run, so it must agree to 1e-12 relative (far inside
the 1e-9 parity bar).  Iteration counts are equal in both.  Two calls in a row, the second
on a smaller batch, check that nothing of the first call's scratch leaks
into the second.  Output buffers are device memory the caller allocates
(hipMalloc through ctypes here), as such a caller would.
"""
import ctypes as C
import os

import numpy as np
import pytest

from polmdqc_amd import water

pytestmark = pytest.mark.gpu


class HipMem:
    """Device buffers of the caller, allocated before the evaluation, through
    the HIP runtime libpmq.so is bound to: the object already loaded under
    the runtime's soname (opening the runtime by file name could load a
    second copy, e.g. one bundled with another package, whose allocations
    the library cannot use)."""

    def __init__(self):
        from polmdqc_amd import _lib

        _lib.lib()  # libpmq.so and its runtime are loaded
        self.hip = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD)
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.ptrs.append(p)
        return p.value

    def get(self, ptr, count, dtype):
        out = np.zeros(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


@pytest.fixture(scope="module")
def engine():
    from polmdqc_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def batch(nclusters, seed, max_mol):
    xyz, off = water.cluster_batch(nclusters, seed, 2, max_mol)
    s = water.water_system(xyz, "water03")
    s["pcgprec"] = 0
    return s, off


def same(a, b, prec, where):
    from polmdqc_amd import _lib as L

    if prec == L.MIXED:
        assert np.array_equal(a, b), where
    else:
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b)), err_msg=str(where))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_equals_batched_compute(engine, precision):
    from polmdqc_amd import _lib as L

    prec = L.PLAIN_FP64 if precision == "fp64" else L.MIXED
    terms = L.TERM_MPOLE | L.TERM_POLAR
    mem = HipMem()
    try:
        for step, (ncl, seed, max_mol) in enumerate(((64, 21, 40), (20, 22, 12))):
            print(f"[gpu] noalloc {precision} call {step}: {ncl} clusters", flush=True)
            s, off = batch(ncl, seed, max_mol)
            n = int(s["n"])
            engine.set_system(s)
            engine.set_batch(off)
            ref = engine.eval(L.GRADIENT, terms, prec)
            rb = engine.fetch_batch()
            if step == 0:
                # buffers sized for the first (larger) call, reused by the second
                cap_c, cap_n = ncl, n
                d_em, d_ep = mem.alloc(8 * cap_c), mem.alloc(8 * cap_c)
                d_it, d_tot, d_grad = mem.alloc(4 * cap_c), mem.alloc(16), mem.alloc(8 * 3 * cap_n)
            assert ncl <= cap_c and n <= cap_n
            for rep in range(2):
                engine.compute_batch_noalloc(L.GRADIENT, terms, prec, d_em, d_ep, d_it, d_tot, d_grad)
                engine.sync()
                em = mem.get(d_em, ncl, np.float64)
                ep = mem.get(d_ep, ncl, np.float64)
                it = mem.get(d_it, ncl, np.int32)
                tot = mem.get(d_tot, 2, np.float64)
                grad = mem.get(d_grad, 3 * n, np.float64).reshape(n, 3)
                same(em, rb["em"], prec, (step, rep))
                same(ep, rb["ep"], prec, (step, rep))
                assert np.array_equal(it, rb["pcg_iters"]), (step, rep)
                assert np.all(it > 0)
                same(tot, np.array([ref["em"], ref["ep"]]), prec, (step, rep))
                same(grad, ref["desum"], prec, (step, rep))
            # energy mode: the gradient array is left alone
            engine.compute_batch_noalloc(L.ENERGY, terms, prec, d_em, 0, 0, d_tot, 0)
            engine.sync()
            same(mem.get(d_em, ncl, np.float64), rb["em"], prec, step)
        engine.set_batch(None)
    finally:
        mem.free()


def test_noalloc_rejects_what_it_cannot_do(engine):
    from polmdqc_amd import _lib as L
    from polmdqc_amd.engine import PmqError

    print("[gpu] noalloc rejections", flush=True)
    s, off = batch(8, 3, 10)
    engine.set_system(s)
    with pytest.raises(PmqError, match="PMQ_ERR_STATE"):  # one system, no batch
        engine.compute_batch_noalloc(L.GRADIENT, L.TERM_MPOLE, L.PLAIN_FP64)
    engine.set_batch(off)
    engine.compute_batch_noalloc(L.ENERGY, L.TERM_MPOLE, L.PLAIN_FP64)
    engine.sync()
    with pytest.raises(PmqError, match="PMQ_ERR_STATE"):  # its results are not fetched
        engine.fetch()
    # more clusters than the fixed scratch holds
    nbig = 65537
    big = water.water_system(np.zeros((3 * nbig, 3)) + np.repeat(np.arange(nbig), 3)[:, None] * 50.0, "water03")
    big["pcgprec"] = 0
    engine.set_system(big)
    engine.set_batch(np.arange(0, 3 * nbig + 1, 3, dtype=np.int32))
    with pytest.raises(PmqError, match="PMQ_ERR_ARG"):
        engine.compute_batch_noalloc(L.ENERGY, L.TERM_MPOLE, L.PLAIN_FP64)
    engine.set_batch(None)
