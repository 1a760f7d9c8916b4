"""ctypes binding of libpmq.so (the C ABI declared in include/pmq.h).

This is the Python side of the drop-in boundary: the structs below mirror the
header field for field.  The library must exist: importing this module on a
machine without the built ``libpmq.so`` raises, it never falls back to a CPU
path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmq.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# enums of include/pmq.h
PMQ_OK = 0
ERRORS = {-1: "PMQ_ERR_ARG", -2: "PMQ_ERR_HIP", -3: "PMQ_ERR_STATE", -4: "PMQ_ERR_UNSUPPORTED",
          -5: "PMQ_ERR_NAN", -6: "PMQ_ERR_NOCONV"}
ENERGY, ANALYSIS, GRADIENT, VIRIAL = 0, 1, 2, 3
TERM_MPOLE, TERM_POLAR, TERM_REPEL, TERM_DISP, TERM_VDW, TERM_CTRN = 1, 2, 4, 8, 16, 32
PLAIN_FP64, MIXED = 0, 1
BOX_NONE, BOX_ORTHO, BOX_MONO, BOX_TRI, BOX_OCTA, BOX_DODECA = range(6)


class Config(C.Structure):
    _fields_ = [("device", C.c_int)]


class Box(C.Structure):
    _fields_ = [("kind", C.c_int), ("xbox", C.c_double), ("ybox", C.c_double), ("zbox", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double)]


class Topology(C.Structure):
    _fields_ = [("n", C.c_int), ("polaxe", _ip), ("zaxis", _ip), ("xaxis", _ip), ("yaxis", _ip),
                ("use", _ip), ("molcule", _ip),
                ("n12", _ip), ("i12", _ip), ("n13", _ip), ("i13", _ip),
                ("n14", _ip), ("i14", _ip), ("n15", _ip), ("i15", _ip),
                ("np11", _ip), ("ip11", _ip), ("np12", _ip), ("ip12", _ip),
                ("np13", _ip), ("ip13", _ip), ("np14", _ip), ("ip14", _ip),
                ("nbond", C.c_int), ("ibnd", _ip), ("nangle", C.c_int), ("iang", _ip), ("balist", _ip)]


class Params(C.Structure):
    _fields_ = [("pole", _dp),
                ("m2scale", C.c_double), ("m3scale", C.c_double), ("m4scale", C.c_double), ("m5scale", C.c_double),
                ("electric", C.c_double), ("dielec", C.c_double),
                ("pentype", C.c_int), ("pcore", _dp), ("pval0", _dp), ("palpha", _dp),
                ("use_chgflx", C.c_int), ("bl", _dp), ("bflx", _dp), ("anat", _dp), ("aflx", _dp), ("abflx", _dp),
                ("polarity", _dp), ("thole", _dp), ("pdamp", _dp),
                ("p2scale", C.c_double), ("p3scale", C.c_double), ("p4scale", C.c_double), ("p5scale", C.c_double),
                ("p2iscale", C.c_double), ("p3iscale", C.c_double), ("p4iscale", C.c_double), ("p5iscale", C.c_double),
                ("d1scale", C.c_double), ("d2scale", C.c_double), ("d3scale", C.c_double), ("d4scale", C.c_double),
                ("u1scale", C.c_double), ("u2scale", C.c_double), ("u3scale", C.c_double), ("u4scale", C.c_double),
                ("poleps", C.c_double), ("politer", C.c_int),
                ("use_ewald", C.c_int), ("mpolecut", C.c_double), ("ewaldcut", C.c_double), ("aewald", C.c_double),
                ("nfft1", C.c_int), ("nfft2", C.c_int), ("nfft3", C.c_int), ("bsorder", C.c_int),
                # HIPPO polarization, dispersion, repulsion
                ("w2scale", C.c_double), ("w3scale", C.c_double), ("w4scale", C.c_double), ("w5scale", C.c_double),
                ("dpequal", C.c_int),
                ("csix", _dp), ("adisp", _dp),
                ("dsp2scale", C.c_double), ("dsp3scale", C.c_double), ("dsp4scale", C.c_double),
                ("dsp5scale", C.c_double),
                ("use_dewald", C.c_int), ("dewald", C.c_double), ("dispcut", C.c_double), ("disptaper", C.c_double),
                ("dfft1", C.c_int), ("dfft2", C.c_int), ("dfft3", C.c_int), ("dbsorder", C.c_int),
                ("sizpr", _dp), ("dmppr", _dp), ("elepr", _dp),
                ("r2scale", C.c_double), ("r3scale", C.c_double), ("r4scale", C.c_double), ("r5scale", C.c_double),
                ("repcut", C.c_double), ("reptaper", C.c_double),
                ("pcgprec", C.c_int), ("usolvcut", C.c_double), ("udiag", C.c_double),
                # van der Waals, charge transfer
                ("nvdw", C.c_int), ("vdwclass", _ip),
                ("radmin", _dp), ("epsilon", _dp), ("radmin4", _dp), ("epsilon4", _dp),
                ("kred", _dp), ("ired", _ip),
                ("v2scale", C.c_double), ("v3scale", C.c_double), ("v4scale", C.c_double), ("v5scale", C.c_double),
                ("vdwcut", C.c_double), ("vdwtaper", C.c_double),
                ("chgct", _dp), ("dmpct", _dp), ("ctrncut", C.c_double), ("ctrntaper", C.c_double)]


class Result(C.Structure):
    _fields_ = [("em", C.c_double), ("ep", C.c_double), ("er", C.c_double), ("edsp", C.c_double),
                ("esum", C.c_double), ("einter", C.c_double), ("nem", C.c_longlong),
                ("vir", C.c_double * 9), ("pcg_iters", C.c_int),
                ("dem", _dp), ("dep", _dp), ("desum", _dp), ("aem", _dp), ("rpole", _dp),
                ("uind", _dp), ("uinp", _dp), ("der", _dp), ("dedsp", _dp), ("nep", C.c_longlong),
                ("aep", _dp), ("ev", C.c_double), ("ect", C.c_double), ("dev", _dp), ("dect", _dp)]


EXPORTS = ["pmq_create", "pmq_destroy", "pmq_last_error", "pmq_abi_version", "pmq_set_topology",
           "pmq_set_params", "pmq_eval", "pmq_set_coords", "pmq_compute", "pmq_fetch", "pmq_sync",
           "pmq_profile_enable", "pmq_profile_read", "pmq_set_batch", "pmq_fetch_batch", "pmq_param_grad",
           "pmq_set_fft", "pmq_fft_choice", "pmq_set_gather_lanes", "pmq_compute_batch_noalloc", "pmq_set_pair_form", "pmq_set_graphs"]

_lib = None


def lib() -> C.CDLL:
    """The loaded libpmq.so with argument types set; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m polmdqc_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.pmq_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.pmq_destroy.argtypes = [vp]
    L.pmq_last_error.argtypes = [vp]
    L.pmq_last_error.restype = C.c_char_p
    L.pmq_abi_version.argtypes = []
    L.pmq_set_topology.argtypes = [vp, C.POINTER(Topology)]
    L.pmq_set_params.argtypes = [vp, C.POINTER(Params)]
    L.pmq_eval.argtypes = [vp, C.c_int, C.c_uint, C.c_int, _dp, _dp, _dp, C.POINTER(Box), C.POINTER(Result)]
    L.pmq_set_coords.argtypes = [vp, _dp, _dp, _dp, C.POINTER(Box)]
    L.pmq_compute.argtypes = [vp, C.c_int, C.c_uint, C.c_int]
    L.pmq_fetch.argtypes = [vp, C.POINTER(Result)]
    L.pmq_sync.argtypes = [vp]
    L.pmq_set_batch.argtypes = [vp, C.c_int, _ip]
    L.pmq_fetch_batch.argtypes = [vp, _dp, _dp, _ip]
    L.pmq_param_grad.argtypes = [vp, _dp, _dp, _dp, _dp]
    # device pointers (integers) for the allocation-free batched evaluation
    L.pmq_compute_batch_noalloc.argtypes = [vp, C.c_int, C.c_uint, C.c_int, vp, vp, vp, vp, vp]
    for name in EXPORTS:
        if name != "pmq_last_error":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L
