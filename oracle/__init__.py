"""ctypes front end of the FP64 CPU oracle (liboracle.so).

Test infrastructure only: imported by tests/, the smoke check and bench.py's
CPU-baseline leg as the checker.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Box(C.Structure):
    _fields_ = [("kind", C.c_int), ("xbox", C.c_double), ("ybox", C.c_double), ("zbox", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double)]


class MpoleSys(C.Structure):
    _fields_ = [("n", C.c_int), ("x", _dp), ("y", _dp), ("z", _dp), ("pole", _dp),
                ("polaxe", _ip), ("zaxis", _ip), ("xaxis", _ip), ("yaxis", _ip),
                ("use", _ip), ("molcule", _ip),
                ("n12", _ip), ("i12", _ip), ("n13", _ip), ("i13", _ip),
                ("n14", _ip), ("i14", _ip), ("n15", _ip), ("i15", _ip),
                ("m2scale", C.c_double), ("m3scale", C.c_double), ("m4scale", C.c_double),
                ("m5scale", C.c_double), ("f", C.c_double), ("off2", C.c_double), ("box", Box),
                ("pentype", C.c_int), ("pcore", _dp), ("pval0", _dp), ("palpha", _dp), ("mono0", _dp),
                ("use_chgflx", C.c_int),
                ("nbond", C.c_int), ("ibnd", _ip), ("bl", _dp), ("bflx", _dp),
                ("nangle", C.c_int), ("iang", _ip), ("anat", _dp), ("aflx", _dp), ("abflx", _dp),
                ("balist", _ip)]


class MpoleOut(C.Structure):
    _fields_ = [("em", C.c_double), ("einter", C.c_double), ("nem", C.c_long),
                ("aem", _dp), ("dem", _dp), ("vir", _dp), ("rpole", _dp), ("pole_out", _dp)]


ENERGY, ANALYSIS, GRADIENT, VIRIAL = 0, 1, 2, 3
BOX_NONE, BOX_ORTHO, BOX_MONO, BOX_TRI, BOX_OCTA, BOX_DODECA = range(6)

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle library missing: build it with `make -C {HERE}`")
        _lib = C.CDLL(LIB)
        _lib.orc_empole.argtypes = [C.POINTER(MpoleSys), C.c_int, C.POINTER(MpoleOut)]
        _lib.orc_empole.restype = C.c_int
        _lib.orc_image.argtypes = [C.POINTER(Box), _dp, _dp, _dp]
        _lib.orc_pair_mpole.argtypes = [C.c_double] * 6 + [_dp, _dp, _dp]
    return _lib


def box_of(fx) -> Box:
    """Box from a fixture's reference globals (boxes.h flags)."""
    if not int(fx["use_bounds"]):
        kind = BOX_NONE
    elif int(fx["octahedron"]):
        kind = BOX_OCTA
    elif int(fx["dodecadron"]):
        kind = BOX_DODECA
    elif int(fx["triclinic"]):
        kind = BOX_TRI
    elif int(fx["monoclinic"]):
        kind = BOX_MONO
    else:
        kind = BOX_ORTHO
    return Box(kind, float(fx["xbox"]), float(fx["ybox"]), float(fx["zbox"]),
               float(fx["alpha"]), float(fx["beta"]), float(fx["gamma"]))


class _Keep:
    """Holds the numpy arrays a ctypes struct points into."""

    def __init__(self):
        self.arrays = []

    def d(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.arrays.append(a)
        return a.ctypes.data_as(_dp)

    def i(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        self.arrays.append(a)
        return a.ctypes.data_as(_ip)


def mpole_sys(fx, keep: _Keep, x=None, yaxis=None) -> MpoleSys:
    """MpoleSys over a fixture dict (keys as written by oracle/ref_dump.cpp)."""
    n = int(fx["n"])
    xyz = (fx["x"], fx["y"], fx["z"]) if x is None else (x[:, 0], x[:, 1], x[:, 2])
    cutoff = float(fx["mpolecut"])
    s = MpoleSys()
    s.n = n
    s.x, s.y, s.z = (keep.d(v) for v in xyz)
    s.pole = keep.d(fx["pole"])
    s.polaxe, s.zaxis, s.xaxis = keep.i(fx["polaxe"]), keep.i(fx["zaxis"]), keep.i(fx["xaxis"])
    ya = np.array(fx["yaxis"] if yaxis is None else yaxis, dtype=np.int32)
    s.yaxis = keep.i(ya)
    s.use, s.molcule = keep.i(fx["use"]), keep.i(fx["molcule"])
    for a in ("12", "13", "14", "15"):
        setattr(s, "n" + a, keep.i(fx["n" + a]))
        setattr(s, "i" + a, keep.i(fx["i" + a] if len(fx["i" + a]) else np.zeros(1)))
    s.m2scale, s.m3scale = float(fx["m2scale"]), float(fx["m3scale"])
    s.m4scale, s.m5scale = float(fx["m4scale"]), float(fx["m5scale"])
    s.f = float(fx["electric"]) / float(fx["dielec"])
    s.off2 = cutoff * cutoff
    s.box = box_of(fx)
    s.pentype = int(fx["pentype"])
    s.pcore, s.pval0 = keep.d(fx["pcore"]), keep.d(fx["pval0"])
    s.palpha, s.mono0 = keep.d(fx["palpha"]), keep.d(fx["mono0"])
    s.use_chgflx = int(fx["use_chgflx"])
    s.nbond = int(fx["nbond"])
    s.ibnd, s.bl = keep.i(fx["ibnd"]), keep.d(fx["bl"])
    s.bflx = keep.d(fx["bflx"] if len(fx["bflx"]) else np.zeros(1))
    s.nangle = int(fx["nangle"])
    s.iang, s.anat = keep.i(fx["iang"] if len(fx["iang"]) else np.zeros(1)), keep.d(
        fx["anat"] if len(fx["anat"]) else np.zeros(1))
    s.aflx = keep.d(fx["aflx"] if len(fx["aflx"]) else np.zeros(1))
    s.abflx = keep.d(fx["abflx"] if len(fx["abflx"]) else np.zeros(1))
    s.balist = keep.i(fx["balist"] if len(fx["balist"]) else np.zeros(1))
    return s


def empole(fx, mode: int = VIRIAL, x=None) -> dict:
    """Run the oracle's empole on a fixture; returns em/einter/nem/aem/dem/vir/rpole."""
    keep = _Keep()
    s = mpole_sys(fx, keep, x=x)
    n = s.n
    aem = np.zeros(n)
    dem = np.zeros(3 * n)
    vir = np.zeros(9)
    rpole = np.zeros(13 * n)
    o = MpoleOut()
    o.aem, o.dem = aem.ctypes.data_as(_dp), dem.ctypes.data_as(_dp)
    o.vir, o.rpole = vir.ctypes.data_as(_dp), rpole.ctypes.data_as(_dp)
    o.pole_out = None
    rc = lib().orc_empole(C.byref(s), mode, C.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_empole failed ({rc})")
    return {"em": o.em, "einter": o.einter, "nem": o.nem, "aem": aem, "dem": dem.reshape(n, 3),
            "vir": vir.reshape(3, 3), "rpole": rpole.reshape(n, 13)}


def pair_mpole(r, mscale, f, mi, mk):
    """One undamped multipole pair: (e, frc[3], trqi[3], trqk[3], vir[6])."""
    out = np.zeros(16)
    mi = np.ascontiguousarray(mi, dtype=np.float64)
    mk = np.ascontiguousarray(mk, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    lib().orc_pair_mpole(float(r @ r), float(r[0]), float(r[1]), float(r[2]), float(mscale), float(f),
                         mi.ctypes.data_as(_dp), mk.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


class ElecSys(C.Structure):
    _fields_ = [("mp", C.POINTER(MpoleSys)), ("use_ewald", C.c_int), ("aewald", C.c_double), ("cut", C.c_double),
                ("nfft1", C.c_int), ("nfft2", C.c_int), ("nfft3", C.c_int), ("bsorder", C.c_int),
                ("use_polar", C.c_int), ("polarity", _dp), ("thole", _dp), ("pdamp", _dp),
                ("np11", _ip), ("ip11", _ip), ("np12", _ip), ("ip12", _ip),
                ("np13", _ip), ("ip13", _ip), ("np14", _ip), ("ip14", _ip)] + \
               [(k, C.c_double) for k in ("p2scale", "p3scale", "p4scale", "p5scale", "p2iscale", "p3iscale",
                                          "p4iscale", "p5iscale", "d1scale", "d2scale", "d3scale", "d4scale",
                                          "u1scale", "u2scale", "u3scale", "u4scale", "poleps")] + \
               [("politer", C.c_int), ("solver", C.c_int)] + \
               [(k, C.c_double) for k in ("w2scale", "w3scale", "w4scale", "w5scale")] + [("dpequal", C.c_int)] + \
               [("pcgprec", C.c_int), ("usolvcut", C.c_double), ("udiag", C.c_double)]


class ElecOut(C.Structure):
    _fields_ = [("em", C.c_double), ("ep", C.c_double), ("em_real", C.c_double), ("em_recip", C.c_double),
                ("em_self", C.c_double), ("iters", C.c_int), ("dem", _dp), ("dep", _dp), ("uind", _dp),
                ("uinp", _dp), ("vir_m", _dp), ("vir_p", _dp), ("aem", _dp), ("aep", _dp), ("nem", C.c_long),
                ("nep", C.c_long), ("einter", C.c_double)]


def charge_flux_delta(fx, x=None) -> np.ndarray:
    """The monopole change of every atom from charge flux (alterchg.cpp:36-169)."""
    L = lib()
    L.orc_charge_flux_delta.argtypes = [C.POINTER(MpoleSys), _dp]
    L.orc_charge_flux_delta.restype = None
    keep = _Keep()
    mp = mpole_sys(fx, keep, x=x)
    out = np.zeros(int(fx["n"]))
    L.orc_charge_flux_delta(C.byref(mp), out.ctypes.data_as(_dp))
    return out


def elec(fx, want_grad: bool = True, solver: int = 0, x=None, poleps: float | None = None,
         polar: bool | None = None) -> dict:
    """Ewald/PME multipoles and AMOEBA polarization on a system dict."""
    L = lib()
    if not hasattr(L, "_elec_ready"):
        L.orc_elec.argtypes = [C.POINTER(ElecSys), C.c_int, C.POINTER(ElecOut)]
        L.orc_elec.restype = C.c_int
        L._elec_ready = True
    keep = _Keep()
    mp = mpole_sys(fx, keep, x=x)
    s = ElecSys()
    s.mp = C.pointer(mp)
    s.use_ewald = int(fx.get("use_ewald", 0))
    s.aewald = float(fx.get("aewald", 0.0))
    s.cut = float(fx["ewaldcut"]) if s.use_ewald else float(fx["mpolecut"])
    s.nfft1, s.nfft2, s.nfft3 = int(fx.get("nfft1", 0)), int(fx.get("nfft2", 0)), int(fx.get("nfft3", 0))
    s.bsorder = int(fx.get("bsorder", 5))
    s.use_polar = int(fx.get("use_polar", 1)) if polar is None else int(polar)
    n = int(fx["n"])
    s.polarity, s.thole, s.pdamp = keep.d(fx["polarity"]), keep.d(fx["thole"]), keep.d(fx["pdamp"])
    for a in ("11", "12", "13", "14"):
        cnt = fx.get("np" + a, np.zeros(n, np.int32))
        lst = fx.get("ip" + a, np.zeros(0, np.int32))
        setattr(s, "np" + a, keep.i(cnt))
        setattr(s, "ip" + a, keep.i(lst if len(lst) else np.zeros(1)))
    for k in ("p2scale", "p3scale", "p4scale", "p5scale", "p2iscale", "p3iscale", "p4iscale", "p5iscale",
              "d1scale", "d2scale", "d3scale", "d4scale", "u1scale", "u2scale", "u3scale", "u4scale"):
        setattr(s, k, float(fx[k]))
    s.poleps = float(fx.get("poleps", 1e-6)) if poleps is None else poleps
    s.politer = int(fx.get("politer", 100))
    s.solver = solver
    for k in ("w2scale", "w3scale", "w4scale", "w5scale"):
        setattr(s, k, float(fx.get(k, 1.0)))
    s.dpequal = int(fx.get("dpequal", 0))
    # PCG preconditioner: on by default with usolvcut 4.5 A and udiag 2, the
    # reference's defaults (kpolar.cpp:84, cutoffs.cpp:61)
    s.pcgprec = int(fx.get("pcgprec", 1))
    s.usolvcut = float(fx.get("usolvcut", 4.5))
    s.udiag = float(fx.get("udiag", 2.0))
    o = ElecOut()
    dem, dep, ud, up = (np.zeros(3 * n) for _ in range(4))
    vm, vp = np.zeros(9), np.zeros(9)
    o.dem, o.dep = dem.ctypes.data_as(_dp), dep.ctypes.data_as(_dp)
    o.uind, o.uinp = ud.ctypes.data_as(_dp), up.ctypes.data_as(_dp)
    o.vir_m, o.vir_p = vm.ctypes.data_as(_dp), vp.ctypes.data_as(_dp)
    aem, aep = np.zeros(n), np.zeros(n)
    o.aem, o.aep = aem.ctypes.data_as(_dp), aep.ctypes.data_as(_dp)
    rc = L.orc_elec(C.byref(s), 1 if want_grad else 0, C.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_elec failed ({rc})")
    return {"em": o.em, "ep": o.ep, "em_real": o.em_real, "em_recip": o.em_recip, "em_self": o.em_self,
            "iters": o.iters, "dem": dem.reshape(n, 3), "dep": dep.reshape(n, 3), "uind": ud.reshape(n, 3),
            "uinp": up.reshape(n, 3), "vir_m": vm.reshape(3, 3), "vir_p": vp.reshape(3, 3),
            "aem": aem, "aep": aep, "nem": int(o.nem), "nep": int(o.nep), "einter": o.einter}


class HippoSys(C.Structure):
    _fields_ = [("mp", C.POINTER(MpoleSys)), ("csix", _dp), ("adisp", _dp)] + \
               [(k, C.c_double) for k in ("dsp2scale", "dsp3scale", "dsp4scale", "dsp5scale")] + \
               [("use_dewald", C.c_int), ("dewald", C.c_double), ("dispcut", C.c_double), ("disptaper", C.c_double),
                ("nfft1", C.c_int), ("nfft2", C.c_int), ("nfft3", C.c_int), ("bsorder", C.c_int),
                ("sizpr", _dp), ("dmppr", _dp), ("elepr", _dp)] + \
               [(k, C.c_double) for k in ("r2scale", "r3scale", "r4scale", "r5scale", "repcut", "reptaper")]


class HippoOut(C.Structure):
    _fields_ = [("edsp", C.c_double), ("er", C.c_double), ("edsp_real", C.c_double), ("edsp_recip", C.c_double),
                ("edsp_self", C.c_double), ("dedsp", _dp), ("der", _dp), ("vir_d", _dp), ("vir_r", _dp)]


def taper_of(fx, key_cut: str, key_taper: str, default: float = 0.9) -> float:
    """Absolute taper distance: a value below 1 is a fraction of the cutoff
    (cutoffs.cpp:259-261)."""
    t = float(fx.get(key_taper, default))
    return t * float(fx[key_cut]) if t < 1.0 else t


def hippo_terms(fx, want_grad: bool = True, x=None) -> dict:
    """HIPPO dispersion (real + PME + self) and Pauli repulsion on a system dict."""
    L = lib()
    if not hasattr(L, "_hippo_ready"):
        L.orc_hippo.argtypes = [C.POINTER(HippoSys), C.c_int, C.POINTER(HippoOut)]
        L.orc_hippo.restype = C.c_int
        L.orc_rep_series.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L._hippo_ready = True
    keep = _Keep()
    mp = mpole_sys(fx, keep, x=x)
    s = HippoSys()
    s.mp = C.pointer(mp)
    s.csix, s.adisp = keep.d(fx["csix"]), keep.d(fx["adisp"])
    for k in ("dsp2scale", "dsp3scale", "dsp4scale", "dsp5scale", "r2scale", "r3scale", "r4scale", "r5scale"):
        setattr(s, k, float(fx[k]))
    s.use_dewald, s.dewald = int(fx.get("use_dewald", 0)), float(fx.get("dewald", 0.0))
    s.dispcut = float(fx["dispcut"])
    s.disptaper = taper_of(fx, "dispcut", "disptaper")
    s.nfft1, s.nfft2, s.nfft3 = int(fx.get("nfft1", 0)), int(fx.get("nfft2", 0)), int(fx.get("nfft3", 0))
    s.bsorder = int(fx.get("bsorder", 5))
    s.sizpr, s.dmppr, s.elepr = keep.d(fx["sizpr"]), keep.d(fx["dmppr"]), keep.d(fx["elepr"])
    s.repcut = float(fx["repcut"])
    s.reptaper = taper_of(fx, "repcut", "reptaper")
    n = int(fx["n"])
    o = HippoOut()
    dd, dr = np.zeros(3 * n), np.zeros(3 * n)
    vd, vr = np.zeros(9), np.zeros(9)
    o.dedsp, o.der = dd.ctypes.data_as(_dp), dr.ctypes.data_as(_dp)
    o.vir_d, o.vir_r = vd.ctypes.data_as(_dp), vr.ctypes.data_as(_dp)
    rc = L.orc_hippo(C.byref(s), 1 if want_grad else 0, C.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_hippo failed ({rc})")
    return {"edsp": o.edsp, "er": o.er, "edsp_real": o.edsp_real, "edsp_recip": o.edsp_recip,
            "edsp_self": o.edsp_self, "dedsp": dd.reshape(n, 3), "der": dr.reshape(n, 3),
            "vir_d": vd.reshape(3, 3), "vir_r": vr.reshape(3, 3)}


def rep_series(r: float, ai: float, ak: float) -> np.ndarray:
    hippo_lib = lib()
    hippo_lib.orc_rep_series.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
    B = np.zeros(6)
    hippo_lib.orc_rep_series(float(r), float(ai), float(ak), B.ctypes.data_as(_dp))
    return B


class VdwCtSys(C.Structure):
    _fields_ = [("mp", C.POINTER(MpoleSys)), ("use_vdw", C.c_int), ("nv", C.c_int), ("jv", _ip),
                ("radmin", _dp), ("epsilon", _dp), ("radmin4", _dp), ("epsilon4", _dp), ("kred", _dp),
                ("ired", _ip)] + [(k, C.c_double) for k in ("v2scale", "v3scale", "v4scale", "v5scale", "vdwcut",
                                                            "vdwtaper")] + \
               [("use_ctrn", C.c_int), ("chgct", _dp), ("dmpct", _dp), ("ctrncut", C.c_double),
                ("ctrntaper", C.c_double)]


class VdwCtOut(C.Structure):
    _fields_ = [("ev", C.c_double), ("ect", C.c_double), ("dev", _dp), ("dect", _dp), ("vir_v", _dp),
                ("vir_c", _dp)]


def vdwct(fx, want_grad: bool = True, x=None) -> dict:
    """Buffered 14-7 van der Waals and charge transfer on a system dict (the
    keys tinker.build_system writes)."""
    L = lib()
    if not hasattr(L, "_vdwct_ready"):
        L.orc_vdwct.argtypes = [C.POINTER(VdwCtSys), C.c_int, C.POINTER(VdwCtOut)]
        L.orc_vdwct.restype = C.c_int
        L._vdwct_ready = True
    keep = _Keep()
    mp = mpole_sys(fx, keep, x=x)
    n = int(fx["n"])
    s = VdwCtSys()
    s.mp = C.pointer(mp)
    s.use_vdw = int(fx.get("use_vdw", 0))
    s.nv = int(fx.get("nvdw", 1))
    z = np.zeros(max(1, n))
    s.jv = keep.i(fx.get("vdwclass", np.zeros(n, np.int32)))
    for k in ("radmin", "epsilon", "radmin4", "epsilon4"):
        setattr(s, k, keep.d(fx.get(k, np.zeros(1))))
    s.kred = keep.d(fx.get("kred", z))
    s.ired = keep.i(fx.get("ired", np.arange(n, dtype=np.int32)))
    for k in ("v2scale", "v3scale", "v4scale", "v5scale"):
        setattr(s, k, float(fx.get(k, 1.0)))
    s.vdwcut, s.vdwtaper = float(fx.get("vdwcut", 9.0)), float(fx.get("vdwtaper", 8.1))
    s.use_ctrn = int(fx.get("use_ctrn", 0))
    s.chgct, s.dmpct = keep.d(fx.get("chgct", z)), keep.d(fx.get("dmpct", z))
    s.ctrncut, s.ctrntaper = float(fx.get("ctrncut", 6.0)), float(fx.get("ctrntaper", 5.4))
    o = VdwCtOut()
    dev, dect, vv, vc = np.zeros(3 * n), np.zeros(3 * n), np.zeros(9), np.zeros(9)
    o.dev, o.dect = dev.ctypes.data_as(_dp), dect.ctypes.data_as(_dp)
    o.vir_v, o.vir_c = vv.ctypes.data_as(_dp), vc.ctypes.data_as(_dp)
    rc = L.orc_vdwct(C.byref(s), 1 if want_grad else 0, C.byref(o))
    if rc != 0:
        raise RuntimeError(f"orc_vdwct failed ({rc})")
    return {"ev": o.ev, "ect": o.ect, "dev": dev.reshape(n, 3), "dect": dect.reshape(n, 3),
            "vir_v": vv.reshape(3, 3), "vir_c": vc.reshape(3, 3)}
