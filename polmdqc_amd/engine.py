"""Python front end over the C ABI: one Engine = one pmq context on one GPU.

Systems are described by a dict whose keys are the reference's global arrays
(``x, y, z, pole, polaxe, zaxis, xaxis, yaxis, use, molcule, n12, i12, ...,
m2scale, electric, dielec, pentype, ...``), i.e. what ``mechanic()`` leaves
in the modules of include/md/*.h.  Arrays are numpy; connectivity lists are
flat (counts ``n1x`` plus concatenated 0-based lists ``i1x``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class PmqError(RuntimeError):
    pass


def box_from(sys: dict) -> L.Box:
    """Box struct from the reference's box globals (boxes.h flags)."""
    if not int(sys.get("use_bounds", 0)):
        kind = L.BOX_NONE
    elif int(sys.get("octahedron", 0)):
        kind = L.BOX_OCTA
    elif int(sys.get("dodecadron", 0)):
        kind = L.BOX_DODECA
    elif int(sys.get("triclinic", 0)):
        kind = L.BOX_TRI
    elif int(sys.get("monoclinic", 0)):
        kind = L.BOX_MONO
    else:
        kind = L.BOX_ORTHO
    return L.Box(kind, float(sys.get("xbox", 0)), float(sys.get("ybox", 0)), float(sys.get("zbox", 0)),
                 float(sys.get("alpha", 90)), float(sys.get("beta", 90)), float(sys.get("gamma", 90)))


class Engine:
    def __init__(self, device: int = 0):
        self._L = L.lib()
        self._ctx = C.c_void_p()
        cfg = L.Config(device)
        rc = self._L.pmq_create(C.byref(cfg), C.byref(self._ctx))
        if rc != 0:
            raise PmqError(f"pmq_create failed: {L.ERRORS.get(rc, rc)}")
        self._keep = []
        self.n = 0
        self.nsys = 0

    # ------------------------------------------------------------------
    def close(self):
        if self._ctx:
            self._L.pmq_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self._L.pmq_last_error(self._ctx).decode(errors="replace")
            raise PmqError(f"{what}: {L.ERRORS.get(rc, rc)}: {msg}")

    def _d(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._keep.append(a)
        return a.ctypes.data_as(_dp)

    def _i(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.int32)
        if a.size == 0:
            a = np.zeros(1, dtype=np.int32)
        self._keep.append(a)
        return a.ctypes.data_as(_ip)

    # ------------------------------------------------------------------
    def set_system(self, s: dict):
        """Upload topology and parameters of a system dict."""
        self._keep = []
        n = int(s["n"])
        t = L.Topology()
        t.n = n
        t.polaxe, t.zaxis, t.xaxis, t.yaxis = (self._i(s[k]) for k in ("polaxe", "zaxis", "xaxis", "yaxis"))
        t.use = self._i(s.get("use"))
        t.molcule = self._i(s.get("molcule"))
        for a in ("12", "13", "14", "15"):
            setattr(t, "n" + a, self._i(s["n" + a]))
            setattr(t, "i" + a, self._i(s["i" + a]))
        for a in ("11", "12", "13", "14"):
            if "np" + a in s:
                setattr(t, "np" + a, self._i(s["np" + a]))
                setattr(t, "ip" + a, self._i(s["ip" + a]))
        t.nbond = int(s.get("nbond", 0))
        t.ibnd = self._i(s.get("ibnd", np.zeros(0)))
        t.nangle = int(s.get("nangle", 0))
        t.iang = self._i(s.get("iang", np.zeros(0)))
        bal = s.get("balist")
        t.balist = self._i(bal) if bal is not None and len(bal) else None
        self._check(self._L.pmq_set_topology(self._ctx, C.byref(t)), "pmq_set_topology")
        self.nsys = 0

        p = L.Params()
        p.pole = self._d(s["pole"])
        p.m2scale, p.m3scale = float(s["m2scale"]), float(s["m3scale"])
        p.m4scale, p.m5scale = float(s["m4scale"]), float(s["m5scale"])
        p.electric, p.dielec = float(s["electric"]), float(s["dielec"])
        p.pentype = int(s.get("pentype", 0))
        if p.pentype:
            p.pcore, p.pval0, p.palpha = self._d(s["pcore"]), self._d(s["pval0"]), self._d(s["palpha"])
        p.use_chgflx = int(s.get("use_chgflx", 0))
        if p.use_chgflx:
            pole = np.array(s["pole"], dtype=np.float64).reshape(n, 13).copy()
            pole[:, 0] = s["mono0"]
            p.pole = self._d(pole)
            for k in ("bl", "bflx", "anat", "aflx", "abflx"):
                setattr(p, k, self._d(s[k]))
        for k in ("polarity", "thole", "pdamp"):
            if k in s:
                setattr(p, k, self._d(s[k]))
        for k in ("p2scale", "p3scale", "p4scale", "p5scale", "p2iscale", "p3iscale", "p4iscale", "p5iscale",
                  "d1scale", "d2scale", "d3scale", "d4scale", "u1scale", "u2scale", "u3scale", "u4scale",
                  "poleps"):
            if k in s:
                setattr(p, k, float(s[k]))
        p.politer = int(s.get("politer", 100))
        p.use_ewald = int(s.get("use_ewald", 0))
        p.mpolecut = float(s.get("mpolecut", 1e12))
        p.ewaldcut = float(s.get("ewaldcut", 7.0))
        p.aewald = float(s.get("aewald", 0.0))
        p.nfft1, p.nfft2, p.nfft3 = int(s.get("nfft1", 0)), int(s.get("nfft2", 0)), int(s.get("nfft3", 0))
        p.bsorder = int(s.get("bsorder", 5))
        # HIPPO: w scales and d-equals-p (polarization with charge
        # penetration), dispersion and repulsion; tapers below 1 are
        # fractions of their cutoff (cutoffs.cpp:259-261)
        for k in ("w2scale", "w3scale", "w4scale", "w5scale"):
            setattr(p, k, float(s.get(k, 1.0)))
        p.dpequal = int(s.get("dpequal", 0))
        for k, d in (("dsp2scale", 0.0), ("dsp3scale", 0.0), ("dsp4scale", 1.0), ("dsp5scale", 1.0),
                     ("r2scale", 0.0), ("r3scale", 0.0), ("r4scale", 1.0), ("r5scale", 1.0)):
            setattr(p, k, float(s.get(k, d)))

        def taper(cut, key):
            t = float(s.get(key, 0.9))
            return t * cut if t < 1.0 else t

        csix = s.get("csix")
        if csix is not None and np.any(np.asarray(csix) != 0.0):
            p.csix, p.adisp = self._d(csix), self._d(s["adisp"])
            p.use_dewald, p.dewald = int(s.get("use_dewald", 0)), float(s.get("dewald", 0.0))
            p.dispcut = float(s.get("dispcut", 9.0))
            p.disptaper = taper(p.dispcut, "disptaper")
            p.dfft1, p.dfft2, p.dfft3 = (int(s.get(k, s.get(k.replace("d", "n", 1), 0))) for k in
                                         ("dfft1", "dfft2", "dfft3"))
            p.dbsorder = int(s.get("dbsorder", s.get("bsorder", 5)))
        sizpr = s.get("sizpr")
        if sizpr is not None and np.any(np.asarray(sizpr) != 0.0):
            p.sizpr, p.dmppr, p.elepr = self._d(sizpr), self._d(s["dmppr"]), self._d(s["elepr"])
            p.repcut = float(s.get("repcut", 6.0))
            p.reptaper = taper(p.repcut, "reptaper")
        if int(s.get("nvdw", 0)) > 0:
            p.nvdw = int(s["nvdw"])
            p.vdwclass, p.ired = self._i(s["vdwclass"]), self._i(s["ired"])
            p.radmin, p.epsilon = self._d(s["radmin"]), self._d(s["epsilon"])
            p.radmin4, p.epsilon4 = self._d(s["radmin4"]), self._d(s["epsilon4"])
            p.kred = self._d(s["kred"])
            for k, d in (("v2scale", 0.0), ("v3scale", 0.0), ("v4scale", 1.0), ("v5scale", 1.0)):
                setattr(p, k, float(s.get(k, d)))
            p.vdwcut = float(s.get("vdwcut", 9.0))
            p.vdwtaper = taper(p.vdwcut, "vdwtaper")
        chgct = s.get("chgct")
        if chgct is not None and np.any(np.asarray(chgct) != 0.0):
            p.chgct, p.dmpct = self._d(chgct), self._d(s["dmpct"])
            p.ctrncut = float(s.get("ctrncut", 6.0))
            p.ctrntaper = taper(p.ctrncut, "ctrntaper")
        # the reference's PCG defaults: preconditioner on, 4.5 A, udiag 2
        p.pcgprec = int(s.get("pcgprec", 1))
        p.usolvcut, p.udiag = float(s.get("usolvcut", 4.5)), float(s.get("udiag", 2.0))
        self._check(self._L.pmq_set_params(self._ctx, C.byref(p)), "pmq_set_params")
        self.n = n
        self.box = box_from(s)
        self.set_coords(np.stack([s["x"], s["y"], s["z"]], 1))

    def set_coords(self, xyz: np.ndarray, box: L.Box | None = None):
        xyz = np.asarray(xyz, dtype=np.float64)
        if box is not None:
            self.box = box
        x, y, z = (np.ascontiguousarray(xyz[:, a]) for a in range(3))
        self._check(self._L.pmq_set_coords(self._ctx, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                                           z.ctypes.data_as(_dp), C.byref(self.box)), "pmq_set_coords")

    def compute(self, mode: int = L.VIRIAL, terms: int = L.TERM_MPOLE, precision: int = L.PLAIN_FP64):
        """Enqueue one evaluation on the device (asynchronous)."""
        self._check(self._L.pmq_compute(self._ctx, mode, terms, precision), "pmq_compute")

    def sync(self):
        self._check(self._L.pmq_sync(self._ctx), "pmq_sync")

    def fetch(self, want_rpole: bool = False) -> dict:
        n = self.n
        dem = np.zeros(3 * n)
        aem = np.zeros(n)
        rpole = np.zeros(13 * n) if want_rpole else None
        uind = np.zeros(3 * n)
        uinp = np.zeros(3 * n)
        dep = np.zeros(3 * n)
        desum = np.zeros(3 * n)
        der = np.zeros(3 * n)
        dedsp = np.zeros(3 * n)
        aep = np.zeros(n)
        dev, dect = np.zeros(3 * n), np.zeros(3 * n)
        r = L.Result()
        r.dev, r.dect = dev.ctypes.data_as(_dp), dect.ctypes.data_as(_dp)
        r.aep = aep.ctypes.data_as(_dp)
        r.der, r.dedsp = der.ctypes.data_as(_dp), dedsp.ctypes.data_as(_dp)
        r.dem, r.aem = dem.ctypes.data_as(_dp), aem.ctypes.data_as(_dp)
        r.dep, r.desum = dep.ctypes.data_as(_dp), desum.ctypes.data_as(_dp)
        r.rpole = rpole.ctypes.data_as(_dp) if want_rpole else None
        r.uind, r.uinp = uind.ctypes.data_as(_dp), uinp.ctypes.data_as(_dp)
        self._check(self._L.pmq_fetch(self._ctx, C.byref(r)), "pmq_fetch")
        out = {"em": r.em, "ep": r.ep, "er": r.er, "edsp": r.edsp, "esum": r.esum, "einter": r.einter,
               "nem": int(r.nem), "vir": np.array(r.vir[:]).reshape(3, 3), "pcg_iters": r.pcg_iters,
               "dem": dem.reshape(n, 3), "dep": dep.reshape(n, 3), "desum": desum.reshape(n, 3), "aem": aem,
               "uind": uind.reshape(n, 3), "uinp": uinp.reshape(n, 3), "der": der.reshape(n, 3),
               "dedsp": dedsp.reshape(n, 3), "nep": int(r.nep), "aep": aep, "ev": r.ev, "ect": r.ect,
               "dev": dev.reshape(n, 3), "dect": dect.reshape(n, 3)}
        if want_rpole:
            out["rpole"] = rpole.reshape(n, 13)
        return out

    def param_grad(self) -> dict:
        """dE/d(parameters) of the last FP64 GRADIENT evaluation, per atom:
        dpole (n, 13) in the local frame, dpolarity (fixed pdamp), dpdamp, dthole."""
        n = self.n
        dpole, dpol, dpd, dth = np.zeros(13 * n), np.zeros(n), np.zeros(n), np.zeros(n)
        self._check(self._L.pmq_param_grad(self._ctx, dpole.ctypes.data_as(_dp), dpol.ctypes.data_as(_dp),
                                           dpd.ctypes.data_as(_dp), dth.ctypes.data_as(_dp)), "pmq_param_grad")
        return {"dpole": dpole.reshape(n, 13), "dpolarity": dpol, "dpdamp": dpd, "dthole": dth}

    def set_batch(self, offsets):
        """Treat the system as independent non-periodic clusters: cluster c is
        atoms offsets[c]:offsets[c+1] (offsets[0] = 0, offsets[-1] = n);
        None returns to one system."""
        if offsets is None:
            self._check(self._L.pmq_set_batch(self._ctx, 0, None), "pmq_set_batch")
            self.nsys = 0
            return
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        self._check(self._L.pmq_set_batch(self._ctx, len(off) - 1, off.ctypes.data_as(_ip)), "pmq_set_batch")
        self.nsys = len(off) - 1

    def set_batch_param_grad(self, on: bool = True):
        """Batched GRADIENT evaluations also compute parameter gradients
        (read with param_grad())."""
        self._L.pmq_set_batch_param_grad.argtypes = [C.c_void_p, C.c_int]
        self._L.pmq_set_batch_param_grad.restype = C.c_int
        self._check(self._L.pmq_set_batch_param_grad(self._ctx, 1 if on else 0), "pmq_set_batch_param_grad")

    def set_batch_shape(self, precision: int, bucket: int, lanes: int, wgs_per_cu: int = 1):
        """Fix one size bucket's shape (lanes = 0: back to the autotuner)."""
        self._L.pmq_set_batch_shape.argtypes = [C.c_void_p] + [C.c_int] * 4
        self._L.pmq_set_batch_shape.restype = C.c_int
        self._check(self._L.pmq_set_batch_shape(self._ctx, precision, bucket, lanes, wgs_per_cu), "pmq_set_batch_shape")

    def batch_tuning(self, precision: int) -> list:
        """The autotuner's shape per cluster-size bucket: dicts of rows, lanes
        per row, workgroups per CU, cluster count and timed ms (-1: none)."""
        nb = 3
        rows, lanes, wpc, ncl = [(C.c_int * nb)() for _ in range(4)]
        ms = (C.c_float * nb)()
        self._L.pmq_batch_tuning.argtypes = [C.c_void_p, C.c_int] + [C.POINTER(C.c_int)] * 4 + [C.POINTER(C.c_float)]
        self._L.pmq_batch_tuning.restype = C.c_int
        self._check(self._L.pmq_batch_tuning(self._ctx, precision, rows, lanes, wpc, ncl, ms), "pmq_batch_tuning")
        return [{"rows": rows[b], "lanes_per_row": lanes[b], "wgs_per_cu": wpc[b], "clusters": ncl[b],
                 "ms": round(float(ms[b]), 4)} for b in range(nb)]

    def fetch_batch(self) -> dict:
        ns = self.nsys
        em, ep, it = np.zeros(ns), np.zeros(ns), np.zeros(ns, dtype=np.int32)
        self._check(self._L.pmq_fetch_batch(self._ctx, em.ctypes.data_as(_dp), ep.ctypes.data_as(_dp),
                                            it.ctypes.data_as(_ip)), "pmq_fetch_batch")
        return {"em": em, "ep": ep, "pcg_iters": it}

    def compute_batch_noalloc(self, mode: int, terms: int, precision: int, em=0, ep=0, iters=0, etot=0,
                              grad=0):
        """The batched evaluation without allocation (pmq_compute_batch_noalloc):
        outputs are device addresses (integers, 0 = not wanted) of em, ep
        (double, one per cluster), iters (int32, one per cluster), etot
        (double, 2) and grad (double, 3n).  Asynchronous: sync() waits."""
        self._check(self._L.pmq_compute_batch_noalloc(self._ctx, mode, terms, precision, em or None, ep or None,
                                                      iters or None, etot or None, grad or None),
                    "pmq_compute_batch_noalloc")

    def eval(self, mode: int = L.VIRIAL, terms: int = L.TERM_MPOLE, precision: int = L.PLAIN_FP64,
             xyz: np.ndarray | None = None, want_rpole: bool = False) -> dict:
        if xyz is not None:
            self.set_coords(xyz)
        self.compute(mode, terms, precision)
        return self.fetch(want_rpole)


def _profile_bind(L):
    if getattr(L, "_prof_ready", False):
        return
    L.pmq_profile_enable.argtypes = [C.c_void_p, C.c_int]
    L.pmq_profile_enable.restype = C.c_int
    L.pmq_profile_read.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
    L.pmq_profile_read.restype = C.c_int
    L._prof_ready = True


KERNEL_FAMILIES = ("neighbors", "mpole_tiles", "rep_tiles", "disp_tiles", "pol_field_tiles", "pol_pair_fill",
                   "pol_mutual_spmv", "pol_spmv_spread", "pol_grad_tiles", "pme_spread", "pme_spread_mpole", "pme_fft", "pme_fft1",
                   "pme_gather_field", "pme_gather", "batch_clusters")


FFT_AUTO, FFT_LIBRARY, FFT_DIRECT = 0, 1, 2
FFT_GRIDS_FP32, FFT_GRIDS_FP64 = 16, 32


def set_fft(eng: Engine, engine: int = FFT_AUTO):
    """PME transform engine: a timed choice (FFT_AUTO, the default), rocFFT
    (FFT_LIBRARY), or the direct LDS transforms where the grid fits them
    (FFT_DIRECT); FFT_GRIDS_FP32 or FFT_GRIDS_FP64 or'ed in fixes the mixed
    form's grid precision."""
    L = eng._L
    L.pmq_set_fft.argtypes = [C.c_void_p, C.c_int]
    L.pmq_set_fft.restype = C.c_int
    eng._check(L.pmq_set_fft(eng._ctx, int(engine)), "pmq_set_fft")


PAIR_CARTESIAN, PAIR_QI = 0, 1


def set_pair_form(eng: Engine, form: int = PAIR_CARTESIAN):
    """Pair algebra of the multipole tiles: Cartesian (PAIR_CARTESIAN) or in
    each pair's quasi-internal frame (PAIR_QI)."""
    L = eng._L
    L.pmq_set_pair_form.argtypes = [C.c_void_p, C.c_int]
    L.pmq_set_pair_form.restype = C.c_int
    eng._check(L.pmq_set_pair_form(eng._ctx, int(form)), "pmq_set_pair_form")


def set_graphs(eng: Engine, enable: bool = True):
    """HIP graphs for repeated evaluations of one system (off by default:
    plain enqueueing on two streams is faster)."""
    L = eng._L
    L.pmq_set_graphs.argtypes = [C.c_void_p, C.c_int]
    L.pmq_set_graphs.restype = C.c_int
    eng._check(L.pmq_set_graphs(eng._ctx, int(bool(enable))), "pmq_set_graphs")


GATHER_AUTO, GATHER_ROWS, GATHER_PLANES = 0, 1, 2


def set_gather_lanes(eng: Engine, mode: int = GATHER_AUTO):
    """PME gather layout: by system size (GATHER_AUTO), one lane per stencil
    row (GATHER_ROWS) or one per stencil plane (GATHER_PLANES)."""
    L = eng._L
    L.pmq_set_gather_lanes.argtypes = [C.c_void_p, C.c_int]
    L.pmq_set_gather_lanes.restype = C.c_int
    eng._check(L.pmq_set_gather_lanes(eng._ctx, int(mode)), "pmq_set_gather_lanes")


def fft_choice(eng: Engine) -> dict:
    """The transforms chosen for the current grid (after a PME evaluation)."""
    L = eng._L
    L.pmq_fft_choice.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 3
    L.pmq_fft_choice.restype = C.c_int
    v = [C.c_int() for _ in range(3)]
    eng._check(L.pmq_fft_choice(eng._ctx, *[C.byref(x) for x in v]), "pmq_fft_choice")
    name = {-1: None, 0: "rocFFT", 1: "direct"}
    return {"fp64_grids": name[v[0].value], "mixed_grids": {-1: None, 0: "fp64", 1: "fp32"}[v[1].value],
            "fp32_grids": name[v[2].value]}


def profile_enable(eng: Engine, on: bool = True):
    _profile_bind(eng._L)
    eng._check(eng._L.pmq_profile_enable(eng._ctx, 1 if on else 0), "pmq_profile_enable")


def profile_read(eng: Engine) -> dict:
    """{family: (total_ms, launches)} of the instrumented kernel families."""
    _profile_bind(eng._L)
    out = {}
    for name in KERNEL_FAMILIES:
        ms, cnt = C.c_double(), C.c_longlong()
        eng._check(eng._L.pmq_profile_read(eng._ctx, name.encode(), C.byref(ms), C.byref(cnt)), "pmq_profile_read")
        out[name] = (ms.value, cnt.value)
    return out
