"""FP64 numpy restatement of the HIPPO dispersion term (test infrastructure:
used only by tests, smoke() and the CPU baseline, never by the library).

The reference computes nothing for this term: `edisp` is absent
(src/md/energy.cpp:280, mechanic.cpp:111); only its parameters exist
(`dspsix`/`dspdmp` per class, kdsp.h:18-19; `dsp2..5scale`, dsppot.h:21-25;
DISPERSION lines parsed at readprm.cpp:1047).  The physics follows the paper
the parameter file cites (Rackers et al., JCP 149, 084115 (2018)):

    E = - sum_{i<k} s_ik C_i C_k f_ik(r)^2 / r^6,   f = 1.5 d5 - 0.5 d3

where d3, d5 are the r^-3 and r^-5 factors of the Gordon overlap damping of
two exponential densities with exponents adisp_i, adisp_k (the same
functions as the multipole charge-penetration damping of damping.h:32-196),
and s_ik the 1-2..1-5 scale (a later class overrides an earlier one).

With dispersion Ewald (coefficient a), r^-6 = [g(ar) + (1 - g(ar))] r^-6 with
g(x) = (1 + x^2 + x^4/2) exp(-x^2):
    real space   - C_i C_k [s f^2 - 1 + g(a r)] / r^6      (pairs within the cutoff)
    reciprocal   - (1/2V) sum_k psi(k) |S(k)|^2,  S(k) = sum_j C_j exp(i k.r_j),
                 psi(k) = (2 pi^1.5 a^3 / 3) [sqrt(pi) b^3 erfc(b) + (1/2 - b^2) exp(-b^2)],
                 b = |k| / (2a), the k = 0 term included (psi(0) = pi^1.5 a^3 / 3)
    self         + a^6 / 12 sum_i C_i^2
psi is the Fourier transform of (1 - g) r^-6 (checked here by quadrature in
the tests).  Parity with the reference is unpinned (it has no such term);
the tests pin the restatement by the exact lattice sum, independence of a,
and finite differences.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfc


def gordon_d3_d5_d7(r, ai, ak):
    """r^-3, r^-5, r^-7 damping factors of two exponential densities (Gordon
    form; equal exponents within 1e-3 use the one-exponent limit)."""
    r, ai, ak = np.broadcast_arrays(np.asarray(r, float), np.asarray(ai, float), np.asarray(ak, float))
    di, dk = ai * r, ak * r
    ei, ek = np.exp(-di), np.exp(-dk)
    same = np.abs(ai - ak) < 1e-3
    with np.errstate(divide="ignore", invalid="ignore"):
        ai2, ak2 = ai * ai, ak * ak
        ti = np.where(same, 0.0, ak2 / (ak2 - ai2))
        tk = np.where(same, 0.0, ai2 / (ai2 - ak2))
    ti2, tk2 = ti * ti, tk * tk
    d3u = (1 - ti2 * (1 + di + 0.5 * di**2) * ei - tk2 * (1 + dk + 0.5 * dk**2) * ek
           - 2 * ti2 * tk * (1 + di) * ei - 2 * tk2 * ti * (1 + dk) * ek)
    d5u = (1 - ti2 * (1 + di + 0.5 * di**2 + di**3 / 6) * ei - tk2 * (1 + dk + 0.5 * dk**2 + dk**3 / 6) * ek
           - 2 * ti2 * tk * (1 + di + di**2 / 3) * ei - 2 * tk2 * ti * (1 + dk + dk**2 / 3) * ek)
    d7u = (1 - ti2 * (1 + di + 0.5 * di**2 + di**3 / 6 + di**4 / 30) * ei
           - tk2 * (1 + dk + 0.5 * dk**2 + dk**3 / 6 + dk**4 / 30) * ek
           - 2 * ti2 * tk * (1 + di + 2 * di**2 / 5 + di**3 / 15) * ei
           - 2 * tk2 * ti * (1 + dk + 2 * dk**2 / 5 + dk**3 / 15) * ek)
    b5 = 1 + di + 0.5 * di**2 + di**3 / 6 + di**4 / 24 + di**5 / 120 + di**6 / 720
    d3s = 1 - (1 + di + 0.5 * di**2 + 7 * di**3 / 48 + di**4 / 48) * ei
    d5s = 1 - (1 + di + 0.5 * di**2 + di**3 / 6 + di**4 / 24 + di**5 / 144) * ei
    d7s = 1 - b5 * ei
    return np.where(same, d3s, d3u), np.where(same, d5s, d5u), np.where(same, d7s, d7u)


def pair_scales(s, key: str):
    """Dense n x n scale matrix from the 1-2..1-5 lists (key: 'dsp' or 'r')."""
    n = int(s["n"])
    sc = np.ones((n, n))
    for q, name in ((2, "12"), (3, "13"), (4, "14"), (5, "15")):
        cnt, lst = np.asarray(s["n" + name]), np.asarray(s["i" + name])
        val = float(s[f"{key}{q}scale"])
        o = np.concatenate([[0], np.cumsum(cnt)])
        for i in range(n):
            for k in lst[o[i]:o[i + 1]]:
                sc[i, k] = sc[k, i] = val
    return sc


def _box(s):
    L = np.array([float(s["xbox"]), float(s["ybox"]), float(s["zbox"])])
    return L if int(s.get("use_bounds", 0)) else None


def edisp(s, want_grad: bool = True, kcut: float = 8.0):
    """Dispersion energy (kcal/mol) and gradient (n, 3) of an orthogonal or
    non-periodic system dict."""
    n = int(s["n"])
    xyz = np.stack([s["x"], s["y"], s["z"]], 1).astype(float)
    C = np.asarray(s["csix"], float)
    A = np.asarray(s["adisp"], float)
    L = _box(s)
    ew = int(s.get("use_dewald", 0)) and L is not None
    a = float(s.get("dewald", 0.0))
    cut = float(s.get("dispcut", 1e12))
    sc = pair_scales(s, "dsp")
    iu, ku = np.triu_indices(n, 1)
    d = xyz[ku] - xyz[iu]
    if L is not None:
        d -= L * np.round(d / L)
    r2 = np.sum(d * d, 1)
    m = r2 <= cut * cut
    iu, ku, d, r2 = iu[m], ku[m], d[m], r2[m]
    r = np.sqrt(r2)
    d3, d5, d7 = gordon_d3_d5_d7(r, A[iu], A[ku])
    f = 1.5 * d5 - 0.5 * d3
    fp = 1.5 * 5 * (d5 - d7) / r - 0.5 * 3 * (d3 - d5) / r      # df/dr
    cc = C[iu] * C[ku]
    sik = sc[iu, ku]
    r6 = r2**3
    if ew:
        x2 = (a * r) ** 2
        ex = np.exp(-x2)
        g = (1 + x2 + 0.5 * x2 * x2) * ex
        gp = -a * a * r * x2 * x2 * ex * 2 / 2                     # dg/dr = -a^2 r x^4 e^{-x}
        e_pair = -cc * (sik * f * f - 1 + g) / r6
        de = -cc * ((2 * sik * f * fp + gp) / r6 - 6 * (sik * f * f - 1 + g) / (r6 * r))
    else:
        e_pair = -cc * sik * f * f / r6
        de = -cc * sik * (2 * f * fp / r6 - 6 * f * f / (r6 * r))
    e = float(e_pair.sum())
    grad = np.zeros((n, 3))
    if want_grad:
        gv = (de / r)[:, None] * d          # dE/dr_k along d = r_k - r_i
        np.add.at(grad, ku, gv)
        np.add.at(grad, iu, -gv)
    if ew:
        V = float(np.prod(L))
        mmax = np.ceil(kcut * a * L / np.pi).astype(int)
        ms = np.stack(np.meshgrid(*[np.arange(-mm, mm + 1) for mm in mmax], indexing="ij"), -1).reshape(-1, 3)
        kv = 2 * np.pi * ms / L
        kk = np.sqrt(np.sum(kv * kv, 1))
        b = kk / (2 * a)
        psi = 2 * np.pi**1.5 * a**3 / 3 * (np.sqrt(np.pi) * b**3 * erfc(b) + (0.5 - b * b) * np.exp(-b * b))
        ph = xyz @ kv.T                     # (n, nk)
        cs, sn = np.cos(ph), np.sin(ph)
        Sr, Si = C @ cs, C @ sn
        e += float(-np.sum(psi * (Sr * Sr + Si * Si)) / (2 * V))
        e += a**6 / 12 * float(np.sum(C * C))
        if want_grad:
            # dE/dr_j = (1/V) sum_k psi k C_j Im(S* e^{ik.r_j})
            im = Sr[None, :] * sn - Si[None, :] * cs
            grad += (C[:, None] / V) * ((im * psi[None, :]) @ kv)
    return e, grad


def edisp_lattice(s, rmax: float):
    """Direct lattice sum of the dispersion energy of a small periodic box
    over all images within rmax (undamped tail ignored beyond rmax)."""
    n = int(s["n"])
    xyz = np.stack([s["x"], s["y"], s["z"]], 1).astype(float)
    C = np.asarray(s["csix"], float)
    A = np.asarray(s["adisp"], float)
    L = _box(s)
    sc = pair_scales(s, "dsp")
    nm = np.ceil(rmax / L).astype(int)
    e = 0.0
    for ix in range(-nm[0], nm[0] + 1):
        for iy in range(-nm[1], nm[1] + 1):
            for iz in range(-nm[2], nm[2] + 1):
                sh = np.array([ix, iy, iz]) * L
                home = ix == 0 and iy == 0 and iz == 0
                d = xyz[None, :, :] + sh - xyz[:, None, :]
                r = np.sqrt(np.sum(d * d, 2))
                mask = (r <= rmax) & (r > 0)
                if not mask.any():
                    continue
                i, k = np.nonzero(mask)
                d3, d5, _ = gordon_d3_d5_d7(r[i, k], A[i], A[k])
                f = 1.5 * d5 - 0.5 * d3
                w = np.where(home, sc[i, k], 1.0)
                e += float(np.sum(-0.5 * w * C[i] * C[k] * f * f / r[i, k] ** 6))
    return e
