"""Pins the oracle's two damping functions that no reference test covers.

AMOEBA Thole damping (oracle/pmq_oracle_elec.c thole_lambda; the reference
only sets the parameters, kpolar.cpp:461-537).  Thole's exponential model
smears one site of a pair into the normalized density
rho(s) = 3a/(4 pi A^3) exp(-a (s/A)^3), A = (alpha_i alpha_k)^(1/6)
(= pdamp_i pdamp_k, pdamp = alpha^(1/6) at kpolar.cpp:474); the damped
radial series is that of the potential phi of rho: B_0 = phi,
B_(n+1) = -(1/r) dB_n/dr, lambda_(2n+1) = B_n r^(2n+1) / (2n-1)!!.  Here phi
is built from the enclosed charge Q(r), taken by quadrature, and the shell
theorem (phi' = -Q/r^2, so the higher derivatives need only Q and the
density's own derivatives): nothing of the oracle's polynomial in
D = a u^3 is reused.  The pair's a is min(a_i, a_k), or the larger when the
smaller is 0 (kpolar.cpp:516-517); every AMOEBA parameter file used here
gives all types a = 0.39, so a two-atom system with unequal values checks
the rule end to end, together with u, lambda_3 in the permanent field and
lambda_3, lambda_5 in the mutual field, against a closed-form 2x2 solve.

HIPPO dispersion damping (orc_hippo; f = 1.5 d5 - 0.5 d3 of the adisp
exponents, the pair energy -C_i C_k f^2 / r^6).  d3 and d5 are the damping
factors of the Coulomb interaction V(r) of two normalized Slater densities
rho(s) = a^3/(8 pi) exp(-a s) (d3 = -r^2 V', d5 = (r^3 V'' - r^2 V')/3), and
f = (r^3/2) V'' is the damping of the dipole-dipole tensor component along
the pair axis (T_zz = V'', -2/r^3 undamped).  V is the potential of the
convolution sigma = rho_i * rho_k, taken by nested quadrature; by the shell
theorem d3 = Q_sigma(r), d5 = Q_sigma - (4 pi/3) r^3 sigma(r),
f = Q_sigma - 2 pi r^3 sigma(r).  A two-atom system checks the oracle's
dispersion energy end to end.  A wrong exponent (a/2), the other sign of
the d3 term or f instead of f^2 misses these by far more than the bars.
"""
import ctypes as C

import mpmath as mp
import numpy as np
import pytest
from scipy.integrate import quad

import oracle
from polmdqc_amd import water

_dp = C.POINTER(C.c_double)


def lib():
    L = oracle.lib()
    L.orc_thole_lambda.argtypes = [C.c_double] * 4 + [_dp]
    L.orc_damp_gordon1.argtypes = [C.c_double] * 3 + [_dp] * 3
    return L


def oracle_thole(r, pdi, pdk, a):
    lam = (C.c_double * 6)()
    lib().orc_thole_lambda(r, pdi, pdk, a, lam)
    return np.array(lam[1:5])


def thole_model(r, A, a):
    """lambda_3,5,7,9 of Thole's exponential density, from its enclosed charge."""
    mp.mp.dps = 30
    r = mp.mpf(r)
    rho = lambda s: 3 * a / (4 * mp.pi * A ** 3) * mp.exp(-a * (s / A) ** 3)
    Q = mp.quad(lambda s: 4 * mp.pi * s * s * rho(s), [0, r])
    p0, p1, p2 = rho(r), mp.diff(rho, r, 1), mp.diff(rho, r, 2)
    d1 = -Q / r ** 2                                   # phi'  (shell theorem)
    d2 = 2 * Q / r ** 3 - 4 * mp.pi * p0               # phi''
    d3 = 8 * mp.pi * p0 / r - 6 * Q / r ** 4 - 4 * mp.pi * p1
    d4 = 24 * Q / r ** 5 - 32 * mp.pi * p0 / r ** 2 + 8 * mp.pi * p1 / r - 4 * mp.pi * p2
    B = [-d1 / r, d2 / r ** 2 - d1 / r ** 3, -d3 / r ** 3 + 3 * d2 / r ** 4 - 3 * d1 / r ** 5,
         d4 / r ** 4 - 6 * d3 / r ** 5 + 15 * d2 / r ** 6 - 15 * d1 / r ** 7]
    return np.array([float(B[n] * r ** (2 * n + 3) / [1, 3, 15, 105][n]) for n in range(4)])


@pytest.mark.parametrize("ai,ak,a", [(0.837, 0.496, 0.39), (0.837, 0.837, 0.39), (1.334, 0.496, 0.2)])
def test_thole_factors_are_the_smeared_charge_field(ai, ak, a):
    pdi, pdk = ai ** (1 / 6), ak ** (1 / 6)
    for r in (0.6, 1.0, 1.6, 2.5, 4.0):
        model = thole_model(r, pdi * pdk, a)
        got = oracle_thole(r, pdi, pdk, a)
        np.testing.assert_allclose(got, model, rtol=0, atol=1e-13)
    # the factors are not trivially 1 in the range tested
    assert thole_model(1.0, pdi * pdk, a)[0] < 0.9


def two_sites(x_b, q, alpha, thole, pen=None):
    """Two unbonded point sites (frame NONE), charges q, polarizabilities alpha."""
    s = dict(water.water_system(water.cluster(1, np.random.default_rng(0)), "water03"))
    n = 2
    s.update({"n": n, "x": np.array([0.0, 0.0]), "y": np.array([0.0, 0.0]), "z": np.array([0.0, x_b]),
              "pole": np.zeros((n, 13)), "polaxe": np.zeros(n, np.int32), "zaxis": np.zeros(n, np.int32),
              "xaxis": np.zeros(n, np.int32), "yaxis": np.zeros(n, np.int32), "use": np.ones(n + 1, np.int32),
              "molcule": np.arange(n, dtype=np.int32), "pcore": np.zeros(n), "pval0": np.zeros(n),
              "palpha": np.ones(n), "mono0": np.asarray(q, float).copy(), "type": np.array([1, 2], np.int32),
              "polarity": np.asarray(alpha, float), "thole": np.asarray(thole, float),
              "pdamp": np.asarray(alpha, float) ** (1 / 6), "poleps": 1e-13, "pcgprec": 0,
              "csix": np.zeros(n), "adisp": np.ones(n), "sizpr": np.zeros(n), "dmppr": np.ones(n),
              "elepr": np.zeros(n)})
    s["pole"][:, 0] = q
    s["pole"] = s["pole"].ravel()
    s["use"][0] = 0
    for k in ("12", "13", "14", "15"):
        s["n" + k], s["i" + k] = np.zeros(n, np.int32), np.zeros(0, np.int32)
    s["np11"], s["ip11"] = np.ones(n, np.int32), np.arange(n, dtype=np.int32)
    for k in ("12", "13", "14"):
        s["np" + k], s["ip" + k] = np.zeros(n, np.int32), np.zeros(0, np.int32)
    return s


@pytest.mark.parametrize("r", [1.2, 2.0, 3.0])
def test_two_site_polarization_energy_closed_form(r):
    q, alpha, thole = np.array([0.7, -0.4]), np.array([0.9, 1.3]), np.array([0.39, 0.2])
    s = two_sites(r, q, alpha, thole)
    a = min(thole)  # the pair's Thole value (kpolar.cpp:516-517)
    l3, l5 = thole_model(r, (alpha[0] * alpha[1]) ** (1 / 6), a)[:2]
    # on the axis: E_A = -q_B l3/r^2, E_B = q_A l3/r^2, mutual coupling t = (3 l5 - l3)/r^3
    EA, EB, t = -q[1] * l3 / r ** 2, q[0] * l3 / r ** 2, (3 * l5 - l3) / r ** 3
    M = np.array([[1.0 / alpha[0], -t], [-t, 1.0 / alpha[1]]])
    mu = np.linalg.solve(M, np.array([EA, EB]))
    expect = -0.5 * float(s["electric"]) * (mu[0] * EA + mu[1] * EB)
    got = oracle.elec(s, want_grad=False)
    assert abs(got["ep"] - expect) <= 1e-11 * abs(expect)
    # with a = max instead the energy differs clearly (the rule matters)
    l3m, l5m = thole_model(r, (alpha[0] * alpha[1]) ** (1 / 6), max(thole))[:2]
    assert abs(l3m - l3) > 1e-4


def slater(a):
    return lambda s: a ** 3 / (8 * np.pi) * np.exp(-a * s)


def sigma(t, ri, rk):
    """(rho_i * rho_k)(t) by quadrature over the sphere of radius |s|."""
    def inner(s):
        I = quad(lambda w: w * rk(w), abs(t - s), t + s, epsabs=0, epsrel=1e-13, limit=200)[0]
        return 2 * np.pi * s * ri(s) * I / t
    return quad(inner, 0, np.inf, epsabs=0, epsrel=1e-13, limit=200)[0]


def overlap_damping(r, ai, ak):
    """d3, d5 and f of two normalized Slater densities (shell theorem on sigma)."""
    ri, rk = slater(ai), slater(ak)
    Q = quad(lambda t: 4 * np.pi * t * t * sigma(t, ri, rk), 0, r, epsabs=0, epsrel=1e-12, limit=200)[0]
    sr = sigma(r, ri, rk)
    return Q, Q - 4 * np.pi / 3 * r ** 3 * sr, Q - 2 * np.pi * r ** 3 * sr


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
@pytest.mark.parametrize("ai,ak", [(4.4, 4.4), (4.4, 3.2)])
def test_dispersion_damping_is_the_slater_overlap(ai, ak):
    for r in (1.0, 2.2):
        d3, d5, f = overlap_damping(r, ai, ak)
        dmpi, dmpk, dmpik = (C.c_double * 5)(), (C.c_double * 5)(), (C.c_double * 6)()
        lib().orc_damp_gordon1(r, ai, ak, dmpi, dmpk, dmpik)
        assert abs(dmpik[1] - d3) < 1e-10 and abs(dmpik[2] - d5) < 1e-10
        assert abs(1.5 * dmpik[2] - 0.5 * dmpik[1] - f) < 1e-10
        # a halved exponent or the other sign of the d3 term would not pass
        lib().orc_damp_gordon1(r, ai / 2, ak / 2, dmpi, dmpk, dmpik)
        assert abs(1.5 * dmpik[2] - 0.5 * dmpik[1] - f) > 1e-3
        assert abs(1.5 * d5 + 0.5 * d3 - f) > 1e-3


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
def test_two_site_dispersion_energy():
    r, ai, ak, ci, ck = 2.2, 4.4, 3.2, 12.0, 7.5
    s = dict(water.water_system(water.cluster(1, np.random.default_rng(0)), "hippo19"))
    t = two_sites(r, [0.0, 0.0], [0.0, 0.0], [0.0, 0.0])
    t.update({"csix": np.array([ci, ck]), "adisp": np.array([ai, ak]), "dispcut": 50.0, "disptaper": 50.0,
              "repcut": 6.0, "use_dewald": 0})
    for k in ("dsp2scale", "dsp3scale", "dsp4scale", "dsp5scale", "r2scale", "r3scale", "r4scale", "r5scale"):
        t[k] = s[k]
    f = overlap_damping(r, ai, ak)[2]
    expect = -ci * ck * f * f / r ** 6
    got = oracle.hippo_terms(t, want_grad=False)["edsp"]
    assert abs(got - expect) <= 1e-10 * abs(expect)
    assert abs(got - (-ci * ck * f / r ** 6)) > 1e-3 * abs(expect)  # f^2, not f
