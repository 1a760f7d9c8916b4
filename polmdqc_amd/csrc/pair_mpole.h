// pair_mpole.h -- device pair algebra of the point-multipole interaction.
//
// One routine serves every radial law: the five angular "terms" of the
// charge/dipole/quadrupole pair energy (pairMpole.h:48-73) are contracted
// with a radial series R[0..5] that must satisfy dR_n/dr = -r R_{n+1}.
// Undamped (rr1..rr11, pairMpole.h:36-46), Gordon-damped (damping.h) and
// Ewald-screened series all do, so energy, gradient (pairMpole.h:91-106) and
// torques (pairMpole.h:108-164) come from the same code.
//
// Convention: r = x_k - x_i; frc = dE/dx_i (the reference adds it to dem[i]
// and subtracts it from dem[k]); trq = torque on each site's multipole.
#pragma once

#include <hip/hip_runtime.h>

namespace pmq {

template <class R>
struct Mp {  // c, d[3], q{xx,xy,xz,yy,yz,zz}
    R c, dx, dy, dz, qxx, qxy, qxz, qyy, qyz, qzz;
};

// component c of a multipole (c, dx, dy, dz, qxx, qxy, qxz, qyy, qyz, qzz)
template <class R>
__device__ __forceinline__ R mp_get(const Mp<R> &m, int c)
{
    switch (c) {
    case 0: return m.c;
    case 1: return m.dx;
    case 2: return m.dy;
    case 3: return m.dz;
    case 4: return m.qxx;
    case 5: return m.qxy;
    case 6: return m.qxz;
    case 7: return m.qyy;
    case 8: return m.qyz;
    default: return m.qzz;
    }
}

template <class R>
__device__ __forceinline__ void mp_set(Mp<R> &m, int c, R v)
{
    switch (c) {
    case 0: m.c = v; break;
    case 1: m.dx = v; break;
    case 2: m.dy = v; break;
    case 3: m.dz = v; break;
    case 4: m.qxx = v; break;
    case 5: m.qxy = v; break;
    case 6: m.qxz = v; break;
    case 7: m.qyy = v; break;
    case 8: m.qyz = v; break;
    default: m.qzz = v; break;
    }
}

template <class R>
struct PairOut {
    R e;
    R fx, fy, fz;
    R tix, tiy, tiz, tkx, tky, tkz;
    R phii, phik; // dE/dc_i, dE/dc_k (charge-flux potentials)
};

template <class R, bool GRAD, bool ACCUM = false>
__device__ __forceinline__ void pair_series(const R *Rs, const Mp<R> &a, const Mp<R> &b, R xr, R yr, R zr,
                                            PairOut<R> &o)
{
    // Q r for both sites
    const R qix = a.qxx * xr + a.qxy * yr + a.qxz * zr;
    const R qiy = a.qxy * xr + a.qyy * yr + a.qyz * zr;
    const R qiz = a.qxz * xr + a.qyz * yr + a.qzz * zr;
    const R qkx = b.qxx * xr + b.qxy * yr + b.qxz * zr;
    const R qky = b.qxy * xr + b.qyy * yr + b.qyz * zr;
    const R qkz = b.qxz * xr + b.qyz * yr + b.qzz * zr;
    const R dir = a.dx * xr + a.dy * yr + a.dz * zr;
    const R dkr = b.dx * xr + b.dy * yr + b.dz * zr;
    const R qir = qix * xr + qiy * yr + qiz * zr;
    const R qkr = qkx * xr + qky * yr + qkz * zr;
    const R dik = a.dx * b.dx + a.dy * b.dy + a.dz * b.dz;
    const R qik = qix * qkx + qiy * qky + qiz * qkz;
    const R diqk = a.dx * qkx + a.dy * qky + a.dz * qkz;
    const R dkqi = b.dx * qix + b.dy * qiy + b.dz * qiz;
    const R qiqk = R(2) * (a.qxy * b.qxy + a.qxz * b.qxz + a.qyz * b.qyz) + a.qxx * b.qxx + a.qyy * b.qyy
                 + a.qzz * b.qzz;
    const R t1 = a.c * b.c;
    const R t2 = b.c * dir - a.c * dkr + dik;
    const R t3 = a.c * qkr + b.c * qir - dir * dkr + R(2) * (dkqi - diqk + qiqk);
    const R t4 = dir * qkr - dkr * qir - R(4) * qik;
    const R t5 = qir * qkr;
    const R e = t1 * Rs[0] + t2 * Rs[1] + t3 * Rs[2] + t4 * Rs[3] + t5 * Rs[4];
    const R phii = b.c * Rs[0] - dkr * Rs[1] + qkr * Rs[2];
    const R phik = a.c * Rs[0] + dir * Rs[1] + qir * Rs[2];
    if (ACCUM) { o.e += e; o.phii += phii; o.phik += phik; }
    else { o.e = e; o.phii = phii; o.phik = phik; }
    if (!GRAD) return;

    const R de = t1 * Rs[1] + t2 * Rs[2] + t3 * Rs[3] + t4 * Rs[4] + t5 * Rs[5];
    const R g1 = -b.c * Rs[1] + dkr * Rs[2] - qkr * Rs[3];
    const R g2 = a.c * Rs[1] + dir * Rs[2] + qir * Rs[3];
    const R g3 = R(2) * Rs[2];
    const R g4 = R(2) * (-b.c * Rs[2] + dkr * Rs[3] - qkr * Rs[4]);
    const R g5 = R(2) * (-a.c * Rs[2] - dir * Rs[3] - qir * Rs[4]);
    const R g6 = R(4) * Rs[3];

    // Q_i Q_k r, Q_k Q_i r, Q_k d_i, Q_i d_k
    const R qixk = a.qxx * qkx + a.qxy * qky + a.qxz * qkz;
    const R qiyk = a.qxy * qkx + a.qyy * qky + a.qyz * qkz;
    const R qizk = a.qxz * qkx + a.qyz * qky + a.qzz * qkz;
    const R qkxi = b.qxx * qix + b.qxy * qiy + b.qxz * qiz;
    const R qkyi = b.qxy * qix + b.qyy * qiy + b.qyz * qiz;
    const R qkzi = b.qxz * qix + b.qyz * qiy + b.qzz * qiz;
    const R diqkx = a.dx * b.qxx + a.dy * b.qxy + a.dz * b.qxz;
    const R diqky = a.dx * b.qxy + a.dy * b.qyy + a.dz * b.qyz;
    const R diqkz = a.dx * b.qxz + a.dy * b.qyz + a.dz * b.qzz;
    const R dkqix = b.dx * a.qxx + b.dy * a.qxy + b.dz * a.qxz;
    const R dkqiy = b.dx * a.qxy + b.dy * a.qyy + b.dz * a.qyz;
    const R dkqiz = b.dx * a.qxz + b.dy * a.qyz + b.dz * a.qzz;

    const R fx = de * xr + g1 * a.dx + g2 * b.dx + g3 * (diqkx - dkqix) + g4 * qix + g5 * qkx + g6 * (qixk + qkxi);
    const R fy = de * yr + g1 * a.dy + g2 * b.dy + g3 * (diqky - dkqiy) + g4 * qiy + g5 * qky + g6 * (qiyk + qkyi);
    const R fz = de * zr + g1 * a.dz + g2 * b.dz + g3 * (diqkz - dkqiz) + g4 * qiz + g5 * qkz + g6 * (qizk + qkzi);

    // cross products
    const R dikx = a.dy * b.dz - a.dz * b.dy, diky = a.dz * b.dx - a.dx * b.dz, dikz = a.dx * b.dy - a.dy * b.dx;
    const R dirx = a.dy * zr - a.dz * yr, diry = a.dz * xr - a.dx * zr, dirz = a.dx * yr - a.dy * xr;
    const R dkrx = b.dy * zr - b.dz * yr, dkry = b.dz * xr - b.dx * zr, dkrz = b.dx * yr - b.dy * xr;
    const R qirx = qiz * yr - qiy * zr, qiry = qix * zr - qiz * xr, qirz = qiy * xr - qix * yr;
    const R qkrx = qkz * yr - qky * zr, qkry = qkx * zr - qkz * xr, qkrz = qky * xr - qkx * yr;
    const R qikx = qky * qiz - qkz * qiy, qiky = qkz * qix - qkx * qiz, qikz = qkx * qiy - qky * qix;
    const R qikrx = qizk * yr - qiyk * zr, qikry = qixk * zr - qizk * xr, qikrz = qiyk * xr - qixk * yr;
    const R qkirx = qkzi * yr - qkyi * zr, qkiry = qkxi * zr - qkzi * xr, qkirz = qkyi * xr - qkxi * yr;
    const R diqkrx = diqkz * yr - diqky * zr, diqkry = diqkx * zr - diqkz * xr, diqkrz = diqky * xr - diqkx * yr;
    const R dkqirx = dkqiz * yr - dkqiy * zr, dkqiry = dkqix * zr - dkqiz * xr, dkqirz = dkqiy * xr - dkqix * yr;
    // antisymmetric part of Q_i Q_k contracted with epsilon, plus the d x Qr cross terms
    const R dqikx = a.dy * qkz - a.dz * qky + b.dy * qiz - b.dz * qiy
                  - R(2) * (a.qxy * b.qxz + a.qyy * b.qyz + a.qyz * b.qzz - a.qxz * b.qxy - a.qyz * b.qyy - a.qzz * b.qyz);
    const R dqiky = a.dz * qkx - a.dx * qkz + b.dz * qix - b.dx * qiz
                  - R(2) * (a.qxz * b.qxx + a.qyz * b.qxy + a.qzz * b.qxz - a.qxx * b.qxz - a.qxy * b.qyz - a.qxz * b.qzz);
    const R dqikz = a.dx * qky - a.dy * qkx + b.dx * qiy - b.dy * qix
                  - R(2) * (a.qxx * b.qxy + a.qxy * b.qyy + a.qxz * b.qyz - a.qxy * b.qxx - a.qyy * b.qxy - a.qyz * b.qxz);

    const R tix = -Rs[1] * dikx + g1 * dirx + g3 * (dqikx + dkqirx) - g4 * qirx - g6 * (qikrx + qikx);
    const R tiy = -Rs[1] * diky + g1 * diry + g3 * (dqiky + dkqiry) - g4 * qiry - g6 * (qikry + qiky);
    const R tiz = -Rs[1] * dikz + g1 * dirz + g3 * (dqikz + dkqirz) - g4 * qirz - g6 * (qikrz + qikz);
    const R tkx = Rs[1] * dikx + g2 * dkrx - g3 * (dqikx + diqkrx) - g5 * qkrx - g6 * (qkirx - qikx);
    const R tky = Rs[1] * diky + g2 * dkry - g3 * (dqiky + diqkry) - g5 * qkry - g6 * (qkiry - qiky);
    const R tkz = Rs[1] * dikz + g2 * dkrz - g3 * (dqikz + diqkrz) - g5 * qkrz - g6 * (qkirz - qikz);
    if (ACCUM) {
        o.fx += fx; o.fy += fy; o.fz += fz;
        o.tix += tix; o.tiy += tiy; o.tiz += tiz;
        o.tkx += tkx; o.tky += tky; o.tkz += tkz;
    } else {
        o.fx = fx; o.fy = fy; o.fz = fz;
        o.tix = tix; o.tiy = tiy; o.tiz = tiz;
        o.tkx = tkx; o.tky = tky; o.tkz = tkz;
    }
}

// undamped series with the exclusion scale and unit factor in rr1
template <class R>
__device__ __forceinline__ void rr_series(R r2, R scale_f, R *Rs)
{
    const R rinv = R(1) / sqrt(r2);
    const R rinv2 = rinv * rinv;
    Rs[0] = scale_f * rinv;
    Rs[1] = Rs[0] * rinv2;
    Rs[2] = R(3) * Rs[1] * rinv2;
    Rs[3] = R(5) * Rs[2] * rinv2;
    Rs[4] = R(7) * Rs[3] * rinv2;
    Rs[5] = R(9) * Rs[4] * rinv2;
}

// valence-valence (overlap) factors of the Gordon1 model, orders 1..11
// (damping.h:32-196): shared by charge penetration, the HIPPO dipole-dipole
// damping and the dispersion damping
template <class R>
__device__ __forceinline__ void overlap_gordon1(R ai_, R ak_, R di_, R dk_, R ei_, R ek_, bool same, R *dmpik)
{
    // symmetric in the pair: evaluated with the smaller exponent first so
    // that both orders of a pair round alike (the unequal-exponent form
    // cancels terms up to ~1e5 times the result when the exponents are close)
    const bool sw = ai_ > ak_;
    const R ai = sw ? ak_ : ai_, ak = sw ? ai_ : ak_, di = sw ? dk_ : di_, dk = sw ? di_ : dk_;
    const R ei = sw ? ek_ : ei_, ek = sw ? ei_ : ek_;
    const R di2 = di * di, di3 = di2 * di, di4 = di2 * di2, di5 = di4 * di;
    const R dk2 = dk * dk, dk3 = dk2 * dk, dk4 = dk2 * dk2, dk5 = dk4 * dk;
    if (same) {
        const R di6 = di3 * di3, di7 = di3 * di4, di8 = di4 * di4;
        const R b5 = R(1) + di + R(0.5) * di2 + di3 / R(6) + di4 / R(24) + di5 / R(120) + di6 / R(720);
        dmpik[0] = R(1) - (R(1) + R(11) * di / R(16) + R(3) * di2 / R(16) + di3 / R(48)) * ei;
        dmpik[1] = R(1) - (R(1) + di + R(0.5) * di2 + R(7) * di3 / R(48) + di4 / R(48)) * ei;
        dmpik[2] = R(1) - (R(1) + di + R(0.5) * di2 + di3 / R(6) + di4 / R(24) + di5 / R(144)) * ei;
        dmpik[3] = R(1) - b5 * ei;
        dmpik[4] = R(1) - (b5 + di7 / R(5040)) * ei;
        dmpik[5] = R(1) - (b5 + di7 / R(5040) + di8 / R(45360)) * ei;
    } else {
        const R ai2 = ai * ai, ak2 = ak * ak;
        const R ti = ak2 / (ak2 - ai2), tk = ai2 / (ai2 - ak2);
        const R ti2 = ti * ti, tk2 = tk * tk;
        const R di6 = di3 * di3, dk6 = dk3 * dk3;
        dmpik[0] = R(1) - ti2 * (R(1) + R(2) * tk + R(0.5) * di) * ei - tk2 * (R(1) + R(2) * ti + R(0.5) * dk) * ek;
        dmpik[1] = R(1) - ti2 * (R(1) + di + R(0.5) * di2) * ei - tk2 * (R(1) + dk + R(0.5) * dk2) * ek
                 - R(2) * ti2 * tk * (R(1) + di) * ei - R(2) * tk2 * ti * (R(1) + dk) * ek;
        dmpik[2] = R(1) - ti2 * (R(1) + di + R(0.5) * di2 + di3 / R(6)) * ei
                 - tk2 * (R(1) + dk + R(0.5) * dk2 + dk3 / R(6)) * ek
                 - R(2) * ti2 * tk * (R(1) + di + di2 / R(3)) * ei - R(2) * tk2 * ti * (R(1) + dk + dk2 / R(3)) * ek;
        dmpik[3] = R(1) - ti2 * (R(1) + di + R(0.5) * di2 + di3 / R(6) + di4 / R(30)) * ei
                 - tk2 * (R(1) + dk + R(0.5) * dk2 + dk3 / R(6) + dk4 / R(30)) * ek
                 - R(2) * ti2 * tk * (R(1) + di + R(2) * di2 / R(5) + di3 / R(15)) * ei
                 - R(2) * tk2 * ti * (R(1) + dk + R(2) * dk2 / R(5) + dk3 / R(15)) * ek;
        dmpik[4] = R(1) - ti2 * (R(1) + di + R(0.5) * di2 + di3 / R(6) + R(4) * di4 / R(105) + di5 / R(210)) * ei
                 - tk2 * (R(1) + dk + R(0.5) * dk2 + dk3 / R(6) + R(4) * dk4 / R(105) + dk5 / R(210)) * ek
                 - R(2) * ti2 * tk * (R(1) + di + R(3) * di2 / R(7) + R(2) * di3 / R(21) + di4 / R(105)) * ei
                 - R(2) * tk2 * ti * (R(1) + dk + R(3) * dk2 / R(7) + R(2) * dk3 / R(21) + dk4 / R(105)) * ek;
        dmpik[5] = R(1)
                 - ti2 * (R(1) + di + R(0.5) * di2 + di3 / R(6) + R(5) * di4 / R(126) + R(2) * di5 / R(315) + di6 / R(1890)) * ei
                 - tk2 * (R(1) + dk + R(0.5) * dk2 + dk3 / R(6) + R(5) * dk4 / R(126) + R(2) * dk5 / R(315) + dk6 / R(1890)) * ek
                 - R(2) * ti2 * tk * (R(1) + di + R(4) * di2 / R(9) + di3 / R(9) + di4 / R(63) + di5 / R(945)) * ei
                 - R(2) * tk2 * ti * (R(1) + dk + R(4) * dk2 / R(9) + dk3 / R(9) + dk4 / R(63) + dk5 / R(945)) * ek;
    }
}

// Gordon charge-penetration damping factors (damping.h:32-196): dmpi damps
// the valence multipoles of i seen by core k, dmpk the reverse, dmpik the
// valence-valence interaction.  Orders 1..11 of the radial series.
template <class R, int PEN>
__device__ __forceinline__ void damp_gordon(R r, R ai, R ak, R *dmpi, R *dmpk, R *dmpik)
{
    const R di = ai * r, dk = ak * r;
    const R ei = exp(-di), ek = exp(-dk);
    const bool same = fabs(ai - ak) < R(1e-3);
    const R di2 = di * di, di3 = di2 * di, di4 = di2 * di2, di5 = di4 * di;
    const R dk2 = dk * dk, dk3 = dk2 * dk, dk4 = dk2 * dk2, dk5 = dk4 * dk;
    if (PEN == 1) {
        dmpi[0] = R(1) - (R(1) + R(0.5) * di) * ei;
        dmpi[1] = R(1) - (R(1) + di + R(0.5) * di2) * ei;
        dmpi[2] = R(1) - (R(1) + di + R(0.5) * di2 + di3 / R(6)) * ei;
        dmpi[3] = R(1) - (R(1) + di + R(0.5) * di2 + di3 / R(6) + di4 / R(30)) * ei;
        dmpi[4] = R(1) - (R(1) + di + R(0.5) * di2 + di3 / R(6) + R(4) * di4 / R(105) + di5 / R(210)) * ei;
        if (same) {
            for (int j = 0; j < 5; j++) dmpk[j] = dmpi[j];
        } else {
            dmpk[0] = R(1) - (R(1) + R(0.5) * dk) * ek;
            dmpk[1] = R(1) - (R(1) + dk + R(0.5) * dk2) * ek;
            dmpk[2] = R(1) - (R(1) + dk + R(0.5) * dk2 + dk3 / R(6)) * ek;
            dmpk[3] = R(1) - (R(1) + dk + R(0.5) * dk2 + dk3 / R(6) + dk4 / R(30)) * ek;
            dmpk[4] = R(1) - (R(1) + dk + R(0.5) * dk2 + dk3 / R(6) + R(4) * dk4 / R(105) + dk5 / R(210)) * ek;
        }
        // in FP64 whatever R is: the unequal-exponent form cancels terms up to
        // ~1e5 times its result when the exponents are close (O and H of
        // HIPPO water differ by 1%), which FP32 cannot carry
        double d6[6];
        const double rd = (double)r, aid = (double)ai, akd = (double)ak;
        overlap_gordon1<double>(aid, akd, aid * rd, akd * rd, exp(-aid * rd), exp(-akd * rd), same, d6);
        for (int j = 0; j < 6; j++) dmpik[j] = R(d6[j]);
    } else {
        dmpi[0] = R(1) - ei;
        dmpi[1] = R(1) - (R(1) + di) * ei;
        dmpi[2] = R(1) - (R(1) + di + di2 / R(3)) * ei;
        dmpi[3] = R(1) - (R(1) + di + R(0.4) * di2 + di3 / R(15)) * ei;
        dmpi[4] = R(0);
        if (same) {
            for (int j = 0; j < 5; j++) dmpk[j] = dmpi[j];
        } else {
            dmpk[0] = R(1) - ek;
            dmpk[1] = R(1) - (R(1) + dk) * ek;
            dmpk[2] = R(1) - (R(1) + dk + dk2 / R(3)) * ek;
            dmpk[3] = R(1) - (R(1) + dk + R(0.4) * dk2 + dk3 / R(15)) * ek;
            dmpk[4] = R(0);
        }
        if (same) {
            const R di6 = di3 * di3;
            dmpik[0] = R(1) - (R(1) + R(0.5) * di) * ei;
            dmpik[1] = R(1) - (R(1) + di + R(0.5) * di2) * ei;
            dmpik[2] = R(1) - (R(1) + di + R(0.5) * di2 + di3 / R(6)) * ei;
            dmpik[3] = R(1) - (R(1) + di + R(0.5) * di2 + di3 / R(6) + di4 / R(30)) * ei;
            dmpik[4] = R(1) - (R(1) + di + R(0.5) * di2 + di3 / R(6) + R(4) * di4 / R(105) + di5 / R(210)) * ei;
            dmpik[5] = R(1) - (R(1) + di + R(0.5) * di2 + di3 / R(6) + R(5) * di4 / R(126) + R(2) * di5 / R(315)
                               + di6 / R(1890)) * ei;
        } else {
            const R ai2 = ai * ai, ak2 = ak * ak;
            const R ti = ak2 / (ak2 - ai2), tk = ai2 / (ai2 - ak2);
            dmpik[0] = R(1) - ti * ei - tk * ek;
            dmpik[1] = R(1) - ti * (R(1) + di) * ei - tk * (R(1) + dk) * ek;
            dmpik[2] = R(1) - ti * (R(1) + di + di2 / R(3)) * ei - tk * (R(1) + dk + dk2 / R(3)) * ek;
            dmpik[3] = R(1) - ti * (R(1) + di + R(0.4) * di2 + di3 / R(15)) * ei
                     - tk * (R(1) + dk + R(0.4) * dk2 + dk3 / R(15)) * ek;
            dmpik[4] = R(1) - ti * (R(1) + di + R(3) * di2 / R(7) + R(2) * di3 / R(21) + di4 / R(105)) * ei
                     - tk * (R(1) + dk + R(3) * dk2 / R(7) + R(2) * dk3 / R(21) + dk4 / R(105)) * ek;
            dmpik[5] = R(1) - ti * (R(1) + di + R(4) * di2 / R(9) + di3 / R(9) + di4 / R(63) + di5 / R(945)) * ei
                     - tk * (R(1) + dk + R(4) * dk2 / R(9) + dk3 / R(9) + dk4 / R(63) + dk5 / R(945)) * ek;
        }
    }
}

// Complements 1 - dmp of the Gordon damping factors (damping.h:32-196),
// formed directly (no 1 - (1 - x) cancellation): xi/xk the one-site factors
// of i and k (orders 1..9), xik the overlap factors (orders 1..11, in FP64
// whatever R is: the unequal-exponent form cancels terms up to ~1e5 times its
// result when the exponents are close).
template <class R, int PEN>
__device__ __forceinline__ void damp_gordon_comp(R r, R ai, R ak, R *xi, R *xk, R *xik)
{
    const bool same = fabs(ai - ak) < R(1e-3);
    auto onesite = [](R d, R *x) {
        const R e = exp(-d);
        const R d2 = d * d, d3 = d2 * d, d4 = d2 * d2, d5 = d4 * d;
        if (PEN == 1) {
            x[0] = (R(1) + R(0.5) * d) * e;
            x[1] = (R(1) + d + R(0.5) * d2) * e;
            x[2] = (R(1) + d + R(0.5) * d2 + d3 / R(6)) * e;
            x[3] = (R(1) + d + R(0.5) * d2 + d3 / R(6) + d4 / R(30)) * e;
            x[4] = (R(1) + d + R(0.5) * d2 + d3 / R(6) + R(4) * d4 / R(105) + d5 / R(210)) * e;
        } else {
            x[0] = e;
            x[1] = (R(1) + d) * e;
            x[2] = (R(1) + d + d2 / R(3)) * e;
            x[3] = (R(1) + d + R(0.4) * d2 + d3 / R(15)) * e;
            x[4] = R(1); // no order-9 factor in this model (dmp = 0)
        }
    };
    onesite(ai * r, xi);
    if (same) {
        for (int j = 0; j < 5; j++) xk[j] = xi[j];
    } else {
        onesite(ak * r, xk);
    }
    const double rd = (double)r, aid = (double)ai, akd = (double)ak;
    const double di = aid * rd, dk = akd * rd;
    double d6[6];
    if (PEN == 1) {
        overlap_gordon1<double>(aid, akd, di, dk, exp(-di), exp(-dk), same, d6);
        for (int j = 0; j < 6; j++) xik[j] = R(1.0 - d6[j]);
    } else {
        // symmetric: the smaller exponent first so both orders round alike
        const bool sw = aid > akd;
        const double a1 = sw ? akd : aid, a2 = sw ? aid : akd, p = a1 * rd, q = a2 * rd;
        const double ep = exp(-p), eq = exp(-q);
        const double p2 = p * p, p3 = p2 * p, p4 = p2 * p2, p5 = p4 * p, q2 = q * q, q3 = q2 * q, q4 = q2 * q2,
                     q5 = q4 * q;
        if (same) {
            xik[0] = R((1.0 + 0.5 * p) * ep);
            xik[1] = R((1.0 + p + 0.5 * p2) * ep);
            xik[2] = R((1.0 + p + 0.5 * p2 + p3 / 6.0) * ep);
            xik[3] = R((1.0 + p + 0.5 * p2 + p3 / 6.0 + p4 / 30.0) * ep);
            xik[4] = R((1.0 + p + 0.5 * p2 + p3 / 6.0 + 4.0 * p4 / 105.0 + p5 / 210.0) * ep);
            xik[5] = R((1.0 + p + 0.5 * p2 + p3 / 6.0 + 5.0 * p4 / 126.0 + 2.0 * p5 / 315.0 + p3 * p3 / 1890.0) * ep);
        } else {
            const double t1 = a2 * a2 / (a2 * a2 - a1 * a1), t2 = a1 * a1 / (a1 * a1 - a2 * a2);
            xik[0] = R(t1 * ep + t2 * eq);
            xik[1] = R(t1 * (1.0 + p) * ep + t2 * (1.0 + q) * eq);
            xik[2] = R(t1 * (1.0 + p + p2 / 3.0) * ep + t2 * (1.0 + q + q2 / 3.0) * eq);
            xik[3] = R(t1 * (1.0 + p + 0.4 * p2 + p3 / 15.0) * ep + t2 * (1.0 + q + 0.4 * q2 + q3 / 15.0) * eq);
            xik[4] = R(t1 * (1.0 + p + 3.0 * p2 / 7.0 + 2.0 * p3 / 21.0 + p4 / 105.0) * ep
                       + t2 * (1.0 + q + 3.0 * q2 / 7.0 + 2.0 * q3 / 21.0 + q4 / 105.0) * eq);
            xik[5] = R(t1 * (1.0 + p + 4.0 * p2 / 9.0 + p3 / 9.0 + p4 / 63.0 + p5 / 945.0) * ep
                       + t2 * (1.0 + q + 4.0 * q2 / 9.0 + q3 / 9.0 + q4 / 63.0 + q5 / 945.0) * eq);
        }
    }
}

// Charge-penetration pair (pairMpole.h:186-389): a.c/b.c hold the valence
// charges, qi/qk the total charges.  The model's four parts (core-core
// undamped, core_k with the valence multipoles of i damped by dmpi, core_i
// with those of k by dmpk, valence-valence by dmpik) are summed as
//   [total multipoles, undamped] - core_k x val_i (1-dmpi) - core_i x val_k
//   (1-dmpk) - val_i x val_k (1-dmpik)
// which is the same sum without the large core/valence cancellation (cores
// of 6 and valence charges of -6.4 for water O), so it stays accurate in
// FP32.  Each series is scale_f x (factor) x rr; with Ewald (EW) the
// undamped part uses B instead of rr: the real-space part of a damped
// interaction whose reciprocal part is the undamped PME sum.
template <class R, bool GRAD, int PEN, bool EW = false>
__device__ __forceinline__ void pair_cp(R r2, R xr, R yr, R zr, R scale_f, R corei, R qi, R alphai, const Mp<R> &a,
                                        R corek, R qk, R alphak, const Mp<R> &b, PairOut<R> &o,
                                        const R *bn = nullptr)
{
    R rr[6], xi[5], xk[5], xik[6], Rs[6];
    rr_series<R>(r2, R(1), rr);
    damp_gordon_comp<R, PEN>(sqrt(r2), alphai, alphak, xi, xk, xik);
    Mp<R> ta = a, tb = b, ci{}, ck{};
    ta.c = qi;
    tb.c = qk;
    ci.c = corei;
    ck.c = corek;
    for (int j = 0; j < 6; j++) Rs[j] = scale_f * (EW ? bn[j] : rr[j]);
    pair_series<R, GRAD, false>(Rs, ta, tb, xr, yr, zr, o);  // phii/phik: derivatives in the valence charges
    for (int j = 0; j < 6; j++) Rs[j] = -scale_f * xik[j] * rr[j];
    pair_series<R, GRAD, true>(Rs, a, b, xr, yr, zr, o);
    PairOut<R> t;
    for (int j = 0; j < 6; j++) Rs[j] = j < 5 ? -scale_f * xi[j] * rr[j] : R(0);
    pair_series<R, GRAD, false>(Rs, a, ck, xr, yr, zr, t);
    o.e += t.e;
    o.phii += t.phii;
    if (GRAD) {
        o.fx += t.fx; o.fy += t.fy; o.fz += t.fz;
        o.tix += t.tix; o.tiy += t.tiy; o.tiz += t.tiz;
    }
    for (int j = 0; j < 6; j++) Rs[j] = j < 5 ? -scale_f * xk[j] * rr[j] : R(0);
    pair_series<R, GRAD, false>(Rs, ci, b, xr, yr, zr, t);
    o.e += t.e;
    o.phik += t.phik;
    if (GRAD) {
        o.fx += t.fx; o.fy += t.fy; o.fz += t.fz;
        o.tkx += t.tkx; o.tky += t.tky; o.tkz += t.tkz;
    }
}

// ---------------------------------------------------------------------------
// Quasi-internal frame of a pair (the formulation of the reference's
// pairMpoleQI.h:22-278 and qi.h:11-64, which it includes but never calls):
// rows ex, ey, ez of an orthonormal frame with ez along r; any ex normal to
// r serves.  In that frame r = (0, 0, |r|), so every pair-algebra term that
// carries the two normal components of r vanishes; the caller rotates both
// multipoles in (qi_rotate), evaluates the ordinary pair algebra at
// (0, 0, |r|) with literal zeros (the kernels are built without signed
// zeros or non-finite values, so the compiler drops those terms), and
// rotates the force and torques back (qi_back): they are vectors, so the
// result equals the Cartesian form's up to rounding.
template <class R>
struct QiFrame {
    R m[3][3]; // rows ex, ey, ez
    R len;     // |r|
};

template <class R>
__device__ __forceinline__ QiFrame<R> qi_frame(R xr, R yr, R zr, R r2)
{
    QiFrame<R> f;
    const R rinv = R(1) / sqrt(r2);
    f.len = r2 * rinv;
    const R z[3] = {xr * rinv, yr * rinv, zr * rinv};
    // ex, ey completing ez to a right-handed orthonormal basis with one
    // division and no square root (Duff, Burgess, Christensen, Hery, Kensler,
    // Liani and Villemin, "Building an orthonormal basis, revisited",
    // J. Computer Graphics Techniques 6(1), 2017)
    const R sg = z[2] >= R(0) ? R(1) : R(-1);
    const R a = R(-1) / (sg + z[2]);
    const R b = z[0] * z[1] * a;
    const R x[3] = {R(1) + sg * z[0] * z[0] * a, sg * b, -sg * z[0]};
    const R y[3] = {b, sg + z[1] * z[1] * a, -z[1]};
#pragma unroll
    for (int c = 0; c < 3; c++) { f.m[0][c] = x[c]; f.m[1][c] = y[c]; f.m[2][c] = z[c]; }
    return f;
}

// a multipole in the frame: d' = M d, Q' = M Q M^T (Q' traceless by
// construction: its zz is minus the sum of the other two diagonal entries)
template <class R>
__device__ __forceinline__ Mp<R> qi_rotate(const QiFrame<R> &f, const Mp<R> &a)
{
    Mp<R> o;
    o.c = a.c;
    const R (*M)[3] = f.m;
    o.dx = M[0][0] * a.dx + M[0][1] * a.dy + M[0][2] * a.dz;
    o.dy = M[1][0] * a.dx + M[1][1] * a.dy + M[1][2] * a.dz;
    o.dz = M[2][0] * a.dx + M[2][1] * a.dy + M[2][2] * a.dz;
    // rows of M Q: (Q m_p) for p = x, y, z
    R q[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        q[p][0] = a.qxx * M[p][0] + a.qxy * M[p][1] + a.qxz * M[p][2];
        q[p][1] = a.qxy * M[p][0] + a.qyy * M[p][1] + a.qyz * M[p][2];
        q[p][2] = a.qxz * M[p][0] + a.qyz * M[p][1] + a.qzz * M[p][2];
    }
    o.qxx = q[0][0] * M[0][0] + q[0][1] * M[0][1] + q[0][2] * M[0][2];
    o.qxy = q[0][0] * M[1][0] + q[0][1] * M[1][1] + q[0][2] * M[1][2];
    o.qxz = q[0][0] * M[2][0] + q[0][1] * M[2][1] + q[0][2] * M[2][2];
    o.qyy = q[1][0] * M[1][0] + q[1][1] * M[1][1] + q[1][2] * M[1][2];
    o.qyz = q[1][0] * M[2][0] + q[1][1] * M[2][1] + q[1][2] * M[2][2];
    o.qzz = -(o.qxx + o.qyy);
    return o;
}

// force and torques back to the global frame (v = M^T v')
template <class R>
__device__ __forceinline__ void qi_back(const QiFrame<R> &f, PairOut<R> &o)
{
    const R (*M)[3] = f.m;
    auto back = [&](R &x, R &y, R &z) {
        const R a = x, b = y, c = z;
        x = M[0][0] * a + M[1][0] * b + M[2][0] * c;
        y = M[0][1] * a + M[1][1] * b + M[2][1] * c;
        z = M[0][2] * a + M[1][2] * b + M[2][2] * c;
    };
    back(o.fx, o.fy, o.fz);
    back(o.tix, o.tiy, o.tiz);
    back(o.tkx, o.tky, o.tkz);
}

} // namespace pmq

namespace pmq {

// Ewald-screened series B_0..B_5 (B_0 = erfc(a r)/r and
// B_n = ((2n-1) B_{n-1} + (2a^2)^n e^{-a^2 r^2} / (a sqrt(pi))) / r^2):
// the real-space part of each rr_n
template <class R>
__device__ __forceinline__ void ewald_series(R r2, R aewald, R *bn)
{
    const R r = sqrt(r2);
    const R ra = aewald * r;
    const R r2i = R(1) / r2;
    bn[0] = erfc(ra) / r;
    const R ex = exp(-ra * ra);
    const R a2 = R(2) * aewald * aewald;
    R an = R(1) / (R(1.77245385090551602730) * aewald);
#pragma unroll
    for (int j = 1; j < 6; j++) {
        an *= a2;
        bn[j] = (R(2 * j - 1) * bn[j - 1] + an * ex) * r2i;
    }
}

// Thole damping factors of the polarization interactions (lambda_3..9 in
// lam[1..4]); D = a (r / (pd_i pd_k))^3, the factor set that keeps
// d(lambda_n rr_n)/dr = -r lambda_{n+2} rr_{n+2}
template <class R>
__device__ __forceinline__ void thole_lambda(R r, R pdi, R pdk, R ti, R tk, R *lam)
{
    lam[0] = lam[1] = lam[2] = lam[3] = lam[4] = lam[5] = R(1);
    const R damp = pdi * pdk;
    if (damp == R(0)) return;
    const R mn = fmin(ti, tk);
    const R pg = mn != R(0) ? mn : fmax(ti, tk);
    const R u = r / damp;
    const R D = pg * u * u * u;
    if (D >= R(50)) return;
    const R e = exp(-D);
    lam[1] = R(1) - e;
    lam[2] = R(1) - (R(1) + D) * e;
    lam[3] = R(1) - (R(1) + D + R(0.6) * D * D) * e;
    lam[4] = R(1) - (R(1) + D + R(18.0 / 35.0) * D * D + R(9.0 / 35.0) * D * D * D) * e;
}

// series of a polarization interaction at scale 1: Ewald + Thole, or Thole
template <bool EW, class R_ = double>
__device__ __forceinline__ void pol_series(R_ r2, R_ aewald, const R_ *lam, R_ *R)
{
    R_ rr[6];
    rr_series<R_>(r2, R_(1), rr);
    if (EW) {
        R_ bn[6];
        ewald_series<R_>(r2, aewald, bn);
        for (int q = 0; q < 6; q++) R[q] = bn[q] - (R_(1) - lam[q]) * rr[q];
    } else {
        for (int q = 0; q < 6; q++) R[q] = lam[q] * rr[q];
    }
}

// correction series (s - 1) lambda rr for a scaled pair
template <class R_ = double>
__device__ __forceinline__ void corr_series(R_ r2, R_ s, const R_ *lam, R_ *R)
{
    R_ rr[6];
    rr_series<R_>(r2, R_(1), rr);
    for (int q = 0; q < 6; q++) R[q] = (s - R_(1)) * lam[q] * rr[q];
}

} // namespace pmq

namespace pmq {

// the overlap (valence-valence) factors of two exponential densities at r,
// evaluated in FP64 whatever R is (see damp_gordon)
template <class R>
__device__ __forceinline__ void overlap_damp(R r, R ai, R ak, R *dmpik)
{
    const double rd = (double)r, aid = (double)ai, akd = (double)ak;
    double d6[6];
    overlap_gordon1<double>(aid, akd, aid * rd, akd * rd, exp(-aid * rd), exp(-akd * rd), fabs(aid - akd) < 1e-3, d6);
    for (int j = 0; j < 6; j++) dmpik[j] = R(d6[j]);
}

// one-site factors of one exponential density (Gordon1), orders 1..9: the
// field of a valence multipole of exponent a seen by a point at distance r
template <class R>
__device__ __forceinline__ void onesite_damp(R r, R a, R *dmp)
{
    const R d = a * r, e = exp(-d);
    const R d2 = d * d, d3 = d2 * d, d4 = d2 * d2, d5 = d4 * d;
    dmp[0] = R(1) - (R(1) + R(0.5) * d) * e;
    dmp[1] = R(1) - (R(1) + d + R(0.5) * d2) * e;
    dmp[2] = R(1) - (R(1) + d + R(0.5) * d2 + d3 / R(6)) * e;
    dmp[3] = R(1) - (R(1) + d + R(0.5) * d2 + d3 / R(6) + d4 / R(30)) * e;
    dmp[4] = R(1) - (R(1) + d + R(0.5) * d2 + d3 / R(6) + R(4) * d4 / R(105) + d5 / R(210)) * e;
}

// HIPPO repulsion radial series B_n = D^n (S^2), n = 0..5, D = -(1/r) d/dr,
// S the overlap of two normalized 1s Slater functions of exponents ai/2 and
// ak/2 (Rackers and Ponder, JCP 150, 084104 (2019)).  D obeys the product
// rule, so B_n = sum_k C(n,k) S_k S_{n-k} with S_k = D^k S.
//   equal exponents z (within 1e-3 in a; their mean is used):
//     S = (1 + z r + z^2 r^2/3) e^{-zr}, S_1 = z^2/3 (1 + z r) e^{-zr},
//     S_2 = z^4/3 e^{-zr}, S_3 = z^5/3 x e^{-zr}, S_4 = z^5/3 (x^3 + z x^2) e^{-zr},
//     S_5 = z^5/3 (3 x^5 + 3 z x^4 + z^2 x^3) e^{-zr}   (x = 1/r)
//   unequal a, b: S = A [(a - t x) e^{-br} + (b + t x) e^{-ar}],
//     t = 4ab/(a^2-b^2), A = 8 (ab)^{3/2}/(a^2-b^2)^2, and D maps the
//     coefficient q_m of x^m e^{-cr} to m q_m x^{m+2} + c q_m x^{m+1}.
template <class R>
__device__ __forceinline__ void rep_series(R r, R ai, R ak, R *B)
{
    R S[6];
    const R x = R(1) / r;
    if (fabs(ai - ak) < R(1e-3)) {
        const R z = R(0.25) * (ai + ak);
        const R e = exp(-z * r), z2 = z * z, z5 = z2 * z2 * z;
        const R x2 = x * x, x3 = x2 * x;
        S[0] = (R(1) + z * r + z2 * r * r / R(3)) * e;
        S[1] = z2 / R(3) * (R(1) + z * r) * e;
        S[2] = z2 * z2 / R(3) * e;
        S[3] = z5 / R(3) * x * e;
        S[4] = z5 / R(3) * (x3 + z * x2) * e;
        S[5] = z5 / R(3) * (R(3) * x3 * x2 + R(3) * z * x2 * x2 + z2 * x3) * e;
    } else {
        const R a = R(0.5) * ai, b = R(0.5) * ak;
        const R d = a * a - b * b, t = R(4) * a * b / d;
        const R ab = a * b;
        const R A = R(8) * ab * sqrt(ab) / (d * d);
#pragma unroll
        for (int n = 0; n < 6; n++) S[n] = R(0);
#pragma unroll
        for (int term = 0; term < 2; term++) {
            const R c = term == 0 ? b : a;
            R q[12];
#pragma unroll
            for (int m = 0; m < 12; m++) q[m] = R(0);
            q[0] = A * (term == 0 ? a : b);
            q[1] = A * (term == 0 ? -t : t);
            const R e = exp(-c * r);
#pragma unroll
            for (int n = 0; n < 6; n++) {
                // value of sum_m q_m x^m (powers up to 1 + 2n)
                R v = R(0);
#pragma unroll
                for (int m = 11; m >= 0; m--)
                    if (m <= 1 + 2 * n) v = v * x + q[m];
                S[n] += v * e;
                if (n == 5) break;
                R nq[12];
#pragma unroll
                for (int m = 0; m < 12; m++) nq[m] = R(0);
#pragma unroll
                for (int m = 0; m < 12; m++) {
                    if (m > 1 + 2 * n) continue;
                    if (m + 2 < 12) nq[m + 2] += R(m) * q[m];
                    if (m + 1 < 12) nq[m + 1] += c * q[m];
                }
#pragma unroll
                for (int m = 0; m < 12; m++) q[m] = nq[m];
            }
        }
    }
    B[0] = S[0] * S[0];
    B[1] = R(2) * S[0] * S[1];
    B[2] = R(2) * (S[0] * S[2] + S[1] * S[1]);
    B[3] = R(2) * (S[0] * S[3] + R(3) * S[1] * S[2]);
    B[4] = R(2) * (S[0] * S[4] + R(4) * S[1] * S[3]) + R(6) * S[2] * S[2];
    B[5] = R(2) * (S[0] * S[5] + R(5) * S[1] * S[4] + R(10) * S[2] * S[3]);
}

// quintic multiplicative switch T(r) = c0 + c1 r + ... + c5 r^5 between the
// taper distance and the cutoff (cutoffSwitch.cpp:125-133), and dT/dr
struct Taper {
    double cut, c[6];
    double w; // 1 / (off - cut)
};

// quintic switch coefficients between cut and off (cutoffSwitch.cpp:125-133);
// no switch when cut >= off (k_hippo.hip)
Taper make_taper(double cut, double off);

// the switch T(r) and dT/dr beyond cut.  FP64 sums the power series of
// cutoffSwitch.cpp; FP32 the same polynomial in u = (r - cut) / (off - cut),
// T = 1 - 10 u^3 + 15 u^4 - 6 u^5 (the one quintic with T = 1 at cut, 0 at
// off and zero first and second derivatives at both), whose terms stay near
// 1 where the power series' terms reach off^5 / (off - cut)^5 and cancel
template <class R>
__device__ __forceinline__ void taper_eval(const Taper &tp, R r, R &T, R &dT)
{
    T = R(1);
    dT = R(0);
    if (r <= R(tp.cut)) return;
    if constexpr (sizeof(R) == 4) {
        const R w = R(tp.w), u = (r - R(tp.cut)) * w, u2 = u * u;
        T = R(1) + u2 * u * (R(-10) + u * (R(15) - R(6) * u));
        dT = u2 * (R(-30) + u * (R(60) - R(30) * u)) * w;
    } else {
        const R c0 = R(tp.c[0]), c1 = R(tp.c[1]), c2 = R(tp.c[2]), c3 = R(tp.c[3]), c4 = R(tp.c[4]), c5 = R(tp.c[5]);
        T = c0 + r * (c1 + r * (c2 + r * (c3 + r * (c4 + r * c5))));
        dT = c1 + r * (R(2) * c2 + r * (R(3) * c3 + r * (R(4) * c4 + r * R(5) * c5)));
    }
}

} // namespace pmq
