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

// Charge-penetration pair (pairMpole.h:186-389): a.c/b.c hold the valence
// charges; the interaction is core-core (undamped) + core_k with the
// valence multipoles of i (dmpi) + core_i with those of k (dmpk) +
// valence-valence (dmpik).  Each class X uses the series sf*dmp_X*rr, or
// with Ewald (EW) f*(B - (1 - dmp_X) rr): the real-space part of a damped
// interaction whose reciprocal part is the undamped PME sum.
template <class R, bool GRAD, int PEN, bool EW = false>
__device__ __forceinline__ void pair_cp(R r2, R xr, R yr, R zr, R scale_f, R corei, R alphai, const Mp<R> &a,
                                        R corek, R alphak, const Mp<R> &b, PairOut<R> &o, const R *bn = nullptr)
{
    R rr[6], dmpi[5], dmpk[5], dmpik[6], Rs[6];
    rr_series<R>(r2, R(1), rr);
    damp_gordon<R, PEN>(sqrt(r2), alphai, alphak, dmpi, dmpk, dmpik);
    auto series = [&](const R *dmp, int nd) {
        for (int j = 0; j < 6; j++) {
            const R d = j < nd ? dmp[j] : R(0);
            Rs[j] = EW ? scale_f * (bn[j] - (R(1) - d) * rr[j]) : scale_f * d * rr[j];
        }
    };
    Mp<R> ci{}, ck{};
    ci.c = corei;
    ck.c = corek;
    const R one[6] = {R(1), R(1), R(1), R(1), R(1), R(1)};
    series(one, 6);
    pair_series<R, GRAD, false>(Rs, ci, ck, xr, yr, zr, o);
    // core-core has no potential on the valence charges
    o.phii = R(0);
    o.phik = R(0);
    series(dmpik, 6);
    pair_series<R, GRAD, true>(Rs, a, b, xr, yr, zr, o);
    {
        PairOut<R> t;
        series(dmpi, 5);
        pair_series<R, GRAD, false>(Rs, a, ck, xr, yr, zr, t);
        o.e += t.e;
        o.phii += t.phii;
        if (GRAD) {
            o.fx += t.fx; o.fy += t.fy; o.fz += t.fz;
            o.tix += t.tix; o.tiy += t.tiy; o.tiz += t.tiz;
        }
        series(dmpk, 5);
        pair_series<R, GRAD, false>(Rs, ci, b, xr, yr, zr, t);
        o.e += t.e;
        o.phik += t.phik;
        if (GRAD) {
            o.fx += t.fx; o.fy += t.fy; o.fz += t.fz;
            o.tkx += t.tkx; o.tky += t.tky; o.tkz += t.tkz;
        }
    }
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
