// param_pair.h -- one pair's contributions to the derivatives of the energy
// with respect to the parameters (multipole components, Thole damping),
// shared by the one-system tiles (k_params.hip) and the batched clusters
// (k_batch.hip).
//
// Every pair energy is bilinear in the two multipoles, so the derivative of a
// pair's energy with respect to component c of M_i is the pair energy with
// M_i replaced by the unit multipole e_c; the same holds for the direct fields
// E^d and E^p.  The variational polarization energy E_pol = -(f/2) mu_d.E^p
// with mu_p as the adjoint of mu_d gives
//     dE_pol = -(f/2) [ mu_p . dE^d + mu_d . dE^p + mu_p . dT mu_d ]
// (the 1/alpha part is per site).  The Thole factors depend on
// D = a_ik (r / (pd_i pd_k))^3; dE/dD of the pair comes from the derivative
// series dl_n rr_n in the direct fields and in T.
#pragma once

#include "pair_mpole.h"

namespace pmq {

template <class T>
__device__ __forceinline__ Mp<T> unit_mp(int c)
{
    Mp<T> m{};
    mp_set(m, c, T(1));
    return m;
}

// dlambda_n / dD in dl[1..4] (dl[0] = dl[5] = 0); zero where the factors are 1.
// Returns D and the pair's Thole parameter in *Dout, *pgout (0 when undamped).
template <class T>
__device__ __forceinline__ void thole_dlambda(T r, T pdi, T pdk, T ti, T tk, T *dl, T *Dout, T *pgout)
{
    for (int q = 0; q < 6; q++) dl[q] = T(0);
    *Dout = T(0);
    *pgout = T(0);
    const T damp = pdi * pdk;
    if (damp == T(0)) return;
    const T mn = fmin(ti, tk);
    const T pg = mn != T(0) ? mn : fmax(ti, tk);
    const T uu = r / damp;
    const T D = pg * uu * uu * uu;
    if (D >= T(50)) return;
    const T e = exp(-D);
    dl[1] = e;
    dl[2] = D * e;
    dl[3] = D * (T(3) * D - T(1)) / T(5) * e;
    dl[4] = D * (T(9) * D * D - T(9) * D - T(1)) / T(35) * e;
    *Dout = D;
    *pgout = pg;
}

// share of d a_ik / d a_i for a_ik = min(a_i, a_k) (max when the min is 0)
template <class T>
__device__ __forceinline__ T thole_share(T ti, T tk)
{
    const T mn = fmin(ti, tk);
    if (mn != T(0)) return ti < tk ? T(1) : (ti > tk ? T(0) : T(0.5));
    return ti > tk ? T(1) : (ti < tk ? T(0) : T(0.5));
}

// T v at i from dipole v at k with series R (the dipole-field operator)
template <class T>
__device__ __forceinline__ void tdip(const T *R, const T *v, T xr, T yr, T zr, T *o)
{
    const T vr = v[0] * xr + v[1] * yr + v[2] * zr;
    o[0] = vr * R[2] * xr - R[1] * v[0];
    o[1] = vr * R[2] * yr - R[1] * v[1];
    o[2] = vr * R[2] * zr - R[1] * v[2];
}

// One pair's contributions, scaled: sm (multipole energy), sd / sp (direct
// fields E^d / E^p), su (T).  Rm: the multipole series (f included); R: the
// polarization series at scale 1; rr: the undamped series; dl: dlambda/dD.
// gi (and, with BOTH, gk) += dE/dM of the two sites; *dEdD += dE/dD.
template <class T, bool BOTH>
__device__ void param_pair(bool do_mpole, bool do_polar, double f, T xr, T yr, T zr, const Mp<T> &mi,
                           const Mp<T> &mk, const T *Rm, const T *R, const T *rr, const T *dl, T sm, T sd, T sp, T su,
                           const T *udi, const T *upi, const T *udk, const T *upk, double *gi, double *gk,
                           double *dEdD)
{
    if (do_mpole && sm != T(0)) {
        T Rs[6];
        for (int q = 0; q < 6; q++) Rs[q] = sm * Rm[q];
#pragma unroll
        for (int c = 0; c < 10; c++) {
            const Mp<T> e = unit_mp<T>(c);
            PairOut<T> o;
            pair_series<T, false>(Rs, e, mk, xr, yr, zr, o);
            gi[c] += (double)o.e;
            if (BOTH) {
                pair_series<T, false>(Rs, mi, e, xr, yr, zr, o);
                gk[c] += (double)o.e;
            }
        }
    }
    if (!do_polar) return;
    const double hf = 0.5 * f;
    // direct fields: the field at k of e_c at i weighted by (sd mu_p,k + sp mu_d,k) for
    // site i, and the field at i of e_c at k for site k
    if (sd != T(0) || sp != T(0)) {
        const T wi[3] = {sd * upi[0] + sp * udi[0], sd * upi[1] + sp * udi[1], sd * upi[2] + sp * udi[2]};
        const T wk[3] = {sd * upk[0] + sp * udk[0], sd * upk[1] + sp * udk[1], sd * upk[2] + sp * udk[2]};
#pragma unroll
        for (int c = 0; c < 10; c++) {
            const Mp<T> e = unit_mp<T>(c);
            const T qx = e.qxx * xr + e.qxy * yr + e.qxz * zr, qy = e.qxy * xr + e.qyy * yr + e.qyz * zr,
                    qz = e.qxz * xr + e.qyz * yr + e.qzz * zr;
            const T dr = e.dx * xr + e.dy * yr + e.dz * zr;
            const T qr = qx * xr + qy * yr + qz * zr;
            // field at k from e_c at i
            const T sk = e.c * R[1] + dr * R[2] + qr * R[3];
            const T Ek[3] = {sk * xr - R[1] * e.dx - T(2) * R[2] * qx, sk * yr - R[1] * e.dy - T(2) * R[2] * qy,
                             sk * zr - R[1] * e.dz - T(2) * R[2] * qz};
            gi[c] -= hf * (double)(wk[0] * Ek[0] + wk[1] * Ek[1] + wk[2] * Ek[2]);
            if (BOTH) {
                // field at i from e_c at k
                const T si = e.c * R[1] - dr * R[2] + qr * R[3];
                const T Ei[3] = {-si * xr - R[1] * e.dx + T(2) * R[2] * qx, -si * yr - R[1] * e.dy + T(2) * R[2] * qy,
                                 -si * zr - R[1] * e.dz + T(2) * R[2] * qz};
                gk[c] -= hf * (double)(wi[0] * Ei[0] + wi[1] * Ei[1] + wi[2] * Ei[2]);
            }
        }
    }
    // damping: dE/dD of this pair
    T dR[6];
    for (int q = 0; q < 6; q++) dR[q] = dl[q] * rr[q];
    if (dR[1] == T(0) && dR[2] == T(0) && dR[3] == T(0)) return;
    T acc = T(0);
    if (sd != T(0) || sp != T(0)) {
        // fields of M_k at i and of M_i at k with the derivative series
        const T qix = mi.qxx * xr + mi.qxy * yr + mi.qxz * zr, qiy = mi.qxy * xr + mi.qyy * yr + mi.qyz * zr,
                qiz = mi.qxz * xr + mi.qyz * yr + mi.qzz * zr;
        const T qkx = mk.qxx * xr + mk.qxy * yr + mk.qxz * zr, qky = mk.qxy * xr + mk.qyy * yr + mk.qyz * zr,
                qkz = mk.qxz * xr + mk.qyz * yr + mk.qzz * zr;
        const T dir = mi.dx * xr + mi.dy * yr + mi.dz * zr, dkr = mk.dx * xr + mk.dy * yr + mk.dz * zr;
        const T qir = qix * xr + qiy * yr + qiz * zr, qkr = qkx * xr + qky * yr + qkz * zr;
        const T si = mk.c * dR[1] - dkr * dR[2] + qkr * dR[3];
        const T sk = mi.c * dR[1] + dir * dR[2] + qir * dR[3];
        const T Ei[3] = {-si * xr - dR[1] * mk.dx + T(2) * dR[2] * qkx, -si * yr - dR[1] * mk.dy + T(2) * dR[2] * qky,
                         -si * zr - dR[1] * mk.dz + T(2) * dR[2] * qkz};
        const T Ek[3] = {sk * xr - dR[1] * mi.dx - T(2) * dR[2] * qix, sk * yr - dR[1] * mi.dy - T(2) * dR[2] * qiy,
                         sk * zr - dR[1] * mi.dz - T(2) * dR[2] * qiz};
        for (int c = 0; c < 3; c++)
            acc += (sd * upi[c] + sp * udi[c]) * Ei[c] + (sd * upk[c] + sp * udk[c]) * Ek[c];
    }
    if (su != T(0)) {
        T t1[3], t2[3];
        tdip<T>(dR, udk, xr, yr, zr, t1); // at i from mu_d,k
        tdip<T>(dR, udi, xr, yr, zr, t2); // at k from mu_d,i
        for (int c = 0; c < 3; c++) acc += su * (upi[c] * t1[c] + upk[c] * t2[c]);
    }
    *dEdD += -hf * (double)acc;
}

} // namespace pmq
