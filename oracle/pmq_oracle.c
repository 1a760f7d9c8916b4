/*
 * pmq_oracle.c -- FP64 CPU restatement of the reference's multipole energy.
 *
 * TEST INFRASTRUCTURE ONLY (see pmq_oracle.h).  Plain C, single thread,
 * written from the reference's algorithm, not from its code: every routine
 * names the reference lines whose arithmetic it restates.
 *
 * Multipole layout used throughout: m[10] = {c, dx, dy, dz, qxx, qxy, qxz,
 * qyy, qyz, qzz}, taken from the 13-component reference rows at indices
 * 0,1,2,3,4,5,6,8,9,12 (empole.cpp:206-215).
 */
#include "pmq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RADIAN 57.29577951308232088 /* mathConst.h:28 */

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
/* y = Q v for the symmetric quadrupole stored as {qxx,qxy,qxz,qyy,qyz,qzz} */
static inline void qmul(const double *q, const double *v, double *y)
{
    y[0] = q[0] * v[0] + q[1] * v[1] + q[2] * v[2];
    y[1] = q[1] * v[0] + q[3] * v[1] + q[4] * v[2];
    y[2] = q[2] * v[0] + q[4] * v[1] + q[5] * v[2];
}

static void pick10(const double *row13, double *m)
{
    static const int idx[10] = {0, 1, 2, 3, 4, 5, 6, 8, 9, 12};
    for (int j = 0; j < 10; j++) m[j] = row13[idx[j]];
}

/* ------------------------------------------------------------------------
 * Generic multipole pair with an arbitrary radial series.
 *
 * R[0..5] multiply the five energy "terms" (and the gradient uses the next
 * member of the series), exactly as rr1..rr11 do in pairMpole.h:36-173;
 * any series with dR_n/dr = -r R_{n+2} (undamped, Ewald-erfc, Thole- or
 * Gordon-damped) gives the energy, its gradient and the torques.
 *
 * Sign conventions follow pairMpole.h: r = x_k - x_i, frc = dE/dx_i,
 * trq = torque on the site's multipole.  phi_i = dE/dc_i and
 * phi_k = dE/dc_k are the potentials the charge-flux chain rule needs
 * (pairMpole.h:304-313).
 * ---------------------------------------------------------------------- */
typedef struct {
    double e;
    double frc[3];
    double trqi[3], trqk[3];
    double phi_i, phi_k;
} pair_out;

static void pair_series(const double *R, const double *mi, const double *mk, const double *r,
                        int do_g, pair_out *o)
{
    const double ci = mi[0], ck = mk[0];
    const double *di = mi + 1, *dk = mk + 1, *qi = mi + 4, *qk = mk + 4;
    double qir_v[3], qkr_v[3];
    qmul(qi, r, qir_v);
    qmul(qk, r, qkr_v);
    const double dir = dot3(di, r), dkr = dot3(dk, r);
    const double qir = dot3(qir_v, r), qkr = dot3(qkr_v, r);
    const double dik = dot3(di, dk), qik = dot3(qir_v, qkr_v);
    const double diqk = dot3(di, qkr_v), dkqi = dot3(dk, qir_v);
    const double qiqk = 2.0 * (qi[1] * qk[1] + qi[2] * qk[2] + qi[4] * qk[4])
                      + qi[0] * qk[0] + qi[3] * qk[3] + qi[5] * qk[5];

    const double t1 = ci * ck;
    const double t2 = ck * dir - ci * dkr + dik;
    const double t3 = ci * qkr + ck * qir - dir * dkr + 2.0 * (dkqi - diqk + qiqk);
    const double t4 = dir * qkr - dkr * qir - 4.0 * qik;
    const double t5 = qir * qkr;
    o->e = t1 * R[0] + t2 * R[1] + t3 * R[2] + t4 * R[3] + t5 * R[4];
    o->phi_i = ck * R[0] - dkr * R[1] + qkr * R[2];
    o->phi_k = ci * R[0] + dir * R[1] + qir * R[2];
    if (!do_g) return;

    const double de = t1 * R[1] + t2 * R[2] + t3 * R[3] + t4 * R[4] + t5 * R[5];
    const double g1 = -ck * R[1] + dkr * R[2] - qkr * R[3];   /* multiplies d_i */
    const double g2 = ci * R[1] + dir * R[2] + qir * R[3];    /* multiplies d_k */
    const double g3 = 2.0 * R[2];
    const double g4 = 2.0 * (-ck * R[2] + dkr * R[3] - qkr * R[4]);
    const double g5 = 2.0 * (-ci * R[2] - dir * R[3] - qir * R[4]);
    const double g6 = 4.0 * R[3];

    double qk_di[3], qi_dk[3], qi_qkr[3], qk_qir[3];
    qmul(qk, di, qk_di);
    qmul(qi, dk, qi_dk);
    qmul(qi, qkr_v, qi_qkr);
    qmul(qk, qir_v, qk_qir);
    for (int a = 0; a < 3; a++)
        o->frc[a] = de * r[a] + g1 * di[a] + g2 * dk[a] + g3 * (qk_di[a] - qi_dk[a])
                  + g4 * qir_v[a] + g5 * qkr_v[a] + g6 * (qi_qkr[a] + qk_qir[a]);

    /* torque pieces, pairMpole.h:110-164 */
    double di_x_dk[3], di_x_r[3], dk_x_r[3], r_x_qir[3], r_x_qkr[3], qkr_x_qir[3];
    double r_x_qiqkr[3], r_x_qkqir[3], r_x_qkdi[3], r_x_qidk[3], di_x_qkr[3], dk_x_qir[3];
    cross3(di, dk, di_x_dk);
    cross3(di, r, di_x_r);
    cross3(dk, r, dk_x_r);
    cross3(r, qir_v, r_x_qir);
    cross3(r, qkr_v, r_x_qkr);
    cross3(qkr_v, qir_v, qkr_x_qir);
    cross3(r, qi_qkr, r_x_qiqkr);
    cross3(r, qk_qir, r_x_qkqir);
    cross3(r, qk_di, r_x_qkdi);
    cross3(r, qi_dk, r_x_qidk);
    cross3(di, qkr_v, di_x_qkr);
    cross3(dk, qir_v, dk_x_qir);
    /* (Qi Qk) antisymmetric part contracted with epsilon: x = (QiQk)_yz - (QiQk)_zy, ... */
    double A[3][3];
    const double Qi[3][3] = {{qi[0], qi[1], qi[2]}, {qi[1], qi[3], qi[4]}, {qi[2], qi[4], qi[5]}};
    const double Qk[3][3] = {{qk[0], qk[1], qk[2]}, {qk[1], qk[3], qk[4]}, {qk[2], qk[4], qk[5]}};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) A[a][b] = Qi[a][0] * Qk[0][b] + Qi[a][1] * Qk[1][b] + Qi[a][2] * Qk[2][b];
    const double epsA[3] = {A[1][2] - A[2][1], A[2][0] - A[0][2], A[0][1] - A[1][0]};
    for (int a = 0; a < 3; a++) {
        const double dqik = di_x_qkr[a] + dk_x_qir[a] - 2.0 * epsA[a];
        o->trqi[a] = -R[1] * di_x_dk[a] + g1 * di_x_r[a] + g3 * (dqik + r_x_qidk[a])
                   - g4 * r_x_qir[a] - g6 * (r_x_qiqkr[a] + qkr_x_qir[a]);
        o->trqk[a] = R[1] * di_x_dk[a] + g2 * dk_x_r[a] - g3 * (dqik + r_x_qkdi[a])
                   - g5 * r_x_qkr[a] - g6 * (r_x_qkqir[a] - qkr_x_qir[a]);
    }
}

/* undamped series rr1..rr11 with the exclusion scale and unit factor folded
 * into rr1, pairMpole.h:36-46 */
static void rr_series(double r2, double mscale, double f, double *R)
{
    const double r = sqrt(r2);
    const double rinv2 = 1.0 / r2;
    R[0] = mscale * f / r;
    R[1] = R[0] * rinv2;
    R[2] = 3.0 * R[1] * rinv2;
    R[3] = 5.0 * R[2] * rinv2;
    R[4] = 7.0 * R[3] * rinv2;
    R[5] = 9.0 * R[4] * rinv2;
}

/* ------------------------------------------------------------------------
 * Gordon charge-penetration damping, damping.h:32-196 (Slipchenko & Gordon
 * 2007; Rackers et al. 2017).  dmpi/dmpk damp core(k)-valence(i) and
 * core(i)-valence(k), dmpik valence-valence; all 6 orders are filled.
 * ---------------------------------------------------------------------- */
static void damp_gordon1(double r, double ai, double ak, double *dmpi, double *dmpk, double *dmpik)
{
    const double di = ai * r, dk = ak * r;
    const double ei = exp(-di), ek = exp(-dk);
    const int same = fabs(ai - ak) < 1e-3;
    double p[9], q[9];
    p[0] = 1.0; q[0] = 1.0;
    for (int j = 1; j < 9; j++) { p[j] = p[j - 1] * di; q[j] = q[j - 1] * dk; }
    /* core-valence, one exponential */
#define CV1(P, E, out)                                                                          \
    out[0] = 1.0 - (1.0 + 0.5 * P[1]) * E;                                                      \
    out[1] = 1.0 - (1.0 + P[1] + 0.5 * P[2]) * E;                                               \
    out[2] = 1.0 - (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0) * E;                                  \
    out[3] = 1.0 - (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + P[4] / 30.0) * E;                    \
    out[4] = 1.0 - (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + 4.0 * P[4] / 105.0 + P[5] / 210.0) * E;
    CV1(p, ei, dmpi);
    if (same) memcpy(dmpk, dmpi, 5 * sizeof(double));
    else { CV1(q, ek, dmpk); }
#undef CV1
    /* valence-valence factors: symmetric in the pair; evaluated with the
     * smaller exponent first so that both orders of a pair round alike (the
     * unequal-exponent form cancels terms up to ~1e5 times larger than the
     * result when the exponents are close) */
    const int sw = ai > ak;
    const double *P = sw ? q : p, *Q = sw ? p : q;
    const double eI = sw ? ek : ei, eK = sw ? ei : ek, aI = sw ? ak : ai, aK = sw ? ai : ak;
    if (same) {
        dmpik[0] = 1.0 - (1.0 + 11.0 * P[1] / 16.0 + 3.0 * P[2] / 16.0 + P[3] / 48.0) * eI;
        dmpik[1] = 1.0 - (1.0 + P[1] + 0.5 * P[2] + 7.0 * P[3] / 48.0 + P[4] / 48.0) * eI;
        dmpik[2] = 1.0 - (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + P[4] / 24.0 + P[5] / 144.0) * eI;
        dmpik[3] = 1.0 - (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + P[4] / 24.0 + P[5] / 120.0 + P[6] / 720.0) * eI;
        dmpik[4] = 1.0 - (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + P[4] / 24.0 + P[5] / 120.0 + P[6] / 720.0
                          + P[7] / 5040.0) * eI;
        dmpik[5] = 1.0 - (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + P[4] / 24.0 + P[5] / 120.0 + P[6] / 720.0
                          + P[7] / 5040.0 + P[8] / 45360.0) * eI;
    } else {
        const double ai2 = aI * aI, ak2 = aK * aK;
        const double ti = ak2 / (ak2 - ai2), tk = ai2 / (ai2 - ak2);
        const double ti2 = ti * ti, tk2 = tk * tk;
        dmpik[0] = 1.0 - ti2 * (1.0 + 2.0 * tk + 0.5 * P[1]) * eI - tk2 * (1.0 + 2.0 * ti + 0.5 * Q[1]) * eK;
        dmpik[1] = 1.0 - ti2 * (1.0 + P[1] + 0.5 * P[2]) * eI - tk2 * (1.0 + Q[1] + 0.5 * Q[2]) * eK
                 - 2.0 * ti2 * tk * (1.0 + P[1]) * eI - 2.0 * tk2 * ti * (1.0 + Q[1]) * eK;
        dmpik[2] = 1.0 - ti2 * (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0) * eI
                 - tk2 * (1.0 + Q[1] + 0.5 * Q[2] + Q[3] / 6.0) * eK
                 - 2.0 * ti2 * tk * (1.0 + P[1] + P[2] / 3.0) * eI
                 - 2.0 * tk2 * ti * (1.0 + Q[1] + Q[2] / 3.0) * eK;
        dmpik[3] = 1.0 - ti2 * (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + P[4] / 30.0) * eI
                 - tk2 * (1.0 + Q[1] + 0.5 * Q[2] + Q[3] / 6.0 + Q[4] / 30.0) * eK
                 - 2.0 * ti2 * tk * (1.0 + P[1] + 2.0 * P[2] / 5.0 + P[3] / 15.0) * eI
                 - 2.0 * tk2 * ti * (1.0 + Q[1] + 2.0 * Q[2] / 5.0 + Q[3] / 15.0) * eK;
        dmpik[4] = 1.0 - ti2 * (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + 4.0 * P[4] / 105.0 + P[5] / 210.0) * eI
                 - tk2 * (1.0 + Q[1] + 0.5 * Q[2] + Q[3] / 6.0 + 4.0 * Q[4] / 105.0 + Q[5] / 210.0) * eK
                 - 2.0 * ti2 * tk * (1.0 + P[1] + 3.0 * P[2] / 7.0 + 2.0 * P[3] / 21.0 + P[4] / 105.0) * eI
                 - 2.0 * tk2 * ti * (1.0 + Q[1] + 3.0 * Q[2] / 7.0 + 2.0 * Q[3] / 21.0 + Q[4] / 105.0) * eK;
        dmpik[5] = 1.0 - ti2 * (1.0 + P[1] + 0.5 * P[2] + P[3] / 6.0 + 5.0 * P[4] / 126.0 + 2.0 * P[5] / 315.0
                                + P[6] / 1890.0) * eI
                 - tk2 * (1.0 + Q[1] + 0.5 * Q[2] + Q[3] / 6.0 + 5.0 * Q[4] / 126.0 + 2.0 * Q[5] / 315.0
                          + Q[6] / 1890.0) * eK
                 - 2.0 * ti2 * tk * (1.0 + P[1] + 4.0 * P[2] / 9.0 + P[3] / 9.0 + P[4] / 63.0 + P[5] / 945.0) * eI
                 - 2.0 * tk2 * ti * (1.0 + Q[1] + 4.0 * Q[2] / 9.0 + Q[3] / 9.0 + Q[4] / 63.0 + Q[5] / 945.0) * eK;
    }
}

static void damp_gordon2(double r, double ai, double ak, double *dmpi, double *dmpk, double *dmpik)
{
    const double di = ai * r, dk = ak * r;
    const double ei = exp(-di), ek = exp(-dk);
    const int same = fabs(ai - ak) < 1e-3;
    double p[7], q[7];
    p[0] = 1.0; q[0] = 1.0;
    for (int j = 1; j < 7; j++) { p[j] = p[j - 1] * di; q[j] = q[j - 1] * dk; }
#define CV2(P, E, out)                                                     \
    out[0] = 1.0 - E;                                                      \
    out[1] = 1.0 - (1.0 + P[1]) * E;                                       \
    out[2] = 1.0 - (1.0 + P[1] + P[2] / 3.0) * E;                          \
    out[3] = 1.0 - (1.0 + P[1] + 0.4 * P[2] + P[3] / 15.0) * E;            \
    out[4] = 0.0;
    CV2(p, ei, dmpi);
    if (same) memcpy(dmpk, dmpi, 5 * sizeof(double));
    else { CV2(q, ek, dmpk); }
#undef CV2
    if (same) {
        dmpik[0] = 1.0 - (1.0 + 0.5 * p[1]) * ei;
        dmpik[1] = 1.0 - (1.0 + p[1] + 0.5 * p[2]) * ei;
        dmpik[2] = 1.0 - (1.0 + p[1] + 0.5 * p[2] + p[3] / 6.0) * ei;
        dmpik[3] = 1.0 - (1.0 + p[1] + 0.5 * p[2] + p[3] / 6.0 + p[4] / 30.0) * ei;
        dmpik[4] = 1.0 - (1.0 + p[1] + 0.5 * p[2] + p[3] / 6.0 + 4.0 * p[4] / 105.0 + p[5] / 210.0) * ei;
        dmpik[5] = 1.0 - (1.0 + p[1] + 0.5 * p[2] + p[3] / 6.0 + 5.0 * p[4] / 126.0 + 2.0 * p[5] / 315.0
                          + p[6] / 1890.0) * ei;
    } else {
        const double ai2 = ai * ai, ak2 = ak * ak;
        const double ti = ak2 / (ak2 - ai2), tk = ai2 / (ai2 - ak2);
        dmpik[0] = 1.0 - ti * ei - tk * ek;
        dmpik[1] = 1.0 - ti * (1.0 + p[1]) * ei - tk * (1.0 + q[1]) * ek;
        dmpik[2] = 1.0 - ti * (1.0 + p[1] + p[2] / 3.0) * ei - tk * (1.0 + q[1] + q[2] / 3.0) * ek;
        dmpik[3] = 1.0 - ti * (1.0 + p[1] + 0.4 * p[2] + p[3] / 15.0) * ei
                 - tk * (1.0 + q[1] + 0.4 * q[2] + q[3] / 15.0) * ek;
        dmpik[4] = 1.0 - ti * (1.0 + p[1] + 3.0 * p[2] / 7.0 + 2.0 * p[3] / 21.0 + p[4] / 105.0) * ei
                 - tk * (1.0 + q[1] + 3.0 * q[2] / 7.0 + 2.0 * q[3] / 21.0 + q[4] / 105.0) * ek;
        dmpik[5] = 1.0 - ti * (1.0 + p[1] + 4.0 * p[2] / 9.0 + p[3] / 9.0 + p[4] / 63.0 + p[5] / 945.0) * ei
                 - tk * (1.0 + q[1] + 4.0 * q[2] / 9.0 + q[3] / 9.0 + q[4] / 63.0 + q[5] / 945.0) * ek;
    }
}

/* pairMpoleCP, pairMpole.h:186-389: the interaction splits into core-core
 * (undamped), core(k) with the valence charge/dipole/quadrupole of i (dmpi),
 * core(i) with those of k (dmpk) and valence-valence (dmpik). */
static void pair_cp(int pentype, double r2, const double *r, double mscale, double f,
                    double corei, double alphai, const double *mi, /* mi[0] = valence */
                    double corek, double alphak, const double *mk, int do_g, pair_out *o)
{
    double rr[6], Ri[6], Rk[6], Rik[6], dmpi[6] = {0}, dmpk[6] = {0}, dmpik[6];
    rr_series(r2, mscale, f, rr);
    const double rad = sqrt(r2);
    if (pentype == 1) damp_gordon1(rad, alphai, alphak, dmpi, dmpk, dmpik);
    else damp_gordon2(rad, alphai, alphak, dmpi, dmpk, dmpik);
    for (int j = 0; j < 6; j++) {
        Ri[j] = dmpi[j < 5 ? j : 4] * rr[j];
        Rk[j] = dmpk[j < 5 ? j : 4] * rr[j];
        Rik[j] = dmpik[j] * rr[j];
    }
    const double cci[10] = {corei, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double cck[10] = {corek, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    pair_out a, b, c, d;
    pair_series(rr, cci, cck, r, do_g, &a);
    pair_series(Ri, mi, cck, r, do_g, &b);
    pair_series(Rk, cci, mk, r, do_g, &c);
    pair_series(Rik, mi, mk, r, do_g, &d);
    o->e = a.e + d.e + b.e + c.e;
    o->phi_i = b.phi_i + d.phi_i;
    o->phi_k = c.phi_k + d.phi_k;
    if (!do_g) return;
    for (int j = 0; j < 3; j++) {
        o->frc[j] = a.frc[j] + b.frc[j] + c.frc[j] + d.frc[j];
        o->trqi[j] = b.trqi[j] + c.trqi[j] + d.trqi[j];
        o->trqk[j] = b.trqk[j] + c.trqk[j] + d.trqk[j];
    }
}

void orc_pair_mpole(double r2, double xr, double yr, double zr, double mscale, double f,
                    const double *mi, const double *mk, double *out)
{
    double R[6], r[3] = {xr, yr, zr};
    pair_out o;
    rr_series(r2, mscale, f, R);
    pair_series(R, mi, mk, r, 1, &o);
    out[0] = o.e;
    for (int j = 0; j < 3; j++) { out[1 + j] = o.frc[j]; out[4 + j] = o.trqi[j]; out[7 + j] = o.trqk[j]; }
    out[10] = -xr * o.frc[0];
    out[11] = -0.5 * (yr * o.frc[0] + xr * o.frc[1]);
    out[12] = -0.5 * (zr * o.frc[0] + xr * o.frc[2]);
    out[13] = -yr * o.frc[1];
    out[14] = -0.5 * (zr * o.frc[1] + yr * o.frc[2]);
    out[15] = -zr * o.frc[2];
}

/* ------------------------------------------------------------------------
 * Minimum image, image.cpp:37-124 (lattice constants as lattice.cpp sets them)
 * ---------------------------------------------------------------------- */
typedef struct {
    double beta_sin, beta_cos, gamma_sin, gamma_cos, beta_term, gamma_term;
} box_trig;

static box_trig box_trig_of(const orc_box *b)
{
    box_trig t = {1.0, 0.0, 1.0, 0.0, 0.0, 1.0};
    if (b->kind == ORC_BOX_TRI) {
        const double ac = cos(b->alpha / RADIAN);
        t.beta_sin = sin(b->beta / RADIAN);
        t.beta_cos = cos(b->beta / RADIAN);
        t.gamma_sin = sin(b->gamma / RADIAN);
        t.gamma_cos = cos(b->gamma / RADIAN);
        t.beta_term = (ac - t.beta_cos * t.gamma_cos) / t.gamma_sin;
        t.gamma_term = sqrt(t.beta_sin * t.beta_sin - t.beta_term * t.beta_term);
    } else if (b->kind == ORC_BOX_MONO) {
        t.beta_sin = sin(b->beta / RADIAN);
        t.beta_cos = cos(b->beta / RADIAN);
        t.gamma_term = t.beta_sin;
    }
    return t;
}

static inline void wrap_while(double *v, double cell)
{
    const double half = 0.5 * cell;
    while (fabs(*v) > half) *v -= copysign(cell, *v);
}

void orc_image(const orc_box *b, double *xr, double *yr, double *zr)
{
    const box_trig t = box_trig_of(b);
    double x = *xr, y = *yr, z = *zr;
    switch (b->kind) {
    case ORC_BOX_ORTHO:
        wrap_while(&x, b->xbox); wrap_while(&y, b->ybox); wrap_while(&z, b->zbox);
        break;
    case ORC_BOX_MONO:
        z = z / t.beta_sin;
        x = x - z * t.beta_cos;
        wrap_while(&x, b->xbox); wrap_while(&y, b->ybox); wrap_while(&z, b->zbox);
        x = x + z * t.beta_cos;
        z = z * t.beta_sin;
        break;
    case ORC_BOX_TRI:
        z = z / t.gamma_term;
        y = (y - z * t.beta_term) / t.gamma_sin;
        x = x - y * t.gamma_cos - z * t.beta_cos;
        wrap_while(&x, b->xbox); wrap_while(&y, b->ybox); wrap_while(&z, b->zbox);
        x = x + y * t.gamma_cos + z * t.beta_cos;
        y = y * t.gamma_sin + z * t.beta_term;
        z = z * t.gamma_term;
        break;
    case ORC_BOX_OCTA:
        wrap_while(&x, b->xbox); wrap_while(&y, b->ybox); wrap_while(&z, b->zbox);
        if (fabs(x) + fabs(y) + fabs(z) > 0.75 * b->xbox) {
            x -= copysign(0.5 * b->xbox, x);
            y -= copysign(0.5 * b->ybox, y);
            z -= copysign(0.5 * b->zbox, z);
        }
        break;
    case ORC_BOX_DODECA: {
        const double root2 = 1.41421356237309504880;
        wrap_while(&x, b->xbox); wrap_while(&y, b->ybox);
        z = z - root2 * b->zbox * round(z / (b->zbox * root2));
        const double corr = 0.5 * b->xbox * (double)(int)(fabs(x / b->xbox) + fabs(y / b->ybox)
                                                           + fabs(root2 * z / b->zbox));
        x -= copysign(corr, x);
        y -= copysign(corr, y);
        z -= copysign(corr, z) * root2;
        break;
    }
    default:
        break;
    }
    *xr = x; *yr = y; *zr = z;
}

/* ------------------------------------------------------------------------
 * Local frames: chkpole.cpp:44-89, rotmat rotpole.cpp:102-345,
 * rotsite rotpole.cpp:356-425
 * ---------------------------------------------------------------------- */
void orc_chkpole(orc_mpole_sys *s, double *pole)
{
    for (int i = 0; i < s->n; i++) {
        const int k = s->yaxis[i];
        if (s->polaxe[i] != ORC_FRAME_ZTHENX || k == 0) continue;
        const int ib = s->zaxis[i] - 1, ic = s->xaxis[i] - 1, id = abs(k) - 1;
        const double ad[3] = {s->x[i] - s->x[id], s->y[i] - s->y[id], s->z[i] - s->z[id]};
        const double bd[3] = {s->x[ib] - s->x[id], s->y[ib] - s->y[id], s->z[ib] - s->z[id]};
        const double cd[3] = {s->x[ic] - s->x[id], s->y[ic] - s->y[id], s->z[ic] - s->z[id]};
        /* signed volume ad . (bd x cd), chkpole.cpp:73-78 */
        const double c1 = bd[1] * cd[2] - bd[2] * cd[1];
        const double c2 = cd[1] * ad[2] - cd[2] * ad[1];
        const double c3 = ad[1] * bd[2] - ad[2] * bd[1];
        const double vol = ad[0] * c1 + bd[0] * c2 + cd[0] * c3;
        if ((k < 0 && vol > 0.0) || (k > 0 && vol < 0.0)) {
            s->yaxis[i] = -k;
            double *p = pole + 13 * i;
            p[2] = -p[2]; p[5] = -p[5]; p[7] = -p[7]; p[9] = -p[9]; p[11] = -p[11];
        }
    }
}

static void unit_from(const orc_mpole_sys *s, int from, int to, double *u, double *len)
{
    u[0] = s->x[to] - s->x[from];
    u[1] = s->y[to] - s->y[from];
    u[2] = s->z[to] - s->z[from];
    *len = sqrt(dot3(u, u));
}

/* rows a[0] = x axis, a[1] = y axis, a[2] = z axis of the local frame */
static void rotmat(const orc_mpole_sys *s, int i, double a[3][3], int *planar)
{
    const int iz = s->zaxis[i] - 1, ix = s->xaxis[i] - 1, iy = abs(s->yaxis[i]) - 1;
    double d[3], r, dot;
    memset(a, 0, 9 * sizeof(double));
    a[0][0] = a[1][1] = a[2][2] = 1.0;
    *planar = 0;
    switch (s->polaxe[i]) {
    case ORC_FRAME_ZONLY: {
        unit_from(s, i, iz, d, &r);
        for (int j = 0; j < 3; j++) a[2][j] = d[j] / r;
        double v[3] = {1.0, 0.0, 0.0};
        dot = a[2][0];
        if (fabs(dot) > 0.707) { v[0] = 0.0; v[1] = 1.0; dot = a[2][1]; }
        for (int j = 0; j < 3; j++) v[j] -= dot * a[2][j];
        r = sqrt(dot3(v, v));
        for (int j = 0; j < 3; j++) a[0][j] = v[j] / r;
        break;
    }
    case ORC_FRAME_ZTHENX:
        unit_from(s, i, iz, d, &r);
        for (int j = 0; j < 3; j++) a[2][j] = d[j] / r;
        unit_from(s, i, ix, d, &r);
        dot = dot3(d, a[2]);
        for (int j = 0; j < 3; j++) d[j] -= dot * a[2][j];
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) a[0][j] = d[j] / r;
        break;
    case ORC_FRAME_BISECTOR: {
        double u1[3], u2[3], r1, r2;
        unit_from(s, i, iz, u1, &r1);
        unit_from(s, i, ix, u2, &r2);
        for (int j = 0; j < 3; j++) { u1[j] /= r1; u2[j] /= r2; d[j] = u1[j] + u2[j]; }
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) a[2][j] = d[j] / r;
        dot = dot3(u2, a[2]);
        for (int j = 0; j < 3; j++) d[j] = u2[j] - dot * a[2][j];
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) a[0][j] = d[j] / r;
        break;
    }
    case ORC_FRAME_ZBISECT: {
        double u1[3], u2[3], r1, r2;
        unit_from(s, i, iz, d, &r);
        for (int j = 0; j < 3; j++) a[2][j] = d[j] / r;
        unit_from(s, i, ix, u1, &r1);
        unit_from(s, i, iy, u2, &r2);
        for (int j = 0; j < 3; j++) { u1[j] /= r1; u2[j] /= r2; d[j] = u1[j] + u2[j]; }
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) d[j] /= r;
        dot = dot3(d, a[2]);
        if (180.0 - RADIAN * acos(dot) < 0.0) { /* planar fallback, rotpole.cpp:253-265 */
            *planar = 1;
            cross3(u1, u2, d);
            dot = dot3(d, a[2]);
            if (dot < 0.0) { for (int j = 0; j < 3; j++) d[j] = -d[j]; dot = -dot; }
        }
        for (int j = 0; j < 3; j++) d[j] -= dot * a[2][j];
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) a[0][j] = d[j] / r;
        break;
    }
    case ORC_FRAME_THREEFOLD: {
        double u1[3], u2[3], u3[3], r1, r2, r3;
        unit_from(s, i, iz, u1, &r1);
        unit_from(s, i, ix, u2, &r2);
        unit_from(s, i, iy, u3, &r3);
        for (int j = 0; j < 3; j++) { u1[j] /= r1; u2[j] /= r2; u3[j] /= r3; d[j] = u1[j] + u2[j] + u3[j]; }
        r = sqrt(dot3(d, d));
        if (r < 0.0) { /* planar fallback, rotpole.cpp:304-327 (never taken: r >= 0) */
            *planar = 1;
            const double p2[3] = {s->x[ix] - s->x[iz], s->y[ix] - s->y[iz], s->z[ix] - s->z[iz]};
            const double p3[3] = {s->x[iy] - s->x[iz], s->y[iy] - s->y[iz], s->z[iy] - s->z[iz]};
            double p4[3];
            cross3(p2, p3, p4);
            const double sg = dot3(p4, d) > 0.0 ? 1.0 : -1.0;
            for (int j = 0; j < 3; j++) d[j] = sg * p4[j];
            r = sqrt(dot3(d, d));
        }
        for (int j = 0; j < 3; j++) a[2][j] = d[j] / r;
        dot = dot3(u1, a[2]);
        for (int j = 0; j < 3; j++) d[j] = u1[j] - dot * a[2][j];
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) a[0][j] = d[j] / r;
        break;
    }
    default:
        break;
    }
    /* y = z x x, rotpole.cpp:342-344 */
    a[1][0] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
    a[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    a[1][2] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
}

/* global = A^T local for dipoles; global Q = A^T Q_local A (rotsite) */
static void rotsite(int frame, const double a[3][3], int planar, const double *in, double *out)
{
    double sp[13];
    memcpy(sp, in, sizeof sp);
    if (planar) {
        if (frame == ORC_FRAME_ZBISECT) {
            sp[1] = 0.0; sp[6] = 0.0; sp[10] = 0.0;
            sp[4] = 0.5 * (sp[4] + sp[8]);
            sp[8] = sp[4];
        } else if (frame == ORC_FRAME_THREEFOLD) {
            for (int j = 1; j < 13; j++) sp[j] = 0.0;
        }
    }
    out[0] = sp[0];
    for (int g = 0; g < 3; g++) out[1 + g] = sp[1] * a[0][g] + sp[2] * a[1][g] + sp[3] * a[2][g];
    /* local quadrupole M[l][m] = sp[4 + 3l + m] */
    for (int g = 0; g < 3; g++)
        for (int h = 0; h < 3; h++) {
            double v = 0.0;
            for (int l = 0; l < 3; l++)
                for (int m = 0; m < 3; m++) v += a[l][g] * a[m][h] * sp[4 + 3 * l + m];
            out[4 + 3 * g + h] = v;
        }
}

void orc_rotpole(const orc_mpole_sys *s, const double *pole, double *rpole)
{
    for (int i = 0; i < s->n; i++) {
        double a[3][3];
        int planar;
        rotmat(s, i, a, &planar);
        rotsite(s->polaxe[i], a, planar, pole + 13 * i, rpole + 13 * i);
    }
}

/* ------------------------------------------------------------------------
 * Torque -> force on the frame atoms, torque.cpp:42-397
 * ---------------------------------------------------------------------- */
static void unitv(double *v, double *len)
{
    *len = sqrt(dot3(v, v));
    v[0] /= *len; v[1] /= *len; v[2] /= *len;
}

void orc_torque(const orc_mpole_sys *s, const double *te, double *de, double *vir)
{
    for (int i = 0; i < s->n; i++) {
        const int frame = s->polaxe[i];
        if (frame == ORC_FRAME_NONE) continue;
        const double *t = te + 3 * i;
        const int ia = s->zaxis[i] - 1, ib = i, ic = s->xaxis[i] - 1, id = abs(s->yaxis[i]) - 1;
        double u[3], v[3], w[3], usiz, vsiz, wsiz;
        u[0] = s->x[ia] - s->x[ib]; u[1] = s->y[ia] - s->y[ib]; u[2] = s->z[ia] - s->z[ib];
        usiz = sqrt(dot3(u, u));
        if (frame != ORC_FRAME_ZONLY) {
            v[0] = s->x[ic] - s->x[ib]; v[1] = s->y[ic] - s->y[ib]; v[2] = s->z[ic] - s->z[ib];
            vsiz = sqrt(dot3(v, v));
        } else {
            v[0] = 1.0; v[1] = 0.0; v[2] = 0.0;
            vsiz = 1.0;
            if (fabs(u[0] / usiz) > 0.866) { v[0] = 0.0; v[1] = 1.0; }
        }
        if (frame == ORC_FRAME_ZBISECT || frame == ORC_FRAME_THREEFOLD) {
            w[0] = s->x[id] - s->x[ib]; w[1] = s->y[id] - s->y[ib]; w[2] = s->z[id] - s->z[ib];
        } else {
            cross3(u, v, w);
        }
        wsiz = sqrt(dot3(w, w));
        for (int j = 0; j < 3; j++) { u[j] /= usiz; v[j] /= vsiz; w[j] /= wsiz; }

        double rr[3], ss[3], rsiz, ssiz;
        if (frame == ORC_FRAME_ZBISECT) {
            for (int j = 0; j < 3; j++) rr[j] = v[j] + w[j];
            rsiz = sqrt(dot3(rr, rr));
            cross3(u, rr, ss);
            ssiz = sqrt(dot3(ss, ss));
            for (int j = 0; j < 3; j++) { rr[j] /= rsiz; ss[j] /= ssiz; }
        }
        const double dphidu = -dot3(t, u), dphidv = -dot3(t, v), dphidw = -dot3(t, w);
        double dphidr = 0.0, dphids = 0.0;
        if (frame == ORC_FRAME_ZBISECT) { dphidr = -dot3(t, rr); dphids = -dot3(t, ss); }

        double uv[3], uw[3], vw[3], l;
        cross3(v, u, uv); unitv(uv, &l);
        cross3(w, u, uw); unitv(uw, &l);
        cross3(w, v, vw); unitv(vw, &l);
        double ur[3], us[3], vs[3], ws[3];
        if (frame == ORC_FRAME_ZBISECT) {
            cross3(rr, u, ur); unitv(ur, &l);
            cross3(ss, u, us); unitv(us, &l);
            cross3(ss, v, vs); unitv(vs, &l);
            cross3(ss, w, ws); unitv(ws, &l);
        }
        const double uvcos = dot3(u, v), uvsin = sqrt(1.0 - uvcos * uvcos);
        double ursin = 0, vscos = 0, vssin = 0, wscos = 0, wssin = 0;
        if (frame == ORC_FRAME_ZBISECT) {
            const double urcos = dot3(u, rr);
            ursin = sqrt(1.0 - urcos * urcos);
            vscos = dot3(v, ss); vssin = sqrt(1.0 - vscos * vscos);
            wscos = dot3(w, ss); wssin = sqrt(1.0 - wscos * wscos);
        }
        double t1[3], t2[3], ut1sin = 0, ut2sin = 0;
        if (frame == ORC_FRAME_ZBISECT) {
            for (int j = 0; j < 3; j++) { t1[j] = v[j] - ss[j] * vscos; t2[j] = w[j] - ss[j] * wscos; }
            unitv(t1, &l);
            unitv(t2, &l);
            const double c1 = dot3(u, t1), c2 = dot3(u, t2);
            ut1sin = sqrt(1.0 - c1 * c1);
            ut2sin = sqrt(1.0 - c2 * c2);
        }

        double fz[3] = {0, 0, 0}, fx[3] = {0, 0, 0}, fy[3] = {0, 0, 0};
        if (frame == ORC_FRAME_ZONLY) {
            for (int j = 0; j < 3; j++) {
                const double du = uv[j] * dphidv / (usiz * uvsin) + uw[j] * dphidw / usiz;
                fz[j] = du;
            }
        } else if (frame == ORC_FRAME_ZTHENX) {
            for (int j = 0; j < 3; j++) {
                fz[j] = uv[j] * dphidv / (usiz * uvsin) + uw[j] * dphidw / usiz;
                fx[j] = -uv[j] * dphidu / (vsiz * uvsin);
            }
        } else if (frame == ORC_FRAME_BISECTOR) {
            for (int j = 0; j < 3; j++) {
                fz[j] = uv[j] * dphidv / (usiz * uvsin) + 0.5 * uw[j] * dphidw / usiz;
                fx[j] = -uv[j] * dphidu / (vsiz * uvsin) + 0.5 * vw[j] * dphidw / vsiz;
            }
        } else if (frame == ORC_FRAME_ZBISECT) {
            for (int j = 0; j < 3; j++) {
                fz[j] = ur[j] * dphidr / (usiz * ursin) + us[j] * dphids / usiz;
                fx[j] = (vssin * ss[j] - vscos * t1[j]) * dphidu / (vsiz * (ut1sin + ut2sin));
                fy[j] = (wssin * ss[j] - wscos * t2[j]) * dphidu / (wsiz * (ut1sin + ut2sin));
            }
        } else if (frame == ORC_FRAME_THREEFOLD) {
            double p[3], psiz, del[3], eps[3], dsiz, r2v[3], rs;
            for (int j = 0; j < 3; j++) p[j] = u[j] + v[j] + w[j];
            unitv(p, &psiz);
            const double wpcos = dot3(w, p), upcos = dot3(u, p), vpcos = dot3(v, p);
            /* w-axis piece */
            for (int j = 0; j < 3; j++) r2v[j] = u[j] + v[j];
            unitv(r2v, &rs);
            double c = dot3(r2v, w), sn = sqrt(1.0 - c * c);
            double dr = -dot3(t, r2v);
            cross3(r2v, w, del); unitv(del, &dsiz);
            double dd = -dot3(t, del);
            cross3(del, w, eps);
            for (int j = 0; j < 3; j++) fy[j] = del[j] * dr / (wsiz * sn) + eps[j] * dd * wpcos / (wsiz * psiz);
            /* u-axis piece */
            for (int j = 0; j < 3; j++) r2v[j] = v[j] + w[j];
            unitv(r2v, &rs);
            c = dot3(r2v, u); sn = sqrt(1.0 - c * c);
            dr = -dot3(t, r2v);
            cross3(r2v, u, del); unitv(del, &dsiz);
            dd = -dot3(t, del);
            cross3(del, u, eps);
            for (int j = 0; j < 3; j++) fz[j] = del[j] * dr / (usiz * sn) + eps[j] * dd * upcos / (usiz * psiz);
            /* v-axis piece */
            for (int j = 0; j < 3; j++) r2v[j] = u[j] + w[j];
            unitv(r2v, &rs);
            c = dot3(r2v, v); sn = sqrt(1.0 - c * c);
            dr = -dot3(t, r2v);
            cross3(r2v, v, del); unitv(del, &dsiz);
            dd = -dot3(t, del);
            cross3(del, v, eps);
            for (int j = 0; j < 3; j++) fx[j] = del[j] * dr / (vsiz * sn) + eps[j] * dd * vpcos / (vsiz * psiz);
        }
        for (int j = 0; j < 3; j++) {
            de[3 * ia + j] += fz[j];
            if (ic >= 0 && frame != ORC_FRAME_ZONLY) de[3 * ic + j] += fx[j];
            if (id >= 0 && (frame == ORC_FRAME_ZBISECT || frame == ORC_FRAME_THREEFOLD)) de[3 * id + j] += fy[j];
            de[3 * ib + j] -= fz[j] + fx[j] + fy[j];
        }
        if (vir) {
            const int jz = ia >= 0 ? ia : i, jx = ic >= 0 ? ic : i, jy = id >= 0 ? id : i;
            const double rz[3] = {s->x[jz] - s->x[i], s->y[jz] - s->y[i], s->z[jz] - s->z[i]};
            const double rx[3] = {s->x[jx] - s->x[i], s->y[jx] - s->y[i], s->z[jx] - s->z[i]};
            const double ry[3] = {s->x[jy] - s->x[i], s->y[jy] - s->y[i], s->z[jy] - s->z[i]};
            double W[3][3];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) W[a][b] = rx[a] * fx[b] + ry[a] * fy[b] + rz[a] * fz[b];
            vir[0] += W[0][0];
            vir[4] += W[1][1];
            vir[8] += W[2][2];
            const double vxy = 0.5 * (W[1][0] + W[0][1]);
            const double vxz = 0.5 * (W[2][0] + W[0][2]);
            const double vyz = 0.5 * (W[2][1] + W[1][2]);
            vir[1] += vxy; vir[3] += vxy;
            vir[2] += vxz; vir[6] += vxz;
            vir[5] += vyz; vir[7] += vyz;
        }
    }
}

/* ------------------------------------------------------------------------
 * Charge flux: alterchg.cpp:36-169 and dcflux.cpp:31-225
 * ---------------------------------------------------------------------- */
void orc_charge_flux_delta(const orc_mpole_sys *s, double *pdelta)
{
    memset(pdelta, 0, s->n * sizeof(double));
    for (int b = 0; b < s->nbond; b++) {
        const int ia = s->ibnd[2 * b], ib = s->ibnd[2 * b + 1];
        const double d[3] = {s->x[ia] - s->x[ib], s->y[ia] - s->y[ib], s->z[ia] - s->z[ib]};
        const double dq = s->bflx[b] * (sqrt(dot3(d, d)) - s->bl[b]);
        pdelta[ia] -= dq;
        pdelta[ib] += dq;
    }
    for (int g = 0; g < s->nangle; g++) {
        const int ia = s->iang[3 * g], ib = s->iang[3 * g + 1], ic = s->iang[3 * g + 2];
        const double ab[3] = {s->x[ia] - s->x[ib], s->y[ia] - s->y[ib], s->z[ia] - s->z[ib]};
        const double cb[3] = {s->x[ic] - s->x[ib], s->y[ic] - s->y[ib], s->z[ic] - s->z[ib]};
        const double rab = sqrt(dot3(ab, ab)), rcb = sqrt(dot3(cb, cb));
        if (rab > 0.0 && rcb > 0.0) {
            double cosine = dot3(ab, cb) / (rab * rcb);
            cosine = fmin(1.0, fmax(-1.0, cosine));
            const double angle = RADIAN * acos(cosine);
            const double dth = (angle - s->anat[g]) / RADIAN;
            const double rab0 = s->bl[s->balist[2 * g]], rcb0 = s->bl[s->balist[2 * g + 1]];
            const double dq1 = s->abflx[2 * g] * (rcb - rcb0) + s->aflx[2 * g] * dth;
            const double dq2 = s->abflx[2 * g + 1] * (rab - rab0) + s->aflx[2 * g + 1] * dth;
            pdelta[ia] += dq1;
            pdelta[ib] -= dq1 + dq2;
            pdelta[ic] += dq2;
        }
    }
}

void orc_charge_flux_grad(const orc_mpole_sys *s, const double *pot, double *de, double *vir)
{
    const int n = s->n;
    double *fc = (double *)calloc(3 * (size_t)n, sizeof(double));
    for (int b = 0; b < s->nbond; b++) {
        const int ia = s->ibnd[2 * b], ib = s->ibnd[2 * b + 1];
        const double d[3] = {s->x[ia] - s->x[ib], s->y[ia] - s->y[ib], s->z[ia] - s->z[ib]};
        const double pb = s->bflx[b] * (pot[ib] - pot[ia]) / sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) { fc[3 * ia + j] += pb * d[j]; fc[3 * ib + j] -= pb * d[j]; }
    }
    for (int g = 0; g < s->nangle; g++) {
        const int ia = s->iang[3 * g], ib = s->iang[3 * g + 1], ic = s->iang[3 * g + 2];
        const double ab[3] = {s->x[ia] - s->x[ib], s->y[ia] - s->y[ib], s->z[ia] - s->z[ib]};
        const double cb[3] = {s->x[ic] - s->x[ib], s->y[ic] - s->y[ib], s->z[ic] - s->z[ib]};
        const double rab2 = dot3(ab, ab), rcb2 = dot3(cb, cb);
        const double rab = sqrt(rab2), rcb = sqrt(rcb2);
        if (!(rab > 0.0 && rcb > 0.0)) continue;
        const double dpa = pot[ia] - pot[ib], dpc = pot[ic] - pot[ib];
        const double pb1 = dpa * s->abflx[2 * g], pb2 = dpc * s->abflx[2 * g + 1];
        double fa1[3], fc1[3];
        for (int j = 0; j < 3; j++) { fa1[j] = pb2 * ab[j] / rab; fc1[j] = pb1 * cb[j] / rcb; }
        const double rabc = rab * rcb, dotp = dot3(ab, cb);
        const double dra3 = dotp / (rab2 * rab * rcb), drc3 = dotp / (rab * rcb2 * rcb);
        const double term = -rabc / sqrt(rab2 * rcb2 - dotp * dotp);
        const double fterm = term * (dpa * s->aflx[2 * g] + dpc * s->aflx[2 * g + 1]);
        for (int j = 0; j < 3; j++) {
            const double fa2 = fterm * (cb[j] / rabc - ab[j] * dra3);
            const double fc2 = fterm * (ab[j] / rabc - cb[j] * drc3);
            fc[3 * ia + j] += fa1[j] + fa2;
            fc[3 * ic + j] += fc1[j] + fc2;
            fc[3 * ib + j] += -fa1[j] - fc1[j] - fa2 - fc2;
        }
    }
    for (int i = 0; i < n; i++) {
        const double fx = fc[3 * i], fy = fc[3 * i + 1], fz = fc[3 * i + 2];
        de[3 * i] += fx; de[3 * i + 1] += fy; de[3 * i + 2] += fz;
        if (vir) {
            /* dcflux.cpp:204-223 uses absolute coordinates */
            const double xi = s->x[i], yi = s->y[i], zi = s->z[i];
            vir[0] += xi * fx;
            vir[1] += yi * fx; vir[3] += yi * fx;
            vir[2] += zi * fx; vir[6] += zi * fx;
            vir[4] += yi * fy;
            vir[5] += zi * fy; vir[7] += zi * fy;
            vir[8] += zi * fz;
        }
    }
    free(fc);
}

/* ------------------------------------------------------------------------
 * empole_a<Mode, PenType, use_cf>, empole.cpp:96-357
 * ---------------------------------------------------------------------- */
int orc_empole(orc_mpole_sys *s, int mode, orc_mpole_out *out)
{
    const int n = s->n;
    const int do_a = mode == ORC_ANALYSIS;
    const int do_g = mode == ORC_GRADIENT || mode == ORC_VIRIAL;
    const int do_v = mode == ORC_VIRIAL;
    const int cp = s->pentype == 1 || s->pentype == 2;
    const int cf = cp && s->use_chgflx;

    double *pole = (double *)malloc(13 * (size_t)n * sizeof(double));
    double *rpole = (double *)malloc(13 * (size_t)n * sizeof(double));
    double *pval = (double *)malloc((size_t)n * sizeof(double));
    double *mscale = (double *)malloc((size_t)n * sizeof(double));
    double *te = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *pot = (double *)calloc((size_t)n, sizeof(double));
    int *o12 = (int *)malloc((n + 1) * sizeof(int)), *o13 = (int *)malloc((n + 1) * sizeof(int));
    int *o14 = (int *)malloc((n + 1) * sizeof(int)), *o15 = (int *)malloc((n + 1) * sizeof(int));
    if (!pole || !rpole || !pval || !mscale || !te || !pot || !o12 || !o13 || !o14 || !o15) return -1;
    o12[0] = o13[0] = o14[0] = o15[0] = 0;
    for (int i = 0; i < n; i++) {
        o12[i + 1] = o12[i] + s->n12[i];
        o13[i + 1] = o13[i] + s->n13[i];
        o14[i + 1] = o14[i] + s->n14[i];
        o15[i + 1] = o15[i] + s->n15[i];
    }
    memcpy(pole, s->pole, 13 * (size_t)n * sizeof(double));
    for (int i = 0; i < n; i++) pval[i] = s->pval0 ? s->pval0[i] : 0.0;

    /* charge flux alters monopoles and valence charges, energy.cpp:241 */
    if (s->use_chgflx) {
        double *pdelta = (double *)malloc((size_t)n * sizeof(double));
        orc_charge_flux_delta(s, pdelta);
        for (int i = 0; i < n; i++) {
            pole[13 * i] = s->mono0[i] + pdelta[i];
            pval[i] = s->pval0[i] + pdelta[i];
        }
        free(pdelta);
    }
    orc_chkpole(s, pole);
    orc_rotpole(s, pole, rpole);

    double em = 0.0, einter = 0.0;
    long nem = 0;
    double *dem = out->dem, *vir = out->vir;
    if (do_a && out->aem) memset(out->aem, 0, n * sizeof(double));
    if (do_g && dem) memset(dem, 0, 3 * (size_t)n * sizeof(double));
    double vloc[9] = {0};
    for (int i = 0; i < n; i++) mscale[i] = 1.0;

    for (int i = 0; i < n - 1; i++) {
        const int iz = s->zaxis[i] - 1, ix = s->xaxis[i] - 1, iy = abs(s->yaxis[i]) - 1;
        const int usei = s->use[i + 1] || s->use[iz + 1] || s->use[ix + 1] || s->use[iy + 1];
        double mi[10];
        pick10(rpole + 13 * i, mi);
        if (cp) mi[0] = pval[i];
        for (int j = o12[i]; j < o12[i + 1]; j++) mscale[s->i12[j]] = s->m2scale;
        for (int j = o13[i]; j < o13[i + 1]; j++) mscale[s->i13[j]] = s->m3scale;
        for (int j = o14[i]; j < o14[i + 1]; j++) mscale[s->i14[j]] = s->m4scale;
        for (int j = o15[i]; j < o15[i + 1]; j++) mscale[s->i15[j]] = s->m5scale;
        for (int k = i + 1; k < n; k++) {
            const int kz = s->zaxis[k] - 1, kx = s->xaxis[k] - 1, ky = abs(s->yaxis[k]) - 1;
            const int usek = s->use[k + 1] || s->use[kz + 1] || s->use[kx + 1] || s->use[ky + 1];
            if (!(usei || usek)) continue;
            double r[3] = {s->x[k] - s->x[i], s->y[k] - s->y[i], s->z[k] - s->z[i]};
            if (s->box.kind != ORC_BOX_NONE) orc_image(&s->box, &r[0], &r[1], &r[2]);
            const double r2 = dot3(r, r);
            if (r2 > s->off2) continue;
            double mk[10];
            pick10(rpole + 13 * k, mk);
            pair_out o;
            if (cp) {
                mk[0] = pval[k];
                pair_cp(s->pentype, r2, r, mscale[k], s->f, s->pcore[i], s->palpha[i], mi,
                        s->pcore[k], s->palpha[k], mk, do_g, &o);
            } else {
                double R[6];
                rr_series(r2, mscale[k], s->f, R);
                pair_series(R, mi, mk, r, do_g, &o);
            }
            if (do_a) {
                if (o.e != 0.0) nem++;
                if (out->aem) { out->aem[i] += 0.5 * o.e; out->aem[k] += 0.5 * o.e; }
                if (s->molcule[i] != s->molcule[k]) einter += o.e;
            }
            em += o.e;
            if (do_g) {
                for (int j = 0; j < 3; j++) {
                    if (dem) { dem[3 * i + j] += o.frc[j]; dem[3 * k + j] -= o.frc[j]; }
                    te[3 * i + j] += o.trqi[j];
                    te[3 * k + j] += o.trqk[j];
                }
                if (do_v) {
                    const double *f = o.frc;
                    const double vxy = -0.5 * (r[1] * f[0] + r[0] * f[1]);
                    const double vxz = -0.5 * (r[2] * f[0] + r[0] * f[2]);
                    const double vyz = -0.5 * (r[2] * f[1] + r[1] * f[2]);
                    vloc[0] += -r[0] * f[0];
                    vloc[4] += -r[1] * f[1];
                    vloc[8] += -r[2] * f[2];
                    vloc[1] += vxy; vloc[3] += vxy;
                    vloc[2] += vxz; vloc[6] += vxz;
                    vloc[5] += vyz; vloc[7] += vyz;
                }
                if (cf) { pot[i] += o.phi_i; pot[k] += o.phi_k; }
            }
        }
        for (int j = o12[i]; j < o12[i + 1]; j++) mscale[s->i12[j]] = 1.0;
        for (int j = o13[i]; j < o13[i + 1]; j++) mscale[s->i13[j]] = 1.0;
        for (int j = o14[i]; j < o14[i + 1]; j++) mscale[s->i14[j]] = 1.0;
        for (int j = o15[i]; j < o15[i + 1]; j++) mscale[s->i15[j]] = 1.0;
    }
    if (do_g && dem) {
        orc_torque(s, te, dem, do_v ? vloc : NULL);
        if (cf) orc_charge_flux_grad(s, pot, dem, do_v ? vloc : NULL);
    }
    out->em = em;
    out->einter = einter;
    out->nem = nem;
    if (vir) memcpy(vir, vloc, sizeof vloc);
    if (out->rpole) memcpy(out->rpole, rpole, 13 * (size_t)n * sizeof(double));
    if (out->pole_out) memcpy(out->pole_out, pole, 13 * (size_t)n * sizeof(double));
    free(pole); free(rpole); free(pval); free(mscale); free(te); free(pot);
    free(o12); free(o13); free(o14); free(o15);
    return 0;
}

void orc_damp_gordon1(double r, double ai, double ak, double *dmpi, double *dmpk, double *dmpik)
{
    damp_gordon1(r, ai, ak, dmpi, dmpk, dmpik);
}
