/*
 * pmq_oracle_elec.c -- FP64 CPU restatement of the electrostatics the
 * reference declares but does not compute: Ewald-summed permanent
 * multipoles (real space, smooth PME reciprocal space, self energy) and
 * mutual induced-dipole polarization (direct field, preconditioned CG,
 * energy and analytic gradient), for AMOEBA (Thole damping).
 *
 * TEST INFRASTRUCTURE ONLY (see pmq_oracle.h).
 *
 * Nothing here can be compared with the reference, which computes none of
 * it (empole.cpp:53-60 Ewald branches are empty; epolar is commented out at
 * energy.cpp:267).  The reference contributes the state these terms read:
 * polarizabilities and Thole factors (kpolar.cpp:461-537), polarization
 * groups (kpolar.cpp:592-853), p/d/u scale factors (polpot.h:47-66), Ewald
 * coefficient and grid (kewald.cpp:40-349), PCG defaults (kpolar.cpp:80-103).
 * The algorithms restated are the published ones the reference cites:
 *   - smooth PME for multipoles: Essmann et al., J. Chem. Phys. 103, 8577
 *     (1995); Sagui, Pedersen and Darden, J. Chem. Phys. 120, 73 (2004);
 *   - AMOEBA polarization with Thole damping and mutual induction:
 *     Ren and Ponder, J. Phys. Chem. B 107, 5933 (2003).
 * Every piece is pinned by tests/test_oracle_elec.py: direct lattice sums,
 * independence of the Ewald coefficient, a dense linear solve, central
 * finite differences of the energy, and the variational property.
 *
 * Virial (W_ab = dE/d(strain_ab) with every position and the box strained
 * and the local-frame multipoles carried by their frame atoms, symmetrized;
 * the reference's convention of empole.cpp:310-320): real-space pairs give
 * -(r_a g_b + r_b g_a)/2 with g the pair gradient on the first atom; the
 * frame rotation gives torque.cpp:361-395's frame virial; the reciprocal
 * energy of a distribution X, (f/2) sum_m G(m) |rho_X(m)|^2, gives
 *   (f/2) sum_m G |rho_X|^2 [2 h_a h_b (pi^2/a^2 + 1/h^2) - delta_ab]
 *   - f sum_j [d_jb phi_a + 2 sum_c q_jbc phi_ac]  (symmetrized),
 * the second line because dipoles and quadrupoles keep their Cartesian
 * components while the reciprocal vectors h strain (phi: X's reciprocal
 * potential at site j); cross energies between two distributions A, B
 * follow from E(A+B) - E(A) - E(B).  The self energies do not depend on
 * the box.  Pinned by central differences of the energy under strain.
 *
 * Energy definitions used (f = electric/dielec):
 *   E_mpole = real space [f (B_n - (1-m) rr_n) series] + reciprocal + self,
 *     self = -f a/sqrt(pi) sum_i [c^2 + (2a^2/3) d.d + (8a^4/5) q:q]
 *     (q in the reference's traceless x1/3 convention);
 *   E_pol = -(f/2) sum_i mu_d,i . E^p_i, with mu_d solving
 *     mu_d = alpha (E^d + T mu_d) and mu_p = alpha (E^p + T mu_p),
 *     E^d/E^p the permanent field with d/p exclusion scaling and T the
 *     u-scaled dipole field operator, both Thole damped and Ewald summed.
 */
#include "pmq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SQRTPI 1.77245385090551602730
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* ---------------------------------------------------------------------
 * generic multipole pair (same algebra as pair_series in pmq_oracle.c;
 * restated here so this file stands alone)
 * ------------------------------------------------------------------- */
typedef struct {
    double e, frc[3], trqi[3], trqk[3];
} gpair;

static inline void qmul(const double *q, const double *v, double *y)
{
    y[0] = q[0] * v[0] + q[1] * v[1] + q[2] * v[2];
    y[1] = q[1] * v[0] + q[3] * v[1] + q[4] * v[2];
    y[2] = q[2] * v[0] + q[4] * v[1] + q[5] * v[2];
}
static inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static void gpair_eval(const double *R, const double *mi, const double *mk, const double *r, int do_g, gpair *o)
{
    const double ci = mi[0], ck = mk[0];
    const double *di = mi + 1, *dk = mk + 1, *qi = mi + 4, *qk = mk + 4;
    double qirv[3], qkrv[3];
    qmul(qi, r, qirv);
    qmul(qk, r, qkrv);
    const double dir = dot3(di, r), dkr = dot3(dk, r), qir = dot3(qirv, r), qkr = dot3(qkrv, r);
    const double dik = dot3(di, dk), qik = dot3(qirv, qkrv), diqk = dot3(di, qkrv), dkqi = dot3(dk, qirv);
    const double qiqk = 2.0 * (qi[1] * qk[1] + qi[2] * qk[2] + qi[4] * qk[4]) + qi[0] * qk[0] + qi[3] * qk[3] + qi[5] * qk[5];
    const double t1 = ci * ck, t2 = ck * dir - ci * dkr + dik;
    const double t3 = ci * qkr + ck * qir - dir * dkr + 2.0 * (dkqi - diqk + qiqk);
    const double t4 = dir * qkr - dkr * qir - 4.0 * qik, t5 = qir * qkr;
    o->e = t1 * R[0] + t2 * R[1] + t3 * R[2] + t4 * R[3] + t5 * R[4];
    if (!do_g) return;
    const double de = t1 * R[1] + t2 * R[2] + t3 * R[3] + t4 * R[4] + t5 * R[5];
    const double g1 = -ck * R[1] + dkr * R[2] - qkr * R[3];
    const double g2 = ci * R[1] + dir * R[2] + qir * R[3];
    const double g3 = 2.0 * R[2];
    const double g4 = 2.0 * (-ck * R[2] + dkr * R[3] - qkr * R[4]);
    const double g5 = 2.0 * (-ci * R[2] - dir * R[3] - qir * R[4]);
    const double g6 = 4.0 * R[3];
    double qkdi[3], qidk[3], qiqkr[3], qkqir[3];
    qmul(qk, di, qkdi);
    qmul(qi, dk, qidk);
    qmul(qi, qkrv, qiqkr);
    qmul(qk, qirv, qkqir);
    for (int a = 0; a < 3; a++)
        o->frc[a] = de * r[a] + g1 * di[a] + g2 * dk[a] + g3 * (qkdi[a] - qidk[a]) + g4 * qirv[a] + g5 * qkrv[a]
                  + g6 * (qiqkr[a] + qkqir[a]);
    double dixdk[3], dixr[3], dkxr[3], rxqir[3], rxqkr[3], qkrxqir[3], rxqiqkr[3], rxqkqir[3], rxqkdi[3], rxqidk[3];
    double dixqkr[3], dkxqir[3];
    cross3(di, dk, dixdk); cross3(di, r, dixr); cross3(dk, r, dkxr);
    cross3(r, qirv, rxqir); cross3(r, qkrv, rxqkr); cross3(qkrv, qirv, qkrxqir);
    cross3(r, qiqkr, rxqiqkr); cross3(r, qkqir, rxqkqir); cross3(r, qkdi, rxqkdi); cross3(r, qidk, rxqidk);
    cross3(di, qkrv, dixqkr); cross3(dk, qirv, dkxqir);
    const double Qi[3][3] = {{qi[0], qi[1], qi[2]}, {qi[1], qi[3], qi[4]}, {qi[2], qi[4], qi[5]}};
    const double Qk[3][3] = {{qk[0], qk[1], qk[2]}, {qk[1], qk[3], qk[4]}, {qk[2], qk[4], qk[5]}};
    double A[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) A[a][b] = Qi[a][0] * Qk[0][b] + Qi[a][1] * Qk[1][b] + Qi[a][2] * Qk[2][b];
    const double eA[3] = {A[1][2] - A[2][1], A[2][0] - A[0][2], A[0][1] - A[1][0]};
    for (int a = 0; a < 3; a++) {
        const double dq = dixqkr[a] + dkxqir[a] - 2.0 * eA[a];
        o->trqi[a] = -R[1] * dixdk[a] + g1 * dixr[a] + g3 * (dq + rxqidk[a]) - g4 * rxqir[a] - g6 * (rxqiqkr[a] + qkrxqir[a]);
        o->trqk[a] = R[1] * dixdk[a] + g2 * dkxr[a] - g3 * (dq + rxqkdi[a]) - g5 * rxqkr[a] - g6 * (rxqkqir[a] - qkrxqir[a]);
    }
}

/* ---------------------------------------------------------------------
 * radial series
 * ------------------------------------------------------------------- */
/* undamped rr1..rr11 (with the (2n-1)!! factors) */
static void series_rr(double r2, double *rr)
{
    const double rinv = 1.0 / sqrt(r2), r2i = rinv * rinv;
    rr[0] = rinv;
    for (int j = 1; j < 6; j++) rr[j] = (2 * j - 1) * rr[j - 1] * r2i;
}
/* Ewald-screened B_0..B_5: B_0 = erfc(a r)/r, B_n = ((2n-1) B_{n-1} + (2a^2)^n e^{-a^2 r^2}/(a sqrt(pi)))/r^2 */
static void series_ewald(double r2, double a, double *bn)
{
    const double r = sqrt(r2);
    bn[0] = erfc(a * r) / r;
    const double ex = exp(-a * a * r2);
    double an = 1.0 / (SQRTPI * a);
    for (int j = 1; j < 6; j++) {
        an *= 2.0 * a * a;
        bn[j] = ((2 * j - 1) * bn[j - 1] + an * ex) / r2;
    }
}
/* Thole damping factors lambda_3,5,7,9 (index 1..4; index 0 unused = 1):
 * u = r/(pd_i pd_k), D = a u^3, lambda_3 = 1 - e^{-D}, and each higher
 * order makes d(lambda_n/r^n)/dr = -n lambda_{n+2}/r^{n+1} hold. */
static void thole_lambda(double r, double pdi, double pdk, double pgamma, double *lam)
{
    for (int j = 0; j < 6; j++) lam[j] = 1.0;
    const double damp = pdi * pdk;
    if (damp == 0.0) return;
    const double u = r / damp;
    const double D = pgamma * u * u * u;
    if (D >= 50.0) return;
    const double e = exp(-D);
    lam[1] = 1.0 - e;
    lam[2] = 1.0 - (1.0 + D) * e;
    lam[3] = 1.0 - (1.0 + D + 0.6 * D * D) * e;
    lam[4] = 1.0 - (1.0 + D + (18.0 / 35.0) * D * D + (9.0 / 35.0) * D * D * D) * e;
    lam[5] = 1.0; /* unused by dipole interactions */
}

/* the Thole factors of one pair (tests): lam[0..5] as thole_lambda */
void orc_thole_lambda(double r, double pdi, double pdk, double pgamma, double *lam)
{
    thole_lambda(r, pdi, pdk, pgamma, lam);
}

/* ---------------------------------------------------------------------
 * system, scaling
 * ------------------------------------------------------------------- */
typedef struct {
    int n;
    orc_elec_sys *s;
    const orc_mpole_sys *mp;
    double *rp;      /* 10n global multipoles */
    double f, a;     /* electric/dielec, Ewald coefficient (0 = none) */
    double cut2;
    int ewald;
    /* connectivity offsets */
    int *o12, *o13, *o14, *o15, *g11, *g12, *g13, *g14;
    /* charge penetration (HIPPO): core charges, valence exponents */
    int cp;
    const double *core, *alpha;
    /* PME */
    int K[3], order;
    double A[3][3];  /* u = A r */
    double *bsmod[3];
    double *theta;   /* n * 3 * order * 4 */
    int *igrid;      /* n * 3 first grid index */
} ctx_t;

static int *offsets(int n, const int *cnt)
{
    int *o = (int *)malloc((n + 1) * sizeof(int));
    o[0] = 0;
    for (int i = 0; i < n; i++) o[i + 1] = o[i] + (cnt ? cnt[i] : 0);
    return o;
}

/* scale factors of atom i against every k: m (mpole), p, d, u (polpot.h:47-66) */
static void set_scales(const ctx_t *c, int i, double *ms, double *ps, double *ds, double *us, double *ws, double v)
{
    const orc_mpole_sys *m = c->mp;
    const orc_elec_sys *s = c->s;
    if (v == 1.0) {
        /* reset */
        for (int q = 0; q < 4; q++) {
            const int *off = q == 0 ? c->o12 : q == 1 ? c->o13 : q == 2 ? c->o14 : c->o15;
            const int *lst = q == 0 ? m->i12 : q == 1 ? m->i13 : q == 2 ? m->i14 : m->i15;
            for (int j = off[i]; j < off[i + 1]; j++) { ms[lst[j]] = 1.0; ps[lst[j]] = 1.0; ws[lst[j]] = 1.0; }
        }
        for (int q = 0; q < 4; q++) {
            const int *off = q == 0 ? c->g11 : q == 1 ? c->g12 : q == 2 ? c->g13 : c->g14;
            const int *lst = q == 0 ? s->ip11 : q == 1 ? s->ip12 : q == 2 ? s->ip13 : s->ip14;
            if (!lst) continue;
            for (int j = off[i]; j < off[i + 1]; j++) { ds[lst[j]] = 1.0; us[lst[j]] = 1.0; }
        }
        return;
    }
    const double mv[4] = {m->m2scale, m->m3scale, m->m4scale, m->m5scale};
    const double pv[4] = {s->p2scale, s->p3scale, s->p4scale, s->p5scale};
    const double piv[4] = {s->p2iscale, s->p3iscale, s->p4iscale, s->p5iscale};
    const double dv[4] = {s->d1scale, s->d2scale, s->d3scale, s->d4scale};
    const double uv[4] = {s->u1scale, s->u2scale, s->u3scale, s->u4scale};
    const double wv[4] = {s->w2scale, s->w3scale, s->w4scale, s->w5scale};
    for (int q = 0; q < 4; q++) {
        const int *off = q == 0 ? c->o12 : q == 1 ? c->o13 : q == 2 ? c->o14 : c->o15;
        const int *lst = q == 0 ? m->i12 : q == 1 ? m->i13 : q == 2 ? m->i14 : m->i15;
        for (int j = off[i]; j < off[i + 1]; j++) {
            const int k = lst[j];
            ms[k] = mv[q];
            ps[k] = pv[q];
            ws[k] = wv[q];
            /* same polarization group: the "intra" value */
            if (s->ip11)
                for (int g = c->g11[i]; g < c->g11[i + 1]; g++)
                    if (s->ip11[g] == k) ps[k] = piv[q];
        }
    }
    for (int q = 0; q < 4; q++) {
        const int *off = q == 0 ? c->g11 : q == 1 ? c->g12 : q == 2 ? c->g13 : c->g14;
        const int *lst = q == 0 ? s->ip11 : q == 1 ? s->ip12 : q == 2 ? s->ip13 : s->ip14;
        if (!lst) continue;
        for (int j = off[i]; j < off[i + 1]; j++) { ds[lst[j]] = dv[q]; us[lst[j]] = uv[q]; }
    }
}

/* the pairs within the cutoff with their minimum-image separation */
typedef struct {
    int i, k;
    double r[3], r2;
    double m, p, d, u, w; /* scales */
} pairrec;

static pairrec *pair_list(ctx_t *c, long *np)
{
    const int n = c->n;
    const orc_mpole_sys *m = c->mp;
    double *ms = (double *)malloc(n * sizeof(double)), *ps = (double *)malloc(n * sizeof(double));
    double *ds = (double *)malloc(n * sizeof(double)), *us = (double *)malloc(n * sizeof(double));
    double *ws = (double *)malloc(n * sizeof(double));
    for (int i = 0; i < n; i++) ms[i] = ps[i] = ds[i] = us[i] = ws[i] = 1.0;
    long cap = 1024, cnt = 0;
    pairrec *L = (pairrec *)malloc(cap * sizeof(pairrec));
    for (int i = 0; i < n - 1; i++) {
        set_scales(c, i, ms, ps, ds, us, ws, 0.0);
        for (int k = i + 1; k < n; k++) {
            double r[3] = {m->x[k] - m->x[i], m->y[k] - m->y[i], m->z[k] - m->z[i]};
            if (m->box.kind != ORC_BOX_NONE) orc_image(&m->box, &r[0], &r[1], &r[2]);
            const double r2 = dot3(r, r);
            if (r2 > c->cut2) continue;
            if (cnt == cap) { cap *= 2; L = (pairrec *)realloc(L, cap * sizeof(pairrec)); }
            pairrec *q = &L[cnt++];
            q->i = i; q->k = k;
            memcpy(q->r, r, sizeof r);
            q->r2 = r2;
            q->m = ms[k]; q->p = ps[k]; q->d = c->s->dpequal ? ps[k] : ds[k]; q->u = us[k]; q->w = ws[k];
        }
        set_scales(c, i, ms, ps, ds, us, ws, 1.0);
    }
    free(ms); free(ps); free(ds); free(us); free(ws);
    *np = cnt;
    return L;
}

/* ---------------------------------------------------------------------
 * PME: cardinal B-splines, spreading, FFT (direct separable DFT), gather
 * ------------------------------------------------------------------- */
static double bspline_m(int n, double x)
{
    /* cardinal B-spline M_n(x), support [0, n] */
    if (x <= 0.0 || x >= n) return 0.0;
    if (n == 2) return 1.0 - fabs(x - 1.0);
    return (x * bspline_m(n - 1, x) + (n - x) * bspline_m(n - 1, x - 1.0)) / (n - 1);
}
/* d-th derivative of M_n at x: sum_j (-1)^j C(d,j) M_{n-d}(x - j) */
static double bspline_deriv(int n, int d, double x)
{
    static const int C[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 1, 0}, {1, 3, 3, 1}};
    double v = 0.0;
    for (int j = 0; j <= d; j++) v += ((j & 1) ? -1.0 : 1.0) * C[d][j] * bspline_m(n - d, x - j);
    return v;
}

static void pme_setup(ctx_t *c)
{
    orc_elec_sys *s = c->s;
    const orc_box *b = &c->mp->box;
    c->K[0] = s->nfft1; c->K[1] = s->nfft2; c->K[2] = s->nfft3;
    c->order = s->bsorder;
    /* lattice vectors (lattice.cpp conventions) and reciprocal rows */
    const double radian = 57.29577951308232088;
    double bs = 1, bc = 0, gs = 1, gc = 0, bt = 0, gt = 1;
    if (b->kind == ORC_BOX_TRI) {
        const double ac = cos(b->alpha / radian);
        bs = sin(b->beta / radian); bc = cos(b->beta / radian);
        gs = sin(b->gamma / radian); gc = cos(b->gamma / radian);
        bt = (ac - bc * gc) / gs;
        gt = sqrt(bs * bs - bt * bt);
    } else if (b->kind == ORC_BOX_MONO) {
        bs = sin(b->beta / radian); bc = cos(b->beta / radian);
        gt = bs;
    }
    const double av[3] = {b->xbox, 0, 0};
    const double bv[3] = {b->ybox * gc, b->ybox * gs, 0};
    const double cv[3] = {b->zbox * bc, b->zbox * bt, b->zbox * gt};
    double bxc[3], cxa[3], axb[3];
    cross3(bv, cv, bxc); cross3(cv, av, cxa); cross3(av, bv, axb);
    const double vol = dot3(av, bxc);
    for (int j = 0; j < 3; j++) {
        c->A[0][j] = c->K[0] * bxc[j] / vol;
        c->A[1][j] = c->K[1] * cxa[j] / vol;
        c->A[2][j] = c->K[2] * axb[j] / vol;
    }
    /* B-spline moduli |sum_k M_n(k+1) e^{2 pi i m k / K}|^2, zeros replaced
     * by the mean of their neighbours */
    for (int a = 0; a < 3; a++) {
        const int K = c->K[a];
        c->bsmod[a] = (double *)malloc(K * sizeof(double));
        for (int mm = 0; mm < K; mm++) {
            double re = 0, im = 0;
            for (int k = 0; k < c->order - 1; k++) {
                const double w = bspline_m(c->order, k + 1.0);
                const double arg = 2.0 * M_PI * mm * k / K;
                re += w * cos(arg);
                im += w * sin(arg);
            }
            c->bsmod[a][mm] = re * re + im * im;
        }
        for (int mm = 0; mm < K; mm++)
            if (c->bsmod[a][mm] < 1e-7)
                c->bsmod[a][mm] = 0.5 * (c->bsmod[a][(mm - 1 + K) % K] + c->bsmod[a][(mm + 1) % K]);
    }
    /* per-atom spline values and derivatives */
    const int n = c->n, ord = c->order;
    c->theta = (double *)malloc((size_t)n * 3 * ord * 4 * sizeof(double));
    c->igrid = (int *)malloc((size_t)n * 3 * sizeof(int));
    for (int i = 0; i < n; i++) {
        const double p[3] = {c->mp->x[i], c->mp->y[i], c->mp->z[i]};
        for (int a = 0; a < 3; a++) {
            const double u = dot3(c->A[a], p);
            const double fl = floor(u);
            const double w = u - fl;
            int i0 = (int)fl - ord + 1;
            i0 %= c->K[a];
            if (i0 < 0) i0 += c->K[a];
            c->igrid[3 * i + a] = i0;
            for (int k = 0; k < ord; k++)
                for (int d = 0; d < 4; d++)
                    c->theta[(((size_t)i * 3 + a) * ord + k) * 4 + d] = bspline_deriv(ord, d, w + ord - 1 - k);
        }
    }
}

static void pme_free(ctx_t *c)
{
    for (int a = 0; a < 3; a++) free(c->bsmod[a]);
    free(c->theta);
    free(c->igrid);
}

/* derivative operator indices of the 20 potential components, by
 * (d1,d2,d3) derivative orders along u1,u2,u3 */
static const int DER[20][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
                               {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0},
                               {2, 0, 1}, {1, 2, 0}, {1, 0, 2}, {0, 2, 1}, {0, 1, 2}, {1, 1, 1}};

/* spread fractional multipoles fm (10 per atom: c, d1..3, q11,q22,q33,
 * 2q12,2q13,2q23) onto the grid: rho(g) += sum_j fm_j D_j M(u - g) */
static void pme_spread(const ctx_t *c, const double *fm, int ncomp, double *grid)
{
    const int K0 = c->K[0], K1 = c->K[1], K2 = c->K[2], ord = c->order;
    memset(grid, 0, (size_t)K0 * K1 * K2 * sizeof(double));
    for (int i = 0; i < c->n; i++) {
        const double *f = fm + 10 * i;
        const double *t0 = c->theta + (((size_t)i * 3 + 0) * ord) * 4;
        const double *t1 = c->theta + (((size_t)i * 3 + 1) * ord) * 4;
        const double *t2 = c->theta + (((size_t)i * 3 + 2) * ord) * 4;
        for (int kz = 0; kz < ord; kz++) {
            const int gz = (c->igrid[3 * i + 2] + kz) % K2;
            for (int ky = 0; ky < ord; ky++) {
                const int gy = (c->igrid[3 * i + 1] + ky) % K1;
                for (int kx = 0; kx < ord; kx++) {
                    const int gx = (c->igrid[3 * i] + kx) % K0;
                    double v = 0.0;
                    for (int j = 0; j < ncomp; j++)
                        v += f[j] * t0[kx * 4 + DER[j][0]] * t1[ky * 4 + DER[j][1]] * t2[kz * 4 + DER[j][2]];
                    grid[((size_t)gz * K1 + gy) * K0 + gx] += v;
                }
            }
        }
    }
}

/* fphi[j] = sum_g phi(g) D_j M(u - g), j < ncomp (<= 20) */
static void pme_gather(const ctx_t *c, const double *grid, int ncomp, double *fphi)
{
    const int K0 = c->K[0], K1 = c->K[1], ord = c->order;
    for (int i = 0; i < c->n; i++) {
        double *o = fphi + 20 * i;
        for (int j = 0; j < 20; j++) o[j] = 0.0;
        const double *t0 = c->theta + (((size_t)i * 3 + 0) * ord) * 4;
        const double *t1 = c->theta + (((size_t)i * 3 + 1) * ord) * 4;
        const double *t2 = c->theta + (((size_t)i * 3 + 2) * ord) * 4;
        for (int kz = 0; kz < ord; kz++) {
            const int gz = (c->igrid[3 * i + 2] + kz) % c->K[2];
            for (int ky = 0; ky < ord; ky++) {
                const int gy = (c->igrid[3 * i + 1] + ky) % K1;
                for (int kx = 0; kx < ord; kx++) {
                    const int gx = (c->igrid[3 * i] + kx) % K0;
                    const double g = grid[((size_t)gz * K1 + gy) * K0 + gx];
                    for (int j = 0; j < ncomp; j++)
                        o[j] += g * t0[kx * 4 + DER[j][0]] * t1[ky * 4 + DER[j][1]] * t2[kz * 4 + DER[j][2]];
                }
            }
        }
    }
}

/* one separable DFT pass along axis a of a complex grid (sign = +1/-1) */
static void dft_axis(const ctx_t *c, double *re, double *im, int a, int sign)
{
    const int K[3] = {c->K[0], c->K[1], c->K[2]};
    const int L = K[a];
    double *cs = (double *)malloc(L * sizeof(double)), *sn = (double *)malloc(L * sizeof(double));
    for (int j = 0; j < L; j++) { cs[j] = cos(2.0 * M_PI * j / L); sn[j] = sign * sin(2.0 * M_PI * j / L); }
    double *tr = (double *)malloc(L * sizeof(double)), *ti = (double *)malloc(L * sizeof(double));
    const size_t stride = a == 0 ? 1 : a == 1 ? (size_t)K[0] : (size_t)K[0] * K[1];
    const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
    for (int p = 0; p < K[o1]; p++)
        for (int q = 0; q < K[o2]; q++) {
            int idx[3] = {0, 0, 0};
            idx[o1] = p; idx[o2] = q;
            const size_t base = ((size_t)idx[2] * K[1] + idx[1]) * K[0] + idx[0];
            for (int m = 0; m < L; m++) {
                double sr = 0, si = 0;
                for (int j = 0; j < L; j++) {
                    const int w = (int)(((long)m * j) % L);
                    const double xr = re[base + j * stride], xi = im[base + j * stride];
                    sr += xr * cs[w] - xi * sn[w];
                    si += xr * sn[w] + xi * cs[w];
                }
                tr[m] = sr; ti[m] = si;
            }
            for (int m = 0; m < L; m++) { re[base + m * stride] = tr[m]; im[base + m * stride] = ti[m]; }
        }
    free(cs); free(sn); free(tr); free(ti);
}

/* the reciprocal-space kernel G(m) = exp(-pi^2 h^2/a^2) / (pi V h^2 prod bsmod) */
static double pme_volume(const ctx_t *c)
{
    /* V = K0 K1 K2 / det(A) */
    const double (*A)[3] = c->A;
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
                     + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    return (double)c->K[0] * c->K[1] * c->K[2] / det;
}

/* a reciprocal potential and its derivatives at one site */
typedef struct {
    double phi, g[3], H[3][3], T[3][3][3];
} sitepot;

/* symmetrized pair virial of a pair gradient g (on the first atom) at
 * separation r = x_k - x_i, times s (empole.cpp:310-320, pairMpole.h:166-173) */
static void pair_vir(double *v, const double *r, const double *g, double s)
{
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) v[3 * a + b] -= 0.5 * s * (r[a] * g[b] + r[b] * g[a]);
}

/* the strain term of the dipole and quadrupole of multipole mA (10, the
 * reference's traceless x1/3 quadrupole) in a reciprocal potential pB,
 * times s: -s sum [dA_b phi_a + 2 q_bc phi_ac], symmetrized */
static void cterm_vir(double *v, const double *mA, const sitepot *pB, double s)
{
    const double Q[3][3] = {{mA[4], mA[5], mA[6]}, {mA[5], mA[7], mA[8]}, {mA[6], mA[8], mA[9]}};
    double W[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double w = mA[1 + b] * pB->g[a];
            for (int c3 = 0; c3 < 3; c3++) w += 2.0 * Q[b][c3] * pB->H[a][c3];
            W[a][b] = w;
        }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) v[3 * a + b] -= 0.5 * s * (W[a][b] + W[b][a]);
}

/* phi = IDFT[G DFT[rho]] in place; returns sum_m G |rho^(m)|^2 and, when
 * vk is given, adds sum_m G |rho|^2 [2 h_a h_b (pi^2/a^2 + 1/h^2) - delta_ab]
 * to vk (9) */
static double pme_convolve(const ctx_t *c, double *grid, double *vk)
{
    const int K0 = c->K[0], K1 = c->K[1], K2 = c->K[2];
    const size_t N = (size_t)K0 * K1 * K2;
    double *im = (double *)calloc(N, sizeof(double));
    for (int a = 0; a < 3; a++) dft_axis(c, grid, im, a, +1);
    const double V = pme_volume(c);
    double e = 0.0;
    for (int m3 = 0; m3 < K2; m3++)
        for (int m2 = 0; m2 < K1; m2++)
            for (int m1 = 0; m1 < K0; m1++) {
                const size_t id = ((size_t)m3 * K1 + m2) * K0 + m1;
                if (m1 == 0 && m2 == 0 && m3 == 0) { grid[id] = 0; im[id] = 0; continue; }
                const int f1 = m1 > K0 / 2 ? m1 - K0 : m1, f2 = m2 > K1 / 2 ? m2 - K1 : m2, f3 = m3 > K2 / 2 ? m3 - K2 : m3;
                /* h = sum_a m_a (A row a)/K_a */
                double h[3];
                for (int j = 0; j < 3; j++)
                    h[j] = f1 * c->A[0][j] / K0 + f2 * c->A[1][j] / K1 + f3 * c->A[2][j] / K2;
                const double h2 = dot3(h, h);
                const double G = exp(-M_PI * M_PI * h2 / (c->a * c->a)) /
                                 (M_PI * V * h2 * c->bsmod[0][m1] * c->bsmod[1][m2] * c->bsmod[2][m3]);
                const double t = G * (grid[id] * grid[id] + im[id] * im[id]);
                e += t;
                if (vk) {
                    const double w = 2.0 * (M_PI * M_PI / (c->a * c->a) + 1.0 / h2);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) vk[3 * a + b] += t * (w * h[a] * h[b] - (a == b ? 1.0 : 0.0));
                }
                grid[id] *= G;
                im[id] *= G;
            }
    for (int a = 0; a < 3; a++) dft_axis(c, grid, im, a, -1);
    free(im);
    return e;
}

/* Cartesian multipole (10) -> fractional: d^f = A d, q^f = A q A^T, with
 * off-diagonal q^f doubled */
static void cart_to_frac(const ctx_t *c, const double *m, double *fm)
{
    const double (*A)[3] = c->A;
    const double Q[3][3] = {{m[4], m[5], m[6]}, {m[5], m[7], m[8]}, {m[6], m[8], m[9]}};
    fm[0] = m[0];
    for (int a = 0; a < 3; a++) fm[1 + a] = A[a][0] * m[1] + A[a][1] * m[2] + A[a][2] * m[3];
    double F[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double v = 0;
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++) v += A[a][p] * Q[p][q] * A[b][q];
            F[a][b] = v;
        }
    fm[4] = F[0][0]; fm[5] = F[1][1]; fm[6] = F[2][2];
    fm[7] = 2 * F[0][1]; fm[8] = 2 * F[0][2]; fm[9] = 2 * F[1][2];
}

/* fractional derivatives (20) -> Cartesian potential phi, gradient g[3],
 * Hessian H[3][3], third derivatives T[3][3][3] */
static void frac_to_cart(const ctx_t *c, const double *fp, double *phi, double *g, double H[3][3], double T[3][3][3])
{
    const double (*A)[3] = c->A;
    double F1[3], F2[3][3], F3[3][3][3];
    F1[0] = fp[1]; F1[1] = fp[2]; F1[2] = fp[3];
    F2[0][0] = fp[4]; F2[1][1] = fp[5]; F2[2][2] = fp[6];
    F2[0][1] = F2[1][0] = fp[7]; F2[0][2] = F2[2][0] = fp[8]; F2[1][2] = F2[2][1] = fp[9];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int d = 0; d < 3; d++) {
                int cnt[3] = {0, 0, 0};
                cnt[a]++; cnt[b]++; cnt[d]++;
                int idx = -1;
                for (int j = 10; j < 20; j++)
                    if (DER[j][0] == cnt[0] && DER[j][1] == cnt[1] && DER[j][2] == cnt[2]) idx = j;
                F3[a][b][d] = fp[idx];
            }
    *phi = fp[0];
    for (int x = 0; x < 3; x++) {
        g[x] = 0;
        for (int a = 0; a < 3; a++) g[x] += A[a][x] * F1[a];
    }
    if (H)
        for (int x = 0; x < 3; x++)
            for (int y = 0; y < 3; y++) {
                double v = 0;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) v += A[a][x] * A[b][y] * F2[a][b];
                H[x][y] = v;
            }
    if (T)
        for (int x = 0; x < 3; x++)
            for (int y = 0; y < 3; y++)
                for (int z = 0; z < 3; z++) {
                    double v = 0;
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++)
                            for (int d = 0; d < 3; d++) v += A[a][x] * A[b][y] * A[d][z] * F3[a][b][d];
                    T[x][y][z] = v;
                }
}

/* potential derivatives at every site from a set of Cartesian multipoles
 * (10 per atom); returns the reciprocal energy sum G|rho|^2 / 2 (no f) */

static double pme_potential(const ctx_t *c, const double *mpoles, sitepot *out, double *vk)
{
    const int n = c->n;
    const size_t N = (size_t)c->K[0] * c->K[1] * c->K[2];
    double *fm = (double *)malloc(10 * (size_t)n * sizeof(double));
    double *fp = (double *)malloc(20 * (size_t)n * sizeof(double));
    double *grid = (double *)malloc(N * sizeof(double));
    for (int i = 0; i < n; i++) cart_to_frac(c, mpoles + 10 * i, fm + 10 * i);
    pme_spread(c, fm, 10, grid);
    const double e = 0.5 * pme_convolve(c, grid, vk);
    if (out) {
        pme_gather(c, grid, 20, fp);
        for (int i = 0; i < n; i++) frac_to_cart(c, fp + 20 * i, &out[i].phi, out[i].g, out[i].H, out[i].T);
    }
    free(fm); free(fp); free(grid);
    return e;
}

/* energy, gradient and torque of multipole m (10) in a potential p:
 * U = c phi + d.g + q:H */
static double site_energy(const double *m, const sitepot *p, double *grad, double *trq)
{
    const double Q[3][3] = {{m[4], m[5], m[6]}, {m[5], m[7], m[8]}, {m[6], m[8], m[9]}};
    double U = m[0] * p->phi;
    for (int a = 0; a < 3; a++) U += m[1 + a] * p->g[a];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) U += Q[a][b] * p->H[a][b];
    if (grad)
        for (int x = 0; x < 3; x++) {
            double v = m[0] * p->g[x];
            for (int a = 0; a < 3; a++) v += m[1 + a] * p->H[a][x];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) v += Q[a][b] * p->T[a][b][x];
            grad[x] = v;
        }
    if (trq) {
        /* tau = d x E - 2 eps:(q H), E = -g */
        double E[3] = {-p->g[0], -p->g[1], -p->g[2]};
        cross3(m + 1, E, trq);
        double QH[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) QH[a][b] = Q[a][0] * p->H[0][b] + Q[a][1] * p->H[1][b] + Q[a][2] * p->H[2][b];
        trq[0] -= 2.0 * (QH[1][2] - QH[2][1]);
        trq[1] -= 2.0 * (QH[2][0] - QH[0][2]);
        trq[2] -= 2.0 * (QH[0][1] - QH[1][0]);
    }
    return U;
}

/* ---------------------------------------------------------------------
 * fields
 * ------------------------------------------------------------------- */
/* field at i from multipole mk (and at k from mi): E_i = -dU/dmu_i */
static void pair_fields(const double *R, const double *mi, const double *mk, const double *r, double *Ei, double *Ek)
{
    double qirv[3], qkrv[3];
    qmul(mi + 4, r, qirv);
    qmul(mk + 4, r, qkrv);
    const double dir = dot3(mi + 1, r), dkr = dot3(mk + 1, r), qir = dot3(qirv, r), qkr = dot3(qkrv, r);
    const double si = mk[0] * R[1] - dkr * R[2] + qkr * R[3];
    const double sk = mi[0] * R[1] + dir * R[2] + qir * R[3];
    for (int a = 0; a < 3; a++) {
        Ei[a] = -si * r[a] - R[1] * mk[1 + a] + 2.0 * R[2] * qkrv[a];
        Ek[a] = sk * r[a] - R[1] * mi[1 + a] - 2.0 * R[2] * qirv[a];
    }
}

/* damped series: Ewald real space bn - (1 - s dmp) rr, or s dmp rr */
static void damped_series(const ctx_t *c, const pairrec *q, double scale, const double *dmp, int nd, double *R)
{
    double rr[6], bn[6];
    series_rr(q->r2, rr);
    if (c->ewald) series_ewald(q->r2, c->a, bn);
    for (int j = 0; j < 6; j++) {
        const double d = j < nd ? dmp[j] : 1.0;
        R[j] = c->ewald ? bn[j] - (1.0 - scale * d) * rr[j] : scale * d * rr[j];
    }
}

/* series for a polarization interaction: Ewald + scaled, Thole-damped; with
 * charge penetration the overlap damping of the two valence densities
 * (the dipole-dipole damping of HIPPO) */
static void pol_series(const ctx_t *c, const pairrec *q, double scale, double *R)
{
    const orc_elec_sys *s = c->s;
    double rr[6], lam[6];
    if (c->cp) {
        double dmpi[5], dmpk[5], dmpik[6];
        orc_damp_gordon1(sqrt(q->r2), c->alpha[q->i], c->alpha[q->k], dmpi, dmpk, dmpik);
        damped_series(c, q, scale, dmpik, 6, R);
        return;
    }
    series_rr(q->r2, rr);
    const double pgamma = fmin(s->thole[q->i], s->thole[q->k]) != 0.0 ? fmin(s->thole[q->i], s->thole[q->k])
                                                                    : fmax(s->thole[q->i], s->thole[q->k]);
    thole_lambda(sqrt(q->r2), s->pdamp[q->i], s->pdamp[q->k], pgamma, lam);
    if (c->ewald) {
        double bn[6];
        series_ewald(q->r2, c->a, bn);
        for (int j = 0; j < 6; j++) R[j] = bn[j] - (1.0 - scale * lam[j]) * rr[j];
    } else {
        for (int j = 0; j < 6; j++) R[j] = scale * lam[j] * rr[j];
    }
}

/* permanent fields E^d and E^p (3n each), field units (no f) */
static void direct_field(ctx_t *c, const pairrec *L, long np, const sitepot *prec, double *Ed, double *Ep)
{
    const int n = c->n;
    memset(Ed, 0, 3 * (size_t)n * sizeof(double));
    memset(Ep, 0, 3 * (size_t)n * sizeof(double));
    for (long t = 0; t < np; t++) {
        const pairrec *q = &L[t];
        double R[6], Ei[3], Ek[3];
        if (c->cp) {
            /* field at i: core of k (point) + valence multipole of k damped by
             * its one-site function (and the reverse at k) */
            double dmpi[5], dmpk[5], dmpik[6], one[1] = {1.0};
            orc_damp_gordon1(sqrt(q->r2), c->alpha[q->i], c->alpha[q->k], dmpi, dmpk, dmpik);
            double ci[10] = {0}, ck[10] = {0}, vi[10], vk[10], z[10] = {0};
            memcpy(vi, c->rp + 10 * q->i, sizeof vi);
            memcpy(vk, c->rp + 10 * q->k, sizeof vk);
            ci[0] = c->core[q->i]; ck[0] = c->core[q->k];
            vi[0] -= ci[0]; vk[0] -= ck[0];
            for (int set = 0; set < 2; set++) {
                double *E = set ? Ep : Ed;
                const double sc = set ? q->p : q->d;
                double Ei2[3], Ek2[3];
                damped_series(c, q, sc, one, 1, R);
                pair_fields(R, ci, ck, q->r, Ei, Ek);
                damped_series(c, q, sc, dmpk, 5, R);
                pair_fields(R, z, vk, q->r, Ei2, Ek2);
                for (int a = 0; a < 3; a++) Ei[a] += Ei2[a];
                damped_series(c, q, sc, dmpi, 5, R);
                pair_fields(R, vi, z, q->r, Ei2, Ek2);
                for (int a = 0; a < 3; a++) Ek[a] += Ek2[a];
                for (int a = 0; a < 3; a++) { E[3 * q->i + a] += Ei[a]; E[3 * q->k + a] += Ek[a]; }
            }
            continue;
        }
        pol_series(c, q, q->d, R);
        pair_fields(R, c->rp + 10 * q->i, c->rp + 10 * q->k, q->r, Ei, Ek);
        for (int a = 0; a < 3; a++) { Ed[3 * q->i + a] += Ei[a]; Ed[3 * q->k + a] += Ek[a]; }
        pol_series(c, q, q->p, R);
        pair_fields(R, c->rp + 10 * q->i, c->rp + 10 * q->k, q->r, Ei, Ek);
        for (int a = 0; a < 3; a++) { Ep[3 * q->i + a] += Ei[a]; Ep[3 * q->k + a] += Ek[a]; }
    }
    if (c->ewald) {
        const double self = 4.0 / 3.0 * c->a * c->a * c->a / SQRTPI;
        for (int i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                const double v = -prec[i].g[a] + self * c->rp[10 * i + 1 + a];
                Ed[3 * i + a] += v;
                Ep[3 * i + a] += v;
            }
    }
}

/* mutual field T mu for two dipole sets (3n each) */
static void mutual_field(ctx_t *c, const pairrec *L, long np, const double *ud, const double *up, double *Fd, double *Fp)
{
    const int n = c->n;
    memset(Fd, 0, 3 * (size_t)n * sizeof(double));
    memset(Fp, 0, 3 * (size_t)n * sizeof(double));
    for (long t = 0; t < np; t++) {
        const pairrec *q = &L[t];
        double R[6];
        pol_series(c, q, c->cp ? q->w : q->u, R);
        for (int set = 0; set < 2; set++) {
            const double *u = set ? up : ud;
            double *F = set ? Fp : Fd;
            const double *ui = u + 3 * q->i, *uk = u + 3 * q->k;
            const double uir = dot3(ui, q->r), ukr = dot3(uk, q->r);
            for (int a = 0; a < 3; a++) {
                F[3 * q->i + a] += ukr * R[2] * q->r[a] - R[1] * uk[a];
                F[3 * q->k + a] += uir * R[2] * q->r[a] - R[1] * ui[a];
            }
        }
    }
    if (c->ewald) {
        double *m = (double *)calloc(10 * (size_t)n, sizeof(double));
        sitepot *p = (sitepot *)malloc(n * sizeof(sitepot));
        const double self = 4.0 / 3.0 * c->a * c->a * c->a / SQRTPI;
        for (int set = 0; set < 2; set++) {
            const double *u = set ? up : ud;
            double *F = set ? Fp : Fd;
            for (int i = 0; i < n; i++)
                for (int a = 0; a < 3; a++) m[10 * i + 1 + a] = u[3 * i + a];
            pme_potential(c, m, p, NULL);
            for (int i = 0; i < n; i++)
                for (int a = 0; a < 3; a++) F[3 * i + a] += -p[i].g[a] + self * u[3 * i + a];
        }
        free(m);
        free(p);
    }
}

/* ---------------------------------------------------------------------
 * induced dipoles
 * ------------------------------------------------------------------- */
#define DEBYE 4.80321

/* the preconditioner z = M r for both sets: udiag a r + a T_s (a r) with T_s
 * the real-space dipole field (no Ewald) of the pairs within usolvcut,
 * damped and scaled like the mutual field; without pcgprec z = a r */
static void precondition(ctx_t *c, const pairrec *L, long np, const double *poli, const double *rd,
                         const double *rp, double *zd, double *zp)
{
    const orc_elec_sys *s = c->s;
    const int n = c->n, n3 = 3 * n;
    if (!s->pcgprec) {
        for (int j = 0; j < n3; j++) { zd[j] = poli[j / 3] * rd[j]; zp[j] = poli[j / 3] * rp[j]; }
        return;
    }
    double *wd = malloc(n3 * sizeof(double)), *wp = malloc(n3 * sizeof(double));
    double *fd = calloc(n3, sizeof(double)), *fp = calloc(n3, sizeof(double));
    for (int j = 0; j < n3; j++) { wd[j] = poli[j / 3] * rd[j]; wp[j] = poli[j / 3] * rp[j]; }
    const double us2 = s->usolvcut * s->usolvcut;
    const int ew = c->ewald;
    c->ewald = 0; /* the series without Ewald */
    for (long t = 0; t < np; t++) {
        const pairrec *q = &L[t];
        if (q->r2 > us2) continue;
        double R[6];
        pol_series(c, q, c->cp ? q->w : q->u, R);
        for (int set = 0; set < 2; set++) {
            const double *u = set ? wp : wd;
            double *F = set ? fp : fd;
            const double *ui = u + 3 * q->i, *uk = u + 3 * q->k;
            const double uir = dot3(ui, q->r), ukr = dot3(uk, q->r);
            for (int a = 0; a < 3; a++) {
                F[3 * q->i + a] += ukr * R[2] * q->r[a] - R[1] * uk[a];
                F[3 * q->k + a] += uir * R[2] * q->r[a] - R[1] * ui[a];
            }
        }
    }
    c->ewald = ew;
    for (int j = 0; j < n3; j++) {
        zd[j] = s->udiag * wd[j] + poli[j / 3] * fd[j];
        zp[j] = s->udiag * wp[j] + poli[j / 3] * fp[j];
    }
    free(wd); free(wp); free(fd); free(fp);
}

/* preconditioned conjugate gradient on (1/alpha - T) mu = E for both sets,
 * stop when the rms residual (Debye) < poleps.  First guess: the direct
 * induced dipoles a E; with the sparse preconditioner (pcgprec) those
 * dipoles' near field added once, mu0 = a (E + T_s a E) with T_s the
 * preconditioner's operator (pairs within usolvcut, no Ewald) -- one
 * Jacobi step on the near field, which on the 12,288-atom water box saves
 * one iteration (8 -> 7 to 1e-6 D); the GPU solver takes the same guess */
static int solve_pcg(ctx_t *c, const pairrec *L, long np, const double *Ed, const double *Ep, double *ud, double *up)
{
    const int n = c->n, n3 = 3 * n;
    const orc_elec_sys *s = c->s;
    double *poli = (double *)malloc(n * sizeof(double));
    int npolar = 0;
    for (int i = 0; i < n; i++) {
        poli[i] = fmax(1e-8, s->polarity[i]);
        if (s->polarity[i] != 0.0) npolar++;
    }
    double *rd = malloc(n3 * sizeof(double)), *rp_ = malloc(n3 * sizeof(double));
    double *zd = malloc(n3 * sizeof(double)), *zp = malloc(n3 * sizeof(double));
    double *pd = malloc(n3 * sizeof(double)), *pp = malloc(n3 * sizeof(double));
    double *Fd = malloc(n3 * sizeof(double)), *Fp = malloc(n3 * sizeof(double));
    /* guess: direct induced dipoles */
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            ud[3 * i + a] = s->polarity[i] * Ed[3 * i + a];
            up[3 * i + a] = s->polarity[i] * Ep[3 * i + a];
        }
    if (s->pcgprec) {
        /* ... plus their near field: the preconditioner applied to E with
         * a unit diagonal, a E + a T_s (a E) */
        orc_elec_sys *sw = (orc_elec_sys *)s;
        const double udiag = sw->udiag;
        sw->udiag = 1.0;
        precondition(c, L, np, poli, Ed, Ep, ud, up);
        sw->udiag = udiag;
        for (int j = 0; j < n3; j++)
            if (s->polarity[j / 3] == 0.0) ud[j] = up[j] = 0.0;
    }
    mutual_field(c, L, np, ud, up, Fd, Fp);
    for (int j = 0; j < n3; j++) {
        const int i = j / 3;
        const double pol = s->polarity[i];
        rd[j] = pol != 0.0 ? Ed[j] + Fd[j] - ud[j] / poli[i] : 0.0;
        rp_[j] = pol != 0.0 ? Ep[j] + Fp[j] - up[j] / poli[i] : 0.0;
    }
    precondition(c, L, np, poli, rd, rp_, zd, zp);
    for (int j = 0; j < n3; j++) { pd[j] = zd[j]; pp[j] = zp[j]; }
    double sd = 0, sp = 0;
    for (int j = 0; j < n3; j++) { sd += rd[j] * zd[j]; sp += rp_[j] * zp[j]; }
    int iter = 0, done = 0;
    while (!done && iter < s->politer) {
        iter++;
        mutual_field(c, L, np, pd, pp, Fd, Fp);
        double ad = 0, ap = 0;
        for (int j = 0; j < n3; j++) {
            const int i = j / 3;
            const double vd = s->polarity[i] != 0.0 ? pd[j] / poli[i] - Fd[j] : 0.0;
            const double vp = s->polarity[i] != 0.0 ? pp[j] / poli[i] - Fp[j] : 0.0;
            Fd[j] = vd; Fp[j] = vp;
            ad += pd[j] * vd;
            ap += pp[j] * vp;
        }
        ad = ad != 0.0 ? sd / ad : 0.0;
        ap = ap != 0.0 ? sp / ap : 0.0;
        double sd2 = 0, sp2 = 0, epsd = 0, epsp = 0;
        for (int j = 0; j < n3; j++) {
            ud[j] += ad * pd[j];
            up[j] += ap * pp[j];
            rd[j] -= ad * Fd[j];
            rp_[j] -= ap * Fp[j];
        }
        precondition(c, L, np, poli, rd, rp_, zd, zp);
        for (int j = 0; j < n3; j++) {
            sd2 += rd[j] * zd[j];
            sp2 += rp_[j] * zp[j];
            epsd += rd[j] * rd[j];
            epsp += rp_[j] * rp_[j];
        }
        const double bd = sd != 0.0 ? sd2 / sd : 0.0, bp = sp != 0.0 ? sp2 / sp : 0.0;
        for (int j = 0; j < n3; j++) { pd[j] = zd[j] + bd * pd[j]; pp[j] = zp[j] + bp * pp[j]; }
        sd = sd2;
        sp = sp2;
        const double eps = DEBYE * sqrt(fmax(epsd, epsp) / (npolar > 0 ? npolar : 1));
        if (eps < s->poleps) done = 1;
    }
    free(poli); free(rd); free(rp_); free(zd); free(zp); free(pd); free(pp); free(Fd); free(Fp);
    return done ? iter : -iter;
}

/* dense solve of (1/alpha - T) mu = E (small systems; tests) */
static int solve_dense(ctx_t *c, const pairrec *L, long np, const double *Ed, const double *Ep, double *ud, double *up)
{
    const int n = c->n, n3 = 3 * n;
    const orc_elec_sys *s = c->s;
    double *M = (double *)malloc((size_t)n3 * n3 * sizeof(double));
    double *e = calloc(n3, sizeof(double)), *Fd = malloc(n3 * sizeof(double)), *Fp = malloc(n3 * sizeof(double));
    double *z = calloc(n3, sizeof(double));
    /* columns of T by applying the field operator to unit vectors */
    for (int col = 0; col < n3; col++) {
        e[col] = 1.0;
        mutual_field(c, L, np, e, z, Fd, Fp);
        e[col] = 0.0;
        for (int row = 0; row < n3; row++) M[(size_t)row * n3 + col] = -Fd[row];
    }
    for (int j = 0; j < n3; j++) {
        const double pol = s->polarity[j / 3];
        M[(size_t)j * n3 + j] += pol != 0.0 ? 1.0 / pol : 1.0;
    }
    /* Gaussian elimination with partial pivoting, two right-hand sides */
    double *B = (double *)malloc((size_t)n3 * 2 * sizeof(double));
    for (int j = 0; j < n3; j++) {
        const double pol = s->polarity[j / 3];
        B[2 * j] = pol != 0.0 ? Ed[j] : 0.0;
        B[2 * j + 1] = pol != 0.0 ? Ep[j] : 0.0;
    }
    for (int k = 0; k < n3; k++) {
        int piv = k;
        for (int r = k + 1; r < n3; r++)
            if (fabs(M[(size_t)r * n3 + k]) > fabs(M[(size_t)piv * n3 + k])) piv = r;
        if (piv != k) {
            for (int q = 0; q < n3; q++) { double t = M[(size_t)k * n3 + q]; M[(size_t)k * n3 + q] = M[(size_t)piv * n3 + q]; M[(size_t)piv * n3 + q] = t; }
            for (int q = 0; q < 2; q++) { double t = B[2 * k + q]; B[2 * k + q] = B[2 * piv + q]; B[2 * piv + q] = t; }
        }
        for (int r = k + 1; r < n3; r++) {
            const double fct = M[(size_t)r * n3 + k] / M[(size_t)k * n3 + k];
            if (fct == 0.0) continue;
            for (int q = k; q < n3; q++) M[(size_t)r * n3 + q] -= fct * M[(size_t)k * n3 + q];
            B[2 * r] -= fct * B[2 * k];
            B[2 * r + 1] -= fct * B[2 * k + 1];
        }
    }
    for (int k = n3 - 1; k >= 0; k--) {
        for (int q = 0; q < 2; q++) {
            double v = B[2 * k + q];
            for (int col = k + 1; col < n3; col++) v -= M[(size_t)k * n3 + col] * B[2 * col + q];
            B[2 * k + q] = v / M[(size_t)k * n3 + k];
        }
    }
    for (int j = 0; j < n3; j++) {
        const double pol = s->polarity[j / 3];
        ud[j] = pol != 0.0 ? B[2 * j] : 0.0;
        up[j] = pol != 0.0 ? B[2 * j + 1] : 0.0;
    }
    free(M); free(e); free(Fd); free(Fp); free(z); free(B);
    return 1;
}

/* ---------------------------------------------------------------------
 * driver
 * ------------------------------------------------------------------- */

int orc_elec(orc_elec_sys *s, int want_grad, orc_elec_out *o)
{
    orc_mpole_sys *mp = s->mp;
    const int n = mp->n;
    ctx_t c;
    memset(&c, 0, sizeof c);
    c.n = n;
    c.s = s;
    c.mp = mp;
    c.f = mp->f;
    c.ewald = s->use_ewald;
    c.a = s->use_ewald ? s->aewald : 0.0;
    c.cut2 = s->cut * s->cut;
    c.cp = mp->pentype != 0;
    if (c.cp && mp->pentype != 1) return -3; /* Gordon1 (HIPPO) only */
    c.core = mp->pcore;
    c.alpha = mp->palpha;
    if (c.ewald && !(mp->box.kind == ORC_BOX_ORTHO || mp->box.kind == ORC_BOX_MONO || mp->box.kind == ORC_BOX_TRI))
        return -2;
    c.o12 = offsets(n, mp->n12); c.o13 = offsets(n, mp->n13);
    c.o14 = offsets(n, mp->n14); c.o15 = offsets(n, mp->n15);
    c.g11 = offsets(n, s->np11); c.g12 = offsets(n, s->np12);
    c.g13 = offsets(n, s->np13); c.g14 = offsets(n, s->np14);

    /* global multipoles; charge flux alters the monopoles (and so the
     * valence charges: total minus core) before the frames, energy.cpp:241 */
    double *pole = (double *)malloc(13 * (size_t)n * sizeof(double));
    double *rp13 = (double *)malloc(13 * (size_t)n * sizeof(double));
    memcpy(pole, mp->pole, 13 * (size_t)n * sizeof(double));
    if (mp->use_chgflx) {
        double *pdelta = (double *)malloc((size_t)n * sizeof(double));
        orc_charge_flux_delta(mp, pdelta);
        for (int i = 0; i < n; i++) pole[13 * i] = mp->mono0[i] + pdelta[i];
        free(pdelta);
    }
    /* charge-flux gradient (with charge penetration, as empole.cpp:66-80
     * selects it): pot_m[i] = dE_m/dq_i, pot_p[i] = dE_p/dq_i, each turned
     * into forces through the monopole changes */
    const int cf = c.cp && mp->use_chgflx && want_grad;
    double *pot_m = cf ? (double *)calloc((size_t)n, sizeof(double)) : NULL;
    double *pot_p = cf ? (double *)calloc((size_t)n, sizeof(double)) : NULL;
    static const double e0[10] = {1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    orc_chkpole(mp, pole);
    orc_rotpole(mp, pole, rp13);
    c.rp = (double *)malloc(10 * (size_t)n * sizeof(double));
    static const int idx[10] = {0, 1, 2, 3, 4, 5, 6, 8, 9, 12};
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 10; j++) c.rp[10 * i + j] = rp13[13 * i + idx[j]];
    if (c.ewald) pme_setup(&c);

    long np = 0;
    pairrec *L = pair_list(&c, &np);
    double *te = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *dem = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *dep = (double *)calloc(3 * (size_t)n, sizeof(double));
    const double f = c.f;
    double vm[9] = {0}, vp[9] = {0}, vkm[9] = {0};
    const int want_vir = want_grad && (o->vir_m || o->vir_p);

    /* ---- permanent multipoles: real space */
    double em_real = 0.0;
    /* analysis: pair energies halved onto the atoms, counted when nonzero;
     * the same-molecule pairs' unscreened energies (Ewald) or the
     * intermolecular pair energies (without) */
    double *aem = o->aem;
    if (aem) memset(aem, 0, n * sizeof(double));
    long nem = 0;
    double eintra = 0.0, einter = 0.0;
    ctx_t c0 = c; /* the same pairs without Ewald */
    c0.ewald = 0;
    for (long t = 0; t < np; t++) {
        const pairrec *q = &L[t];
        double R[6], rr[6];
        const int same = mp->molcule[q->i] == mp->molcule[q->k];
        if (c.cp) {
            /* core-core (undamped), core k with valence i (dmpi), core i with
             * valence k (dmpk), valence-valence (dmpik) */
            double dmpi[5], dmpk[5], dmpik[6], one[1] = {1.0};
            orc_damp_gordon1(sqrt(q->r2), c.alpha[q->i], c.alpha[q->k], dmpi, dmpk, dmpik);
            double ci[10] = {0}, ck[10] = {0}, vi[10], vk[10];
            memcpy(vi, c.rp + 10 * q->i, sizeof vi);
            memcpy(vk, c.rp + 10 * q->k, sizeof vk);
            ci[0] = c.core[q->i]; ck[0] = c.core[q->k];
            vi[0] -= ci[0]; vk[0] -= ck[0];
            const double *dm[4] = {one, dmpi, dmpk, dmpik};
            const int nd[4] = {1, 5, 5, 6};
            const double *A[4] = {ci, vi, ci, vi}, *B[4] = {ck, ck, vk, vk};
            double epair = 0.0, efull = 0.0;
            for (int part = 0; part < 4; part++) {
                damped_series(&c, q, q->m, dm[part], nd[part], R);
                for (int j = 0; j < 6; j++) R[j] *= f;
                gpair g;
                gpair_eval(R, A[part], B[part], q->r, want_grad, &g);
                em_real += g.e;
                epair += g.e;
                if (cf) {
                    /* the pair energy is linear in each valence charge */
                    gpair gq;
                    if (A[part] == vi) { gpair_eval(R, e0, B[part], q->r, 0, &gq); pot_m[q->i] += gq.e; }
                    if (B[part] == vk) { gpair_eval(R, A[part], e0, q->r, 0, &gq); pot_m[q->k] += gq.e; }
                }
                if (c.ewald && same) {
                    double R0[6];
                    gpair g0;
                    damped_series(&c0, q, q->m, dm[part], nd[part], R0);
                    for (int j = 0; j < 6; j++) R0[j] *= f;
                    gpair_eval(R0, A[part], B[part], q->r, 0, &g0);
                    efull += g0.e;
                }
                if (want_vir) pair_vir(vm, q->r, g.frc, 1.0);
                if (want_grad)
                    for (int a = 0; a < 3; a++) {
                        dem[3 * q->i + a] += g.frc[a];
                        dem[3 * q->k + a] -= g.frc[a];
                        te[3 * q->i + a] += g.trqi[a];
                        te[3 * q->k + a] += g.trqk[a];
                    }
            }
            if (epair != 0.0) nem++;
            if (aem) { aem[q->i] += 0.5 * epair; aem[q->k] += 0.5 * epair; }
            if (c.ewald) { if (same) eintra += efull; }
            else if (!same) einter += epair;
            continue;
        }
        series_rr(q->r2, rr);
        if (c.ewald) {
            double bn[6];
            series_ewald(q->r2, c.a, bn);
            for (int j = 0; j < 6; j++) R[j] = f * (bn[j] - (1.0 - q->m) * rr[j]);
        } else {
            for (int j = 0; j < 6; j++) R[j] = f * q->m * rr[j];
        }
        gpair g;
        gpair_eval(R, c.rp + 10 * q->i, c.rp + 10 * q->k, q->r, want_grad, &g);
        em_real += g.e;
        if (g.e != 0.0) nem++;
        if (aem) { aem[q->i] += 0.5 * g.e; aem[q->k] += 0.5 * g.e; }
        if (c.ewald) {
            if (same) {
                double R0[6];
                gpair g0;
                for (int j = 0; j < 6; j++) R0[j] = f * q->m * rr[j];
                gpair_eval(R0, c.rp + 10 * q->i, c.rp + 10 * q->k, q->r, 0, &g0);
                eintra += g0.e;
            }
        } else if (!same) {
            einter += g.e;
        }
        if (want_vir) pair_vir(vm, q->r, g.frc, 1.0);
        if (want_grad)
            for (int a = 0; a < 3; a++) {
                dem[3 * q->i + a] += g.frc[a];
                dem[3 * q->k + a] -= g.frc[a];
                te[3 * q->i + a] += g.trqi[a];
                te[3 * q->k + a] += g.trqk[a];
            }
    }
    /* ---- reciprocal space and self energy */
    double em_recip = 0.0, em_self = 0.0;
    sitepot *pm = NULL;
    if (c.ewald) {
        pm = (sitepot *)malloc(n * sizeof(sitepot));
        em_recip = f * pme_potential(&c, c.rp, pm, want_vir ? vkm : NULL);
        if (want_vir) {
            for (int j = 0; j < 9; j++) vm[j] += 0.5 * f * vkm[j];
            for (int i = 0; i < n; i++) cterm_vir(vm, c.rp + 10 * i, &pm[i], f);
        }
        for (int i = 0; i < n; i++) {
            const double *m = c.rp + 10 * i;
            const double qq = m[4] * m[4] + m[7] * m[7] + m[9] * m[9] + 2.0 * (m[5] * m[5] + m[6] * m[6] + m[8] * m[8]);
            const double dd = m[1] * m[1] + m[2] * m[2] + m[3] * m[3];
            const double a2 = c.a * c.a;
            const double es = -f * c.a / SQRTPI * (m[0] * m[0] + 2.0 * a2 / 3.0 * dd + 8.0 * a2 * a2 / 5.0 * qq);
            em_self += es;
            int site = 0;
            for (int j = 0; j < 10; j++) site |= m[j] != 0.0;
            nem += site;
            if (aem) aem[i] += es + 0.5 * f * site_energy(m, &pm[i], NULL, NULL);
            if (want_grad) {
                double g[3], tq[3];
                site_energy(m, &pm[i], g, tq);
                for (int a = 0; a < 3; a++) { dem[3 * i + a] += f * g[a]; te[3 * i + a] += f * tq[a]; }
            }
            /* (f/2) sum M.phi and the self energy in the charge of i */
            if (cf) pot_m[i] += f * pm[i].phi - 2.0 * f * c.a / SQRTPI * m[0];
        }
    }
    o->em = em_real + em_recip + em_self;
    o->nem = nem;
    o->einter = c.ewald ? o->em - eintra : einter;
    o->em_real = em_real;
    o->em_recip = em_recip;
    o->em_self = em_self;
    if (want_grad && (o->dem || want_vir)) {
        orc_torque(mp, te, dem, want_vir ? vm : NULL);
        if (cf) orc_charge_flux_grad(mp, pot_m, dem, want_vir ? vm : NULL);
        if (o->dem) memcpy(o->dem, dem, 3 * (size_t)n * sizeof(double));
    }

    /* ---- polarization */
    o->ep = 0.0;
    o->iters = 0;
    if (s->use_polar) {
        const int n3 = 3 * n;
        double *Ed = malloc(n3 * sizeof(double)), *Ep = malloc(n3 * sizeof(double));
        double *ud = o->uind ? o->uind : malloc(n3 * sizeof(double));
        double *up = o->uinp ? o->uinp : malloc(n3 * sizeof(double));
        direct_field(&c, L, np, pm, Ed, Ep);
        o->iters = s->solver == 1 ? solve_dense(&c, L, np, Ed, Ep, ud, up) : solve_pcg(&c, L, np, Ed, Ep, ud, up);
        double ep = 0.0;
        for (int j = 0; j < n3; j++) ep += ud[j] * Ep[j];
        o->ep = -0.5 * f * ep;
        o->nep = 0;
        for (int i = 0; i < n; i++) {
            const double ei = -0.5 * f * (ud[3 * i] * Ep[3 * i] + ud[3 * i + 1] * Ep[3 * i + 1] + ud[3 * i + 2] * Ep[3 * i + 2]);
            if (o->aep) o->aep[i] = ei;
            o->nep += ei != 0.0;
        }
        if (want_grad && (o->dep || want_vir)) {
            memset(te, 0, n3 * sizeof(double));
            /* real space: (1/2)[U_p(mu_d, M) + U_d(mu_p, M) + U_u(mu_d, mu_p) + U_u(mu_p, mu_d)] */
            for (long t = 0; t < np; t++) {
                const pairrec *q = &L[t];
                const int i = q->i, k = q->k;
                const double *Mi = c.rp + 10 * i, *Mk = c.rp + 10 * k;
                double mdi[10] = {0}, mdk[10] = {0}, mpi[10] = {0}, mpk[10] = {0};
                for (int a = 0; a < 3; a++) {
                    mdi[1 + a] = ud[3 * i + a]; mdk[1 + a] = ud[3 * k + a];
                    mpi[1 + a] = up[3 * i + a]; mpk[1 + a] = up[3 * k + a];
                }
                double R[6];
                gpair g1, g2;
                double frc[3] = {0, 0, 0}, ti[3] = {0, 0, 0}, tk[3] = {0, 0, 0};
                if (c.cp) {
                    /* the induced dipoles are points: each sees the core of the
                     * other site undamped and its valence multipole damped by
                     * that site's one-site function */
                    double dmpi[5], dmpk[5], dmpik[6], one[1] = {1.0};
                    orc_damp_gordon1(sqrt(q->r2), c.alpha[i], c.alpha[k], dmpi, dmpk, dmpik);
                    double ci[10] = {0}, ck[10] = {0}, vi[10], vk[10];
                    memcpy(vi, Mi, sizeof vi);
                    memcpy(vk, Mk, sizeof vk);
                    ci[0] = c.core[i]; ck[0] = c.core[k];
                    vi[0] -= ci[0]; vk[0] -= ck[0];
                    for (int set = 0; set < 2; set++) {
                        const double sc = set ? q->d : q->p;  /* mu_d sees E^p, mu_p sees E^d */
                        const double *di_ = set ? mpi : mdi, *dk_ = set ? mpk : mdk;
                        damped_series(&c, q, sc, one, 1, R);
                        gpair_eval(R, di_, ck, q->r, 1, &g1);
                        gpair_eval(R, ci, dk_, q->r, 1, &g2);
                        for (int a = 0; a < 3; a++) frc[a] += g1.frc[a] + g2.frc[a];
                        damped_series(&c, q, sc, dmpk, 5, R);
                        gpair_eval(R, di_, vk, q->r, 1, &g1);
                        for (int a = 0; a < 3; a++) { frc[a] += g1.frc[a]; tk[a] += g1.trqk[a]; }
                        if (cf) { gpair gq; gpair_eval(R, di_, e0, q->r, 0, &gq); pot_p[k] += 0.5 * f * gq.e; }
                        damped_series(&c, q, sc, dmpi, 5, R);
                        gpair_eval(R, vi, dk_, q->r, 1, &g2);
                        for (int a = 0; a < 3; a++) { frc[a] += g2.frc[a]; ti[a] += g2.trqi[a]; }
                        if (cf) { gpair gq; gpair_eval(R, e0, dk_, q->r, 0, &gq); pot_p[i] += 0.5 * f * gq.e; }
                    }
                    pol_series(&c, q, q->w, R);
                    gpair_eval(R, mdi, mpk, q->r, 1, &g1);
                    gpair_eval(R, mpi, mdk, q->r, 1, &g2);
                    for (int a = 0; a < 3; a++) frc[a] += g1.frc[a] + g2.frc[a];
                    if (want_vir) pair_vir(vp, q->r, frc, 0.5 * f);
                    for (int a = 0; a < 3; a++) {
                        dep[3 * i + a] += 0.5 * f * frc[a];
                        dep[3 * k + a] -= 0.5 * f * frc[a];
                        te[3 * i + a] += 0.5 * f * ti[a];
                        te[3 * k + a] += 0.5 * f * tk[a];
                    }
                    continue;
                }
                /* mu_d with p-scaled permanent field */
                pol_series(&c, q, q->p, R);
                gpair_eval(R, mdi, Mk, q->r, 1, &g1);
                gpair_eval(R, Mi, mdk, q->r, 1, &g2);
                for (int a = 0; a < 3; a++) { frc[a] += g1.frc[a] + g2.frc[a]; tk[a] += g1.trqk[a]; ti[a] += g2.trqi[a]; }
                /* mu_p with d-scaled permanent field */
                pol_series(&c, q, q->d, R);
                gpair_eval(R, mpi, Mk, q->r, 1, &g1);
                gpair_eval(R, Mi, mpk, q->r, 1, &g2);
                for (int a = 0; a < 3; a++) { frc[a] += g1.frc[a] + g2.frc[a]; tk[a] += g1.trqk[a]; ti[a] += g2.trqi[a]; }
                /* dipole-dipole, u-scaled */
                pol_series(&c, q, q->u, R);
                gpair_eval(R, mdi, mpk, q->r, 1, &g1);
                gpair_eval(R, mpi, mdk, q->r, 1, &g2);
                for (int a = 0; a < 3; a++) frc[a] += g1.frc[a] + g2.frc[a];
                if (want_vir) pair_vir(vp, q->r, frc, 0.5 * f);
                for (int a = 0; a < 3; a++) {
                    dep[3 * i + a] += 0.5 * f * frc[a];
                    dep[3 * k + a] -= 0.5 * f * frc[a];
                    te[3 * i + a] += 0.5 * f * ti[a];
                    te[3 * k + a] += 0.5 * f * tk[a];
                }
            }
            if (c.ewald) {
                /* reciprocal space: potentials of mu_d and mu_p */
                double *md = calloc(10 * (size_t)n, sizeof(double)), *mq = calloc(10 * (size_t)n, sizeof(double));
                for (int i = 0; i < n; i++)
                    for (int a = 0; a < 3; a++) { md[10 * i + 1 + a] = ud[3 * i + a]; mq[10 * i + 1 + a] = up[3 * i + a]; }
                sitepot *pd = (sitepot *)malloc(n * sizeof(sitepot)), *pq = (sitepot *)malloc(n * sizeof(sitepot));
                double vkd[9] = {0}, vkp[9] = {0}, vka[9] = {0};
                pme_potential(&c, md, pd, want_vir ? vkd : NULL);
                pme_potential(&c, mq, pq, want_vir ? vkp : NULL);
                if (want_vir) {
                    /* the reciprocal part of the polarization energy, (1/2) E(mu_d + mu_p, M)
                     * + (1/2) E(mu_d, mu_p) as cross energies: its k-space sum is
                     * (1/2)[V(M + mu_d + mu_p) - V(M) - V(mu_d) - V(mu_p)] */
                    double *ma = (double *)malloc(10 * (size_t)n * sizeof(double));
                    for (size_t j = 0; j < 10 * (size_t)n; j++) ma[j] = c.rp[j] + md[j] + mq[j];
                    pme_potential(&c, ma, NULL, vka);
                    free(ma);
                    for (int j = 0; j < 9; j++) vp[j] += 0.25 * f * (vka[j] - vkm[j] - vkd[j] - vkp[j]);
                }
                const double self = 4.0 / 3.0 * c.a * c.a * c.a / SQRTPI;
                for (int i = 0; i < n; i++) {
                    double g[3], tq[3];
                    /* induced dipoles in the permanent potential and in each other's */
                    for (int x = 0; x < 3; x++) {
                        double v = 0.0;
                        for (int a = 0; a < 3; a++)
                            v += (ud[3 * i + a] + up[3 * i + a]) * pm[i].H[a][x] + ud[3 * i + a] * pq[i].H[a][x]
                               + up[3 * i + a] * pd[i].H[a][x];
                        dep[3 * i + x] += 0.5 * f * v;
                    }
                    /* permanent multipoles in the induced dipoles' potential */
                    sitepot ps;
                    ps.phi = pd[i].phi + pq[i].phi;
                    for (int a = 0; a < 3; a++) {
                        ps.g[a] = pd[i].g[a] + pq[i].g[a];
                        for (int b = 0; b < 3; b++) {
                            ps.H[a][b] = pd[i].H[a][b] + pq[i].H[a][b];
                            for (int d = 0; d < 3; d++) ps.T[a][b][d] = pd[i].T[a][b][d] + pq[i].T[a][b][d];
                        }
                    }
                    site_energy(c.rp + 10 * i, &ps, g, tq);
                    for (int a = 0; a < 3; a++) { dep[3 * i + a] += 0.5 * f * g[a]; te[3 * i + a] += 0.5 * f * tq[a]; }
                    if (cf) pot_p[i] += 0.5 * f * ps.phi;
                    if (want_vir) {
                        /* strain terms of the cross energies: (1/2)[C(mu_d + mu_p in phi_M)
                         * + C(M in phi_(mu_d + mu_p)) + C(mu_d in phi_p) + C(mu_p in phi_d)] */
                        double mu[10] = {0};
                        for (int a = 0; a < 3; a++) mu[1 + a] = ud[3 * i + a] + up[3 * i + a];
                        cterm_vir(vp, mu, &pm[i], 0.5 * f);
                        cterm_vir(vp, c.rp + 10 * i, &ps, 0.5 * f);
                        cterm_vir(vp, md + 10 * i, &pq[i], 0.5 * f);
                        cterm_vir(vp, mq + 10 * i, &pd[i], 0.5 * f);
                    }
                    /* self term: -(f/2) self (mu_d + mu_p) . d_i gives a torque on d_i */
                    double Eeff[3], tt[3];
                    for (int a = 0; a < 3; a++) Eeff[a] = 0.5 * f * self * (ud[3 * i + a] + up[3 * i + a]);
                    cross3(c.rp + 10 * i + 1, Eeff, tt);
                    for (int a = 0; a < 3; a++) te[3 * i + a] += tt[a];
                }
                free(md); free(mq); free(pd); free(pq);
            }
            orc_torque(mp, te, dep, want_vir ? vp : NULL);
            if (cf) orc_charge_flux_grad(mp, pot_p, dep, want_vir ? vp : NULL);
            if (o->dep) memcpy(o->dep, dep, n3 * sizeof(double));
        }
        free(Ed); free(Ep);
        if (!o->uind) free(ud);
        if (!o->uinp) free(up);
    }

    if (want_vir && o->vir_m) memcpy(o->vir_m, vm, sizeof vm);
    if (want_vir && o->vir_p) memcpy(o->vir_p, vp, sizeof vp);
    free(L); free(te); free(dem); free(dep); free(pole); free(rp13); free(c.rp); free(pm); free(pot_m); free(pot_p);
    free(c.o12); free(c.o13); free(c.o14); free(c.o15); free(c.g11); free(c.g12); free(c.g13); free(c.g14);
    if (c.ewald) pme_free(&c);
    return 0;
}


/* =====================================================================
 * Part 3: HIPPO dispersion and Pauli repulsion
 *
 * Dispersion (Rackers et al., JCP 149, 084115 (2018)): the C6 pair energy
 * -C_i C_k f^2 / r^6 with f = 1.5 d5 - 0.5 d3, d3/d5 the Gordon overlap
 * damping factors of the adisp exponents (dmpik orders 3 and 5 of
 * damping.h); scales dsp2..5 by connectivity.  With dispersion Ewald
 * (coefficient a, g(x) = (1 + x^2 + x^4/2) e^{-x^2}):
 *   real   -C_i C_k [s f^2 - 1 + g(a r)] / r^6,
 *   recip  (1/2) sum_m G(m) |rho(m)|^2, G = -psi(2 pi |h|) / (V bsmod),
 *          psi(k) = (2 pi^1.5 a^3 / 3)[sqrt(pi) b^3 erfc(b) + (1/2 - b^2) e^{-b^2}],
 *          b = k / (2a), m = 0 included,
 *   self   + a^6 / 12 sum C^2.
 * Without Ewald the pair energy is switched off between disptaper and
 * dispcut by the quintic of cutoffSwitch.cpp:125-133.
 *
 * Repulsion (Rackers and Ponder, JCP 150, 084104 (2019), cited at
 * erepel.cpp:36-40): E = K_i K_k s_ik / r * M(B), M the multipole pair
 * contraction (valence count elepr as monopole, the rotated dipoles and
 * quadrupoles) over the radial series B_n = D^n (S^2), D = -(1/r) d/dr, S
 * the overlap of two normalized 1s Slater functions of exponents
 * dmppr/2 (equal exponents within 1e-3 use their mean).  The pair energy is
 * switched off between reptaper and repcut as above.
 * ===================================================================== */

/* S = sum_c P_c(r) e^{-c r} with Laurent polynomials P_c (powers -12..2) */
#define RS_LO (-12)
#define RS_N 15
typedef struct {
    double c[2];
    double p[2][RS_N]; /* p[t][m - RS_LO] */
    int nt;
} rs_fn;

static void rs_overlap(double ai, double ak, rs_fn *S)
{
    const double a = 0.5 * ai, b = 0.5 * ak;
    memset(S, 0, sizeof *S);
    if (fabs(ai - ak) < 1e-3) {
        const double z = 0.25 * (ai + ak);
        S->nt = 1;
        S->c[0] = z;
        S->p[0][0 - RS_LO] = 1.0;
        S->p[0][1 - RS_LO] = z;
        S->p[0][2 - RS_LO] = z * z / 3.0;
        return;
    }
    const double d = a * a - b * b, t = 4.0 * a * b / d, A = 8.0 * pow(a * b, 1.5) / (d * d);
    S->nt = 2;
    S->c[0] = b;
    S->p[0][0 - RS_LO] = A * a;
    S->p[0][-1 - RS_LO] = -A * t;
    S->c[1] = a;
    S->p[1][0 - RS_LO] = A * b;
    S->p[1][-1 - RS_LO] = A * t;
}

/* D (P e^{-c r}) = sum_m (-m p_m r^{m-2} + c p_m r^{m-1}) e^{-c r} */
static void rs_apply_D(rs_fn *S)
{
    for (int t = 0; t < S->nt; t++) {
        double q[RS_N] = {0};
        for (int j = 0; j < RS_N; j++) {
            const double v = S->p[t][j];
            if (v == 0.0) continue;
            const int m = j + RS_LO;
            if (j >= 2) q[j - 2] += -m * v;
            if (j >= 1) q[j - 1] += S->c[t] * v;
        }
        memcpy(S->p[t], q, sizeof q);
    }
}

static double rs_eval(const rs_fn *S, double r)
{
    double v = 0.0;
    for (int t = 0; t < S->nt; t++) {
        double s = 0.0, rp = pow(r, RS_LO);
        for (int j = 0; j < RS_N; j++) { s += S->p[t][j] * rp; rp *= r; }
        v += s * exp(-S->c[t] * r);
    }
    return v;
}

void orc_rep_series(double r, double ai, double ak, double *B)
{
    rs_fn S;
    rs_overlap(ai, ak, &S);
    double Sn[6];
    for (int n = 0; n < 6; n++) {
        Sn[n] = rs_eval(&S, r);
        rs_apply_D(&S);
    }
    static const double C[6][6] = {{1}, {1, 1}, {1, 2, 1}, {1, 3, 3, 1}, {1, 4, 6, 4, 1}, {1, 5, 10, 10, 5, 1}};
    for (int n = 0; n < 6; n++) {
        double v = 0.0;
        for (int k = 0; k <= n; k++) v += C[n][k] * Sn[k] * Sn[n - k];
        B[n] = v;
    }
}

/* multiplicative quintic switch (cutoffSwitch.cpp:125-133) and d/dr */
static void taper5(double r, double cut, double off, double *T, double *dT)
{
    *T = 1.0;
    *dT = 0.0;
    if (!(cut < off) || r <= cut) return;
    const double den = pow(off - cut, 5), off2 = off * off, cut2 = cut * cut;
    const double c0 = off * off2 * (off2 - 5.0 * off * cut + 10.0 * cut2) / den;
    const double c1 = -30.0 * off2 * cut2 / den;
    const double c2 = 30.0 * (off2 * cut + off * cut2) / den;
    const double c3 = -10.0 * (off2 + 4.0 * off * cut + cut2) / den;
    const double c4 = 15.0 * (off + cut) / den;
    const double c5 = -6.0 / den;
    *T = c0 + r * (c1 + r * (c2 + r * (c3 + r * (c4 + r * c5))));
    *dT = c1 + r * (2.0 * c2 + r * (3.0 * c3 + r * (4.0 * c4 + r * 5.0 * c5)));
}

/* connectivity scale of atom i against every k: vals for 1-2..1-5 (a later
 * class overrides an earlier one); v = 1 resets */
static void conn_scales(const ctx_t *c, int i, const double *vals, double *sc, int reset)
{
    const orc_mpole_sys *m = c->mp;
    for (int q = 0; q < 4; q++) {
        const int *off = q == 0 ? c->o12 : q == 1 ? c->o13 : q == 2 ? c->o14 : c->o15;
        const int *lst = q == 0 ? m->i12 : q == 1 ? m->i13 : q == 2 ? m->i14 : m->i15;
        for (int j = off[i]; j < off[i + 1]; j++) sc[lst[j]] = reset ? 1.0 : vals[q];
    }
}

/* dispersion reciprocal: phi = IDFT[G DFT[rho]], returns sum_m G |rho|^2;
 * with vk adds its strain derivative sum_m |rho|^2 dG/d(strain_ab): G =
 * -psi(b)/(V bsmod), b = pi |h|/a, dV = V delta_ab, d|h| = -h_a h_b/|h|,
 * psi'(b) = 3 pre b (sqrt(pi) b erfc(b) - exp(-b^2)) */
static double disp_convolve(const ctx_t *c, double *grid, double a, double *vk)
{
    const int K0 = c->K[0], K1 = c->K[1], K2 = c->K[2];
    const size_t N = (size_t)K0 * K1 * K2;
    double *im = (double *)calloc(N, sizeof(double));
    for (int ax = 0; ax < 3; ax++) dft_axis(c, grid, im, ax, +1);
    const double V = pme_volume(c);
    const double pre = 2.0 * pow(M_PI, 1.5) * a * a * a / 3.0;
    double e = 0.0;
    for (int m3 = 0; m3 < K2; m3++)
        for (int m2 = 0; m2 < K1; m2++)
            for (int m1 = 0; m1 < K0; m1++) {
                const size_t id = ((size_t)m3 * K1 + m2) * K0 + m1;
                const int f1 = m1 > K0 / 2 ? m1 - K0 : m1, f2 = m2 > K1 / 2 ? m2 - K1 : m2, f3 = m3 > K2 / 2 ? m3 - K2 : m3;
                double h[3];
                for (int j = 0; j < 3; j++)
                    h[j] = f1 * c->A[0][j] / K0 + f2 * c->A[1][j] / K1 + f3 * c->A[2][j] / K2;
                const double b = M_PI * sqrt(dot3(h, h)) / a;
                const double psi = pre * (SQRTPI * b * b * b * erfc(b) + (0.5 - b * b) * exp(-b * b));
                const double bs = c->bsmod[0][m1] * c->bsmod[1][m2] * c->bsmod[2][m3];
                const double G = -psi / (V * bs);
                const double rho2 = grid[id] * grid[id] + im[id] * im[id];
                e += G * rho2;
                if (vk) {
                    const double w = 3.0 * pre * M_PI * M_PI / (a * a) * (SQRTPI * b * erfc(b) - exp(-b * b)) / (V * bs);
                    for (int p = 0; p < 3; p++)
                        for (int q = 0; q < 3; q++)
                            vk[3 * p + q] += rho2 * (w * h[p] * h[q] - (p == q ? G : 0.0));
                }
                grid[id] *= G;
                im[id] *= G;
            }
    for (int ax = 0; ax < 3; ax++) dft_axis(c, grid, im, ax, -1);
    free(im);
    return e;
}

int orc_hippo(orc_hippo_sys *s, int want_grad, orc_hippo_out *o)
{
    orc_mpole_sys *mp = s->mp;
    const int n = mp->n;
    ctx_t c;
    memset(&c, 0, sizeof c);
    c.n = n;
    c.mp = mp;
    c.o12 = offsets(n, mp->n12); c.o13 = offsets(n, mp->n13);
    c.o14 = offsets(n, mp->n14); c.o15 = offsets(n, mp->n15);
    const int periodic = mp->box.kind != ORC_BOX_NONE;
    const int ew = s->use_dewald && periodic;
    if (ew && !(mp->box.kind == ORC_BOX_ORTHO || mp->box.kind == ORC_BOX_MONO || mp->box.kind == ORC_BOX_TRI))
        return -2;
    double *sc = (double *)malloc(n * sizeof(double));
    double *dd = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *dr = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *te = (double *)calloc(3 * (size_t)n, sizeof(double));
    double vd[9] = {0}, vr[9] = {0};
    const int want_vir = want_grad && (o->vir_d || o->vir_r);
    /* ---- dispersion, real space */
    const double a = s->dewald, a2 = a * a;
    const double dvals[4] = {s->dsp2scale, s->dsp3scale, s->dsp4scale, s->dsp5scale};
    const double dcut2 = s->dispcut * s->dispcut;
    double ed_real = 0.0;
    for (int i = 0; i < n; i++) sc[i] = 1.0;
    for (int i = 0; i < n - 1; i++) {
        conn_scales(&c, i, dvals, sc, 0);
        for (int k = i + 1; k < n; k++) {
            double r3[3] = {mp->x[k] - mp->x[i], mp->y[k] - mp->y[i], mp->z[k] - mp->z[i]};
            if (periodic) orc_image(&mp->box, &r3[0], &r3[1], &r3[2]);
            const double r2 = dot3(r3, r3);
            if (r2 > dcut2) continue;
            const double r = sqrt(r2), cc = s->csix[i] * s->csix[k];
            double dmpi[5], dmpk[5], dmpik[6];
            orc_damp_gordon1(r, s->adisp[i], s->adisp[k], dmpi, dmpk, dmpik);
            const double fd = 1.5 * dmpik[2] - 0.5 * dmpik[1];
            const double fp = 7.5 * (dmpik[2] - dmpik[3]) / r - 1.5 * (dmpik[1] - dmpik[2]) / r;
            const double r6 = r2 * r2 * r2;
            double e, de;
            if (ew) {
                const double x2 = a2 * r2, ex = exp(-x2);
                const double g = (1.0 + x2 + 0.5 * x2 * x2) * ex, gp = -a2 * r * x2 * x2 * ex;
                e = -cc * (sc[k] * fd * fd - 1.0 + g) / r6;
                de = -cc * ((2.0 * sc[k] * fd * fp + gp) / r6 - 6.0 * (sc[k] * fd * fd - 1.0 + g) / (r6 * r));
            } else {
                double T, dT;
                taper5(r, s->disptaper, s->dispcut, &T, &dT);
                const double e0 = -cc * sc[k] * fd * fd / r6;
                const double de0 = -cc * sc[k] * (2.0 * fd * fp / r6 - 6.0 * fd * fd / (r6 * r));
                e = e0 * T;
                de = de0 * T + e0 * dT;
            }
            ed_real += e;
            if (want_vir) {
                const double g[3] = {-de * r3[0] / r, -de * r3[1] / r, -de * r3[2] / r};
                pair_vir(vd, r3, g, 1.0);
            }
            if (want_grad)
                for (int q = 0; q < 3; q++) {
                    dd[3 * k + q] += de * r3[q] / r;
                    dd[3 * i + q] -= de * r3[q] / r;
                }
        }
        conn_scales(&c, i, dvals, sc, 1);
    }
    /* ---- dispersion, reciprocal and self */
    double ed_recip = 0.0, ed_self = 0.0;
    if (ew) {
        orc_elec_sys es;
        memset(&es, 0, sizeof es);
        es.nfft1 = s->nfft1; es.nfft2 = s->nfft2; es.nfft3 = s->nfft3; es.bsorder = s->bsorder;
        c.s = &es;
        pme_setup(&c);
        const size_t N = (size_t)c.K[0] * c.K[1] * c.K[2];
        double *fm = (double *)calloc(10 * (size_t)n, sizeof(double));
        double *fp = (double *)malloc(20 * (size_t)n * sizeof(double));
        double *grid = (double *)malloc(N * sizeof(double));
        for (int i = 0; i < n; i++) fm[10 * i] = s->csix[i];
        pme_spread(&c, fm, 1, grid);
        double vkd[9] = {0};
        ed_recip = 0.5 * disp_convolve(&c, grid, a, want_vir ? vkd : NULL);
        for (int j = 0; j < 9; j++) vd[j] += 0.5 * vkd[j];
        pme_gather(&c, grid, 4, fp);
        for (int i = 0; i < n; i++) {
            ed_self += a2 * a2 * a2 / 12.0 * s->csix[i] * s->csix[i];
            if (want_grad) {
                double phi, g[3], H[3][3], T[3][3][3];
                double f20[20] = {0};
                for (int j = 0; j < 4; j++) f20[j] = fp[20 * i + j];
                frac_to_cart(&c, f20, &phi, g, H, T);
                for (int q = 0; q < 3; q++) dd[3 * i + q] += s->csix[i] * g[q];
            }
        }
        free(fm); free(fp); free(grid);
        pme_free(&c);
    }
    o->edsp_real = ed_real;
    o->edsp_recip = ed_recip;
    o->edsp_self = ed_self;
    o->edsp = ed_real + ed_recip + ed_self;

    /* ---- repulsion */
    double *pole = (double *)malloc(13 * (size_t)n * sizeof(double));
    double *rp13 = (double *)malloc(13 * (size_t)n * sizeof(double));
    memcpy(pole, mp->pole, 13 * (size_t)n * sizeof(double));
    orc_chkpole(mp, pole);
    orc_rotpole(mp, pole, rp13);
    static const int idx[10] = {0, 1, 2, 3, 4, 5, 6, 8, 9, 12};
    const double rvals[4] = {s->r2scale, s->r3scale, s->r4scale, s->r5scale};
    const double rcut2 = s->repcut * s->repcut;
    double er = 0.0;
    for (int i = 0; i < n; i++) sc[i] = 1.0;
    for (int i = 0; i < n - 1; i++) {
        if (s->sizpr[i] == 0.0) continue;
        conn_scales(&c, i, rvals, sc, 0);
        for (int k = i + 1; k < n; k++) {
            if (s->sizpr[k] == 0.0 || sc[k] == 0.0) continue;
            double r3[3] = {mp->x[k] - mp->x[i], mp->y[k] - mp->y[i], mp->z[k] - mp->z[i]};
            if (periodic) orc_image(&mp->box, &r3[0], &r3[1], &r3[2]);
            const double r2 = dot3(r3, r3);
            if (r2 > rcut2) continue;
            const double r = sqrt(r2);
            double mi[10], mk[10], B[6];
            for (int j = 0; j < 10; j++) { mi[j] = rp13[13 * i + idx[j]]; mk[j] = rp13[13 * k + idx[j]]; }
            mi[0] = s->elepr[i];
            mk[0] = s->elepr[k];
            orc_rep_series(r, s->dmppr[i], s->dmppr[k], B);
            gpair g;
            gpair_eval(B, mi, mk, r3, want_grad, &g);
            const double K = s->sizpr[i] * s->sizpr[k] * sc[k];
            double T, dT;
            taper5(r, s->reptaper, s->repcut, &T, &dT);
            const double e0 = K * g.e / r;
            er += e0 * T;
            if (want_grad) {
                double fv[3];
                for (int q = 0; q < 3; q++) {
                    /* d/dx_i of K M / r: K (dM/dx_i / r + M r_vec / r^3); taper d/dx_i = -T' r_vec / r */
                    const double f = T * K * (g.frc[q] / r + g.e * r3[q] / (r2 * r)) - e0 * dT * r3[q] / r;
                    fv[q] = f;
                    dr[3 * i + q] += f;
                    dr[3 * k + q] -= f;
                    te[3 * i + q] += T * K * g.trqi[q] / r;
                    te[3 * k + q] += T * K * g.trqk[q] / r;
                }
                if (want_vir) pair_vir(vr, r3, fv, 1.0);
            }
        }
        conn_scales(&c, i, rvals, sc, 1);
    }
    o->er = er;
    if (want_grad) {
        orc_torque(mp, te, dr, want_vir ? vr : NULL);
        if (o->dedsp) memcpy(o->dedsp, dd, 3 * (size_t)n * sizeof(double));
        if (want_vir && o->vir_d) memcpy(o->vir_d, vd, sizeof vd);
        if (want_vir && o->vir_r) memcpy(o->vir_r, vr, sizeof vr);
        if (o->der) memcpy(o->der, dr, 3 * (size_t)n * sizeof(double));
    }
    free(sc); free(dd); free(dr); free(te); free(pole); free(rp13);
    free(c.o12); free(c.o13); free(c.o14); free(c.o15);
    return 0;
}

/* ---------------------------------------------------------------------
 * Part 4: buffered 14-7 van der Waals and charge transfer (Tinker's
 * published ehal1 and echgtrn1 forms; see pmq_oracle.h)
 * ------------------------------------------------------------------- */
static void hal_pair(double r, double rv, double eps, double *e, double *de)
{
    const double dhal = 0.07, ghal = 0.12;
    const double rho = r / rv, rho6 = pow(rho, 6), rho7 = rho6 * rho;
    const double s1 = 1.0 / (rho + dhal), s2 = 1.0 / (rho7 + ghal);
    const double t1 = pow((1.0 + dhal) * s1, 7), t2 = (1.0 + ghal) * s2;
    *e = eps * t1 * (t2 - 2.0);
    /* dt1/drho = -7 t1 s1, dt2/drho = -7 rho^6 t2 s2 */
    *de = eps * (-7.0 * t1 * s1 * (t2 - 2.0) - 7.0 * t1 * rho6 * t2 * s2) / rv;
}

int orc_vdwct(orc_vdwct_sys *s, int want_grad, orc_vdwct_out *o)
{
    orc_mpole_sys *m = s->mp;
    const int n = m->n;
    int *o12 = offsets(n, m->n12), *o13 = offsets(n, m->n13), *o14 = offsets(n, m->n14), *o15 = offsets(n, m->n15);
    double *vs = malloc(n * sizeof(double)), *ms = malloc(n * sizeof(double));
    int *is14 = calloc(n, sizeof(int));
    double *xr = malloc(3 * (size_t)n * sizeof(double));
    double *gv = calloc(3 * (size_t)n, sizeof(double)), *gc = calloc(3 * (size_t)n, sizeof(double));
    double vv[9] = {0}, vc[9] = {0};
    for (int i = 0; i < n; i++) vs[i] = ms[i] = 1.0;
    /* reduced vdW sites */
    for (int i = 0; i < n; i++) {
        const double p[3] = {m->x[i], m->y[i], m->z[i]};
        if (s->use_vdw && s->ired[i] != i) {
            const int j = s->ired[i];
            const double q[3] = {m->x[j], m->y[j], m->z[j]};
            for (int a = 0; a < 3; a++) xr[3 * i + a] = s->kred[i] * (p[a] - q[a]) + q[a];
        } else {
            for (int a = 0; a < 3; a++) xr[3 * i + a] = p[a];
        }
    }
    const double f = m->f;
    const double vsc[4] = {s->v2scale, s->v3scale, s->v4scale, s->v5scale};
    const double msc[4] = {m->m2scale, m->m3scale, m->m4scale, m->m5scale};
    double ev = 0.0, ect = 0.0;
    for (int i = 0; i < n - 1; i++) {
        for (int q = 0; q < 4; q++) {
            const int *off = q == 0 ? o12 : q == 1 ? o13 : q == 2 ? o14 : o15;
            const int *lst = q == 0 ? m->i12 : q == 1 ? m->i13 : q == 2 ? m->i14 : m->i15;
            for (int j = off[i]; j < off[i + 1]; j++) { vs[lst[j]] = vsc[q]; ms[lst[j]] = msc[q]; }
        }
        for (int j = o14[i]; j < o14[i + 1]; j++) is14[m->i14[j]] = 1;
        const int *u = m->use;
        const int ui = !u || u[i + 1], uvi = ui || (s->use_vdw && u[s->ired[i] + 1]);
        for (int k = i + 1; k < n; k++) {
            /* a pair counts when either atom is active; for vdW an atom's
             * reduced site is also active when its reduction partner is */
            const int uk = !u || u[k + 1];
            if (s->use_vdw && vs[k] != 0.0 && (uvi || uk || u[s->ired[k] + 1])) {
                double d[3] = {xr[3 * k] - xr[3 * i], xr[3 * k + 1] - xr[3 * i + 1], xr[3 * k + 2] - xr[3 * i + 2]};
                if (m->box.kind != ORC_BOX_NONE) orc_image(&m->box, &d[0], &d[1], &d[2]);
                const double r2 = dot3(d, d);
                const int t = s->jv[i] * s->nv + s->jv[k];
                const double rv = is14[k] ? s->radmin4[t] : s->radmin[t];
                const double ep = (is14[k] ? s->epsilon4[t] : s->epsilon[t]) * vs[k];
                /* a pair with no well depth or no radius has no energy */
                if (r2 <= s->vdwcut * s->vdwcut && ep != 0.0 && rv > 0.0) {
                    const double r = sqrt(r2);
                    double e, de, T, dT;
                    hal_pair(r, rv, ep, &e, &de);
                    taper5(r, s->vdwtaper, s->vdwcut, &T, &dT);
                    de = de * T + e * dT;
                    e *= T;
                    ev += e;
                    if (want_grad) {
                        const double g[3] = {de * d[0] / r, de * d[1] / r, de * d[2] / r}; /* dE/dx_k */
                        const int pi[2] = {i, k};
                        for (int side = 0; side < 2; side++) {
                            const int a = pi[side], b = s->ired[a];
                            const double sg = side ? 1.0 : -1.0, kr = s->kred[a];
                            for (int c = 0; c < 3; c++) {
                                if (b != a) {
                                    gv[3 * a + c] += sg * kr * g[c];
                                    gv[3 * b + c] += sg * (1.0 - kr) * g[c];
                                } else {
                                    gv[3 * a + c] += sg * g[c];
                                }
                            }
                        }
                        pair_vir(vv, d, g, -1.0);
                    }
                }
            }
            if (s->use_ctrn && ms[k] != 0.0 && (ui || uk)) {
                double d[3] = {m->x[k] - m->x[i], m->y[k] - m->y[i], m->z[k] - m->z[i]};
                if (m->box.kind != ORC_BOX_NONE) orc_image(&m->box, &d[0], &d[1], &d[2]);
                const double r2 = dot3(d, d);
                if (r2 <= s->ctrncut * s->ctrncut) {
                    const double r = sqrt(r2);
                    const double ai = s->dmpct[i] != 0.0 ? s->dmpct[i] : 1000.0;
                    const double ak = s->dmpct[k] != 0.0 ? s->dmpct[k] : 1000.0;
                    const double ei = exp(-ai * r), ek = exp(-ak * r);
                    const double sc = f * ms[k];
                    double e = -sc * (s->chgct[i] * ek + s->chgct[k] * ei);
                    double de = sc * (s->chgct[i] * ak * ek + s->chgct[k] * ai * ei);
                    double T, dT;
                    taper5(r, s->ctrntaper, s->ctrncut, &T, &dT);
                    de = de * T + e * dT;
                    e *= T;
                    ect += e;
                    if (want_grad) {
                        const double g[3] = {de * d[0] / r, de * d[1] / r, de * d[2] / r};
                        for (int c = 0; c < 3; c++) { gc[3 * k + c] += g[c]; gc[3 * i + c] -= g[c]; }
                        pair_vir(vc, d, g, -1.0);
                    }
                }
            }
        }
        for (int q = 0; q < 4; q++) {
            const int *off = q == 0 ? o12 : q == 1 ? o13 : q == 2 ? o14 : o15;
            const int *lst = q == 0 ? m->i12 : q == 1 ? m->i13 : q == 2 ? m->i14 : m->i15;
            for (int j = off[i]; j < off[i + 1]; j++) { vs[lst[j]] = 1.0; ms[lst[j]] = 1.0; }
        }
        for (int j = o14[i]; j < o14[i + 1]; j++) is14[m->i14[j]] = 0;
    }
    o->ev = ev;
    o->ect = ect;
    if (want_grad) {
        if (o->dev) memcpy(o->dev, gv, 3 * (size_t)n * sizeof(double));
        if (o->dect) memcpy(o->dect, gc, 3 * (size_t)n * sizeof(double));
        if (o->vir_v) memcpy(o->vir_v, vv, sizeof vv);
        if (o->vir_c) memcpy(o->vir_c, vc, sizeof vc);
    }
    free(o12); free(o13); free(o14); free(o15); free(vs); free(ms); free(is14); free(xr); free(gv); free(gc);
    return 0;
}
