/*
 * pmq_oracle.h -- FP64 CPU restatement of the polarizable-electrostatics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against; it is never linked into, loaded by or called from the
 * product library (polmdqc_amd/libpmq.so).  Only tests/, the smoke check and
 * bench.py's CPU-baseline leg may use it.
 *
 * Part 1 restates the reference's multipole energy (src/md/empole.cpp:96-357
 * and what it calls: rotpole.cpp, chkpole.cpp, torque.cpp, image.cpp,
 * alterchg.cpp, dcflux.cpp, include/md/pairMpole.h, damping.h) on plain
 * arrays with the reference's conventions (1-based frame atoms, 0 = none,
 * signed yaxis for chirality, use[] of length n+1).
 *
 * Part 2 (pmq_oracle_ewald.c) adds what the reference declares but does not
 * compute: Ewald real space, PME, polarization (direct field, mutual PCG,
 * energy and gradient) and the HIPPO terms.
 */
#ifndef PMQ_ORACLE_H
#define PMQ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* periodic box, boxes.h:46-67 conventions */
enum { ORC_BOX_NONE = 0, ORC_BOX_ORTHO, ORC_BOX_MONO, ORC_BOX_TRI, ORC_BOX_OCTA, ORC_BOX_DODECA };

typedef struct {
    int kind;
    double xbox, ybox, zbox;
    double alpha, beta, gamma; /* degrees */
} orc_box;

/* frame types, kmulti.h:23-31 */
enum { ORC_FRAME_NONE = 0, ORC_FRAME_ZTHENX, ORC_FRAME_ZONLY, ORC_FRAME_BISECTOR,
       ORC_FRAME_ZBISECT, ORC_FRAME_THREEFOLD };

typedef struct {
    int n;
    const double *x, *y, *z;
    const double *pole;        /* n*13 local-frame multipoles (kmpole units) */
    const int *polaxe;         /* n */
    const int *zaxis, *xaxis;  /* n, 1-based, 0 = none */
    int *yaxis;                /* n, 1-based signed; chkpole may flip the sign */
    const int *use;            /* n+1, use[0] = 0 */
    const int *molcule;        /* n */
    const int *n12, *i12, *n13, *i13, *n14, *i14, *n15, *i15; /* counts and flat lists */
    double m2scale, m3scale, m4scale, m5scale;
    double f;                  /* electric / dielec */
    double off2;               /* squared cutoff */
    orc_box box;
    int pentype;               /* 0 none, 1 Gordon1, 2 Gordon2 (mplpot.h:25) */
    const double *pcore, *pval0, *palpha, *mono0; /* n each (charge penetration) */
    int use_chgflx;
    int nbond;  const int *ibnd; const double *bl, *bflx;          /* 2*nbond, nbond, nbond */
    int nangle; const int *iang; const double *anat, *aflx, *abflx; /* 3*nangle, nangle, 2*, 2* */
    const int *balist;         /* 2*nangle: bond indices of the angle's two bonds */
} orc_mpole_sys;

/* calculation modes, calcMode.h:35-49 */
enum { ORC_ENERGY = 0, ORC_ANALYSIS = 1, ORC_GRADIENT = 2, ORC_VIRIAL = 3 };

typedef struct {
    double em;
    double einter;
    long nem;
    double *aem;    /* n, analysis mode (may be NULL) */
    double *dem;    /* 3n AoS, gradient modes (may be NULL) */
    double *vir;    /* 9, row-major (may be NULL) */
    double *rpole;  /* 13n global-frame multipoles (may be NULL) */
    double *pole_out; /* 13n local multipoles after chkpole/charge flux (may be NULL) */
} orc_mpole_out;

/* empole_a<Mode, PenType, use_cf>: empole.cpp:96-357 */
int orc_empole(orc_mpole_sys *s, int mode, orc_mpole_out *out);

/* charge flux (alterchg.cpp:36-169, dcflux.cpp:31-225): the monopole change
 * of every atom from its bonds and angles, and the gradient (+ virial, may
 * be NULL) of pot[i] = dE/dq_i through those changes, added to de / vir */
void orc_charge_flux_delta(const orc_mpole_sys *s, double *pdelta);
void orc_charge_flux_grad(const orc_mpole_sys *s, const double *pot, double *de, double *vir);

/* minimum image of one displacement, image.cpp:37-124 */
void orc_image(const orc_box *b, double *xr, double *yr, double *zr);

/* rotpole(Mpole) for all atoms: rotpole.cpp:22-49 (after chkpole) */
void orc_rotpole(const orc_mpole_sys *s, const double *pole, double *rpole);

/* one pair of the undamped point-multipole interaction, pairMpole.h:21-175;
 * out = {e, frc[3], trqi[3], trqk[3], vir[6]} */
void orc_pair_mpole(double r2, double xr, double yr, double zr, double mscale, double f,
                    const double *mi, const double *mk, double *out);

/* chkpole (chkpole.cpp:24-89): flips pole rows and yaxis signs in place */
void orc_chkpole(orc_mpole_sys *s, double *pole);

/* torque -> gradient on frame atoms, torque.cpp:42-397 (vir may be NULL) */
void orc_torque(const orc_mpole_sys *s, const double *te, double *de, double *vir);

/* Gordon1 damping factors (damping.h:32-196) of orders 1..9 (dmpi, dmpk: 5)
 * and 1..11 (dmpik: 6) */
void orc_damp_gordon1(double r, double ai, double ak, double *dmpi, double *dmpk, double *dmpik);

/* ---------------------------------------------------------------------
 * Part 2: Ewald multipoles and AMOEBA polarization (pmq_oracle_elec.c)
 * ------------------------------------------------------------------- */
typedef struct {
    orc_mpole_sys *mp;      /* multipoles, frames, connectivity, box, f */
    int use_ewald;
    double aewald;          /* Ewald coefficient, 1/A */
    double cut;             /* real-space cutoff, A (ewaldcut, or mpolecut without Ewald) */
    int nfft1, nfft2, nfft3, bsorder;
    int use_polar;
    const double *polarity, *thole, *pdamp; /* n each */
    const int *np11, *ip11, *np12, *ip12, *np13, *ip13, *np14, *ip14; /* polarization groups */
    double p2scale, p3scale, p4scale, p5scale;
    double p2iscale, p3iscale, p4iscale, p5iscale;
    double d1scale, d2scale, d3scale, d4scale;
    double u1scale, u2scale, u3scale, u4scale;
    double poleps;
    int politer;
    int solver;             /* 0 preconditioned CG, 1 dense direct solve */
    /* with charge penetration (mp->pentype != 0) polarization is HIPPO's:
     * direct field of core + damped valence, mutual field with overlap
     * damping and the w scales (1-2..1-5), d = p when dpequal */
    double w2scale, w3scale, w4scale, w5scale;
    int dpequal;
    /* PCG preconditioner (pcgprec of polpcg.h; usolvcut 4.5 A at
     * cutoffs.cpp:61): z = udiag a r + a T_s (a r), T_s the dipole field of
     * the pairs within usolvcut, damped and scaled as the mutual field,
     * without Ewald; pcgprec = 0 gives the diagonal z = a r (the GPU
     * solver's choice) */
    int pcgprec;
    double usolvcut, udiag;
} orc_elec_sys;

typedef struct {
    double em, ep;
    double em_real, em_recip, em_self;
    int iters;              /* CG iterations (negative if not converged) */
    double *dem, *dep;      /* 3n gradients (may be NULL) */
    double *uind, *uinp;    /* 3n induced dipoles (may be NULL) */
    /* virial of em and of ep, 9 each, row-major, the reference's convention
     * (empole.cpp:310-320, torque.cpp:361-395: W_ab = sum_i x_ia dE/dx_ib,
     * i.e. dE/d(strain) with positions and box strained); may be NULL;
     * computed only with want_grad */
    double *vir_m, *vir_p;
    /* analysis (aem, aep may be NULL; the counts and einter are always
     * formed): nem = real-space pairs with a nonzero energy + one self term
     * per multipole site with Ewald; aem[i] = half of each pair energy of i
     * + (Ewald) its reciprocal energy (f/2) M_i.phi_i and self energy;
     * einter = em - the unscreened scaled energies of the same-molecule
     * pairs within the cutoff (without Ewald: the intermolecular pair
     * energies); aep[i] = -(f/2) mu_d(i).E^p(i), nep = sites with aep != 0 */
    double *aem, *aep;
    long nem, nep;
    double einter;
} orc_elec_out;

int orc_elec(orc_elec_sys *s, int want_grad, orc_elec_out *o);

/* Thole damping factors of one pair (tests): lam[1..4] = lambda_3,5,7,9 for
 * u = r/(pdi pdk) and parameter pgamma (lam[0] = lam[5] = 1) */
void orc_thole_lambda(double r, double pdi, double pdk, double pgamma, double *lam);

/* ---------------------------------------------------------------------
 * Part 3: HIPPO dispersion and Pauli repulsion (pmq_oracle_elec.c).  The
 * reference has parameters but no energy for either (energy.cpp:279-280,
 * erepel.cpp:149-151).
 * ------------------------------------------------------------------- */
typedef struct {
    orc_mpole_sys *mp;                /* frames, connectivity, box */
    /* dispersion: E = -sum s C_i C_k f^2 / r^6, f = 1.5 d5 - 0.5 d3 */
    const double *csix, *adisp;
    double dsp2scale, dsp3scale, dsp4scale, dsp5scale;
    int use_dewald;
    double dewald, dispcut, disptaper; /* taper (absolute) used without Ewald only */
    int nfft1, nfft2, nfft3, bsorder;
    /* repulsion: E = K_i K_k s / r * M(B), B_n = D^n S^2 */
    const double *sizpr, *dmppr, *elepr;
    double r2scale, r3scale, r4scale, r5scale;
    double repcut, reptaper;
} orc_hippo_sys;

typedef struct {
    double edsp, er;
    double edsp_real, edsp_recip, edsp_self;
    double *dedsp, *der;   /* 3n (may be NULL) */
    double *vir_d, *vir_r; /* virial of edsp and of er, 9 each (may be NULL; with want_grad) */
} orc_hippo_out;

int orc_hippo(orc_hippo_sys *s, int want_grad, orc_hippo_out *o);

/* ---------------------------------------------------------------------
 * Part 4: AMOEBA buffered 14-7 van der Waals (ehal) and HIPPO charge
 * transfer (ectrn), after Tinker's published ehal/echgtrn; the reference
 * has their parameters (kvdw.cpp, kchgtrn.cpp) and no energy
 * (energy.cpp:268,276).
 *   vdW: rho = r/R (r between the reduced sites, R = radmin of the pair's
 *   classes), E = s eps (1.07/(rho+0.07))^7 (1.12/(rho^7+0.12) - 2), s the
 *   1-n vdW scale (the 1-4 tables for 1-4 pairs), switched off between
 *   vdwtaper and vdwcut; hydrogen sites moved to kred x_i + (1-kred) x_ired
 *   and their gradient split the same way.
 *   charge transfer: E = -s f (c_i e^{-a_k r} + c_k e^{-a_i r}), s the 1-n
 *   multipole scale, f = electric/dielec, a = 0 read as 1000, switched off
 *   between ctrntaper and ctrncut.
 *   A pair counts when either atom is active (mp->use); for vdW an atom
 *   also counts as active when its reduction partner is.
 * ------------------------------------------------------------------- */
typedef struct {
    orc_mpole_sys *mp;           /* coordinates, connectivity, box, m-scales, f */
    int use_vdw;
    int nv;                      /* vdW classes */
    const int *jv;               /* n: class index of each atom */
    const double *radmin, *epsilon, *radmin4, *epsilon4; /* nv x nv */
    const double *kred;          /* n reduction factors */
    const int *ired;             /* n: 0-based reduction partner (itself when none) */
    double v2scale, v3scale, v4scale, v5scale;
    double vdwcut, vdwtaper;     /* absolute */
    int use_ctrn;
    const double *chgct, *dmpct; /* n */
    double ctrncut, ctrntaper;   /* absolute */
} orc_vdwct_sys;

typedef struct {
    double ev, ect;
    double *dev, *dect;          /* 3n (may be NULL) */
    double *vir_v, *vir_c;       /* 9 each (may be NULL) */
} orc_vdwct_out;

int orc_vdwct(orc_vdwct_sys *s, int want_grad, orc_vdwct_out *o);

/* the overlap series B_0..B_5 of one pair (tests) */
void orc_rep_series(double r, double ai, double ak, double *B);

#ifdef __cplusplus
}
#endif
#endif
