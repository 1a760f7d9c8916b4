/*
 * pmq.h -- C ABI of the MI355X polarizable-electrostatics library (libpmq.so).
 *
 * The reference has no FFI: its hot path is a set of argument-less C++
 * templates over global arrays, `energy<CalcMode>()` (src/md/energy.cpp:31)
 * calling `empole<CalcMode>()` (src/md/empole.cpp:50) and, by Tinker
 * convention, `epolar/erepel/edisp<CalcMode>()` (energy.cpp:267,279,280).
 * This header is the boundary those bodies cross: plain pointers and sizes,
 * no C++ or torch types.  A host shim that keeps the reference's globals
 * (polmdqc_amd/csrc/host/) fills these structs from them and maps the return
 * codes onto the reference's fatal() (fatal.cpp:21-28).
 *
 * Conventions (identical to the reference's globals):
 *   - atoms are 0-based in arrays; frame atoms (zaxis/xaxis/yaxis) are 1-based
 *     with 0 meaning "none" and the sign of yaxis carrying chirality
 *     (mpole.h:30-41, chkpole.cpp:44-89);
 *   - pole is n*13 per atom in the local frame with kmpole's units
 *     (dipoles x bohr, quadrupoles x bohr^2/3, kmpole.cpp:543-552);
 *   - use[] has n+1 entries with use[0] = 0 (active.cpp:41,51);
 *   - gradients are AoS double[3n] in kcal/mol/A, virial double[9] row-major.
 *
 * Every function returns 0 on success or a negative PMQ_ERR_* code, and never
 * exits; pmq_last_error() describes the last failure.  One context per host
 * thread; calls on one context are not re-entrant; no caller pointer is kept
 * past a call.
 */
#ifndef PMQ_H
#define PMQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMQ_ABI_VERSION 6

#if defined(__GNUC__) || defined(__clang__)
#define PMQ_API __attribute__((visibility("default")))
#else
#define PMQ_API
#endif

enum {
    PMQ_OK = 0,
    PMQ_ERR_ARG = -1,         /* bad argument or size */
    PMQ_ERR_HIP = -2,         /* a HIP runtime call failed */
    PMQ_ERR_STATE = -3,       /* call out of order (e.g. compute before set_params) */
    PMQ_ERR_UNSUPPORTED = -4, /* option not built */
    PMQ_ERR_NAN = -5,         /* non-finite total energy (energy.cpp:329-332) */
    PMQ_ERR_NOCONV = -6       /* induced dipoles did not converge in politer steps */
};

/* CalcMode, include/md/calcMode.h:18-49 */
enum { PMQ_ENERGY = 0, PMQ_ANALYSIS = 1, PMQ_GRADIENT = 2, PMQ_VIRIAL = 3 };

/* energy terms (bit mask) */
enum {
    PMQ_TERM_MPOLE = 1u << 0, /* empole: permanent multipoles */
    PMQ_TERM_POLAR = 1u << 1, /* epolar: induced dipoles (mutual, PCG) */
    PMQ_TERM_REPEL = 1u << 2, /* erepel: HIPPO Pauli repulsion */
    PMQ_TERM_DISP = 1u << 3,  /* edisp:  HIPPO damped dispersion */
    PMQ_TERM_VDW = 1u << 4,   /* ehal:   AMOEBA buffered 14-7 van der Waals */
    PMQ_TERM_CTRN = 1u << 5   /* echgtrn: HIPPO charge transfer */
};

/* arithmetic of the pair/PME kernels */
enum {
    PMQ_PLAIN_FP64 = 0, /* FP64 throughout (parity form) */
    PMQ_MIXED = 1       /* FP32 pair/PME math, int64 fixed-point forces, FP64 reductions */
};

/* periodic box, boxes.h:46-67 */
enum { PMQ_BOX_NONE = 0, PMQ_BOX_ORTHO, PMQ_BOX_MONO, PMQ_BOX_TRI, PMQ_BOX_OCTA, PMQ_BOX_DODECA };

/* local frame types, kmulti.h:23-31 */
enum { PMQ_FRAME_NONE = 0, PMQ_FRAME_ZTHENX, PMQ_FRAME_ZONLY, PMQ_FRAME_BISECTOR, PMQ_FRAME_ZBISECT,
       PMQ_FRAME_THREEFOLD };

/* charge penetration, mplpot.h:25 */
enum { PMQ_PEN_NONE = 0, PMQ_PEN_GORDON1 = 1, PMQ_PEN_GORDON2 = 2 };

typedef struct pmq_ctx pmq_ctx;

typedef struct {
    int device; /* HIP device ordinal */
} pmq_config;

typedef struct {
    int kind;                  /* PMQ_BOX_* */
    double xbox, ybox, zbox;   /* A */
    double alpha, beta, gamma; /* degrees */
} pmq_box;

/* Static topology: replaces what mechanic() leaves in couple.h, mpole.h,
 * usage.h, molcul.h, polgrp.h, bndstr.h and angbnd.h. */
typedef struct {
    int n;
    const int *polaxe;                 /* n, PMQ_FRAME_* */
    const int *zaxis, *xaxis, *yaxis;  /* n, 1-based, signed yaxis */
    const int *use;                    /* n+1, or NULL = every atom active */
    const int *molcule;                /* n, or NULL = one molecule */
    /* 1-2 .. 1-5 connectivity: counts n1x[n] and flat 0-based lists */
    const int *n12, *i12, *n13, *i13, *n14, *i14, *n15, *i15;
    /* polarization groups (polgrp.h:28-39), counts and flat lists, may be NULL */
    const int *np11, *ip11, *np12, *ip12, *np13, *ip13, *np14, *ip14;
    /* charge-flux geometry (bndstr.h, angbnd.h, atmlst.h) */
    int nbond;
    const int *ibnd;   /* 2*nbond */
    int nangle;
    const int *iang;   /* 3*nangle */
    const int *balist; /* 2*nangle */
} pmq_topology;

/* Parameters: what kmpole/kpolar/kchgflx/kewald/cutoffs assign. */
typedef struct {
    const double *pole; /* n*13 */
    double m2scale, m3scale, m4scale, m5scale;
    double electric, dielec;
    /* charge penetration (chgpen.h) */
    int pentype;
    const double *pcore, *pval0, *palpha; /* n each */
    /* charge flux (cflux.h) */
    int use_chgflx;
    const double *bl, *bflx;            /* nbond */
    const double *anat, *aflx, *abflx;  /* nangle, 2*nangle, 2*nangle */
    /* polarization (polar.h, polpot.h) */
    const double *polarity, *thole, *pdamp; /* n each */
    double p2scale, p3scale, p4scale, p5scale;
    double p2iscale, p3iscale, p4iscale, p5iscale;
    double d1scale, d2scale, d3scale, d4scale;
    double u1scale, u2scale, u3scale, u4scale;
    double poleps; /* rms induced-dipole change, Debye */
    int politer;
    /* method and cutoffs (limits.h, ewald.h, pme.h) */
    int use_ewald;
    double mpolecut; /* non-Ewald multipole cutoff, A */
    double ewaldcut; /* Ewald real-space cutoff, A */
    double aewald;   /* Ewald coefficient, 1/A */
    int nfft1, nfft2, nfft3, bsorder;
    /* HIPPO polarization: with charge penetration (pentype != NONE) the
     * polarization uses HIPPO damping, as kpolar.cpp:577 turns Thole off
     * when chgpen sites exist: the direct field sees the core of each site
     * undamped and its valence multipole damped by the one-site function of
     * its palpha; the dipole-dipole field is damped by the overlap of the two
     * valence densities and scaled by w (1-2..1-5, polpot.h:47-66); dpequal
     * (d-equals-p, prmkey.cpp:531) makes the direct-field scales the p ones. */
    double w2scale, w3scale, w4scale, w5scale;
    int dpequal;
    /* HIPPO dispersion (kdsp.h:18-19, dsppot.h:21-25, ewald.h, pme.h):
     * E = -sum s C_i C_k f^2 / r^6, f = 1.5 d5 - 0.5 d3 of the adisp
     * exponents; with use_dewald real space + smooth PME (grid dfft, order
     * dbsorder, which must equal the multipole grid when both are on) +
     * self.  disptaper (absolute, A) starts the quintic switch to dispcut
     * when dispersion Ewald is off. */
    const double *csix, *adisp;  /* n each (NULL = no dispersion) */
    double dsp2scale, dsp3scale, dsp4scale, dsp5scale;
    int use_dewald;
    double dewald, dispcut, disptaper;
    int dfft1, dfft2, dfft3, dbsorder;
    /* HIPPO Pauli repulsion (repel.h:25-32, reppot.h:20-23, krepel.cpp:92-153):
     * size, damping exponent and valence count (stored negative) per atom;
     * switched off between reptaper (absolute) and repcut (cutoffSwitch.cpp). */
    const double *sizpr, *dmppr, *elepr; /* n each (NULL = no repulsion) */
    double r2scale, r3scale, r4scale, r5scale;
    double repcut, reptaper;
    /* PCG preconditioner (polpcg.h:21-25, kpolar.cpp:84, cutoffs.cpp:61; the
     * reference declares the flag and the cutoff but has no solver, so udiag
     * follows the published Tinker method, default 2):
     * with pcgprec the residual r is preconditioned as
     * z = udiag a r + a T_s (a r), a = max(polarity, 1e-8) and T_s the
     * damped, u-scaled (w-scaled with charge penetration) real-space
     * dipole field tensor without Ewald of the pairs within usolvcut (taken
     * no longer than the real-space cutoff), and the solve starts from
     * a (E + T_s a E) instead of a E; without it z = a r.  Batched clusters
     * take either. */
    int pcgprec;
    double usolvcut, udiag;
    /* AMOEBA buffered 14-7 van der Waals (vdw.h, vdwpot.h, kvdw.cpp; the
     * reference assigns these and has no energy, energy.cpp:268; the form is
     * Tinker's published ehal):
     *   rho = r/R, E = s eps (1.07/(rho + 0.07))^7 (1.12/(rho^7 + 0.12) - 2),
     * r between the reduced sites (atom i at kred_i (x_i - x_ired) + x_ired),
     * R and eps from the nvdw x nvdw tables of the two atoms' classes (the
     * 1-4 tables for pairs listed 1-4), s the v2..v5 scale of the pair's
     * connectivity (later classes override earlier, as the m scales),
     * switched off between vdwtaper and vdwcut (absolute, A).  A pair counts
     * when either reduced site is active (the atom or its partner in use[]).
     * The gradient of a reduced site goes kred to the atom and 1 - kred to
     * its partner.  vdwcut may exceed the electrostatic cutoff: the term
     * builds its own pair structure over the reduced sites. */
    int nvdw;                  /* classes (0 = no vdW) */
    const int *vdwclass;       /* n: class index 0..nvdw-1 */
    const double *radmin, *epsilon, *radmin4, *epsilon4; /* nvdw*nvdw, row-major */
    const double *kred;        /* n reduction factors */
    const int *ired;           /* n: 0-based reduction partner (the atom itself: none) */
    double v2scale, v3scale, v4scale, v5scale;
    double vdwcut, vdwtaper;
    /* HIPPO charge transfer (ctrpot.h, kchgtrn.cpp; the reference assigns
     * these and has no energy, energy.cpp:276; Tinker's published echgtrn):
     *   E = -s f (chgct_i e^{-a_k r} + chgct_k e^{-a_i r}),
     * a = dmpct (0 read as 1000), f = electric/dielec, s the m2..m5 scale of
     * the pair, switched off between ctrntaper and ctrncut (absolute, A;
     * ctrncut may not exceed the electrostatic cutoff when multipoles or
     * polarization run in a periodic box).  A pair counts when either atom
     * is active. */
    const double *chgct, *dmpct; /* n each (NULL = no charge transfer) */
    double ctrncut, ctrntaper;
} pmq_params;

typedef struct {
    /* energies, kcal/mol */
    double em, ep, er, edsp, esum;
    /* analysis of em (empole.cpp:285-292 without Ewald; with Ewald, as
     * Tinker's Ewald analysis is organized): nem counts the real-space
     * pairs with a nonzero energy, plus one self term per multipole site
     * with Ewald; aem gives each atom half of each of its pair energies and,
     * with Ewald, its own reciprocal (f/2 M.phi) and self energy, so aem sums
     * to em; einter is em minus the unscreened, scaled energies of the pairs
     * within the cutoff inside one molecule (without Ewald: the sum of the
     * intermolecular pair energies) */
    double einter;     /* intermolecular part of em (analysis) */
    long long nem;     /* multipole interactions counted (analysis) */
    double vir[9];     /* row-major */
    int pcg_iters;     /* iterations the induced-dipole solve took */
    /* optional outputs (NULL = not wanted) */
    double *dem;   /* 3n gradient of em */
    double *dep;   /* 3n gradient of ep */
    double *desum; /* 3n total gradient */
    double *aem;   /* n per-atom em partition (analysis) */
    double *rpole; /* 13n global-frame multipoles */
    double *uind, *uinp; /* 3n induced dipoles (d and p fields) */
    double *der;   /* 3n gradient of er (HIPPO repulsion) */
    double *dedsp; /* 3n gradient of edsp (HIPPO dispersion) */
    /* analysis of the polarization energy (no counterpart in the reference,
     * which has no polarization energy): ep = sum of aep, aep[i] =
     * -(f/2) mu_d(i) . E^p(i), nep = the sites with a nonzero aep */
    long long nep;
    double *aep;   /* n per-atom ep partition (analysis) */
    /* van der Waals and charge transfer (esum and desum include them) */
    double ev, ect;
    double *dev;   /* 3n gradient of ev */
    double *dect;  /* 3n gradient of ect */
} pmq_result;

PMQ_API int pmq_create(const pmq_config *cfg, pmq_ctx **ctx);
PMQ_API int pmq_destroy(pmq_ctx *ctx);
PMQ_API const char *pmq_last_error(const pmq_ctx *ctx);
PMQ_API int pmq_abi_version(void);

/* Static state, called after mechanic() and whenever topology changes. */
PMQ_API int pmq_set_topology(pmq_ctx *ctx, const pmq_topology *top);
PMQ_API int pmq_set_params(pmq_ctx *ctx, const pmq_params *prm);

/* One full evaluation from host coordinates into host outputs:
 * upload x,y,z and box, compute `terms` in `mode` with `precision`, fetch. */
PMQ_API int pmq_eval(pmq_ctx *ctx, int mode, unsigned terms, int precision,
             const double *x, const double *y, const double *z, const pmq_box *box, pmq_result *res);

/* The same evaluation split so coordinates can stay resident in HBM:
 * pmq_set_coords uploads, pmq_compute only enqueues device work on the
 * context's stream (asynchronous), pmq_fetch waits and copies results out. */
PMQ_API int pmq_set_coords(pmq_ctx *ctx, const double *x, const double *y, const double *z, const pmq_box *box);
PMQ_API int pmq_compute(pmq_ctx *ctx, int mode, unsigned terms, int precision);
PMQ_API int pmq_fetch(pmq_ctx *ctx, pmq_result *res);
PMQ_API int pmq_sync(pmq_ctx *ctx);

/* Parameter gradients of the last evaluation (plain FP64, GRADIENT, one
 * system; no charge penetration or flux; or a batch, see
 * pmq_set_batch_param_grad).  No reference counterpart: the
 * reference has no derivatives with respect to parameters.  Per atom, in
 * caller order: dE/dpole[13n] with respect to the local-frame components as
 * passed in pmq_params.pole (each stored component on its own), and, with
 * polarization, dE/dpolarity[n] at fixed pdamp, dE/dpdamp[n] and dE/dthole[n]
 * (the Thole parameter of a pair is the smaller of the two, so equal values
 * share the derivative half and half).  E is em + ep of that evaluation.  The
 * polarization part uses the variational form (mu_p is the adjoint of mu_d):
 * no further solve.  NULL outputs are skipped. */
PMQ_API int pmq_param_grad(pmq_ctx *ctx, double *dpole, double *dpolarity, double *dpdamp, double *dthole);

/* Batched clusters.  No counterpart in the reference: the batched workload of
 * this path (many small non-periodic systems).  After pmq_set_batch the
 * context's system is a set of independent clusters, cluster c being atoms
 * [offsets[c], offsets[c+1]) (at most 256 atoms each); no pair across clusters
 * interacts, no cutoff applies, and excluded pairs must lie inside a cluster.
 * pmq_compute (ENERGY or GRADIENT, terms MPOLE|POLAR, either precision, no
 * charge penetration or flux) then solves every cluster, each to its own PCG
 * convergence, in one persistent launch; the evaluation is captured in a HIP
 * graph after the first call and replayed.  Totals and gradients come from
 * pmq_fetch as for one system; pmq_fetch_batch gives the per-cluster energies
 * and PCG iteration counts (negative: not converged).  nsys = 0 returns the
 * context to one system. */
PMQ_API int pmq_set_batch(pmq_ctx *ctx, int nsys, const int *offsets);
PMQ_API int pmq_fetch_batch(pmq_ctx *ctx, double *em, double *ep, int *pcg_iters);

/* Parameter gradients of batched clusters.  No counterpart in the reference
 * (it has no parameter derivatives).  With enable = 1 every GRADIENT
 * evaluation of a batch, in either precision, also computes dE/d(parameters)
 * of every atom inside the same persistent launch (row-owned sums over each
 * cluster's pairs, the scaled pairs corrected afterwards; the mixed form
 * evaluates the pair terms in FP32 and sums in FP64) and the per-site step
 * in the same graph; pmq_param_grad then returns them, with the meaning and
 * layout it has for one system (E = em + ep of all clusters; each atom's
 * derivatives come from its own cluster).  pmq_compute_batch_noalloc does
 * not compute them. */
PMQ_API int pmq_set_batch_param_grad(pmq_ctx *ctx, int enable);

/* The batched-cluster autotuner's choice.  No counterpart in the reference.
 * Clusters are bucketed by size (more than 128 atoms, more than 64, the rest;
 * workgroups of 256, 128 and 64 rows), each bucket one persistent launch, the
 * three side by side on forked streams.  On the first evaluation of a batch
 * in each precision the tuner walks the buckets (largest clusters first) and
 * times every admissible shape of one bucket -- lanes per row (1-2, 1-4, 1-8
 * by bucket) times 1, 2 or 4 persistent workgroups per compute unit -- on the
 * whole batch with the other buckets at their current shapes, keeping the
 * fastest; then one merged launch of all clusters in 256-row workgroups is
 * timed in each of its shapes and kept if faster (reported as bucket 0
 * holding every cluster).  Per bucket: rows per workgroup, lanes per row, workgroups per CU,
 * the bucket's cluster count and the whole batch's timed launch in ms when
 * that bucket's shape was chosen (-1 where no bucket or not yet tuned).
 * Arrays of PMQ_BATCH_BUCKETS; NULL outputs are skipped. */
#define PMQ_BATCH_BUCKETS 3
PMQ_API int pmq_batch_tuning(pmq_ctx *ctx, int precision, int *rows, int *lanes, int *wgs_per_cu, int *nclusters,
                             float *ms);

/* Fix one bucket's shape instead of timing it (for tests and reproducible
 * runs): lanes per row and workgroups per CU among those the bucket admits
 * (PMQ_ERR_ARG otherwise); lanes = 0 returns the bucket to the autotuner.
 * A fixed shape keeps the buckets apart (no merged launch).  Kept across
 * pmq_set_batch calls.  No counterpart in the reference. */
PMQ_API int pmq_set_batch_shape(pmq_ctx *ctx, int precision, int bucket, int lanes, int wgs_per_cu);

/* The batched evaluation without allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory while they
 * run (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  After pmq_set_topology, pmq_set_params, pmq_set_coords
 * and pmq_set_batch (which allocate), this enqueues the same evaluation as
 * pmq_compute on a batch -- same kernels, the lanes-per-row variant a previous
 * pmq_compute tuned, else the first -- on the context's stream, calling no
 * hipMalloc, hipFree or allocating routine.  Its scratch (work-queue counter,
 * per-cluster energies and iteration counts, energy totals) lives in one
 * fixed device array per GPU, held until the process ends, for at most
 * PMQ_NOALLOC_MAX_CLUSTERS (65536) clusters; larger batches return
 * PMQ_ERR_ARG.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (two contexts, two threads or streams) race.  Outputs
 * are caller DEVICE pointers, each may be NULL: em, ep and pcg_iters per
 * cluster (nsys), etot = {em, ep} totals (2), grad = 3n total gradient
 * (dem + dep, GRADIENT only).  The call is asynchronous: pmq_sync waits for
 * it.  pmq_fetch and pmq_fetch_batch do not see its results. */
#define PMQ_NOALLOC_MAX_CLUSTERS 65536
PMQ_API int pmq_compute_batch_noalloc(pmq_ctx *ctx, int mode, unsigned terms, int precision, double *em, double *ep,
                                      int *pcg_iters, double *etot, double *grad);

/* Kernel timing for measurement.  No counterpart in the reference (it has no
 * timers on this path); for benchmark and profiling harnesses.  While
 * enabled, the context brackets every launch of each instrumented kernel
 * family with HIP events on its own stream.  pmq_profile_read waits for the
 * stream, returns the summed elapsed time and launch count of one family
 * ("neighbors", "mpole_tiles", "rep_tiles", "disp_tiles", "pol_field_tiles",
 * "pol_pair_fill", "pol_mutual_spmv" (the pair-matrix product T p alone),
 * "pol_spmv_spread" (T p and the spreading of p in one launch, with Ewald),
 * "pol_grad_tiles", "pme_spread" (the
 * two induced-dipole sets), "pme_spread_mpole" (multipoles or dispersion
 * coefficients), "pme_fft" (one two-set convolution: forward transforms,
 * kernel, inverse transforms), "pme_fft1" (a one-set convolution),
 * "pme_gather_field" (the induced-dipole fields of a PCG step), "pme_gather"
 * (the other gathers and their per-atom post-processing), "batch_clusters")
 * and resets it. */
PMQ_API int pmq_profile_enable(pmq_ctx *ctx, int on);
PMQ_API int pmq_profile_read(pmq_ctx *ctx, const char *name, double *total_ms, long long *launches);

/* PME transform engine.  No counterpart in the reference (it has no FFT
 * code); for tuning and tests.  PMQ_FFT_AUTO (the default) times, once per
 * grid, rocFFT against the direct LDS transforms (three launches per
 * convolution, the reciprocal kernel applied between the z transforms;
 * grids whose planes fit in LDS) and keeps the faster, separately for FP64
 * and FP32 grids (a profiler that adds a cost per launch biases that
 * choice); PMQ_FFT_LIBRARY always takes rocFFT; PMQ_FFT_DIRECT takes the
 * direct transforms wherever they fit.  One of PMQ_FFT_GRIDS_FP32 and
 * PMQ_FFT_GRIDS_FP64 may be or'ed into the engine to fix the grid precision
 * of the mixed form (timed against each other otherwise).  pmq_fft_choice
 * reports the choice made for the current grid: direct transforms for FP64
 * grids, FP32 grids in the mixed form, direct transforms for FP32 grids
 * (each 0/1, or -1 while not yet chosen). */
#define PMQ_FFT_AUTO 0
#define PMQ_FFT_LIBRARY 1
#define PMQ_FFT_DIRECT 2
#define PMQ_FFT_GRIDS_FP32 16
#define PMQ_FFT_GRIDS_FP64 32
PMQ_API int pmq_set_fft(pmq_ctx *ctx, int engine);
PMQ_API int pmq_fft_choice(pmq_ctx *ctx, int *fp64_direct, int *fp32_grids, int *fp32_direct);

/* PME gather lanes.  No counterpart in the reference; for tuning and tests.
 * PMQ_GATHER_AUTO (the default) gives each atom 8 lanes, one stencil plane
 * each, from 12,288 atoms on and one lane per stencil row below; the other
 * two values force one layout. */
#define PMQ_GATHER_AUTO 0
#define PMQ_GATHER_ROWS 1
#define PMQ_GATHER_PLANES 2
PMQ_API int pmq_set_gather_lanes(pmq_ctx *ctx, int mode);

/* Pair algebra of the multipole tiles.  PMQ_PAIR_CARTESIAN (the default)
 * evaluates each pair as pairMpole.h:21-175 does; PMQ_PAIR_QI rotates each
 * pair into its quasi-internal frame (z along the pair vector) first, the
 * formulation of pairMpoleQI.h:22-278 (qi.h:11-64), which the reference
 * includes but never calls.  The two give the same result up to rounding;
 * the scaled (excluded) pairs always take the Cartesian form. */
#define PMQ_PAIR_CARTESIAN 0
#define PMQ_PAIR_QI 1
PMQ_API int pmq_set_pair_form(pmq_ctx *ctx, int form);

/* HIP graphs for the evaluations of one system (no counterpart in the
 * reference, which has no GPU path).  On: once an evaluation runs without
 * host round trips, the second one with the same mode, terms, precision,
 * box and PCG batch is captured into a graph and later ones replay it; any
 * setter other than pmq_set_coords with an unchanged box, and any
 * evaluation that has to size buffers, starts over.  Off (the default):
 * every evaluation is enqueued launch by launch, which is faster once the
 * permanent multipoles' reciprocal part runs on a second stream beside the
 * pair kernels (a graph with that branch replays more slowly than the same
 * launches enqueued on two streams).  Results are the same. */
PMQ_API int pmq_set_graphs(pmq_ctx *ctx, int enable);

#ifdef __cplusplus
}
#endif
#endif
