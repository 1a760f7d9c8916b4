// pmq_host.hpp -- the reference's operator interface for this path, in C++,
// on top of the C ABI of libpmq.so.
//
// The reference computes these terms in argument-less templates over global
// module arrays: energy<CalcMode>() (src/md/energy.cpp:31-333) calls
// empole<CalcMode>() (src/md/empole.cpp:49-83) and, by Tinker convention,
// epolar<CalcMode>() (commented out at energy.cpp:267).  This header keeps that
// shape: the globals are the members of `pmq_host::g` (names and layouts as
// in include/md/atoms.h, mpole.h, couple.h, usage.h, molcul.h, polar.h,
// polpot.h, polgrp.h, mplpot.h, chgpen.h, boxes.h, ewald.h, pme.h, limits.h,
// energi.h, deriv.h, virial.h, analyz.h, action.h, inter.h), and the
// templates read the inputs from them and write the outputs back, on one GPU.
//
// Errors: every failure of the library ends in pmq_host::fatal(), which
// prints the library's message and exits like the reference's fatal()
// (src/md/fatal.cpp:21-28); callers that prefer a return code use
// pmq_host::try_energy().
#pragma once

#include <string>
#include <vector>

namespace pmq_host {

// include/md/calcMode.h:18-49
enum class CalcMode { Energy, Analysis, Gradient, Virial };

struct Globals {
    // atoms.h
    int n = 0;
    std::vector<double> x, y, z;
    // mpole.h (pole: n x 13, local frame, kmpole units)
    std::vector<double> pole;
    std::vector<int> polaxe, zaxis, xaxis, yaxis; // frame atoms 1-based, 0 = none, signed yaxis
    // usage.h (n + 1 entries, use[0] = 0), molcul.h
    std::vector<int> use, molcule;
    // couple.h: per-atom lists (0-based)
    std::vector<std::vector<int>> i12, i13, i14, i15;
    // polgrp.h
    std::vector<std::vector<int>> ip11, ip12, ip13, ip14;
    // mplpot.h, chgpen.h
    double m2scale = 0, m3scale = 0, m4scale = 0.4, m5scale = 0.8;
    int pentype = 0;
    std::vector<double> pcore, pval0, palpha;
    // cflux.h, bndstr.h, angbnd.h, atmlst.h (charge flux); mono0: permanent monopoles before flux
    bool use_chgflx = false;
    std::vector<double> mono0, bl, bflx, anat, aflx, abflx;
    std::vector<int> ibnd, iang, balist; // 2 x nbond, 3 x nangle, 2 x nangle (0-based)
    // units / chgpot.h
    double electric = 332.0637133, dielec = 1.0;
    // polar.h, polpot.h
    std::vector<double> polarity, thole, pdamp;
    double p2scale = 0, p3scale = 0, p4scale = 1, p5scale = 1;
    double p2iscale = 0, p3iscale = 0, p4iscale = 0.5, p5iscale = 1;
    double d1scale = 0, d2scale = 1, d3scale = 1, d4scale = 1;
    double u1scale = 1, u2scale = 1, u3scale = 1, u4scale = 1;
    double w2scale = 1, w3scale = 1, w4scale = 1, w5scale = 1;   // initprm.cpp:390-393
    bool dpequal = false;
    double poleps = 1e-6;
    int politer = 100;
    bool pcgprec = true;    // polpcg.h:24, default kpolar.cpp:84
    double usolvcut = 4.5;  // cutoffs.cpp:61
    double udiag = 2.0;     // diagonal factor of the sparse preconditioner (the reference declares pcgprec, not the solver)
    // HIPPO dispersion: kdsp.h (csix, adisp per atom), dsppot.h, ewald.h, pme.h, limits.h
    std::vector<double> csix, adisp;
    double dsp2scale = 0, dsp3scale = 0, dsp4scale = 1, dsp5scale = 1; // initprm.cpp:338-341
    bool use_dewald = false;
    double dewald = 0, dewaldcut = 7.0, dispcut = 9.0, disptaper = 0.9; // taper < 1: fraction of the cutoff
    int dfft1 = 0, dfft2 = 0, dfft3 = 0, dbsorder = 5;
    // HIPPO repulsion: repel.h, reppot.h
    std::vector<double> sizpr, dmppr, elepr;
    double r2scale = 0, r3scale = 0, r4scale = 1, r5scale = 1;       // initprm.cpp:332-335
    double repcut = 6.0, reptaper = 0.9;                 // cutoffs.cpp:57,65
    // potent.h
    bool use_mpole = true, use_polar = false, use_repel = false, use_disp = false;
    // boxes.h
    bool use_bounds = false, monoclinic = false, triclinic = false, octahedron = false, dodecadron = false;
    double xbox = 0, ybox = 0, zbox = 0, alpha = 90, beta = 90, gamma = 90;
    // ewald.h, pme.h, limits.h
    bool use_ewald = false;
    double aewald = 0, mpolecut = 9.0, ewaldcut = 7.0;
    int nfft1 = 0, nfft2 = 0, nfft3 = 0, bsorder = 5;
    // outputs: energi.h, deriv.h (n x 3), virial.h, analyz.h, action.h, inter.h, polar.h
    double em = 0, ep = 0, er = 0, edsp = 0, esum = 0, einter = 0;
    int nem = 0;
    std::vector<double> dem, dep, der, dedsp, desum, aem, uind, uinp;
    double vir[3][3] = {{0}};
    int pcg_iters = 0;
    // precision of the GPU evaluation (PMQ_PLAIN_FP64 = 0, PMQ_MIXED = 1)
    int precision = 0;
};

extern Globals g;

// call after changing anything but x, y, z (the reference's mechanic())
void mechanic_changed();

template <CalcMode M> void empole();
template <CalcMode M> void epolar();
template <CalcMode M> void erepel();
template <CalcMode M> void edisp();
template <CalcMode M> void energy();

// energy<M>() with the library's status instead of an exit
int try_energy(CalcMode m, std::string *err);

[[noreturn]] void fatal(const std::string &msg);

// release the GPU context
void shutdown();

} // namespace pmq_host
