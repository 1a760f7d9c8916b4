// pmq_host.cpp -- energy/empole/epolar<CalcMode>() over pmq_host::g, through
// the C ABI (include/pmq.h).  See pmq_host.hpp.
#include "pmq_host.hpp"

#include <cstdio>
#include <cstdlib>

#include "../../include/pmq.h"

namespace pmq_host {

Globals g;

namespace {

pmq_ctx *ctx = nullptr;
bool uploaded = false;

void check(int rc, const char *what)
{
    if (rc != PMQ_OK) fatal(std::string(what) + ": " + (ctx ? pmq_last_error(ctx) : "no context"));
}

void flatten(const std::vector<std::vector<int>> &lists, int n, std::vector<int> &cnt, std::vector<int> &flat)
{
    cnt.assign(n, 0);
    flat.clear();
    for (int i = 0; i < n && i < (int)lists.size(); i++) {
        cnt[i] = (int)lists[i].size();
        flat.insert(flat.end(), lists[i].begin(), lists[i].end());
    }
    if (flat.empty()) flat.push_back(0);
}

pmq_box box_of(const Globals &s)
{
    pmq_box b{};
    if (!s.use_bounds) b.kind = PMQ_BOX_NONE;
    else if (s.octahedron) b.kind = PMQ_BOX_OCTA;
    else if (s.dodecadron) b.kind = PMQ_BOX_DODECA;
    else if (s.triclinic) b.kind = PMQ_BOX_TRI;
    else if (s.monoclinic) b.kind = PMQ_BOX_MONO;
    else b.kind = PMQ_BOX_ORTHO;
    b.xbox = s.xbox; b.ybox = s.ybox; b.zbox = s.zbox;
    b.alpha = s.alpha; b.beta = s.beta; b.gamma = s.gamma;
    return b;
}

// topology and parameters from the globals (once per mechanic_changed())
int upload(std::string *err)
{
    int rc;
    if (!ctx) {
        pmq_config cfg{0};
        if (const char *d = std::getenv("PMQ_DEVICE")) cfg.device = std::atoi(d);
        if ((rc = pmq_create(&cfg, &ctx))) {
            if (err) *err = "pmq_create failed";
            return rc;
        }
    }
    const int n = g.n;
    std::vector<int> c12, l12, c13, l13, c14, l14, c15, l15, p11, q11, p12, q12, p13, q13, p14, q14;
    flatten(g.i12, n, c12, l12); flatten(g.i13, n, c13, l13);
    flatten(g.i14, n, c14, l14); flatten(g.i15, n, c15, l15);
    flatten(g.ip11, n, p11, q11); flatten(g.ip12, n, p12, q12);
    flatten(g.ip13, n, p13, q13); flatten(g.ip14, n, p14, q14);
    pmq_topology t{};
    t.n = n;
    t.polaxe = g.polaxe.data(); t.zaxis = g.zaxis.data(); t.xaxis = g.xaxis.data(); t.yaxis = g.yaxis.data();
    t.use = g.use.empty() ? nullptr : g.use.data();
    t.molcule = g.molcule.empty() ? nullptr : g.molcule.data();
    t.n12 = c12.data(); t.i12 = l12.data(); t.n13 = c13.data(); t.i13 = l13.data();
    t.n14 = c14.data(); t.i14 = l14.data(); t.n15 = c15.data(); t.i15 = l15.data();
    const bool groups = !g.ip11.empty();
    if (groups) {
        t.np11 = p11.data(); t.ip11 = q11.data(); t.np12 = p12.data(); t.ip12 = q12.data();
        t.np13 = p13.data(); t.ip13 = q13.data(); t.np14 = p14.data(); t.ip14 = q14.data();
    }
    std::vector<int> ibnd = g.ibnd, iang = g.iang, balist = g.balist;
    if (g.use_chgflx) {
        t.nbond = (int)g.ibnd.size() / 2;
        t.nangle = (int)g.iang.size() / 3;
        if (ibnd.empty()) ibnd.push_back(0);
        if (iang.empty()) iang.push_back(0);
        t.ibnd = ibnd.data();
        t.iang = iang.data();
        t.balist = balist.empty() ? nullptr : balist.data();
    }
    if ((rc = pmq_set_topology(ctx, &t))) return rc;
    pmq_params p{};
    // with charge flux the monopoles are recomputed from mono0 every evaluation
    std::vector<double> pole = g.pole;
    if (g.use_chgflx && (int)g.mono0.size() == n)
        for (int i = 0; i < n; i++) pole[13 * (size_t)i] = g.mono0[i];
    p.pole = pole.data();
    p.use_chgflx = g.use_chgflx;
    if (g.use_chgflx) {
        p.bl = g.bl.data(); p.bflx = g.bflx.data();
        p.anat = g.anat.data(); p.aflx = g.aflx.data(); p.abflx = g.abflx.data();
    }
    p.m2scale = g.m2scale; p.m3scale = g.m3scale; p.m4scale = g.m4scale; p.m5scale = g.m5scale;
    p.electric = g.electric; p.dielec = g.dielec;
    p.pentype = g.pentype;
    if (g.pentype) { p.pcore = g.pcore.data(); p.pval0 = g.pval0.data(); p.palpha = g.palpha.data(); }
    if (!g.polarity.empty()) { p.polarity = g.polarity.data(); p.thole = g.thole.data(); p.pdamp = g.pdamp.data(); }
    p.p2scale = g.p2scale; p.p3scale = g.p3scale; p.p4scale = g.p4scale; p.p5scale = g.p5scale;
    p.p2iscale = g.p2iscale; p.p3iscale = g.p3iscale; p.p4iscale = g.p4iscale; p.p5iscale = g.p5iscale;
    p.d1scale = g.d1scale; p.d2scale = g.d2scale; p.d3scale = g.d3scale; p.d4scale = g.d4scale;
    p.u1scale = g.u1scale; p.u2scale = g.u2scale; p.u3scale = g.u3scale; p.u4scale = g.u4scale;
    p.w2scale = g.w2scale; p.w3scale = g.w3scale; p.w4scale = g.w4scale; p.w5scale = g.w5scale;
    p.dpequal = g.dpequal;
    p.poleps = g.poleps; p.politer = g.politer;
    p.pcgprec = g.pcgprec ? 1 : 0; p.usolvcut = g.usolvcut; p.udiag = g.udiag;
    auto taper = [](double t, double cut) { return t < 1.0 ? t * cut : t; }; // cutoffs.cpp:259-261
    if (!g.csix.empty()) {
        p.csix = g.csix.data(); p.adisp = g.adisp.data();
        p.dsp2scale = g.dsp2scale; p.dsp3scale = g.dsp3scale; p.dsp4scale = g.dsp4scale; p.dsp5scale = g.dsp5scale;
        p.use_dewald = g.use_dewald; p.dewald = g.dewald;
        p.dispcut = g.use_dewald ? g.dewaldcut : g.dispcut; // cutoffs.cpp:254-256
        p.disptaper = taper(g.disptaper, p.dispcut);
        p.dfft1 = g.dfft1 ? g.dfft1 : g.nfft1; p.dfft2 = g.dfft2 ? g.dfft2 : g.nfft2;
        p.dfft3 = g.dfft3 ? g.dfft3 : g.nfft3; p.dbsorder = g.dbsorder;
    }
    if (!g.sizpr.empty()) {
        p.sizpr = g.sizpr.data(); p.dmppr = g.dmppr.data(); p.elepr = g.elepr.data();
        p.r2scale = g.r2scale; p.r3scale = g.r3scale; p.r4scale = g.r4scale; p.r5scale = g.r5scale;
        p.repcut = g.repcut; p.reptaper = taper(g.reptaper, g.repcut);
    }
    p.use_ewald = g.use_ewald; p.mpolecut = g.mpolecut; p.ewaldcut = g.ewaldcut; p.aewald = g.aewald;
    p.nfft1 = g.nfft1; p.nfft2 = g.nfft2; p.nfft3 = g.nfft3; p.bsorder = g.bsorder;
    if ((rc = pmq_set_params(ctx, &p))) return rc;
    uploaded = true;
    return PMQ_OK;
}

int mode_of(CalcMode m)
{
    switch (m) {
    case CalcMode::Energy: return PMQ_ENERGY;
    case CalcMode::Analysis: return PMQ_ANALYSIS;
    case CalcMode::Gradient: return PMQ_GRADIENT;
    default: return PMQ_VIRIAL;
    }
}

// one evaluation of `terms`; outputs are added to the globals like the
// reference's term routines add into em/dem/vir (energy() zeroes them first)
int run(CalcMode m, unsigned terms, std::string *err)
{
    int rc;
    if (!uploaded && (rc = upload(err))) {
        if (err && ctx) *err = pmq_last_error(ctx);
        return rc;
    }
    const int n = g.n;
    const bool grad = m == CalcMode::Gradient || m == CalcMode::Virial;
    std::vector<double> dm(3 * (size_t)n), dp(3 * (size_t)n), am(n), ui(3 * (size_t)n), up(3 * (size_t)n);
    std::vector<double> dr(3 * (size_t)n), dd(3 * (size_t)n);
    pmq_result r{};
    r.dem = dm.data(); r.dep = dp.data(); r.aem = am.data(); r.uind = ui.data(); r.uinp = up.data();
    r.der = dr.data(); r.dedsp = dd.data();
    const pmq_box b = box_of(g);
    rc = pmq_eval(ctx, mode_of(m), terms, g.precision, g.x.data(), g.y.data(), g.z.data(), &b, &r);
    if (rc) {
        if (err) *err = pmq_last_error(ctx);
        return rc;
    }
    if (terms & PMQ_TERM_MPOLE) {
        g.em += r.em;
        if (m == CalcMode::Analysis) {
            g.nem += (int)r.nem;
            g.einter += r.einter;
            g.aem.resize(n, 0.0);
            for (int i = 0; i < n; i++) g.aem[i] += am[i];
        }
        if (grad) {
            g.dem.resize(3 * (size_t)n, 0.0);
            for (size_t j = 0; j < 3 * (size_t)n; j++) g.dem[j] += dm[j];
        }
    }
    if (terms & PMQ_TERM_POLAR) {
        g.ep += r.ep;
        g.pcg_iters = r.pcg_iters;
        g.uind = ui;
        g.uinp = up;
        if (grad) {
            g.dep.resize(3 * (size_t)n, 0.0);
            for (size_t j = 0; j < 3 * (size_t)n; j++) g.dep[j] += dp[j];
        }
    }
    if (terms & PMQ_TERM_REPEL) {
        g.er += r.er;
        if (grad) {
            g.der.resize(3 * (size_t)n, 0.0);
            for (size_t j = 0; j < 3 * (size_t)n; j++) g.der[j] += dr[j];
        }
    }
    if (terms & PMQ_TERM_DISP) {
        g.edsp += r.edsp;
        if (grad) {
            g.dedsp.resize(3 * (size_t)n, 0.0);
            for (size_t j = 0; j < 3 * (size_t)n; j++) g.dedsp[j] += dd[j];
        }
    }
    if (m == CalcMode::Virial)
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) g.vir[a][c] += r.vir[3 * a + c];
    return PMQ_OK;
}

void zero_outputs()
{
    const size_t n3 = 3 * (size_t)g.n;
    g.em = g.ep = g.er = g.edsp = g.esum = g.einter = 0.0;
    g.nem = 0;
    g.dem.assign(n3, 0.0);
    g.dep.assign(n3, 0.0);
    g.der.assign(n3, 0.0);
    g.dedsp.assign(n3, 0.0);
    g.desum.assign(n3, 0.0);
    g.aem.assign(g.n, 0.0);
    for (auto &row : g.vir)
        for (double &v : row) v = 0.0;
}

} // namespace

void fatal(const std::string &msg)
{
    std::fprintf(stderr, " Fatal error: %s\n", msg.c_str());
    std::exit(1);
}

void mechanic_changed() { uploaded = false; }

void shutdown()
{
    if (ctx) pmq_destroy(ctx);
    ctx = nullptr;
    uploaded = false;
}

template <CalcMode M> void empole()
{
    std::string err;
    if (run(M, PMQ_TERM_MPOLE, &err)) fatal("empole: " + err);
}

template <CalcMode M> void epolar()
{
    std::string err;
    if (run(M, PMQ_TERM_POLAR, &err)) fatal("epolar: " + err);
}

// erepel<CalcMode> replaces the reference's empty pair loop (erepel.cpp:42-153);
// edisp<CalcMode> the absent dispersion term (energy.cpp:280)
template <CalcMode M> void erepel()
{
    std::string err;
    if (run(M, PMQ_TERM_REPEL, &err)) fatal("erepel: " + err);
}

template <CalcMode M> void edisp()
{
    std::string err;
    if (run(M, PMQ_TERM_DISP, &err)) fatal("edisp: " + err);
}

int try_energy(CalcMode m, std::string *err)
{
    zero_outputs();
    unsigned terms = 0;
    if (g.use_mpole) terms |= PMQ_TERM_MPOLE;
    if (g.use_polar) terms |= PMQ_TERM_POLAR;
    if (g.use_repel) terms |= PMQ_TERM_REPEL;
    if (g.use_disp) terms |= PMQ_TERM_DISP;
    if (!terms) return PMQ_OK;
    int rc = run(m, terms, err);
    if (rc) return rc;
    g.esum = g.em + g.ep + g.er + g.edsp;
    for (size_t j = 0; j < g.desum.size(); j++) g.desum[j] = g.dem[j] + g.dep[j] + g.der[j] + g.dedsp[j];
    return PMQ_OK;
}

template <CalcMode M> void energy()
{
    std::string err;
    if (try_energy(M, &err)) fatal("energy: " + err);
}

template void empole<CalcMode::Energy>();
template void empole<CalcMode::Analysis>();
template void empole<CalcMode::Gradient>();
template void empole<CalcMode::Virial>();
template void epolar<CalcMode::Energy>();
template void epolar<CalcMode::Gradient>();
template void epolar<CalcMode::Virial>();
template void erepel<CalcMode::Energy>();
template void erepel<CalcMode::Gradient>();
template void erepel<CalcMode::Virial>();
template void edisp<CalcMode::Energy>();
template void edisp<CalcMode::Gradient>();
template void edisp<CalcMode::Virial>();
template void energy<CalcMode::Energy>();
template void energy<CalcMode::Analysis>();
template void energy<CalcMode::Gradient>();
template void energy<CalcMode::Virial>();

} // namespace pmq_host
