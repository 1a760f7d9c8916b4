// ref_dump -- golden-vector generator built against the compiled reference.
//
// Test infrastructure only (see oracle/ref.mk).  It runs the reference's own
// setup (initial -> getcart -> mechanic, as analyze.cpp:60-62 does), writes
// every array the electrostatics path reads, then runs the reference's
// energy<Analysis>() and energy<Virial>() (energy.cpp:31) and writes what they
// produce.  The JSON it prints is the fixture format read by tests/golden.
//
//   usage: ref_dump OUT.json XYZFILE [-k KEYFILE]
//
// Run it single-threaded (OMP_NUM_THREADS=1): the reference's clang/libomp
// build gives wrong energies with more than one thread.

#include "action.h"
#include "analyz.h"
#include "angbnd.h"
#include "atmlst.h"
#include "atomid.h"
#include "atoms.h"
#include "bndstr.h"
#include "bound.h"
#include "boxes.h"
#include "calcMode.h"
#include "cell.h"
#include "cflux.h"
#include "chgpen.h"
#include "chgpot.h"
#include "couple.h"
#include "deriv.h"
#include "energi.h"
#include "energy.h"
#include "ewald.h"
#include "files.h"
#include "getcart.h"
#include "inform.h"
#include "initial.h"
#include "inter.h"
#include "kmulti.h"
#include "mechanic.h"
#include "mdqclimits.h"
#include "molcul.h"
#include "mplpot.h"
#include "mpole.h"
#include "pme.h"
#include "polar.h"
#include "polgrp.h"
#include "polpot.h"
#include "potent.h"
#include "repel.h"
#include "reppot.h"
#include "shunt.h"
#include "usage.h"
#include "virial.h"

#include <cstdio>
#include <string>
#include <vector>

using namespace polmdqc;

static FILE* out;
static bool first_key;

static void key(const char* k)
{
    std::fprintf(out, "%s\n  \"%s\": ", first_key ? "" : ",", k);
    first_key = false;
}
static void num(const char* k, double v) { key(k); std::fprintf(out, "%.17g", v); }
static void inum(const char* k, long v) { key(k); std::fprintf(out, "%ld", v); }
static void str(const char* k, const std::string& v) { key(k); std::fprintf(out, "\"%s\"", v.c_str()); }

template <class F>
static void arr(const char* k, int len, F get)
{
    key(k);
    std::fprintf(out, "[");
    for (int i = 0; i < len; i++) std::fprintf(out, "%s%.17g", i ? "," : "", (double)get(i));
    std::fprintf(out, "]");
}

template <class F>
static void iarr(const char* k, int len, F get)
{
    key(k);
    std::fprintf(out, "[");
    for (int i = 0; i < len; i++) std::fprintf(out, "%s%ld", i ? "," : "", (long)get(i));
    std::fprintf(out, "]");
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: ref_dump OUT.json XYZFILE [-k KEYFILE]\n");
        return 2;
    }
    std::vector<char*> av;
    av.push_back(argv[0]);
    for (int i = 2; i < argc; i++) av.push_back(argv[i]);
    test = true;
    initial((int)av.size(), av.data());
    getcart(ffile);
    mechanic();

    out = std::fopen(argv[1], "w");
    if (!out) return 3;
    first_key = true;
    std::fprintf(out, "{");

    // --- inputs exactly as mechanic() left them -------------------------
    inum("n", n);
    arr("x", n, [](int i) { return x[i]; });
    arr("y", n, [](int i) { return y[i]; });
    arr("z", n, [](int i) { return z[i]; });
    iarr("type", n, [](int i) { return type[i]; });
    iarr("atomic", n, [](int i) { return atomic[i]; });
    arr("pole", 13 * n, [](int i) { return pole[i / 13][i % 13]; });
    iarr("polaxe", n, [](int i) { return (int)polaxe[i]; });
    iarr("zaxis", n, [](int i) { return zaxis[i]; });
    iarr("xaxis", n, [](int i) { return xaxis[i]; });
    iarr("yaxis", n, [](int i) { return yaxis[i]; });
    iarr("pollist", n, [](int i) { return pollist[i]; });
    iarr("use", n + 1, [](int i) { return (int)use[i]; });
    iarr("molcule", n, [](int i) { return molcule[i]; });
    iarr("n12", n, [](int i) { return n12[i]; });
    iarr("n13", n, [](int i) { return n13[i]; });
    iarr("n14", n, [](int i) { return n14[i]; });
    iarr("n15", n, [](int i) { return n15[i]; });
    {
        // flatten the connectivity lists as CSR (0-based atom indices)
        std::vector<int> v12, v13, v14, v15;
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n12[i]; j++) v12.push_back(i12[i][j]);
            for (int j = 0; j < n13[i]; j++) v13.push_back(i13[i][j]);
            for (int j = 0; j < n14[i]; j++) v14.push_back(i14[i][j]);
            for (int j = 0; j < n15[i]; j++) v15.push_back(i15[i][j]);
        }
        iarr("i12", (int)v12.size(), [&](int i) { return v12[i]; });
        iarr("i13", (int)v13.size(), [&](int i) { return v13[i]; });
        iarr("i14", (int)v14.size(), [&](int i) { return v14[i]; });
        iarr("i15", (int)v15.size(), [&](int i) { return v15[i]; });
    }
    num("m2scale", m2scale);
    num("m3scale", m3scale);
    num("m4scale", m4scale);
    num("m5scale", m5scale);
    num("electric", electric);
    num("dielec", dielec);
    inum("pentype", (int)pentype);
    inum("use_chgflx", use_chgflx);
    inum("use_mpole", use_mpole);
    inum("use_polar", use_polar);
    inum("use_repel", use_repel);
    inum("use_ewald", use_ewald);
    inum("use_bounds", use_bounds);
    num("mpolecut", mpolecut);
    num("ewaldcut", ewaldcut);
    num("usolvcut", usolvcut);
    num("aewald", aewald);
    inum("nfft1", nfft1);
    inum("nfft2", nfft2);
    inum("nfft3", nfft3);
    inum("bsorder", bsorder);
    arr("pcore", n, [](int i) { return pcore[i]; });
    arr("pval0", n, [](int i) { return pval0[i]; });
    arr("palpha", n, [](int i) { return palpha[i]; });
    arr("mono0", n, [](int i) { return mono0[i]; });

    // periodic box
    inum("orthogonal", orthogonal);
    inum("monoclinic", monoclinic);
    inum("triclinic", triclinic);
    inum("octahedron", octahedron);
    inum("dodecadron", dodecadron);
    num("xbox", xbox);
    num("ybox", ybox);
    num("zbox", zbox);
    num("alpha", alphaA);
    num("beta", betaA);
    num("gamma", gammaA);

    // charge flux topology and parameters
    inum("nbond", nbond);
    iarr("ibnd", 2 * nbond, [](int i) { return ibnd[i / 2][i % 2]; });
    arr("bl", nbond, [](int i) { return bl[i]; });
    arr("bflx", use_chgflx ? nbond : 0, [](int i) { return bflx[i]; });
    inum("nangle", nangle);
    iarr("iang", 3 * nangle, [](int i) { return iang[i / 3][i % 3]; });
    arr("anat", nangle, [](int i) { return anat[i]; });
    arr("aflx", use_chgflx ? 2 * nangle : 0, [](int i) { return aflx[i / 2][i % 2]; });
    arr("abflx", use_chgflx ? 2 * nangle : 0, [](int i) { return abflx[i / 2][i % 2]; });
    iarr("balist", use_chgflx ? 2 * nangle : 0, [](int i) { return balist[i / 2][i % 2]; });

    // polarization setup (kpolar.cpp); consumed by no reference energy term
    arr("polarity", n, [](int i) { return polarity[i]; });
    arr("thole", n, [](int i) { return thole[i]; });
    arr("pdamp", n, [](int i) { return pdamp[i]; });
    num("p2scale", p2scale); num("p3scale", p3scale);
    num("p4scale", p4scale); num("p5scale", p5scale);
    num("p2iscale", p2iscale); num("p3iscale", p3iscale);
    num("p4iscale", p4iscale); num("p5iscale", p5iscale);
    num("d1scale", d1scale); num("d2scale", d2scale);
    num("d3scale", d3scale); num("d4scale", d4scale);
    num("u1scale", u1scale); num("u2scale", u2scale);
    num("u3scale", u3scale); num("u4scale", u4scale);
    num("poleps", poleps);
    str("poltyp", poltyp);
    iarr("np11", n, [](int i) { return np11[i]; });
    iarr("np12", n, [](int i) { return np12[i]; });
    iarr("np13", n, [](int i) { return np13[i]; });
    iarr("np14", n, [](int i) { return np14[i]; });
    {
        std::vector<int> v11, v12, v13, v14;
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < np11[i]; j++) v11.push_back(ip11[i][j]);
            for (int j = 0; j < np12[i]; j++) v12.push_back(ip12[i][j]);
            for (int j = 0; j < np13[i]; j++) v13.push_back(ip13[i][j]);
            for (int j = 0; j < np14[i]; j++) v14.push_back(ip14[i][j]);
        }
        iarr("ip11", (int)v11.size(), [&](int i) { return v11[i]; });
        iarr("ip12", (int)v12.size(), [&](int i) { return v12[i]; });
        iarr("ip13", (int)v13.size(), [&](int i) { return v13[i]; });
        iarr("ip14", (int)v14.size(), [&](int i) { return v14[i]; });
    }
    num("w2scale", w2scale); num("w3scale", w3scale);
    num("w4scale", w4scale); num("w5scale", w5scale);
    // repulsion (krepel.cpp) and dispersion (kdsp.h) setup
    arr("sizpr", n, [](int i) { return sizpr[i]; });
    arr("dmppr", n, [](int i) { return dmppr[i]; });
    arr("elepr", n, [](int i) { return elepr[i]; });
    num("r2scale", r2scale); num("r3scale", r3scale);
    num("r4scale", r4scale); num("r5scale", r5scale);

    // --- reference outputs ----------------------------------------------
    energy<CalcMode::Analysis>();
    num("em_analysis", em);
    inum("nem", nem);
    num("einter", einter);
    num("esum_analysis", esum);
    arr("aem", n, [](int i) { return aem[i]; });
    arr("rpole", 13 * n, [](int i) { return rpole[i / 13][i % 13]; });
    // chkpole inside empole may flip yaxis signs; record the final state too
    iarr("yaxis_after", n, [](int i) { return yaxis[i]; });

    energy<CalcMode::Virial>();
    num("em", em);
    num("esum", esum);
    arr("dem", 3 * n, [](int i) { return dem[i]; });
    arr("desum", 3 * n, [](int i) { return desum[i]; });
    arr("vir", 9, [](int i) { return vir[i / 3][i % 3]; });

    std::fprintf(out, "\n}\n");
    std::fclose(out);
    return 0;
}
