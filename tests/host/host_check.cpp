// host_check.cpp -- drives pmq_host::energy<CalcMode>() the way the reference's
// own tests drive energy<CalcMode>() (test/*.h): fill the globals, call, read
// em/dem/vir back.  Input: a text system file (`name count v0 v1 ...` per
// line, written by tests/test_gpu_host_shim.py); output: a text file of
// results.  Usage: host_check SYSTEM OUT MODE(0..3) TERMS(bit 1 mpole, 2 polar) PRECISION
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../polmdqc_amd/host/pmq_host.hpp"

using Arr = std::vector<double>;

static std::map<std::string, Arr> read_system(const char *path)
{
    std::map<std::string, Arr> m;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string name;
        size_t cnt = 0;
        if (!(is >> name >> cnt)) continue;
        Arr v(cnt);
        for (size_t i = 0; i < cnt; i++) is >> v[i];
        m[name] = v;
    }
    return m;
}

static std::vector<int> ints(const Arr &a) { return std::vector<int>(a.begin(), a.end()); }

static std::vector<std::vector<int>> lists(const std::map<std::string, Arr> &m, const std::string &cnt,
                                           const std::string &flat, int n)
{
    std::vector<std::vector<int>> out(n);
    auto ic = m.find(cnt), il = m.find(flat);
    if (ic == m.end() || il == m.end()) return {};
    size_t k = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < (int)ic->second[i]; j++) out[i].push_back((int)il->second[k++]);
    return out;
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        std::fprintf(stderr, "usage: host_check SYSTEM OUT MODE TERMS PRECISION\n");
        return 2;
    }
    auto m = read_system(argv[1]);
    auto s = [&](const char *k, double dflt) { auto it = m.find(k); return it == m.end() ? dflt : it->second[0]; };
    auto &g = pmq_host::g;
    g.n = (int)s("n", 0);
    g.x = m["x"]; g.y = m["y"]; g.z = m["z"];
    g.pole = m["pole"];
    g.polaxe = ints(m["polaxe"]); g.zaxis = ints(m["zaxis"]); g.xaxis = ints(m["xaxis"]); g.yaxis = ints(m["yaxis"]);
    if (m.count("use")) g.use = ints(m["use"]);
    if (m.count("molcule")) g.molcule = ints(m["molcule"]);
    g.i12 = lists(m, "n12", "i12", g.n); g.i13 = lists(m, "n13", "i13", g.n);
    g.i14 = lists(m, "n14", "i14", g.n); g.i15 = lists(m, "n15", "i15", g.n);
    g.ip11 = lists(m, "np11", "ip11", g.n); g.ip12 = lists(m, "np12", "ip12", g.n);
    g.ip13 = lists(m, "np13", "ip13", g.n); g.ip14 = lists(m, "np14", "ip14", g.n);
    g.m2scale = s("m2scale", 0); g.m3scale = s("m3scale", 0); g.m4scale = s("m4scale", 0); g.m5scale = s("m5scale", 0);
    g.electric = s("electric", 332.0637133); g.dielec = s("dielec", 1);
    g.pentype = (int)s("pentype", 0);
    if (g.pentype) { g.pcore = m["pcore"]; g.pval0 = m["pval0"]; g.palpha = m["palpha"]; }
    g.use_chgflx = s("use_chgflx", 0) != 0;
    if (g.use_chgflx) {
        g.mono0 = m["mono0"]; g.bl = m["bl"]; g.bflx = m["bflx"]; g.anat = m["anat"]; g.aflx = m["aflx"];
        g.abflx = m["abflx"];
        g.ibnd = ints(m["ibnd"]); g.iang = ints(m["iang"]);
        if (m.count("balist")) g.balist = ints(m["balist"]);
    }
    if (m.count("polarity")) { g.polarity = m["polarity"]; g.thole = m["thole"]; g.pdamp = m["pdamp"]; }
    g.p2scale = s("p2scale", 0); g.p3scale = s("p3scale", 0); g.p4scale = s("p4scale", 1); g.p5scale = s("p5scale", 1);
    g.p2iscale = s("p2iscale", 0); g.p3iscale = s("p3iscale", 0); g.p4iscale = s("p4iscale", 0.5);
    g.p5iscale = s("p5iscale", 1);
    g.d1scale = s("d1scale", 0); g.d2scale = s("d2scale", 1); g.d3scale = s("d3scale", 1); g.d4scale = s("d4scale", 1);
    g.u1scale = s("u1scale", 1); g.u2scale = s("u2scale", 1); g.u3scale = s("u3scale", 1); g.u4scale = s("u4scale", 1);
    g.poleps = s("poleps", 1e-6); g.politer = (int)s("politer", 100);
    g.use_bounds = s("use_bounds", 0) != 0; g.monoclinic = s("monoclinic", 0) != 0; g.triclinic = s("triclinic", 0) != 0;
    g.octahedron = s("octahedron", 0) != 0; g.dodecadron = s("dodecadron", 0) != 0;
    g.xbox = s("xbox", 0); g.ybox = s("ybox", 0); g.zbox = s("zbox", 0);
    g.alpha = s("alpha", 90); g.beta = s("beta", 90); g.gamma = s("gamma", 90);
    g.use_ewald = s("use_ewald", 0) != 0; g.aewald = s("aewald", 0);
    g.mpolecut = s("mpolecut", 1e12); g.ewaldcut = s("ewaldcut", 7);
    g.nfft1 = (int)s("nfft1", 0); g.nfft2 = (int)s("nfft2", 0); g.nfft3 = (int)s("nfft3", 0); g.bsorder = (int)s("bsorder", 5);
    const int mode = std::atoi(argv[3]), terms = std::atoi(argv[4]);
    g.use_mpole = terms & 1;
    g.use_polar = (terms & 2) != 0;
    g.precision = std::atoi(argv[5]);
    pmq_host::mechanic_changed();
    const pmq_host::CalcMode cm[4] = {pmq_host::CalcMode::Energy, pmq_host::CalcMode::Analysis,
                                      pmq_host::CalcMode::Gradient, pmq_host::CalcMode::Virial};
    switch (mode) {
    case 0: pmq_host::energy<pmq_host::CalcMode::Energy>(); break;
    case 1: pmq_host::energy<pmq_host::CalcMode::Analysis>(); break;
    case 2: pmq_host::energy<pmq_host::CalcMode::Gradient>(); break;
    default: pmq_host::energy<pmq_host::CalcMode::Virial>(); break;
    }
    (void)cm;
    FILE *o = std::fopen(argv[2], "w");
    std::fprintf(o, "em 1 %.17g\nep 1 %.17g\nesum 1 %.17g\nnem 1 %d\neinter 1 %.17g\npcg_iters 1 %d\n", g.em, g.ep,
                 g.esum, g.nem, g.einter, g.pcg_iters);
    auto dump = [&](const char *name, const Arr &a) {
        std::fprintf(o, "%s %zu", name, a.size());
        for (double v : a) std::fprintf(o, " %.17g", v);
        std::fprintf(o, "\n");
    };
    dump("desum", g.desum);
    dump("aem", g.aem);
    Arr vir(9);
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) vir[3 * a + c] = g.vir[a][c];
    dump("vir", vir);
    std::fclose(o);
    pmq_host::shutdown();
    return 0;
}
