// Test-infrastructure driver for the reference's 2D simplex fit (oracle/_ref only).
//
// Links the reference's own sources from the reference checkout: the console
// parameter layer SC_CalcCons (Source/sc_calcCons.cpp), the key table and
// setter glue SC_Calc_GENERIC (Source/sc_calc_generic.cpp), the simplex
// SasCalc_SimplexFit2D (Source/sc_simplexfit2d.cpp) and the engine
// (Source/sc_calc_generic_cpu.cpp, Source/sc_calc_generic_gpu.cu as C++), all
// built with CONSOLENPROG as the console program is.  It runs one
// doSimplexFit2D (Source/sc_simplexfit2d.cpp:254-611) the way the console's
// autofit does (Source/sc_mainCons.cpp:1660-1720): a parameter hash with
// used and fixed entries, the fit data handed over with setImageInfo, and the
// progress log collected.  Nothing here is part of the product.
//
// usage: ref_fit <params.txt> <out.bin>
//   params.txt: "key=value" lines, keys as in the reference key table
//   (Source/sc_calc_generic.cpp:36-138), plus driver keys:
//     threads=<n>                    numThreads of the fit (default 1)
//     fitdata=<file> fitw=<W> fith=<H>  raw little-endian doubles, W*H values
//     x0,x1,y0,y1                    bounds of the data image (default -GridPoints..GridPoints)
//     bstop=<n> border=<n>           fit masks
//     nofit<i>=x0,y0,x1,y1           no-fit rectangles (i < 4; others unused)
//     stp=<s> maxit=<n> ftol=<t>     doSimplexFit2D arguments (stp as a fraction)
//     fitpar<i>=key,used,min,max     entries of the parameter hash (i = 0, 1, ...);
//                                    fitstart is the parameter's current value
//   out.bin records: fitres (per fitpar<i>, input order), fitstart, repetitions,
//   numImgCalc, aborted, retinfo (text), log (text: the progress lines, '\n'-separated),
//   fqs0, cnt0, fitfqs0: the objective and its per-pixel terms at the start values.

#include "sc_calcCons.h"
#include "sc_simplexfit2d.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::map<std::string, std::string> g_kv;
std::string g_log;

double num(const char* key, double dflt) {
    auto it = g_kv.find(key);
    if (it == g_kv.end()) return dflt;
    return std::strtod(it->second.c_str(), nullptr);
}

bool logger(char* msg) {
    g_log += msg;
    g_log += '\n';
    return false;  // never abort
}

struct Writer {
    FILE* f;
    void put(const char* name, char t, const void* data, long long n, int esz) {
        std::fwrite(name, 1, std::strlen(name) + 1, f);
        std::fwrite(&t, 1, 1, f);
        std::fwrite(&n, sizeof(n), 1, f);
        if (n > 0) std::fwrite(data, esz, n, f);
    }
    void d(const char* name, const double* p, long long n) { put(name, 'd', p, n, 8); }
    void s(const char* name, const std::string& v) { put(name, 's', v.c_str(), (long long)v.size(), 1); }
    void scalar(const char* name, double v) { d(name, &v, 1); }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s params.txt out.bin\n", argv[0]);
        return 2;
    }
    {
        std::ifstream in(argv[1]);
        std::string line;
        while (std::getline(in, line)) {
            auto p = line.find('=');
            if (p == std::string::npos || line.empty() || line[0] == '#') continue;
            g_kv[line.substr(0, p)] = line.substr(p + 1);
        }
    }

    SC_CalcCons calc;
    // Parameter values as the console sets them from a key=value list: the method's own
    // keys through updateParamValue, the others into the global value table.
    for (const auto& kv : g_kv) {
        const QString k = QString::fromStdString(kv.first);
        if (SC_CalcCons::params.contains(k)) calc.updateParamValue(k, std::strtod(kv.second.c_str(), nullptr));
    }
    for (const char* g : {"RadioButtonQ1", "RadioButtonQ2", "RadioButtonQ4", "ExpandImage"})
        if (g_kv.count(g)) SC_CalcCons::inpValues.insert(g, num(g, 0));
    calc.prepareData(true, false);  // fills the global defaults

    const int fw = (int)num("fitw", 0), fh = (int)num("fith", 0);
    std::vector<double> data((size_t)fw * fh, 0.0);
    {
        std::ifstream din(g_kv["fitdata"], std::ios::binary);
        din.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(data.size() * 8));
    }

    // The parameter hash of the autofit (Source/sc_mainCons.cpp:1906-1935).
    _param2fitval p2f;
    std::vector<std::string> order;
    for (int i = 0;; ++i) {
        const std::string k = "fitpar" + std::to_string(i);
        if (!g_kv.count(k)) break;
        std::stringstream ss(g_kv[k]);
        std::string name, used, lo, hi;
        std::getline(ss, name, ',');
        std::getline(ss, used, ',');
        std::getline(ss, lo, ',');
        std::getline(ss, hi, ',');
        _fitLimits* fl = new _fitLimits;
        fl->orgval = fl->fitstart = calc.currentParamValue(QString::fromStdString(name));
        fl->fitres = 0.0;
        fl->fitType = fitNumeric;
        fl->fitvalid = false;
        fl->used = std::atoi(used.c_str()) != 0;
        fl->min = std::strtod(lo.c_str(), nullptr);
        fl->max = std::strtod(hi.c_str(), nullptr);
        p2f.insert(QString::fromStdString(name), fl);
        order.push_back(name);
    }

    // No-fit rectangles as the GUI sets them before a fit, unused ones as -1
    // (Source/sc_maingui.cpp:4177-4184).  The console never sets them, which leaves
    // noFitX0..noFitY1 (Source/sc_calc_generic_gpu.h:682) uninitialised.
    for (int i = 0; i < 4; ++i) {
        const std::string k = "nofit" + std::to_string(i);
        int x0 = -1, y0 = -1, x1 = -1, y1 = -1;
        if (g_kv.count(k)) std::sscanf(g_kv[k].c_str(), "%d,%d,%d,%d", &x0, &y0, &x1, &y1);
        calc.getCalcPtr()->setNoFitRect(i, x0, y0, x1, y1);
    }

    // The objective at the start values, pixel by pixel (doFitCalculation through the console
    // layer, as fehlerquadratsumme calls it, Source/sc_simplexfit2d.cpp:45-62).
    long cnt0 = 0, nan0 = 0;
    calc.setFitData(fw, fh, data.data());
    const double fqs0 = calc.doFitCalculation((int)num("threads", 1), (int)num("bstop", 0), (int)num("border", 0),
                                              cnt0, nan0);
    const std::vector<double> fitfqs0(calc.getCalcPtr()->arrFitFqs, calc.getCalcPtr()->arrFitFqs + (size_t)fw * fh);

    SasCalc_SimplexFit2D fit(&calc);
    // Bounds of the data image (the console passes those of the data file); default: the
    // centred Q4 grid of GridPoints.
    const int gp = (int)num("GridPoints", 64);
    fit.setImageInfo((int)num("x0", -gp), (int)num("x1", gp), (int)num("y0", -gp), (int)num("y1", gp),
                     (int)num("BeamPosX", 0), (int)num("BeamPosY", 0), data.data());
    QString retinfo;
    fit.doSimplexFit2D((int)num("threads", 1), num("stp", 0.01), (int)num("maxit", 100), num("ftol", 1e-3),
                       (int)num("border", 0), (int)num("bstop", 0), false, &logger, &p2f, retinfo);

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) { std::perror("open out"); return 3; }
    Writer w{out};
    std::vector<double> res, start;
    for (const auto& name : order) {
        res.push_back(p2f.value(QString::fromStdString(name))->fitres);
        start.push_back(p2f.value(QString::fromStdString(name))->fitstart);
    }
    w.d("fitres", res.data(), (long long)res.size());
    w.d("fitstart", start.data(), (long long)start.size());
    w.scalar("repetitions", fit.repetitions);
    w.scalar("numImgCalc", fit.numImgCalc);
    w.scalar("aborted", fit.aborted ? 1.0 : 0.0);
    w.s("retinfo", retinfo.toStdString());
    w.s("log", g_log);
    w.scalar("fqs0", fqs0);
    w.scalar("cnt0", (double)cnt0);
    w.d("fitfqs0", fitfqs0.data(), (long long)fitfqs0.size());
    std::fclose(out);
    for (auto* fl : p2f) delete fl;
    return 0;
}
