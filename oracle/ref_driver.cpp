// Test-infrastructure driver for the reference CPU engine (oracle/_ref only).
//
// Links the reference's own engine sources, unchanged, from the reference
// checkout (Source/sc_calc_generic_cpu.cpp, Source/sc_calc_generic_gpu.cu
// compiled as C++) and drives them through their public API: the grouped
// setters (Source/sc_calc_generic_gpu.h:108-353), doCalculation
// (Source/sc_calc_generic_cpu.cpp:1151) and doFitCalculation
// (Source/sc_calc_generic_cpu.cpp:1334).  The only output is a binary dump of
// the prepared tables and the image / fit residuals, used to write the golden
// fixtures under tests/golden/.  Nothing here is part of the product.
//
// usage: ref_driver <params.txt> <out.bin>
//   params.txt: one "key=value" per line, keys as in the reference key table
//   (Source/sc_calc_generic.cpp:36-138) plus driver keys:
//     threads=<n>          numThreads for doCalculation (default 1)
//     mode=image|fit       (default image)
//     fitdata=<file>       raw little-endian doubles, W*H values (fit mode)
//     fitw=<W> fith=<H>    fit array size (fit mode)
//     bstop=<n> border=<n> fit masks (fit mode)
//     nofit<i>=x0,y0,x1,y1 no-fit rectangles (fit mode)
//     ignorecbs=0|1        bIgnoreCbs for doCalculation (default 1)
//     repeat=<n>           repeat the calculation n times (timing)
//     use1d=1              1D mode: EditQsteps pixels, q from EditQmin..EditQmax
//     rphi=<s> bsx= bsy=   r,phi transform of the image (s x s, centre bsx,bsy) and its log
//                          scaling; rphiscale/rphiclip/rphiclip40 = flags
//     fft=<type>           calculateIFFT of the log-scaled r,phi image (0 re, 1 im, 2 abs,
//                          3 spec); fftforward, fftscale, fftclip, fftclip40, fftswap flags,
//                          fftcutx/fftcuty cut-out
//     scaleint=0|1         scaleIntensity of the image (1 linear, 0 log)
//     tpv=<k=f|k=f|...>    tpvPerformRandom(list) before each of tpvn=<n> images
//                          (srand=<s> seeds rand() first; default: the process's seed 1)

#include "sc_calc_generic_gpu.h"
#include "sc_postproc.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <list>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::map<std::string, std::string> g_kv;

double num(const char* key, double dflt) {
    auto it = g_kv.find(key);
    if (it == g_kv.end()) return dflt;
    return std::strtod(it->second.c_str(), nullptr);
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
    void f32(const char* name, const float* p, long long n) { put(name, 'f', p, n, 4); }
    void i32(const char* name, const int* p, long long n) { put(name, 'i', p, n, 4); }
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

    SasCalc_GENERIC_calculation* c = new SasCalc_GENERIC_calculation();

    // Every setter, as the console's prepareData calls them before each image
    // (Source/sc_calc_generic.cpp:142-250).
    auto apply_params = [&]() {
        // Defaults are the key-table defaults (Source/sc_calc_generic.cpp:36-138).
        c->setLatticeParams((int)num("LType", 12), num("uca", 30), num("ucb", 30), num("ucc", 30),
                            num("ucalpha", 90), num("ucbeta", 90), num("ucgamma", 90),
                            num("EditCeff", 1), num("CheckBoxTwinned", 0) != 0, num("EditCeffcyl", 0),
                            num("reff", 0), num("acpl", 0), num("bcpl", 0),
                            num("ifluc", 0), num("rfluc", 0));
        c->setParticleParams((int)num("ComboBoxParticle", 0), (int)num("Ordis", 7),
                             (int)num("ComboBoxInterior", 0), num("EditRadius", 1.0),
                             num("EditRadiusi", 0.5), num("EditSigma", 0.07), num("EditDbeta", 0.4),
                             num("Length", 1), num("SigmaL", 0.06), num("Alpha", 0), num("EditRho", 0),
                             num("RotAlpha", 0));
        c->setPeakShapeParams((int)num("ComboBoxPeak", 1), num("EditPeakPar", 0),
                              num("EditDebyeWaller", 2), num("EditAzi", 180),
                              num("EditDomainSize", 400), num("RadButDebyeScherrer", 0) != 0,
                              num("RadioButtonPara", 0) != 0);
        c->setPeakShapeMatrix(num("VAx1", 1), num("VAx2", 0), num("VAx3", 0),
                              num("Ay1", 0), num("Ay2", 1), num("Ay3", 0),
                              num("Az1", 0), num("Az2", 0), num("Az3", 1),
                              num("SigX", 40), num("SigY", 40), num("SigZ", 40));
        c->setOrientationParams(num("ucpsi", 0), num("ucn1", 1), num("ucn2", 1), num("ucn3", 1),
                                num("theta", 90), num("phi", 0), num("rotTheta", 90), num("rotPhi", 90));
        int q1 = (int)num("RadioButtonQ1", 0), q2 = (int)num("RadioButtonQ2", 0);
        int q4 = (int)num("RadioButtonQ4", (q1 || q2) ? 0 : 1);
        // 1D mode (use1d=1): the grid is EditQsteps and EditQmin is set, as the console's
        // prepareData does for use1d (Source/sc_calc_generic.cpp:202-206)
        const bool d1 = num("use1d", 0) != 0;
        c->setCalculationParams(d1 ? (int)num("EditQsteps", 200) : (int)num("GridPoints", 64), (int)num("HKLmax", 5),
                                num("EditQmax", 1), q1 != 0, q2 != 0, q4 != 0, num("ExpandImage", 0) != 0, d1);
        if (d1) c->setQMin(num("EditQmin", 0));
        c->setExperimentParams(num("EditPixelNoX", 2048), num("EditPixelNoY", 2048),
                               num("EditPixelX", 1), num("EditPixelY", 1), num("EditDet", 1),
                               num("EditWavelength", 0.09499), num("BeamPosX", 0), num("BeamPosY", 0),
                               num("CenterBeam", 0) != 0);
        c->setControlParams(num("EditBFactor", 1), num("CheckBoxWAXS", 0) != 0, num("P1", 0));
        c->setPixelManipulationParams(num("iso", 0), num("I0", 10000), num("Base", 0));
    };
    apply_params();

    const int threads = (int)num("threads", 1);
    const int repeat = (int)num("repeat", 1);
    const bool fitmode = g_kv.count("mode") && g_kv["mode"] == "fit";

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) { std::perror("open out"); return 3; }
    Writer w{out};

    double fqs = 0.0;
    long cnt = 0, nancnt = 0;
    std::vector<double> calcms, prepms;
    if (fitmode) {
        int fw = (int)num("fitw", 0), fh = (int)num("fith", 0);
        std::vector<double> data((size_t)fw * fh, 0.0);
        std::ifstream din(g_kv["fitdata"], std::ios::binary);
        din.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(data.size() * 8));
        for (int i = 0; i < 4; ++i) {
            std::string k = "nofit" + std::to_string(i);
            int x0 = -1, y0 = -1, x1 = -1, y1 = -1;
            if (g_kv.count(k)) std::sscanf(g_kv[k].c_str(), "%d,%d,%d,%d", &x0, &y0, &x1, &y1);
            c->setNoFitRect(i, x0, y0, x1, y1);
        }
        c->setFitData(fw, fh, data.data());
        for (int r = 0; r < repeat; ++r) {
            fqs = c->doFitCalculation(threads, (int)num("bstop", 0), (int)num("border", 0), cnt, nancnt);
            calcms.push_back(c->higResTimerElapsedCalc);
            prepms.push_back(c->higResTimerElapsedPrep);
        }
        w.scalar("fqs", fqs);
        w.scalar("cnt", (double)cnt);
        w.scalar("nancnt", (double)nancnt);
        w.d("fitfqs", c->arrFitFqs, (long long)fw * fh);
    } else if (g_kv.count("tpv")) {
        // tpvPerformRandom before each of tpvn images (Source/sc_mainCons.cpp:1236-1246), with
        // the process's rand() state as the reference run has it (seed 1 unless srand=<n>).
        if (g_kv.count("srand")) std::srand((unsigned)num("srand", 1));
        std::list<std::string> ids;
        std::stringstream ss(g_kv["tpv"]);
        std::string item;
        while (std::getline(ss, item, '|')) ids.push_back(item);
        const int n = (int)num("tpvn", 1);
        std::vector<double> imgs;
        std::string reports;
        for (int r = 0; r < n; ++r) {
            apply_params();
            reports += c->tpvPerformRandom(ids) + "#";
            c->doCalculation(threads, false, num("ignorecbs", 1) != 0);
            const long long np = (long long)(c->maxX() - c->minX()) * (c->maxY() - c->minY());
            imgs.insert(imgs.end(), c->data(), c->data() + np);
        }
        w.d("tpv_images", imgs.data(), (long long)imgs.size());
        w.put("tpv_report", 's', reports.c_str(), (long long)reports.size(), 1);
        double b[4] = {(double)c->minX(), (double)c->maxX(), (double)c->minY(), (double)c->maxY()};
        w.d("bounds", b, 4);
    } else {
        for (int r = 0; r < repeat; ++r) {
            c->doCalculation(threads, false, num("ignorecbs", 1) != 0);
            calcms.push_back(c->higResTimerElapsedCalc);
            prepms.push_back(c->higResTimerElapsedPrep);
        }
        const long long n = (long long)(c->maxX() - c->minX()) * (c->maxY() - c->minY());
        w.d("image", c->data(), n);
        if (g_kv.count("rphi")) {
            // generateRphi (Source/sc_postproc.cpp:53-360) of the image, then the log scaling
            // the AI-data loop applies to it (Source/sc_mainCons.cpp:1385-1395)
            const int rs = (int)num("rphi", 64);
            double* rp = SasCalc_PostProc::inst()->generateRphi(c->minX(), c->maxX(), c->minY(), c->maxY(),
                                                               (int)num("bsx", 0), (int)num("bsy", 0), rs, c->data(),
                                                               num("rphiscale", 0) != 0, num("rphiclip", 0) != 0,
                                                               num("rphiclip40", 0) != 0);
            w.d("rphi", rp, (long long)rs * rs);
            SasCalc_PostProc::inst()->scaleAndClipData(rp, rs * rs, false, false, false, true);
            w.d("rphi_log", rp, (long long)rs * rs);
            if (g_kv.count("fft")) {
                // calculateIFFT of the r,phi image (Source/sc_mainCons.cpp:1400-1420): backward,
                // output size = r,phi size
                const int cx = (int)num("fftcutx", 0), cy = (int)num("fftcuty", 0);
                double* f = SasCalc_PostProc::inst()->calculateIFFT(
                    num("fftforward", 0) != 0, 0, rs, 0, rs, rp, rs,
                    static_cast<SasCalc_PostProc::_outType>((int)num("fft", 0)), num("fftscale", 0) != 0,
                    num("fftclip", 0) != 0, num("fftclip40", 0) != 0, num("fftswap", 0) != 0, cx, cy);
                const bool cut = cx > 0 && cy > 0 && cx < rs && cy < rs;
                if (f) w.d("fft", f, cut ? (long long)cx * cy : (long long)rs * rs);
            }
        }
        if (g_kv.count("scaleint")) {
            // scaleIntensity (Source/sc_calc_generic_cpu.cpp:11195-11236): 1 linear, 0 log
            c->scaleIntensity(num("scaleint", 0) != 0);
            w.d("scaled", c->data(), n);
        }
        double b[4] = {(double)c->minX(), (double)c->maxX(), (double)c->minY(), (double)c->maxY()};
        w.d("bounds", b, 4);
        std::string err = c->getLastErrorMessage();
        w.put("error", 's', err.c_str(), (long long)err.size(), 1);
    }
    w.d("calc_ms", calcms.data(), (long long)calcms.size());
    w.d("prep_ms", prepms.data(), (long long)prepms.size());

    // Prepared tables (public members of the engine).
    auto* P = &c->params;
    if (P->CR != nullptr) {
        const auto* CR = P->CR;
        w.d("carr1p", CR->carr1p, coeffarray_len);
        w.d("carr2p", CR->carr2p, coeffarray_len);
        w.d("carr3p", CR->carr3p, coeffarray_len);
        w.d("carr4p", CR->carr4p, coeffarray_len);
        w.d("carr1f", CR->carr1f, coeffarray_len);
        w.d("carr2f", CR->carr2f, coeffarray_len);
        w.d("carr4f", CR->carr4f, coeffarray_len);
        w.d("carr11pm", &CR->carr11pm[0][0], (long long)imax2d_len * imax2d_len);
    }
    double lim[18] = {P->limq1, P->limq2, P->limq3, P->limq4, P->limq5, P->limq6, P->limq7,
                      P->limq8, P->limq9, P->limq1f, P->limq2f, P->limq3f, P->limq4f,
                      P->limq5f, P->limq6f, P->limq7f, P->limq8f, P->limq9f};
    w.d("limq", lim, 18);
    w.scalar("norm", P->norm);
    w.scalar("por", P->por);
    w.scalar("orcase", (double)P->orcase);
    w.scalar("polPhi", P->polPhi);
    w.scalar("ucvol", c->ucvol);
    w.scalar("dwfactor", c->dwfactor);
    w.scalar("phiwidth", c->phiwidth);
    w.scalar("cosphic", c->cosphic);
    w.scalar("sinphic", c->sinphic);
    if (c->latpar1ptr != nullptr) {
        const int pm1 = c->latpar1ptr[1 * 6 + 0];
        const int pm2 = c->latpar2ptr[1 * 6 + 0];
        w.scalar("peakmax1", pm1);
        w.scalar("peakmax2", pm2);
        const int n1 = (pm1 + 1) * 6, n2 = (pm2 + 1) * 6;
        w.i32("latpar1", c->latpar1ptr, n1);
        w.i32("latpar2", c->latpar2ptr, n2);
        const int n3 = (pm2 + 1) * 17;
        w.f32("latpar3", c->latpar3ptr, n3);
        double ri[9] = {c->ri11, c->ri12, c->ri13, c->ri21, c->ri22, c->ri23, c->ri31, c->ri32, c->ri33};
        w.d("ri", ri, 9);
    }
    std::fclose(out);
    return 0;
}
