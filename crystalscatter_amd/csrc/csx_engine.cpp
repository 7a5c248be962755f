// SasCalc_GENERIC_calculation facade over the C ABI (csx_engine_*).
//
// Mirrors the reference engine's public surface used by sc_calcCons and
// sc_simplexfit2d (Source/sc_calc_generic_gpu.h:108-598): parameters by key,
// doCalculation, setFitData / setNoFitRect / doFitCalculation, prepared tables.
// Host preparation runs here in C++; every pixel runs on the GPU.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "csx.h"
#include "csx_prep.h"
#include "csx_rand.h"

namespace {

// zuf() of the reference (Source/sc_calc_generic_cpu.cpp:1085, Source/sc_mainCons.cpp:850-853):
// uniform on [-1, 1] from one rand() value.
double zuf(csx_rand* r) {
    const int32_t v = csr::value(csr::step(r->s));
    return ((static_cast<double>(v) / 2147483647.0 * 200.0) - 100.0) / 100.0;
}

// The variables tpvPerformRandom can vary (Source/sc_calc_generic_cpu.cpp:1057-1081, in its
// order; the table lists EditDbeta twice and the map keeps the first).  norm: old * (1 + f zuf);
// uca: the current uca; phi: 2 + 43 (zuf + 1) degrees; azi: phiwidth is varied, but
// prepareCalculation recomputes it from the unchanged azimuthal width (cpu:413), so only the
// random number is used up; dw: the Debye-Waller factor d^2/3 times (1 + f zuf).
enum TpvFlag { TPV_NORM, TPV_UCA, TPV_PHI, TPV_AZI, TPV_DW };
struct TpvVar {
    const char* key;
    double csx::ParamSet::*field;
    TpvFlag flag;
};
const TpvVar kTpv[] = {
    {"I0", &csx::ParamSet::i0, TPV_NORM},           {"EditRadius", &csx::ParamSet::radius, TPV_NORM},
    {"EditSigma", &csx::ParamSet::sigma, TPV_NORM},  {"Length", &csx::ParamSet::length, TPV_NORM},
    {"SigmaL", &csx::ParamSet::sigmal, TPV_NORM},    {"uca", &csx::ParamSet::uca, TPV_NORM},
    {"ucb", &csx::ParamSet::ucb, TPV_UCA},           {"ucc", &csx::ParamSet::ucc, TPV_UCA},
    {"ucpsi", &csx::ParamSet::ucpsi, TPV_NORM},      {"EditDbeta", &csx::ParamSet::dbeta, TPV_NORM},
    {"EditDomainSize", &csx::ParamSet::width_zuf, TPV_NORM}, {"EditAzi", &csx::ParamSet::azi, TPV_AZI},
    {"EditDebyeWaller", &csx::ParamSet::displacement, TPV_DW}, {"phi", &csx::ParamSet::phi, TPV_PHI},
    {"Base", &csx::ParamSet::base, TPV_NORM},        {"EditRho", &csx::ParamSet::rho, TPV_NORM},
};
constexpr int kNTpv = sizeof(kTpv) / sizeof(kTpv[0]);

}  // namespace

struct csx_engine {
    csx::ParamSet p;        // parameters as set (the reference's parameter store)
    csx::ParamSet pt;       // the values a TPV call derived from them (used while tpv_on)
    bool tpv_on = false;
    double tpv_old[kNTpv];  // values at the first TPV call after a clear
    csx::PreparedSet P;
    csx_ctx* ctx = nullptr;
    int device = 0;
    bool prepared = false;
    bool ignore_cbs = false;
    std::string err;
    // fit state (setFitData)
    int fw = 0, fh = 0;
    double* d_fit = nullptr;
    int rects[16];
    double* d_img = nullptr;
    size_t img_n = 0;
    double* d_block = nullptr;  // computed block of an ExpandImage set
    size_t block_n = 0;
};

const csx::ParamSet* csx_engine_paramset(const csx_engine* e) { return e ? &e->p : nullptr; }

extern "C" {

csx_engine* csx_engine_new(int device) {
    csx_engine* e = new csx_engine();
    e->device = device;
    // device < 0: host-only engine (preparation and tables, no GPU work).
    if (device >= 0 && csx_create(device, &e->ctx) != CSX_OK) { delete e; return nullptr; }
    for (int i = 0; i < 16; ++i) e->rects[i] = -1;
    return e;
}

void csx_engine_free(csx_engine* e) {
    if (!e) return;
    if (e->ctx) {
        (void)hipSetDevice(e->device);
        if (e->d_fit) (void)hipFree(e->d_fit);
        if (e->d_img) (void)hipFree(e->d_img);
        if (e->d_block) (void)hipFree(e->d_block);
        csx_destroy(e->ctx);
    }
    delete e;
}

int csx_engine_set(csx_engine* e, const char* key, double v) {
    if (!e || !key) return CSX_E_ARG;
    if (std::strcmp(key, "ignore_cbs") == 0) { e->ignore_cbs = v != 0; e->prepared = false; return CSX_OK; }
    if (!e->p.set(key, v)) { e->err = std::string("unknown key ") + key; return CSX_E_ARG; }
    if (e->tpv_on) e->pt.set(key, v);  // a setter overwrites the randomized value until the next call
    e->prepared = false;
    return CSX_OK;
}

double csx_engine_get(csx_engine* e, const char* key) {
    double v = 0;
    if (!e || !key || !e->p.get(key, &v)) return 0.0 / 0.0;
    return v;
}

static int engine_prepare(csx_engine* e, bool check) {
    if (check) {
        const std::string m = csx::check_parameters(e->tpv_on ? e->pt : e->p, e->ignore_cbs);
        if (!m.empty()) { e->err = m; return CSX_E_PARAM; }
    }
    const std::string m = csx::prepare(e->tpv_on ? e->pt : e->p, e->P);
    if (!m.empty()) { e->err = m; return CSX_E_ARG; }
    const csx::PreparedSet& P = e->P;
    if (!e->ctx) { e->prepared = true; return CSX_OK; }  // host-only engine
    const int rc = csx_upload(e->ctx, &P.s, P.carr.data(), P.carr11pm.empty() ? nullptr : P.carr11pm.data(),
                              P.latpar1.data(), P.latpar2.data(), P.latpar3.data());
    if (rc != CSX_OK) {
        char b[256];
        csx_last_error(e->ctx, b, sizeof b);
        e->err = b;
        return rc;
    }
    e->prepared = true;
    return CSX_OK;
}

int csx_engine_prepare(csx_engine* e) { return e ? engine_prepare(e, true) : CSX_E_ARG; }

// Checks and host preparation only (the tables go to a batch, not to the engine's context).
static int engine_prepare_host(csx_engine* e) {
    csx_ctx* keep = e->ctx;
    e->ctx = nullptr;
    const int rc = engine_prepare(e, true);
    e->ctx = keep;
    e->prepared = false;  // the engine's own context has not received these tables
    return rc;
}

int csx_engine_calc_device(csx_engine* e, int precision, double* d_out, int32_t* b4) {
    if (!e || !d_out) return CSX_E_ARG;
    if (!e->ctx) { e->err = "host-only engine has no device"; return CSX_E_STATE; }
    int rc = e->prepared ? CSX_OK : engine_prepare(e, true);  // tables stay in HBM until a parameter changes
    if (rc != CSX_OK) return rc;
    const csx_scalars& s = e->P.s;
    csx_grid g{s.xmin, s.xmax, s.ymin, s.ymax};
    if (b4) { b4[0] = s.axmin; b4[1] = s.axmax; b4[2] = s.aymin; b4[3] = s.aymax; }
    const bool expand = s.axmin != s.xmin || s.axmax != s.xmax || s.aymin != s.ymin || s.aymax != s.ymax;
    if (expand) {
        // ExpandImage: the computed block, then the whole image from it
        const size_t nb = (size_t)(s.xmax - s.xmin) * (size_t)(s.ymax - s.ymin);
        (void)hipSetDevice(e->device);
        if (nb > e->block_n) {
            if (csx_synchronize(e->ctx) != CSX_OK) { e->err = "synchronize"; return CSX_E_HIP; }
            if (e->d_block) (void)hipFree(e->d_block);
            e->d_block = nullptr;
            e->block_n = 0;
            if (hipMalloc(&e->d_block, nb * sizeof(double)) != hipSuccess) { e->err = "hipMalloc block"; return CSX_E_NOMEM; }
            e->block_n = nb;
        }
        rc = csx_render(e->ctx, &g, precision, e->d_block);
        if (rc == CSX_OK) rc = csx_expand_image(e->ctx, e->d_block, d_out);
    } else {
        rc = csx_render(e->ctx, &g, precision, d_out);
    }
    if (rc != CSX_OK) {
        char b[256];
        csx_last_error(e->ctx, b, sizeof b);
        e->err = b;
    }
    return rc;
}

int csx_engine_calc(csx_engine* e, int precision, double* h_out, int32_t* b4) {
    if (!e || !h_out) return CSX_E_ARG;
    if (!e->ctx) { e->err = "host-only engine has no device"; return CSX_E_STATE; }
    int rc = e->prepared ? CSX_OK : engine_prepare(e, true);
    if (rc != CSX_OK) return rc;
    const csx_scalars& s = e->P.s;
    const size_t n = (size_t)(s.axmax - s.axmin) * (size_t)(s.aymax - s.aymin);
    (void)hipSetDevice(e->device);
    if (n > e->img_n) {
        if (e->d_img) (void)hipFree(e->d_img);
        e->d_img = nullptr;
        if (hipMalloc(&e->d_img, n * sizeof(double)) != hipSuccess) { e->err = "hipMalloc image"; e->img_n = 0; return CSX_E_NOMEM; }
        e->img_n = n;
    }
    rc = csx_engine_calc_device(e, precision, e->d_img, b4);
    if (rc != CSX_OK) return rc;
    if (csx_synchronize(e->ctx) != CSX_OK ||
        hipMemcpy(h_out, e->d_img, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        e->err = "copy image to host";
        return CSX_E_HIP;
    }
    return CSX_OK;
}

int csx_engine_set_fit_data(csx_engine* e, int w, int h, const double* h_data) {
    if (!e) return CSX_E_ARG;
    if (!e->ctx) { e->err = "host-only engine has no device"; return CSX_E_STATE; }
    (void)hipSetDevice(e->device);
    if (e->d_fit) (void)hipFree(e->d_fit);
    e->d_fit = nullptr;
    e->fw = w;
    e->fh = h;
    if (!h_data || w <= 0 || h <= 0) return CSX_E_ARG;
    if (hipMalloc(&e->d_fit, (size_t)w * h * sizeof(double)) != hipSuccess) { e->err = "hipMalloc fit data"; return CSX_E_NOMEM; }
    if (hipMemcpy(e->d_fit, h_data, (size_t)w * h * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        e->err = "copy fit data";
        return CSX_E_HIP;
    }
    return CSX_OK;
}

int csx_engine_set_nofit_rect(csx_engine* e, int id, int x0, int y0, int x1, int y1) {
    if (!e || id < 0 || id >= 4) return CSX_E_ARG;
    e->rects[4 * id] = x0; e->rects[4 * id + 1] = y0; e->rects[4 * id + 2] = x1; e->rects[4 * id + 3] = y1;
    return CSX_OK;
}

int csx_engine_fit(csx_engine* e, int precision, int bstop, int border, double* h_sum, int64_t* h_cnt,
                   int64_t* h_nan) {
    if (!e || !h_sum) return CSX_E_ARG;
    *h_sum = 1e20;
    if (!e->d_fit) { e->err = "no fit data"; return CSX_E_STATE; }
    int rc = e->prepared ? CSX_OK : engine_prepare(e, false);  // no checkAllParameters in the fit path
    if (rc != CSX_OK) return rc;
    csx_mask m{};
    std::memcpy(m.rects, e->rects, sizeof m.rects);
    m.bstop = bstop;
    m.border = border;
    rc = csx_fit_fqs(e->ctx, e->fw, e->fh, &m, precision, e->d_fit, h_sum, h_cnt, h_nan, nullptr);
    if (rc != CSX_OK) {
        char b[256];
        csx_last_error(e->ctx, b, sizeof b);
        e->err = b;
        *h_sum = 1e20;
    }
    return rc;
}

const void* csx_engine_table(csx_engine* e, const char* name, int64_t* n, int* kind) {
    if (!e || !name || !n || !kind) return nullptr;
    static const char* carr_names[18] = {"carr1p", "carr2p", "carr3p", "carr4p", "carr5p", "carr6p",
                                         "carr7p", "carr8p", "carr9p", "carr1f", "carr2f", "carr3f",
                                         "carr4f", "carr5f", "carr6f", "carr7f", "carr8f", "carr9f"};
    for (int a = 0; a < 18; ++a)
        if (std::strcmp(name, carr_names[a]) == 0) { *n = CSX_NCOEF; *kind = 0; return e->P.carr.data() + a * CSX_NCOEF; }
    if (std::strcmp(name, "carr11pm") == 0) { *n = (int64_t)e->P.carr11pm.size(); *kind = 0; return e->P.carr11pm.data(); }
    if (std::strcmp(name, "limq") == 0) { *n = 18; *kind = 0; return e->P.s.limq; }
    if (std::strcmp(name, "latpar1") == 0) { *n = (int64_t)(e->P.s.peakmax1 + 1) * 6; *kind = 1; return e->P.latpar1.data(); }
    if (std::strcmp(name, "latpar2") == 0) { *n = (int64_t)(e->P.s.peakmax2 + 1) * 6; *kind = 1; return e->P.latpar2.data(); }
    if (std::strcmp(name, "latpar3") == 0) { *n = (int64_t)(e->P.s.peakmax2 + 1) * 17; *kind = 2; return e->P.latpar3.data(); }
    *n = 0;
    return nullptr;
}

csx_rand* csx_rand_new(uint32_t seed) {
    csx_rand* r = new csx_rand();
    csr::seed(r->s, seed);
    return r;
}

void csx_rand_free(csx_rand* r) { delete r; }

int32_t csx_rand_next(csx_rand* r) { return r ? csr::value(csr::step(r->s)) : 0; }

int csx_rand_skip(csx_rand* r, int64_t n) {
    if (!r || n < 0) return CSX_E_ARG;
    if (n < 4 * csr::kDeg) {
        for (int64_t i = 0; i < n; ++i) csr::step(r->s);
    } else {
        std::vector<uint32_t> rows(csr::kRows);
        csr::jump_rows((uint64_t)n, rows.data());
        csr::advance(r->s, rows.data());
    }
    return CSX_OK;
}

int csx_rand_zuf(csx_rand* r, double* out, int64_t n) {
    if (!r || (!out && n > 0) || n < 0) return CSX_E_ARG;
    for (int64_t i = 0; i < n; ++i) out[i] = zuf(r);
    return CSX_OK;
}

int csx_engine_tpv(csx_engine* e, const char* ids, csx_rand* rng, char* report, int cap) {
    if (!e) return CSX_E_ARG;
    std::string rep;
    if (!ids || !*ids) {  // clear (cpu:1033-1049)
        e->tpv_on = false;
        e->p.width_zuf = 1.0;
        e->prepared = false;
        rep = "TPV: clear";
    } else {
        if (!rng) return CSX_E_ARG;
        csx::ParamSet& t = e->pt;
        const double width_zuf = e->tpv_on ? t.width_zuf : e->p.width_zuf;
        t = e->p;  // the console sets every parameter from its store before each call
        t.width_zuf = width_zuf;  // not a parameter: keeps the last random factor
        t.dw_zuf = 1.0;
        if (!e->tpv_on) {
            for (int k = 0; k < kNTpv; ++k) e->tpv_old[k] = e->p.*kTpv[k].field;
            e->tpv_old[10] = width_zuf;
        }
        const char* q = ids;
        while (*q) {
            const char* nl = std::strchr(q, '\n');
            std::string line = nl ? std::string(q, nl) : std::string(q);
            q = nl ? nl + 1 : q + line.size();
            if (line.empty()) continue;
            const size_t pos = line.find('=');
            const std::string key = line.substr(0, pos);
            std::string val = pos == std::string::npos ? line : line.substr(pos + 1);
            double dval;
            try {
                dval = std::stod(val);
                if (dval < 0.00001 && val.length() > 2) {  // decimal comma written by another locale
                    size_t pk = val.find('.');
                    if (pk != std::string::npos) val[pk] = ',';
                    else if ((pk = val.find(',')) != std::string::npos) val[pk] = '.';
                    dval = std::stod(val);
                }
            } catch (const std::exception&) {
                e->err = "TPV: not a number in '" + line + "'";
                return CSX_E_ARG;
            }
            int k = 0;
            while (k < kNTpv && key != kTpv[k].key) ++k;
            if (k == kNTpv) continue;  // a name the list does not have
            double shown = 0.0;
            switch (kTpv[k].flag) {
            case TPV_NORM: shown = t.*kTpv[k].field = e->tpv_old[k] * (1 + dval * zuf(rng)); break;
            case TPV_UCA: shown = t.*kTpv[k].field = t.uca; break;
            case TPV_PHI: shown = t.phi = 2.0 + (zuf(rng) + 1) * 43.0; break;
            case TPV_AZI: shown = 4.0 / e->tpv_old[k] * (1 + dval * zuf(rng)); break;
            case TPV_DW: {
                const double z = zuf(rng);
                t.dw_zuf = 1 + dval * z;
                shown = e->tpv_old[k] * e->tpv_old[k] / 3.0 * t.dw_zuf;
                break;
            }
            }
            char num[64];
            std::snprintf(num, sizeof num, "%f", shown);
            // EOL of Source/sc_math.h:15-19: "\r\n" unless Q_OS_LINUX is defined before it is
            // included, which the engine's include order does not do
            rep += key + " = " + num + "\r\n";
        }
        // consistency of the particle's dimensions (cpu:1136-1140)
        if ((t.interior == 1 || t.particle == 3) && t.length <= t.radius) t.length = t.radius + t.length;
        if (t.particle == 2 && t.length <= 2. * t.radius) t.length = 2. * t.radius + t.length;
        e->tpv_on = true;
        e->prepared = false;
    }
    if (report && cap > 0) std::snprintf(report, (size_t)cap, "%s", rep.c_str());
    return CSX_OK;
}

double csx_engine_tpv_value(csx_engine* e, const char* key) {
    double v = 0;
    if (!e || !key || !(e->tpv_on ? e->pt : e->p).get(key, &v)) return 0.0 / 0.0;
    return v;
}

int csx_engine_stage_image(csx_engine* e, csx_ctx* ctx, double* d_out, int64_t cap) {
    if (!e || !ctx || !d_out) return CSX_E_ARG;
    int rc = e->prepared ? CSX_OK : engine_prepare_host(e);
    if (rc != CSX_OK) return rc;
    const csx::PreparedSet& P = e->P;
    const int w = P.s.xmax - P.s.xmin, h = P.s.ymax - P.s.ymin;
    if ((int64_t)w * h > cap) { e->err = "image larger than its output buffer"; return CSX_E_ARG; }
    csx_mask m{};
    for (int i = 0; i < 16; ++i) m.rects[i] = -1;
    rc = csx_fit_batch_stage(ctx, &P.s, P.carr.data(), P.latpar1.empty() ? nullptr : P.latpar1.data(),
                             P.latpar2.empty() ? nullptr : P.latpar2.data(), P.latpar3.empty() ? nullptr : P.latpar3.data(),
                             w, h, &m, d_out);
    if (rc != CSX_OK) {
        char b[256];
        csx_last_error(ctx, b, sizeof b);
        e->err = b;
    }
    return rc;
}

int csx_engine_tpv_batch(csx_engine* e, csx_ctx* ctx, const char* ids, csx_rand* rng, int n, double* d_out,
                         int64_t stride, double* d_noise, int nextra, double* h_extra, int threads) {
    if (!e || !ctx || !ids || !*ids || !rng || n <= 0 || !d_out || nextra < 0 || (nextra && !h_extra))
        return CSX_E_ARG;
    // the post-processing calls take at most 65535 images per batch (grid y)
    if (n > 65535) { e->err = "at most 65535 images per batch"; return CSX_E_ARG; }
    // the image size of the parameters as set (the random variation keeps the grid)
    int rc = e->prepared ? CSX_OK : engine_prepare_host(e);
    if (rc != CSX_OK) return rc;
    const int64_t npx = (int64_t)(e->P.s.xmax - e->P.s.xmin) * (e->P.s.ymax - e->P.s.ymin);
    if (npx > stride) { e->err = "image larger than its output stride"; return CSX_E_ARG; }
    // On any failure below the generator and the random-variation state are put back as they
    // were, so that a retry draws the reference's sequence from the same point.
    const csr::State s0 = rng->s;
    const csx::ParamSet pt0 = e->pt;
    const bool on0 = e->tpv_on;
    double old0[kNTpv];
    std::copy(e->tpv_old, e->tpv_old + kNTpv, old0);
    auto restore = [&](int code) {
        rng->s = s0;
        e->pt = pt0;
        e->tpv_on = on0;
        std::copy(old0, old0 + kNTpv, e->tpv_old);
        return code;
    };
    // draws in the AI-data loop's order, image after image: variation, noise, then the extra
    // values (beam stop, lines) of the caller's post-processing.  The noise is jumped over
    // here, its start state kept, and drawn on the device.
    std::vector<csx::ParamSet> sets((size_t)n);
    std::vector<uint32_t> wins, rows;
    if (d_noise) {
        wins.resize((size_t)n * csr::kDeg);
        rows.resize(csr::kRows);
        csr::jump_rows((uint64_t)npx, rows.data());
    }
    for (int k = 0; k < n; ++k) {
        if ((rc = csx_engine_tpv(e, ids, rng, nullptr, 0)) != CSX_OK) return restore(rc);
        sets[(size_t)k] = e->pt;
        if (d_noise) {
            csr::window(rng->s, &wins[(size_t)k * csr::kDeg]);
            csr::advance(rng->s, rows.data());
        }
        if (nextra) csx_rand_zuf(rng, h_extra + (size_t)k * nextra, nextra);
    }
    // preparation and the host part of staging on threads
    std::vector<csx::PreparedSet> preps((size_t)n);
    std::vector<std::unique_ptr<csx_staged, void (*)(csx_staged*)>> st;
    st.reserve((size_t)n);
    for (int k = 0; k < n; ++k) st.emplace_back(csx_staged_new(), csx_staged_free);
    std::vector<std::string> errs((size_t)n);
    csx_mask m{};
    for (int i = 0; i < 16; ++i) m.rects[i] = -1;
    std::atomic<int> next{0};
    // lattice sets: the device builds the reflection tables (csx_set_device_prep)
    const bool dev = csx_ctx_device_prep(ctx) != 0;
    auto work = [&] {
        for (int k = next++; k < n; k = next++) {
            const csx::ParamSet& ps = sets[(size_t)k];
            const bool rows = dev && ps.ltype != 12 && ps.peakidx + 1 == 2;  // Gaussian peaks
            std::string msg = csx::check_parameters(ps, e->ignore_cbs);
            if (msg.empty()) msg = csx::prepare(ps, preps[(size_t)k], rows);
            if (!msg.empty()) { errs[(size_t)k] = msg; continue; }
            const csx::PreparedSet& P = preps[(size_t)k];
            const int w = P.s.xmax - P.s.xmin, h = P.s.ymax - P.s.ymin;
            if (rows) {
                csx_stage_build_rows(st[(size_t)k].get(), P, w, h, &m, d_out + (size_t)k * stride);
                continue;
            }
            csx_stage_build(st[(size_t)k].get(), &P.s, P.carr.data(), P.latpar1.empty() ? nullptr : P.latpar1.data(),
                            P.latpar2.empty() ? nullptr : P.latpar2.data(),
                            P.latpar3.empty() ? nullptr : P.latpar3.data(), w, h, &m, d_out + (size_t)k * stride);
        }
    };
    const int nt = std::max(1, std::min(threads > 0 ? threads : 1, n));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    for (int k = 0; k < n; ++k)
        if (!errs[(size_t)k].empty()) { e->err = errs[(size_t)k]; return restore(CSX_E_PARAM); }
    std::vector<const csx_staged*> built((size_t)n);
    for (int k = 0; k < n; ++k) built[(size_t)k] = st[(size_t)k].get();
    rc = csx_stage_commit_many(ctx, n, built.data(), nt);
    if (rc == CSX_OK && d_noise) rc = csx_post_rand_zuf(ctx, wins.data(), n, npx, d_noise);
    if (rc != CSX_OK) {
        char b[256];
        csx_last_error(ctx, b, sizeof b);
        e->err = b;
        return restore(rc);
    }
    return rc;
}

int csx_engine_stage_points(csx_engine* e, csx_ctx* ctx, int n, int npar, const char* const* keys, const double* x,
                            double* d_out, int64_t stride, int threads) {
    if (!e || !ctx || n <= 0 || npar <= 0 || !keys || !x || !d_out || stride <= 0) return CSX_E_ARG;
    for (int j = 0; j < npar; ++j)
        if (!keys[j] || !*keys[j]) return CSX_E_ARG;
    std::vector<csx::PreparedSet> preps((size_t)n);
    std::vector<std::unique_ptr<csx_staged, void (*)(csx_staged*)>> st;
    st.reserve((size_t)n);
    for (int k = 0; k < n; ++k) st.emplace_back(csx_staged_new(), csx_staged_free);
    std::vector<std::string> errs((size_t)n);
    csx_mask m{};
    for (int i = 0; i < 16; ++i) m.rects[i] = -1;
    const bool dev = csx_ctx_device_prep(ctx) != 0;
    std::atomic<int> next{0};
    auto work = [&] {
        for (int k = next++; k < n; k = next++) {
            csx::ParamSet ps = e->p;
            for (int j = 0; j < npar; ++j) {  // a key "a,b" sets both names
                std::string name;
                for (const char* c = keys[j];; ++c) {
                    if (*c == ',' || *c == 0) {
                        if (!name.empty() && !ps.set(name, x[(size_t)k * npar + j])) errs[(size_t)k] = "unknown key " + name;
                        name.clear();
                        if (*c == 0) break;
                    } else {
                        name += *c;
                    }
                }
            }
            if (!errs[(size_t)k].empty()) continue;
            const bool rows = dev && ps.ltype != 12 && ps.peakidx + 1 == 2;  // Gaussian peaks
            std::string msg = csx::check_parameters(ps, e->ignore_cbs);
            if (msg.empty()) msg = csx::prepare(ps, preps[(size_t)k], rows);
            if (!msg.empty()) { errs[(size_t)k] = msg; continue; }
            const csx::PreparedSet& P = preps[(size_t)k];
            const int w = P.s.xmax - P.s.xmin, h = P.s.ymax - P.s.ymin;
            if ((int64_t)w * h > stride) { errs[(size_t)k] = "image larger than its output stride"; continue; }
            if (rows) {
                csx_stage_build_rows(st[(size_t)k].get(), P, w, h, &m, d_out + (size_t)k * stride);
                continue;
            }
            csx_stage_build(st[(size_t)k].get(), &P.s, P.carr.data(), P.latpar1.empty() ? nullptr : P.latpar1.data(),
                            P.latpar2.empty() ? nullptr : P.latpar2.data(),
                            P.latpar3.empty() ? nullptr : P.latpar3.data(), w, h, &m, d_out + (size_t)k * stride);
        }
    };
    const int nt = std::max(1, std::min(threads > 0 ? threads : 1, n));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    for (int k = 0; k < n; ++k)
        if (!errs[(size_t)k].empty()) { e->err = errs[(size_t)k]; return CSX_E_PARAM; }
    std::vector<const csx_staged*> built((size_t)n);
    for (int k = 0; k < n; ++k) built[(size_t)k] = st[(size_t)k].get();
    const int rc = csx_stage_commit_many(ctx, n, built.data(), nt);
    if (rc != CSX_OK) {
        char b[256];
        csx_last_error(ctx, b, sizeof b);
        e->err = b;
    }
    return rc;
}

const csx_scalars* csx_engine_scalars(csx_engine* e) { return e ? &e->P.s : nullptr; }
csx_ctx* csx_engine_ctx(csx_engine* e) { return e ? e->ctx : nullptr; }

int csx_engine_last_error(csx_engine* e, char* buf, int len) {
    if (!e || !buf || len <= 0) return CSX_E_ARG;
    std::snprintf(buf, (size_t)len, "%s", e->err.c_str());
    return CSX_OK;
}

}  // extern "C"
