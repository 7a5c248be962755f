// SasCalc_GENERIC_calculation facade over the C ABI (csx_engine_*).
//
// Mirrors the reference engine's public surface used by sc_calcCons and
// sc_simplexfit2d (Source/sc_calc_generic_gpu.h:108-598): parameters by key,
// doCalculation, setFitData / setNoFitRect / doFitCalculation, prepared tables.
// Host preparation runs here in C++; every pixel runs on the GPU.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "csx.h"
#include "csx_prep.h"

struct csx_engine {
    csx::ParamSet p;
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
        csx_destroy(e->ctx);
    }
    delete e;
}

int csx_engine_set(csx_engine* e, const char* key, double v) {
    if (!e || !key) return CSX_E_ARG;
    if (std::strcmp(key, "ignore_cbs") == 0) { e->ignore_cbs = v != 0; e->prepared = false; return CSX_OK; }
    if (!e->p.set(key, v)) { e->err = std::string("unknown key ") + key; return CSX_E_ARG; }
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
        const std::string m = csx::check_parameters(e->p, e->ignore_cbs);
        if (!m.empty()) { e->err = m; return CSX_E_PARAM; }
    }
    const std::string m = csx::prepare(e->p, e->P);
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

int csx_engine_calc_device(csx_engine* e, int precision, double* d_out, int32_t* b4) {
    if (!e || !d_out) return CSX_E_ARG;
    if (!e->ctx) { e->err = "host-only engine has no device"; return CSX_E_STATE; }
    int rc = e->prepared ? CSX_OK : engine_prepare(e, true);  // tables stay in HBM until a parameter changes
    if (rc != CSX_OK) return rc;
    const csx_scalars& s = e->P.s;
    csx_grid g{s.xmin, s.xmax, s.ymin, s.ymax};
    if (b4) { b4[0] = s.xmin; b4[1] = s.xmax; b4[2] = s.ymin; b4[3] = s.ymax; }
    rc = csx_render(e->ctx, &g, precision, d_out);
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
    const size_t n = (size_t)(s.xmax - s.xmin) * (size_t)(s.ymax - s.ymin);
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

const csx_scalars* csx_engine_scalars(csx_engine* e) { return e ? &e->P.s : nullptr; }
csx_ctx* csx_engine_ctx(csx_engine* e) { return e ? e->ctx : nullptr; }

int csx_engine_last_error(csx_engine* e, char* buf, int len) {
    if (!e || !buf || len <= 0) return CSX_E_ARG;
    std::snprintf(buf, (size_t)len, "%s", e->err.c_str());
    return CSX_OK;
}

}  // extern "C"
