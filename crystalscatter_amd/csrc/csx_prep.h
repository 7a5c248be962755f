// Host-side preparation of one parameter set for the device path.
//
// Produces exactly what the reference's prepareCalculation leaves for the
// per-pixel code (Source/sc_calc_generic_cpu.cpp:319-988): the series
// coefficient tables and their switch points (coefficients(), :2099-8090),
// the FCC/BCT reflection tables (ButtonHKLClick :8246, fhkl_c :10299,
// corotations :1487, table fill :838-954) and the per-pixel scalars, in the
// flat layout of include/csx.h.  Runs in FP64 with x87 long double where the
// reference's recurrences use it, so the tables agree with the reference's.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "csx.h"

namespace csx {

// The reference key table (Source/sc_calc_generic.cpp:36-138) with its defaults.
struct ParamSet {
    int ltype = 12, particle = 0, ordis = 7, interior = 0, peakidx = 1, hklmax = 5, gridpoints = 64;
    int radq1 = 0, radq2 = 0, radq4 = 1, expand = 0, centerbeam = 0, twinned = 0, dsch = 0, para = 0;
    int waxs = 0, fix_orcase = 0;
    int use1d = 0, qsteps = 200;  // 1D mode (the console's prepareData argument) and EditQsteps
    double uca = 30, ucb = 30, ucc = 30, ucalpha = 90, ucbeta = 90, ucgamma = 90, ceff = 1;
    double ceffcyl = 0, reff = 0, acpl = 0, bcpl = 0, ifluc = 0, rfluc = 0;
    double radius = 1.0, radiusi = 0.5, sigma = 0.07, dbeta = 0.4, length = 1, sigmal = 0.06;
    double alphash = 0, rho = 0, rotalpha = 0;
    double peakpar = 0, displacement = 2, azi = 180, domainsize = 400;
    double ucpsi = 0, ucn1 = 1, ucn2 = 1, ucn3 = 1, theta = 90, phi = 0, rottheta = 90, rotphi = 90;
    double qmax = 1, qmin = 0, pixnox = 2048, pixnoy = 2048, pixx_mm = 1, pixy_mm = 1, det = 1;
    double wave = 0.09499, beamx = 0, beamy = 0, bfactor = 1, p1 = 0, iso = 0, i0 = 10000, base = 0;
    double width_zuf = 1;
    double dw_zuf = 1;  // random factor on the Debye-Waller term (TPV), 1 = none
    // Peak-shape axes and domain sizes (setPeakShapeMatrix): accepted and kept for a complete
    // key table; none of the device combinations reads them.
    double vax1 = 1, vax2 = 0, vax3 = 0, ay1 = 0, ay2 = 1, ay3 = 0, az1 = 0, az2 = 0, az3 = 1;
    double sigx = 40, sigy = 40, sigz = 40;

    // Returns false for an unknown key.
    bool set(const std::string& key, double v);
    bool get(const std::string& key, double* v) const;
    uint32_t combo_key() const;
};

// Reflection lists of one lattice type and HKLmax, the part of the tables that does not
// depend on the lattice constants: latpar1 ((n1+1) x 6) and latpar2 ((n2+2) x 6) exactly as
// prepare fills them, and each entry's sphno.
struct LatticeLists {
    int ltype = -1, hklmax = -1, n1 = 0, n2 = 0;
    std::vector<int32_t> l1, l2;
    std::vector<double> s1, s2;
};
// Built once per (ltype, HKLmax) in the process; thread-safe.
std::shared_ptr<const LatticeLists> lattice_lists(int ltype, int hklmax);

struct PreparedSet {
    csx_scalars s{};
    std::array<double, 18 * CSX_NCOEF> carr{};  // carr1p..carr9p, carr1f..carr9f
    std::vector<double> carr11pm;               // CSX_NIM2*CSX_NIM2 (Gaussian cylinders)
    std::vector<int32_t> latpar1, latpar2;      // (peakmax+1) x 6
    std::vector<float> latpar3;                 // (peakmax2+1) x 17
    double order = 0, ucvol = 0, sphno = 1;
    int xmin = 0, xmax = 0, ymin = 0, ymax = 0; // computed block (ihex, i)
    int axmin = 0, axmax = 0, aymin = 0, aymax = 0;  // image array (whole image when expanded)
    int cdim = 3;
    // Lattice kinds: the reflection lists and the rotated reciprocal bases (rows a*, b*, c* of
    // corotations' matrix and of its twinned counterpart) from which latpar3 is computed.
    std::shared_ptr<const LatticeLists> lists;
    double rot[9] = {}, rott[9] = {};
};

// checkAllParameters (Source/sc_calc_generic_cpu.cpp:198-317): empty string when valid.
std::string check_parameters(const ParamSet& p, bool ignore_cbs);

// prepareCalculation.  Returns empty string on success, else the reason.  With
// rows_on_device, lattice kinds leave latpar1/2/3 empty: the caller builds the tables on the
// device from out.lists, out.rot and out.rott (csx_stage_build_rows).
std::string prepare(const ParamSet& p, PreparedSet& out, bool rows_on_device = false);

// Reference Gamma (Source/sc_calc_generic_gpu.cu:613-673), needed on host and device.
double ref_gamma(double z);
// gaussnorm3 (Source/sc_calc_generic_gpu.cu:838-854).
double gauss_norm3(double a);
// lorentznorm3 (.cu:809-826), pearsonnorm3 (.cu:868-946), and the peak norm of latpar3 entry 11
// for peak shape shp (Source/sc_calc_generic_cpu.cpp:899-928).
double lorentz_norm3(double a);
double pearson_norm3(double a, double b);
double peak_norm(int shp, double x2phi, double beta);

}  // namespace csx

// Parameters of a facade engine (csx_engine.cpp), for the batched fit driver.
struct csx_engine;
const csx::ParamSet* csx_engine_paramset(const csx_engine* e);

// Staging of one set of a batched objective in two steps (csx_device.hip):
// csx_stage_build does the host work (checks, fast peak tables, the set's bytes) and touches
// no context, so the batched fit driver runs it on many threads; csx_stage_commit appends
// the result to the context's pending batch.  csx_fit_batch_stage is the two in a row.
struct csx_staged;
csx_staged* csx_staged_new();
void csx_staged_free(csx_staged* s);
void csx_stage_build(csx_staged* out, const csx_scalars* s, const double* h_carr, const int32_t* h_l1,
                     const int32_t* h_l2, const float* h_l3, int w, int h, const csx_mask* mask,
                     const double* d_data);
int csx_stage_commit(csx_ctx* c, const csx_staged* st);
// The same for a lattice set prepared with rows_on_device: its tables are built on the device
// (k_prep_sets) from P.lists, P.rot and P.rott when the batch runs.
void csx_stage_build_rows(csx_staged* out, const csx::PreparedSet& P, int w, int h, const csx_mask* mask,
                          const double* d_data);
// Whether the batched fit driver stages lattice sets for device preparation (csx_set_device_prep).
int csx_ctx_device_prep(const csx_ctx* c);
// n built sets at once, their bytes copied on up to `threads` threads.
int csx_stage_commit_many(csx_ctx* c, int n, const csx_staged* const* sts, int threads);
// csx_fit_batch_run in two halves: enqueue without waiting, then wait for the sums (the
// batched fit driver prepares another group of fits in between).
int csx_fit_batch_launch(csx_ctx* c, int precision);
int csx_fit_batch_wait(csx_ctx* c, double* h_sums, int64_t* h_cnts);
// ExpandImage: the image array of the uploaded parameter set from its computed block
// (csx_device.hip), ordered on the context's stream.
int csx_expand_image(csx_ctx* c, const double* d_region, double* d_full);
// A second context on the same device, with the same batch settings (kernel variant, graphs,
// device preparation).
int csx_create_like(const csx_ctx* c, csx_ctx** out);
// The context's device and stream, and its error message (for the post-processing kernels
// of csx_post.hip, which live outside csx_device.hip).
int csx_ctx_device(const csx_ctx* c);
void* csx_ctx_stream(const csx_ctx* c);
int csx_ctx_fail(csx_ctx* c, int code, const char* msg);
// Device scratch `which` (0 or 1) of the context, at least `bytes` long; work on the
// context's stream is ordered, so its calls can reuse it.  Null on failure.
void* csx_ctx_scratch(csx_ctx* c, int which, size_t bytes);
// nseq runs of len zuf() values on the device, run k from the generator state whose 31 terms
// (oldest first) are h_win[31k .. 31k+30], into d_out + k*len; ordered on the context's
// stream, the host arrays free to reuse on return.
int csx_post_rand_zuf(csx_ctx* c, const uint32_t* h_win, int64_t nseq, int64_t len, double* d_out);
