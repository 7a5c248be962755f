// Host preparation of one parameter set (see csx_prep.h for the contract).
//
// Arithmetic follows the reference's expressions and operand types where
// they decide the rounded table values (long double recurrences, float
// rotation entries, float reflection table); structure and names are this
// project's own.
#include "csx_prep.h"
#include "csx_rows.h"

#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>

namespace csx {

namespace {

constexpr int kNN = 120;              // nnmax, Source/sc_calc_generic_cpu.cpp:2103
constexpr double kPi = 3.14159265358979323846;
constexpr double kEps9 = 0.000000001; // Source/sc_math.h:45

inline double sq(double x) { return x * x; }

// ---------------------------------------------------------------------------
// Key table
// ---------------------------------------------------------------------------
struct KeyDef {
    const char* name;
    int ParamSet::*i;
    double ParamSet::*d;
};

const KeyDef kKeys[] = {
    {"LType", &ParamSet::ltype, nullptr}, {"uca", nullptr, &ParamSet::uca},
    {"ucb", nullptr, &ParamSet::ucb}, {"ucc", nullptr, &ParamSet::ucc},
    {"ucalpha", nullptr, &ParamSet::ucalpha}, {"ucbeta", nullptr, &ParamSet::ucbeta},
    {"ucgamma", nullptr, &ParamSet::ucgamma}, {"EditCeff", nullptr, &ParamSet::ceff},
    {"CheckBoxTwinned", &ParamSet::twinned, nullptr}, {"EditCeffcyl", nullptr, &ParamSet::ceffcyl},
    {"reff", nullptr, &ParamSet::reff}, {"acpl", nullptr, &ParamSet::acpl},
    {"bcpl", nullptr, &ParamSet::bcpl}, {"ifluc", nullptr, &ParamSet::ifluc},
    {"rfluc", nullptr, &ParamSet::rfluc}, {"ComboBoxParticle", &ParamSet::particle, nullptr},
    {"Ordis", &ParamSet::ordis, nullptr}, {"ComboBoxInterior", &ParamSet::interior, nullptr},
    {"EditRadius", nullptr, &ParamSet::radius}, {"EditRadiusi", nullptr, &ParamSet::radiusi},
    {"EditSigma", nullptr, &ParamSet::sigma}, {"EditDbeta", nullptr, &ParamSet::dbeta},
    {"Length", nullptr, &ParamSet::length}, {"SigmaL", nullptr, &ParamSet::sigmal},
    {"Alpha", nullptr, &ParamSet::alphash}, {"EditRho", nullptr, &ParamSet::rho},
    {"RotAlpha", nullptr, &ParamSet::rotalpha}, {"ComboBoxPeak", &ParamSet::peakidx, nullptr},
    {"EditPeakPar", nullptr, &ParamSet::peakpar}, {"EditDebyeWaller", nullptr, &ParamSet::displacement},
    {"EditAzi", nullptr, &ParamSet::azi}, {"EditDomainSize", nullptr, &ParamSet::domainsize},
    {"RadButDebyeScherrer", &ParamSet::dsch, nullptr}, {"RadioButtonPara", &ParamSet::para, nullptr},
    {"ucpsi", nullptr, &ParamSet::ucpsi}, {"ucn1", nullptr, &ParamSet::ucn1},
    {"ucn2", nullptr, &ParamSet::ucn2}, {"ucn3", nullptr, &ParamSet::ucn3},
    {"theta", nullptr, &ParamSet::theta}, {"phi", nullptr, &ParamSet::phi},
    {"rotTheta", nullptr, &ParamSet::rottheta}, {"rotPhi", nullptr, &ParamSet::rotphi},
    {"GridPoints", &ParamSet::gridpoints, nullptr}, {"HKLmax", &ParamSet::hklmax, nullptr},
    {"EditQmax", nullptr, &ParamSet::qmax}, {"EditQmin", nullptr, &ParamSet::qmin},
    {"RadioButtonQ1", &ParamSet::radq1, nullptr}, {"RadioButtonQ2", &ParamSet::radq2, nullptr},
    {"RadioButtonQ4", &ParamSet::radq4, nullptr}, {"ExpandImage", &ParamSet::expand, nullptr},
    {"EditPixelNoX", nullptr, &ParamSet::pixnox}, {"EditPixelNoY", nullptr, &ParamSet::pixnoy},
    {"EditPixelX", nullptr, &ParamSet::pixx_mm}, {"EditPixelY", nullptr, &ParamSet::pixy_mm},
    {"EditDet", nullptr, &ParamSet::det}, {"EditWavelength", nullptr, &ParamSet::wave},
    {"BeamPosX", nullptr, &ParamSet::beamx}, {"BeamPosY", nullptr, &ParamSet::beamy},
    {"CenterBeam", &ParamSet::centerbeam, nullptr}, {"EditBFactor", nullptr, &ParamSet::bfactor},
    {"CheckBoxWAXS", &ParamSet::waxs, nullptr}, {"P1", nullptr, &ParamSet::p1},
    {"iso", nullptr, &ParamSet::iso}, {"I0", nullptr, &ParamSet::i0}, {"Base", nullptr, &ParamSet::base},
    {"width_zuf", nullptr, &ParamSet::width_zuf}, {"dw_zuf", nullptr, &ParamSet::dw_zuf},
    {"fix_orcase", &ParamSet::fix_orcase, nullptr}, {"use1d", &ParamSet::use1d, nullptr},
    {"EditQsteps", &ParamSet::qsteps, nullptr},
    {"VAx1", nullptr, &ParamSet::vax1}, {"VAx2", nullptr, &ParamSet::vax2}, {"VAx3", nullptr, &ParamSet::vax3},
    {"Ay1", nullptr, &ParamSet::ay1}, {"Ay2", nullptr, &ParamSet::ay2}, {"Ay3", nullptr, &ParamSet::ay3},
    {"Az1", nullptr, &ParamSet::az1}, {"Az2", nullptr, &ParamSet::az2}, {"Az3", nullptr, &ParamSet::az3},
    {"SigX", nullptr, &ParamSet::sigx}, {"SigY", nullptr, &ParamSet::sigy}, {"SigZ", nullptr, &ParamSet::sigz},
};

// ---------------------------------------------------------------------------
// Romberg integration with the reference's refinement and stopping rules
// (qrombdeltac/trapzddeltac, Source/sc_calc_generic_gpu.cu:956-1488): trapezoid
// levels refined by midpoints, k=5-point polynomial extrapolation to h=0,
// relative tolerance eps, at most 20 levels.
// ---------------------------------------------------------------------------
void neville0(const double* xa, const double* ya, int n, double& y, double& dy) {
    double c[8], d[8];
    int ns = 1;
    double dif = std::fabs(xa[1]);
    for (int i = 1; i <= n; ++i) {
        const double dift = std::fabs(xa[i]);
        if (dift < dif) { ns = i; dif = dift; }
        c[i] = ya[i];
        d[i] = ya[i];
    }
    y = ya[ns];
    ns = ns - 1;
    for (int m = 1; m <= n - 1; ++m) {
        for (int i = 1; i <= n - m; ++i) {
            const double ho = xa[i], hp = xa[i + m];
            const double w = c[i + 1] - d[i];
            double den = ho - hp;
            if (den != 0) den = w / den;
            d[i] = hp * den;
            c[i] = ho * den;
        }
        if (2 * ns < n - m) dy = c[ns + 1];
        else { dy = d[ns]; ns = ns - 1; }
        y = y + dy;
    }
}

template <class F>
double romberg(F&& f, double a, double b, double eps, double tiny) {
    const int jmax = 20, k = 5;
    double h[24], s[24], xa[8], ya[8];
    double res = 0, err = 0;
    int cnt = 1;
    h[1] = 1.0;
    for (int j = 1; j <= jmax; ++j) {
        if (j == 1) {
            s[j] = 0.5 * (b - a) * (f(a) + f(b));
            cnt = 1;
        } else {
            const double tnm = cnt, del = (b - a) / tnm;
            double x = a + 0.5 * del, sum = 0.0;
            for (int i = 1; i <= cnt; ++i) { sum = sum + f(x); x = x + del; }
            s[j] = 0.5 * (s[j] + (b - a) * sum / tnm);
            cnt = 2 * cnt;
        }
        if (j >= k) {
            for (int i = 1; i <= k; ++i) { xa[i] = h[j - k + i]; ya[i] = s[j - k + i]; }
            neville0(xa, ya, k, res, err);
            if (std::fabs(err) < (eps * std::fabs(res))) break;
            if (std::fabs(res) < tiny) break;
        }
        s[j + 1] = s[j];
        h[j + 1] = 0.25 * h[j];
    }
    return res;
}

// Gaussian orientation weight integrals over delta in [0, delmax] (float limit),
// as qrombdeltac for i0 = 1/2 (weight) and 3 (order parameter).
double delta_weight_integral(double dbeta_deg, bool order_param) {
    const double dbeta = dbeta_deg * kPi / 180.0;
    double delmax = dbeta * std::sqrt(std::log(1.0 / 0.00001));
    if (delmax > kPi / 2.0) delmax = kPi / 2.0;
    const float lo = 0.0f, hi = static_cast<float>(delmax);
    auto w = [&](double x) {
        if (order_param) return std::sin(x) * std::exp(-x * x / (dbeta * dbeta)) * (3 * std::cos(x) * std::cos(x) - 1) / 2.0;
        return std::sin(x) * std::exp(-x * x / (dbeta * dbeta));
    };
    return romberg(w, lo, hi, 1.0e-4, 1.0e-100);
}

// ---------------------------------------------------------------------------
// Series coefficients (coefficients(), the branches the six kernels use)
// ---------------------------------------------------------------------------
struct CoefState {
    double* arr[18];      // carr1p..carr9p, carr1f..carr9f
    int nidx[18];
};

// Homogeneous sphere, Source/sc_calc_generic_cpu.cpp:2509-2543.
void coef_sphere(PreparedSet& P, double z, double xr2z, CoefState& cs) {
    long double fkv[kNN + 2], gam3[kNN + 2], z12v[kNN + 1];
    fkv[0] = 1; gam3[0] = std::sqrt(kPi) / 2.0; z12v[0] = 1.0;
    long double xrn = 1.0;
    double* c4p = cs.arr[3];
    double* c4f = cs.arr[12];
    for (int n = 1; n <= kNN; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xrn = -xrn * xr2z;
        c4p[n] = 9 * std::sqrt(kPi) * std::pow(4.0, n) * z12v[n] * xrn /
                 (2.0 * (n + 3) * (n + 2) * (n + 3 / 2.0) * gam3[n] * fkv[n]);
        double bin = 0.0;
        for (int m = 0; m <= n; ++m)
            bin += z12v[m] * z12v[n - m] /
                   ((m + 3 / 2.0) * gam3[m] * (n - m + 3 / 2.0) * gam3[n - m] * fkv[m] * fkv[n - m]);
        c4f[n] = 9 * kPi * xrn * bin / 16.0;
        if (std::fabs(c4p[n]) < 1E-30 && n < cs.nidx[3]) cs.nidx[3] = n;
        if (std::fabs(c4f[n]) < 1E-30 && n < cs.nidx[12]) cs.nidx[12] = n;
    }
    (void)P;
}

// Cube, homogeneous, isotropic: Source/sc_calc_generic_cpu.cpp:2757-2829.
// P(q): carr4p[n] = sqrt(pi) 4^n z12v[n] (-xr2z)^n (u*u*u)[n] / (16 gam3[n]) with
// u[n] = 1/((n+1/2)(n+1) n!) (u[0] = 2), a triple self-convolution that depends on n only.
// <F>^2: carr4f[n] = pi^2 (-xr2z)^n T[n] / (128 gam3[n]), T a triple convolution of
// fsum[n] gam3[n]/(n+1/2) that depends on z only (O(n^3), cached per thread for the last z).
// Storage types follow the reference: fsum, u, v and the convolution accumulators are
// doubles, fkv, gam3 and z12v long doubles, so the tables equal the reference's bit for bit.
struct CubeSums {
    double z = -1.0;
    double tf[kNN + 1];
    bool have_p = false;
    double tp[kNN + 1];  // (u*u*u)[n]
};

void coef_cube(PreparedSet& P, double z, double xr2z, double radius, CoefState& cs) {
    static thread_local CubeSums cache;
    const double area = 6.0 * 4.0 * sq(radius);
    const double vol = 8.0 * radius * sq(radius);
    P.s.por = 2 * kPi * std::pow(z + 1, 4) * area / (z * (z - 1) * (z - 2) * (z - 3) * vol * vol);
    long double fkv[kNN + 2], gam3[kNN + 2], z12v[kNN + 1];
    fkv[0] = 1; gam3[0] = std::sqrt(kPi) / 2.0; z12v[0] = 1.0;
    for (int n = 1; n <= kNN; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
    }
    if (!cache.have_p) {
        double u[kNN + 1], v[kNN + 1];
        u[0] = 2;
        for (int n = 1; n <= kNN; ++n) u[n] = 1 / ((n + 1 / 2.0) * (n + 1) * fkv[n]);
        for (int n = 0; n <= kNN; ++n) {
            double acc = 0.0;
            for (int m = 0; m <= n; ++m) acc += u[n - m] * u[m];
            v[n] = acc;
        }
        for (int n = 1; n <= kNN; ++n) {
            double acc = 0.0;
            for (int m = 0; m <= n; ++m) acc += u[n - m] * v[m];
            cache.tp[n] = acc;
        }
        cache.have_p = true;
    }
    if (cache.z != z) {
        double fsum[kNN + 1];
        fsum[0] = 1;
        for (int n = 1; n <= kNN; ++n) {
            long double acc = 0.0;
            for (int m = 0; m <= n; ++m) acc += z12v[n - m] * z12v[m] / (gam3[n - m] * gam3[m] * fkv[n - m] * fkv[m]);
            fsum[n] = static_cast<double>(acc);
            // T[n] = sum_m w[n-m] (sum_k fsum[m-k] fsum[k] gam3[m-k] gam3[k] / ((k+1/2)(m-k+1/2)))
            double t = 0.0;
            for (int m = 0; m <= n; ++m) {
                double inner = 0.0;
                for (int k = 0; k <= m; ++k)
                    inner += fsum[m - k] * fsum[k] * gam3[m - k] * gam3[k] / ((k + 1 / 2.0) * (m - k + 1 / 2.0));
                t += inner * fsum[n - m] * gam3[n - m] / (n - m + 1 / 2.0);
            }
            cache.tf[n] = t;
        }
        cache.z = z;
    }
    double* c4p = cs.arr[3];
    double* c4f = cs.arr[12];
    long double xrn = 1.0;
    for (int n = 1; n <= kNN; ++n) {
        xrn = -xrn * xr2z;
        c4p[n] = std::sqrt(kPi) * std::pow(4.0, n) * z12v[n] * xrn * cache.tp[n] / (16.0 * gam3[n]);
        c4f[n] = kPi * kPi * xrn * cache.tf[n] / (128.0 * gam3[n]);
        if (std::fabs(c4p[n]) < 1E-30 && n < cs.nidx[3]) cs.nidx[3] = n;
        if (std::fabs(c4f[n]) < 1E-30 && n < cs.nidx[12]) cs.nidx[12] = n;
    }
}

// The O(n^2) inner sums of the <F>^2 coefficients (carr1f, carr4f) depend only on the
// Schulz parameters z and zl, not on the sizes: kept per thread for the last (z, zl), so a
// sequence of preparations that varies R, L or the lattice (a fit) computes them once.  The
// cached values are the same doubles the loop below would produce.
struct FSums {
    double z = -1.0, zl = -1.0;
    double acc1[kNN + 1], acc4[kNN + 1];
    // the factors of carr1p / carr4p without the powers of xl2z, xr2z (numerator, denominator)
    long double a1[kNN + 1], d1[kNN + 1], b4[kNN + 1], e4[kNN + 1];
};

// Cube, perfect orientation, homogeneous (Source/sc_calc_generic_cpu.cpp:2880-2910): the
// double-series table carr11pm[n][m], m <= n (the other entries keep their initial 1, which the
// pixel reads as well), and carr4p/carr4f.
void coef_cube_zdir(PreparedSet& P, double z, double xr2z, CoefState& cs) {
    long double fkv[kNN + 2], gam3[kNN + 2], z12v[kNN + 1];
    fkv[0] = 1; gam3[0] = std::sqrt(kPi) / 2.0; z12v[0] = 1.0;
    long double xrn = 1.0;
    double* pm = P.carr11pm.data();
    double* c4p = cs.arr[3];
    double* c4f = cs.arr[12];
    for (int n = 1; n <= kNN; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xrn = -xrn * xr2z;
        for (int m = 0; m <= n; ++m)
            pm[n * CSX_NIM2 + m] = (kPi / 4.0) * std::pow(4.0, n) * z12v[n - m] * z12v[m] * xrn /
                                   (gam3[n - m] * gam3[m] * (n - m + 1) * fkv[n - m] * (m + 1) * fkv[m]);
        c4p[n] = std::sqrt(kPi) * std::pow(4.0, n) * xrn / (16.0 * gam3[n]);
        c4f[n] = kPi * kPi * xrn / (128.0 * gam3[n]);
        if (std::fabs(c4p[n]) < 1E-30 && n < cs.nidx[3]) cs.nidx[3] = n;
        if (std::fabs(c4f[n]) < 1E-30 && n < cs.nidx[12]) cs.nidx[12] = n;
    }
}

// Cylinder, isotropic, homogeneous, L/R >= 2 (the factorisation branch,
// Source/sc_calc_generic_cpu.cpp:3686-3746): P|| and P- series and the <F>^2 series.
void coef_cylinder_iso(double z, double zl, double xl2z, double xr2z, CoefState& cs) {
    long double fkv[kNN + 2], fk2v[kNN + 1], gam3[kNN + 2], z12v[kNN + 1], z12vl[kNN + 1];
    fkv[0] = 1; fk2v[0] = 1; gam3[0] = std::sqrt(kPi) / 2.0; z12v[0] = 1.0; z12vl[0] = 1;
    long double xln = 1.0;
    double xrn = 1.0;
    double *c1p = cs.arr[0], *c4p = cs.arr[3], *c1f = cs.arr[9], *c4f = cs.arr[12];
    for (int n = 1; n <= kNN; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        z12vl[n] = z12vl[n - 1] * ((zl + 1) - 2 + 2 * n) * ((zl + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        fk2v[n] = fk2v[n - 1] * (2 * n - 1) * (2 * n);
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xln = -xln * xl2z;
        xrn = -xrn * xr2z;
        c1p[n] = std::sqrt(kPi) * std::pow(4.0, n) * z12vl[n] * xln / (2.0 * (2 * n + 1) * (n + 1) * gam3[n] * fkv[n]);
        c4p[n] = 4 * (n + 1 / 2.0) * fk2v[n] * z12v[n] * xrn /
                 ((n + 2) * (n + 1) * fkv[n] * fkv[n] * (n + 1) * fkv[n] * fkv[n]);
        double acc = 0.0;
        for (int m = 0; m <= n; ++m) acc += z12vl[m] * z12vl[n - m] / (gam3[m] * gam3[n - m] * fkv[m] * fkv[n - m]);
        c1f[n] = kPi * xln * acc / (4.0 * (2 * n + 1));
        acc = 0.0;
        for (int m = 0; m <= n; ++m)
            acc += z12v[m] * z12v[n - m] / ((m + 1) * fkv[m] * (n - m + 1) * fkv[n - m] * fkv[m] * fkv[n - m]);
        c4f[n] = xrn * acc;
        if (std::fabs(c1p[n]) < 1E-30 && n < cs.nidx[0]) cs.nidx[0] = n;
        if (std::fabs(c1f[n]) < 1E-30 && n < cs.nidx[9]) cs.nidx[9] = n;
        if (std::fabs(c4p[n]) < 1E-30 && n < cs.nidx[3]) cs.nidx[3] = n;
        if (std::fabs(c4f[n]) < 1E-30 && n < cs.nidx[12]) cs.nidx[12] = n;
    }
}

// Cylinder, perfect orientation, homogeneous: Source/sc_calc_generic_cpu.cpp:4124-4212.
void coef_cylinder_zdir(double z, double zl, double xl2z, double xr2z, CoefState& cs) {
    // Everything but the powers of xl2z and xr2z depends on z and zl only, and is kept per
    // thread while they stay the same (fit vertices and training images change R and L far
    // more often than the widths).  The products are the same operations in the same order as
    // when formed inline, so the coefficients do not change by a bit.
    static thread_local FSums fs;
    if (!(fs.z == z && fs.zl == zl)) {
        long double fkv[kNN + 2], fk2v[kNN + 1], gam3[kNN + 2], z12vl[kNN + 1];
        double z12v[kNN + 1];
        fkv[0] = 1; fk2v[0] = 1; gam3[0] = std::sqrt(kPi) / 2.0; z12vl[0] = 1; z12v[0] = 1.0;
        double p4 = 1.0;  // 4^n, exact
        const double sqrt_pi = std::sqrt(kPi);
        for (int n = 1; n <= kNN; ++n) {
            p4 *= 4.0;
            z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
            z12vl[n] = z12vl[n - 1] * ((zl + 1) - 2 + 2 * n) * ((zl + 1) - 1 + 2 * n);
            fkv[n] = fkv[n - 1] * n;
            fk2v[n] = fk2v[n - 1] * (2 * n - 1) * (2 * n);
            gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
            fs.a1[n] = sqrt_pi * p4 * z12vl[n];
            fs.d1[n] = 2.0 * (n + 1) * gam3[n] * fkv[n];
            fs.b4[n] = 4 * (n + 1 / 2.0) * fk2v[n] * z12v[n];
            fs.e4[n] = (n + 2) * (n + 1) * fkv[n] * fkv[n] * (n + 1) * fkv[n];
            double acc = 0.0;
            for (int m = 0; m <= n; ++m)
                acc += z12vl[m] * z12vl[n - m] / (gam3[m] * gam3[n - m] * fkv[m] * fkv[n - m] * fkv[m] * fkv[n - m]);
            fs.acc1[n] = acc;
            acc = 0.0;
            for (int m = 0; m <= n; ++m)
                acc += z12v[m] * z12v[n - m] / ((m + 1) * fkv[m] * (n - m + 1) * fkv[n - m] * fkv[m] * fkv[n - m]);
            fs.acc4[n] = acc;
        }
        fs.z = z;
        fs.zl = zl;
    }
    long double xln = 1.0;
    double xrn = 1.0;
    bool find1 = true, find4 = true;
    double *c1p = cs.arr[0], *c4p = cs.arr[3], *c1f = cs.arr[9], *c4f = cs.arr[12];
    for (int n = 1; n <= kNN; ++n) {
        xln = -xln * xl2z;
        xrn = -xrn * xr2z;
        c1p[n] = fs.a1[n] * xln / fs.d1[n];
        c4p[n] = fs.b4[n] * xrn / fs.e4[n];
        c1f[n] = kPi * xln * fs.acc1[n] / 4.0;
        c4f[n] = xrn * fs.acc4[n];
        // The reference's switch-point search stops at the first coefficient below 1e-50
        // with its sign, i.e. at the first negative term (n=1).
        if (find1 && c1p[n] < 1e-50) { cs.nidx[0] = n; find1 = false; }
        if (find4 && c4p[n] < 1e-50) { cs.nidx[3] = n; find4 = false; }
        if (std::fabs(c1f[n]) < 1E-30 && n < cs.nidx[9]) cs.nidx[9] = n;
        if (std::fabs(c4f[n]) < 1E-30 && n < cs.nidx[12]) cs.nidx[12] = n;
    }
}

// Cylinder, Gaussian orientation, general case: Source/sc_calc_generic_cpu.cpp:4648-4720 and
// the second general-orientation block :5549-5618 (which overwrites carr1p and fills carr3p).
void coef_cylinder_gauss(PreparedSet& P, const ParamSet& p, double z, double zl, double xl2z,
                         double xr2z, double polTheta, double& polPhi, CoefState& cs) {
    const double norm = delta_weight_integral(p.dbeta, false);
    P.s.norm = norm;
    P.order = delta_weight_integral(p.dbeta, true) / norm;
    if (P.s.orcase != 0) return;
    polPhi = 0;  // phi is forced to 0 for the general case (:4657)
    const double p11 = -std::cos(polPhi * kPi / 180.0) * std::cos(polTheta * kPi / 180.0);
    const double p13 = std::cos(polPhi * kPi / 180.0) * std::sin(polTheta * kPi / 180.0);
    const double p21 = -std::cos(polPhi * kPi / 180.0);
    // The reference integrates intlar[n][m] with the weight only (the cos^2n sin^2m factor
    // of the original is absent), so every entry equals this one integral.
    const double intl = delta_weight_integral(p.dbeta, false);

    long double fkv[kNN + 2], fk2v[kNN + 1], gam3[kNN + 2], z12vl[kNN + 1];
    double z12v[kNN + 1];
    fkv[0] = 1; fk2v[0] = 1; gam3[0] = std::sqrt(kPi) / 2.0; z12vl[0] = 1; z12v[0] = 1.0;
    double xrn = 1.0;
    long double xln = 1.0;
    bool find1 = true, find4 = true;
    double *c1p = cs.arr[0], *c2p = cs.arr[1], *c4p = cs.arr[3], *c1f = cs.arr[9], *c2f = cs.arr[10],
           *c4f = cs.arr[12];
    double* pm = P.carr11pm.data();
    for (int n = 1; n <= kNN; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        z12vl[n] = z12vl[n - 1] * ((zl + 1) - 2 + 2 * n) * ((zl + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        fk2v[n] = fk2v[n - 1] * (2 * n - 1) * (2 * n);
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xln = -xln * xl2z;
        xrn = -xrn * xr2z;
        for (int m = 0; m <= n; ++m) {
            double acc = 0.0;
            for (int ll = 0; ll <= m; ++ll)
                acc += std::pow(p11 * p11, ll) * std::pow(p13 * p13, m - ll) * intl /
                       (std::pow(4.0, ll) * fk2v[m - ll] * fkv[n - m + ll] * fkv[ll] * norm);
            pm[n * CSX_NIM2 + m] = std::pow(4.0, m) * std::pow(sq(p21), n - m) * acc / (fkv[n - m]);
        }
        c1p[n] = std::pow(4.0, n) * z12vl[n] * xln / ((2.0 * n + 1) * (n + 1));
        c2p[n] = std::pow(4.0, n + 1) * gam3[n] * z12vl[n] * xln /
                 (std::sqrt(kPi) * (n + 2) * (n + 1) * (n + 1) * fkv[n] * fkv[n] * fkv[n]);
        double acc = 0.0;
        for (int m = 0; m <= n; ++m)
            acc = acc + z12vl[m] * z12vl[n - m] / (gam3[m] * gam3[n - m] * fkv[m] * fkv[n - m]);
        c1f[n] = fk2v[n] * xln * acc / (std::pow(4.0, n));
        c2f[n] = c2p[n];
        c4p[n] = 4 * (n + 1 / 2.0) * fk2v[n] * z12v[n] * xrn /
                 ((n + 2) * (n + 1) * fkv[n] * fkv[n] * (n + 1) * fkv[n]);
        acc = 0.0;
        for (int m = 0; m <= n; ++m)
            acc = acc + z12v[m] * z12v[n - m] / ((m + 1) * fkv[m] * (n - m + 1) * fkv[n - m] * fkv[m] * fkv[n - m]);
        c4f[n] = xrn * acc;
        if (find1 && c1p[n] < 1e-50) { cs.nidx[0] = n; find1 = false; }
        if (find4 && c4p[n] < 1e-50) { cs.nidx[3] = n; find4 = false; }
        if (std::fabs(c1f[n]) < 1E-30 && n < cs.nidx[9]) cs.nidx[9] = n;
        if (std::fabs(c4f[n]) < 1E-30 && n < cs.nidx[12]) cs.nidx[12] = n;
    }
    // Second general-orientation block.
    double* c3p = cs.arr[2];
    double b1sv = 1.0, zv = 1.0, xr = 1.0;
    long double xl = 1.0;
    bool find3 = true;
    for (int n = 1; n <= kNN; ++n) {
        zv = zv * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        b1sv = b1sv * (2.0 - 1 + n);
        xl = -xl * xl2z;
        xr = -xr * xr2z;
        c1p[n] = std::pow(4.0, 2 * n) * zv * xl / ((2.0 * n + 1) * (n + 1));
        c3p[n] = 4 * (n + 1 / 2.0) * fk2v[n] * zv * xr / ((n + 2) * (n + 1) * fkv[n] * fkv[n] * b1sv * fkv[n]);
        if (find3 && c3p[n] < 1e-50) { cs.nidx[2] = n; find3 = false; }
    }
}

// ---------------------------------------------------------------------------
// Reflection tables
// ---------------------------------------------------------------------------
inline bool ev(int v) { return (v % 2) == 0; }

// Pattern of the index triple (zero / equal indices), Source/sc_calc_generic_cpu.cpp:9780-9803.
int pattern(int h, int k, int l) {
    enum { P000, P0KL, PH0L, PHK0, PH00, P0K0, P00L, PHHL, PHLH, PLHH, PHH0, PH0H, P0HH, PHHH, PHKL };
    int c = PHKL;
    if (h == 0 && k == 0 && l == 0) c = P000;
    if (h == 0 && k != 0 && l != 0 && k != l) c = P0KL;
    if (k == 0 && h != 0 && l != 0 && h != l) c = PH0L;
    if (l == 0 && h != 0 && k != 0 && h != k) c = PHK0;
    if (h != 0 && k == 0 && l == 0) c = PH00;
    if (k != 0 && h == 0 && l == 0) c = P0K0;
    if (l != 0 && h == 0 && k == 0) c = P00L;
    if (h != 0 && h == k && l != 0 && l != h) c = PHHL;
    if (h != 0 && h == l && k != 0 && k != h) c = PHLH;
    if (k != 0 && k == l && h != 0 && h != k) c = PLHH;
    if (h != 0 && h == k && l == 0) c = PHH0;
    if (h != 0 && h == l && k == 0) c = PH0H;
    if (k != 0 && k == l && h == 0) c = P0HH;
    if (h != 0 && h == k && h == l) c = PHHH;
    return c;
}

// Multiplicities of the cubic (48/24/12/8/6) and tetragonal (16/8/4/2) systems,
// indexed by pattern (Source/sc_calc_generic_cpu.cpp:9815-9862).
const int kMultCubic[15] = {0, 24, 24, 24, 6, 6, 6, 24, 24, 24, 12, 12, 12, 8, 48};
const int kMultTetra[15] = {0, 8, 8, 8, 4, 4, 2, 8, 8, 8, 4, 8, 8, 8, 16};

// Cubic space groups by their reflection conditions (ButtonHKLClick, Source/sc_calc_generic_cpu.cpp:
// 8444-8500 picks the group of each lattice type; the codes are SETCX's c1, c2, c7, c10 at
// :8529-8600; the conditions are :9941-9998):
//   integral  -1 none, 0 h+k+l even (I), 6 h+k, h+l, k+l even (F);
//   zonal     0kl with k, l the two indices after the zero one:  0 k+l even, 3 k+l = 4n with k
//             and l even, 4 k and l even, 5 the index that follows the zero one cyclically even;
//   serial    00l: 0 l even, 1 l = 4n;
//   hhl       two equal indices r and a third u != 0:  0 u even, 2 r+u even, 4 2r+u = 4n with u
//             even, 5 r and u even;
// -1 leaves a class unconditioned.  Every condition applies to each cyclic permutation.
struct CubicGroup {
    int integral, zonal, serial, hhl;
};

bool cubic_group(int ltype, CubicGroup& g) {
    switch (ltype) {
    case 4: case 11: g = {0, 0, 0, 0}; return true;  // BCC, Im3m: I---
    case 5: g = {6, 4, 0, 2}; return true;           // FCC: F---
    case 7: g = {-1, -1, -1, -1}; return true;       // SC: P---
    case 9: g = {0, 4, 1, 4}; return true;           // Ia3d: Ia-d
    case 10: g = {-1, 0, 0, -1}; return true;        // Pn3m: Pn--
    case 17: g = {6, 3, 1, 2}; return true;          // Fd3m: Fd--
    default: return false;
    }
}

bool cubic_allowed(const CubicGroup& g, int h, int k, int l) {
    const int v[3] = {h, k, l};
    const bool integral = g.integral == -1 || (g.integral == 0 && ev(h + k + l)) ||
                          (g.integral == 6 && ev(h + k) && ev(h + l) && ev(k + l));
    if (!integral) return false;
    for (int z = 0; z < 3; ++z) {
        // zonal: v[z] == 0, the others in index order (0kl: k,l; h0l: h,l; hk0: h,k)
        if (v[z] == 0) {
            const int a = v[z == 0 ? 1 : 0], b = v[z == 2 ? 1 : 2], next = v[(z + 1) % 3];
            switch (g.zonal) {
            case 0: if (!ev(a + b)) return false; break;
            case 3: if (!((a + b) % 4 == 0 && ev(a) && ev(b))) return false; break;
            case 4: if (!(ev(a) && ev(b))) return false; break;
            case 5: if (!ev(next)) return false; break;
            default: break;
            }
        }
        // serial: the two indices other than v[z] are zero
        if (v[(z + 1) % 3] == 0 && v[(z + 2) % 3] == 0) {
            if (g.serial == 0 && !ev(v[z])) return false;
            if (g.serial == 1 && v[z] % 4 != 0) return false;
        }
        // hhl: the two indices other than v[z] are equal, v[z] != 0
        const int r = v[(z + 1) % 3], u = v[z];
        if (r == v[(z + 2) % 3] && u != 0) {
            switch (g.hhl) {
            case 0: if (!ev(u)) return false; break;
            case 2: if (!ev(r + u)) return false; break;
            case 4: if (!((2 * r + u) % 4 == 0 && ev(u))) return false; break;
            case 5: if (!(ev(r) && ev(u))) return false; break;
            default: break;
            }
        }
    }
    return true;
}

// Reflection conditions of a lattice type: the cubic groups above; BCT = tetragonal I---
// (hkl: h+k+l; 0kl: k+l; hk0: h+k; hhl: l; 00l: l; 0k0: k).  HCP (hexagonal system) and the
// orthorhombic lattice admit no reflection: the reference's condition blocks of those systems
// are commented out, so its acceptance flag keeps its initial 0 (ButtonHKLClick,
// Source/sc_calc_generic_cpu.cpp:8256, :10059-10145, :10249).
bool allowed(int ltype, int h, int k, int l) {
    if (h == 0 && k == 0 && l == 0) return false;
    if (ltype >= 1 && ltype <= 3) return true;  // 2D systems accept every reflection (:9936)
    if (ltype == 6 || ltype == 18) return false;
    CubicGroup g;
    if (cubic_group(ltype, g)) return cubic_allowed(g, h, k, l);
    if (!ev(h + k + l)) return false;
    if ((h == 0 && !ev(k + l)) || (k == 0 && !ev(h + l))) return false;
    if (l == 0 && !ev(h + k)) return false;
    if (h == k && l != 0 && !ev(l)) return false;
    if ((h == 0 && l == 0 && !ev(k)) || (k == 0 && l == 0 && !ev(h))) return false;
    if (h == 0 && k == 0 && !ev(l)) return false;
    return true;
}

// |F(hkl)|^2 of the lattice's basis with float coordinates: FCC 4 atoms, Fd3m 8 (the FCC four
// and the four at quarter positions), BCC, BCT, Ia3d and Im3m 2 (body-centred); SC, Pn3m and
// the hexagonal and square 2D lattices 1; the centred rectangular 2D lattice 2, in the plane
// (fhkl_c, Source/sc_calc_generic_cpu.cpp:10330-10371, 10374-10563).
double structure_factor(int ltype, int h, int k, int l, double& sphno) {
    if (ltype == 7 || ltype == 10 || ltype == 1 || ltype == 2) {
        sphno = 1;
        return 1;
    }
    if (ltype == 3) {
        const float r = 0.5f;
        const double arg = 2 * kPi * (h * r + k * r);
        const double c = 1.0 + std::cos(arg), sn = 0.0 + std::sin(arg);
        sphno = 2;
        return std::round(c * c + sn * sn);
    }
    static const float diamond[8][3] = {{0, 0, 0}, {0.5f, 0.5f, 0}, {0, 0.5f, 0.5f}, {0.5f, 0, 0.5f},
                                        {0.25f, 0.25f, 0.25f}, {0.75f, 0.75f, 0.25f},
                                        {0.25f, 0.75f, 0.75f}, {0.75f, 0.25f, 0.75f}};
    static const float body[2][3] = {{0, 0, 0}, {0.5f, 0.5f, 0.5f}};
    const float(*r)[3] = (ltype == 5 || ltype == 17) ? diamond : body;
    const int n = ltype == 17 ? 8 : ltype == 5 ? 4 : 2;
    sphno = n;
    double c = 0.0, s = 0.0;
    for (int i = 0; i < n; ++i) {
        const double arg = 2 * kPi * (h * r[i][0] + k * r[i][1] + l * r[i][2]);
        c = c + std::cos(arg);
        s = s + std::sin(arg);
    }
    return std::round(c * c + s * s);
}

// Reciprocal basis rotated so the crystal direction (u,v,w) is along the beam, with the
// reference's float intermediates (corotations, Source/sc_calc_generic_cpu.cpp:1487-2098).
// Out: R (row-major reciprocal vectors a*,b*,c*), RT (twinned cell), vol.
// lat2d (the 2D lattices): R is the rotated in-plane reciprocal basis of the cell (a, b,
// gamma) with c* = (0, 0, 1), and vol the cell area (:1613-1615, :1676-1729, :1935-1943,
// :2060-2066); RT stays the 3D twin.
void rotated_reciprocal_basis(double a, double b, double c, double al, double be, double ga,
                              double u, double v, double w, double psi, bool lat2d, double R[9], double RT[9],
                              double& vol) {
    al *= kPi / 180.0; be *= kPi / 180.0; ga *= kPi / 180.0; psi *= kPi / 180.0;
    const double ca = std::cos(al), cb = std::cos(be), cg = std::cos(ga), sb = std::sin(be);
    const double m1 = (cg - cb * ca) / sb, m2 = std::sqrt(1.0 - m1 * m1 - ca * ca);
    // cell matrix rows a(l), b(m), c(n) and its inverse (twice), giving M
    const double S[9] = {a * sb, a * 0, a * cb, b * m1, b * m2, b * ca, c * 0, c * 0, c * 1};
    auto inv3 = [](const double* A, double* Bo) {
        const double d = A[0] * A[4] * A[8] + A[1] * A[5] * A[6] + A[2] * A[3] * A[7] - A[6] * A[4] * A[2] -
                         A[7] * A[5] * A[0] - A[8] * A[3] * A[1];
        Bo[0] = (A[4] * A[8] - A[5] * A[7]) / d; Bo[1] = (A[2] * A[7] - A[1] * A[8]) / d;
        Bo[2] = (A[1] * A[5] - A[2] * A[4]) / d; Bo[3] = (A[5] * A[6] - A[3] * A[8]) / d;
        Bo[4] = (A[0] * A[8] - A[2] * A[6]) / d; Bo[5] = (A[2] * A[3] - A[0] * A[5]) / d;
        Bo[6] = (A[3] * A[7] - A[4] * A[6]) / d; Bo[7] = (A[1] * A[6] - A[0] * A[7]) / d;
        Bo[8] = (A[0] * A[4] - A[1] * A[3]) / d;
    };
    double Ms[9], MsT[9], Mt[9];
    inv3(S, Ms);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) MsT[3 * i + j] = Ms[3 * j + i];
    inv3(MsT, Mt);
    double M[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[3 * i + j] = Mt[3 * j + i];
    // cell vectors (columns of M) and twinned cell, rounded to float
    auto col = [&](float x, float y, float z, float* o) {
        o[0] = M[0] * x + M[1] * y + M[2] * z;
        o[1] = M[3] * x + M[4] * y + M[5] * z;
        o[2] = M[6] * x + M[7] * y + M[8] * z;
    };
    float ae[3], bev[3], ce[3], aet[3], bet[3], cet[3];
    col(1, 0, 0, ae); col(0, 1, 0, bev); col(0, 0, 1, ce);
    col(2 / 3.0, 2 / 3.0, -1 / 3.0, aet); col(-1 / 3.0, 2 / 3.0, 2 / 3.0, bet); col(2 / 3.0, -1 / 3.0, 2 / 3.0, cet);
    vol = ae[0] * (bev[1] * ce[2] - bev[2] * ce[1]) + ae[1] * (bev[2] * ce[0] - bev[0] * ce[2]) +
          ae[2] * (bev[0] * ce[1] - bev[1] * ce[0]);
    // rotation taking the (u,v,w) direction onto the beam (float arithmetic)
    const float nx = M[0] * u + M[1] * v + M[2] * w, ny = M[3] * u + M[4] * v + M[5] * w,
                nz = M[6] * u + M[7] * v + M[8] * w;
    const float len = std::sqrt(nx * nx + ny * ny + nz * nz);
    const float nxn = nx / len, nzn = nz / len;
    const float th = std::acos(-nzn);
    const float ep = std::asin(nxn / std::sin(th));
    const float r[9] = {float(std::cos(ep) * std::cos(psi) - std::sin(ep) * std::cos(th) * std::sin(psi)),
                        float(-std::cos(ep) * std::sin(psi) - std::sin(ep) * std::cos(th) * std::cos(psi)),
                        float(std::sin(ep) * std::sin(th)),
                        float(std::sin(ep) * std::cos(psi) + std::cos(ep) * std::cos(th) * std::sin(psi)),
                        float(-std::sin(ep) * std::sin(psi) + std::cos(ep) * std::cos(th) * std::cos(psi)),
                        float(-std::cos(ep) * std::sin(th)),
                        float(std::sin(th) * std::sin(psi)), float(std::sin(th) * std::cos(psi)), float(std::cos(th))};
    const double dr = r[0] * r[4] * r[8] + r[1] * r[5] * r[6] + r[2] * r[3] * r[7] - r[6] * r[4] * r[2] -
                      r[7] * r[5] * r[0] - r[8] * r[3] * r[1];
    const double q[9] = {(r[4] * r[8] - r[5] * r[7]) / dr, (r[2] * r[7] - r[1] * r[8]) / dr,
                         (r[1] * r[5] - r[2] * r[4]) / dr, (r[5] * r[6] - r[3] * r[8]) / dr,
                         (r[0] * r[8] - r[2] * r[6]) / dr, (r[2] * r[3] - r[0] * r[5]) / dr,
                         (r[3] * r[7] - r[4] * r[6]) / dr, (r[1] * r[6] - r[0] * r[7]) / dr,
                         (r[0] * r[4] - r[1] * r[3]) / dr};
    auto rot_f = [&](const float* x, float* o) {
        o[0] = q[0] * x[0] + q[1] * x[1] + q[2] * x[2];
        o[1] = q[3] * x[0] + q[4] * x[1] + q[5] * x[2];
        o[2] = q[6] * x[0] + q[7] * x[1] + q[8] * x[2];
    };
    auto rot_d = [&](const float* x, double* o) {
        o[0] = q[0] * x[0] + q[1] * x[1] + q[2] * x[2];
        o[1] = q[3] * x[0] + q[4] * x[1] + q[5] * x[2];
        o[2] = q[6] * x[0] + q[7] * x[1] + q[8] * x[2];
    };
    float ad[3], bd[3], cd[3];
    double adt[3], bdt[3], cdt[3];
    rot_f(ae, ad); rot_f(bev, bd); rot_f(ce, cd);
    rot_d(aet, adt); rot_d(bet, bdt); rot_d(cet, cdt);
    // reciprocal vectors by cross products over the volume (float products, double division)
    R[0] = (bd[1] * cd[2] - bd[2] * cd[1]) / vol; R[1] = (bd[2] * cd[0] - bd[0] * cd[2]) / vol;
    R[2] = (bd[0] * cd[1] - bd[1] * cd[0]) / vol; R[3] = (ad[2] * cd[1] - ad[1] * cd[2]) / vol;
    R[4] = (ad[0] * cd[2] - ad[2] * cd[0]) / vol; R[5] = (ad[1] * cd[0] - ad[0] * cd[1]) / vol;
    R[6] = (ad[1] * bd[2] - ad[2] * bd[1]) / vol; R[7] = (ad[2] * bd[0] - ad[0] * bd[2]) / vol;
    R[8] = (ad[0] * bd[1] - ad[1] * bd[0]) / vol;
    RT[0] = (bdt[1] * cdt[2] - bdt[2] * cdt[1]) / vol; RT[1] = (bdt[2] * cdt[0] - bdt[0] * cdt[2]) / vol;
    RT[2] = (bdt[0] * cdt[1] - bdt[1] * cdt[0]) / vol; RT[3] = (adt[2] * cdt[1] - adt[1] * cdt[2]) / vol;
    RT[4] = (adt[0] * cdt[2] - adt[2] * cdt[0]) / vol; RT[5] = (adt[1] * cdt[0] - adt[0] * cdt[1]) / vol;
    RT[6] = (adt[1] * bdt[2] - adt[2] * bdt[1]) / vol; RT[7] = (adt[2] * bdt[0] - adt[0] * bdt[2]) / vol;
    RT[8] = (adt[0] * bdt[1] - adt[1] * bdt[0]) / vol;
    if (lat2d) {
        // the plane cell's vectors (a, 0) and (b cos g, b sin g) through the unit axes (float),
        // their reciprocals over the area, rotated by q
        const double sg = std::sin(ga);
        const float one = 1.0f, zero = 0.0f;
        const double aex = a * one + (b * cg) * zero + 0.0 * zero, aey = 0.0 * one + (b * sg) * zero + 0.0 * zero;
        const double bex = a * zero + (b * cg) * one + 0.0 * zero, bey = 0.0 * zero + (b * sg) * one + 0.0 * zero;
        const double area = a * b * sg;
        const double as[3] = {bey / area, -bex / area, 0.0}, bs[3] = {-aey / area, aex / area, 0.0},
                     cs[3] = {0.0, 0.0, 1.0};
        const double* rs[3] = {as, bs, cs};
        for (int r3 = 0; r3 < 3; ++r3)
            for (int i = 0; i < 3; ++i)
                R[3 * r3 + i] = q[3 * i] * rs[r3][0] + q[3 * i + 1] * rs[r3][1] + q[3 * i + 2] * rs[r3][2];
        vol = area;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
bool ParamSet::set(const std::string& key, double v) {
    for (const auto& k : kKeys) {
        if (key == k.name) {
            if (k.i) this->*(k.i) = static_cast<int>(v);
            else this->*(k.d) = v;
            return true;
        }
    }
    return false;
}

bool ParamSet::get(const std::string& key, double* v) const {
    for (const auto& k : kKeys) {
        if (key == k.name) {
            *v = k.i ? static_cast<double>(this->*(k.i)) : this->*(k.d);
            return true;
        }
    }
    return false;
}

uint32_t ParamSet::combo_key() const {
    const uint32_t shp = static_cast<uint32_t>(peakidx + 1);
    return (static_cast<uint32_t>(particle) << 24) | (ltype == 12 ? 0u : (shp << 16)) |
           (static_cast<uint32_t>(ltype) << 8) | static_cast<uint32_t>(ordis);
}

double ref_gamma(double z) {
    static const double bm[9] = {0.0, -0.577191652, 0.988205891, -0.897056937, 0.918206857,
                                 -0.756704078, 0.482199394, -0.193527818, 0.035868343};
    static const double cm[27] = {
        0.0, 1.0000000000000000, 0.5772156649015329, -0.6558780715202538, -0.0420026350340952,
        0.1665386113822915, -0.0421977345555443, -0.0096219715278770, 0.00721894322466630,
        -0.0011651675918591, -0.0002152416741149, 0.0001280502823882, -0.0000201348547807,
        -0.0000012504934821, 0.0000011330272320, -0.0000002056338417, 0.0000000061160950,
        0.0000000050020075, -0.0000000011812746, 0.0000000001043427, 0.0000000000077823,
        -0.0000000000036968, 0.0000000000005100, -0.0000000000000206, -0.0000000000000054,
        0.0000000000000014, 0.0000000000000001};
    auto fracp = [](double x) { double ip; return std::modf(x, &ip); };
    bool neg;
    if (std::fabs(z) < kEps9) z = z + kEps9;
    if (z > 0) neg = false;
    else { neg = true; if (std::fabs(fracp(z)) < kEps9) z = z + kEps9; }
    const double az = std::fabs(z);
    if (az > 0.0 && az <= 1) {
        double s = 0;
        for (int i = 1; i <= 26; ++i) s += cm[i] * std::pow(az, i);
        return neg ? -kPi * s / (az * std::sin(kPi * az)) : 1. / s;
    }
    if (az > 1 && az <= 1000) {
        if (std::fabs(fracp(az)) < kEps9) {
            double f = 1;
            for (int i = 1; i <= std::trunc(az) - 1; ++i) f = f * i;
            return neg ? -kPi / (az * f * std::sin(kPi * az)) : f;
        }
        const double y = az - 1, di = std::trunc(y), x = az - std::trunc(az);
        double fak = y;
        if (di == 0) fak = 1;
        else for (int i = 1; i < di; ++i) fak = fak * (y - i);
        double f = 1;
        for (int i = 1; i <= 8; ++i) f = f + bm[i] * std::pow(x, i);
        return neg ? -kPi / (az * fak * f * std::sin(kPi * az)) : fak * f;
    }
    const double s = std::exp(-az) * std::exp((az - 0.5) * std::log(az)) * std::sqrt(2 * kPi);
    return neg ? -kPi / (az * s * std::sin(kPi * az)) : s;
}

double gauss_norm3(double a) {
    return rows::gauss_norm3_from(a, rows::gauss_norm3_uses_exp(a) ? std::exp(rows::gauss_norm3_exp_arg(a)) : 0.0);
}

// lorentznorm3 (Source/sc_calc_generic_gpu.cu:809-826): the integral of sin x / (1 + a x^2) over
// [0, pi], asymptotic form for a > 0.01, Taylor below.
double lorentz_norm3(double a) {
    if (a > 0.01) {
        const double b0 = std::log(1 + a * kPi * kPi) / 2.;
        const double b1 = (a * kPi * kPi - std::log(1 + a * kPi * kPi)) / 12.;
        const double b2 = (-2 * a * kPi * kPi + a * a * kPi * kPi * kPi * kPi + 2 * std::log(1 + a * kPi * kPi)) / 480.;
        const double b3 = (a * kPi * kPi * (6 - 3 * a * kPi * kPi + 2 * a * a * kPi * kPi * kPi * kPi) -
                           6 * std::log(1 + a * kPi * kPi)) / 60480.;
        return b0 / a - b1 / (a * a) + b2 / (a * a * a) - b3 / (a * a * a * a);
    }
    const double a1 = (kPi * kPi - 4);
    const double a2 = (18 - 12 * kPi * kPi + kPi * kPi * kPi * kPi) / 2.;
    const double a3 = (-1440 + 360 * kPi * kPi - 30 * kPi * kPi * kPi * kPi + kPi * kPi * kPi * kPi * kPi * kPi) / 6.;
    return 2 - a1 * a + a2 * a * a - a3 * a * a * a;
}

// pearsonnorm3 (Source/sc_calc_generic_gpu.cu:868-946): the integral of sin x / (1 + a x^2)^b
// over [0, pi] with b nudged by 1e-6: closed forms for a > 20 and a <= 0.01, else trapezoid
// levels (end points, then the midpoints of the doubling intervals) until two agree to 1e-3,
// at most 10.
double pearson_norm3(double a, double b) {
    b += 0.000001;
    if (a > 20.) return (1 - std::exp((1 - b) * std::log(1 + a * kPi * kPi))) / (2 * (b - 1)) / a;
    if (a <= 0.01) {
        const double a1 = (kPi * kPi - 4) * b;
        const double a2 = (18 - 12 * kPi * kPi + kPi * kPi * kPi * kPi) * b * (b + 1) / 2.;
        const double a3 = (-1440 + 360 * kPi * kPi - 30 * kPi * kPi * kPi * kPi + kPi * kPi * kPi * kPi * kPi * kPi) *
                          b * (b + 1) * (b + 2) / 6.;
        return 2. - a1 * a + a2 * a * a - a3 * a * a * a;
    }
    auto f = [&](double x) { return std::sin(x) / std::exp(b * std::log(1 + a * x * x)); };
    int n = 1;
    double s = 0.5 * (kPi - 0) * (f(0) + f(kPi));
    double old = s;
    for (int level = 2; level <= 10; ++level) {
        const double del = (kPi - 0) / n;
        double x = 0 + 0.5 * del, sum = 0.0;
        for (int j = 1; j <= n; ++j) {
            sum += f(x);
            x = x + del;
        }
        s = 0.5 * (s + (kPi - 0) * sum / n);
        n = 2 * n;
        if (std::fabs(1 - s / old) < 0.001) break;
        old = s;
    }
    return s;
}

// The peak norm of latpar3 entry 11 for peak shape shp (Source/sc_calc_generic_cpu.cpp:899-928).
double peak_norm(int shp, double x2phi, double beta) {
    switch (shp) {
    case 1: case 5: return lorentz_norm3(x2phi);
    case 2: case 7: return gauss_norm3(x2phi);
    case 3: return pearson_norm3(x2phi, 2);
    case 4: return pearson_norm3(x2phi, 3 / 2.0);
    case 6: return pearson_norm3(x2phi, beta);
    default: return 0.0;
    }
}

// Reflection lists of one lattice type and HKLmax (ButtonHKLClick :8246-10296, fhkl_c
// :10299-10500): every allowed (h, k, l) of the isotropic octant and of the whole box, its
// multiplicity and rounded structure factor, in latpar1/latpar2 layout.  Built once per
// (ltype, HKLmax) in the process and shared.
std::shared_ptr<const LatticeLists> lattice_lists(int ltype, int hklmax) {
    static std::mutex mu;
    static std::vector<std::shared_ptr<const LatticeLists>> cache;
    std::lock_guard<std::mutex> lock(mu);
    for (const auto& c : cache)
        if (c->ltype == ltype && c->hklmax == hklmax) return c;
    auto L = std::make_shared<LatticeLists>();
    L->ltype = ltype;
    L->hklmax = hklmax;
    CubicGroup g;
    const bool cubic = cubic_group(ltype, g);
    // 2D lattices: l runs to 0 (pass 0: l = 0; pass 1: l in [-H, 0]), multiplicity 1
    // (ButtonHKLClick :9585-9593, :9720-9722, :9812)
    const bool sys2d = ltype >= 1 && ltype <= 3;
    const int H = hklmax;
    std::vector<int> v[2];
    std::vector<double> f[2];
    for (int pass = 0; pass < 2; ++pass) {
        const int lo = pass == 0 ? 0 : -H, lmax = sys2d ? 0 : H;
        for (int h = lo; h <= H; ++h)
            for (int k = lo; k <= H; ++k)
                for (int l = lo; l <= lmax; ++l)
                    if (allowed(ltype, h, k, l)) {
                        const int pat = pattern(h, k, l);
                        v[pass].push_back(h); v[pass].push_back(k); v[pass].push_back(l);
                        v[pass].push_back(sys2d ? 1 : cubic ? kMultCubic[pat] : kMultTetra[pat]);
                        double sph;
                        f[pass].push_back(structure_factor(ltype, h, k, l, sph));
                        (pass == 0 ? L->s1 : L->s2).push_back(sph);
                    }
    }
    const int n1 = static_cast<int>(v[0].size() / 4), n2 = static_cast<int>(v[1].size() / 4);
    L->n1 = n1;
    L->n2 = n2;
    L->l1.assign(static_cast<size_t>(std::max(n1, 1) + 1) * 6, 0);  // row 1 holds the count
    L->l2.assign(static_cast<size_t>(n2 + 2) * 6, 0);
    L->l1[6] = n1;
    L->l2[6] = n2;
    for (int i = 1; i <= n1; ++i) {
        int32_t* r = &L->l1[static_cast<size_t>(i) * 6];
        for (int c = 0; c < 4; ++c) r[1 + c] = v[0][4 * (i - 1) + c];
        r[5] = static_cast<int32_t>(std::round(f[0][static_cast<size_t>(i - 1)]));
    }
    for (int i = 1; i <= n2; ++i) {
        int32_t* r = &L->l2[static_cast<size_t>(i) * 6];
        for (int c = 0; c < 3; ++c) r[1 + c] = v[1][4 * (i - 1) + c];
        r[4] = 1;
        r[5] = static_cast<int32_t>(std::round(f[1][static_cast<size_t>(i - 1)]));
    }
    cache.push_back(L);
    return L;
}

std::string check_parameters(const ParamSet& p, bool ignore_cbs) {
    if (p.uca <= 0.01) return "uca<=0.01";
    if (p.ucb <= 0.01) return "ucb<=0.01";
    if (p.ucc <= 0.01) return "ucb<=0.01";
    if (ignore_cbs) return "";
    if (p.ltype < 0 || p.ltype > 19) return "INVALID ltype value";
    if (p.particle < 0 || p.particle > 10) return "INVALID Particle value";
    if (p.ordis < 0 || p.ordis > 13) return "INVALID Ordis value";
    if (p.interior < 0 || p.interior > 4) return "INVALID Interior value";
    const int shp = p.peakidx + 1;
    if (shp < 1 || shp > 8) return "INVALID Peak value";
    // Particle forced by the lattice type (isCbsValid, Source/sc_calc_generic_cpu.cpp:119-196).
    static const int fixed_part[20] = {2, 1, 1, 1, 0, 0, 0, 0, 0, 3, 3, 3, -1, 0, 0, 0, 2, 0, 0, 0};
    const int fp = fixed_part[p.ltype];
    if (fp != -1 && p.particle != fp) return "INVALID CBParticle due to LatticeType";
    if ((p.ltype == 13 || p.ltype == 19) && shp != 7) return "INVALID CBPeak due to LatticeType";
    return "";
}

std::string prepare(const ParamSet& p, PreparedSet& P, bool rows_on_device) {
    csx_scalars& s = P.s;
    std::memset(&s, 0, sizeof(s));
    s.combo_key = p.combo_key();
    s.zmax = p.use1d ? p.qsteps : p.gridpoints;  // setGridPoints(EditQsteps, true) in 1D
    s.use1d = p.use1d;
    s.qmin = p.qmin;
    s.centerbeam = p.centerbeam;
    s.fix_orcase = p.fix_orcase;
    s.ordis = p.ordis;
    s.dsch = p.dsch; s.para = p.para; s.twin = p.twinned;
    s.qmax = p.qmax; s.pixnox = p.pixnox; s.pixnoy = p.pixnoy;
    s.pixx_m = p.pixx_mm / 1000.; s.pixy_m = p.pixy_mm / 1000.;
    s.det = p.det; s.wave = p.wave; s.beamx = p.beamx; s.beamy = p.beamy;
    s.radius = p.radius; s.sigma = p.sigma; s.length = p.length; s.sigmal = p.sigmal; s.dbeta = p.dbeta;
    s.base = p.base; s.i0 = p.i0; s.iso = p.iso; s.ifluc = p.ifluc; s.rfluc = p.rfluc;
    s.dwfactor = p.displacement * p.displacement / 3.0 * p.dw_zuf;
    s.uca = p.uca; s.ucb = p.ucb; s.domainsize = p.domainsize; s.width_zuf = p.width_zuf;
    s.phiwidth = 4. / p.azi;
    // peak-shape parameters: setPeakPar (Source/sc_calc_generic_gpu.h:224) and the width
    // constants of prepareCalculation (Source/sc_calc_generic_cpu.cpp:585-590)
    s.peak_eta = p.peakpar / 100.;
    s.peak_beta = p.peakpar;
    s.peak_c[0] = 1;
    s.peak_c[1] = 4;
    s.peak_c[2] = 4 * (std::sqrt(2.0) - 1);
    s.peak_c[3] = 4 * (std::exp(2 * std::log(2.0) / 3.0) - 1);
    s.peak_c[4] = (s.peak_beta == 0) ? 1 : std::exp(std::log(2.) / s.peak_beta) - 1;
    s.peak_gamma_b = ref_gamma(s.peak_beta);
    s.peak_gamma_bh = ref_gamma(s.peak_beta - 0.5);
    s.ceff = p.ceff;

    // quadrant ranges (:373-394); 1D: one row of zmax pixels, never expanded
    const int quad = p.radq1 ? 1 : (p.radq2 ? 2 : 4);
    if (p.use1d) {
        P.xmin = 0;
        P.xmax = s.zmax;
        P.ymin = 0;
        P.ymax = 1;
    } else {
        P.xmin = (quad == 4) ? -p.gridpoints : 0;
        P.xmax = p.gridpoints;
        P.ymin = (quad == 1) ? 0 : -p.gridpoints;
        P.ymax = p.gridpoints;
    }
    // The image array: the whole image when ExpandImage fills it from the computed quadrant
    // or half (checkArrays(-zmax, zmax, -zmax, zmax), :979-986), else the computed block.
    const bool expand = p.expand && quad != 4 && !p.use1d;
    P.axmin = expand ? -p.gridpoints : P.xmin;
    P.axmax = expand ? p.gridpoints : P.xmax;
    P.aymin = expand ? -p.gridpoints : P.ymin;
    P.aymax = expand ? p.gridpoints : P.ymax;

    double polPhi = p.phi;
    const double polTheta = p.theta;
    s.cosphi = std::cos(polPhi * kPi / 180.0); s.sinphi = std::sin(polPhi * kPi / 180.0);
    s.costheta = std::cos(polTheta * kPi / 180.0); s.sintheta = std::sin(polTheta * kPi / 180.0);

    switch (p.particle) {
    case 0: P.cdim = 3; break;
    case 1: P.cdim = 1; break;
    case 2: P.cdim = 2; break;
    default: P.cdim = p.particle; break;
    }
    switch (p.ltype) {
    case 5: case 13: case 17: case 19: s.dist = std::sqrt(2.0) * p.uca / 2.0; break;
    case 4: case 9: case 11: s.dist = std::sqrt(3.0) * p.uca / 2.0; break;
    default: s.dist = p.uca; break;
    }
    s.reldis = p.displacement / s.dist;

    // ---- coefficients() ----
    double* base = P.carr.data();
    for (int i = 0; i < 18 * CSX_NCOEF; ++i) base[i] = 1;
    CoefState cs;
    for (int a = 0; a < 18; ++a) { cs.arr[a] = base + a * CSX_NCOEF; cs.nidx[a] = kNN; }
    s.orcase = 0;
    if (polPhi == 0 && polTheta == 90) s.orcase = 1;
    if (polPhi == 90 && polTheta == 90) s.orcase = 2;
    if (polPhi == 0 && polTheta == 0) s.orcase = 3;
    if (polPhi == 90 && polTheta == 0) s.orcase = 3;
    const double z = (1.0 - sq(p.sigma)) / sq(p.sigma);
    const double zl = (1.0 - sq(p.sigmal)) / sq(p.sigmal);
    const double xl2z = sq(p.length / (2.0 * (zl + 1)));
    const double xr2z = sq(p.radius / (2.0 * (z + 1)));
    s.norm = 1;
    bool done = false;
    s.cs = p.interior;
    if (p.ordis == 6 && P.cdim == 4) {
        // perfectly oriented cubes: homogeneous only (the reference leaves the form factor of
        // the other interiors undefined)
        if (p.interior != 0) return "oriented cubes need a homogeneous interior";
        P.order = 1;
        P.carr11pm.assign(CSX_NIM2 * CSX_NIM2, 1.0);
        coef_cube_zdir(P, z, xr2z, cs);
        done = true;
    } else if (p.ordis == 7 && P.cdim == 4) {
        // Isotropic cubes: only the homogeneous interior has coefficients; the pixel of any
        // other interior uses P(q) = 1 (Source/sc_gpu_pCub_lNon_oIso.h:37-45).
        P.order = 0;
        if (p.interior == 0) coef_cube(P, z, xr2z, p.radius, cs);
        done = true;
    } else if (p.interior == 0) {
        if (P.cdim == 3) {
            coef_sphere(P, z, xr2z, cs);
            done = true;
        } else if (p.ordis == 7 && P.cdim == 1) {
            // isotropic cylinders: the factorisation branch (L/R >= 2); the exact-average branch
            // for short cylinders (:3592-3685) is not on the device path
            if (p.length / p.radius < 2.0) return "isotropic cylinders need L/R >= 2 on the device path";
            P.order = 0;
            coef_cylinder_iso(z, zl, xl2z, xr2z, cs);
            done = true;
        } else if (p.ordis == 6 && P.cdim == 1) {
            P.order = 1;
            coef_cylinder_zdir(z, zl, xl2z, xr2z, cs);
            done = true;
        } else if (p.ordis == 0 && P.cdim == 1) {
            P.carr11pm.assign(CSX_NIM2 * CSX_NIM2, 1.0);
            coef_cylinder_gauss(P, p, z, zl, xl2z, xr2z, polTheta, polPhi, cs);
            done = (s.orcase == 0);
        }
    }
    if (!done) return "combination not supported by the device path";
    for (int a = 0; a < 18; ++a) {
        int& n = cs.nidx[a];
        const double* c = cs.arr[a];
        while (n > 1 && (std::isinf(c[n]) || std::isnan(c[n]) || std::fabs(c[n]) < 1e-99)) n--;
        s.limq[a] = std::pow(std::fabs(c[n]), -1 / (2.0 * n));
    }

    // ---- lattice tables ----
    const double cosa = std::cos(p.ucalpha * kPi / 180.0), cosb = std::cos(p.ucbeta * kPi / 180.0),
                 cosg = std::cos(p.ucgamma * kPi / 180.0);
    P.ucvol = p.uca * p.ucb * p.ucc * std::sqrt(1.0 - sq(cosa) - sq(cosb) - sq(cosg) + 2 * cosa * cosb * cosg);
    P.sphno = 1;
    double lat_vol = P.ucvol, lat_sph = 1;
    P.latpar1.assign(2 * 6, 0);
    P.latpar2.assign(2 * 6, 0);
    P.latpar3.assign(2 * 17, 0.f);
    if (p.ltype != 12) {
        CubicGroup g;
        if (!cubic_group(p.ltype, g) && !(p.ltype >= 1 && p.ltype <= 3) && p.ltype != 6 && p.ltype != 8 &&
            p.ltype != 18)
            return "lattice type not supported by the device path";
        // The reflection lists, multiplicities and structure factors depend only on the
        // lattice type and HKLmax: shared by every set of that lattice (a fit's vertices and a
        // training set's images change only the lattice constants).
        static thread_local std::shared_ptr<const LatticeLists> tl;
        if (!tl || tl->ltype != p.ltype || tl->hklmax != p.hklmax) tl = lattice_lists(p.ltype, p.hklmax);
        const LatticeLists& L = *tl;
        const int n1 = L.n1, n2 = L.n2;
        if (n1 >= CSX_LATLEN || n2 >= CSX_LATLEN) return "reflection table exceeds latparlen";
        s.peakmax1 = n1;
        s.peakmax2 = n2;
        double R[9], RT[9], vol;
        rotated_reciprocal_basis(p.uca, p.ucb, p.ucc, p.ucalpha, p.ucbeta, p.ucgamma, p.ucn1, p.ucn2,
                                 p.ucn3, p.ucpsi, p.ltype >= 1 && p.ltype <= 3, R, RT, vol);
        P.ucvol = vol;
        lat_vol = P.ucvol;
        lat_sph = n1 > 0 ? L.s1[static_cast<size_t>(n1 - 1)] : 1;
        if (n2 > 0) P.sphno = L.s2[static_cast<size_t>(n2 - 1)];
        P.lists = tl;
        std::memcpy(P.rot, R, sizeof R);
        std::memcpy(P.rott, RT, sizeof RT);
        const int shp = p.peakidx + 1;
        if (rows_on_device) {
            // the caller's device builds latpar3 from rot/rott (csx_rows.h); latpar1/2 are
            // P.lists' arrays
            P.latpar1.clear();
            P.latpar2.clear();
            P.latpar3.clear();
        } else {
            P.latpar1 = L.l1;
            P.latpar2 = L.l2;
            P.latpar3.assign(static_cast<size_t>(n2 + 1) * 17, 0.f);
            for (int i = 1; i <= n2; ++i) {
                const int32_t* Lr = &L.l2[static_cast<size_t>(i) * 6];
                float* T = &P.latpar3[static_cast<size_t>(i) * 17];
                rows::Geometry g;
                rows::geometry(R, RT, Lr[1], Lr[2], Lr[3], g);
                T[1] = 2 * kPi / p.uca;
                T[2] = g.qx; T[3] = g.qy; T[4] = g.qz; T[5] = g.qh; T[6] = g.qh / (2.0 * kPi);
                T[7] = g.tx; T[8] = g.ty; T[9] = g.tz;
                T[10] = rows::g3(g.qh);
                T[11] = peak_norm(shp, rows::x2phi(g.qh, s.phiwidth), s.peak_beta);
                T[12] = 0;
                T[13] = g.qxy;
            }
        }
        if (shp < 1 || shp > 6) return "peak shape not supported by the device path";
        // the generic kernel leaves P(q) of a cube on a lattice unset for the other interiors
        if (p.particle == 4 && p.interior != 0) return "cubes on a lattice need a homogeneous interior";
        if (p.twinned) return "twinning not supported by the device path";
    }
    s.sphno_cubevol = lat_vol * lat_sph;
    s.polPhi = polPhi;
    s.polTheta = polTheta;
    s.sinphic = std::sin(-polPhi * kPi / 180.);
    s.cosphic = std::cos(-polPhi * kPi / 180.);
    s.xmin = P.xmin; s.xmax = P.xmax; s.ymin = P.ymin; s.ymax = P.ymax;
    s.iimin = P.ymin;
    s.axmin = P.axmin; s.axmax = P.axmax; s.aymin = P.aymin; s.aymax = P.aymax;
    return "";
}

}  // namespace csx
