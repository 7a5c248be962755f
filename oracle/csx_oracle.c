/*
 * csx_oracle.c -- FP64 CPU oracle for the CrystalScatter per-pixel 2D path.
 *
 * TEST INFRASTRUCTURE ONLY (see csx_oracle.h).  The product never loads it.
 *
 * A restatement, not a copy: every routine follows the reference lines it
 * cites, keeping the reference's arithmetic order and operand types (float,
 * double, x87 long double) wherever they decide the rounded result, so that
 * the oracle reproduces the reference's FP64 output to the last bits on the
 * same inputs.  Quirks of the reference are reproduced on purpose and marked
 * "QUIRK"; the only switch away from reference behaviour is `fix_orcase`
 * (default off), which maps the 0-based orientation-case enum to the 1-based
 * numbering the form factors compare against (see QUIRK-ORCASE below).
 *
 * Compile with -ffp-contract=off (oracle/Makefile): no FMA contraction, as in
 * the reference's x86-64 g++ build.
 */
#define _GNU_SOURCE
#include "csx_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define NCOEF 151  /* coeffarray_len, Source/sc_calc_generic_gpu.h:443 */
#define NIM2 131   /* imax2d_len,     Source/sc_calc_generic_gpu.h:444 */
#define NNMAX 120  /* nnmax,          Source/sc_calc_generic_cpu.cpp:2103 */
#define LATLEN 2500 /* latparlen,     Source/sc_calc_generic_gpu.h:543 */

static const double kEps3 = 0.001, kEps9 = 0.000000001; /* Source/sc_math.h:37-45 */

static inline double sq(double x) { return x * x; }

/* Orientation cases.  QUIRK-ORCASE: the reference enum is 0-based
 * (Source/sc_calc_generic_gpu.h:486-492: orcGeneral=0 .. orcZaxis=3) while
 * formpq/formfq_partCylinder and trapzdchid compare against the Pascal numbers
 * 1..4 (Source/sc_lib_formpq_partCylinder.h:16144,16182; .cu trapzdchid i1==1).
 * With fix_orcase=0 the oracle compares the raw enum value, as the reference does. */
enum { ORC_GENERAL = 0, ORC_X = 1, ORC_Y = 2, ORC_Z = 3 };

typedef struct {
    double c1p[NCOEF], c2p[NCOEF], c3p[NCOEF], c4p[NCOEF], c5p[NCOEF], c6p[NCOEF], c7p[NCOEF],
        c8p[NCOEF], c9p[NCOEF];
    double c1f[NCOEF], c2f[NCOEF], c3f[NCOEF], c4f[NCOEF], c5f[NCOEF], c6f[NCOEF], c7f[NCOEF],
        c8f[NCOEF], c9f[NCOEF];
    double c11pm[NIM2][NIM2];
} coeff_tables;

struct csxo_engine {
    /* ---- key-table inputs (Source/sc_calc_generic.cpp:36-138) ---- */
    int ltype, particle, ordis, interior, peakidx, hklmax, gridpoints;
    int radq1, radq2, radq4, expand, centerbeam, twinned, dsch, para, waxs;
    double uca, ucb, ucc, ucalpha, ucbeta, ucgamma, ceff, ceffcyl, reff, acpl, bcpl, ifluc, rfluc;
    double radius, radiusi, sigma, dbeta, length, sigmal, alphash, rho, rotalpha;
    double peakpar, displacement, azi, domainsize;
    double vax[3], vay[3], vaz[3], sigxyz[3];
    double ucpsi, ucn1, ucn2, ucn3, theta, phi, rottheta, rotphi;
    double qmax, qmin, pixnox, pixnoy, pixx_mm, pixy_mm, det, wave, beamx, beamy;
    double bfactor, p1in, iso, i0, base, width_zuf, dw_zuf;
    int fix_orcase;
    /* ---- prepared state ---- */
    int prepared, shp, cdim, cs, lattice, twin, zmax, use1d;
    int zzmin, zzmax, iimin, iimax, xmin, xmax, ymin, ymax;
    double p1, phiwidth, dwfactor, dist, reldis, polPhi, polTheta, cosphi, sinphi, costheta,
        sintheta, cosphic, sinphic, norm, order, por, pixx_m, pixy_m;
    int orcase;
    double limq[18];
    coeff_tables C;
    int peakmax1, peakmax2;
    int latpar1[LATLEN][6], latpar2[LATLEN][6];
    float latpar3[LATLEN][17];
    double ucvol, sphno, lat_ucvol, lat_sphno;
    double ri[9], rt[9];
    double eta, beta, c0, c1, c2, c3, c4; /* peak-shape parameters (setPeakShapeParams, cpu:585-590) */
};

/* ========================================================================== */
/* Parameters                                                                  */
/* ========================================================================== */

csxo_engine* csxo_new(void)
{
    csxo_engine* e = (csxo_engine*)calloc(1, sizeof(csxo_engine));
    if (!e) return NULL;
    /* Defaults of the key table, Source/sc_calc_generic.cpp:36-138. */
    e->ltype = 12; e->uca = e->ucb = e->ucc = 30; e->ucalpha = e->ucbeta = e->ucgamma = 90;
    e->ceff = 1; e->particle = 0; e->ordis = 7; e->interior = 0; e->radius = 1.0;
    e->radiusi = 0.5; e->sigma = 0.07; e->dbeta = 0.4; e->length = 1; e->sigmal = 0.06;
    e->peakidx = 1; e->displacement = 2; e->azi = 180; e->domainsize = 400;
    e->vax[0] = 1; e->vay[1] = 1; e->vaz[2] = 1; e->sigxyz[0] = e->sigxyz[1] = e->sigxyz[2] = 40;
    e->ucn1 = e->ucn2 = e->ucn3 = 1; e->theta = 90; e->phi = 0; e->rottheta = 90; e->rotphi = 90;
    e->gridpoints = 64; e->hklmax = 5; e->qmax = 1; e->radq4 = 1;
    e->pixnox = e->pixnoy = 2048; e->pixx_mm = e->pixy_mm = 1; e->det = 1; e->wave = 0.09499;
    e->bfactor = 1; e->i0 = 10000; e->width_zuf = 1; e->dw_zuf = 1;
    return e;
}

void csxo_free(csxo_engine* e) { free(e); }

typedef struct { const char* key; size_t off; int isint; } keydef;
#define KD(k, f) {k, offsetof(csxo_engine, f), 0}
#define KI(k, f) {k, offsetof(csxo_engine, f), 1}
#include <stddef.h>
static const keydef kKeys[] = {
    KI("LType", ltype), KD("uca", uca), KD("ucb", ucb), KD("ucc", ucc), KD("ucalpha", ucalpha),
    KD("ucbeta", ucbeta), KD("ucgamma", ucgamma), KD("EditCeff", ceff),
    KI("CheckBoxTwinned", twinned), KD("EditCeffcyl", ceffcyl), KD("reff", reff),
    KD("acpl", acpl), KD("bcpl", bcpl), KD("ifluc", ifluc), KD("rfluc", rfluc),
    KI("ComboBoxParticle", particle), KI("Ordis", ordis), KI("ComboBoxInterior", interior),
    KD("EditRadius", radius), KD("EditRadiusi", radiusi), KD("EditSigma", sigma),
    KD("EditDbeta", dbeta), KD("Length", length), KD("SigmaL", sigmal), KD("Alpha", alphash),
    KD("EditRho", rho), KD("RotAlpha", rotalpha), KI("ComboBoxPeak", peakidx),
    KD("EditPeakPar", peakpar), KD("EditDebyeWaller", displacement), KD("EditAzi", azi),
    KD("EditDomainSize", domainsize), KI("RadButDebyeScherrer", dsch),
    KI("RadioButtonPara", para), KD("VAx1", vax[0]), KD("VAx2", vax[1]), KD("VAx3", vax[2]),
    KD("Ay1", vay[0]), KD("Ay2", vay[1]), KD("Ay3", vay[2]), KD("Az1", vaz[0]),
    KD("Az2", vaz[1]), KD("Az3", vaz[2]), KD("SigX", sigxyz[0]), KD("SigY", sigxyz[1]),
    KD("SigZ", sigxyz[2]), KD("ucpsi", ucpsi), KD("ucn1", ucn1), KD("ucn2", ucn2),
    KD("ucn3", ucn3), KD("theta", theta), KD("phi", phi), KD("rotTheta", rottheta),
    KD("rotPhi", rotphi), KI("GridPoints", gridpoints), KI("HKLmax", hklmax),
    KD("EditQmax", qmax), KD("EditQmin", qmin), KI("RadioButtonQ1", radq1),
    KI("RadioButtonQ2", radq2), KI("RadioButtonQ4", radq4), KI("ExpandImage", expand),
    KD("EditPixelNoX", pixnox), KD("EditPixelNoY", pixnoy), KD("EditPixelX", pixx_mm),
    KD("EditPixelY", pixy_mm), KD("EditDet", det), KD("EditWavelength", wave),
    KD("BeamPosX", beamx), KD("BeamPosY", beamy), KI("CenterBeam", centerbeam),
    KD("EditBFactor", bfactor), KI("CheckBoxWAXS", waxs), KD("P1", p1in), KD("iso", iso),
    KD("I0", i0), KD("Base", base), KD("width_zuf", width_zuf), KD("dw_zuf", dw_zuf), KI("fix_orcase", fix_orcase),
};

int csxo_set(csxo_engine* e, const char* key, double v)
{
    for (size_t i = 0; i < sizeof(kKeys) / sizeof(kKeys[0]); ++i) {
        if (strcmp(kKeys[i].key, key) == 0) {
            char* p = (char*)e + kKeys[i].off;
            if (kKeys[i].isint) *(int*)p = (int)v; else *(double*)p = v;
            e->prepared = 0;
            return 0;
        }
    }
    return -1;
}

double csxo_get(const csxo_engine* e, const char* key)
{
    for (size_t i = 0; i < sizeof(kKeys) / sizeof(kKeys[0]); ++i) {
        if (strcmp(kKeys[i].key, key) == 0) {
            const char* p = (const char*)e + kKeys[i].off;
            return kKeys[i].isint ? (double)*(const int*)p : *(const double*)p;
        }
    }
    return NAN;
}

/* ========================================================================== */
/* Special functions                                                           */
/* ========================================================================== */

/* Reference Gamma, Source/sc_calc_generic_gpu.cu:613-673: a 26-term series of
 * 1/Gamma on (0,1], the recursion times an 8-term polynomial on (1,1000],
 * Stirling above, reflection for negative arguments.  Coefficient tables are
 * the reference's (they decide the values). */
static const double kGamBm[9] = {0.0, -0.577191652, 0.988205891, -0.897056937, 0.918206857,
                                 -0.756704078, 0.482199394, -0.193527818, 0.035868343};
static const double kGamCm[27] = {
    0.0, 1.0000000000000000, 0.5772156649015329, -0.6558780715202538, -0.0420026350340952,
    0.1665386113822915, -0.0421977345555443, -0.0096219715278770, 0.00721894322466630,
    -0.0011651675918591, -0.0002152416741149, 0.0001280502823882, -0.0000201348547807,
    -0.0000012504934821, 0.0000011330272320, -0.0000002056338417, 0.0000000061160950,
    0.0000000050020075, -0.0000000011812746, 0.0000000001043427, 0.0000000000077823,
    -0.0000000000036968, 0.0000000000005100, -0.0000000000000206, -0.0000000000000054,
    0.0000000000000014, 0.0000000000000001};

static double frac_part(double x) { double ip; return modf(x, &ip); }

double csxo_gamma(double z)
{
    int neg;
    if (fabs(z) < kEps9) z = z + kEps9;
    if (z > 0) neg = 0;
    else { neg = 1; if (fabs(frac_part(z)) < kEps9) z = z + kEps9; }
    const double az = fabs(z);
    if (az > 0.0 && az <= 1) {
        double s = 0;
        for (int i = 1; i <= 26; ++i) s += kGamCm[i] * pow(az, i);
        if (neg) return -M_PI * s / (az * sin(M_PI * az));
        return 1. / s;
    }
    if (az > 1 && az <= 1000) {
        if (fabs(frac_part(az)) < kEps9) {
            double f = 1;
            for (int i = 1; i <= trunc(az) - 1; ++i) f = f * i;
            if (neg) return -M_PI / (az * f * sin(M_PI * az));
            return f;
        }
        const double y = az - 1;
        const double di = trunc(y);
        const double x = az - trunc(az);
        double fak = y;
        if (di == 0) fak = 1;
        else for (int i = 1; i < di; ++i) fak = fak * (y - i);
        double f = 1;
        for (int i = 1; i <= 8; ++i) f = f + kGamBm[i] * pow(x, i);
        if (neg) return -M_PI / (az * fak * f * sin(M_PI * az));
        return fak * f;
    }
    const double s = exp(-az) * exp((az - 0.5) * log(az)) * sqrt(2 * M_PI);
    if (neg) return -M_PI / (az * s * sin(M_PI * az));
    return s;
}

/* Integral of exp(-a x^2) sin x over [0,pi] as the reference approximates it,
 * Source/sc_calc_generic_gpu.cu:838-854 (asymptotic form for a>0.01, Taylor below). */
double csxo_gaussnorm3(double a)
{
    if (a > 0.01) {
        const double ap = a * M_PI * M_PI, ex = exp(-a * M_PI * M_PI);
        (void)ap;
        const double b0 = (1 - exp(-a * M_PI * M_PI)) / 2.;
        const double b1 = (1 - ex * (1 + a * M_PI * M_PI)) / 12.;
        const double b2 = (2 - ex * (2 + 2 * a * M_PI * M_PI + a * a * M_PI * M_PI * M_PI * M_PI)) / 240.;
        const double b3 = (6 - ex * (6 + 6 * a * M_PI * M_PI + 3 * a * a * M_PI * M_PI * M_PI * M_PI +
                                     a * a * a * M_PI * M_PI * M_PI * M_PI * M_PI * M_PI)) / 10080.;
        return b0 / a - b1 / (a * a) + b2 / (a * a * a) - b3 / (a * a * a * a);
    }
    const double a0 = 2;
    const double a1 = (M_PI * M_PI - 4);
    const double a2 = (18 - 12 * M_PI * M_PI + M_PI * M_PI * M_PI * M_PI) / 2.;
    const double a3 = (-1440 + 360 * M_PI * M_PI - 30 * M_PI * M_PI * M_PI * M_PI +
                       M_PI * M_PI * M_PI * M_PI * M_PI * M_PI) / 6.;
    return a0 - a1 * a + a2 * a * a - a3 * a * a * a;
}

/* Integral of sin x / (1 + a x^2) over [0,pi] as the reference approximates it,
 * Source/sc_calc_generic_gpu.cu:809-826 (asymptotic form for a>0.01, Taylor below). */
double csxo_lorentznorm3(double a)
{
    if (a > 0.01) {
        const double b0 = log(1 + a * M_PI * M_PI) / 2.;
        const double b1 = (a * M_PI * M_PI - log(1 + a * M_PI * M_PI)) / 12.;
        const double b2 = (-2 * a * M_PI * M_PI + a * a * M_PI * M_PI * M_PI * M_PI + 2 * log(1 + a * M_PI * M_PI)) / 480.;
        const double b3 = (a * M_PI * M_PI * (6 - 3 * a * M_PI * M_PI + 2 * a * a * M_PI * M_PI * M_PI * M_PI) -
                           6 * log(1 + a * M_PI * M_PI)) / 60480.;
        return b0 / a - b1 / (a * a) + b2 / (a * a * a) - b3 / (a * a * a * a);
    }
    const double a0 = 2;
    const double a1 = (M_PI * M_PI - 4);
    const double a2 = (18 - 12 * M_PI * M_PI + M_PI * M_PI * M_PI * M_PI) / 2.;
    const double a3 = (-1440 + 360 * M_PI * M_PI - 30 * M_PI * M_PI * M_PI * M_PI +
                       M_PI * M_PI * M_PI * M_PI * M_PI * M_PI) / 6.;
    return a0 - a1 * a + a2 * a * a - a3 * a * a * a;
}

/* One trapezoid refinement of the integral of sin x / (1 + a x^2)^b over [at, bt]
 * (pearsonintegral3, Source/sc_calc_generic_gpu.cu:918-946): level 1 the end points, then
 * the midpoints of the *it current intervals, which double. */
static void pearson_trapz(double at, double bt, double a, double b, double* sx, int nx, int* it)
{
    if (nx == 1) {
        const double fa = sin(at) / exp(b * log(1 + a * at * at));
        const double fb = sin(bt) / exp(b * log(1 + a * bt * bt));
        *sx = 0.5 * (bt - at) * (fa + fb);
        *it = 1;
        return;
    }
    const double tnm = *it;
    const double del = (bt - at) / tnm;
    double xt = at + 0.5 * del, sum = 0.0;
    for (int j = 1; j <= *it; ++j) {
        sum += sin(xt) / exp(b * log(1 + a * xt * xt));
        xt = xt + del;
    }
    *sx = 0.5 * (*sx + (bt - at) * sum / tnm);
    *it = 2 * *it;
}

/* Integral of sin x / (1 + a x^2)^b over [0,pi] (pearsonnorm3, Source/sc_calc_generic_gpu.cu:
 * 868-899): b nudged by 1e-6, closed forms for a > 20 and a <= 0.01, else trapezoid levels
 * until two agree to 1e-3 (at most 10). */
double csxo_pearsonnorm3(double a, double b)
{
    b += 0.000001;
    if (a > 20.) {
        const double b0 = (1 - exp((1 - b) * log(1 + a * M_PI * M_PI))) / (2 * (b - 1));
        return b0 / a;
    }
    if (a <= 0.01) {
        const double a0 = 2.;
        const double a1 = (M_PI * M_PI - 4) * b;
        const double a2 = (18 - 12 * M_PI * M_PI + M_PI * M_PI * M_PI * M_PI) * b * (b + 1) / 2.;
        const double a3 = (-1440 + 360 * M_PI * M_PI - 30 * M_PI * M_PI * M_PI * M_PI +
                           M_PI * M_PI * M_PI * M_PI * M_PI * M_PI) * b * (b + 1) * (b + 2) / 6.;
        return a0 - a1 * a + a2 * a * a - a3 * a * a * a;
    }
    int it = 0;
    double s = 0;
    pearson_trapz(0, M_PI, a, b, &s, 1, &it);
    double old = s;
    for (int itf = 2; itf <= 10; ++itf) {
        pearson_trapz(0, M_PI, a, b, &s, itf, &it);
        if (fabs(1 - s / old) < kEps3) break;
        old = s;
    }
    return s;
}

/* Neville extrapolation to h=0 over k points, 1-based arrays, as polint
 * (Source/sc_calc_generic_gpu.cu:1280-1335; x=0 so |x-xa| = |xa|). */
static void neville0(const double* xa, const double* ya, int n, double* y, double* dy)
{
    double c[8], d[8];
    int ns = 1;
    double dif = fabs(xa[1]);
    for (int i = 1; i <= n; ++i) {
        const double dift = fabs(xa[i]);
        if (dift < dif) { ns = i; dif = dift; }
        c[i] = ya[i];
        d[i] = ya[i];
    }
    *y = ya[ns];
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
        if (2 * ns < n - m) *dy = c[ns + 1];
        else { *dy = d[ns]; ns = ns - 1; }
        *y = *y + *dy;
    }
}

/* ========================================================================== */
/* Orientation averages: nested Romberg (chi inner, delta outer)               */
/* ========================================================================== */

typedef struct {
    double l, sigma, dbeta;         /* length, Schulz sigma of the length, dbeta [rad] */
    double p11, p12, p13, p21, p22, p23; /* float-rounded rotation entries */
    double qx, qy;
    int i1;                         /* orientation case as passed (see QUIRK-ORCASE) */
} chi_ctx;

/* Longitudinal Schulz-averaged P(q) of a rod at argument arg,
 * trapzdchid i3==0 (Source/sc_calc_generic_gpu.cu:1640-1652). */
static inline double rod_pq(double z, double arg)
{
    const double a1 = 1 / (2.0 * z * (z - 1));
    return (a1 / (arg * arg)) * (1 - cos((z - 1) * atan(2.0 * arg)) / pow(1.0 + 4 * arg * arg, (z - 1) / 2.0));
}

/* Integrand of trapzdchid for the cylinder cases i1=1..4 at azimuth x and polar delta
 * (.cu:1634-1720 for n==1, .cu:2451-2520 for n>1).  Any other i1 contributes 0:
 * pa/pb/px start at 0 in the reference and no branch assigns them. */
static double chi_integrand(const chi_ctx* c, double x, double delta)
{
    const double eps = 0.0000000001;
    const double z = (1 - c->sigma * c->sigma) / (c->sigma * c->sigma);
    double arg;
    switch (c->i1) {
    case 1: {
        const double ml1 = c->p11 * cos(x) * sin(delta) + c->p12 * sin(x) * sin(delta) + c->p13 * cos(delta);
        const double ml2 = c->p21 * sin(x) * sin(delta) + c->p22 * cos(x) * sin(delta) + c->p23 * cos(delta);
        arg = (c->qx * ml1 + c->qy * ml2 + eps) * c->l / (z + 1);
        break;
    }
    case 2: arg = (c->qx * cos(delta) - c->qy * sin(x) * sin(delta) + eps) * c->l / (z + 1); break;
    case 3: arg = (c->qx * sin(x) * sin(delta) + c->qy * cos(delta) + eps) * c->l / (z + 1); break;
    case 4: arg = (c->qx * sin(x) * sin(delta) - c->qy * cos(x) * sin(delta) + eps) * c->l / (z + 1); break;
    default: return 0.0;
    }
    return rod_pq(z, arg);
}

/* qrombchid (Source/sc_calc_generic_gpu.cu:1495-1583) with trapzdchid's refinement
 * (.cu:1590-2961): integral over chi in [0, float(2*pi)], eps 1e-4, jmax 20, k 5. */
static double romb_chi(const chi_ctx* c, double delta)
{
    const double eps = 1.0e-4;
    const int jmax = 20, k = 5;
    double hp[24], sp[24], cp[8], dp[8];
    const float alim = 0.0f, blim = (float)(2 * M_PI);
    const double a = alim, b = blim;
    double pq = 0, dssp = 0;
    int cnt = 1;
    hp[1] = 1.0;
    for (int j = 1; j <= jmax; ++j) {
        if (j == 1) {
            const double pa = chi_integrand(c, a, delta), pb = chi_integrand(c, b, delta);
            sp[j] = 0.5 * (b - a) * (pa + pb);
            cnt = 1;
        } else {
            const double tnm = cnt, del = (b - a) / tnm;
            double x = a + 0.5 * del, sum = 0.0;
            for (int i = 1; i <= cnt; ++i) { sum += chi_integrand(c, x, delta); x = x + del; }
            sp[j] = 0.5 * (sp[j] + (b - a) * sum / tnm);
            cnt = 2 * cnt;
        }
        if (j >= k) {
            for (int i = 1; i <= k; ++i) { cp[i] = hp[j - k + i]; dp[i] = sp[j - k + i]; }
            neville0(cp, dp, k, &pq, &dssp);
            if (fabs(dssp) < (eps * fabs(pq))) break;
            if (fabs(pq) < 1.0e-300) break;
        }
        sp[j + 1] = sp[j];
        hp[j + 1] = 0.25 * hp[j];
    }
    return pq;
}

/* trapzddeltac integrands for i0 = 1 (delta weight; QUIRK: the cos^i3 sin^i4
 * factor of the Pascal original is dropped, .cu:2190), 2 (norm, .cu:2197),
 * 3 (order parameter, .cu:2204), 4 (cylinder form factor, .cu:2211-2212). */
static double delta_integrand(int i0, double x, double dbeta, const chi_ctx* c)
{
    switch (i0) {
    case 1:
    case 2: return sin(x) * exp(-x * x / (dbeta * dbeta));
    case 3: return sin(x) * exp(-x * x / (dbeta * dbeta)) * (3 * cos(x) * cos(x) - 1) / 2.0;
    case 4: {
        const double px = romb_chi(c, x);
        return sin(x) * exp(-x * x / (dbeta * dbeta)) * px / (2.0 * M_PI);
    }
    default: return 0.0;
    }
}

/* qrombdeltac (Source/sc_calc_generic_gpu.cu:1370-1488) with trapzddeltac
 * (.cu:956-1272), for i2 = 0 (Gaussian): delta in [0, float(min(dbeta*sqrt(ln 1e5), pi/2))].
 * sigma is the Schulz sigma forwarded to the chi integrand (sigmal for cylinders). */
static double romb_delta(const csxo_engine* e, double sigma, double theta_deg, double phi_deg,
                         double qx, double qy, int i0, int i1)
{
    const double eps = 1.0e-4, prmin = 0.00001;
    const int jmax = 20, k = 5;
    double hp[24], sp[24], cp[8], dp[8];
    const double dbeta = e->dbeta * M_PI / 180.0;
    const double th = theta_deg * M_PI / 180.0, ph = phi_deg * M_PI / 180.0;
    double delmax = dbeta * sqrt(log(1.0 / prmin));
    if (delmax > M_PI / 2.0) delmax = M_PI / 2.0;
    const float alim = 0.0f, blim = (float)delmax;
    const float p11 = -cos(ph) * cos(th), p12 = sin(ph), p13 = cos(ph) * sin(th);
    const float p21 = -cos(ph), p22 = -sin(ph) * cos(th), p23 = sin(ph) * sin(th);
    chi_ctx c = {e->length, sigma, dbeta, p11, p12, p13, p21, p22, p23, qx, qy, i1};
    const double a = alim, b = blim;
    double pq = 0, dssp = 0;
    int cnt = 1;
    hp[1] = 1.0;
    for (int j = 1; j <= jmax; ++j) {
        if (j == 1) {
            const double pa = delta_integrand(i0, a, dbeta, &c), pb = delta_integrand(i0, b, dbeta, &c);
            sp[j] = 0.5 * (b - a) * (pa + pb);
            cnt = 1;
        } else {
            const double tnm = cnt, del = (b - a) / tnm;
            double x = a + 0.5 * del, sum = 0.0;
            for (int i = 1; i <= cnt; ++i) { sum = sum + delta_integrand(i0, x, dbeta, &c); x = x + del; }
            sp[j] = 0.5 * (sp[j] + (b - a) * sum / tnm);
            cnt = 2 * cnt;
        }
        if (j >= k) {
            for (int i = 1; i <= k; ++i) { cp[i] = hp[j - k + i]; dp[i] = sp[j - k + i]; }
            neville0(cp, dp, k, &pq, &dssp);
            if (fabs(dssp) < (eps * fabs(pq))) break;
            if (fabs(pq) < 1.0e-100) break;
        }
        sp[j + 1] = sp[j];
        hp[j + 1] = 0.25 * hp[j];
    }
    return pq;
}

/* ========================================================================== */
/* Series coefficients: coefficients() branches on the hot path                */
/* ========================================================================== */

/* Sphere, homogeneous (Source/sc_calc_generic_cpu.cpp:2509-2543). */
static void coef_sphere(csxo_engine* e, double z, double xr2z, int* n4, int* n4f)
{
    long double fkv[NNMAX + 2], gam3[NNMAX + 2], z12v[NNMAX + 1];
    fkv[0] = 1; gam3[0] = sqrt(M_PI) / 2.0; z12v[0] = 1.0;
    long double xrn = 1.0;
    const double tiny = 1E-30;
    e->norm = 1;
    e->order = 0;
    for (int n = 1; n <= NNMAX; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xrn = -xrn * xr2z;
        e->C.c4p[n] = 9 * sqrt(M_PI) * pow(4.0, n) * z12v[n] * xrn /
                      (2.0 * (n + 3) * (n + 2) * (n + 3 / 2.0) * gam3[n] * fkv[n]);
        double bs = 0.0;
        for (int m = 0; m <= n; ++m)
            bs += z12v[m] * z12v[n - m] /
                  ((m + 3 / 2.0) * gam3[m] * (n - m + 3 / 2.0) * gam3[n - m] * fkv[m] * fkv[n - m]);
        e->C.c4f[n] = 9 * M_PI * xrn * bs / 16.0;
        if (fabs(e->C.c4p[n]) < tiny && n < *n4) *n4 = n;
        if (fabs(e->C.c4f[n]) < tiny && n < *n4f) *n4f = n;
    }
}

/* Cube, homogeneous, isotropic (Source/sc_calc_generic_cpu.cpp:2757-2829): Porod constant,
 * P(q) coefficients from the double self-convolution of u_n = 1/((n+1/2)(n+1) n!), and the
 * <F>^2 coefficients from the triple convolution of the sphere-like fsum.  fsum, u and v are
 * doubles and fkv, gam3, z12v long doubles, as the reference declares them (:2127-2130). */
static void coef_cube(csxo_engine* e, double z, double xr2z, int* n4, int* n4f)
{
    long double fkv[NNMAX + 2], gam3[NNMAX + 2], z12v[NNMAX + 1];
    double fsum[NNMAX + 1], u1[NNMAX + 1], v1[NNMAX + 1];
    const double tiny = 1E-30;
    e->norm = 1;
    e->order = 0;
    const double area = 6.0 * 4.0 * sq(e->radius);
    const double vol = 8.0 * e->radius * sq(e->radius);
    e->por = 2 * M_PI * pow(z + 1, 4) * area / (z * (z - 1) * (z - 2) * (z - 3) * vol * vol);
    fkv[0] = 1; gam3[0] = sqrt(M_PI) / 2.0; z12v[0] = 1.0; fsum[0] = 1; u1[0] = 2;
    for (int n = 1; n <= NNMAX; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        u1[n] = 1 / ((n + 1 / 2.0) * (n + 1) * fkv[n]);
    }
    for (int n = 0; n <= NNMAX; ++n) {
        double s = 0.0;
        for (int m = 0; m <= n; ++m) s += u1[n - m] * u1[m];
        v1[n] = s;
    }
    long double xrn = 1.0;
    for (int n = 1; n <= NNMAX; ++n) {
        xrn = -xrn * xr2z;
        long double sf = 0.0;
        for (int m = 0; m <= n; ++m) sf += z12v[n - m] * z12v[m] / (gam3[n - m] * gam3[m] * fkv[n - m] * fkv[m]);
        fsum[n] = sf;
        double sp = 0.0;
        for (int m = 0; m <= n; ++m) sp += u1[n - m] * v1[m];
        e->C.c4p[n] = sqrt(M_PI) * pow(4.0, n) * z12v[n] * xrn * sp / (16.0 * gam3[n]);
        sp = 0.0;
        for (int m = 0; m <= n; ++m) {
            double s1 = 0.0;
            for (int k = 0; k <= m; ++k) s1 += fsum[m - k] * fsum[k] * gam3[m - k] * gam3[k] / ((k + 1 / 2.0) * (m - k + 1 / 2.0));
            sp += s1 * fsum[n - m] * gam3[n - m] / (n - m + 1 / 2.0);
        }
        e->C.c4f[n] = M_PI * M_PI * xrn * sp / (128.0 * gam3[n]);
        if (fabs(e->C.c4p[n]) < tiny && n < *n4) *n4 = n;
        if (fabs(e->C.c4f[n]) < tiny && n < *n4f) *n4f = n;
    }
}

/* Cube, perfectly oriented (z-dir), homogeneous (Source/sc_calc_generic_cpu.cpp:2880-2910):
 * carr11pm[n][m] for m <= n (the rest keeps its initial 1), carr4p and carr4f. */
static void coef_cube_zdir(csxo_engine* e, double z, double xr2z, int* n4, int* n4f)
{
    long double fkv[NNMAX + 2], gam3[NNMAX + 2], z12v[NNMAX + 1];
    fkv[0] = 1; gam3[0] = sqrt(M_PI) / 2.0; z12v[0] = 1.0;
    long double xrn = 1.0;
    const double tiny = 1E-30;
    e->norm = 1;
    e->order = 1;
    for (int n = 1; n <= NNMAX; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xrn = -xrn * xr2z;
        for (int m = 0; m <= n; ++m)
            e->C.c11pm[n][m] = (M_PI / 4.0) * pow(4.0, n) * z12v[n - m] * z12v[m] * xrn /
                               (gam3[n - m] * gam3[m] * (n - m + 1) * fkv[n - m] * (m + 1) * fkv[m]);
        e->C.c4p[n] = sqrt(M_PI) * pow(4.0, n) * xrn / (16.0 * gam3[n]);
        e->C.c4f[n] = M_PI * M_PI * xrn / (128.0 * gam3[n]);
        if (fabs(e->C.c4p[n]) < tiny && n < *n4) *n4 = n;
        if (fabs(e->C.c4f[n]) < tiny && n < *n4f) *n4f = n;
    }
}

/* Cylinder, isotropic, homogeneous, axial ratio L/R >= 2 (Source/sc_calc_generic_cpu.cpp:3578-3583,
 * 3686-3746, the factorisation branch): P|| and P- series, <F>^2 series, with the 1e-30 switch
 * points.  z12v, z12vl, fkv, fk2v, gam3 and xln are long double, xrn double, as declared there. */
static void coef_cyl_iso(csxo_engine* e, double z, double zl, double xl2z, double xr2z, int* n1, int* n4,
                         int* n1f, int* n4f)
{
    long double fkv[NNMAX + 2], fk2v[NNMAX + 1], gam3[NNMAX + 2], z12v[NNMAX + 1], z12vl[NNMAX + 1];
    fkv[0] = 1; fk2v[0] = 1; gam3[0] = sqrt(M_PI) / 2.0; z12v[0] = 1.0; z12vl[0] = 1;
    long double xln = 1.0;
    double xrn = 1.0;
    const double tiny = 1E-30;
    e->norm = 1;
    e->order = 0;
    for (int n = 1; n <= NNMAX; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        z12vl[n] = z12vl[n - 1] * ((zl + 1) - 2 + 2 * n) * ((zl + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        fk2v[n] = fk2v[n - 1] * (2 * n - 1) * (2 * n);
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xln = -xln * xl2z;
        xrn = -xrn * xr2z;
        e->C.c1p[n] = sqrt(M_PI) * pow(4.0, n) * z12vl[n] * xln / (2.0 * (2 * n + 1) * (n + 1) * gam3[n] * fkv[n]);
        e->C.c4p[n] = 4 * (n + 1 / 2.0) * fk2v[n] * z12v[n] * xrn /
                      ((n + 2) * (n + 1) * fkv[n] * fkv[n] * (n + 1) * fkv[n] * fkv[n]);
        double bs = 0.0;
        for (int m = 0; m <= n; ++m) bs += z12vl[m] * z12vl[n - m] / (gam3[m] * gam3[n - m] * fkv[m] * fkv[n - m]);
        e->C.c1f[n] = M_PI * xln * bs / (4.0 * (2 * n + 1));
        bs = 0.0;
        for (int m = 0; m <= n; ++m)
            bs += z12v[m] * z12v[n - m] / ((m + 1) * fkv[m] * (n - m + 1) * fkv[n - m] * fkv[m] * fkv[n - m]);
        e->C.c4f[n] = xrn * bs;
        if (fabs(e->C.c1p[n]) < tiny && n < *n1) *n1 = n;
        if (fabs(e->C.c1f[n]) < tiny && n < *n1f) *n1f = n;
        if (fabs(e->C.c4p[n]) < tiny && n < *n4) *n4 = n;
        if (fabs(e->C.c4f[n]) < tiny && n < *n4f) *n4f = n;
    }
}

/* Cylinder, perfect (z-dir) orientation, homogeneous
 * (Source/sc_calc_generic_cpu.cpp:4124-4212 incl. the signed first-sign-change searches). */
static void coef_cyl_zdir(csxo_engine* e, double z, double zl, double xl2z, double xr2z, int* n1,
                          int* n4, int* n1f, int* n4f)
{
    long double fkv[NNMAX + 2], fk2v[NNMAX + 1], gam3[NNMAX + 2], z12vl[NNMAX + 1];
    double z12v[NNMAX + 1];
    fkv[0] = 1; fk2v[0] = 1; gam3[0] = sqrt(M_PI) / 2.0; z12vl[0] = 1; z12v[0] = 1.0;
    long double xln = 1.0;
    double xrn = 1.0;
    int s1 = 1, s4 = 1;
    const double tiny = 1E-30;
    e->norm = 1;
    e->order = 1;
    for (int n = 1; n <= NNMAX; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        z12vl[n] = z12vl[n - 1] * ((zl + 1) - 2 + 2 * n) * ((zl + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        fk2v[n] = fk2v[n - 1] * (2 * n - 1) * (2 * n);
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xln = -xln * xl2z;
        xrn = -xrn * xr2z;
        e->C.c1p[n] = sqrt(M_PI) * pow(4.0, n) * z12vl[n] * xln / (2.0 * (n + 1) * gam3[n] * fkv[n]);
        e->C.c4p[n] = 4 * (n + 1 / 2.0) * fk2v[n] * z12v[n] * xrn /
                      ((n + 2) * (n + 1) * fkv[n] * fkv[n] * (n + 1) * fkv[n]);
        double s = 0.0;
        for (int m = 0; m <= n; ++m)
            s += z12vl[m] * z12vl[n - m] / (gam3[m] * gam3[n - m] * fkv[m] * fkv[n - m] * fkv[m] * fkv[n - m]);
        e->C.c1f[n] = M_PI * xln * s / 4.0;
        s = 0.0;
        for (int m = 0; m <= n; ++m)
            s += z12v[m] * z12v[n - m] / ((m + 1) * fkv[m] * (n - m + 1) * fkv[n - m] * fkv[m] * fkv[n - m]);
        e->C.c4f[n] = xrn * s;
        /* QUIRK: signed test (coefficient < 1e-50) -> the first negative term, n=1. */
        if (s1 && e->C.c1p[n] < 1e-50) { *n1 = n; s1 = 0; }
        if (s4 && e->C.c4p[n] < 1e-50) { *n4 = n; s4 = 0; }
        if (fabs(e->C.c1f[n]) < tiny && n < *n1f) *n1f = n;
        if (fabs(e->C.c4f[n]) < tiny && n < *n4f) *n4f = n;
    }
}

/* Cylinder, Gaussian orientation, homogeneous (Source/sc_calc_generic_cpu.cpp:4648-4720):
 * norm and order parameter by nested Romberg, phi forced to 0 for the general case,
 * intlar[n][m] (all the same integral, QUIRK .cu:2190), carr11pm, carr1p/2p/1f/2f/4p/4f. */
static void coef_cyl_gauss(csxo_engine* e, double z, double zl, double xl2z, double xr2z, int* n1,
                           int* n4, int* n1f, int* n4f)
{
    const double tiny = 1E-30;
    e->norm = romb_delta(e, e->sigma, e->polTheta, e->polPhi, 1, 1, 2, 0);
    double ord = romb_delta(e, e->sigma, e->polTheta, e->polPhi, 1, 1, 3, 0);
    e->order = ord / e->norm;
    if (e->orcase != ORC_GENERAL) return; /* only the general case fills the tables */

    e->polPhi = 0; /* QUIRK: phi zeroed here, after orcase was fixed (:4657) */
    const double p11 = -cos(e->polPhi * M_PI / 180.0) * cos(e->polTheta * M_PI / 180.0);
    const double p13 = cos(e->polPhi * M_PI / 180.0) * sin(e->polTheta * M_PI / 180.0);
    const double p21 = -cos(e->polPhi * M_PI / 180.0);
    /* intlar[n][m]: qrombdeltac(i0=1, i3=2n, i4=2m) -- one value for every (n,m). */
    const double intl = romb_delta(e, e->sigma, e->polTheta, e->polPhi, 1, 1, 1, 0);

    long double fkv[NNMAX + 2], fk2v[NNMAX + 1], gam3[NNMAX + 2], z12vl[NNMAX + 1];
    double z12v[NNMAX + 1];
    fkv[0] = 1; fk2v[0] = 1; gam3[0] = sqrt(M_PI) / 2.0; z12vl[0] = 1; z12v[0] = 1.0;
    /* fkv/fk2v are needed up to index nmax before the n loop fills them in order;
     * the carr11pm sum reads fk2v[m-ll], fkv[n-m+ll], fkv[ll] with indices <= n. */
    double xrn = 1.0;
    long double xln = 1.0;
    int s1 = 1, s4 = 1;
    for (int n = 1; n <= NNMAX; ++n) {
        z12v[n] = z12v[n - 1] * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
        z12vl[n] = z12vl[n - 1] * ((zl + 1) - 2 + 2 * n) * ((zl + 1) - 1 + 2 * n);
        fkv[n] = fkv[n - 1] * n;
        fk2v[n] = fk2v[n - 1] * (2 * n - 1) * (2 * n);
        gam3[n] = gam3[n - 1] * (2 * n + 1) / 2.0;
        xln = -xln * xl2z;
        xrn = -xrn * xr2z;
        for (int m = 0; m <= n; ++m) {
            double s = 0.0;
            for (int ll = 0; ll <= m; ++ll)
                s += pow(p11 * p11, ll) * pow(p13 * p13, m - ll) * intl /
                     (pow(4.0, ll) * fk2v[m - ll] * fkv[n - m + ll] * fkv[ll] * e->norm);
            e->C.c11pm[n][m] = pow(4.0, m) * pow(sq(p21), n - m) * s / (fkv[n - m]);
        }
        e->C.c1p[n] = pow(4.0, n) * z12vl[n] * xln / ((2.0 * n + 1) * (n + 1));
        e->C.c2p[n] = pow(4.0, n + 1) * gam3[n] * z12vl[n] * xln /
                      (sqrt(M_PI) * (n + 2) * (n + 1) * (n + 1) * fkv[n] * fkv[n] * fkv[n]);
        double s = 0.0;
        for (int m = 0; m <= n; ++m)
            s = s + z12vl[m] * z12vl[n - m] / (gam3[m] * gam3[n - m] * fkv[m] * fkv[n - m]);
        e->C.c1f[n] = fk2v[n] * xln * s / (pow(4.0, n));
        e->C.c2f[n] = e->C.c2p[n];
        e->C.c4p[n] = 4 * (n + 1 / 2.0) * fk2v[n] * z12v[n] * xrn /
                      ((n + 2) * (n + 1) * fkv[n] * fkv[n] * (n + 1) * fkv[n]);
        s = 0.0;
        for (int m = 0; m <= n; ++m)
            s = s + z12v[m] * z12v[n - m] / ((m + 1) * fkv[m] * (n - m + 1) * fkv[n - m] * fkv[m] * fkv[n - m]);
        e->C.c4f[n] = xrn * s;
        if (s1 && e->C.c1p[n] < 1e-50) { *n1 = n; s1 = 0; }
        if (s4 && e->C.c4p[n] < 1e-50) { *n4 = n; s4 = 0; }
        if (fabs(e->C.c1f[n]) < tiny && n < *n1f) *n1f = n;
        if (fabs(e->C.c4f[n]) < tiny && n < *n4f) *n4f = n;
    }
    /* QUIRK: a second general-orientation block (Source/sc_calc_generic_cpu.cpp:5549-5618)
     * then overwrites carr1p (with the radial Schulz z and 4^(2n)) and fills carr3p; its
     * qrombdeltac(i0=1,i1=5) loop only discards results and is omitted.  n1 keeps the
     * value found above (search1 is already off); n3 = first negative carr3p. */
    {
        const double b1s = 2; /* cylinder cross-section (:5584) */
        double b1sv = 1.0, z12vn = 1.0, xr = 1.0;
        long double xl = 1.0;
        int s3 = 1;
        for (int n = 1; n <= NNMAX; ++n) {
            z12vn = z12vn * ((z + 1) - 2 + 2 * n) * ((z + 1) - 1 + 2 * n);
            b1sv = b1sv * (b1s - 1 + n);
            xl = -xl * xl2z;
            xr = -xr * xr2z;
            e->C.c1p[n] = pow(4.0, 2 * n) * z12vn * xl / ((2.0 * n + 1) * (n + 1));
            e->C.c3p[n] = 4 * (n + 1 / 2.0) * fk2v[n] * z12vn * xr /
                          ((n + 2) * (n + 1) * fkv[n] * fkv[n] * b1sv * fkv[n]);
            if (s3 && e->C.c3p[n] < 1e-50) { n1[2] = n; s3 = 0; } /* n1[2] is n3 */
        }
    }
}

static int bad_coef(double v) { return isinf(v) || isnan(v) || fabs(v) < 1e-99; }

/* coefficients() driver (Source/sc_calc_generic_cpu.cpp:2099-2140, 9682-9697) and the
 * Label99 back-off + limq (:8046-8090).  Returns 0 if a branch filled the tables. */
static int coefficients(csxo_engine* e)
{
    coeff_tables* C = &e->C;
    double* arrs[18] = {C->c1p, C->c2p, C->c3p, C->c4p, C->c5p, C->c6p, C->c7p, C->c8p, C->c9p,
                        C->c1f, C->c2f, C->c3f, C->c4f, C->c5f, C->c6f, C->c7f, C->c8f, C->c9f};
    for (int a = 0; a < 18; ++a)
        for (int n = 0; n < NCOEF; ++n) arrs[a][n] = 1;
    for (int n = 0; n < NIM2; ++n)
        for (int m = 0; m < NIM2; ++m) C->c11pm[n][m] = 1;

    const double z = (1.0 - sq(e->sigma)) / sq(e->sigma);
    const double zl = (1.0 - sq(e->sigmal)) / sq(e->sigmal);
    const double xlz = e->length / (2.0 * (zl + 1)), xl2z = sq(xlz);
    const double xrz = e->radius / (2.0 * (z + 1)), xr2z = sq(xrz);
    int nidx[18];
    for (int a = 0; a < 18; ++a) nidx[a] = NNMAX;

    e->orcase = ORC_GENERAL;
    if ((e->polPhi == 0) && (e->polTheta == 90)) e->orcase = ORC_X;
    if ((e->polPhi == 90) && (e->polTheta == 90)) e->orcase = ORC_Y;
    if ((e->polPhi == 0) && (e->polTheta == 0)) e->orcase = ORC_Z;
    if ((e->polPhi == 90) && (e->polTheta == 0)) e->orcase = ORC_Z;

    int done = 0;
    if (e->cdim == 3) {
        if (e->cs == 0) { coef_sphere(e, z, xr2z, &nidx[3], &nidx[12]); done = 1; }
    } else if (e->ordis == 7 && e->cdim == 4) {
        /* Only the homogeneous cube has a form factor; the pixel of any other interior uses
         * P(q) = 1 (Source/sc_gpu_pCub_lNon_oIso.h:37-45), so its tables stay unused. */
        if (e->cs == 0) coef_cube(e, z, xr2z, &nidx[3], &nidx[12]);
        done = 1;
    } else if (e->ordis == 6 && e->cdim == 4) {
        /* perfectly oriented cubes: the homogeneous interior only (the pixel of the others is
         * undefined in the reference) */
        if (e->cs == 0) { coef_cube_zdir(e, z, xr2z, &nidx[3], &nidx[12]); done = 1; }
    } else if (e->ordis == 7 && e->cdim == 1) {
        /* short cylinders (L/R < 2, the exact-average branch :3592-3685) are not restated */
        if (e->cs == 0 && e->length / e->radius >= 2.0) {
            coef_cyl_iso(e, z, zl, xl2z, xr2z, &nidx[0], &nidx[3], &nidx[9], &nidx[12]);
            done = 1;
        }
    } else if (e->ordis == 6 && e->cdim != 3) {
        if (e->cs == 0) { coef_cyl_zdir(e, z, zl, xl2z, xr2z, &nidx[0], &nidx[3], &nidx[9], &nidx[12]); done = 1; }
    } else if (e->ordis == 0 && e->cdim != 3) {
        if (e->cdim == 1 && e->cs == 0) {
            coef_cyl_gauss(e, z, zl, xl2z, xr2z, &nidx[0], &nidx[3], &nidx[9], &nidx[12]);
            done = (e->orcase == ORC_GENERAL);
        }
    }
    if (!done) return -1;

    for (int a = 0; a < 18; ++a) {
        while (nidx[a] > 1 && bad_coef(arrs[a][nidx[a]])) nidx[a]--;
        e->limq[a] = pow(fabs(arrs[a][nidx[a]]), -1 / (2.0 * nidx[a]));
    }
    return 0;
}

/* ========================================================================== */
/* Lattice: reflection lists (ButtonHKLClick), fhkl_c, corotations            */
/* ========================================================================== */

/* Multiplicity by index pattern (Source/sc_calc_generic_cpu.cpp:9780-9870) for the
 * cubic and tetragonal systems; pattern classification as in :9780-9803. */
enum { P000, P0KL, PH0L, PHK0, PH00, P0K0, P00L, PHHL, PHLH, PLHH, PHH0, PH0H, P0HH, PHHH, PHKL };

static int hkl_pattern(int h, int k, int l)
{
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

static int mult_cubic(int pat)
{
    switch (pat) {
    case PHKL: return 48;
    case PHHL: case PHLH: case PHK0: case PH0L: case P0KL: case PLHH: return 24;
    case PHH0: case PH0H: case P0HH: return 12;
    case PHHH: return 8;
    case PH00: case P0K0: case P00L: return 6;
    default: return 0;
    }
}

static int mult_tetragonal(int pat)
{
    switch (pat) {
    case PHKL: return 16;
    case PHHL: case PHLH: case PLHH: case PH0H: case P0HH: case PHHH: case PHK0: case PH0L: case P0KL: return 8;
    case PHH0: case PH00: case P0K0: return 4;
    case P00L: return 2;
    default: return 0;
    }
}

static int even(int v) { return (v % 2) == 0; }

/* BCT, tetragonal group I--- (ComboBoxTetragonal 23: c1=0, c2=0, c4=0, c6=0, c7=0,
 * c8=-1, c10=0; Source/sc_calc_generic_cpu.cpp:8367, :8833-8835, conditions :9992-10056). */
static int allowed_bct(int h, int k, int l)
{
    int c1 = 0, c2 = 1, c4 = 1, c6 = 1, c7 = 1, c10 = 1;
    if (even(h + k + l)) c1 = 1;
    if (h == 0 && !even(k + l)) c2 = 0;
    if (k == 0 && !even(h + l)) c2 = 0;
    if (l == 0 && !even(h + k)) c4 = 0;
    if (h == k && l != 0 && !even(l)) c10 = 0;
    if (h == 0 && l == 0 && !even(k)) c6 = 0;
    if (k == 0 && l == 0 && !even(h)) c6 = 0;
    if (h == 0 && k == 0 && !even(l)) c7 = 0;
    return c1 * c2 * c4 * c6 * c7 * c10;
}

/* The cubic system's reflection conditions (Source/sc_calc_generic_cpu.cpp:9941-9990) with the
 * codes of the space group ButtonHKLClick selects for the lattice (SETCX, :8529-8600):
 * c1 integral (-1 none, 0 h+k+l, 6 h+k,h+l,k+l), c2 zonal (0 k+l, 3 k+l=4n and k,l, 4 k,l,
 * 5 k cyclic), c7 serial (0 l, 1 l=4n), c10 hhl (0 l, 2 h+l, 4 2h+l=4n and l, 5 h,l). */
static int allowed_cubic(int c1, int c2, int c7, int c10, int h, int k, int l)
{
    int c1case = 0, c2case = 1, c7case = 1, c10case = 1;
    if (c1 == -1) c1case = 1;
    if (c1 == 0 && even(h + k + l)) c1case = 1;
    if (c1 == 6 && even(h + k) && even(h + l) && even(k + l)) c1case = 1;
    if (c2 == 0 && h == 0 && !even(k + l)) c2case = 0;
    if (c2 == 0 && k == 0 && !even(h + l)) c2case = 0;
    if (c2 == 0 && l == 0 && !even(h + k)) c2case = 0;
    if (c2 == 3 && h == 0 && !(((k + l) % 4) == 0 && even(k) && even(l))) c2case = 0;
    if (c2 == 3 && k == 0 && !(((h + l) % 4) == 0 && even(h) && even(l))) c2case = 0;
    if (c2 == 3 && l == 0 && !(((h + k) % 4) == 0 && even(h) && even(k))) c2case = 0;
    if (c2 == 4 && h == 0 && !(even(k) && even(l))) c2case = 0;
    if (c2 == 4 && k == 0 && !(even(h) && even(l))) c2case = 0;
    if (c2 == 4 && l == 0 && !(even(h) && even(k))) c2case = 0;
    if (c2 == 5 && h == 0 && !even(k)) c2case = 0;
    if (c2 == 5 && k == 0 && !even(l)) c2case = 0;
    if (c2 == 5 && l == 0 && !even(h)) c2case = 0;
    if (c10 == 0 && h == k && l != 0 && !even(l)) c10case = 0;
    if (c10 == 0 && h == l && k != 0 && !even(k)) c10case = 0;
    if (c10 == 0 && k == l && h != 0 && !even(h)) c10case = 0;
    if (c10 == 2 && h == k && l != 0 && !even(h + l)) c10case = 0;
    if (c10 == 2 && h == l && k != 0 && !even(h + k)) c10case = 0;
    if (c10 == 2 && k == l && h != 0 && !even(h + k)) c10case = 0;
    if (c10 == 4 && h == k && l != 0 && !(((2 * h + l) % 4) == 0 && even(l))) c10case = 0;
    if (c10 == 4 && h == l && k != 0 && !(((2 * h + k) % 4) == 0 && even(k))) c10case = 0;
    if (c10 == 4 && k == l && h != 0 && !(((2 * k + h) % 4) == 0 && even(h))) c10case = 0;
    if (c10 == 5 && h == k && l != 0 && !(even(h) && even(l))) c10case = 0;
    if (c10 == 5 && h == l && k != 0 && !(even(h) && even(k))) c10case = 0;
    if (c10 == 5 && k == l && h != 0 && !(even(k) && even(h))) c10case = 0;
    if (c7 == 0 && h == 0 && k == 0 && !even(l)) c7case = 0;
    if (c7 == 0 && h == 0 && l == 0 && !even(k)) c7case = 0;
    if (c7 == 0 && k == 0 && l == 0 && !even(h)) c7case = 0;
    if (c7 == 1 && h == 0 && k == 0 && !((l % 4) == 0)) c7case = 0;
    if (c7 == 1 && h == 0 && l == 0 && !((k % 4) == 0)) c7case = 0;
    if (c7 == 1 && k == 0 && l == 0 && !((h % 4) == 0)) c7case = 0;
    return c1case * c2case * c7case * c10case;
}

/* The cubic lattices' space-group codes (ButtonHKLClick :8444-8500 picks ComboBoxCubic per
 * lattice type; SETCX :8529-8600): BCC and Im3m I--- (7), FCC F--- (12), SC P--- (0), Ia3d
 * Ia-d (11), Pn3m Pn-- (5), Fd3m Fd-- (15).  Returns 0 for a lattice that is not cubic. */
static int cubic_codes(int ltype, int* c)
{
    switch (ltype) {
    case 4: case 11: c[0] = 0; c[1] = 0; c[2] = 0; c[3] = 0; return 1;
    case 5: c[0] = 6; c[1] = 4; c[2] = 0; c[3] = 2; return 1;
    case 7: c[0] = -1; c[1] = -1; c[2] = -1; c[3] = -1; return 1;
    case 9: c[0] = 0; c[1] = 4; c[2] = 1; c[3] = 4; return 1;
    case 10: c[0] = -1; c[1] = 0; c[2] = 0; c[3] = -1; return 1;
    case 17: c[0] = 6; c[1] = 3; c[2] = 1; c[3] = 2; return 1;
    default: return 0;
    }
}

/* hkl loops (Source/sc_calc_generic_cpu.cpp:9712-10291): pass 1 (isotropic) over
 * [0,H]^3 with multiplicity, pass 2 (anisotropic) over [-H,H]^3, h-major. */
static void build_reflections(csxo_engine* e)
{
    const int H = e->hklmax;
    int cc[4];
    const int cubic = cubic_codes(e->ltype, cc);
    /* 2D lattices (:9585-9593, :9720-9722, :9812, :9936): l runs to lmax = 0 (pass 1: l = 0;
     * pass 2: l in [-H, 0]), every (h, k, l) but (0, 0, 0) is accepted, multiplicity 1 */
    const int sys2d = e->ltype >= 1 && e->ltype <= 3;
    for (int pass = 1; pass <= 2; ++pass) {
        const int lo = (pass == 1) ? 0 : -H;
        const int lmax = sys2d ? 0 : H;
        int ct = 1;
        int(*lp)[6] = (pass == 1) ? e->latpar1 : e->latpar2;
        lp[1][0] = 0;
        for (int h = lo; h <= H; ++h)
            for (int k = lo; k <= H; ++k)
                for (int l = lo; l <= lmax; ++l) {
                    const int pat = hkl_pattern(h, k, l);
                    const int mult = sys2d ? 1 : cubic ? mult_cubic(pat) : mult_tetragonal(pat);
                    /* HCP and orthorhombic: their systems' condition blocks are commented out in
                     * the reference, so ccase keeps its initial 0 (:8256, :10059-10145) */
                    int ok = sys2d ? 1
                             : cubic ? allowed_cubic(cc[0], cc[1], cc[2], cc[3], h, k, l)
                                     : (e->ltype == 6 || e->ltype == 18) ? 0 : allowed_bct(h, k, l);
                    if (h == 0 && k == 0 && l == 0) ok = 0;
                    if (ok) {
                        lp[1][0] = ct;
                        lp[ct][1] = h; lp[ct][2] = k; lp[ct][3] = l; lp[ct][4] = mult;
                        ct++;
                    }
                }
    }
}

/* Structure factor |sum exp(2 pi i h.r)|^2 over the basis (fhkl_c, BCC :10374-10393, FCC
 * :10397-10420, SC :10449-10455, BCT :10457-10478, Ia3d :10480-10501, Pn3m :10503-10509, Im3m
 * :10511-10532, Fd3m :10534-10563; basis coordinates in float as in the reference). */
static double fhkl_value(int ltype, int h, int k, int l, double* sphno)
{
    float r[9][4];
    int nn;
    memset(r, 0, sizeof(r));
    if (ltype == 7 || ltype == 10 || ltype == 1 || ltype == 2) { /* SC, Pn3m, hex and square 2D */
        *sphno = 1;
        return 1;
    }
    if (ltype == 3) { /* centred rectangular 2D: (0,0), (1/2,1/2) in the plane (:10352-10371) */
        const float r2 = 1 / 2.;
        double sc = 1.0, ss = 0.0; /* atom 1 at the origin: cos 0, sin 0 */
        sc = sc + cos(2 * M_PI * (h * r2 + k * r2));
        ss = ss + sin(2 * M_PI * (h * r2 + k * r2));
        *sphno = 2;
        (void)l;
        return round(sc * sc + ss * ss);
    }
    if (ltype == 5) {
        *sphno = 4; nn = 4;
        r[2][1] = 1 / 2.; r[2][2] = 1 / 2.;
        r[3][2] = 1 / 2.; r[3][3] = 1 / 2.;
        r[4][1] = 1 / 2.; r[4][3] = 1 / 2.;
    } else if (ltype == 17) { /* Fd3m, 8 atoms (:10537-10548) */
        *sphno = 8; nn = 8;
        r[2][1] = 1 / 2.; r[2][2] = 1 / 2.;
        r[3][2] = 1 / 2.; r[3][3] = 1 / 2.;
        r[4][1] = 1 / 2.; r[4][3] = 1 / 2.;
        r[5][1] = 1 / 4.; r[5][2] = 1 / 4.; r[5][3] = 1 / 4.;
        r[6][1] = 3 / 4.; r[6][2] = 3 / 4.; r[6][3] = 1 / 4.;
        r[7][1] = 1 / 4.; r[7][2] = 3 / 4.; r[7][3] = 3 / 4.;
        r[8][1] = 3 / 4.; r[8][2] = 1 / 4.; r[8][3] = 3 / 4.;
    } else { /* BCC, BCT, Ia3d, Im3m: 2 atoms */
        *sphno = 2; nn = 2;
        r[2][1] = 1 / 2.; r[2][2] = 1 / 2.; r[2][3] = 1 / 2.;
    }
    double sc = 0.0, ss = 0.0;
    for (int i = 1; i <= nn; ++i) {
        const double arg = 2 * M_PI * (h * r[i][1] + k * r[i][2] + l * r[i][3]);
        sc = sc + cos(arg);
        ss = ss + sin(arg);
    }
    return round(sc * sc + ss * ss);
}

/* corotations (Source/sc_calc_generic_cpu.cpp:1487-2098): the reciprocal basis rotated so
 * the crystal axis (u,v,w) points along the beam, with the reference's float intermediates.
 * Outputs ri (row-major m11..m33), rt (twin) and the unit-cell volume vvol.  For the 2D
 * lattices (lat2d) ri is the rotated in-plane reciprocal basis a*, b* of the cell (a, b,
 * gamma) and c* = (0, 0, 1), and vvol the cell area (:1613-1615, :1676-1729, :1935-1943,
 * :2060-2066); rt stays the 3D twin. */
static void corotations(double a, double b, double c, double alpha, double beta, double gamma,
                        double u, double v, double w, double ephi, int lat2d, double* M, double* MT, double* vvol)
{
    alpha = alpha * M_PI / 180.0; beta = beta * M_PI / 180.0;
    gamma = gamma * M_PI / 180.0; ephi = ephi * M_PI / 180.0;
    const double ca = cos(alpha), cb = cos(beta), cg = cos(gamma), sb = sin(beta);
    const double n1 = 0, n2 = 0, n3 = 1, l1 = sb, l2 = 0, l3 = cb;
    const double m1 = (cg - cb * ca) / sb;
    const double m2 = sqrt(1.0 - m1 * m1 - ca * ca), m3 = ca;
    /* Msi and its inverse Ms */
    const double s11 = a * l1, s12 = a * l2, s13 = a * l3;
    const double s21 = b * m1, s22 = b * m2, s23 = b * m3;
    const double s31 = c * n1, s32 = c * n2, s33 = c * n3;
    const double ds = s11 * s22 * s33 + s12 * s23 * s31 + s13 * s21 * s32 - s31 * s22 * s13 - s32 * s23 * s11 - s33 * s21 * s12;
    const double ms11 = (s22 * s33 - s23 * s32) / ds, ms12 = (s13 * s32 - s12 * s33) / ds;
    const double ms13 = (s12 * s23 - s13 * s22) / ds, ms21 = (s23 * s31 - s21 * s33) / ds;
    const double ms22 = (s11 * s33 - s13 * s31) / ds, ms23 = (s13 * s21 - s11 * s23) / ds;
    const double ms31 = (s21 * s32 - s22 * s31) / ds, ms32 = (s12 * s31 - s11 * s32) / ds;
    const double ms33 = (s11 * s22 - s12 * s21) / ds;
    /* transpose, invert */
    const double t11 = ms11, t12 = ms21, t13 = ms31, t21 = ms12, t22 = ms22, t23 = ms32, t31 = ms13, t32 = ms23, t33 = ms33;
    const double dt = t11 * t22 * t33 + t12 * t23 * t31 + t13 * t21 * t32 - t31 * t22 * t13 - t32 * t23 * t11 - t33 * t21 * t12;
    const double mt11 = (t22 * t33 - t23 * t32) / dt, mt12 = (t13 * t32 - t12 * t33) / dt;
    const double mt13 = (t12 * t23 - t13 * t22) / dt, mt21 = (t23 * t31 - t21 * t33) / dt;
    const double mt22 = (t11 * t33 - t13 * t31) / dt, mt23 = (t13 * t21 - t11 * t23) / dt;
    const double mt31 = (t21 * t32 - t22 * t31) / dt, mt32 = (t12 * t31 - t11 * t32) / dt;
    const double mt33 = (t11 * t22 - t12 * t21) / dt;
    /* M = mt^T (the reference stores it in the output references m11..m33 first) */
    const double M11 = mt11, M12 = mt21, M13 = mt31, M21 = mt12, M22 = mt22, M23 = mt32, M31 = mt13, M32 = mt23, M33 = mt33;
    /* base vectors a,b,c and the twinned cell in xyz (float) */
    const float aax = 1, aay = 0, aaz = 0, bax = 0, bay = 1, baz = 0, cax = 0, cay = 0, caz = 1;
    const float aaxt = 2 / 3.0, aayt = 2 / 3.0, aazt = -1 / 3.0, baxt = -1 / 3.0, bayt = 2 / 3.0, bazt = 2 / 3.0;
    const float caxt = 2 / 3.0, cayt = -1 / 3.0, cazt = 2 / 3.0;
    const float aex = M11 * aax + M12 * aay + M13 * aaz, aey = M21 * aax + M22 * aay + M23 * aaz, aez = M31 * aax + M32 * aay + M33 * aaz;
    const float bex = M11 * bax + M12 * bay + M13 * baz, bey = M21 * bax + M22 * bay + M23 * baz, bez = M31 * bax + M32 * bay + M33 * baz;
    const float cex = M11 * cax + M12 * cay + M13 * caz, cey = M21 * cax + M22 * cay + M23 * caz, cez = M31 * cax + M32 * cay + M33 * caz;
    const float aext = M11 * aaxt + M12 * aayt + M13 * aazt, aeyt = M21 * aaxt + M22 * aayt + M23 * aazt, aezt = M31 * aaxt + M32 * aayt + M33 * aazt;
    const float bext = M11 * baxt + M12 * bayt + M13 * bazt, beyt = M21 * baxt + M22 * bayt + M23 * bazt, bezt = M31 * baxt + M32 * bayt + M33 * bazt;
    const float cext = M11 * caxt + M12 * cayt + M13 * cazt, ceyt = M21 * caxt + M22 * cayt + M23 * cazt, cezt = M31 * caxt + M32 * cayt + M33 * cazt;
    const double vol = aex * (bey * cez - bez * cey) + aey * (bez * cex - bex * cez) + aez * (bex * cey - bey * cex);
    /* beam direction (u,v,w) in xyz (float), unit director */
    const float nx = M11 * u + M12 * v + M13 * w, ny = M21 * u + M22 * v + M23 * w, nz = M31 * u + M32 * v + M33 * w;
    /* C++ overload resolution in the reference picks the float versions of
     * sqrt/acos/asin/sin/cos for float arguments; ephi stays double. */
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    const float nxn = nx / len, nzn = nz / len;
    const float eth = acosf(-nzn);
    const float eps_ = asinf(nxn / sinf(eth));
    const float r11 = cosf(eps_) * cos(ephi) - sinf(eps_) * cosf(eth) * sin(ephi);
    const float r12 = -cosf(eps_) * sin(ephi) - sinf(eps_) * cosf(eth) * cos(ephi);
    const float r13 = sinf(eps_) * sinf(eth);
    const float r21 = sinf(eps_) * cos(ephi) + cosf(eps_) * cosf(eth) * sin(ephi);
    const float r22 = -sinf(eps_) * sin(ephi) + cosf(eps_) * cosf(eth) * cos(ephi);
    const float r23 = -cosf(eps_) * sinf(eth);
    const float r31 = sinf(eth) * sin(ephi), r32 = sinf(eth) * cos(ephi), r33 = cosf(eth);
    /* inverse rotation (double) */
    const double dr = r11 * r22 * r33 + r12 * r23 * r31 + r13 * r21 * r32 - r31 * r22 * r13 - r32 * r23 * r11 - r33 * r21 * r12;
    const double q11 = (r22 * r33 - r23 * r32) / dr, q12 = (r13 * r32 - r12 * r33) / dr, q13 = (r12 * r23 - r13 * r22) / dr;
    const double q21 = (r23 * r31 - r21 * r33) / dr, q22 = (r11 * r33 - r13 * r31) / dr, q23 = (r13 * r21 - r11 * r23) / dr;
    const double q31 = (r21 * r32 - r22 * r31) / dr, q32 = (r12 * r31 - r11 * r32) / dr, q33 = (r11 * r22 - r12 * r21) / dr;
    /* rotated base vectors (float) */
    const float adx = q11 * aex + q12 * aey + q13 * aez, ady = q21 * aex + q22 * aey + q23 * aez, adz = q31 * aex + q32 * aey + q33 * aez;
    const float bdx = q11 * bex + q12 * bey + q13 * bez, bdy = q21 * bex + q22 * bey + q23 * bez, bdz = q31 * bex + q32 * bey + q33 * bez;
    const float cdx = q11 * cex + q12 * cey + q13 * cez, cdy = q21 * cex + q22 * cey + q23 * cez, cdz = q31 * cex + q32 * cey + q33 * cez;
    const double adxt = q11 * aext + q12 * aeyt + q13 * aezt, adyt = q21 * aext + q22 * aeyt + q23 * aezt, adzt = q31 * aext + q32 * aeyt + q33 * aezt;
    const double bdxt = q11 * bext + q12 * beyt + q13 * bezt, bdyt = q21 * bext + q22 * beyt + q23 * bezt, bdzt = q31 * bext + q32 * beyt + q33 * bezt;
    const double cdxt = q11 * cext + q12 * ceyt + q13 * cezt, cdyt = q21 * cext + q22 * ceyt + q23 * cezt, cdzt = q31 * cext + q32 * ceyt + q33 * cezt;
    /* reciprocal vectors from cross products / vol (double) */
    M[0] = (bdy * cdz - bdz * cdy) / vol; M[1] = (bdz * cdx - bdx * cdz) / vol; M[2] = (bdx * cdy - bdy * cdx) / vol;
    M[3] = (adz * cdy - ady * cdz) / vol; M[4] = (adx * cdz - adz * cdx) / vol; M[5] = (ady * cdx - adx * cdy) / vol;
    M[6] = (ady * bdz - adz * bdy) / vol; M[7] = (adz * bdx - adx * bdz) / vol; M[8] = (adx * bdy - ady * bdx) / vol;
    MT[0] = (bdyt * cdzt - bdzt * cdyt) / vol; MT[1] = (bdzt * cdxt - bdxt * cdzt) / vol; MT[2] = (bdxt * cdyt - bdyt * cdxt) / vol;
    MT[3] = (adzt * cdyt - adyt * cdzt) / vol; MT[4] = (adxt * cdzt - adzt * cdxt) / vol; MT[5] = (adyt * cdxt - adxt * cdyt) / vol;
    MT[6] = (adyt * bdzt - adzt * bdyt) / vol; MT[7] = (adzt * bdxt - adxt * bdzt) / vol; MT[8] = (adxt * bdyt - adyt * bdxt) / vol;
    *vvol = vol;
    if (lat2d) {
        const double sg = sin(gamma);
        const double m11 = a, m12 = b * cg, m13 = 0, m21 = 0, m22 = b * sg, m23 = 0;
        const double aexd = m11 * aax + m12 * aay + m13 * aaz, aeyd = m21 * aax + m22 * aay + m23 * aaz;
        const double bexd = m11 * bax + m12 * bay + m13 * baz, beyd = m21 * bax + m22 * bay + m23 * baz;
        const double area = a * b * sg;
        const double asx = beyd / area, asy = -bexd / area, asz = 0;
        const double bsx = -aeyd / area, bsy = aexd / area, bsz = 0;
        const double csx = 0, csy = 0, csz = 1;
        M[0] = q11 * asx + q12 * asy + q13 * asz; M[1] = q21 * asx + q22 * asy + q23 * asz;
        M[2] = q31 * asx + q32 * asy + q33 * asz;
        M[3] = q11 * bsx + q12 * bsy + q13 * bsz; M[4] = q21 * bsx + q22 * bsy + q23 * bsz;
        M[5] = q31 * bsx + q32 * bsy + q33 * bsz;
        M[6] = q11 * csx + q12 * csy + q13 * csz; M[7] = q21 * csx + q22 * csy + q23 * csz;
        M[8] = q31 * csx + q32 * csy + q33 * csz;
        *vvol = area;
    }
}

/* ========================================================================== */
/* prepareCalculation                                                          */
/* ========================================================================== */

int csxo_prepare(csxo_engine* e)
{
    /* Source/sc_calc_generic_cpu.cpp:319-988, the parts that feed the hot path. */
    e->shp = e->peakidx + 1; /* setComboBoxPeak: shp = index+1 (.h:254) */
    /* setDisplacement (.h:257); dw_zuf: the TPV factor applied afterwards (cpu:1126-1130) */
    e->dwfactor = e->displacement * e->displacement / 3.0 * e->dw_zuf;
    e->pixx_m = e->pixx_mm / 1000.; e->pixy_m = e->pixy_mm / 1000.;
    e->zmax = e->gridpoints;
    e->use1d = 0;
    e->lattice = (e->ltype != 12);
    e->twin = e->twinned;
    e->polPhi = e->phi; e->polTheta = e->theta;
    const int quad = e->radq1 ? 1 : (e->radq2 ? 2 : 4);
    e->zzmin = (quad == 4) ? -e->zmax : 0;
    e->zzmax = e->zmax;
    e->iimin = (quad == 1) ? 0 : -e->zmax;
    e->iimax = e->zmax;
    e->phiwidth = 4. / e->azi;
    /* peak-shape parameters: setPeakPar (.h:224), prepareCalculation cpu:585-590 */
    e->eta = e->peakpar / 100.;
    e->beta = e->peakpar;
    e->c0 = 1;
    e->c1 = 4;
    e->c2 = 4 * (sqrt(2.0) - 1);
    e->c3 = 4 * (exp(2 * log(2.0) / 3.0) - 1);
    e->c4 = (e->beta == 0) ? 1 : exp(log(2.) / e->beta) - 1;
    e->p1 = e->radius / e->radiusi;
    e->cosphi = cos(e->polPhi * M_PI / 180.0); e->sinphi = sin(e->polPhi * M_PI / 180.0);
    e->costheta = cos(e->polTheta * M_PI / 180.0); e->sintheta = sin(e->polTheta * M_PI / 180.0);
    const double cosa = cos(e->ucalpha * M_PI / 180.0), cosb = cos(e->ucbeta * M_PI / 180.0);
    const double cosg = cos(e->ucgamma * M_PI / 180.0);
    e->ucvol = e->uca * e->ucb * e->ucc * sqrt(1.0 - sq(cosa) - sq(cosb) - sq(cosg) + 2 * cosa * cosb * cosg);

    switch (e->particle) {
    case 0: e->cdim = 3; break;
    case 1: e->cdim = 1; break;
    case 2: e->cdim = 2; break;
    default: e->cdim = e->particle; break;
    }
    e->cs = e->interior;
    switch (e->ltype) { /* dist by lattice (:734-767) */
    case 5: case 13: case 17: case 19: e->dist = sqrt(2.0) * e->uca / 2.0; break;
    case 4: case 9: case 11: e->dist = sqrt(3.0) * e->uca / 2.0; break;
    default: e->dist = e->uca; break;
    }
    e->reldis = e->displacement / e->dist;

    if (coefficients(e) != 0) return -2;

    e->peakmax1 = e->peakmax2 = 0;
    e->lat_ucvol = e->ucvol; e->lat_sphno = 1;
    if (e->lattice) {
        {
            int cc[4];
            if (!cubic_codes(e->ltype, cc) && !(e->ltype >= 1 && e->ltype <= 3) && e->ltype != 6 &&
                e->ltype != 8 && e->ltype != 18)
                return -3; /* 2D lattices, cubic lattices, HCP, BCT, orthorhombic */
        }
        build_reflections(e);
        e->peakmax1 = e->latpar1[1][0];
        e->peakmax2 = e->latpar2[1][0];
        double vv;
        corotations(e->uca, e->ucb, e->ucc, e->ucalpha, e->ucbeta, e->ucgamma, e->ucn1, e->ucn2,
                    e->ucn3, e->ucpsi, e->ltype >= 1 && e->ltype <= 3, e->ri, e->rt, &vv);
        e->ucvol = vv; /* corotations writes the cell volume back (:1049-1052) */
        double sphno = 1;
        for (int p = 1; p <= e->peakmax1; ++p) {
            const double f = fhkl_value(e->ltype, e->latpar1[p][1], e->latpar1[p][2], e->latpar1[p][3], &sphno);
            e->latpar1[p][5] = (int)round(f);
            e->lat_ucvol = e->ucvol;
            e->lat_sphno = sphno;
        }
        if (e->peakmax1 == 0) { e->lat_ucvol = e->ucvol; e->lat_sphno = 1; }
        const double* R = e->ri;
        const double* T = e->rt;
        for (int p = 1; p <= e->peakmax2; ++p) {
            const int h = e->latpar2[p][1], k = e->latpar2[p][2], l = e->latpar2[p][3];
            const double f = fhkl_value(e->ltype, h, k, l, &sphno);
            e->latpar2[p][5] = (int)round(f);
            e->latpar2[p][4] = 1;
            float* L3 = e->latpar3[p];
            L3[1] = 2 * M_PI / e->uca;
            const double qx = (2 * M_PI) * (R[0] * h + R[3] * k + R[6] * l);
            const double qy = (2 * M_PI) * (R[1] * h + R[4] * k + R[7] * l);
            const double qz = (2 * M_PI) * (R[2] * h + R[5] * k + R[8] * l);
            const double qxy = (2 * M_PI) * sqrt(h * h * (R[0] * R[0] + R[1] * R[1] + R[2] * R[2]) +
                                                 k * k * (R[3] * R[3] + R[4] * R[4] + R[5] * R[5]));
            const double qh = sqrt(qx * qx + qy * qy + qz * qz);
            L3[2] = qx; L3[3] = qy; L3[4] = qz; L3[5] = qh; L3[6] = qh / (2.0 * M_PI);
            L3[7] = (2 * M_PI) * (T[0] * h + T[3] * k + T[6] * l);
            L3[8] = (2 * M_PI) * (T[1] * h + T[4] * k + T[7] * l);
            L3[9] = (2 * M_PI) * (T[2] * h + T[5] * k + T[8] * l);
            const double g3 = 4 * M_PI * M_PI / (2.0 * M_PI * qh * qh);
            const double x2phi = 4 * qh * qh / (M_PI * e->phiwidth * e->phiwidth);
            double pn1 = 0; /* the peak norm of each shape (cpu:899-928) */
            switch (e->shp) {
            case 1: case 5: pn1 = csxo_lorentznorm3(x2phi); break;
            case 2: case 7: pn1 = csxo_gaussnorm3(x2phi); break;
            case 3: pn1 = csxo_pearsonnorm3(x2phi, 2); break;
            case 4: pn1 = csxo_pearsonnorm3(x2phi, 3 / 2.0); break;
            case 6: pn1 = csxo_pearsonnorm3(x2phi, e->beta); break;
            default: break;
            }
            L3[10] = g3; L3[11] = pn1; L3[12] = 0; L3[13] = qxy;
        }
        e->sphno = sphno;
    }
    e->sinphic = sin(-e->polPhi * M_PI / 180.);
    e->cosphic = cos(-e->polPhi * M_PI / 180.);
    e->xmin = e->zzmin; e->xmax = e->zzmax; e->ymin = e->iimin; e->ymax = e->iimax;
    e->prepared = 1;
    return 0;
}

unsigned csxo_combo_key(const csxo_engine* e)
{
    const int shp = e->peakidx + 1;
    return ((unsigned)e->particle << 24) | (e->ltype == 12 ? 0u : ((unsigned)shp << 16)) |
           ((unsigned)e->ltype << 8) | (unsigned)e->ordis;
}

void csxo_bounds(const csxo_engine* e, int* b)
{
    b[0] = e->xmin; b[1] = e->xmax; b[2] = e->ymin; b[3] = e->ymax;
}

/* ========================================================================== */
/* Per-pixel physics                                                           */
/* ========================================================================== */

/* Series with the reference's exit rule: relative change < 1e-4 (e.g.
 * Source/sc_lib_formpq_partSphere.h:15247-15259). */
static double series_q2(const double* c, double q, int nmax)
{
    double s = 1.0, old = 0.0, xn = 1.0;
    for (int n = 1; n <= nmax; ++n) {
        xn = xn * q * q; /* (xn*q)*q, as the reference rounds it */
        s = s + c[n] * xn;
        const double d = fabs((s - old) / s);
        if (d < 0.0001) break;
        old = s;
    }
    return s;
}

/* formpq_partSphere, cs=0 (Source/sc_lib_formpq_partSphere.h:11-75). */
static double pq_sphere(const csxo_engine* e, double q)
{
    double zr = (1 - sq(e->sigma)) / (sq(e->sigma));
    if (q < 0.4 * e->limq[3]) {
        double s = 1.0, old = 0.0, qq = 1.0;
        for (int n = 1; n <= 100; ++n) {
            qq = qq * q * q;
            s = s + e->C.c4p[n] * qq;
            const double d = fabs((s - old) / s);
            if (d < 0.0001) break;
            old = s;
        }
        return s;
    }
    const double argq = q * e->radius / (zr + 1);
    if (zr * (zr - 1) * (zr - 2) * (zr - 3) * (zr - 4) * (zr - 5) == 0) zr += kEps3;
    const double pqr = (1 / (2.0 * zr * (zr - 1) * (zr - 2) * (zr - 3))) * pow(argq, -4);
    const double pq1 = pqr * (1 + cos((zr - 3) * atan(2.0 * argq)) / pow(1.0 + 4 * argq * argq, (zr - 3) / 2.0));
    const double pq2 = (pqr / ((zr - 4) * argq)) * sin((zr - 4) * atan(2.0 * argq)) / pow(1.0 + 4 * argq * argq, (zr - 4) / 2.0);
    /* QUIRK: exponent zr-5 then /2.0 on the whole ratio (:15268) */
    const double pq3 = (pqr / ((zr - 4) * (zr - 5) * argq * argq)) * (1 - cos((zr - 5) * atan(2.0 * argq)) / pow(1.0 + 4 * argq * argq, zr - 5) / 2.0);
    return 9.0 * (pq1 - 2.0 * pq2 + pq3);
}

/* formpq_partCube, isotropic and homogeneous (Source/sc_lib_formpq_partCube.h:27-61): the
 * series up to 120 terms below 1.2 limq4, the Porod law por/q^4 above. */
static double pq_cube(const csxo_engine* e, double q)
{
    if (q < 1.2 * e->limq[3]) {
        double s = 1.0, old = 0.0, qq = 1.0;
        for (int n = 1; n <= 120; ++n) {
            qq = qq * q * q;
            s += e->C.c4p[n] * qq;
            const double d = fabs((s - old) / s);
            if (d < 0.0001) break;
            old = s;
        }
        return s;
    }
    return e->por / (q * q * q * q);
}

/* formpq_partCube, perfect orientation (Source/sc_lib_formpq_partCube.h:70-106): the double
 * series over carr11pm below q = 1/R (read as carr11pm[m][n-m]), else the product of the two
 * edge factors. */
static double pq_cube_zdir(const csxo_engine* e, double qxs, double qys, double q)
{
    const double zr = (1 - sq(e->sigma)) / (sq(e->sigma));
    double pql;
    if (q < (1.0 / e->radius)) {
        double qxn[121], qyn[121], old = 0.0;
        pql = 1.0;
        qxn[0] = 1.0; qyn[0] = 1.0;
        for (int n = 1; n <= 80; ++n) {
            qxn[n] = qxn[n - 1] * qxs * qxs;
            qyn[n] = qyn[n - 1] * qys * qys;
            double bs = 0.0;
            for (int m = 0; m <= n; ++m) bs += e->C.c11pm[m][n - m] * qxn[n - m] * qyn[m];
            pql += bs;
            const double d = fabs((pql - old) / pql);
            if (d < 0.0001) break;
            old = pql;
        }
    } else {
        const double ax = qxs * e->radius / (zr + 1) + kEps9;
        const double ay = qys * e->radius / (zr + 1) + kEps9;
        const double px = (1 / (2.0 * zr * (zr - 1))) * (1 / (ax * ax)) *
                          (1 - cos((zr - 1) * atan(2.0 * ax)) / pow(1.0 + 4 * ax * ax, (zr - 1) / 2.0));
        const double py = (1 / (2.0 * zr * (zr - 1))) * (1 / (ay * ay)) *
                          (1 - cos((zr - 1) * atan(2.0 * ay)) / pow(1.0 + 4 * ay * ay, (zr - 1) / 2.0));
        pql = px * py;
    }
    return pql;
}

/* formfq_partSphere, cs=0 (Source/sc_lib_formfq_partSphere.h:8-52). */
static double fq_sphere(const csxo_engine* e, double q)
{
    const double zr = (1 - sq(e->sigma)) / sq(e->sigma);
    if (q < (0.4 * e->limq[12])) {
        double s = 1.0, old = 0.0, qq = 1.0;
        for (int n = 1; n <= 100; ++n) {
            qq = qq * q * q;
            s = s + e->C.c4f[n] * qq;
            const double d = fabs((s - old) / s);
            if (d < 0.0001) break;
            old = s;
        }
        return s;
    }
    const double argq = q * e->radius / (zr + 1);
    const double pqr = (1 / (zr * (zr - 1)) * pow(argq, -2));
    const double pq1 = pqr * cos((zr - 1) * atan(argq)) / pow(1.0 + argq * argq, (zr - 1) / 2.0);
    const double pq2 = (pqr / ((zr - 2) * argq)) * sin((zr - 2) * atan(argq)) / pow(1.0 + argq * argq, (zr - 2) / 2.0);
    const double pq3 = 3 * (pq2 - pq1);
    return pq3 * pq3;
}

static int orc_is(const csxo_engine* e, int pascal_case)
{
    /* QUIRK-ORCASE: compare the raw 0-based enum against the Pascal number,
     * unless fix_orcase maps it (general=1, x=2, y=3, z=4). */
    return (e->fix_orcase ? e->orcase + 1 : e->orcase) == pascal_case;
}

/* formpq_partCylinder, ordis 0 and 6, cs 0 (Source/sc_lib_formpq_partCylinder.h:8-257). */
static double pq_cylinder(const csxo_engine* e, double qxs, double qys, double q)
{
    const double zl = e->sigmal == 1.0 ? kEps9 : (1 - sq(e->sigmal)) / sq(e->sigmal);
    const double zr = e->sigma == 1.0 ? kEps9 : (1 - sq(e->sigma)) / (sq(e->sigma));
    double pql = 0, pqr;
    if (e->ordis == 7) { /* isotropic, factorisation (L/R >= 2; Source/sc_lib_formpq_partCylinder.h:30-90) */
        if (q < (0.6 * e->limq[0])) {
            pql = 1.0;
            double old = 0.0, qq = 1.0;
            for (int n = 1; n <= 120; ++n) {
                qq = qq * q * q;
                pql += e->C.c1p[n] * qq;
                const double d = fabs((pql - old) / pql);
                if (d < 0.0001) break;
                old = pql;
            }
        } else {
            const double arglq = q * e->length / (zl + 1);
            pql = (M_PI / (2.0 * zl)) * (1 / arglq);
            pql = pql - (1 / (2.0 * zl * (zl - 1) * arglq * arglq)) * cos((zl - 1) * atan(2.0 * arglq)) /
                            pow(1.0 + 4 * arglq * arglq, (zl - 1) / 2.0);
        }
    } else if (e->ordis == 6) {
        if (orc_is(e, 4)) pql = 1.0;
        else if (q < (0.6 * e->limq[0])) {
            pql = 1.0;
            double old = 0.0, qq = 1.0;
            for (int n = 1; n <= 120; ++n) {
                qq = qq * (qxs + qys) * (qxs + qys);
                pql += e->C.c1p[n] * qq;
                const double d = fabs((pql - old) / pql);
                if (d < 0.0001) break;
                old = pql;
            }
        } else {
            const double arglq = (qxs + qys + kEps9) * e->length / (zl + 1);
            pql = (1 / (2.0 * zl * (zl - 1))) * (1 / (arglq * arglq)) *
                  (1 - cos((zl - 1) * atan(2.0 * arglq)) / pow(1.0 + 4 * arglq * arglq, (zl - 1) / 2.0));
        }
    } else if (e->ordis == 0) {
        if (orc_is(e, 4)) pql = 1.0;
        else if (q < (0.4 * e->limq[0])) {
            pql = 1.0;
            double old = 0.0, qxn[121], qyn[121];
            qxn[0] = 1.0; qyn[0] = 1.0;
            const int oc = e->fix_orcase ? e->orcase + 1 : e->orcase;
            if (oc >= 1 && oc <= 3) {
                for (int n = 1; n <= 120; ++n) {
                    qxn[n] = qxn[n - 1] * qxs * qxs;
                    qyn[n] = qyn[n - 1] * qys * qys;
                    double bs = 0.0;
                    for (int m = 0; m <= n; ++m) {
                        if (oc == 3) bs += e->C.c11pm[n - m][m] * qyn[m] * qxn[n - m];
                        else bs += e->C.c11pm[n - m][m] * qxn[m] * qyn[n - m];
                    }
                    pql += e->C.c1p[n] * bs;
                    const double d = fabs((pql - old) / pql);
                    if (d < 0.0001) break;
                    old = pql;
                }
            }
        } else {
            const int i1 = e->fix_orcase ? e->orcase + 1 : e->orcase;
            pql = romb_delta(e, e->sigmal, e->polTheta, 0, qxs, qys, 4, i1);
            pql = pql / e->norm;
        }
    }
    /* transverse part, homogeneous (:16246-16279) */
    if ((e->length / e->radius) < 2) pqr = 1;
    else if (q < (0.65 * e->limq[3])) {
        pqr = series_q2(e->C.c4p, q, 120);
    } else {
        const double argpq = q * e->radius / (zr + 1);
        const double pqr1 = (1 / (zr * (zr - 1) * (zr - 2))) * pow(argpq, -3);
        const double pqr2 = (1 / (zr * (zr - 1) * (zr - 2))) * pow(argpq, -3) * sin((zr - 2) * atan(2.0 * argpq)) / pow(1.0 + 4 * argpq * argpq, (zr - 2) / 2.0);
        const double pqr3 = (1 / (zr * (zr - 1) * (zr - 2) * (zr - 3))) * pow(argpq, -4) * cos((zr - 3) * atan(2.0 * argpq)) / pow(1.0 + 4 * argpq * argpq, (zr - 3) / 2.0);
        pqr = (4. / M_PI) * (pqr1 - pqr2 - (9. / 8.0) * pqr3);
    }
    return pql * pqr;
}

/* formfq_partCylinder, ordis 6, cs 0 (Source/sc_lib_formfq_partCylinder.h:8-206). */
static double fq_cylinder(const csxo_engine* e, double limql, double qxs, double qys, double q)
{
    const double zl = (1 - sq(e->sigmal)) / sq(e->sigmal);
    const double zr = (1 - sq(e->sigma)) / sq(e->sigma);
    double pql = 1.0, pqr;
    if (e->ordis == 7) { /* isotropic (Source/sc_lib_formfq_partCylinder.h:27-53) */
        if (q < (0.6 * e->limq[9])) {
            double s = 1.0, old = 0.0, qq = 1.0;
            for (int n = 1; n <= 120; ++n) {
                qq = qq * q * q;
                s += e->C.c1f[n] * qq;
                const double d = fabs((s - old) / s);
                if (d < 0.0001) break;
                old = s;
            }
            pql = s;
        } else {
            const double arglq = q * e->length / (zl + 1);
            pql = (M_PI / (2.0 * zl)) * (1 / arglq);
            pql = pql - (1 / (2.0 * zl * (zl - 1) * arglq * arglq)) * cos((zl - 1) * atan(2.0 * arglq)) /
                            pow(1.0 + 4 * arglq * arglq, (zl - 1) / 2.0);
        }
    }
    if (e->ordis == 6) {
        if (orc_is(e, 4)) pql = 1.0;
        else if (limql < (0.6 * e->limq[9])) {
            double s = 1.0, old = 0.0, qq = 1.0;
            for (int n = 1; n <= 120; ++n) {
                qq = qq * (qxs + qys) * (qxs + qys);
                s = s + e->C.c1f[n] * qq;
                const double d = fabs((s - old) / s);
                if (d < 0.0001) break;
                old = s;
            }
            pql = s;
        } else {
            const double arglq = (qxs + qys + kEps9) * e->length / (zl + 1);
            pql = (1 / (2.0 * zl * (zl - 1))) * (1 / (arglq * arglq)) *
                  (1 - cos((zl - 1) * atan(2.0 * arglq)) / pow(1.0 + 4 * arglq * arglq, (zl - 1) / 2.0));
        }
    }
    if (q < (1.0 * e->limq[12])) {
        double s = 1.0, old = 0.0, qq = 1.0;
        for (int n = 1; n <= 120; ++n) {
            qq = qq * q * q;
            s += e->C.c4f[n] * qq;
            const double d = fabs((s - old) / s);
            if (d < 0.0001) break;
            old = s;
        }
        pqr = s;
    } else {
        const double argpq = q * e->radius / (zr + 1);
        const double pqr1 = (csxo_gamma(zr - 1 / 2.0) / csxo_gamma(zr + 1)) * pow(argpq, -3 / 2.0) *
                            (sin((zr - 1 / 2.0) * atan(argpq)) - cos((zr - 1 / 2.0) * atan(argpq))) /
                            pow(1.0 + argpq * argpq, (zr - 1 / 2.0) / 2.0);
        const double pqr2 = (csxo_gamma(zr - 3 / 2.0) / csxo_gamma(zr + 1)) * pow(argpq, -5 / 2.0) *
                            (sin((zr - 3 / 2.0) * atan(argpq)) + cos((zr - 3 / 2.0) * atan(argpq))) /
                            pow(1.0 + argpq * argpq, (zr - 3 / 2.0) / 2.0);
        const double pqr3 = (2 / sqrt(M_PI)) * (pqr1 + (9 / 16.0) * pqr2);
        pqr = pqr3 * pqr3;
    }
    return pql * pqr;
}

/* myacos macro, Source/sc_calc_generic_gpu.cu:31 -- QUIRK: out-of-range arguments
 * map to 0 for x<-1 and to pi for x>1. */
static inline double ref_acos(double x) { return (x < -1) ? 0.0 : ((x > +1) ? M_PI : acos(x)); }

/* Lattice factor of calc_GENERIC for the other peak shapes (Source/sc_gpu_generic.h:743-880):
 * the isotropic Gaussian sum and the anisotropic sum in the generic kernel's order (a
 * reflection with |q_hkl| <= 0 or NaN is skipped, the term is sq*mhkl*fhkl*psiord, a sum that
 * turns inf or NaN keeps its previous value and ends the loop).  Twinning is not served. */
static void lattice_sums_shape(const csxo_engine* e, double qx, double qy, double qz, double q, double width,
                               double* szq_factor_I, double* szqiso)
{
    const double widthiso = 1.0 / e->uca;
    const double sphno_cubevol = e->lat_ucvol * e->lat_sphno;
    double radint = 0.0, inten = 0.0;
    for (int ii = 1; ii <= e->peakmax1; ++ii) {
        const int mhkl = e->latpar1[ii][4], fhkl = e->latpar1[ii][5];
        const double qhkl = e->latpar3[ii][5];
        const double x2 = sq(q - qhkl) / sq(widthiso);
        const double s = exp(-4 * x2 / M_PI) / (M_PI * widthiso / 2.0);
        const double sav = radint;
        radint += s * mhkl * fhkl;
        if (isinf(radint) || isnan(radint)) { radint = sav; break; }
    }
    for (int ii = 1; ii <= e->peakmax2; ++ii) {
        const double qhkl = e->latpar3[ii][5];
        if (qhkl <= 0 || isnan(qhkl)) continue;
        const int mhkl = e->latpar2[ii][4], fhkl = e->latpar2[ii][5];
        const double qxh = e->latpar3[ii][2], qyh = e->latpar3[ii][3], qzh = e->latpar3[ii][4];
        const double g3 = e->latpar3[ii][10], pn1 = e->latpar3[ii][11], pn2 = e->latpar3[ii][12];
        const double yphi = ref_acos((qx * qxh + qy * qyh + qz * qzh) / (q * qhkl));
        const double x2 = (q - qhkl) * (q - qhkl) / (width * width);
        const double x2phi = 4 * q * q / sq(e->phiwidth); /* the Lorentzian shapes' b/c factor */
        double s = 0, psi = 1;
        switch (e->shp) {
        case 1: /* Lorentzian */
            s = sqrt(e->c1) / (M_PI * width * (1.0 + e->c1 * x2));
            psi = g3 / (pn1 * (1 + x2phi * yphi * yphi));
            break;
        case 3: /* mod. 1 Lorentzian */
            s = 2 * sqrt(e->c2) / (M_PI * width * (1 + e->c1 * x2) * (1 + e->c1 * x2));
            psi = g3 / (pn1 * (1 + x2phi * yphi * yphi) * (1 + x2phi * yphi * yphi));
            break;
        case 4: /* mod. 2 Lorentzian */
            s = sqrt(e->c3) / (2 * width * exp(3 * log(1 + e->c1 * x2) / 2.0));
            psi = g3 / (pn1 * exp(3 * log(1 + x2phi * yphi * yphi) / 2.0));
            break;
        case 5: { /* Pseudo-Voigt */
            s = e->eta * sqrt(e->c1) / (M_PI * width * (1 + e->c1 * x2)) +
                (1 - e->eta) * sqrt(e->c0) * exp(-e->c0 * x2) / (sqrt(M_PI) * width);
            const double x2psi = 4 * q * q / (M_PI * sq(e->phiwidth));
            psi = g3 * (e->eta * (1 / (1 + x2phi * yphi * yphi)) + (1 - e->eta) * exp(-x2psi * yphi * yphi)) /
                  (e->eta * pn1 + (1 - e->eta) * pn2);
            break;
        }
        case 6: /* Pearson VII */
            s = csxo_gamma(e->beta) * e->c4 * 2 /
                (csxo_gamma(e->beta - 0.5) * M_PI * width * exp(e->beta * log(1 + 4 * e->c4 * x2)));
            psi = g3 / (pn1 * exp(e->beta * log(1 + x2phi * yphi * yphi)));
            break;
        default:
            break;
        }
        const double sav = inten;
        inten += s * mhkl * fhkl * psi;
        if (isinf(inten) || isnan(inten)) { inten = sav; break; }
    }
    *szqiso = (1 + (1) * (8 * M_PI * M_PI * M_PI * radint / (4 * M_PI * sphno_cubevol) - 1) * exp(-e->ucb * q * q));
    *szq_factor_I = 2 * M_PI * inten / sphno_cubevol;
}

/* Lattice factor of the Gaussian-peak kernels: isotropic and anisotropic peak sums
 * (Source/sc_gpu_pSph_pGau_lFCC_oGau.h:52-146; identical loops in
 * Source/sc_gpu_pCyl_pGau_lBCT_oZDi.h:100-217).  Returns szq and szqiso. */
static void lattice_sums(const csxo_engine* e, double qx, double qy, double qz, double q,
                         double* szq_factor_I, double* szqiso)
{
    double width = 1;
    if (e->dsch) width = 4.0 / e->domainsize;
    if (e->para) width = (4.0 / e->domainsize) + sq(e->reldis) * e->dist * sq(q);
    width *= e->width_zuf;
    const double widthiso = 1.0 / e->uca;
    const double sphno_cubevol = e->lat_ucvol * e->lat_sphno;
    double radint = 0.0, inten = 0.0;
    if (e->shp != 2) {
        lattice_sums_shape(e, qx, qy, qz, q, width, szq_factor_I, szqiso);
        return;
    }
    for (int ii = 1; ii <= e->peakmax1; ++ii) {
        const int mhkl = e->latpar1[ii][4], fhkl = e->latpar1[ii][5];
        if (mhkl == 0 || fhkl == 0) continue;
        const double qhkl = e->latpar3[ii][5]; /* QUIRK: aniso table column (:25765) */
        const double x2 = sq(q - qhkl) / sq(widthiso);
        const double s = exp(-4 * x2 / M_PI) / (M_PI * widthiso / 2.0);
        radint += s * mhkl * fhkl;
    }
    for (int ii = 1; ii <= e->peakmax2; ++ii) {
        const double qhkl = e->latpar3[ii][5];
        const int mf = e->latpar2[ii][4] * e->latpar2[ii][5];
        if (!(qhkl > 0 && mf != 0)) continue;
        const double qxh = e->latpar3[ii][2], qyh = e->latpar3[ii][3], qzh = e->latpar3[ii][4];
        const double g3 = e->latpar3[ii][10], pn1 = e->latpar3[ii][11];
        const double yphi = ref_acos((qx * qxh + qy * qyh + qz * qzh) / (q * qhkl));
        const double x2 = (q - qhkl) * (q - qhkl) / (width * width);
        const double s = exp(-4 * x2 / M_PI) / (M_PI * width / 2.0);
        const double x2phi = 4 * q * q / (M_PI * sq(e->phiwidth));
        double psi = g3 * exp(-x2phi * yphi * yphi) / pn1;
        if (e->twin) {
            const double qxt = e->latpar3[ii][7], qyt = e->latpar3[ii][8], qzt = e->latpar3[ii][9];
            const double qht = e->latpar3[ii][14], g3t = e->latpar3[ii][15], pn1t = e->latpar3[ii][16];
            const double yt = ref_acos((qx * qxt + qy * qyt + qz * qzt) / (q * qht));
            const double phiord = g3t * exp(-x2phi * yt * yt) / pn1t;
            psi = e->ceff * psi + (1 - e->ceff) * phiord;
        }
        const double sav = inten;
        inten += s * mf * psi;
        if (isinf(inten)) { inten = sav; break; }
    }
    *szqiso = (1 + (1) * (8 * M_PI * M_PI * M_PI * radint / (4 * M_PI * sphno_cubevol) - 1) * exp(-e->ucb * q * q));
    *szq_factor_I = 2 * M_PI * inten / sphno_cubevol;
}

/* Spheres on a lattice: P(q), <F>^2 and the lattice factor (Source/sc_gpu_pSph_pGau_lFCC_oGau.h:
 * 15-171; calc_GENERIC's sphere and peak branches, Source/sc_gpu_generic.h:40-52, 590-880). */
static double sphere_lattice_pixel(const csxo_engine* e, double qx, double qy, double qz, double q)
{
    const double pq = pq_sphere(e, q);
    const double fq = fq_sphere(e, q);
    double I, szqiso;
    lattice_sums(e, qx, qy, qz, q, &I, &szqiso);
    const double szq = (1 + (fq / pq) * (I - 1) * exp(-e->dwfactor * q * q));
    return e->base + e->i0 * (szq * pq + e->iso * szqiso * pq) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
}

/* The 2D lattices (hexagonal, square, centred rectangular) and the 3D ones served: the cubic
 * lattices, HCP, BCT, orthorhombic. */
static int served_lattice(unsigned lt)
{
    int cc[4];
    return (lt >= 1 && lt <= 3) || cubic_codes((int)lt, cc) || lt == 6 || lt == 8 || lt == 18;
}

/* Spheres on a lattice with the Lorentzian, Gaussian, mod. 1/2 Lorentzian, Pseudo-Voigt or
 * Pearson VII peak (calc_GENERIC; the FCC Gaussian key has its own kernel with the same sums):
 * neither the sphere's factors nor these peak sums depend on the orientation distribution. */
static int sphere_lattice_shape_key(unsigned key)
{
    const unsigned part = key >> 24, shp = key >> 16 & 0xff, lt = key >> 8 & 0xff, ord = key & 0xff;
    return part == 0 && shp >= 1 && shp <= 6 && served_lattice(lt) && ord <= 13;
}

/* Perfectly oriented (ordis 6) or isotropic (7) cylinders on a lattice with peak shapes 1-6:
 * the BCT z-dir Gaussian key's own kernel (Source/sc_gpu_pCyl_pGau_lBCT_oZDi.h) and
 * calc_GENERIC's cylinder branches (Source/sc_gpu_generic.h:62-121) with the same sums. */
static int cylinder_lattice_key(unsigned key)
{
    const unsigned part = key >> 24, shp = key >> 16 & 0xff, lt = key >> 8 & 0xff, ord = key & 0xff;
    return part == 1 && shp >= 1 && shp <= 6 && served_lattice(lt) && (ord == 6 || ord == 7);
}

/* Isotropic homogeneous cubes on a lattice with peak shapes 1-6 (calc_GENERIC's cube branch
 * sets fq = pq). */
static int cube_lattice_key(unsigned key)
{
    const unsigned part = key >> 24, shp = key >> 16 & 0xff, lt = key >> 8 & 0xff, ord = key & 0xff;
    return part == 4 && shp >= 1 && shp <= 6 && served_lattice(lt) && ord == 7;
}

int csxo_pixel(const csxo_engine* e, double qx, double qy, double qz, double* out)
{
    const double q = sqrt(qx * qx + qy * qy + qz * qz) + kEps9;
    const unsigned key = csxo_combo_key(e);
    if (sphere_lattice_shape_key(key)) {
        if (e->twin) return -1;
        *out = sphere_lattice_pixel(e, qx, qy, qz, q);
        return 0;
    }
    switch (key) {
    case 0x00000C00u: /* spheres without lattice, Gaussian or z-oriented distribution: the generic */
    case 0x00000C06u: /* kernel's sphere branch (Source/sc_gpu_generic.h:40-52, 868-880), the same value */
    case 0x00000C07u: { /* Source/sc_gpu_pSph_lNon_oIso.h:15-51 */
        const double pq = pq_sphere(e, q);
        *out = e->base + e->i0 * (pq + e->iso * pq) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    case 0x04000C06u: { /* calc_GENERIC, cube, perfect orientation (Source/sc_gpu_generic.h:307-314, 868-880) */
        const double pq = pq_cube_zdir(e, qx, qy, q);
        *out = e->base + e->i0 * (pq + e->iso) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    case 0x04000C07u: { /* Source/sc_gpu_pCub_lNon_oIso.h:14-57 */
        const double pq = (e->cs == 0) ? pq_cube(e, q) : 1.0;
        *out = e->base + e->i0 * (pq + e->iso) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    case 0x01000C00u: { /* Source/sc_gpu_pCyl_lNon_oGau.h:15-100 (fq unused without lattice) */
        double pq = 0;
        switch (e->orcase) {
        case ORC_GENERAL: pq = pq_cylinder(e, qx * e->cosphic - qy * e->sinphic, qx * e->sinphic + qy * e->cosphic, q); break;
        default: pq = pq_cylinder(e, qx, qy, q); break;
        }
        *out = e->base + e->i0 * (pq + e->iso) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    case 0x01000C07u: { /* calc_GENERIC, cylinder, isotropic (Source/sc_gpu_generic.h:58-66, 868-880) */
        const double pq = pq_cylinder(e, qx, qy, q);
        *out = e->base + e->i0 * (pq + e->iso) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    case 0x01000C06u: { /* Source/sc_gpu_pCyl_lNon_oZDi.h:17-60 */
        double pq = 0;
        switch (e->orcase) {
        case ORC_GENERAL: pq = pq_cylinder(e, qx * e->cosphi * e->sintheta, qy * e->sinphi * e->sintheta, q); break;
        case ORC_X: pq = pq_cylinder(e, qx, 0, q); break;
        case ORC_Y: pq = pq_cylinder(e, 0, qy, q); break;
        default: pq = pq_cylinder(e, qx, qy, q); break;
        }
        *out = e->base + e->i0 * (pq + e->iso) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    default: break;
    }
    if (cube_lattice_key(key)) { /* Source/sc_gpu_generic.h:290-302, 536-880: fq = pq */
        if (e->cs != 0) return -1;
        const double pq = pq_cube(e, q);
        double I, szqiso;
        lattice_sums(e, qx, qy, qz, q, &I, &szqiso);
        const double szq = (1 + (pq / pq) * (I - 1) * exp(-e->dwfactor * q * q));
        *out = e->base + e->i0 * (szq * pq + e->iso * szqiso * 1.0) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    if (cylinder_lattice_key(key) && e->ordis == 7) { /* Source/sc_gpu_generic.h:64-74, 536-880 */
        const double pq = pq_cylinder(e, qx, qy, q);
        const double fq = fq_cylinder(e, q, qx, qy, q);
        double I, szqiso;
        lattice_sums(e, qx, qy, qz, q, &I, &szqiso);
        const double szq = (1 + (fq / pq) * (I - 1) * exp(-e->dwfactor * q * q));
        *out = e->base + e->i0 * (szq * pq + e->iso * szqiso * 1.0) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    if (cylinder_lattice_key(key)) { /* Source/sc_gpu_pCyl_pGau_lBCT_oZDi.h:16-245 */
        double pq = 0, fq = 0;
        switch (e->orcase) {
        case ORC_GENERAL:
            pq = pq_cylinder(e, qx * e->cosphi * e->sintheta, qy * e->sinphi * e->sintheta, q);
            fq = fq_cylinder(e, sqrt(e->cosphi * qx * qx + e->sinphi * qy * qy + kEps9),
                             qx * e->cosphi * e->sintheta, qy * e->sinphi * e->sintheta, q);
            break;
        case ORC_X: pq = pq_cylinder(e, qx, 0, q); fq = fq_cylinder(e, fabs(qx), qx, 0, q); break;
        case ORC_Y: pq = pq_cylinder(e, 0, qy, q); fq = fq_cylinder(e, fabs(qy), 0, qy, q); break;
        default: pq = pq_cylinder(e, qx, qy, q); fq = fq_cylinder(e, q, qx, qy, q); break;
        }
        double I, szqiso;
        lattice_sums(e, qx, qy, qz, q, &I, &szqiso);
        const double szq = (1 + (fq / pq) * (I - 1) * exp(-e->dwfactor * q * q));
        *out = e->base + e->i0 * (szq * pq + e->iso * szqiso * 1.0) + e->ifluc / (1 + q * q * e->rfluc * e->rfluc);
        return 0;
    }
    return -1;
}

/* convert_idx2q (Source/sc_calc_generic_gpu.cu:249-287). */
void csxo_idx2q(const csxo_engine* e, int ihex, int i, double* q3)
{
    if (e->centerbeam) {
        const double mdet = (ihex) * e->pixnoy / (2.0 * e->zmax);
        const double ndet = (i) * e->pixnox / (2.0 * e->zmax);
        const double xdet = e->pixx_m * (ndet - e->beamx);
        const double ydet = e->pixy_m * (mdet - e->beamy);
        const double rdet = sqrt(xdet * xdet + ydet * ydet);
        const double phidet = atan2(ydet, xdet);
        const double thetadet = atan2(rdet, e->det);
        q3[0] = 2 * M_PI * cos(phidet) * sin(thetadet) / e->wave;
        q3[1] = 2 * M_PI * sin(phidet) * sin(thetadet) / e->wave;
        q3[2] = 2 * M_PI * (1 - cos(thetadet)) / e->wave;
    } else {
        q3[0] = e->qmax * i / (double)(e->zmax);
        q3[1] = e->qmax * ihex / (double)(e->zmax);
        q3[2] = 1e-20;
    }
}

/* ========================================================================== */
/* Drivers                                                                     */
/* ========================================================================== */

typedef struct {
    const csxo_engine* e;
    double* out;
    const int *ih, *ii;
    long n;
    int t, nt, rc;
} work_t;

static void* image_rows(void* arg)
{
    work_t* w = (work_t*)arg;
    const csxo_engine* e = w->e;
    const int W = e->xmax - e->xmin;
    for (int ihex = e->zzmin + w->t; ihex < e->zzmax; ihex += w->nt) {
        for (int i = e->iimin; i < e->iimax; ++i) {
            double q3[3], v;
            csxo_idx2q(e, ihex, i, q3);
            if (csxo_pixel(e, q3[0], q3[1], q3[2], &v) != 0) { w->rc = -1; return NULL; }
            w->out[(size_t)(ihex - e->xmin) + (size_t)W * (size_t)(i - e->ymin)] = v;
        }
    }
    return NULL;
}

static int run_threads(work_t* proto, void* (*fn)(void*), int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    work_t w[256];
    for (int t = 0; t < nthreads; ++t) { w[t] = *proto; w[t].t = t; w[t].nt = nthreads; w[t].rc = 0; }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, fn, &w[t]);
    fn(&w[0]);
    int rc = w[0].rc;
    for (int t = 1; t < nthreads; ++t) { pthread_join(th[t], NULL); if (w[t].rc) rc = w[t].rc; }
    return rc;
}

int csxo_image(const csxo_engine* e, double* out, int nthreads)
{
    if (!e->prepared) return -2;
    work_t p = {e, out, NULL, NULL, 0, 0, 1, 0};
    return run_threads(&p, image_rows, nthreads);
}

static void* image_points(void* arg)
{
    work_t* w = (work_t*)arg;
    for (long k = w->t; k < w->n; k += w->nt) {
        double q3[3], v;
        csxo_idx2q(w->e, w->ih[k], w->ii[k], q3);
        if (csxo_pixel(w->e, q3[0], q3[1], q3[2], &v) != 0) { w->rc = -1; return NULL; }
        w->out[k] = v;
    }
    return NULL;
}

int csxo_image_points(const csxo_engine* e, const int* ihex, const int* ii, long n, double* out, int nthreads)
{
    if (!e->prepared) return -2;
    work_t p = {e, out, ihex, ii, n, 0, 1, 0};
    return run_threads(&p, image_points, nthreads);
}

/* FQSVERGL (Source/sc_globalConfig.h:32). */
static inline double fqs_term(double a, double b) { return sq(log(a) / log(10.) - log(b) / log(10.)); }

typedef struct {
    const csxo_engine* e;
    const double* data;
    double* fqs;
    int w, h, bstop, border;
    const int* rects;
    int t, nt, rc;
} fitwork_t;

/* checkFitRanges (Source/sc_calc_generic_gpu.cu:293-347) + the fit branch of
 * KERNEL_POSTWORK / calcCPU_selection (.cu:64-73, :228-237). */
static void* fit_cols(void* arg)
{
    fitwork_t* w = (fitwork_t*)arg;
    const csxo_engine* e = w->e;
    const int W = e->xmax - e->xmin;
    for (int x = w->t; x < w->w; x += w->nt) {
        for (int y = 0; y < w->h; ++y) {
            int masked = 0;
            for (int r = 0; r < 4 && w->rects; ++r) {
                const int* R = w->rects + 4 * r;
                if (R[0] < 0) continue;
                if (x < R[0] || R[2] < x || y < R[1] || R[3] < y) continue;
                masked = 1;
                break;
            }
            if (!masked) {
                if (w->border > 0 || w->bstop > 0) {
                    if (x < e->xmin + w->border || x >= e->xmax - w->border ||
                        y < e->ymin + w->border || y >= e->ymax - w->border)
                        masked = 1;
                    else if (x >= e->beamx - w->bstop && x < e->beamx + w->bstop &&
                             y >= e->beamy - w->bstop && y < e->beamy + w->bstop)
                        masked = 1;
                } else if (w->bstop == -1) {
                    if (w->data[(size_t)x + (size_t)w->w * y] <= w->data[0]) masked = 1;
                }
            }
            if (masked) { w->fqs[(size_t)x + (size_t)w->w * y] = 0.0; continue; }
            const int ihex = x + e->iimin, i = y + e->iimin;
            double q3[3], v;
            csxo_idx2q(e, ihex, i, q3);
            if (csxo_pixel(e, q3[0], q3[1], q3[2], &v) != 0) { w->rc = -1; return NULL; }
            const size_t idx = (size_t)(ihex - e->xmin) + (size_t)W * (size_t)(i - e->ymin);
            w->fqs[idx] = (v > 0 && w->data[idx] > 0) ? fqs_term(v, w->data[idx]) : 0.0;
        }
    }
    return NULL;
}

double csxo_fit_fqs(const csxo_engine* e, const double* data, int w, int h, int bstop, int border,
                    const int* rects, double* fqs_out, long* cnt, long* nancnt, int nthreads)
{
    if (!e->prepared) return 1e20;
    double* fqs = fqs_out ? fqs_out : (double*)calloc((size_t)w * h, sizeof(double));
    fitwork_t p = {e, data, fqs, w, h, bstop, border, rects, 0, 1, 0};
    if (nthreads < 1) nthreads = 1;
    pthread_t th[256];
    fitwork_t ws[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; ++t) { ws[t] = p; ws[t].t = t; ws[t].nt = nthreads; }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, fit_cols, &ws[t]);
    fit_cols(&ws[0]);
    int rc = ws[0].rc;
    for (int t = 1; t < nthreads; ++t) { pthread_join(th[t], NULL); if (ws[t].rc) rc = ws[t].rc; }
    double sum = 0.0;
    long c = 0, nc = 0;
    if (rc == 0) {
        for (size_t i = 0; i < (size_t)w * h; ++i) {
            if (isnan(fqs[i])) nc++;
            else { if (fqs[i] != 0) c++; sum += fqs[i]; }
        }
    }
    if (cnt) *cnt = c;
    if (nancnt) *nancnt = nc;
    if (!fqs_out) free(fqs);
    return rc == 0 ? sum : 1e20;
}

/* ========================================================================== */
/* Table access                                                                */
/* ========================================================================== */

const void* csxo_table(const csxo_engine* e, const char* name, long* n, int* kind)
{
    *kind = 0; /* 0 double, 1 int, 2 float */
#define TAB(nm, ptr, cnt, k) if (strcmp(name, nm) == 0) { *n = (cnt); *kind = (k); return (ptr); }
    TAB("carr1p", e->C.c1p, NCOEF, 0);
    TAB("carr2p", e->C.c2p, NCOEF, 0);
    TAB("carr3p", e->C.c3p, NCOEF, 0);
    TAB("carr4p", e->C.c4p, NCOEF, 0);
    TAB("carr1f", e->C.c1f, NCOEF, 0);
    TAB("carr2f", e->C.c2f, NCOEF, 0);
    TAB("carr4f", e->C.c4f, NCOEF, 0);
    TAB("carr11pm", &e->C.c11pm[0][0], (long)NIM2 * NIM2, 0);
    TAB("limq", e->limq, 18, 0);
    TAB("latpar1", &e->latpar1[0][0], (long)(e->peakmax1 + 1) * 6, 1);
    TAB("latpar2", &e->latpar2[0][0], (long)(e->peakmax2 + 1) * 6, 1);
    TAB("latpar3", &e->latpar3[0][0], (long)(e->peakmax2 + 1) * 17, 2);
    TAB("ri", e->ri, 9, 0);
#undef TAB
    *n = 0;
    return NULL;
}

double csxo_scalar(const csxo_engine* e, const char* name)
{
#define SC(nm, v) if (strcmp(name, nm) == 0) return (double)(v);
    SC("norm", e->norm); SC("order", e->order); SC("orcase", e->orcase); SC("ucvol", e->ucvol);
    SC("peakmax1", e->peakmax1); SC("peakmax2", e->peakmax2); SC("dwfactor", e->dwfactor);
    SC("phiwidth", e->phiwidth); SC("sinphic", e->sinphic); SC("cosphic", e->cosphic);
    SC("polPhi", e->polPhi); SC("sphno", e->lat_sphno); SC("lat_ucvol", e->lat_ucvol);
    SC("reldis", e->reldis); SC("dist", e->dist); SC("zmax", e->zmax); SC("por", e->por);
#undef SC
    return NAN;
}

/* ========================================================================== */
/* Nelder-Mead (doSimplexFit2D)                                               */
/* ========================================================================== */

typedef struct { csxo_objective f; void* user; int numfit; const double* minv; const double* maxv; long nev; } nm_ctx;

/* Bounds rule of fehlerquadratsumme (Source/sc_simplexfit2d.cpp:25-44). */
static double nm_eval(nm_ctx* c, const double* x)
{
    for (int j = 0; j < c->numfit; ++j) {
        if (c->minv[j] == 0.0 && c->maxv[j] == 360.0) continue;
        if (x[j] < c->minv[j]) return 1e20;
        if (x[j] > c->maxv[j]) return 1e20;
    }
    c->nev++;
    return c->f(c->user, x);
}

#define NAP 31
#define MAP 30

/* amotry (Source/sc_simplexfit2d.cpp:221-251). */
static double nm_try(nm_ctx* c, double p[NAP][MAP], double* y, double* sum, int mmax, int ihi, double fac)
{
    double pt[MAP];
    for (int j = 0; j < mmax; ++j) pt[j] = sum[j];
    const double f1 = (1.0 - fac) / mmax, f2 = f1 - fac;
    for (int j = 0; j < c->numfit; ++j) pt[j] = sum[j] * f1 - p[ihi][j] * f2;
    const double yt = nm_eval(c, pt);
    if (yt < y[ihi]) {
        y[ihi] = yt;
        for (int j = 0; j < c->numfit; ++j) { sum[j] += pt[j] - p[ihi][j]; p[ihi][j] = pt[j]; }
    }
    return yt;
}

int csxo_simplex(csxo_objective f, void* user, int numfit, int mmax, double* x, const double* minv,
                 const double* maxv, double stp, int maxit, double ftol, int* iters_out,
                 double* fbest_out, long* nevals_out)
{
    if (mmax >= NAP || numfit > mmax) return -1;
    nm_ctx c = {f, user, numfit, minv, maxv, 0};
    const double alfa = 1.0, beta = 0.5, gamma = 2.0;
    double y[NAP], pp[MAP], qv[MAP], step[MAP], psum[MAP], p[NAP][MAP];
    memset(p, 0, sizeof(p));
    memset(y, 0, sizeof(y));
    const double nmax = mmax + 1.0;
    const int ndim = mmax, mpts = numfit + 1;
    const double rootn1 = sqrt(nmax) - 1.0, mroot2 = 1.0 / (mmax * sqrt(2.0));
    double vals[MAP];
    for (int i = 0; i < mmax; ++i) {
        vals[i] = x[i];
        psum[i] = x[i];
        p[0][i] = x[i];
        if (i < numfit) {
            if (x[i] == 0) step[i] = (minv[i] == 0) ? stp : minv[i] * stp;
            else step[i] = x[i] * stp;
        } else step[i] = 0.0;
        pp[i] = step[i] * (rootn1 + mmax) * mroot2;
        qv[i] = step[i] * rootn1 * mroot2;
    }
    /* The clearing loop of :384-389 advances i in its inner loop: it zeroes p[i][0] for
     * i = 0..map-1, vertex 0's first entry included, after vertex 0 was set.  y[0] is
     * evaluated at the start values (:392) and the other vertices are built from the zeroed
     * vertex 0 (:411-417). */
    for (int i = 0; i < MAP; ++i) p[i][0] = 0.0;
    y[0] = nm_eval(&c, vals);
    for (int i = 1; i < nmax; ++i) {
        for (int j = 0; j < mmax; ++j) p[i][j] = p[0][j] + qv[j];
        p[i][i - 1] = p[i][i - 1] + pp[i - 1];
    }
    for (int i = 1; i < nmax; ++i) {
        for (int j = 0; j < numfit; ++j) vals[j] = p[i][j];
        if (i > numfit + 1) y[i] = y[numfit + 1];
        else y[i] = nm_eval(&c, vals);
    }
    for (int j = 0; j < ndim; ++j) {
        double s = 0.0;
        for (int i = 0; i < mpts; ++i) s += p[i][j];
        psum[j] = s;
    }
    int ilo = 0, ihi, inhi, rep;
    for (rep = 0; rep < maxit; ++rep) {
        ilo = 0;
        if (y[0] > y[1]) { ihi = 0; inhi = 1; } else { ihi = 1; inhi = 0; }
        for (int i = 0; i < mpts; ++i) {
            if (y[i] < y[ilo]) ilo = i;
            if (y[i] > y[ihi]) { inhi = ihi; ihi = i; }
            else if (y[i] > y[inhi]) { if (i != ihi) inhi = i; }
        }
        if (isnan(y[ihi]) || isnan(y[ilo]) || fabs(y[ihi]) + fabs(y[ilo]) == 0) break;
        const double rtol = 2.0 * fabs(y[ihi] - y[ilo]) / (fabs(y[ihi]) + fabs(y[ilo]));
        if (rtol <= ftol) break;
        double yt = nm_try(&c, p, y, psum, mmax, ihi, -alfa);
        if (yt <= y[ilo]) {
            nm_try(&c, p, y, psum, mmax, ihi, gamma);
        } else if (yt >= y[inhi]) {
            const double ysave = y[ihi];
            yt = nm_try(&c, p, y, psum, mmax, ihi, beta);
            if (yt >= ysave) {
                for (int i = 0; i < mpts; ++i)
                    if (i != ilo) {
                        for (int j = 0; j < ndim; ++j) { psum[j] = 0.5 * (p[i][j] + p[ilo][j]); p[i][j] = psum[j]; }
                        y[i] = nm_eval(&c, psum);
                    }
                for (int j = 0; j < ndim; ++j) {
                    double s = 0.0;
                    for (int i = 0; i < mpts; ++i) s += p[i][j];
                    psum[j] = s;
                }
            }
        }
    }
    for (int i = 0; i < mmax; ++i) x[i] = p[ilo][i];
    if (iters_out) *iters_out = rep;
    if (fbest_out) *fbest_out = y[ilo];
    if (nevals_out) *nevals_out = c.nev;
    return 0;
}


/* ========================================================================== */
/* AI-data post-processing                                                     */
/* ========================================================================== */

void csxo_scale_intensity(double* p, long n, int linlog)
{
    double vmin = 1e20, vmax = 1e-20;
    if (linlog) {
        for (long i = 0; i < n; ++i)
            if (p[i] < vmin) vmin = p[i];
            else if (p[i] > vmax) vmax = p[i];
    } else {
        for (long i = 0; i < n; ++i)
            if (p[i] > 0) {
                if (log10(p[i]) < vmin) vmin = log10(p[i]);
                else if (log10(p[i]) > vmax) vmax = log10(p[i]);
            }
    }
    if (fabs(vmin - vmax) < 1e-6) {
        vmin *= 0.998;
        vmax *= 1.002;
        if (vmax == 0) vmax = 0.2;
    }
    for (long i = 0; i < n; ++i) {
        if (linlog) p[i] = (p[i] - vmin) / (vmax - vmin);
        else p[i] = (p[i] > 0) ? (log10(p[i]) - vmin) / (vmax - vmin) : 0;
    }
}

void csxo_scale_clip(double* data, long len, int scale, int clip, int clip40, int genlog)
{
    if (genlog) {
        double lo = 1e26, hi = 0;
        for (long i = 0; i < len; ++i)
            if (data[i] > 0) {
                if (log10(data[i]) < lo) lo = log10(data[i]);
                else if (log10(data[i]) > hi) hi = log10(data[i]);
            }
        for (long i = 0; i < len; ++i)
            data[i] = (data[i] > 0) ? (log10(data[i]) - lo) / (hi - lo) : 0;
    }
    if (scale) {
        double mx = 0;
        for (long i = 0; i < len; ++i)
            if (mx < data[i]) mx = data[i];
        for (long i = 0; i < len; ++i) data[i] = data[i] / mx;
    }
    if (clip)
        for (long i = 0; i < len; ++i) {
            data[i] = ((log(fabs(data[i]) / 1e-20) / log(10)) + 6.0) / 6.0;
            if (data[i] < 0) data[i] = 0;
        }
    if (clip40)
        for (long i = 0; i < len; ++i) {
            data[i] = 1.0 - (1.0 - data[i]) / (1.0 - 0.4);
            if (data[i] < 0) data[i] = 0;
        }
}

static double dmax(double a, double b) { return (a < b) ? b : a; } /* qMax */

void csxo_rphi(const double* d, int x0, int x1, int y0, int y1, int xc, int yc, int s, double* out,
               int scale, int clip, int clip40)
{
    if (x0 < 0) { x1 -= x0; xc -= x0; x0 = 0; }
    if (y0 < 0) { y1 -= y0; yc -= y0; y0 = 0; }
    const int sx = x1 - x0, sy = y1 - y0;
    const long len = (long)sx * sy, arrlen = (long)s * s;
    const int phisteps = s;
    const double phiber = 2. * M_PI / phisteps;
    double rmax = sqrt(sq(1. - xc) + sq(1. - yc));
    rmax = dmax(rmax, sqrt(sq((sx - 1.) - xc) + sq(1. - yc)));
    rmax = dmax(rmax, sqrt(sq((sx - 1.) - xc) + sq((sy - 1.) - yc)));
    rmax = dmax(rmax, sqrt(sq(1. - xc) + sq((sy - 1.) - yc)));
    const int irmax = (int)(rmax + 0.5);
    const float irstp = (float)irmax / (float)(s - 1);
    for (float ir = 0; ir <= irmax; ir += irstp) {
        const double rad = ir;
        for (int i = 0; i <= phisteps; ++i) {
            double twt = 0.0, avg = 0.0;
            int umf = 1024 - (int)(1024. - phiber * rad);
            if (umf < 1) umf = 1;
            for (int j = 1; j <= umf; ++j) {
                const double phi = (i - 1.5) * phiber + (j - 0.5) * phiber / umf;
                const double xs = xc + rad * cos(phi), ys = yc + rad * sin(phi);
                const int ix1 = (int)xs, iy1 = (int)ys, ix2 = ix1 + 1, iy2 = iy1 + 1;
                const double rex = xs - ix1, rey = ys - iy1;
                /* the four corners in the reference's order; data(x,y) = d[x + sx*y] */
                if (ix1 >= 0 && ix1 < sx && iy1 >= 0 && iy1 < sy) {
                    const double wt = 1 * (1.0 - rex) * (1.0 - rey);
                    twt = twt + wt; avg = avg + wt * d[ix1 + (long)sx * iy1];
                }
                if (ix2 >= 0 && ix2 < sx && iy1 >= 0 && iy1 < sy) {
                    const double wt = 1 * (rex) * (1.0 - rey);
                    twt = twt + wt; avg = avg + wt * d[ix2 + (long)sx * iy1];
                }
                if (ix1 >= 0 && ix1 < sx && iy2 >= 0 && iy2 < sy) {
                    const double wt = 1 * (1.0 - rex) * (rey);
                    twt = twt + wt; avg = avg + wt * d[ix1 + (long)sx * iy2];
                }
                if (ix2 >= 0 && ix2 < sx && iy2 >= 0 && iy2 < sy) {
                    const double wt = 1 * (rex) * (rey);
                    twt = twt + wt; avg = avg + wt * d[ix2 + (long)sx * iy2];
                }
            }
            if (twt > 0.0) avg = avg / twt;
            const int idx = (int)((float)(i * s) + ir / irstp);
            if (idx < arrlen) out[idx] = (avg < 0) ? 0 : avg;
        }
    }
    (void)len;
    csxo_scale_clip(out, arrlen, scale, clip, clip40, 0);
}


/* The radix-2 FFT the reference carries for builds without FFTW (Source/sc_postproc.cpp:833-1000,
 * from P. Bourke's published DFT/FFT notes): bit reversal, then m butterfly levels whose
 * twiddles follow the recurrence u <- u * (c1 + i c2) with c1, c2 halved-angle square roots. */
static void fft1d(int dir, int m, double* x, double* y)
{
    long nn = 1;
    for (int i = 0; i < m; ++i) nn *= 2;
    long i2 = nn >> 1, j = 0;
    for (long i = 0; i < nn - 1; ++i) {
        if (i < j) {
            double tx = x[i], ty = y[i];
            x[i] = x[j]; y[i] = y[j];
            x[j] = tx; y[j] = ty;
        }
        long k = i2;
        while (k <= j) { j -= k; k >>= 1; }
        j += k;
    }
    double c1 = -1.0, c2 = 0.0;
    long l2 = 1;
    for (int l = 0; l < m; ++l) {
        const long l1 = l2;
        l2 <<= 1;
        double u1 = 1.0, u2 = 0.0;
        for (long jj = 0; jj < l1; ++jj) {
            for (long i = jj; i < nn; i += l2) {
                const long i1 = i + l1;
                const double t1 = u1 * x[i1] - u2 * y[i1];
                const double t2 = u1 * y[i1] + u2 * x[i1];
                x[i1] = x[i] - t1;
                y[i1] = y[i] - t2;
                x[i] += t1;
                y[i] += t2;
            }
            const double z = u1 * c1 - u2 * c2;
            u2 = u1 * c2 + u2 * c1;
            u1 = z;
        }
        c2 = sqrt((1.0 - c1) / 2.0);
        if (dir == 1) c2 = -c2;
        c1 = sqrt((1.0 + c1) / 2.0);
    }
    if (dir == 1)
        for (long i = 0; i < nn; ++i) { x[i] /= (double)nn; y[i] /= (double)nn; }
}

/* calculateIndex (:729-752): 0 at the centre when swap is set. */
static long fft_index(int swap, int x, int y, int dim)
{
    if (swap) {
        if (y < dim / 2 && x < dim / 2) return (dim / 2 - x) + (long)dim * (dim / 2 - y);
        if (y >= dim / 2 && x < dim / 2) return (dim / 2 - x) + (long)dim * (dim - 1 - y + dim / 2);
        if (y < dim / 2 && x >= dim / 2) return (dim - 1 - x + dim / 2) + (long)dim * (dim / 2 - y);
        return (dim - 1 - x + dim / 2) + (long)dim * (dim - 1 - y + dim / 2);
    }
    return x + (long)dim * y;
}

long csxo_ifft(const double* d, int s, int forward, int out_type, int scaled, int clip, int clip40, int swap,
               int cutx, int cuty, double* out)
{
    const long len = (long)s * s;
    double mx = 0;
    for (long i = 0; i < len; ++i) if (d[i] > mx) mx = d[i];
    if (mx == 0) return 0;
    int m = 0;
    while ((1L << m) < s) ++m;
    if ((1L << m) != s || s < 4) return 0;
    /* c[first][second]: first = y, second = x of the input (_complex[yc][xc]) */
    double* re = (double*)malloc(sizeof(double) * len);
    double* im = (double*)malloc(sizeof(double) * len);
    double* lr = (double*)malloc(sizeof(double) * s);
    double* li = (double*)malloc(sizeof(double) * s);
    for (int y = 0; y < s; ++y)
        for (int x = 0; x < s; ++x) { re[(long)y * s + x] = d[x + (long)s * y] / mx; im[(long)y * s + x] = 0.0; }
    const int dir = forward ? 1 : -1;
    for (int j = 0; j < s; ++j) { /* FFT2D's "rows": along the first index for each second index */
        for (int i = 0; i < s; ++i) { lr[i] = re[(long)i * s + j]; li[i] = im[(long)i * s + j]; }
        fft1d(dir, m, lr, li);
        for (int i = 0; i < s; ++i) { re[(long)i * s + j] = lr[i]; im[(long)i * s + j] = li[i]; }
    }
    for (int i = 0; i < s; ++i) { /* then along the second index for each first index */
        for (int j = 0; j < s; ++j) { lr[j] = re[(long)i * s + j]; li[j] = im[(long)i * s + j]; }
        fft1d(dir, m, lr, li);
        for (int j = 0; j < s; ++j) { re[(long)i * s + j] = lr[j]; im[(long)i * s + j] = li[j]; }
    }
    double* full = (double*)calloc((size_t)len, sizeof(double)); /* the swap leaves row and column 0 unwritten */
    for (int y = 0; y < s; ++y)
        for (int x = 0; x < s; ++x) {
            const double a = re[(long)x * s + y], b = im[(long)x * s + y]; /* _complex[x][y] */
            double v;
            switch (out_type) {
            case 0: v = a; break;
            case 1: v = b; break;
            case 2: v = sqrt(sq(a) + sq(b)); break;
            default: v = (a + b) * (a - b); break;
            }
            full[fft_index(swap, x, y, s)] = v;
        }
    csxo_scale_clip(full, len, scaled, clip, clip40, 0);
    long n = len;
    if (cutx > 0 && cuty > 0 && cutx < s && cuty < s) {
        const int x0 = s / 2 - cutx / 2, y0 = s / 2 - cuty / 2;
        for (int yo = 0; yo < cuty; ++yo)
            for (int xo = 0; xo < cutx; ++xo) out[xo + (long)cutx * yo] = full[(x0 + xo) + (long)s * (y0 + yo)];
        n = (long)cutx * cuty;
    } else {
        memcpy(out, full, sizeof(double) * len);
    }
    free(re); free(im); free(lr); free(li); free(full);
    return n;
}
