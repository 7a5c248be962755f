// Per-pixel physics of the specialised kernels and of the generic kernel's combinations served
// here, FP64, reference formulas.
//
// Device functions (HIP, gfx950).  Compiled with -ffp-contract=off so the FP64
// arithmetic rounds operation by operation like the reference's CPU build;
// transcendental functions are the device math library's (within an ulp of
// glibc), so REF64 results agree with the CPU oracle to ~1e-15 relative.
//
// Reference functions restated (paths relative to the reference checkout):
//   pq_sphere     formpq_partSphere cs=0     Source/sc_lib_formpq_partSphere.h:11-75
//   pq_cube       formpq_partCube ordis 7    Source/sc_lib_formpq_partCube.h:27-61
//   fq_sphere     formfq_partSphere cs=0     Source/sc_lib_formfq_partSphere.h:8-52
//   pq_cylinder   formpq_partCylinder        Source/sc_lib_formpq_partCylinder.h:8-257
//   fq_cylinder   formfq_partCylinder ord 6  Source/sc_lib_formfq_partCylinder.h:8-206
//   (Gamma values of the cylinder <F>^2 come from the host: csx::ref_gamma, same formula)
//   pixel_q       convert_idx2q              Source/sc_calc_generic_gpu.cu:249-287
#pragma once

#include <hip/hip_runtime.h>

#include "csx.h"

namespace csx {

constexpr double kPiD = 3.14159265358979323846;

// Device copy of one parameter set: scalars by value, tables in HBM.
// One Friedel pair (+hkl, -hkl) of the anisotropic list: unit direction u of q_hkl,
// |q_hkl|, u_z^2, the common weight A = mhkl*fhkl*g3/peaknorm1 and both latpar indices.
struct alignas(32) PairRec {
    float ux, uy, uz, qh;
    float uz2, A;
    int32_t ia, ib;
};
// The same pair in FP64 for the precise pair loop: u = q_hkl / |q_hkl| with the stored
// |q_hkl| (the reference's cosine divides by it), and the weight A in FP64.
struct alignas(16) PairRec64 {
    double ux, uy, uz, A;
    float qh;
    int32_t ia, ib;
    float pad;
};
constexpr int kBands = 8;
// One more segment follows the merged one: the pairs with |u_z| <= kPlaneUz.  For q in the
// detector plane (q_z = 0), |q^.u| <= sqrt(1 - u_z^2) plus rounding, so only those pairs can
// come near |q^.u| = 1.
constexpr float kPlaneUz = 0.003f;
constexpr int kSegments = kBands + 2;
constexpr int kCqDelta = 64;  // polar nodes
constexpr int kCqPsi = 64;    // azimuth nodes on [0, pi] (symmetric pairs: kCqPsi/2 values of t)
constexpr int kCqTab = 4096;  // intervals of the rod-factor table

// Per-|q| values shared by the eight sign/swap images of a midpoint-grid pixel.
struct alignas(32) Radial {
    double q, pq, fq, dwe;  // |q|, P(q), <F>^2, exp(-dwfactor q^2)
};

struct DevSet {
    csx_scalars s;
    const double* carr;      // 4 x kCarrLen: carr1p, carr4p, carr1f, carr4f (carr_of)
    const double* carr11pm;  // CSX_NIM2^2 or null
    const int32_t* latpar1;  // (peakmax1+1) x 6
    const int32_t* latpar2;  // (peakmax2+1) x 6
    const float* latpar3;    // (peakmax2+1) x 17
    // Mixed-precision tables (built at upload, see csx_device.hip).  Friedel pairs in kSegments
    // segments: segment b < kBands holds the pairs with b/kBands <= |u_z| < (b+1)/kBands,
    // segment kBands all pairs, segment kBands + 1 those with |u_z| <= kPlaneUz; each is sorted by |q_hkl|, starts at a multiple of 4 and is
    // padded to one with inert records.  lut[b*(nlut+1)+k] is the index of the first pair of
    // segment b with |q_hkl| >= k/lut_inv_dq (the segment end for k = nlut).
    const struct PairRec* pairs;
    const struct PairRec64* pairs64;  // index-aligned with pairs; lattice kinds with a precise loop only
    const int32_t* lut;
    const int32_t* singles;  // latpar indices of unpaired reflections
    const float2* iso;       // (|q_hkl|, mhkl*fhkl)
    int32_t npairs, nsingle, niso, nlut;
    float lut_inv_dq;
    float cut;               // base-2 exponent below which a pair term is dropped
    float qh_min, qh_max;    // |q_hkl| range of the pairs
    double asum2;            // sum over pairs of 2|A| (infinite when a weight is negative)
    double kp_coef;          // -4/(pi*phiwidth^2)*log2(e): angular exponent = kp_coef*q^2*y^2
    double kr_const;         // -4/(pi*width^2)*log2(e) when the width does not depend on q
    double inv_hpw;          // 1/(pi*width/2) when the width does not depend on q
    double two_pi_over_v;    // 2*pi/sphno_cubevol
    // Per-set constants of the mixed path's sphere closed forms (pq_sphere_fast,
    // fq_sphere_fast): zr, R/(zr+1) and the reciprocal coefficient products.
    double sph_zr, sph_arg, sph_c0, sph_c4, sph_c45;
    double sfq_zr, sfq_arg, sfq_c0, sfq_c2;
    // Gamma(zr+1), Gamma(zr-1/2), Gamma(zr-3/2) of the transverse cylinder <F>^2 (zr from
    // sigma): per-set constants, evaluated once on the host with the reference's Gamma.
    double cyl_g0, cyl_g1, cyl_g2;
    // Orientation quadrature of Gaussian-distributed cylinders in physics mode (built at
    // upload, see csx_device.hip): kCqDelta Gauss-Legendre polar nodes (cos, sin, weight),
    // kCqPsi/2 azimuth nodes t = cos(psi) of the midpoint rule on [0, pi], and the
    // longitudinal rod factor P(a) on [0, amax] as (value, h*slope) at kCqTab+1 knots.
    const float4* cq_delta;
    const float* cq_t;
    const float2* cq_tab;
    float cq_invh;           // kCqTab / amax
    float cq_lz;             // L/(z_L+1)
};

// The coefficient series the pixel bodies read, carr1p, carr4p, carr1f and carr4f, entries
// 0..120 (the longest series sums 120 terms), gathered from the host's 18 x CSX_NCOEF block
// (carr1p..carr9p, carr1f..carr9f) when a set is uploaded or staged.
constexpr int kCarrLen = 121;
constexpr int kCarrArrays[4] = {0, 3, 9, 12};
template <int W>
__device__ __forceinline__ const double* carr_of(const DevSet& d) {
    static_assert(W == 0 || W == 3 || W == 9 || W == 12, "coefficient series not kept on the device");
    return d.carr + (W == 0 ? 0 : W == 3 ? 1 : W == 9 ? 2 : 3) * kCarrLen;
}

__device__ __forceinline__ double sqd(double x) { return x * x; }

// q of pixel (ihex, i), convert_idx2q (.cu:249-287).
__device__ __forceinline__ void pixel_q(const csx_scalars& s, int ihex, int i, double& qx, double& qy, double& qz) {
    if (s.use1d) {
        qx = (s.qmax - s.qmin) / (double)(s.zmax) * (double)ihex;
        qy = 1e-20;
        qz = 1e-20;
    } else if (s.centerbeam) {
        const double mdet = (ihex) * s.pixnoy / (2.0 * s.zmax);
        const double ndet = (i) * s.pixnox / (2.0 * s.zmax);
        const double xdet = s.pixx_m * (ndet - s.beamx);
        const double ydet = s.pixy_m * (mdet - s.beamy);
        const double rdet = sqrt(xdet * xdet + ydet * ydet);
        const double phidet = atan2(ydet, xdet);
        const double thetadet = atan2(rdet, s.det);
        qx = 2 * kPiD * cos(phidet) * sin(thetadet) / s.wave;
        qy = 2 * kPiD * sin(phidet) * sin(thetadet) / s.wave;
        qz = 2 * kPiD * (1 - cos(thetadet)) / s.wave;
    } else {
        qx = s.qmax * i / (double)(s.zmax);
        qy = s.qmax * ihex / (double)(s.zmax);
        qz = 1e-20;
    }
}

// Power series sum_n c[n] x^(2n) with the reference's exit rule (relative change < 1e-4),
// the power built as (xn*x)*x.
__device__ __forceinline__ double series_x2(const double* c, double x, int nmax) {
    double sum = 1.0, old = 0.0, xn = 1.0;
    for (int n = 1; n <= nmax; ++n) {
        xn = xn * x * x;
        sum = sum + c[n] * xn;
        const double d = fabs((sum - old) / sum);
        if (d < 0.0001) break;
        old = sum;
    }
    return sum;
}

// Same series in the longitudinal projection p = qxs+qys: the power is built as (xn*p)*p.
__device__ __forceinline__ double series_proj(const double* c, double p, int nmax) {
    double sum = 1.0, old = 0.0, xn = 1.0;
    for (int n = 1; n <= nmax; ++n) {
        xn = xn * p * p;
        sum += c[n] * xn;
        const double d = fabs((sum - old) / sum);
        if (d < 0.0001) break;
        old = sum;
    }
    return sum;
}

__device__ inline double pq_sphere(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    double zr = (1 - sqd(s.sigma)) / (sqd(s.sigma));
    if (q < 0.4 * s.limq[3]) return series_x2(carr_of<3>(d), q, 100);
    const double argq = q * s.radius / (zr + 1);
    if (zr * (zr - 1) * (zr - 2) * (zr - 3) * (zr - 4) * (zr - 5) == 0) zr += 0.001;
    const double pqr = (1 / (2.0 * zr * (zr - 1) * (zr - 2) * (zr - 3))) * pow(argq, -4);
    const double at = atan(2.0 * argq), b = 1.0 + 4 * argq * argq;
    const double pq1 = pqr * (1 + cos((zr - 3) * at) / pow(b, (zr - 3) / 2.0));
    const double pq2 = (pqr / ((zr - 4) * argq)) * sin((zr - 4) * at) / pow(b, (zr - 4) / 2.0);
    // exponent zr-5 and a trailing /2.0 on the ratio, as the reference writes it
    const double pq3 = (pqr / ((zr - 4) * (zr - 5) * argq * argq)) * (1 - cos((zr - 5) * at) / pow(b, zr - 5) / 2.0);
    return 9.0 * (pq1 - 2.0 * pq2 + pq3);
}

// Isotropic cube (Source/sc_lib_formpq_partCube.h:27-61, Source/sc_gpu_pCub_lNon_oIso.h:37-45):
// the series below 1.2 limq4 (it is summed, with the 1e-4 exit rule, also where it no longer
// converges, between limq4 and 1.2 limq4), the Porod law por/q^4 above; P(q) = 1 for any
// interior other than homogeneous.
__device__ inline double pq_cube(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    if (s.cs != 0) return 1.0;
    if (q < 1.2 * s.limq[3]) return series_x2(carr_of<3>(d), q, 120);
    return s.por / (q * q * q * q);
}

// Perfectly oriented cube (Source/sc_lib_formpq_partCube.h:70-106): below q = 1/R the double
// series sum_n sum_m carr11pm[m][n-m] qx^2(n-m) qy^2m with the powers built as running
// products (kept in private arrays, as the reference indexes them), else the product of the
// two edge factors.
__device__ inline double pq_cube_zdir(const DevSet& d, double qxs, double qys, double q) {
    const csx_scalars& s = d.s;
    const double zr = (1 - sqd(s.sigma)) / (sqd(s.sigma));
    if (q < (1.0 / s.radius)) {
        double qxn[81], qyn[81];
        double pql = 1.0, old = 0.0;
        qxn[0] = 1.0;
        qyn[0] = 1.0;
        for (int n = 1; n <= 80; ++n) {
            qxn[n] = qxn[n - 1] * qxs * qxs;
            qyn[n] = qyn[n - 1] * qys * qys;
            double bs = 0.0;
            for (int m = 0; m <= n; ++m) bs += d.carr11pm[m * CSX_NIM2 + (n - m)] * qxn[n - m] * qyn[m];
            pql += bs;
            const double del = fabs((pql - old) / pql);
            if (del < 0.0001) break;
            old = pql;
        }
        return pql;
    }
    const double ax = qxs * s.radius / (zr + 1) + 0.000000001;
    const double ay = qys * s.radius / (zr + 1) + 0.000000001;
    const double px = (1 / (2.0 * zr * (zr - 1))) * (1 / (ax * ax)) *
                      (1 - cos((zr - 1) * atan(2.0 * ax)) / pow(1.0 + 4 * ax * ax, (zr - 1) / 2.0));
    const double py = (1 / (2.0 * zr * (zr - 1))) * (1 / (ay * ay)) *
                      (1 - cos((zr - 1) * atan(2.0 * ay)) / pow(1.0 + 4 * ay * ay, (zr - 1) / 2.0));
    return px * py;
}

__device__ inline double fq_sphere(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    const double zr = (1 - sqd(s.sigma)) / sqd(s.sigma);
    if (q < (0.4 * s.limq[12])) return series_x2(carr_of<12>(d), q, 100);
    const double argq = q * s.radius / (zr + 1);
    const double pqr = (1 / (zr * (zr - 1)) * pow(argq, -2));
    const double at = atan(argq), b = 1.0 + argq * argq;
    const double pq1 = pqr * cos((zr - 1) * at) / pow(b, (zr - 1) / 2.0);
    const double pq2 = (pqr / ((zr - 2) * argq)) * sin((zr - 2) * at) / pow(b, (zr - 2) / 2.0);
    const double pq3 = 3 * (pq2 - pq1);
    return pq3 * pq3;
}

// Same closed forms with fewer FP64 library calls and one division, for the mixed-precision
// path: the per-set coefficient products are reciprocals taken once (sph_*, sfq_*); the
// powers of one base b come from one log (b^(-(zr-5)/2) and b^(-1/2) by exp), the cosines at
// (zr-3)*at and (zr-5)*at from the sine and cosine at (zr-4)*at by angle addition, with
// cos(at) = 1/sqrt(b) and sin(at) = 2*argq/sqrt(b) for at = atan(2*argq).  Differences from
// the reference's evaluation are rounding-level (phases agree to ~1e-14 rad).
__device__ inline double pq_sphere_fast(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    if (q < 0.4 * s.limq[3]) return series_x2(carr_of<3>(d), q, 100);
    const double zr = d.sph_zr;
    const double argq = q * d.sph_arg;
    const double ia = 1.0 / argq, ia2 = ia * ia;
    const double a2 = argq * argq;
    const double pqr = d.sph_c0 * (ia2 * ia2);
    const double u = 2.0 * argq;
    const double at = atan(u), lb = log(1.0 + 4 * a2);
    const double isb = exp(-0.5 * lb);               // b^(-1/2)
    const double iP = exp(-0.5 * (zr - 5) * lb);     // b^(-(zr-5)/2)
    double s4, c4;
    sincos((zr - 4) * at, &s4, &c4);
    const double sa = u * isb;
    const double cos3 = c4 * isb - s4 * sa, cos5 = c4 * isb + s4 * sa;
    const double pq1 = pqr * (1 + cos3 * (iP * (isb * isb)));
    const double pq2 = pqr * (d.sph_c4 * ia) * s4 * (iP * isb);
    const double pq3 = pqr * (d.sph_c45 * ia2) * (1 - cos5 * (iP * iP) / 2.0);
    return 9.0 * (pq1 - 2.0 * pq2 + pq3);
}

// fq_sphere with one division, one log and the cosine at (zr-1)*at from the sine and cosine
// at (zr-2)*at (at = atan(argq), cos(at) = 1/sqrt(b), sin(at) = argq/sqrt(b)).
__device__ inline double fq_sphere_fast(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    if (q < (0.4 * s.limq[12])) return series_x2(carr_of<12>(d), q, 100);
    const double zr = d.sfq_zr;
    const double argq = q * d.sfq_arg;
    const double ia = 1.0 / argq;
    const double pqr = d.sfq_c0 * (ia * ia);
    const double at = atan(argq), lb = log(1.0 + argq * argq);
    const double isb = exp(-0.5 * lb);            // b^(-1/2)
    const double iP = exp(-0.5 * (zr - 2) * lb);  // b^(-(zr-2)/2)
    double s2, c2;
    sincos((zr - 2) * at, &s2, &c2);
    const double cos1 = c2 * isb - s2 * (argq * isb);
    const double pq1 = pqr * cos1 * (iP * isb);
    const double pq2 = pqr * (d.sfq_c2 * ia) * s2 * iP;
    const double pq3 = 3 * (pq2 - pq1);
    return pq3 * pq3;
}

// Schulz-averaged longitudinal P of a rod, argument (qxs+qys+eps9)*L/(zl+1).
__device__ __forceinline__ double rod_p(double zl, double arg) {
    return (1 / (2.0 * zl * (zl - 1))) * (1 / (arg * arg)) *
           (1 - cos((zl - 1) * atan(2.0 * arg)) / pow(1.0 + 4 * arg * arg, (zl - 1) / 2.0));
}

// Orientation case as the form factors compare it: the reference passes its 0-based enum
// where the comparisons use the 1-based numbering; fix_orcase maps it (physics mode).
__device__ __forceinline__ int orc_cmp(const csx_scalars& s) { return s.fix_orcase ? s.orcase + 1 : s.orcase; }

// Transverse part of the homogeneous cylinder P(q) (:16246-16279).
__device__ inline double pq_cyl_transverse(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    const double zr = s.sigma == 1.0 ? 0.000000001 : (1 - sqd(s.sigma)) / (sqd(s.sigma));
    if ((s.length / s.radius) < 2) return 1;
    if (q < (0.65 * s.limq[3])) return series_x2(carr_of<3>(d), q, 120);
    const double argpq = q * s.radius / (zr + 1);
    const double at = atan(2.0 * argpq), b = 1.0 + 4 * argpq * argpq;
    const double pqr1 = (1 / (zr * (zr - 1) * (zr - 2))) * pow(argpq, -3);
    const double pqr2 = (1 / (zr * (zr - 1) * (zr - 2))) * pow(argpq, -3) * sin((zr - 2) * at) / pow(b, (zr - 2) / 2.0);
    const double pqr3 = (1 / (zr * (zr - 1) * (zr - 2) * (zr - 3))) * pow(argpq, -4) * cos((zr - 3) * at) / pow(b, (zr - 3) / 2.0);
    return (4. / kPiD) * (pqr1 - pqr2 - (9. / 8.0) * pqr3);
}

// Longitudinal part for isotropic orientation, factorisation branch (L/R >= 2,
// Source/sc_lib_formpq_partCylinder.h:30-90): the P|| series below 0.6 limq1, else the
// rod average's closed form.
__device__ inline double pq_cyl_long_iso(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    const double zl = s.sigmal == 1.0 ? 0.000000001 : (1 - sqd(s.sigmal)) / sqd(s.sigmal);
    if (q < (0.6 * s.limq[0])) return series_x2(carr_of<0>(d), q, 120);
    const double arglq = q * s.length / (zl + 1);
    double pql = (kPiD / (2.0 * zl)) * (1 / arglq);
    pql = pql - (1 / (2.0 * zl * (zl - 1) * arglq * arglq)) * cos((zl - 1) * atan(2.0 * arglq)) /
                    pow(1.0 + 4 * arglq * arglq, (zl - 1) / 2.0);
    return pql;
}

// Longitudinal part for perfect (z-dir) orientation, ordis 6 (:16143-16168).
__device__ inline double pq_cyl_long_zdir(const DevSet& d, double qxs, double qys, double q) {
    const csx_scalars& s = d.s;
    if (orc_cmp(s) == 4) return 1.0;
    const double zl = s.sigmal == 1.0 ? 0.000000001 : (1 - sqd(s.sigmal)) / sqd(s.sigmal);
    if (q < (0.6 * s.limq[0])) return series_proj(carr_of<0>(d), qxs + qys, 120);
    return rod_p(zl, (qxs + qys + 0.000000001) * s.length / (zl + 1));
}

// formfq_partCylinder, ordis 6, cs 0 (Source/sc_lib_formfq_partCylinder.h:8-206): the
// longitudinal factor (projection qxs + qys, series below 0.6 limq1f of limql) times the radial
// one (a function of |q| only, kept apart for the radial tables).
__device__ inline double fq_cyl_long(const DevSet& d, double limql, double qxs, double qys) {
    const csx_scalars& s = d.s;
    const double zl = (1 - sqd(s.sigmal)) / sqd(s.sigmal);
    if (orc_cmp(s) == 4) return 1.0;
    if (limql < (0.6 * s.limq[9])) return series_proj(carr_of<9>(d), qxs + qys, 120);
    return rod_p(zl, (qxs + qys + 0.000000001) * s.length / (zl + 1));
}
__device__ inline double fq_cyl_radial(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    const double zr = (1 - sqd(s.sigma)) / sqd(s.sigma);
    double pqr;
    if (q < (1.0 * s.limq[12])) {
        pqr = series_x2(carr_of<12>(d), q, 120);
    } else {
        const double argpq = q * s.radius / (zr + 1);
        const double at = atan(argpq), b = 1.0 + argpq * argpq;
        const double g0 = d.cyl_g0;
        const double pqr1 = (d.cyl_g1 / g0) * pow(argpq, -3 / 2.0) *
                            (sin((zr - 1 / 2.0) * at) - cos((zr - 1 / 2.0) * at)) / pow(b, (zr - 1 / 2.0) / 2.0);
        const double pqr2 = (d.cyl_g2 / g0) * pow(argpq, -5 / 2.0) *
                            (sin((zr - 3 / 2.0) * at) + cos((zr - 3 / 2.0) * at)) / pow(b, (zr - 3 / 2.0) / 2.0);
        const double pqr3 = (2 / sqrt(kPiD)) * (pqr1 + (9 / 16.0) * pqr2);
        pqr = pqr3 * pqr3;
    }
    return pqr;
}
// Longitudinal part of <F>^2 for isotropic orientation (formfq_partCylinder ordis 7,
// Source/sc_lib_formfq_partCylinder.h:27-53): the series below 0.6 limq1f, else the closed form.
__device__ inline double fq_cyl_long_iso(const DevSet& d, double q) {
    const csx_scalars& s = d.s;
    const double zl = (1 - sqd(s.sigmal)) / sqd(s.sigmal);
    if (q < (0.6 * s.limq[9])) return series_x2(carr_of<9>(d), q, 120);
    const double arglq = q * s.length / (zl + 1);
    double pql = (kPiD / (2.0 * zl)) * (1 / arglq);
    pql = pql - (1 / (2.0 * zl * (zl - 1) * arglq * arglq)) * cos((zl - 1) * atan(2.0 * arglq)) /
                    pow(1.0 + 4 * arglq * arglq, (zl - 1) / 2.0);
    return pql;
}
__device__ inline double fq_cyl_zdir(const DevSet& d, double limql, double qxs, double qys, double q) {
    return fq_cyl_long(d, limql, qxs, qys) * fq_cyl_radial(d, q);
}

// pq_cyl_long_zdir and fq_cyl_long at one pixel: the same values, with the rod average
// evaluated once when both take it at the same argument (z_L agrees unless sigma_L = 1).
__device__ inline void cyl_long_both(const DevSet& d, double q, double limql, double qxs, double qys, double& pl,
                                     double& fl) {
    const csx_scalars& s = d.s;
    if (orc_cmp(s) == 4) {
        pl = 1.0;
        fl = 1.0;
        return;
    }
    const bool pser = q < (0.6 * s.limq[0]);
    const bool fser = limql < (0.6 * s.limq[9]);
    const double zlf = (1 - sqd(s.sigmal)) / sqd(s.sigmal);
    const double zlp = s.sigmal == 1.0 ? 0.000000001 : zlf;
    const bool same = zlp == zlf;
    double rod = 0.0;
    if (!fser || (!pser && same)) rod = rod_p(zlf, (qxs + qys + 0.000000001) * s.length / (zlf + 1));
    if (pser) pl = series_proj(carr_of<0>(d), qxs + qys, 120);
    else pl = same ? rod : rod_p(zlp, (qxs + qys + 0.000000001) * s.length / (zlp + 1));
    fl = fser ? series_proj(carr_of<9>(d), qxs + qys, 120) : rod;
}

}  // namespace csx
