// One row of the anisotropic reflection table (latpar3) from the rotated reciprocal bases.
//
// Shared by the host preparation (csx_prep.cpp, g++ on x86-64: SSE2 doubles, no contraction)
// and the device preparation of batched sets (csx_device.hip, hipcc -ffp-contract=off): the
// same operations in the same order, so that both round every table value alike.  Only the
// exponential of gaussnorm3 differs (the C library's on the host, a correctly rounded one on
// the device, csx_ddexp.h), and its callers check that the rounded table entry cannot depend
// on which of the two it was.
//
// Reference: the table fill Source/sc_calc_generic_cpu.cpp:838-954 (q_hkl from the reciprocal
// vectors of corotations, :1487, g3 and the peak norm), gaussnorm3
// Source/sc_calc_generic_gpu.cu:838-854.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define CSX_HD __host__ __device__
#else
#define CSX_HD
#endif

namespace csx {
namespace rows {

constexpr double kPi = 3.14159265358979323846;

// q_hkl = 2 pi (h a* + k b* + l c*) with the reciprocal vectors a*, b*, c* in the rows of R,
// its twinned counterpart from RT, |q_hkl| and 2 pi sqrt(h^2 |a*|^2 + k^2 |b*|^2) (latpar3
// entries 2-9 and 13).
struct Geometry {
    double qx, qy, qz, qh, qxy, tx, ty, tz;
};

CSX_HD inline void geometry(const double* R, const double* RT, int h, int k, int l, Geometry& g) {
    g.qx = (2 * kPi) * (R[0] * h + R[3] * k + R[6] * l);
    g.qy = (2 * kPi) * (R[1] * h + R[4] * k + R[7] * l);
    g.qz = (2 * kPi) * (R[2] * h + R[5] * k + R[8] * l);
    g.qxy = (2 * kPi) * sqrt(h * h * (R[0] * R[0] + R[1] * R[1] + R[2] * R[2]) +
                             k * k * (R[3] * R[3] + R[4] * R[4] + R[5] * R[5]));
    g.qh = sqrt(g.qx * g.qx + g.qy * g.qy + g.qz * g.qz);
    g.tx = (2 * kPi) * (RT[0] * h + RT[3] * k + RT[6] * l);
    g.ty = (2 * kPi) * (RT[1] * h + RT[4] * k + RT[7] * l);
    g.tz = (2 * kPi) * (RT[2] * h + RT[5] * k + RT[8] * l);
}

// g3 = 4 pi^2 / (2 pi |q_hkl|^2) (latpar3 entry 10)
CSX_HD inline double g3(double qh) { return 4 * kPi * kPi / (2.0 * kPi * qh * qh); }

// argument a of the peak norm gaussnorm3(a), a = 4 |q_hkl|^2 / (pi phiwidth^2)
CSX_HD inline double x2phi(double qh, double phiwidth) { return 4 * qh * qh / (kPi * phiwidth * phiwidth); }

// gaussnorm3 uses exp(-a pi^2) when a > 0.01, a cubic in a otherwise
CSX_HD inline bool gauss_norm3_uses_exp(double a) { return a > 0.01; }
CSX_HD inline double gauss_norm3_exp_arg(double a) { return -a * kPi * kPi; }

// gaussnorm3(a) given ex = exp(-a pi^2) (ignored for a <= 0.01)
CSX_HD inline double gauss_norm3_from(double a, double ex) {
    if (a > 0.01) {
        const double b0 = (1 - ex) / 2.;
        const double b1 = (1 - ex * (1 + a * kPi * kPi)) / 12.;
        const double b2 = (2 - ex * (2 + 2 * a * kPi * kPi + a * a * kPi * kPi * kPi * kPi)) / 240.;
        const double b3 = (6 - ex * (6 + 6 * a * kPi * kPi + 3 * a * a * kPi * kPi * kPi * kPi +
                                     a * a * a * kPi * kPi * kPi * kPi * kPi * kPi)) / 10080.;
        return b0 / a - b1 / (a * a) + b2 / (a * a * a) - b3 / (a * a * a * a);
    }
    const double a1 = (kPi * kPi - 4);
    const double a2 = (18 - 12 * kPi * kPi + kPi * kPi * kPi * kPi) / 2.;
    const double a3 = (-1440 + 360 * kPi * kPi - 30 * kPi * kPi * kPi * kPi + kPi * kPi * kPi * kPi * kPi * kPi) / 6.;
    return 2 - a1 * a + a2 * a * a - a3 * a * a * a;
}

}  // namespace rows
}  // namespace csx
