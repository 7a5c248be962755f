// exp(x) in double-double arithmetic, for the device preparation of reflection tables.
//
// The host's tables take exp from the C library, whose result is within 0.511 ulp of e^x
// (glibc 2.28 and later); the device's own exp is within an ulp.  To round the tables' float
// entries alike, the device evaluates e^x to ~2^-95 relative here and returns the correctly
// rounded double together with a flag for the rare arguments where e^x lies so close to the
// midpoint between two doubles that a 0.511-ulp exp may round the other way (then `alt` is
// the other neighbour, and the caller checks whether the choice can matter).
//
// Method: x = k ln2 + r with ln2 in three pieces (the first with 41 significant bits, so k ln2
// exact for |k| < 4096), r reduced by 2^-8, e^(r/256) - 1 from its Taylor series to degree
// 10 in double-double, squared eight times in the form s -> s (2 + s), then 1 + s scaled by
// 2^k.  Double-double sums and products follow Dekker and Knuth (two-sum, fma two-product).
// Valid for -700 <= x <= 0, the range the peak norm needs (below, e^x < 1e-304 cannot change
// a table entry; the flag is then false).
#pragma once

#include <cmath>

#if !defined(CSX_HD)
#if defined(__HIPCC__)
#define CSX_HD __host__ __device__
#else
#define CSX_HD
#endif
#endif

namespace csx {
namespace ddexp {

struct DD {
    double hi, lo;
};

CSX_HD inline DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
CSX_HD inline DD fast_two_sum(double a, double b) {  // |a| >= |b|
    const double s = a + b;
    return {s, b - (s - a)};
}
CSX_HD inline DD add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    return fast_two_sum(s.hi, s.lo);
}
CSX_HD inline DD mul(DD a, DD b) {
    const double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p, e);
}

// Correctly rounded e^x for -700 <= x <= 0.  near_mid: e^x lies within 0.02 ulp of a midpoint;
// alt: the neighbour of the result on e^x's side.
CSX_HD inline double exp_rn(double x, bool& near_mid, double& alt) {
    near_mid = false;
    alt = 0.0;
    if (x < -700.0) {
        const double r = exp(x);
        alt = r;
        return r;
    }
    if (x == 0.0) {
        alt = 1.0;
        return 1.0;
    }
    constexpr double kInvLn2 = 0x1.71547652b82fep+0;
    constexpr double kL1 = 0x1.62e42fefa4000p-1;   // 41 significant bits
    constexpr double kL2 = -0x1.8432a1b0e2634p-43;
    constexpr double kL3 = 0x1.f97b57a079a19p-103;
    // 1/n! as double-doubles, n = 2..10
    constexpr double kFh[11] = {0, 0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
                                0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
                                0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22};
    constexpr double kFl[11] = {0, 0, 0.0, 0x1.5555555555555p-57, 0x1.5555555555555p-59,
                                0x1.1111111111111p-63, -0x1.f49f49f49f49fp-65, 0x1.a01a01a01a01ap-73,
                                0x1.a01a01a01a01ap-76, -0x1.c154f8ddc6c00p-73, 0x1.cbbc05b4fa99ap-76};
    const double k = rint(x * kInvLn2);
    // r = x - k ln2: the first product is exact and cancels exactly against x
    const double t = x - k * kL1;
    const double p2 = k * kL2;
    const double p2e = fma(k, kL2, -p2);
    DD r = two_sum(t, -p2);
    r.lo += -p2e - k * kL3;
    r = fast_two_sum(r.hi, r.lo);
    const DD rs{r.hi * 0x1p-8, r.lo * 0x1p-8};
    // s = e^rs - 1 = rs (1 + rs (1/2! + rs (1/3! + ... + rs/10!)))
    DD s{kFh[10], kFl[10]};
    for (int n = 9; n >= 2; --n) s = add(mul(s, rs), DD{kFh[n], kFl[n]});
    s = add(mul(s, rs), DD{1.0, 0.0});
    s = mul(s, rs);
    for (int i = 0; i < 8; ++i) s = mul(s, add(DD{2.0, 0.0}, s));
    DD e = add(DD{1.0, 0.0}, s);
    const int ki = (int)k;
    const double hi = ldexp(e.hi, ki), lo = ldexp(e.lo, ki);
    // distance to the midpoint on lo's side, in units of the gap to that neighbour
    const double nb = lo > 0 ? nextafter(hi, INFINITY) : nextafter(hi, 0.0);
    const double gap = fabs(nb - hi);
    near_mid = lo != 0.0 && fabs(lo) >= 0.48 * gap;
    alt = nb;
    return hi;
}

}  // namespace ddexp
}  // namespace csx
