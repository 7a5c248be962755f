// csx_device.hip -- HIP kernels (gfx950) and the device half of the C ABI.
//
// One thread per detector pixel, 16x16-pixel tiles per 256-thread block
// (4 wave64s, each a 16x4 strip), coalesced FP64 stores in the reference's IDX
// order.  Per-pixel form factors run in FP64 with the reference formulas; the
// Gaussian peak sums run either in FP64 exactly as the reference writes them
// (CSX_PREC_REF64) or in FP32 over Friedel pairs (CSX_PREC_MIXED): +hkl and
// -hkl share |q^.u| and |q_hkl|, so one acos/sqrt serves both, their angular
// and radial Gaussians fold into one exp2 each, and partial sums go to an FP64
// accumulator.  Pairs whose |q^.u| is within 1e-6 of 1 are evaluated in FP64
// exactly as the reference does (its acos clamp maps out-of-range cosines to
// swapped angles, which decides those pixels).  The pair table is read with
// wave-uniform addresses, i.e. through the scalar cache.
#include <hip/hip_runtime.h>
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "csx.h"
#include "csx_physics.h"
#include "csx_prep.h"
#include "csx_ddexp.h"
#include "csx_rows.h"

namespace csx {

constexpr int kTile = 16;         // pixels per tile edge
constexpr int kBlock = kTile * kTile;
constexpr float kPiF = 3.14159265358979323846f;
constexpr double kLog2e = 1.4426950408889634074;

// K_CYL_GAUSS_Q: Gaussian cylinders in physics mode with the FP32 orientation quadrature
// (mixed precision only; chosen by csx_render, never by the key alone).
enum Kind { K_SPHERE = 0, K_FCC = 1, K_CYL_GAUSS = 2, K_CYL_ZDIR = 3, K_BCT = 4, K_CYL_GAUSS_Q = 5, K_CUBE = 6,
            K_CYL_ISO = 7, K_CUBE_ZDIR = 8, K_ISO_LAT = 9 };
// Kinds with a lattice sum: spheres (and every key whose factors behave alike) on lattices,
// isotropic cylinders on lattices, perfectly oriented cylinders on lattices.
constexpr bool lattice_kind(int kind) { return kind == K_FCC || kind == K_ISO_LAT || kind == K_BCT; }

inline int kind_of(uint32_t key) {
    switch (key) {
    case CSX_KEY_SPHERE_NONE_ISO:
    case CSX_KEY_SPHERE_NONE_GAUSS:
    case CSX_KEY_SPHERE_NONE_ZDIR: return K_SPHERE;
    case CSX_KEY_SPHERE_FCC_GAUSS:
    case CSX_KEY_SPHERE_FCC_ISO:
    case CSX_KEY_SPHERE_BCC_GAUSS:
    case CSX_KEY_SPHERE_BCC_ISO:
    case CSX_KEY_SPHERE_SC_GAUSS:
    case CSX_KEY_SPHERE_SC_ISO:
    case CSX_KEY_SPHERE_BCT_GAUSS:
    case CSX_KEY_SPHERE_BCT_ISO: return K_FCC;
    case CSX_KEY_CYL_NONE_GAUSS: return K_CYL_GAUSS;
    case CSX_KEY_CYL_NONE_ZDIR: return K_CYL_ZDIR;
    case CSX_KEY_CYL_BCT_ZDIR: return K_BCT;
    case CSX_KEY_CUBE_NONE_ISO: return K_CUBE;
    case CSX_KEY_CYL_NONE_ISO: return K_CYL_ISO;
    case CSX_KEY_CUBE_NONE_ZDIR: return K_CUBE_ZDIR;
    default: break;
    }
    // The lattices served: the 2D ones (hexagonal, square, centred rectangular), the cubic ones
    // (BCC, FCC, SC, Ia3d, Pn3m, Im3m, Fd3m), HCP, BCT and the orthorhombic one (HCP and
    // orthorhombic without reflections).  Spheres on them with the Lorentzian, Gaussian, mod.
    // 1/2 Lorentzian, Pseudo-Voigt or Pearson VII peak, any orientation distribution (which
    // enters neither the sphere's factors nor these sums); perfectly oriented cylinders on
    // them (calc_GENERIC, and the BCT Gaussian key's own kernel).
    const uint32_t part = key >> 24, shp = key >> 16 & 0xff, lt = key >> 8 & 0xff, ord = key & 0xff;
    const bool lattice = (lt >= 1 && lt <= 11) || lt == 17 || lt == 18;
    if (part == 0 && shp >= 1 && shp <= 6 && lattice && ord <= 13) return K_FCC;
    if (part == 1 && shp >= 1 && shp <= 6 && lattice && ord == 6) return K_BCT;
    // isotropic cylinders and cubes on them: P(q) and <F>^2 depend on |q| only, so they take
    // the sums of the sphere lattices with the particle's factors (peak shapes 1-6 for all)
    if ((part == 1 || part == 4) && shp >= 1 && shp <= 6 && lattice && ord == 7) return K_ISO_LAT;
    return -1;
}
// Isotropic cubes on a lattice (kind K_ISO_LAT; the other particle of that kind is the cylinder).
__host__ __device__ inline bool lattice_cube(uint32_t key) { return (key >> 24) == 4; }
// Peak shape of a lattice key (2: Gaussian).
inline int shape_of(uint32_t key) { return (int)(key >> 16 & 0xff); }
// The arithmetic a request runs in: sets with a peak shape other than the Gaussian have only
// the FP64 reference-order sum (the mixed path's windows assume the Gaussian's tails).
inline int effective_precision(uint32_t key, int precision) {
    return (precision == CSX_PREC_MIXED && lattice_kind(kind_of(key)) && shape_of(key) != 2) ? CSX_PREC_REF64
                                                                                            : precision;
}
// Kinds whose mixed peak sum runs the precise pair loop (FP64 exponents): their structure
// factor 1 + (fq/pq)(2*pi*I/V - 1)*dwe cancels over much of the pattern (fq/pq reaches 1e8
// for the z-oriented cylinder), so the FP32 loop's ~1e-5 would send a third of the pixels
// to the FP64 re-sum.
constexpr bool precise_kind(int kind) { return kind == K_BCT; }

// ---------------------------------------------------------------------------
// Gaussian peak sums
// ---------------------------------------------------------------------------

// The reference's acos with its out-of-range mapping (Source/sc_calc_generic_gpu.cu:31).
__device__ __forceinline__ double ref_acos(double x) { return (x < -1) ? 0.0 : ((x > +1) ? kPiD : acos(x)); }

// One anisotropic reflection exactly as the reference evaluates it
// (Source/sc_gpu_pSph_pGau_lFCC_oGau.h:85-142), without the 1/(pi*width/2) factor.
__device__ double aniso_term_ref(const DevSet& d, int ii, double qx, double qy, double qz, double q,
                                 double width, double x2phi) {
    const float* L3 = d.latpar3 + ii * 17;
    const int32_t* L2 = d.latpar2 + ii * 6;
    const double qhkl = L3[5];
    const int mf = L2[4] * L2[5];
    if (!(qhkl > 0 && mf != 0)) return 0.0;
    const double yphi = ref_acos((qx * (double)L3[2] + qy * (double)L3[3] + qz * (double)L3[4]) / (q * qhkl));
    const double x2 = (q - qhkl) * (q - qhkl) / (width * width);
    const double sq = exp(-4 * x2 / kPiD) / (kPiD * width / 2.0);
    const double psi = (double)L3[10] * exp(-x2phi * yphi * yphi) / (double)L3[11];
    return sq * mf * psi;
}

// A pair with |q^.u| near 1 in the hot image pass.  The reference takes the angle of each
// reflection as acos of its own FP64 cosine c = q.q_hkl / (|q| |q_hkl|) (|q_hkl| stored in
// FP32), which near 1 is off the true angle by up to ~3e-4 rad, and maps c > 1 to pi and
// c < -1 to 0.  The partner's cosine is exactly -c, so the pair's two terms are the pair
// formula at y = acos(|c|), or at y = 0 when |c| > 1 (the two clamps swap the terms).  The
// FP64 cosine decides; x = 1 - |c| is exact in FP64 and acos(1-x)^2 = 2x + x^2/3 + O(x^3)
// (x < 3e-6 here).
struct NearCtx {
    const DevSet* d;
    int ihex, i;
    double q;
};
template <bool PLANAR>
__device__ __forceinline__ double near_pair_x(const NearCtx& nc, int ia) {
    const DevSet& d = *nc.d;
    double qx, qy, qz;
    if constexpr (PLANAR) {
        // pixel_q without its beam-geometry branch (the detector plane has no centre beam)
        const csx_scalars& s = d.s;
        if (s.use1d) {
            qx = (s.qmax - s.qmin) / (double)(s.zmax) * (double)nc.ihex;
            qy = 1e-20;
            qz = 1e-20;
        } else {
            qx = s.qmax * nc.i / (double)(s.zmax);
            qy = s.qmax * nc.ihex / (double)(s.zmax);
            qz = 1e-20;
        }
    } else {
        pixel_q(d.s, nc.ihex, nc.i, qx, qy, qz);
    }
    const float* L3 = d.latpar3 + ia * 17;
    const double qhkl = L3[5];
    const double c = fabs((qx * (double)L3[2] + qy * (double)L3[3] + qz * (double)L3[4]) / (nc.q * qhkl));
    return c > 1.0 ? 0.0 : 1.0 - c;
}

// Reference-order FP64 sums: intensity over the anisotropic list (stopping at the first
// inf, :131-140) and radintensity over the isotropic list (reading the anisotropic q_hkl
// column, :69-83).
// The other peak shapes of calc_GENERIC (Source/sc_gpu_generic.h:695-849): the isotropic
// Gaussian sum and the anisotropic sum in the generic kernel's order (a reflection with
// |q_hkl| <= 0 or NaN is skipped, the term is sq*mhkl*fhkl*psiord, a sum that turns inf or
// NaN keeps its previous value and ends the loop).
__device__ void lattice_ref64_shape(const DevSet& d, double qx, double qy, double qz, double q, double width,
                                    double& inten, double& radint) {
    const csx_scalars& s = d.s;
    const double widthiso = 1.0 / s.uca;
    radint = 0.0;
    for (int ii = 1; ii <= s.peakmax1; ++ii) {
        const int mhkl = d.latpar1[ii * 6 + 4], fhkl = d.latpar1[ii * 6 + 5];
        const double qhkl = d.latpar3[ii * 17 + 5];
        const double x2 = sqd(q - qhkl) / sqd(widthiso);
        const double sq = exp(-4 * x2 / kPiD) / (kPiD * widthiso / 2.0);
        const double sav = radint;
        radint += sq * mhkl * fhkl;
        if (isinf(radint) || isnan(radint)) { radint = sav; break; }
    }
    const int shp = (int)(s.combo_key >> 16 & 0xff);
    const double eta = s.peak_eta, beta = s.peak_beta;
    const double c0 = s.peak_c[0], c1 = s.peak_c[1], c2 = s.peak_c[2], c3 = s.peak_c[3], c4 = s.peak_c[4];
    inten = 0.0;
    for (int ii = 1; ii <= s.peakmax2; ++ii) {
        const float* L3 = d.latpar3 + ii * 17;
        const double qhkl = L3[5];
        if (qhkl <= 0 || isnan(qhkl)) continue;
        const int mhkl = d.latpar2[ii * 6 + 4], fhkl = d.latpar2[ii * 6 + 5];
        const double g3 = L3[10], pn1 = L3[11], pn2 = L3[12];
        const double yphi = ref_acos((qx * (double)L3[2] + qy * (double)L3[3] + qz * (double)L3[4]) / (q * qhkl));
        const double x2 = (q - qhkl) * (q - qhkl) / (width * width);
        const double x2phi = 4 * q * q / sqd(s.phiwidth);  // the Lorentzian shapes' factor
        double sq = 0.0, psi = 1.0;
        switch (shp) {
        case 1:  // Lorentzian
            sq = sqrt(c1) / (kPiD * width * (1.0 + c1 * x2));
            psi = g3 / (pn1 * (1 + x2phi * yphi * yphi));
            break;
        case 3:  // mod. 1 Lorentzian
            sq = 2 * sqrt(c2) / (kPiD * width * (1 + c1 * x2) * (1 + c1 * x2));
            psi = g3 / (pn1 * (1 + x2phi * yphi * yphi) * (1 + x2phi * yphi * yphi));
            break;
        case 4:  // mod. 2 Lorentzian
            sq = sqrt(c3) / (2 * width * exp(3 * log(1 + c1 * x2) / 2.0));
            psi = g3 / (pn1 * exp(3 * log(1 + x2phi * yphi * yphi) / 2.0));
            break;
        case 5: {  // Pseudo-Voigt
            sq = eta * sqrt(c1) / (kPiD * width * (1 + c1 * x2)) +
                 (1 - eta) * sqrt(c0) * exp(-c0 * x2) / (sqrt(kPiD) * width);
            const double x2psi = 4 * q * q / (kPiD * sqd(s.phiwidth));
            psi = g3 * (eta * (1 / (1 + x2phi * yphi * yphi)) + (1 - eta) * exp(-x2psi * yphi * yphi)) /
                  (eta * pn1 + (1 - eta) * pn2);
            break;
        }
        case 6:  // Pearson VII
            sq = s.peak_gamma_b * c4 * 2 / (s.peak_gamma_bh * kPiD * width * exp(beta * log(1 + 4 * c4 * x2)));
            psi = g3 / (pn1 * exp(beta * log(1 + x2phi * yphi * yphi)));
            break;
        default: break;
        }
        const double sav = inten;
        inten += sq * mhkl * fhkl * psi;
        if (isinf(inten) || isnan(inten)) { inten = sav; break; }
    }
}

__device__ void lattice_ref64(const DevSet& d, double qx, double qy, double qz, double q, double width,
                              double& inten, double& radint) {
    const csx_scalars& s = d.s;
    if ((s.combo_key >> 16 & 0xff) != 2) {
        lattice_ref64_shape(d, qx, qy, qz, q, width, inten, radint);
        return;
    }
    const double widthiso = 1.0 / s.uca;
    radint = 0.0;
    for (int ii = 1; ii <= s.peakmax1; ++ii) {
        const int mhkl = d.latpar1[ii * 6 + 4], fhkl = d.latpar1[ii * 6 + 5];
        if (mhkl == 0 || fhkl == 0) continue;
        const double qhkl = d.latpar3[ii * 17 + 5];
        const double x2 = sqd(q - qhkl) / sqd(widthiso);
        const double sq = exp(-4 * x2 / kPiD) / (kPiD * widthiso / 2.0);
        radint += sq * mhkl * fhkl;
    }
    const double x2phi = 4 * q * q / (kPiD * sqd(s.phiwidth));
    inten = 0.0;
    for (int ii = 1; ii <= s.peakmax2; ++ii) {
        const double t = aniso_term_ref(d, ii, qx, qy, qz, q, width, x2phi);
        const double sav = inten;
        inten += t;
        if (isinf(inten)) { inten = sav; break; }
    }
}

// acos(c)^2 / (1 - c^2) for c in [0, 1], as a degree-9 polynomial in x = 1 - c fitted for
// relative error (max 1.8e-7 in FP32 Horner).  The function is analytic on [0, 1] (its only
// singularity is the pole at c = -1), equal to 1 at c = 1 and (pi/2)^2 at c = 0.
__device__ __forceinline__ float acos2_over_sin2(float c) {
    const float x = 1.0f - c;
    float p = 9.084462374e-02f;
    p = __builtin_fmaf(p, x, -2.716867626e-01f);
    p = __builtin_fmaf(p, x, 4.152069092e-01f);
    p = __builtin_fmaf(p, x, -2.917656898e-01f);
    p = __builtin_fmaf(p, x, 2.068143934e-01f);
    p = __builtin_fmaf(p, x, 6.417080760e-02f);
    p = __builtin_fmaf(p, x, 2.098862827e-01f);
    p = __builtin_fmaf(p, x, 3.772477508e-01f);
    p = __builtin_fmaf(p, x, 6.666826606e-01f);
    p = __builtin_fmaf(p, x, 9.999999404e-01f);
    return p;
}
// The same polynomial for four arguments, Horner step by step across them (four independent
// chains in flight); each result equals acos2_over_sin2 of its argument.
__device__ __forceinline__ void acos2_over_sin2_x4(const float* c, float* out) {
    constexpr float kC[10] = {9.084462374e-02f, -2.716867626e-01f, 4.152069092e-01f, -2.917656898e-01f,
                              2.068143934e-01f, 6.417080760e-02f,  2.098862827e-01f, 3.772477508e-01f,
                              6.666826606e-01f, 9.999999404e-01f};
    float x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        x[j] = 1.0f - c[j];
        out[j] = kC[0];
    }
#pragma unroll
    for (int k = 1; k < 10; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) out[j] = __builtin_fmaf(out[j], x[j], kC[k]);
}

// Wave minimum / maximum through DPP moves (all 64 lanes active): prefix steps within each
// 16-lane row (row_shr 1, 2, 4, 8), then the row_bcast steps carry rows 0..3 into lane 63,
// whose value every lane reads.  A lane whose DPP source is out of its row keeps its value.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROWS, 0xf, false));
}
__device__ __forceinline__ float wave_min_f(float v) {
    v = fminf(v, dpp_move<0x111, 0xf>(v));  // row_shr:1
    v = fminf(v, dpp_move<0x112, 0xf>(v));  // row_shr:2
    v = fminf(v, dpp_move<0x114, 0xf>(v));  // row_shr:4
    v = fminf(v, dpp_move<0x118, 0xf>(v));  // row_shr:8
    v = fminf(v, dpp_move<0x142, 0xa>(v));  // row_bcast:15 into rows 1 and 3
    v = fminf(v, dpp_move<0x143, 0xc>(v));  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_max_f(float v) {
    v = fmaxf(v, dpp_move<0x111, 0xf>(v));
    v = fmaxf(v, dpp_move<0x112, 0xf>(v));
    v = fmaxf(v, dpp_move<0x114, 0xf>(v));
    v = fmaxf(v, dpp_move<0x118, 0xf>(v));
    v = fmaxf(v, dpp_move<0x142, 0xa>(v));
    v = fmaxf(v, dpp_move<0x143, 0xc>(v));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
// Butterfly sums: every lane gets the total, in the same order on every lane.
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}


// Relative error bounds of the two pair loops' sums, used by the conditioning guard.  FP32
// loop: an exponent e carries ~7 ulp relative error plus the FP32 direction error of the
// angular part, about 1e-5 of a term at the 2^-32 cut.  Precise loop: the FP32 exp2 of the
// fraction (1 ulp), its conversion and the polynomial (2.4e-11 of y^2): 5e-7 with margin.
constexpr double kRelErrFP32 = 1.0e-5;
constexpr double kRelErrPrecise = 5.0e-7;

// |q^.u| above this can put the reference's FP64 cosine above 1, where its acos mapping
// returns pi (Source/sc_calc_generic_gpu.cu:31): such pairs take the FP64 formula.
constexpr float kNearOne = 0.999998f;

// Per-pixel constants of the FP32 pair sum (base-2 exponents).
struct PixelConsts {
    float hx, hy, hz;      // q^
    float qf, qlo;         // q = qf + qlo
    float kr;              // radial exponent = kr * (q - q_hkl)^2
    float kp;              // angular exponent = kp * y^2, y = angle(q^, +-u)
    float c1, c2;          // second Friedel term: e2 = e1 + c1 + c2 * y
};

// A pixel's position and, on demand, its exact FP64 q vector (convert_idx2q).
struct PixQ {
    int ihex, i;
    double qx, qy, qz;
    bool have;
};
__device__ __forceinline__ void need_q(const csx_scalars& s, PixQ& p) {
    if (!p.have) {
        pixel_q(s, p.ihex, p.i, p.qx, p.qy, p.qz);
        p.have = true;
    }
}

// Pair records are read-only during a kernel and indexed wave-uniformly: through the constant
// address space the compiler loads them with scalar loads (into SGPRs, no VGPRs, no vector
// memory latency per pair) also when the table pointer itself was loaded from memory, as in
// the batched fit kernels.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* uniform_table(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}
__device__ __forceinline__ PairRec load_rec(const __attribute__((address_space(4))) PairRec* t, int i) {
    PairRec r;
    r.ux = t[i].ux;
    r.uy = t[i].uy;
    r.uz = t[i].uz;
    r.qh = t[i].qh;
    r.uz2 = t[i].uz2;
    r.A = t[i].A;
    r.ia = t[i].ia;
    r.ib = t[i].ib;
    return r;
}
__device__ __forceinline__ PairRec64 load_rec(const __attribute__((address_space(4))) PairRec64* t, int i) {
    PairRec64 r;
    r.ux = t[i].ux;
    r.uy = t[i].uy;
    r.uz = t[i].uz;
    r.A = t[i].A;
    r.qh = t[i].qh;
    r.ia = t[i].ia;
    r.ib = t[i].ib;
    r.pad = 0.0f;
    return r;
}

// Sum over pairs [p0, p1) (p0, p1 multiples of 4, wave-uniform) of A*(2^e1 [+ 2^e2]).
// y^2 = sin^2(y) * acos^2(c)/(1-c^2), with sin^2 y = |q^ x u|^2 taken from the cross
// product: accurate where the Gaussian is narrow (y small), where 1-|q^.u| is not.
// With CHECK, pairs with |q^.u| > kNearOne set `flag` and contribute 0 here (the caller adds
// their reference terms), or with NEAR as well, the pair formula at the reference's angle.
template <bool PLANAR, bool TWO, bool CHECK, bool NEAR = false>
__device__ __forceinline__ void pair_sum(const PairRec* __restrict__ prg, int p0, int p1, const PixelConsts& L,
                                         double& acc, bool& flag, const NearCtx* nc = nullptr) {
    // The four pairs of a group are independent chains, written stage by stage so that the
    // schedule interleaves them; the terms are added to `part` in pair order.
    const auto pr = uniform_table(prg);
    for (int p = p0; p < p1; p += 4) {
        PairRec P[4];
        float c[4], ss[4], y2[4], e1[4], w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) P[j] = load_rec(pr, p + j);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float dt = __builtin_fmaf(L.hy, P[j].uy, L.hx * P[j].ux);
            if constexpr (PLANAR) {
                const float k = __builtin_fmaf(L.hx, P[j].uy, -(L.hy * P[j].ux));
                ss[j] = __builtin_fmaf(k, k, P[j].uz2);
            } else {
                dt = __builtin_fmaf(L.hz, P[j].uz, dt);
                const float k1 = __builtin_fmaf(L.hy, P[j].uz, -(L.hz * P[j].uy));
                const float k2 = __builtin_fmaf(L.hz, P[j].ux, -(L.hx * P[j].uz));
                const float k3 = __builtin_fmaf(L.hx, P[j].uy, -(L.hy * P[j].ux));
                ss[j] = __builtin_fmaf(k1, k1, __builtin_fmaf(k2, k2, k3 * k3));
            }
            c[j] = fabsf(dt);
        }
        acos2_over_sin2_x4(c, y2);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            y2[j] = ss[j] * y2[j];
            const float t = (L.qf - P[j].qh) + L.qlo;
            const float r = (t * L.kr) * t;
            e1[j] = __builtin_fmaf(L.kp, y2[j], r);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] = __builtin_amdgcn_exp2f(e1[j]);
        if constexpr (TWO) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float y = __builtin_amdgcn_sqrtf(y2[j]);
                w[j] += __builtin_amdgcn_exp2f(__builtin_fmaf(L.c2, y, e1[j] + L.c1));
            }
        }
        float part = 0.0f;
        bool fl[4] = {false, false, false, false};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if constexpr (CHECK) {
                fl[j] = c[j] > kNearOne;
                flag = flag || fl[j];
                w[j] = fl[j] ? 0.0f : w[j];
            }
            part = __builtin_fmaf(P[j].A, w[j], part);
        }
        if constexpr (CHECK && NEAR) {
            // the pair formula at the reference's angle (see near_pair_x), after the group's
            // other pairs
            if (__ballot(fl[0] || fl[1] || fl[2] || fl[3])) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (fl[j]) {
                        const float x = (float)near_pair_x<PLANAR>(*nc, P[j].ia);
                        const float yy = x * __builtin_fmaf(x, 1.0f / 3.0f, 2.0f);
                        const float t = (L.qf - P[j].qh) + L.qlo;
                        const float ee = __builtin_fmaf(L.kp, yy, (t * L.kr) * t);
                        float wn = __builtin_amdgcn_exp2f(ee);
                        if constexpr (TWO)
                            wn += __builtin_amdgcn_exp2f(__builtin_fmaf(L.c2, __builtin_amdgcn_sqrtf(yy), ee + L.c1));
                        part = __builtin_fmaf(P[j].A, wn, part);
                    }
                }
            }
        }
        acc += (double)part;
    }
}

// FP64 terms (reference formula) of the pairs that pair_sum left out, for this lane.
template <bool PLANAR>
__device__ double near_one_terms(const DevSet& d, int p0, int p1, const PixelConsts& L, const PixQ& P3, double q,
                                 double width, double x2phi) {
    const double qx = P3.qx, qy = P3.qy, qz = P3.qz;
    double add = 0.0;
    const auto pr = uniform_table(d.pairs);
    for (int p = p0; p < p1; ++p) {
        const PairRec P = load_rec(pr, p);
        float dt = __builtin_fmaf(L.hy, P.uy, L.hx * P.ux);
        if constexpr (!PLANAR) dt = __builtin_fmaf(L.hz, P.uz, dt);
        if (fabsf(dt) > kNearOne)
            add += aniso_term_ref(d, P.ia, qx, qy, qz, q, width, x2phi) +
                   aniso_term_ref(d, P.ib, qx, qy, qz, q, width, x2phi);
    }
    return add;
}

// acos(1-x)^2 / x for x in [0, 1], degree-11 polynomial (Chebyshev-node interpolant of the
// function, which is analytic on [0, 2) with value 2 at x = 0); max relative error 2.4e-11.
__device__ __forceinline__ double acos2_over_x(double x) {
    double p = 5.8703930721466823e-04;
    p = fma(p, x, -2.2269864280509148e-03);
    p = fma(p, x, 4.3534013060061065e-03);
    p = fma(p, x, -4.3610819883067170e-03);
    p = fma(p, x, 3.8076579843268261e-03);
    p = fma(p, x, 9.4124350172984761e-05);
    p = fma(p, x, 4.2622239283164089e-03);
    p = fma(p, x, 1.0083515220786222e-02);
    p = fma(p, x, 2.8579409148493790e-02);
    p = fma(p, x, 8.8888454904565387e-02);
    p = fma(p, x, 3.3333334251157515e-01);
    p = fma(p, x, 1.9999999999680802e+00);
    return p;
}

// 2^e for e <= 0 in FP64 exponent arithmetic with an FP32 exp2 of the fraction: relative
// error about 1 ulp of FP32 whatever |e| (the FP32 loop's error grows with |e|).
__device__ __forceinline__ double exp2_split(double e) {
    const double n = rint(e);
    const float f = (float)(e - n);  // |f| <= 1/2
    return (double)ldexpf(__builtin_amdgcn_exp2f(f), (int)fmax(n, -200.0));
}

// Per-pixel constants of the precise pair loop.
struct PixelConsts64 {
    double hx, hy, hz;  // q / |q| (|q| with the reference's +1e-9)
    double q, kr, kp;   // |q|, radial and angular base-2 exponent coefficients
};

// The precise pair loop: the FP32 loop's sum with FP64 exponents.  The reference's own
// cosine c = q.q_hkl / (|q| |q_hkl|) is formed in FP64 and y^2 = acos(c)^2 = x*g(x), x = 1-c,
// needs no cancellation-free rewrite in FP64.  Same window, same near-one rule.
template <bool PLANAR, bool TWO, bool CHECK, bool NEAR = false>
__device__ __forceinline__ void pair_sum64(const PairRec64* __restrict__ prg, int p0, int p1, const PixelConsts64& L,
                                           double& acc, bool& flag, const NearCtx* nc = nullptr) {
    constexpr double kPi = 3.14159265358979323846;
    const auto pr = uniform_table(prg);
    for (int p = p0; p < p1; ++p) {
        const PairRec64 P = load_rec(pr, p);
        double dt = fma(L.hy, P.uy, L.hx * P.ux);
        if constexpr (!PLANAR) dt = fma(L.hz, P.uz, dt);
        const double c = fabs(dt);
        const double x = 1.0 - c;
        const double y2 = x * acos2_over_x(x);
        const double t = L.q - (double)P.qh;
        const double r = (t * L.kr) * t;
        const double e1 = fma(L.kp, y2, r);
        double w = exp2_split(e1);
        if constexpr (TWO) {
            // partner at pi - y: e2 = e1 + kp*pi*(pi - 2y)
            const double y = (double)__builtin_amdgcn_sqrtf((float)y2);
            w += exp2_split(fma(L.kp * kPi, kPi - 2.0 * y, e1));
        }
        if constexpr (CHECK) {
            const bool fl = c > (double)kNearOne;
            flag = flag || fl;
            w = fl ? 0.0 : w;
            if constexpr (NEAR) {
                if (fl) {
                    const double xx = near_pair_x<PLANAR>(*nc, P.ia);
                    const double yy = xx * acos2_over_x(xx);
                    const double tt = L.q - (double)P.qh;
                    const double ee = fma(L.kp, yy, (tt * L.kr) * tt);
                    w = exp2_split(ee);
                    if constexpr (TWO) {
                        const double y = (double)__builtin_amdgcn_sqrtf((float)yy);
                        w += exp2_split(fma(L.kp * kPi, kPi - 2.0 * y, ee));
                    }
                }
            }
        }
        acc = fma(P.A, w, acc);
    }
}

// near_one_terms for the precise loop (its cosine test in FP64).
template <bool PLANAR>
__device__ double near_one_terms64(const DevSet& d, int p0, int p1, const PixelConsts64& L, const PixQ& P3, double q,
                                   double width, double x2phi) {
    double add = 0.0;
    const auto pr = uniform_table(d.pairs64);
    for (int p = p0; p < p1; ++p) {
        const PairRec64 P = load_rec(pr, p);
        double dt = fma(L.hy, P.uy, L.hx * P.ux);
        if constexpr (!PLANAR) dt = fma(L.hz, P.uz, dt);
        if (fabs(dt) > (double)kNearOne)
            add += aniso_term_ref(d, P.ia, P3.qx, P3.qy, P3.qz, q, width, x2phi) +
                   aniso_term_ref(d, P.ib, P3.qx, P3.qy, P3.qz, q, width, x2phi);
    }
    return add;
}

// Pair-window bookkeeping of one wave: |q| bins [b0, b1] and |u_z| bands [band0, band1].
struct Window {
    int b0, b1, band0, band1;
    bool near;  // some pair of this lane's window has |q^.u| > kNearOne
};

// q^ of the FP32 pair loop: from the exact q vector when the caller has it, else (midpoint
// grid) q^ = (i, ihex) * (qmax/zmax) / |q| in FP32 (q_z = 1e-20 drops out).
__device__ __forceinline__ void pixel_dirs(const csx_scalars& s, const PixQ& P3, double q, PixelConsts& L) {
    if (P3.have) {
        const double rq = 1.0 / q;
        L.hx = (float)(P3.qx * rq);
        L.hy = (float)(P3.qy * rq);
        L.hz = (float)(P3.qz * rq);
    } else {
        const float c = (float)(s.qmax / s.zmax) * __builtin_amdgcn_rcpf((float)q);
        L.hx = (float)P3.i * c;
        L.hy = (float)P3.ihex * c;
        L.hz = 0.0f;
    }
}
__device__ __forceinline__ void pixel_dirs64(const csx_scalars& s, PixQ& P3, double q, PixelConsts64& L64) {
    need_q(s, P3);
    const double rq = 1.0 / q;
    L64.hx = P3.qx * rq;
    L64.hy = P3.qy * rq;
    L64.hz = P3.qz * rq;
}

// A pixel that the hot image pass leaves to the fix pass: its FP32-loop peak sum without the
// near-one terms, the bound of its dropped mass and its wave's window.
struct ColdPix {
    int ihex, i;
    double inten, dropped;
    int b0, b1, band0, band1;
    int flags;  // 1: near-one pairs in the window, 2: unpaired reflections, 4: placeholder lane
    int pad;
};

// The fix pass's queue: one slot per hot wave with cold pixels, slot k holding the wave's
// lane mask and its cold pixels at ent[64 k + lane].  A render uses count[epoch] and clears
// the other count for the next render (kernels of one stream run in order).
struct ColdQueue {
    int* count;
    unsigned long long* mask;
    ColdPix* ent;
    int* next_count;
};

// How a pixel is evaluated.  kFull: every path in one pass.  kHot: the common paths only;
// a pixel that needs an FP64 path (near-one pairs, the conditioning guard, unpaired
// reflections) is reported `cold` and its whole wave is evaluated again by kFix, which runs
// every path and does not repeat the pair-evaluation counts.
enum PixMode { kFull = 0, kHot = 1, kFix = 2 };

// Adds to inten, band by band, the FP64 terms (the reference's formula) of the window's pairs
// with |q^.u| > kNearOne, which the pair loop left out: L (FP32 loop) or L64 (precise loop)
// holds this pixel's q^.
template <bool PLANAR, bool PRECISE>
__device__ __forceinline__ void near_one_add(const DevSet& d, const Window& W, const PixelConsts& L,
                                             const PixelConsts64& L64, PixQ& P3, double q, double width,
                                             double& inten) {
    need_q(d.s, P3);
    const double x2phi = 4 * q * q / (kPiD * sqd(d.s.phiwidth));  // the reference's expression
    for (int band = W.band0; band <= W.band1; ++band) {
        const int32_t* lb = d.lut + band * (d.nlut + 1);
        const int p0 = lb[W.b0] & ~3, p1 = (lb[W.b1] + 3) & ~3;
        if constexpr (PRECISE) inten += near_one_terms64<PLANAR>(d, p0, p1, L64, P3, q, width, x2phi);
        else inten += near_one_terms<PLANAR>(d, p0, p1, L, P3, q, width, x2phi);
    }
}

// Mixed-precision anisotropic peak sum (sum over reflections of sq*mhkl*fhkl*psi).  Must be
// called by all 64 lanes of a wave: the wave's union of per-lane windows selects the pairs.
// Pairs outside the window, and second Friedel terms where they cannot reach 2^cut, are
// dropped: every dropped term is below 2^cut * A.  need_exact: unpaired reflections exist.
template <bool PLANAR, bool PRECISE, int MODE = kFull>
__device__ __forceinline__ double lattice_mixed(const DevSet& d, PixQ& P3, double q, double width, bool& need_exact,
                                double& dropped, Window& W, unsigned long long* stats) {
    const csx_scalars& s = d.s;
    PixelConsts L;
    pixel_dirs(s, P3, q, L);
    L.qf = (float)q;
    L.qlo = (float)(q - (double)L.qf);
    const double kpd = d.kp_coef * q * q;
    const double krd = s.para ? -4.0 / (kPiD * width * width) * kLog2e : d.kr_const;
    const double inv_hpw = s.para ? 1.0 / (kPiD * width / 2.0) : d.inv_hpw;
    L.kr = (float)krd;
    L.kp = (float)kpd;
    L.c1 = (float)(kpd * kPiD * kPiD);
    L.c2 = (float)(-2.0 * kPiD * kpd);
    const float cut = d.cut;
    // Window edges are bounds, computed in FP32 with a relative margin of 1e-4.
    // radial window: |q - q_hkl| <= dq  <=>  radial exponent >= cut
    const float dq = __builtin_amdgcn_sqrtf(cut / L.kr) * 1.0001f + 1e-6f;
    // angular window: y <= ymax  <=>  angular exponent >= cut (y <= pi/2 always); only
    // evaluated when some lane's angular factor can drop below 2^cut
    const float hpi = 1.5707963f;
    const bool narrow = L.kp * (hpi * hpi) < cut;
    const bool any_narrow = __builtin_amdgcn_readfirstlane((int)(__ballot(narrow) != 0ull)) != 0;
    float uz_lo = 0.0f, uz_hi = 1.0f;
    if (any_narrow) {
        const float ymax = narrow ? __builtin_amdgcn_sqrtf(cut / L.kp) * 1.0001f + 1e-6f : hpi;
        if constexpr (PLANAR) {
            uz_hi = (ymax < hpi) ? sinf(ymax) * 1.0001f + 1e-6f : 1.0f;
        } else {
            // angle(q^, +-u) >= | |asin q^_z| - |asin u_z| |
            const float eh = asinf(fminf(fabsf(L.hz), 1.0f));
            uz_lo = (eh - ymax > 0.0f) ? sinf(eh - ymax) * 0.9999f - 1e-6f : 0.0f;
            uz_hi = (eh + ymax < hpi) ? sinf(eh + ymax) * 1.0001f + 1e-6f : 1.0f;
        }
    }
    // second Friedel term: e2 <= kp*(pi/2)^2 for every pair
    const bool two = L.kp * (hpi * hpi) > cut;
    {
        const float lo = wave_min_f(L.qf - dq), hi = wave_max_f(L.qf + dq);
        const float ulo = any_narrow ? wave_min_f(uz_lo) : 0.0f;
        const float uhi = any_narrow ? wave_max_f(uz_hi) : 1.0f;
        // Bound of this lane's dropped mass (sum of A * 2^e over dropped terms).  Pairs left
        // out by the |q| window lie below lo or above hi, and inside [qh_min, qh_max].
        double dmin = HUGE_VAL;
        if (d.qh_min < lo) dmin = fmin(dmin, q - fmin((double)lo, (double)d.qh_max));
        if (d.qh_max > hi) dmin = fmin(dmin, fmax((double)hi, (double)d.qh_min) - q);
        dmin = fmax(dmin, 0.0);
        // 2^max(e, -126) in FP32: an upper bound of 2^e, cheap, never denormal
        dropped = (dmin < HUGE_VAL)
                      ? d.asum2 * (double)__builtin_amdgcn_exp2f(fmaxf((float)(krd * dmin * dmin), -126.0f))
                      : 0.0;
        W.b0 = uniform_i(max(0, min(d.nlut, (int)floorf(lo * d.lut_inv_dq))));
        W.b1 = uniform_i(max(0, min(d.nlut, (int)ceilf(hi * d.lut_inv_dq))));
        W.band0 = uniform_i(max(0, min(kBands - 1, (int)floorf(ulo * kBands))));
        W.band1 = uniform_i(max(0, min(kBands - 1, (int)floorf(uhi * kBands))));
    }
    const bool any_two = __builtin_amdgcn_readfirstlane((int)(__ballot(two) != 0ull)) != 0;
    if (W.band0 > 0 || W.band1 < kBands - 1) {
        dropped += d.asum2 * (double)__builtin_amdgcn_exp2f(cut);
    } else {
        W.band0 = W.band1 = kBands;  // every band: the merged list, in one pass
    }
    if (!any_two)
        dropped += 0.5 * d.asum2 * (double)__builtin_amdgcn_exp2f(fmaxf(L.kp * (hpi * hpi), -126.0f));
    PixelConsts64 L64;
    if constexpr (PRECISE) {
        pixel_dirs64(s, P3, q, L64);
        L64.q = q;
        L64.kr = krd;
        L64.kp = kpd;
    }
    double acc = 0.0;
    bool flag = false;
    long long evals = 0;
    // The near-one test is needed only where some pair of the window can have |q^.u| >
    // kNearOne.  In the detector plane those pairs lie in the in-plane segment: test them first
    // (once per wave, on the records and in the arithmetic of the pair loop that follows: the
    // FP64 records for the precise loop, which are the only ones its tables hold), and drop
    // the test from the pair loop when none is near one.  The loop's 4-aligned range also
    // takes pairs just outside the window; their terms are below 2^cut * A for every lane,
    // like the dropped ones.
    bool check = true;
    if constexpr (PLANAR) {
        check = false;
        if (W.band0 == 0 || W.band0 == kBands) {
            const int32_t* lp = d.lut + (kBands + 1) * (d.nlut + 1);
            const int q0 = uniform_i(lp[W.b0]), q1 = uniform_i(lp[W.b1]);
            bool hit = false;
            if constexpr (PRECISE) {
                const auto pr = uniform_table(d.pairs64);
                for (int p = q0; p < q1; ++p) {
                    const double dt = fma(L64.hy, pr[p].uy, L64.hx * pr[p].ux);
                    hit = hit || fabs(dt) > (double)kNearOne;
                }
            } else {
                const auto pr = uniform_table(d.pairs);
                for (int p = q0; p < q1; ++p) {
                    const float dt = __builtin_fmaf(L.hy, pr[p].uy, L.hx * pr[p].ux);
                    hit = hit || fabsf(dt) > kNearOne;
                }
            }
            check = __ballot(hit) != 0ull;
        }
    }
    // The hot image pass evaluates the near-one pairs inside the loop in the detector plane;
    // with a centre beam it leaves them to the fix pass (whose pixel q needs the FP64 beam
    // geometry).
    constexpr bool kNear = MODE == kHot && PLANAR;
    const NearCtx nc{&d, P3.ihex, P3.i, q};
    for (int band = W.band0; band <= W.band1; ++band) {
        const int32_t* lb = d.lut + band * (d.nlut + 1);
        const int p0 = uniform_i(lb[W.b0] & ~3);
        const int p1 = uniform_i((lb[W.b1] + 3) & ~3);
        if constexpr (PRECISE) {
            if (check) {
                if (any_two) pair_sum64<PLANAR, true, true, kNear>(d.pairs64, p0, p1, L64, acc, flag, &nc);
                else pair_sum64<PLANAR, false, true, kNear>(d.pairs64, p0, p1, L64, acc, flag, &nc);
            } else {
                if (any_two) pair_sum64<PLANAR, true, false>(d.pairs64, p0, p1, L64, acc, flag);
                else pair_sum64<PLANAR, false, false>(d.pairs64, p0, p1, L64, acc, flag);
            }
        } else {
            if (check) {
                if (any_two) pair_sum<PLANAR, true, true, kNear>(d.pairs, p0, p1, L, acc, flag, &nc);
                else pair_sum<PLANAR, false, true, kNear>(d.pairs, p0, p1, L, acc, flag, &nc);
            } else {
                if (any_two) pair_sum<PLANAR, true, false>(d.pairs, p0, p1, L, acc, flag);
                else pair_sum<PLANAR, false, false>(d.pairs, p0, p1, L, acc, flag);
            }
        }
        evals += p1 - p0;
    }
    double inten = acc * inv_hpw;
    W.near = !kNear && flag;
    if (MODE != kHot && flag) near_one_add<PLANAR, PRECISE>(d, W, L, L64, P3, q, width, inten);
    if (MODE != kFix && stats && (threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], (unsigned long long)(64 * evals));
        atomicAdd(&stats[3], (unsigned long long)(any_two ? 1 : 0));
    }
    if (MODE != kFix && stats && flag) atomicAdd(&stats[2], 1ull);
    need_exact = d.nsingle > 0;
    return inten;
}

// Wave-cooperative FP64 re-sums.  Lanes whose FP32 peak sum is too coarse are taken one at
// a time; the 64 lanes share that lane's reflections (stride 64) and a butterfly sum gives
// the total, so a re-sum costs the wave about 1/64 of a serial one instead of stalling it.
// Called by all 64 lanes; `want` is the per-lane request, `inten`/`radint` are replaced
// for the requesting lanes only.
struct PixArgs {
    double qx, qy, qz, q, width;
};
__device__ __forceinline__ PixArgs shfl_args(const PixArgs& a, int lane) {
    return PixArgs{__shfl(a.qx, lane, 64), __shfl(a.qy, lane, 64), __shfl(a.qz, lane, 64), __shfl(a.q, lane, 64),
                   __shfl(a.width, lane, 64)};
}

// Tier 2: both reflections of every pair in the wave's window, with the reference's
// per-reflection formula (acos mapping included).  Differs from the reference's sum by the
// order of the terms and by the window's dropped mass only.
template <bool PRECISE>
__device__ void coop_window_sum64(const DevSet& d, const Window& W, const PixArgs& me, bool want, double& inten) {
    unsigned long long m = __ballot(want);
    const int lane = threadIdx.x & 63;
    const double x2c = 4 / (kPiD * sqd(d.s.phiwidth));
    while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1;
        const PixArgs a = shfl_args(me, l);
        const double x2phi = x2c * a.q * a.q;
        // the requesting lane's window (the same for every lane of an image wave)
        const int b0 = uniform_i(__shfl(W.b0, l, 64)), b1 = uniform_i(__shfl(W.b1, l, 64));
        const int band0 = uniform_i(__shfl(W.band0, l, 64)), band1 = uniform_i(__shfl(W.band1, l, 64));
        double acc = 0.0;
        for (int band = band0; band <= band1; ++band) {
            const int32_t* lb = d.lut + band * (d.nlut + 1);
            const int p0 = uniform_i(lb[b0] & ~3), p1 = uniform_i((lb[b1] + 3) & ~3);
            for (int p = p0 + lane; p < p1; p += 64) {
                // the precise loop's kinds carry only the FP64 records
                const int ia = PRECISE ? d.pairs64[p].ia : d.pairs[p].ia;
                const int ib = PRECISE ? d.pairs64[p].ib : d.pairs[p].ib;
                if (ia == 0) continue;  // padding record
                acc += aniso_term_ref(d, ia, a.qx, a.qy, a.qz, a.q, a.width, x2phi) +
                       aniso_term_ref(d, ib, a.qx, a.qy, a.qz, a.q, a.width, x2phi);
            }
        }
        acc = wave_sum(acc);
        if (lane == l) inten = acc;
    }
}

// Tier 3: the full reflection lists (lattice_ref64's terms).  A non-finite total takes the
// serial reference loop, which stops at the first infinite partial sum.
__device__ void coop_full_sum64(const DevSet& d, const PixArgs& me, bool want, double& inten, double& radint) {
    unsigned long long m = __ballot(want);
    const int lane = threadIdx.x & 63;
    const csx_scalars& s = d.s;
    const double widthiso = 1.0 / s.uca;
    const double x2c = 4 / (kPiD * sqd(s.phiwidth));
    while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1;
        const PixArgs a = shfl_args(me, l);
        const double x2phi = x2c * a.q * a.q;
        double acc = 0.0, rad = 0.0;
        for (int ii = 1 + lane; ii <= s.peakmax2; ii += 64)
            acc += aniso_term_ref(d, ii, a.qx, a.qy, a.qz, a.q, a.width, x2phi);
        if (s.iso != 0) {
            for (int ii = 1 + lane; ii <= s.peakmax1; ii += 64) {
                const int mhkl = d.latpar1[ii * 6 + 4], fhkl = d.latpar1[ii * 6 + 5];
                if (mhkl == 0 || fhkl == 0) continue;
                const double qhkl = d.latpar3[ii * 17 + 5];
                const double x2 = sqd(a.q - qhkl) / sqd(widthiso);
                rad += exp(-4 * x2 / kPiD) / (kPiD * widthiso / 2.0) * mhkl * fhkl;
            }
        }
        acc = wave_sum(acc);
        rad = wave_sum(rad);
        if (lane == l) {
            inten = acc;
            radint = rad;
        }
    }
    if (want && !isfinite(inten)) lattice_ref64(d, me.qx, me.qy, me.qz, me.q, me.width, inten, radint);
}

// Isotropic (radial) peak sum in FP32, only needed when iso != 0.
__device__ double iso_mixed(const DevSet& d, double q) {
    const double wi = 1.0 / d.s.uca;
    const float ki = (float)(-4.0 / (kPiD * wi * wi) * kLog2e);
    const float qf = (float)q;
    const float qlo = (float)(q - (double)qf);
    float ra = 0.0f;
    for (int k = 0; k < d.niso; ++k) {
        const float2 e = d.iso[k];
        const float t = (qf - e.x) + qlo;
        ra = __builtin_fmaf(e.y, __builtin_amdgcn_exp2f((t * ki) * t), ra);
    }
    return (double)ra / (kPiD * wi / 2.0);
}

// ---------------------------------------------------------------------------
// Gaussian-distributed cylinders: the longitudinal orientation average
// ---------------------------------------------------------------------------

// Rod factor as the chi integrand evaluates it (Source/sc_calc_generic_gpu.cu:1640-1652).
__device__ __forceinline__ double chi_rod(double z, double arg) {
    const double a1 = 1 / (2.0 * z * (z - 1));
    return (a1 / (arg * arg)) * (1 - cos((z - 1) * atan(2.0 * arg)) / pow(1.0 + 4 * arg * arg, (z - 1) / 2.0));
}

struct ChiCtx {
    double l, z;
    float p11, p12, p13, p21, p22, p23;  // float-rounded rotation entries (.cu:1402)
    double qx, qy;
    int i1;
};

// trapzdchid integrand for the cylinder cases i1 = 1..4 (.cu:1634-1720); other cases give 0.
__device__ double chi_integrand(const ChiCtx& c, double x, double delta) {
    const double eps = 0.0000000001;
    double arg;
    switch (c.i1) {
    case 1: {
        const double ml1 = c.p11 * cos(x) * sin(delta) + c.p12 * sin(x) * sin(delta) + c.p13 * cos(delta);
        const double ml2 = c.p21 * sin(x) * sin(delta) + c.p22 * cos(x) * sin(delta) + c.p23 * cos(delta);
        arg = (c.qx * ml1 + c.qy * ml2 + eps) * c.l / (c.z + 1);
        break;
    }
    case 2: arg = (c.qx * cos(delta) - c.qy * sin(x) * sin(delta) + eps) * c.l / (c.z + 1); break;
    case 3: arg = (c.qx * sin(x) * sin(delta) + c.qy * cos(delta) + eps) * c.l / (c.z + 1); break;
    case 4: arg = (c.qx * sin(x) * sin(delta) - c.qy * cos(x) * sin(delta) + eps) * c.l / (c.z + 1); break;
    default: return 0.0;
    }
    return chi_rod(c.z, arg);
}

// Neville extrapolation to h = 0 through k = 5 points (polint, .cu:1280-1335); 1-based.
__device__ void neville5(const double* xa, const double* ya, double& y, double& dy) {
    double c[6], d[6];
    int ns = 1;
    double dif = fabs(xa[1]);
    for (int i = 1; i <= 5; ++i) {
        const double dift = fabs(xa[i]);
        if (dift < dif) { ns = i; dif = dift; }
        c[i] = ya[i];
        d[i] = ya[i];
    }
    y = ya[ns];
    ns = ns - 1;
    for (int m = 1; m <= 4; ++m) {
        for (int i = 1; i <= 5 - m; ++i) {
            const double ho = xa[i], hp = xa[i + m];
            const double w = c[i + 1] - d[i];
            double den = ho - hp;
            if (den != 0) den = w / den;
            d[i] = hp * den;
            c[i] = ho * den;
        }
        if (2 * ns < 5 - m) dy = c[ns + 1];
        else { dy = d[ns]; ns = ns - 1; }
        y = y + dy;
    }
}

// Romberg over chi in [0, float(2 pi)] (qrombchid .cu:1495-1583 with trapzdchid's
// midpoint refinement): eps 1e-4, jmax 20, k 5.
__device__ double romb_chi(const ChiCtx& c, double delta) {
    const double eps = 1.0e-4;
    double hp[22], sp[22];
    const double a = 0.0, b = (double)(float)(2 * kPiD);
    double pq = 0, dssp = 0;
    int cnt = 1;
    hp[1] = 1.0;
    for (int j = 1; j <= 20; ++j) {
        if (j == 1) {
            sp[j] = 0.5 * (b - a) * (chi_integrand(c, a, delta) + chi_integrand(c, b, delta));
            cnt = 1;
        } else {
            const double tnm = cnt, del = (b - a) / tnm;
            double x = a + 0.5 * del, sum = 0.0;
            for (int i = 1; i <= cnt; ++i) { sum += chi_integrand(c, x, delta); x = x + del; }
            sp[j] = 0.5 * (sp[j] + (b - a) * sum / tnm);
            cnt = 2 * cnt;
        }
        if (j >= 5) {
            neville5(hp + j - 5, sp + j - 5, pq, dssp);
            if (fabs(dssp) < (eps * fabs(pq))) break;
            if (fabs(pq) < 1.0e-300) break;
        }
        sp[j + 1] = sp[j];
        hp[j + 1] = 0.25 * hp[j];
    }
    return pq;
}

// Romberg over delta (qrombdeltac .cu:1370-1488, i0 = 4, Gaussian weight) of
// sin(d) exp(-d^2/dbeta^2) * romb_chi / (2 pi), in [0, float(min(dbeta sqrt(ln 1e5), pi/2))].
__device__ double romb_delta_cyl(const DevSet& d, double qx, double qy, int i1) {
    const csx_scalars& s = d.s;
    const double eps = 1.0e-4, prmin = 0.00001;
    double hp[22], sp[22];
    const double dbeta = s.dbeta * kPiD / 180.0;
    const double th = s.polTheta * kPiD / 180.0, ph = 0.0;
    double delmax = dbeta * sqrt(log(1.0 / prmin));
    if (delmax > kPiD / 2.0) delmax = kPiD / 2.0;
    ChiCtx c;
    c.l = s.length;
    c.z = (1 - s.sigmal * s.sigmal) / (s.sigmal * s.sigmal);
    c.p11 = (float)(-cos(ph) * cos(th));
    c.p12 = (float)sin(ph);
    c.p13 = (float)(cos(ph) * sin(th));
    c.p21 = (float)(-cos(ph));
    c.p22 = (float)(-sin(ph) * cos(th));
    c.p23 = (float)(sin(ph) * sin(th));
    c.qx = qx;
    c.qy = qy;
    c.i1 = i1;
    const double a = 0.0, b = (double)(float)delmax;
    auto f = [&](double x) {
        return sin(x) * exp(-x * x / (dbeta * dbeta)) * romb_chi(c, x) / (2.0 * kPiD);
    };
    double pq = 0, dssp = 0;
    int cnt = 1;
    hp[1] = 1.0;
    for (int j = 1; j <= 20; ++j) {
        if (j == 1) {
            sp[j] = 0.5 * (b - a) * (f(a) + f(b));
            cnt = 1;
        } else {
            const double tnm = cnt, del = (b - a) / tnm;
            double x = a + 0.5 * del, sum = 0.0;
            for (int i = 1; i <= cnt; ++i) { sum = sum + f(x); x = x + del; }
            sp[j] = 0.5 * (sp[j] + (b - a) * sum / tnm);
            cnt = 2 * cnt;
        }
        if (j >= 5) {
            neville5(hp + j - 5, sp + j - 5, pq, dssp);
            if (fabs(dssp) < (eps * fabs(pq))) break;
            if (fabs(pq) < 1.0e-100) break;
        }
        sp[j + 1] = sp[j];
        hp[j + 1] = 0.25 * hp[j];
    }
    return pq;
}

// Small-q double series of the Gaussian orientation average (formpq_partCylinder ordis 0,
// Source/sc_lib_formpq_partCylinder.h:133-203), orientation cases 1..3:
//   pql = 1 + sum_n c1p[n] sum_m c11pm[n-m][m] X^m Y^(n-m)
// with (X, Y) = (qxs^2, qys^2), swapped for case 3.  The inner sum is evaluated by Horner in
// the ratio of the smaller to the larger of X and Y (no power tables; the reference builds
// X^m and Y^(n-m) as running products, so the two agree to rounding).
__device__ double cyl_gauss_series(const DevSet& d, double qxs, double qys, int oc) {
    const double* c1p = carr_of<0>(d);
    const double* c11 = d.carr11pm;  // [n][m], CSX_NIM2 columns
    const double X = (oc == 3) ? qys * qys : qxs * qxs, Y = (oc == 3) ? qxs * qxs : qys * qys;
    const bool xs = X <= Y;            // Horner variable r = small/large <= 1
    const double big = xs ? Y : X, r = xs ? (Y > 0 ? X / Y : 0.0) : (X > 0 ? Y / X : 0.0);
    double pql = 1.0, old = 0.0, bn = 1.0;
#pragma unroll 1
    for (int n = 1; n <= 120; ++n) {
        bn *= big;
        // sum_m c11[n-m][m] X^m Y^(n-m) = big^n * sum_k a_k r^k, with k = m (X small) or n-m
        double h = xs ? c11[0 * CSX_NIM2 + n] : c11[n * CSX_NIM2 + 0];
#pragma unroll 1
        for (int k = n - 1; k >= 0; --k) {
            const double a = xs ? c11[(n - k) * CSX_NIM2 + k] : c11[k * CSX_NIM2 + (n - k)];
            h = h * r + a;
        }
        pql += c1p[n] * (h * bn);
        const double dd = fabs((pql - old) / pql);
        if (dd < 0.0001) break;
        old = pql;
    }
    return pql;
}

// LDS copies of the physics-mode orientation quadrature tables (filled by k_image for the
// Gaussian-cylinder kernels; 33.9 KB).  File-scope LDS, so that every access compiles to a
// ds_read with a 32-bit offset.
__shared__ float2 g_cq_tab[kCqTab + 1];
__shared__ float4 g_cq_dl[kCqDelta];
__shared__ float g_cq_tt[kCqPsi / 2];

// Rod factor P(|a|) from the LDS table: cubic Hermite on the knot interval (value and
// h*slope at both knots).  |a| is below amax by construction; the clamp only keeps the
// index in range.
__device__ __forceinline__ float cq_rod(float invh, float a) {
    const float u = fminf(fabsf(a) * invh, (float)kCqTab - 0.001f);
    const int i = (int)u;
    const float t = __builtin_amdgcn_fractf(u);
    const float2 A = g_cq_tab[i], B = g_cq_tab[i + 1];
    const float df = B.x - A.x;
    const float c2 = __builtin_fmaf(3.0f, df, -(2.0f * A.y + B.y));
    const float c3 = (A.y + B.y) - 2.0f * df;
    return __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, c3, c2), A.y), A.x);
}

// Physics-mode orientation average of the longitudinal factor, FP32 with a compensated sum
// over the polar nodes.  For a fixed polar angle delta the chi integrand depends on chi only
// through arg = alpha + R cos(psi) (psi = chi - chi0; the integral is shift-invariant over a
// period), so (1/2pi) int_0^2pi dchi = (1/pi) int_0^pi dpsi, taken as a kCqPsi-point midpoint
// rule whose nodes come in pairs +-t.  Polar: kCqDelta-point Gauss-Legendre with the weight
// sin(d) exp(-d^2/dbeta^2), normalised to 1 on the host.
__device__ double cyl_gauss_quad(const DevSet& d, double qa, double qb, int oc) {
    double c1, rq;
    if (oc == 1) {
        const double p11 = (double)(float)(-cos(d.s.polTheta * kPiD / 180.0));
        const double p13 = (double)(float)(sin(d.s.polTheta * kPiD / 180.0));
        // phi = 0 in the reference's call: p12 = p22 = p23 = 0, p21 = -1
        c1 = qa * p13;
        rq = sqrt(sqd(qa * p11) + sqd(qb * -1.0));
    } else if (oc == 2) {
        c1 = qa;
        rq = fabs(qb);
    } else {
        c1 = qb;
        rq = fabs(qa);
    }
    const float lz = d.cq_lz, invh = d.cq_invh;
    const float fc1 = (float)c1, frq = (float)rq;
    float sum = 0.0f, comp = 0.0f;
    for (int k = 0; k < kCqDelta; ++k) {
        const float4 nd = g_cq_dl[k];  // cos, sin, weight
        const float al = __builtin_fmaf(nd.x, fc1, 1.0e-10f) * lz;
        const float rr = nd.y * frq * lz;
        float sk = 0.0f;
#pragma unroll 4
        for (int j = 0; j < kCqPsi / 2; ++j) {
            const float t = g_cq_tt[j];
            sk += cq_rod(invh, __builtin_fmaf(rr, t, al)) + cq_rod(invh, __builtin_fmaf(-rr, t, al));
        }
        const float y = __builtin_fmaf(nd.z, sk, -comp);
        const float tnew = sum + y;
        comp = (tnew - sum) - y;
        sum = tnew;
    }
    return (double)sum;
}

// ---------------------------------------------------------------------------
// Pixel value of each combination
// ---------------------------------------------------------------------------
// P3: the pixel (its exact q vector is computed here unless the caller has it); rv: its
// radial-table entry (midpoint grid, sphere kinds) or null.
// MODE kHot: *cold reports a pixel that needs an FP64 path, with its state in *cp; kFix:
// finishes such a pixel from *cp.
template <int KIND, int PREC, int MODE = kFull>
__device__ __forceinline__ double pixel_value(const DevSet& d, PixQ& P3, unsigned long long* stats = nullptr,
                                              const Radial* rv = nullptr, bool* cold = nullptr,
                                              ColdPix* cp = nullptr) {
    const csx_scalars& s = d.s;
    // The mixed FCC path with a radial entry needs only |q| and the pixel indices.
    if (!(rv && (KIND == K_FCC || KIND == K_ISO_LAT) && PREC == CSX_PREC_MIXED)) need_q(s, P3);
    const double qx = P3.qx, qy = P3.qy, qz = P3.qz;
    const double q = rv ? rv->q : sqrt(qx * qx + qy * qy + qz * qz) + 0.000000001;
    if constexpr (KIND == K_SPHERE) {
        const double pq = rv ? rv->pq : pq_sphere(d, q);
        return s.base + s.i0 * (pq + s.iso * pq) + s.ifluc / (1 + q * q * s.rfluc * s.rfluc);
    } else if constexpr (KIND == K_CUBE) {
        // Source/sc_gpu_pCub_lNon_oIso.h:14-57: no lattice, iso added to P(q) unweighted
        const double pq = rv ? rv->pq : pq_cube(d, q);
        return s.base + s.i0 * (pq + s.iso) + s.ifluc / (1 + q * q * s.rfluc * s.rfluc);
    } else if constexpr (KIND == K_CUBE_ZDIR) {
        // calc_GENERIC, cubes, perfect orientation (Source/sc_gpu_generic.h:307-314, 868-880)
        const double pq = pq_cube_zdir(d, qx, qy, q);
        return s.base + s.i0 * (pq + s.iso) + s.ifluc / (1 + q * q * s.rfluc * s.rfluc);
    } else if constexpr (KIND == K_CYL_ISO) {
        // calc_GENERIC, cylinders, isotropic (Source/sc_gpu_generic.h:58-66, 868-880)
        const double pq = rv ? rv->pq : pq_cyl_long_iso(d, q) * pq_cyl_transverse(d, q);
        return s.base + s.i0 * (pq + s.iso) + s.ifluc / (1 + q * q * s.rfluc * s.rfluc);
    } else if constexpr (KIND == K_CYL_ZDIR) {
        double pq;
        switch (s.orcase) {
        case 0: pq = pq_cyl_long_zdir(d, qx * s.cosphi * s.sintheta, qy * s.sinphi * s.sintheta, q); break;
        case 1: pq = pq_cyl_long_zdir(d, qx, 0, q); break;
        case 2: pq = pq_cyl_long_zdir(d, 0, qy, q); break;
        default: pq = pq_cyl_long_zdir(d, qx, qy, q); break;
        }
        pq *= pq_cyl_transverse(d, q);
        return s.base + s.i0 * (pq + s.iso) + s.ifluc / (1 + q * q * s.rfluc * s.rfluc);
    } else if constexpr (KIND == K_CYL_GAUSS || KIND == K_CYL_GAUSS_Q) {
        // Gaussian orientation (Source/sc_gpu_pCyl_lNon_oGau.h:36-89, formpq_partCylinder
        // ordis 0).  The general case uses q rotated by (cosphic, sinphic).
        double qa = qx, qb = qy;
        if (s.orcase == 0) {
            qa = qx * s.cosphic - qy * s.sinphic;
            qb = qx * s.sinphic + qy * s.cosphic;
        }
        const int oc = orc_cmp(s);
        double pql;
        if (oc == 4) {
            pql = 1.0;
        } else if (q < (0.4 * s.limq[0])) {
            pql = (oc >= 1 && oc <= 3) ? cyl_gauss_series(d, qa, qb, oc) : 1.0;
        } else if constexpr (KIND == K_CYL_GAUSS_Q) {
            pql = cyl_gauss_quad(d, qa, qb, oc);  // physics mode: oc in 1..3, tables in LDS
        } else {
            // the reference's nested Romberg; in reference mode the general case (oc = 0)
            // has an identically zero integrand
            pql = (oc >= 1 && oc <= 4) ? romb_delta_cyl(d, qa, qb, oc) / s.norm : 0.0;
        }
        const double pq = pql * pq_cyl_transverse(d, q);
        return s.base + s.i0 * (pq + s.iso) + s.ifluc / (1 + q * q * s.rfluc * s.rfluc);
    } else {
        double pq, fq;
        if constexpr (KIND == K_FCC) {
            if (rv) {
                pq = rv->pq;
                fq = rv->fq;
            } else if constexpr (PREC == CSX_PREC_MIXED) {
                pq = pq_sphere_fast(d, q);
                fq = fq_sphere_fast(d, q);
            } else {
                pq = pq_sphere(d, q);
                fq = fq_sphere(d, q);
            }
        } else if constexpr (KIND == K_ISO_LAT) {
            // calc_GENERIC, isotropic cylinders (Source/sc_gpu_generic.h:64-74: formpq and formfq
            // of the cylinder at |q|) and homogeneous cubes (:295-302: formpq, and fq = pq)
            if (rv) {
                pq = rv->pq;
                fq = rv->fq;
            } else if (lattice_cube(s.combo_key)) {
                pq = pq_cube(d, q);
                fq = pq;
            } else {
                pq = pq_cyl_long_iso(d, q) * pq_cyl_transverse(d, q);
                fq = fq_cyl_long_iso(d, q) * fq_cyl_radial(d, q);
            }
        } else {
            // the transverse P(q) and the radial factor of <F>^2 depend on |q| only: from the
            // radial table when there is one (the same functions of the same |q|)
            const double pt = rv ? rv->pq : pq_cyl_transverse(d, q);
            const double fr = rv ? rv->fq : fq_cyl_radial(d, q);
            double pl, fl;
            switch (s.orcase) {
            case 0:
                cyl_long_both(d, q, sqrt(s.cosphi * qx * qx + s.sinphi * qy * qy + 0.000000001),
                              qx * s.cosphi * s.sintheta, qy * s.sinphi * s.sintheta, pl, fl);
                break;
            case 1: cyl_long_both(d, q, fabs(qx), qx, 0, pl, fl); break;
            case 2: cyl_long_both(d, q, fabs(qy), 0, qy, pl, fl); break;
            default: cyl_long_both(d, q, q, qx, qy, pl, fl); break;
            }
            pq = pl * pt;
            fq = fl * fr;
        }
        double width = 1;
        if (s.dsch) width = 4.0 / s.domainsize;
        if (s.para) width = (4.0 / s.domainsize) + sqd(s.reldis) * s.dist * sqd(q);
        width *= s.width_zuf;
        double inten, radint;
        const double dwe = rv ? rv->dwe : exp(-s.dwfactor * q * q);
        const double fqpq = fq / pq;
        if constexpr (PREC == CSX_PREC_REF64) {
            lattice_ref64(d, qx, qy, qz, q, width, inten, radint);
        } else {
            bool need_exact = false;
            double dropped = 0.0;
            Window W;
            constexpr bool kPrecise = precise_kind(KIND);
            if constexpr (MODE == kFix) {
                inten = cp->inten;
                dropped = cp->dropped;
                // one queue slot holds the cold pixels of one hot wave: the same window
                W = Window{uniform_i(cp->b0), uniform_i(cp->b1), uniform_i(cp->band0), uniform_i(cp->band1),
                           (cp->flags & 1) != 0};
                need_exact = (cp->flags & 2) != 0;
                if (W.near && !(cp->flags & 4)) {
                    PixelConsts L;
                    PixelConsts64 L64;
                    if constexpr (kPrecise) pixel_dirs64(s, P3, q, L64);
                    else pixel_dirs(s, P3, q, L);
                    if (s.centerbeam) near_one_add<false, kPrecise>(d, W, L, L64, P3, q, width, inten);
                    else near_one_add<true, kPrecise>(d, W, L, L64, P3, q, width, inten);
                    if (stats) atomicAdd(&stats[2], 1ull);
                }
            } else {
                inten = s.centerbeam
                            ? lattice_mixed<false, kPrecise, MODE>(d, P3, q, width, need_exact, dropped, W, stats)
                            : lattice_mixed<true, kPrecise, MODE>(d, P3, q, width, need_exact, dropped, W, stats);
            }
            radint = (s.iso != 0) ? iso_mixed(d, q) : 0.0;
            // Conditioning guard.  szq = 1 + B - (fq/pq)*dwe with B = (fq/pq)*(2*pi*I/V)*dwe.
            // The peak sum I carries up to kRelErrFP32 (FP32 loop) or kRelErrPrecise (precise
            // loop) relative error plus this lane's dropped mass; szq amplifies an error of I
            // by |B|/|szq|.  A pixel where szq could move by
            // more than 3e-5 relative (and 2^-60 absolute, below the rounding of szq's own
            // FP64 terms) takes the same window in FP64, whose only error is the dropped mass;
            // if that is still too coarse, the reference's full FP64 sum, in its order.
            const double bs = fqpq * d.two_pi_over_v * dwe;
            const double inv_hpw = s.para ? 1.0 / (kPiD * width / 2.0) : d.inv_hpw;
            auto loose = [&](double I, double rel) {
                const double szq_est = 1 + fqpq * (I * d.two_pi_over_v - 1) * dwe;
                const double dsz = fabs(bs) * (rel * fabs(I) + dropped * inv_hpw);
                return !(dsz <= fmax(3.0e-5 * fabs(szq_est), 0x1p-60));
            };
            // placeholder lanes of the fix pass take no FP64 path
            const bool live = MODE != kFix || !(cp->flags & 4);
            const bool want2 = live && !need_exact && loose(inten, kPrecise ? kRelErrPrecise : kRelErrFP32);
            if constexpr (MODE == kHot) {
                *cold = W.near || need_exact || want2;
                *cp = ColdPix{P3.ihex, P3.i, inten, dropped, W.b0, W.b1, W.band0, W.band1,
                              (W.near ? 1 : 0) | (need_exact ? 2 : 0), 0};
            } else {
            if (__ballot(want2)) {
                if (want2) need_q(s, P3);
                coop_window_sum64<kPrecise>(d, W, PixArgs{P3.qx, P3.qy, P3.qz, q, width}, want2, inten);
                if (stats && want2) atomicAdd(&stats[1], 1ull);
            }
            const bool want3 = need_exact || (want2 && loose(inten, 1.0e-13));
            if (__ballot(want3)) {
                if (want3) need_q(s, P3);
                coop_full_sum64(d, PixArgs{P3.qx, P3.qy, P3.qz, q, width}, want3, inten, radint);
                if (stats && want3) atomicAdd(&stats[4], 1ull);
            }
            }
        }
        const double szq = (1 + fqpq * (2 * kPiD * inten / s.sphno_cubevol - 1) * dwe);
        const double pqiso = (KIND == K_FCC) ? pq : 1.0;
        // iso == 0 and ifluc == 0 make their terms exactly +0 (szqiso is finite): skipped.
        double isot = 0.0, fluct = 0.0;
        if (s.iso != 0) {
            const double szqiso = (1 + (1) * (8 * kPiD * kPiD * kPiD * radint / (4 * kPiD * s.sphno_cubevol) - 1) *
                                           exp(-s.ucb * q * q));
            isot = s.iso * szqiso * pqiso;
        }
        if (s.ifluc != 0) fluct = s.ifluc / (1 + q * q * s.rfluc * s.rfluc);
        return s.base + s.i0 * (szq * pq + isot) + fluct;
    }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Image: tile (blockIdx.x, blockIdx.y) of the grid box; store in IDX order.
// Pixel (px, py) of a 16x16 tile handled by lane threadIdx.x: each wave covers an 8x8
// quadrant, the most compact 64-pixel footprint (narrowest |q| and angle ranges per wave).
__device__ __forceinline__ void tile_pixel(int& px, int& py) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    px = (w & 1) * 8 + (l & 7);
    py = (w >> 1) * 8 + (l >> 3);
}

// Kinds whose form factors depend on |q| only.
// (BCT cylinders: the transverse P(q) and the radial factor of <F>^2.)
constexpr bool radial_kind(int kind) {
    return kind == K_SPHERE || kind == K_FCC || kind == K_CUBE || kind == K_CYL_ISO || kind == K_BCT ||
           kind == K_ISO_LAT;
}

// Index of (a, b), 0 <= a <= b, in the triangle table of radial values.
__device__ __forceinline__ int tri_index(int a, int b) { return b * (b + 1) / 2 + a; }

// Radial pre-pass (midpoint q grid): in that grid |q| of pixel (ihex, i) depends only on
// {|i|, |ihex|} and is bitwise the same for all eight sign/swap images, so |q|, P(q), <F>^2
// and the Debye-Waller factor are evaluated once per unordered pair and shared.  Entry
// (a, b) uses pixel (ihex=b, i=a).  A block takes 128 entries with two halves of 128 threads
// (two waves each, so no wave mixes them): the first evaluates q, P(q) and dwe, the second
// <F>^2 -- independent FP64 chains, twice the waves in flight.  The halves meet in LDS and
// the block stores whole 32-byte entries, 16 bytes per thread.
template <int KIND, int PREC>
__device__ __forceinline__ void radial_block(const DevSet& d, int m, Radial* __restrict__ tab, int blk) {
    __shared__ double stage[128][4];
    const int k = threadIdx.x & 127;
    const bool second = threadIdx.x >= 128;
    const int n = blk * 128 + k;
    const int total = (m + 1) * (m + 2) / 2;
    if (n < total) {
        int b = (int)((sqrt(8.0 * n + 1.0) - 1.0) * 0.5);
        while (b * (b + 1) / 2 > n) --b;
        while ((b + 1) * (b + 2) / 2 <= n) ++b;
        const int a = n - b * (b + 1) / 2;
        double qx, qy, qz;
        pixel_q(d.s, b, a, qx, qy, qz);
        const double q = sqrt(qx * qx + qy * qy + qz * qz) + 0.000000001;
        if (!second) {
            double pq;
            if constexpr (KIND == K_CUBE) pq = pq_cube(d, q);
            else if constexpr (KIND == K_CYL_ISO) pq = pq_cyl_long_iso(d, q) * pq_cyl_transverse(d, q);
            else if (KIND == K_ISO_LAT && lattice_cube(d.s.combo_key)) pq = pq_cube(d, q);
            else if constexpr (KIND == K_ISO_LAT) pq = pq_cyl_long_iso(d, q) * pq_cyl_transverse(d, q);
            else if constexpr (KIND == K_BCT) pq = pq_cyl_transverse(d, q);
            else if constexpr (KIND == K_SPHERE || PREC == CSX_PREC_REF64) pq = pq_sphere(d, q);
            else pq = pq_sphere_fast(d, q);
            stage[k][0] = q;
            stage[k][1] = pq;
            stage[k][3] = exp(-d.s.dwfactor * q * q);
        } else {
            double fq = 0.0;
            if constexpr (KIND == K_FCC) {
                if constexpr (PREC == CSX_PREC_REF64) fq = fq_sphere(d, q);
                else fq = fq_sphere_fast(d, q);
            } else if constexpr (KIND == K_BCT) {
                fq = fq_cyl_radial(d, q);
            } else if constexpr (KIND == K_ISO_LAT) {
                fq = lattice_cube(d.s.combo_key) ? pq_cube(d, q) : fq_cyl_long_iso(d, q) * fq_cyl_radial(d, q);
            }
            stage[k][2] = fq;
        }
    }
    __syncthreads();
    const int e = threadIdx.x >> 1, h = threadIdx.x & 1;
    const int ne = blk * 128 + e;
    if (ne < total)
        reinterpret_cast<double2*>(tab + ne)[h] = make_double2(stage[e][2 * h], stage[e][2 * h + 1]);
}
template <int KIND, int PREC>
__global__ void __launch_bounds__(256) k_radial(const DevSet d, int m, Radial* __restrict__ tab) {
    radial_block<KIND, PREC>(d, m, tab, blockIdx.x);
}

// Image: tile (blockIdx.x, blockIdx.y) of the grid box; store in IDX order.  rtab: radial
// table of k_radial (midpoint grid, radial kinds) or null.
template <int KIND, int PREC, int MODE = kFull>
__device__ __forceinline__ void image_tile(const DevSet& d, const csx_grid& g, double* __restrict__ out,
                                           unsigned long long* __restrict__ stats, const Radial* __restrict__ rtab,
                                           ColdQueue cq = ColdQueue{}) {
    // Every lane computes (the mixed path needs all 64 lanes of a wave); lanes past the grid
    // edge evaluate the last valid pixel, and only in-grid pixels are stored.
    int px, py;
    tile_pixel(px, py);
    const int ihex0 = g.xmin + blockIdx.x * kTile + px;
    const int i0 = g.ymin + blockIdx.y * kTile + py;
    const int ihex = min(ihex0, g.xmax - 1), i = min(i0, g.ymax - 1);
    PixQ P3{ihex, i, 0.0, 0.0, 0.0, false};
    if constexpr (KIND == K_CYL_GAUSS_Q) {
        for (int k = threadIdx.x; k <= kCqTab; k += kBlock) g_cq_tab[k] = d.cq_tab[k];
        if (threadIdx.x < kCqDelta) g_cq_dl[threadIdx.x] = d.cq_delta[threadIdx.x];
        if (threadIdx.x < kCqPsi / 2) g_cq_tt[threadIdx.x] = d.cq_t[threadIdx.x];
        __syncthreads();
    }
    const Radial* rv = nullptr;
    if (rtab) {
        const int a = abs(i), b = abs(ihex);
        rv = rtab + (a <= b ? tri_index(a, b) : tri_index(b, a));
    }
    bool cold = false;
    ColdPix cpx;
    const double v = pixel_value<KIND, PREC, MODE>(d, P3, stats, rv, &cold, &cpx);
    if constexpr (MODE == kHot) {
        // queue the in-grid cold pixels (lanes past the edge repeat an in-grid pixel)
        const bool mine = cold && ihex0 < g.xmax && i0 < g.ymax;
        const unsigned long long m = __ballot(mine);
        if (m) {
            const int lane = threadIdx.x & 63;
            int slot = 0;
            if (lane == 0) {
                slot = atomicAdd(cq.count, 1);
                cq.mask[slot] = m;
            }
            slot = __shfl(slot, 0, 64);
            if (mine) cq.ent[(size_t)slot * 64 + lane] = cpx;
        }
    }
    // Stage the tile in LDS and store it row by row: each wave then writes four whole rows of
    // 16 doubles (full 128-byte lines) instead of half lines from its 8x8 quadrant.
    __shared__ double tile[kTile][kTile + 1];
    tile[py][px] = v;
    __syncthreads();
    const int r = threadIdx.x / kTile, cx = threadIdx.x % kTile;
    const int oh = g.xmin + blockIdx.x * kTile + cx, oi = g.ymin + blockIdx.y * kTile + r;
    if (oh < g.xmax && oi < g.ymax) {
        const size_t W = (size_t)(g.xmax - g.xmin);
        out[(size_t)(oh - g.xmin) + W * (size_t)(oi - g.ymin)] = tile[r][cx];
    }
}

template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock) k_image(const DevSet d, const csx_grid g, double* __restrict__ out,
                                                  unsigned long long* __restrict__ stats,
                                                  const Radial* __restrict__ rtab) {
    image_tile<KIND, PREC>(d, g, out, stats, rtab);
}

// Lattice kinds in one pass (sets with unpaired reflections, where every pixel takes an FP64
// path): the rarely taken wave-cooperative FP64 re-sums would otherwise set the register
// budget of the whole kernel (about 140 VGPRs, 3 waves per SIMD); capped at 128 (4 waves per
// SIMD), only those cold paths spill.
template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_image_lat(const DevSet d, const csx_grid g, double* __restrict__ out, unsigned long long* __restrict__ stats,
            const Radial* __restrict__ rtab) {
    image_tile<KIND, PREC>(d, g, out, stats, rtab);
}

// Lattice kinds, two passes.  The hot pass carries no FP64 re-sum code (no spills, more waves
// per SIMD) and queues the pixels that need one; the fix pass finishes them one lane each,
// with the same arithmetic as the one-pass kernel, so every pixel is bitwise the same.
template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock)
k_image_lat_hot(const DevSet d, const csx_grid g, double* __restrict__ out, unsigned long long* __restrict__ stats,
                const Radial* __restrict__ rtab, ColdQueue cq) {
    image_tile<KIND, PREC, kHot>(d, g, out, stats, rtab, cq);
}

template <int KIND, int PREC>
__global__ void __launch_bounds__(64)
k_image_lat_fix(const DevSet d, const csx_grid g, double* __restrict__ out, unsigned long long* __restrict__ stats,
                const Radial* __restrict__ rtab, ColdQueue cq) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *cq.next_count = 0;
    const int n = *cq.count;
    const size_t W = (size_t)(g.xmax - g.xmin);
    // one wave per slot (the FP64 re-sums are wave-cooperative); lanes without a cold pixel
    // repeat the slot's first one as placeholders
    for (int slot = blockIdx.x; slot < n; slot += gridDim.x) {
        const unsigned long long m = cq.mask[slot];
        const int lane = (int)threadIdx.x;
        const bool have = (m >> lane) & 1ull;
        ColdPix cp = cq.ent[(size_t)slot * 64 + (have ? lane : __builtin_ctzll(m))];
        if (!have) cp.flags = 4;
        PixQ P3{cp.ihex, cp.i, 0.0, 0.0, 0.0, false};
        const Radial* rv = nullptr;
        if (rtab) {
            const int a = abs(cp.i), b = abs(cp.ihex);
            rv = rtab + (a <= b ? tri_index(a, b) : tri_index(b, a));
        }
        const double v = pixel_value<KIND, PREC, kFix>(d, P3, stats, rv, nullptr, &cp);
        if (have) out[(size_t)(cp.ihex - g.xmin) + W * (size_t)(cp.i - g.ymin)] = v;
    }
}

__global__ void __launch_bounds__(kBlock) k_qgrid(const csx_scalars s, const csx_grid g, double* __restrict__ out) {
    const int ihex = g.xmin + blockIdx.x * kTile + (threadIdx.x & (kTile - 1));
    const int i = g.ymin + blockIdx.y * kTile + (threadIdx.x / kTile);
    if (ihex >= g.xmax || i >= g.ymax) return;
    double qx, qy, qz;
    pixel_q(s, ihex, i, qx, qy, qz);
    const size_t k = (size_t)(ihex - g.xmin) + (size_t)(g.xmax - g.xmin) * (size_t)(i - g.ymin);
    out[3 * k] = qx;
    out[3 * k + 1] = qy;
    out[3 * k + 2] = qz;
}

// checkFitRanges (Source/sc_calc_generic_gpu.cu:293-347): true if pixel (x,y) of the fit
// array is excluded.
__device__ __forceinline__ bool fit_masked(const csx_scalars& s, const csx_mask& m, const double* data, int w,
                                           int x, int y) {
    for (int r = 0; r < 4; ++r) {
        const int* R = m.rects + 4 * r;
        if (R[0] < 0) continue;
        if (x < R[0] || R[2] < x || y < R[1] || R[3] < y) continue;
        return true;
    }
    if (m.border > 0 || m.bstop > 0) {
        if (x < s.xmin + m.border || x >= s.xmax - m.border || y < s.ymin + m.border || y >= s.ymax - m.border)
            return true;
        if (x >= s.beamx - m.bstop && x < s.beamx + m.bstop && y >= s.beamy - m.bstop && y < s.beamy + m.bstop)
            return true;
    } else if (m.bstop == -1) {
        if (data[(size_t)x + (size_t)w * y] <= data[0]) return true;
    }
    return false;
}

// Device preparation of a batched set's lattice tables (k_prep_sets): inputs written at
// staging, outputs written by the kernel before the set is evaluated.
struct PrepIn {
    double R[9], RT[9];       // rotated reciprocal bases (rows a*, b*, c*) and twinned ones
    const int32_t* l2;        // latpar2 of the reflection list
    const int2* pairs_t;      // Friedel pairs (ia, ib) of the list, in list order
    const int2* iso_t;        // isotropic entries (latpar index, mhkl*fhkl)
    float* l3;                // outputs: latpar3, pair records (one of pr, pr64), bin table,
    PairRec* pr;              // isotropic list
    PairRec64* pr64;
    int32_t* lut;
    float2* iso;
    int32_t on, n2, np, niso;
    int32_t flags;            // bit 0: a pair's members differ; bits 8+: ambiguous roundings
    int32_t pad;
};

struct FitSet {
    DevSet d;
    csx_mask m;
    const double* data;
    int32_t w, h;
    int32_t tiles_x, first_tile;  // tile range of this set in the batch
    const Radial* rtab;           // radial table of the set (k_batch_radial), or null
    int32_t rad_m, rad_pad;       // its largest |index|
    PrepIn pin;
};

struct TileSum {
    double sum;
    int32_t cnt, nan;
};


// Pixel value of a set of a batch.  The kinds without a lattice and the isotropic particles on
// lattices read the set's constants from a private copy: read through the reference into the
// sets array, the isotropic cylinder's P(q) came out wrong above its series radius (with and
// without a lattice).  The sphere and z-oriented cylinder lattice kinds read them in place.
template <int KIND, int PREC>
__device__ __forceinline__ double set_pixel(const FitSet& f, PixQ& P3, const Radial* rv) {
    if constexpr (KIND == K_FCC || KIND == K_BCT) {
        return pixel_value<KIND, PREC>(f.d, P3, nullptr, rv);
    } else {
        const DevSet d = f.d;
        return pixel_value<KIND, PREC>(d, P3, nullptr, rv);
    }
}

// One fit tile: per-pixel FQSVERGL terms, NaN counted, block-reduced in FP64.
template <int KIND, int PREC>
__device__ __forceinline__ void fit_tile(const FitSet& f, int tile, double* __restrict__ fqs_out, TileSum* __restrict__ out) {
    __shared__ double s_sum[kBlock / 64];
    __shared__ int s_cnt[kBlock / 64], s_nan[kBlock / 64];
    const int tx = tile % f.tiles_x, ty = tile / f.tiles_x;
    int px, py;
    tile_pixel(px, py);
    const int x = tx * kTile + px;
    const int y = ty * kTile + py;
    double term = 0.0;
    int nz = 0, isnan_ = 0;
    // All lanes evaluate a pixel (wave-wide |q| ranges need all 64 lanes); lanes outside the
    // fit array or masked discard the value.
    const bool inside = x < f.w && y < f.h;
    const int xc = min(x, f.w - 1), yc = min(y, f.h - 1);
    const csx_scalars& s = f.d.s;
    const bool masked = !inside || fit_masked(s, f.m, f.data, f.w, xc, yc);
    const int ihex = xc + s.iimin, i = yc + s.iimin;
    PixQ P3{ihex, i, 0.0, 0.0, 0.0, false};
    const Radial* rv = nullptr;
    if (f.rtab) {
        const int a = abs(i), b = abs(ihex);
        rv = f.rtab + (a <= b ? tri_index(a, b) : tri_index(b, a));
    }
    const double v = set_pixel<KIND, PREC>(f, P3, rv);
    if (inside) {
        const size_t idx = (size_t)(ihex - s.xmin) + (size_t)(s.xmax - s.xmin) * (size_t)(i - s.ymin);
        if (!masked) {
            const double dv = f.data[idx];
            if (v > 0 && dv > 0) {
                const double a = log(v) / log(10.) - log(dv) / log(10.);
                term = a * a;
            }
            if (fqs_out) fqs_out[idx] = term;
        } else if (fqs_out) {
            fqs_out[(size_t)x + (size_t)f.w * y] = 0.0;
        }
        if (isnan(term)) { isnan_ = 1; term = 0.0; }
        else if (term != 0) nz = 1;
    }
    const double ws = wave_sum(term);
    const int wc = wave_sum_i(nz), wn = wave_sum_i(isnan_);
    const int wid = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) { s_sum[wid] = ws; s_cnt[wid] = wc; s_nan[wid] = wn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        int c = 0, n = 0;
        for (int k = 0; k < kBlock / 64; ++k) { t += s_sum[k]; c += s_cnt[k]; n += s_nan[k]; }
        out->sum = t; out->cnt = c; out->nan = n;
    }
    __syncthreads();
}

template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock) k_fit_one(const FitSet f, double* __restrict__ fqs_out, TileSum* __restrict__ out) {
    fit_tile<KIND, PREC>(f, blockIdx.x, fqs_out, out + blockIdx.x);
}

// Batched objective: one launch over the flat tile list of all sets, one block per tile.
// (A grid-stride loop here makes the compiler hoist the set's fields out of the loop into
// vector registers: about 240 VGPRs and 2 waves per SIMD instead of 128 and 4.)
template <int KIND, int PREC>
__device__ __forceinline__ void fit_batch_tiles(const FitSet* __restrict__ sets, const int2* __restrict__ tiles,
                                                int ntiles, TileSum* __restrict__ out) {
    const int t = blockIdx.x;
    if (t < ntiles) {
        const int2 st = tiles[t];  // (set, tile within set)
        fit_tile<KIND, PREC>(sets[st.x], st.y, nullptr, out + t);
    }
}
template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock) k_fit_batch(const FitSet* __restrict__ sets, const int2* __restrict__ tiles,
                                                      int ntiles, TileSum* __restrict__ out) {
    fit_batch_tiles<KIND, PREC>(sets, tiles, ntiles, out);
}
// Lattice kinds, register budget capped as for k_image_lat.
template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_fit_batch_lat(const FitSet* __restrict__ sets, const int2* __restrict__ tiles, int ntiles, TileSum* __restrict__ out) {
    fit_batch_tiles<KIND, PREC>(sets, tiles, ntiles, out);
}

// Variants of the batched objective for the lattice kinds in mixed precision: the same code
// with different register budgets, so the sums do not depend on the variant, only the time
// does; a tuning table picks one per combination and detector size.  Variant 0 is
// k_fit_batch_lat (budget cut for 4 waves per SIMD: cold paths spill), 1 is uncapped (no
// spills, 3 waves), 2 and 3 cut for 5 and 6 waves (more spills, more waves).  Looping over
// several tiles per workgroup was measured too: the compiler keeps a set's fields live across
// the loop (spills of ~400 bytes per lane, or 240 VGPRs), so every variant takes one tile.
constexpr int kTuneVariants = 4;
constexpr int kTuneSizes = 4;  // mean tiles per set: <= 64 (128^2), <= 256, <= 576, more
template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock)
k_fit_batch_free(const FitSet* __restrict__ sets, const int2* __restrict__ tiles, int ntiles, TileSum* __restrict__ out) {
    fit_batch_tiles<KIND, PREC>(sets, tiles, ntiles, out);
}
template <int KIND, int PREC, int WPE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE)))
k_fit_batch_cap(const FitSet* __restrict__ sets, const int2* __restrict__ tiles, int ntiles, TileSum* __restrict__ out) {
    fit_batch_tiles<KIND, PREC>(sets, tiles, ntiles, out);
}

// Batched images: one launch over the flat tile list of many parameter sets, each set's
// image (its whole grid, IDX order) written to the buffer staged as its data pointer.  The
// tile goes through LDS so that every wave stores whole 128-byte rows.
template <int KIND, int PREC>
__device__ __forceinline__ void image_batch_tiles(const FitSet* __restrict__ sets, const int2* __restrict__ tiles,
                                                  int ntiles) {
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int2 st = tiles[t];
    const FitSet& f = sets[st.x];
    const int tx = st.y % f.tiles_x, ty = st.y / f.tiles_x;
    int px, py;
    tile_pixel(px, py);
    const csx_scalars& s = f.d.s;
    const int x = min(tx * kTile + px, f.w - 1), y = min(ty * kTile + py, f.h - 1);
    PixQ P3{s.xmin + x, s.ymin + y, 0.0, 0.0, 0.0, false};
    const Radial* rv = nullptr;
    if (f.rtab) {
        const int a = abs(P3.i), b = abs(P3.ihex);
        rv = f.rtab + (a <= b ? tri_index(a, b) : tri_index(b, a));
    }
    const double v = set_pixel<KIND, PREC>(f, P3, rv);
    __shared__ double tile[kTile][kTile + 1];
    tile[py][px] = v;
    __syncthreads();
    const int r = threadIdx.x / kTile, cx = threadIdx.x % kTile;
    const int ox = tx * kTile + cx, oy = ty * kTile + r;
    if (ox < f.w && oy < f.h) const_cast<double*>(f.data)[(size_t)ox + (size_t)f.w * oy] = tile[r][cx];
}
template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock) k_image_batch(const FitSet* __restrict__ sets, const int2* __restrict__ tiles,
                                                        int ntiles) {
    image_batch_tiles<KIND, PREC>(sets, tiles, ntiles);
}
template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_image_batch_lat(const FitSet* __restrict__ sets, const int2* __restrict__ tiles, int ntiles) {
    image_batch_tiles<KIND, PREC>(sets, tiles, ntiles);
}

// Sums of a set's tile partials: one workgroup per set.  Lane t adds tiles t, t + kBlock, ...
// in order, then the lanes' partials meet in a fixed tree in LDS.  The order depends only on
// the tile count, and every path that sums tiles (csx_fit_fqs, csx_fit_fqs_noalloc, the
// batch) uses this one, so their sums agree bit for bit.
struct SetSum {
    double sum;
    long long cnt, nan;
    long long flags;  // batch: the set's device-preparation flags (PrepIn::flags)
};
__device__ __forceinline__ void block_sum_tiles(const TileSum* __restrict__ part, int n, SetSum* __restrict__ out) {
    __shared__ double s_s[kBlock];
    __shared__ long long s_c[kBlock], s_n[kBlock];
    const int t = threadIdx.x;
    double v = 0.0;
    long long c = 0, nn = 0;
    for (int k = t; k < n; k += kBlock) {
        v += part[k].sum;
        c += part[k].cnt;
        nn += part[k].nan;
    }
    s_s[t] = v;
    s_c[t] = c;
    s_n[t] = nn;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (t < h) {
            s_s[t] += s_s[t + h];
            s_c[t] += s_c[t + h];
            s_n[t] += s_n[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        out->sum = s_s[0];
        out->cnt = s_c[0];
        out->nan = s_n[0];
    }
}

// Batch: set blockIdx.x, tiles [first_tile, first_tile + ntiles_of).
__global__ void __launch_bounds__(kBlock) k_reduce_sets(const FitSet* __restrict__ sets, int nsets,
                                                        const TileSum* __restrict__ part,
                                                        const int* __restrict__ ntiles_of, SetSum* __restrict__ out) {
    const int s = blockIdx.x;
    if (s >= nsets) return;
    block_sum_tiles(part + sets[s].first_tile, ntiles_of[s], out + s);
    if (threadIdx.x == 0) out[s].flags = sets[s].pin.on ? sets[s].pin.flags : 0;
}

// Radial tables of a batch's sets (the radial kinds in the midpoint grid): block b fills 128
// entries of the set s with rblock0[s] <= b < rblock0[s + 1], as k_radial does for one render.
template <int KIND, int PREC>
__global__ void __launch_bounds__(256) k_batch_radial(const FitSet* __restrict__ sets, int nsets,
                                                      const int* __restrict__ rblock0) {
    const int b = blockIdx.x;
    if (b >= rblock0[nsets]) return;
    int lo = 0, hi = nsets;  // the last set whose first block is <= b
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (rblock0[mid] <= b) lo = mid;
        else hi = mid;
    }
    const FitSet& f = sets[lo];
    // the set's constants as a private copy: read through the reference into the sets array,
    // the isotropic cylinders' entries came out wrong (zero P(q) above the series radius)
    const DevSet d = f.d;
    radial_block<KIND, PREC>(d, f.rad_m, const_cast<Radial*>(f.rtab), b - rblock0[lo]);
}

// ---------------------------------------------------------------------------
// Device preparation of batched lattice sets: what the host builds per parameter set from the
// lattice constants (the latpar3 rows, Source/sc_calc_generic_cpu.cpp:838-954, and the
// mixed path's Friedel-pair records, |u_z| segments and |q| bin table), built by one
// workgroup per set from the set's rotated reciprocal bases and its lattice's structural
// lists.  Every value is the host's, bit for bit: the same arithmetic (csx_rows.h), the pairs
// in the host's stable |q| order, the weights summed in the host's order; the peak norm's
// exponential is correctly rounded (csx_ddexp.h) and each entry where the C library's could
// round the other way is checked and counted.
constexpr int kPrepBlock = 256;
constexpr int kMaxPairs = CSX_LATLEN / 2;
constexpr int kLutBinsDev = 256;  // bins of the |q| table (kLutBins on the host side)

// One latpar3 row: the 17 floats into T (if not null), and the values the pairs use.
struct RowVals {
    float qx, qy, qz, qh, t10, t11;
    int amb;
};
__device__ inline void prep_row(const FitSet& f, int ii, float* T, RowVals& v) {
    const csx_scalars& s = f.d.s;
    const int32_t* L = f.pin.l2 + (size_t)ii * 6;
    rows::Geometry g;
    rows::geometry(f.pin.R, f.pin.RT, L[1], L[2], L[3], g);
    const int shp = (int)(s.combo_key >> 16 & 0xff);
    v.amb = 0;
    v.t11 = 0.0f;
    if (shp == 2 || shp == 7) {
        const double a = rows::x2phi(g.qh, s.phiwidth);
        double ex = 0.0;
        if (rows::gauss_norm3_uses_exp(a)) {
            bool near_mid;
            double alt;
            ex = ddexp::exp_rn(rows::gauss_norm3_exp_arg(a), near_mid, alt);
            if (near_mid && (float)rows::gauss_norm3_from(a, alt) != (float)rows::gauss_norm3_from(a, ex)) v.amb = 1;
        }
        v.t11 = (float)rows::gauss_norm3_from(a, ex);
    }
    v.qx = (float)g.qx;
    v.qy = (float)g.qy;
    v.qz = (float)g.qz;
    v.qh = (float)g.qh;
    v.t10 = (float)rows::g3(g.qh);
    if (T) {
        T[0] = 0.0f;
        T[1] = (float)(2 * kPiD / s.uca);
        T[2] = v.qx; T[3] = v.qy; T[4] = v.qz; T[5] = v.qh;
        T[6] = (float)(g.qh / (2.0 * kPiD));
        T[7] = (float)g.tx; T[8] = (float)g.ty; T[9] = (float)g.tz;
        T[10] = v.t10;
        T[11] = v.t11;
        T[12] = 0.0f;
        T[13] = (float)g.qxy;
        T[14] = 0.0f; T[15] = 0.0f; T[16] = 0.0f;
    }
}

__global__ void __launch_bounds__(kPrepBlock) k_prep_sets(FitSet* __restrict__ sets, int nsets,
                                                          const int* __restrict__ ntiles_of, int2* __restrict__ tiles) {
    if ((int)blockIdx.x >= nsets) return;
    FitSet& f = sets[blockIdx.x];
    // the set's entries of the batch's flat tile list: (set, tile within the set)
    for (int k = threadIdx.x; k < ntiles_of[blockIdx.x]; k += kPrepBlock)
        tiles[f.first_tile + k] = make_int2((int)blockIdx.x, k);
    if (!f.pin.on) return;
    __shared__ float s_qh[kMaxPairs], s_t10[kMaxPairs], s_t11[kMaxPairs];
    __shared__ double s_A[kMaxPairs];
    __shared__ int s_order[kMaxPairs];
    __shared__ unsigned char s_seg[kMaxPairs];  // band | 0x10 if |u_z| <= kPlaneUz
    __shared__ int s_cnt[kSegments][kPrepBlock];
    __shared__ int s_hist[kSegments][kLutBinsDev + 1];
    __shared__ int s_total[kSegments], s_start[kSegments + 1];
    __shared__ int s_flags, s_amb;
    __shared__ float s_qtop, s_qbot;
    const int t = threadIdx.x;
    const int n2 = f.pin.n2, np = f.pin.np;
    if (t == 0) {
        s_flags = 0;
        s_amb = 0;
    }
    for (int i = t; i < kSegments * (kLutBinsDev + 1); i += kPrepBlock) (&s_hist[0][0])[i] = 0;
    __syncthreads();
    // latpar3 rows (row 0 zero)
    for (int i = t; i < 17; i += kPrepBlock) f.pin.l3[i] = 0.0f;
    int amb = 0;
    for (int ii = 1 + t; ii <= n2; ii += kPrepBlock) {
        RowVals v;
        prep_row(f, ii, f.pin.l3 + (size_t)ii * 17, v);
        amb += v.amb;
    }
    // pairs in list order: |q_hkl|, weight, segment; the partner must be the exact mirror
    for (int k = t; k < np; k += kPrepBlock) {
        const int2 ab = f.pin.pairs_t[k];
        RowVals a, b;
        prep_row(f, ab.x, nullptr, a);
        prep_row(f, ab.y, nullptr, b);
        const int32_t* La = f.pin.l2 + (size_t)ab.x * 6;
        const int32_t* Lb = f.pin.l2 + (size_t)ab.y * 6;
        const bool same = b.qh == a.qh && b.t10 == a.t10 && b.t11 == a.t11 && La[4] * La[5] == Lb[4] * Lb[5] &&
                          b.qx == -a.qx && b.qy == -a.qy && b.qz == -a.qz && a.qh > 0.0f;
        if (!same) atomicOr(&s_flags, 1);
        const double A = (double)(La[4] * La[5]) * (double)a.t10 / (double)a.t11;
        const double ux = (double)a.qx, uy = (double)a.qy, uz = (double)a.qz;
        const double un = sqrt(ux * ux + uy * uy + uz * uz);
        const int band = min(kBands - 1, (int)(fabs(uz / un) * kBands));
        const bool plane = fabsf((float)(uz / un)) <= kPlaneUz;
        s_qh[k] = a.qh;
        s_t10[k] = a.t10;
        s_t11[k] = a.t11;
        s_A[k] = A;
        s_seg[k] = (unsigned char)(band | (plane ? 0x10 : 0));
    }
    if (amb) atomicAdd(&s_amb, amb);
    __syncthreads();
    // the weights' bound in list order, the |q| range
    if (t == 0) {
        double asum = 0.0;
        bool negative = false;
        float qtop = 0.0f, qbot = np > 0 ? s_qh[0] : 0.0f;
        for (int k = 0; k < np; ++k) {
            asum += 2.0 * fabs(s_A[k]);
            negative = negative || s_A[k] < 0;
            qtop = fmaxf(qtop, s_qh[k]);
            qbot = fminf(qbot, s_qh[k]);
        }
        f.d.asum2 = negative ? HUGE_VAL : asum;
        s_qtop = qtop;
        s_qbot = qbot;
    }
    // stable order by |q_hkl|
    for (int k = t; k < np; k += kPrepBlock) {
        const float q = s_qh[k];
        int r = 0;
        for (int j = 0; j < np; ++j) r += (s_qh[j] < q) || (s_qh[j] == q && j < k);
        s_order[r] = k;
    }
    __syncthreads();
    // segment positions: each lane takes a run of the sorted order
    const int run = (np + kPrepBlock - 1) / kPrepBlock;
    const int p0 = min(np, t * run), p1 = min(np, p0 + run);
    int cnt[kSegments];
    for (int b = 0; b < kSegments; ++b) cnt[b] = 0;
    for (int p = p0; p < p1; ++p) {
        const int sg = s_seg[s_order[p]];
        ++cnt[sg & 0xf];
        ++cnt[kBands];
        if (sg & 0x10) ++cnt[kBands + 1];
    }
    for (int b = 0; b < kSegments; ++b) s_cnt[b][t] = cnt[b];
    __syncthreads();
    if (t < kSegments) {
        int acc = 0;
        for (int j = 0; j < kPrepBlock; ++j) {
            const int c = s_cnt[t][j];
            s_cnt[t][j] = acc;
            acc += c;
        }
        s_total[t] = acc;
    }
    __syncthreads();
    if (t == 0) {
        s_start[0] = 0;
        for (int b = 0; b < kSegments; ++b) s_start[b + 1] = s_start[b] + ((s_total[b] + 3) & ~3);
    }
    __syncthreads();
    int off[kSegments];
    for (int b = 0; b < kSegments; ++b) off[b] = s_start[b] + s_cnt[b][t];
    for (int p = p0; p < p1; ++p) {
        const int k = s_order[p];
        const int2 ab = f.pin.pairs_t[k];
        const int32_t* La = f.pin.l2 + (size_t)ab.x * 6;
        rows::Geometry g;
        rows::geometry(f.pin.R, f.pin.RT, La[1], La[2], La[3], g);
        const double ux = (double)(float)g.qx, uy = (double)(float)g.qy, uz = (double)(float)g.qz;
        const double qh = (double)s_qh[k];
        const double A = s_A[k];
        const int sg = s_seg[k];
        const int dst[3] = {off[sg & 0xf]++, off[kBands]++, (sg & 0x10) ? off[kBands + 1]++ : -1};
        if (f.pin.pr64) {
            PairRec64 r;
            r.ux = ux / qh;
            r.uy = uy / qh;
            r.uz = uz / qh;
            r.A = A;
            r.qh = s_qh[k];
            r.ia = ab.x;
            r.ib = ab.y;
            r.pad = 0.0f;
            for (int j = 0; j < 3; ++j)
                if (dst[j] >= 0) f.pin.pr64[dst[j]] = r;
        } else {
            const double un = sqrt(ux * ux + uy * uy + uz * uz);
            PairRec r;
            r.ux = (float)(ux / un);
            r.uy = (float)(uy / un);
            r.uz = (float)(uz / un);
            r.qh = s_qh[k];
            r.uz2 = (float)((uz / un) * (uz / un));
            r.A = (float)A;
            r.ia = ab.x;
            r.ib = ab.y;
            for (int j = 0; j < 3; ++j)
                if (dst[j] >= 0) f.pin.pr[dst[j]] = r;
        }
    }
    // inert padding to a multiple of 4 per segment
    if (t < kSegments) {
        for (int i = s_start[t] + s_total[t]; i < s_start[t + 1]; ++i) {
            if (f.pin.pr64) f.pin.pr64[i] = PairRec64{0.0, 0.0, 0.0, 0.0, 1.0e30f, 0, 0, 0.0f};
            else f.pin.pr[i] = PairRec{0.0f, 0.0f, 0.0f, 1.0e30f, 0.0f, 0.0f, 0, 0};
        }
    }
    // bin table: lut[b][k] = first index of segment b with (double)|q| >= k dq
    const double span = (double)s_qtop * 1.001 + 1e-3;
    const float inv_dq = (float)(kLutBinsDev / span);
    const double dqb = 1.0 / (double)inv_dq;
    for (int k = t; k < np; k += kPrepBlock) {
        const double q = (double)s_qh[k];
        int kk = max(0, (int)(q / dqb) - 2);
        while (kk < kLutBinsDev && !(q < kk * dqb)) ++kk;
        if (kk < kLutBinsDev) {
            const int sg = s_seg[k];
            atomicAdd(&s_hist[sg & 0xf][kk], 1);
            atomicAdd(&s_hist[kBands][kk], 1);
            if (sg & 0x10) atomicAdd(&s_hist[kBands + 1][kk], 1);
        }
    }
    __syncthreads();
    if (t < kSegments) {
        int32_t* lb = f.pin.lut + (size_t)t * (kLutBinsDev + 1);
        int acc = 0;
        for (int k = 0; k < kLutBinsDev; ++k) {
            acc += s_hist[t][k];
            lb[k] = s_start[t] + acc;
        }
        lb[kLutBinsDev] = s_start[t] + s_total[t];
    }
    // isotropic list: (|q_hkl| of latpar row ii, mhkl*fhkl)
    for (int i = t; i < f.pin.niso; i += kPrepBlock) {
        const int2 e = f.pin.iso_t[i];
        const int32_t* L = f.pin.l2 + (size_t)e.x * 6;
        rows::Geometry g;
        rows::geometry(f.pin.R, f.pin.RT, L[1], L[2], L[3], g);
        f.pin.iso[i] = make_float2((float)g.qh, (float)e.y);
    }
    if (t == 0) {
        f.d.npairs = s_start[kSegments];
        f.d.lut_inv_dq = inv_dq;
        f.d.qh_min = s_qbot;
        f.d.qh_max = s_qtop;
        f.pin.flags = s_flags | (min(s_amb, 0x7fffff) << 8);
    }
}

// ExpandImage (Source/sc_calc_generic_cpu.cpp:1246-1269): the whole image (-z..z)^2 from the
// computed first quadrant (quad 1) or right half (quad 2), as a gather per output pixel of
// what the reference's in-place copy loops leave there.  Quadrant 1: (-x, y), (-x, -y-1) and
// (x, -y-1) take the value of (x, y).  Half: (-x, -y) takes the value of (x, y) for
// y > -z; on the column x = 0 the copy runs in place in increasing y, so there the upper half
// ends up with the mirrored lower half and the lower half keeps its own values.  Pixels the
// loops never write (the column x = -z, and the row y = -z left of the axis for the half)
// keep what the array held before; here they are 0.
__global__ void k_expand(const double* __restrict__ region, int rw, int rxmin, int rymin, int z, int quad,
                         double* __restrict__ full) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int W = 2 * z;
    if (idx >= (long long)W * W) return;
    const int x = (int)(idx % W) - z, y = (int)(idx / W) - z;
    auto at = [&](int sx, int sy) { return region[(size_t)(sx - rxmin) + (size_t)rw * (size_t)(sy - rymin)]; };
    double v = 0.0;
    if (x != -z) {
        if (quad == 1) v = at(x < 0 ? -x : x, y >= 0 ? y : -y - 1);
        else if (x > 0) v = at(x, y);
        else if (x == 0) v = y > 0 ? at(0, -y) : at(0, y);
        else if (y != -z) v = at(-x, -y);
    }
    full[idx] = v;
}

// One set's sum: tiles [0, n) (csx_fit_fqs, csx_fit_fqs_noalloc).
__global__ void __launch_bounds__(kBlock) k_sum_tiles(const TileSum* __restrict__ part, int n, SetSum* __restrict__ out) {
    block_sum_tiles(part, n, out);
}

}  // namespace csx

// Scratch of csx_fit_fqs_noalloc: the tile partials of one fit array (up to kScratchTiles
// 16x16 tiles, 2048 x 2048 pixels) and their total (a SetSum in the two slots after them),
// 256 KiB per GPU for the life of the process.  Host code never uses this name as an
// address: the device address of each GPU's copy comes from hipGetSymbolAddress
// (scratch_address below).
constexpr int kScratchTiles = 16384;
static_assert(sizeof(csx::SetSum) <= 2 * sizeof(csx::TileSum), "the total fits in two tile slots");
static __device__ __align__(256) csx::TileSum scratch[kScratchTiles + 2];

// ===========================================================================
// C ABI
// ===========================================================================
using namespace csx;

namespace {

// Host staging memory of a batch, pinned (asynchronous, full-rate H2D copies); grows by
// reallocation.  Only touched between batched runs, which end with a stream synchronize.
struct PinnedArena {
    char* p = nullptr;
    size_t cap = 0, size = 0;
    hipError_t reserve(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        const size_t nc = std::max(bytes, 2 * cap);
        char* q = nullptr;
        hipError_t e = hipHostMalloc((void**)&q, nc, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        if (size) std::memcpy(q, p, size);
        if (p) (void)hipHostFree(p);
        p = q;
        cap = nc;
        return hipSuccess;
    }
    void clear() { size = 0; }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = size = 0;
    }
};

// The graph of a batched-objective round and its nodes (rebuilt when the kernel changes).
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t h2d = nullptr, prep = nullptr, rad = nullptr, fit = nullptr, red = nullptr, d2h = nullptr;
    const void* fit_fn = nullptr;
    long long launches = 0;
};

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

// Base-2 exponent below which the mixed path drops a pair term (2^-32 ~ 2.3e-10 of the
// pair's weight); each pixel's bound of the dropped mass enters its conditioning guard.
constexpr float kDropExponent = -32.0f;
constexpr int kLutBins = kLutBinsDev;

// Host-side fast tables of one parameter set: Friedel pairs in |u_z| bands sorted by
// |q_hkl|, the bin table, unpaired reflections and the isotropic list.
struct FastTables {
    std::vector<PairRec> pairs;
    std::vector<PairRec64> pairs64;  // built when the kind sums in the precise loop
    std::vector<int32_t> lut;
    std::vector<int32_t> singles;
    std::vector<float2> iso;
    float lut_inv_dq = 0.0f;
    float qh_min = 0.0f, qh_max = 0.0f;
    double asum2 = 0.0;
};

void build_fast_tables(const csx_scalars& s, const int32_t* L1, const int32_t* L2, const float* L3, FastTables& ft) {
    ft = FastTables{};
    const bool precise = precise_kind(kind_of(s.combo_key));
    const int n2 = s.peakmax2;
    std::vector<char> used(n2 + 1, 0);
    std::vector<PairRec> found;
    std::vector<PairRec64> found64;
    std::vector<int> band_of;
    auto valid = [&](int ii) {
        const double qh = L3[ii * 17 + 5];
        return qh > 0 && L2[ii * 6 + 4] * L2[ii * 6 + 5] != 0;
    };
    double asum = 0.0;
    bool negative = false;
    // (h, k, l) -> reflections with those indices, in list order: a dense table over the
    // index box (chained through `next`), falling back to a hash map for a very large box
    int hm = 0;
    for (int ii = 1; ii <= n2; ++ii)
        for (int c = 1; c <= 3; ++c) hm = std::max(hm, std::abs(L2[ii * 6 + c]));
    const int side = 2 * hm + 1;
    const bool dense = (int64_t)side * side * side <= (1 << 22);
    auto key = [](int h, int k, int l) { return ((int64_t)(h + 1024) << 22) | ((int64_t)(k + 1024) << 11) | (l + 1024); };
    auto cell = [&](int h, int k, int l) { return ((h + hm) * side + (k + hm)) * side + (l + hm); };
    static thread_local std::vector<int> head;
    std::vector<int> next(n2 + 1, 0), tail;
    std::unordered_map<int64_t, std::vector<int>> byhkl;
    if (dense) {
        head.assign((size_t)side * side * side, 0);
        tail.assign((size_t)side * side * side, 0);
        for (int ii = 1; ii <= n2; ++ii) {
            const int c = cell(L2[ii * 6 + 1], L2[ii * 6 + 2], L2[ii * 6 + 3]);
            if (!head[c]) head[c] = ii; else next[tail[c]] = ii;
            tail[c] = ii;
        }
    } else {
        byhkl.reserve(2 * n2 + 1);
        for (int ii = 1; ii <= n2; ++ii) byhkl[key(L2[ii * 6 + 1], L2[ii * 6 + 2], L2[ii * 6 + 3])].push_back(ii);
    }
    for (int ii = 1; ii <= n2; ++ii) {
        if (used[ii] || !valid(ii)) continue;
        used[ii] = 1;
        const int h = L2[ii * 6 + 1], k = L2[ii * 6 + 2], l = L2[ii * 6 + 3];
        int jj = 0;
        auto partner = [&](int c) {
            return !used[c] && valid(c) && L3[c * 17 + 5] == L3[ii * 17 + 5] &&
                   L3[c * 17 + 10] == L3[ii * 17 + 10] && L3[c * 17 + 11] == L3[ii * 17 + 11] &&
                   L2[c * 6 + 4] * L2[c * 6 + 5] == L2[ii * 6 + 4] * L2[ii * 6 + 5] &&
                   L3[c * 17 + 2] == -L3[ii * 17 + 2] && L3[c * 17 + 3] == -L3[ii * 17 + 3] &&
                   L3[c * 17 + 4] == -L3[ii * 17 + 4];
        };
        if (dense) {
            for (int c = head[cell(-h, -k, -l)]; c && !jj; c = next[c])
                if (partner(c)) jj = c;
        } else {
            const auto it = byhkl.find(key(-h, -k, -l));
            if (it != byhkl.end())
                for (int c : it->second) {
                    if (jj) break;
                    if (partner(c)) jj = c;
                }
        }
        if (!jj) { ft.singles.push_back(ii); continue; }
        used[jj] = 1;
        const double qh = L3[ii * 17 + 5];
        const double A = (double)(L2[ii * 6 + 4] * L2[ii * 6 + 5]) * (double)L3[ii * 17 + 10] / (double)L3[ii * 17 + 11];
        const double ux = (double)L3[ii * 17 + 2], uy = (double)L3[ii * 17 + 3], uz = (double)L3[ii * 17 + 4];
        const double un = std::sqrt(ux * ux + uy * uy + uz * uz);
        PairRec64 r64;
        r64.ux = ux / qh;
        r64.uy = uy / qh;
        r64.uz = uz / qh;
        r64.A = A;
        r64.qh = L3[ii * 17 + 5];
        r64.ia = ii;
        r64.ib = jj;
        r64.pad = 0.0f;
        found64.push_back(r64);
        PairRec r;
        r.ux = (float)(ux / un);
        r.uy = (float)(uy / un);
        r.uz = (float)(uz / un);
        r.qh = L3[ii * 17 + 5];
        r.uz2 = (float)((uz / un) * (uz / un));
        r.A = (float)A;
        r.ia = ii;
        r.ib = jj;
        found.push_back(r);
        band_of.push_back(std::min(kBands - 1, (int)(std::fabs(uz / un) * kBands)));
        asum += 2.0 * std::fabs(A);
        negative = negative || A < 0;
    }
    float qtop = 0.0f, qbot = found.empty() ? 0.0f : found[0].qh;
    for (const PairRec& r : found) {
        qtop = std::max(qtop, r.qh);
        qbot = std::min(qbot, r.qh);
    }
    ft.qh_min = qbot;
    ft.qh_max = qtop;
    const double span = (double)qtop * 1.001 + 1e-3;
    ft.lut_inv_dq = (float)(kLutBins / span);
    const double dqb = 1.0 / (double)ft.lut_inv_dq;
    ft.lut.assign((size_t)kSegments * (kLutBins + 1), 0);
    const PairRec inert{0.0f, 0.0f, 0.0f, 1.0e30f, 0.0f, 0.0f, 0, 0};
    const PairRec64 inert64{0.0, 0.0, 0.0, 0.0, 1.0e30f, 0, 0, 0.0f};
    // Segments 0..kBands-1: one |u_z| band each; segment kBands: every pair; segment
    // kBands + 1: the pairs with |u_z| <= kPlaneUz.  One stable sort by |q_hkl|; each segment
    // takes its pairs in that order.
    std::vector<int> order(found.size());
    for (size_t k = 0; k < found.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return found[x].qh < found[y].qh; });
    ft.pairs.reserve(3 * found.size() + 4 * kSegments);
    if (precise) ft.pairs64.reserve(3 * found.size() + 4 * kSegments);
    std::vector<int> band;
    band.reserve(found.size());
    for (int b = 0; b < kSegments; ++b) {
        band.clear();
        for (int k : order)
            if (b == kBands || (b < kBands && band_of[k] == b) || (b > kBands && std::fabs(found[k].uz) <= kPlaneUz))
                band.push_back(k);
        const int start = (int)ft.pairs.size();  // a multiple of 4
        for (int k : band) {
            ft.pairs.push_back(found[k]);
            if (precise) ft.pairs64.push_back(found64[k]);
        }
        const int end = (int)ft.pairs.size();
        while (ft.pairs.size() % 4) {
            ft.pairs.push_back(inert);
            if (precise) ft.pairs64.push_back(inert64);
        }
        int32_t* lb = ft.lut.data() + (size_t)b * (kLutBins + 1);
        int idx = start;
        for (int k = 0; k < kLutBins; ++k) {
            const double edge = k * dqb;
            while (idx < end && (double)ft.pairs[idx].qh < edge) ++idx;
            lb[k] = idx;
        }
        lb[kLutBins] = end;
    }
    // With a negative weight the relative error of the FP32 sum is unbounded: force the
    // conditioning guard to take the FP64 sum everywhere.
    ft.asum2 = negative ? HUGE_VAL : asum;
    for (int ii = 1; ii <= s.peakmax1 && ii <= n2; ++ii) {
        const int m = L1[ii * 6 + 4], f = L1[ii * 6 + 5];
        if (m == 0 || f == 0) continue;
        ft.iso.push_back(make_float2(L3[ii * 17 + 5], (float)(m * f)));
    }
}

// Per-set constants of the pixel bodies (all kinds).
void set_consts(DevSet& d) {
    const csx_scalars& s = d.s;
    const double zr = (1 - s.sigma * s.sigma) / (s.sigma * s.sigma);
    if (kind_of(s.combo_key) == K_BCT || kind_of(s.combo_key) == K_ISO_LAT) {
        d.cyl_g0 = csx::ref_gamma(zr + 1);
        d.cyl_g1 = csx::ref_gamma(zr - 1 / 2.0);
        d.cyl_g2 = csx::ref_gamma(zr - 3 / 2.0);
    }
}

// The per-set constants of the lattice pixel bodies that do not depend on the tables.
void set_fast_consts(DevSet& d) {
    const csx_scalars& s = d.s;
    double width = 1;
    if (s.dsch) width = 4.0 / s.domainsize;
    width *= s.width_zuf;
    d.kp_coef = -4.0 / (kPiD * s.phiwidth * s.phiwidth) * kLog2e;
    d.kr_const = -4.0 / (kPiD * width * width) * kLog2e;
    d.inv_hpw = 1.0 / (kPiD * width / 2.0);
    d.two_pi_over_v = 2 * kPiD / s.sphno_cubevol;
    // sphere closed-form constants (pq_sphere_fast: zr nudged as the reference nudges it,
    // after R/(zr+1) is formed; fq_sphere_fast: no nudge)
    {
        const double z0 = (1 - s.sigma * s.sigma) / (s.sigma * s.sigma);
        double zr = z0;
        if (zr * (zr - 1) * (zr - 2) * (zr - 3) * (zr - 4) * (zr - 5) == 0) zr += 0.001;
        d.sph_zr = zr;
        d.sph_arg = s.radius / (z0 + 1);
        d.sph_c0 = 1 / (2.0 * zr * (zr - 1) * (zr - 2) * (zr - 3));
        d.sph_c4 = 1 / (zr - 4);
        d.sph_c45 = 1 / ((zr - 4) * (zr - 5));
        d.sfq_zr = z0;
        d.sfq_arg = s.radius / (z0 + 1);
        d.sfq_c0 = 1 / (z0 * (z0 - 1));
        d.sfq_c2 = 1 / (z0 - 2);
    }
    d.nlut = kLutBins;
    d.cut = kDropExponent;
}

void set_fast(DevSet& d, const FastTables& ft) {
    set_fast_consts(d);
    d.npairs = (int)ft.pairs.size();
    d.nsingle = (int)ft.singles.size();
    d.niso = (int)ft.iso.size();
    d.lut_inv_dq = ft.lut_inv_dq;
    d.qh_min = ft.qh_min;
    d.qh_max = ft.qh_max;
    d.asum2 = ft.asum2;
}

// The structural part of a lattice's fast tables, for the device preparation (k_prep_sets):
// the Friedel pairs (ia, ib) in list order, the unpaired reflections and the isotropic
// entries.  The host's pairing (build_fast_tables) takes for each valid reflection the first
// unused valid reflection with the mirrored q_hkl, the same |q_hkl|, g3, peak norm and
// mhkl*fhkl; for the lists of prepare, that is (-h, -k, -l) whenever it has the same
// mhkl*fhkl, since q_hkl is linear in (h, k, l) with exactly mirrored rounding.  The device
// checks every pair's values and flags a pair whose members differ.
struct RowsTemplate {
    std::shared_ptr<const csx::LatticeLists> lists;
    std::vector<int2> pairs;
    std::vector<int32_t> singles;
    std::vector<int2> iso;  // (latpar index, mhkl*fhkl)
};

std::shared_ptr<const RowsTemplate> rows_template(const std::shared_ptr<const csx::LatticeLists>& L, int peakmax1) {
    static std::mutex mu;
    static std::vector<std::pair<int, std::shared_ptr<const RowsTemplate>>> cache;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (const auto& c : cache)
            if (c.second->lists == L && c.first == peakmax1) return c.second;
    }
    auto T = std::make_shared<RowsTemplate>();
    T->lists = L;
    const int n2 = L->n2;
    const int32_t* L2 = L->l2.data();
    std::vector<char> used(n2 + 1, 0);
    int hm = 0;
    for (int ii = 1; ii <= n2; ++ii)
        for (int c = 1; c <= 3; ++c) hm = std::max(hm, std::abs(L2[ii * 6 + c]));
    const int side = 2 * hm + 1;
    std::unordered_map<int64_t, std::vector<int>> byhkl;
    auto key = [&](int h, int k, int l) { return ((int64_t)(h + hm) * side + (k + hm)) * side + (l + hm); };
    for (int ii = 1; ii <= n2; ++ii) byhkl[key(L2[ii * 6 + 1], L2[ii * 6 + 2], L2[ii * 6 + 3])].push_back(ii);
    auto valid = [&](int ii) {
        return (L2[ii * 6 + 1] | L2[ii * 6 + 2] | L2[ii * 6 + 3]) != 0 && L2[ii * 6 + 4] * L2[ii * 6 + 5] != 0;
    };
    for (int ii = 1; ii <= n2; ++ii) {
        if (used[ii] || !valid(ii)) continue;
        used[ii] = 1;
        int jj = 0;
        const auto it = byhkl.find(key(-L2[ii * 6 + 1], -L2[ii * 6 + 2], -L2[ii * 6 + 3]));
        if (it != byhkl.end())
            for (int c : it->second)
                if (!used[c] && valid(c) && L2[c * 6 + 4] * L2[c * 6 + 5] == L2[ii * 6 + 4] * L2[ii * 6 + 5]) {
                    jj = c;
                    break;
                }
        if (!jj) {
            T->singles.push_back(ii);
            continue;
        }
        used[jj] = 1;
        T->pairs.push_back(make_int2(ii, jj));
    }
    const int32_t* L1 = L->l1.data();
    for (int ii = 1; ii <= peakmax1 && ii <= n2; ++ii) {
        const int m = L1[ii * 6 + 4], f = L1[ii * 6 + 5];
        if (m == 0 || f == 0) continue;
        T->iso.push_back(make_int2(ii, m * f));
    }
    std::lock_guard<std::mutex> lock(mu);
    cache.emplace_back(peakmax1, T);
    return T;
}

// Physics-mode orientation quadrature tables of Gaussian-distributed cylinders.
struct CylQuad {
    std::vector<float4> dl;   // kCqDelta x (cos d, sin d, weight, 0)
    std::vector<float> tt;    // kCqPsi/2 x cos(psi)
    std::vector<float2> tab;  // kCqTab+1 x (P(a), h*P'(a))
    float invh = 0.0f, lz = 0.0f;
};

// Gauss-Legendre nodes and weights on [-1, 1]: Newton iteration on the three-term
// recurrence of P_n from the asymptotic first guesses cos(pi (i + 3/4) / (n + 1/2)).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(kPiD * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;  // P_0, P_1
            for (int k = 2; k <= n; ++k) {
                const double pk = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);  // P_n'
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1 - z * z) * dp * dp);
    }
}

// Schulz-averaged rod factor P(a) = (1 - e^A cos B) / (2 n (n+1) a^2) with n = z-1,
// A = -(n/2) ln(1+4a^2), B = n atan(2a), written as -expm1(A) + 2 e^A sin^2(B/2) (no
// cancellation at small a), and its slope; long double on the host.
void rod_value_slope(double z, long double a, long double& P, long double& dP) {
    if (a == 0) { P = 1; dP = 0; return; }
    const long double n = z - 1, q = 1 + 4 * a * a;
    const long double A = -(n / 2) * log1pl(4 * a * a), B = n * atanl(2 * a);
    const long double eA = expl(A), sB = sinl(B / 2);
    const long double g = -expm1l(A) + 2 * eA * sB * sB;
    const long double Ap = -4 * n * a / q, Bp = 2 * n / q;
    const long double gp = -eA * (Ap * cosl(B) - Bp * sinl(B));
    const long double den = 2 * n * (n + 1);
    P = g / (den * a * a);
    dP = (gp * a - 2 * g) / (den * a * a * a);
}

void build_cyl_quad(const csx_scalars& s, CylQuad& cq) {
    const double zl = (1 - s.sigmal * s.sigmal) / (s.sigmal * s.sigmal);
    const double dbeta = s.dbeta * kPiD / 180.0;
    double delmax = dbeta * std::sqrt(std::log(1.0 / 0.00001));
    if (delmax > kPiD / 2.0) delmax = kPiD / 2.0;
    const double dmax = (double)(float)delmax;  // the reference's float limit
    std::vector<double> x, w;
    gauss_legendre(kCqDelta, x, w);
    std::vector<double> wd(kCqDelta);
    double wsum = 0.0;
    for (int k = 0; k < kCqDelta; ++k) {
        const double dl = 0.5 * dmax * (x[k] + 1);
        wd[k] = 0.5 * dmax * w[k] * std::sin(dl) * std::exp(-dl * dl / (dbeta * dbeta));
        wsum += wd[k];
    }
    cq.dl.resize(kCqDelta);
    for (int k = 0; k < kCqDelta; ++k) {
        const double dl = 0.5 * dmax * (x[k] + 1);
        cq.dl[k] = make_float4((float)std::cos(dl), (float)std::sin(dl), (float)(wd[k] / (wsum * kCqPsi)), 0.0f);
    }
    cq.tt.resize(kCqPsi / 2);
    for (int j = 0; j < kCqPsi / 2; ++j) cq.tt[j] = (float)std::cos(kPiD * (j + 0.5) / kCqPsi);
    // |arg| <= (|q| + 1e-10) L/(z+1): bound |q| over the grid (midpoint: sqrt(2) qmax;
    // detector: 4 pi / lambda)
    const double lz = s.length / (zl + 1);
    const double qb = s.use1d ? std::fabs(s.qmax - s.qmin) : (s.centerbeam ? 4 * kPiD / s.wave : std::sqrt(2.0) * s.qmax);
    const double amax = (qb * 1.001 + 1e-9) * lz;
    const double h = amax / kCqTab;
    cq.tab.resize(kCqTab + 1);
    for (int i = 0; i <= kCqTab; ++i) {
        long double P, dP;
        rod_value_slope(zl, (long double)i * h, P, dP);
        cq.tab[i] = make_float2((float)P, (float)(dP * h));
    }
    cq.invh = (float)(kCqTab / amax);
    cq.lz = (float)lz;
}

}  // namespace

struct csx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    bool timing_pending = false;
    bool timing = false;  // bracket csx_render launches with events (costs a drain per launch)
    std::string err;
    bool uploaded = false;
    DevSet set{};
    double carr_h[4 * kCarrLen];  // the uploaded set's coefficient series (host copy)
    DevBuf carr, carr11pm, lat1, lat2, lat3, pairs, pairs64, lut, singles, iso, tiles, fqs_tmp, stats, cqdl, cqtt, cqtab;
    // Radial pre-pass: two tables used alternately, filled on a second stream so that a
    // render's pre-pass overlaps the previous render's image kernel.  ev_rad[s]: table s is
    // filled; ev_img[s]: the image kernel that read table s is done.
    hipStream_t side = nullptr;
    DevBuf radial[2];
    hipEvent_t ev_rad[2] = {nullptr, nullptr}, ev_img[2] = {nullptr, nullptr};
    bool img_recorded[2] = {false, false};
    int rslot = 0;
    DevBuf coldq;        // fix-pass queue of the two-pass lattice render: two counts, then entries
    int cold_epoch = 0;  // which count the next render uses
    hipEvent_t pass_ev[CSX_PASS_RING][2] = {};  // image-pass timing ring (timing on)
    long long pass_n = 0;                        // renders recorded into the ring
    // batch staging
    std::vector<FitSet> bsets;
    std::vector<int> bntiles;
    int bntiles_total = 0;  // tiles of the pending batch (its flat tile list is built on the device)
    DevBuf btiles_d;
    PinnedArena barena;  // host image of the device arena (tables of all sets), pinned
    PinnedArena bres;    // sums and counts of the batch in flight (size = its set count)
    struct Pending {
        size_t off_carr, off_l1, off_l2, off_l3, off_pairs, off_pairs64, off_lut, off_singles, off_iso;
        size_t off_pairs_t, off_iso_t;             // device preparation: the lattice's template (arena)
        size_t dev_l3, dev_pairs, dev_lut, dev_iso;  // and the set's tables (bdev); SIZE_MAX if none
        size_t dev_rad;                              // the set's radial table (brad); SIZE_MAX if none
    };
    std::vector<Pending> bpend;
    // device preparation (k_prep_sets): each template's arrays in this batch's arena (l1, l2,
    // singles, pairs, iso), the device-built tables' buffer and its use, counters
    struct TmplAt { const RowsTemplate* t; size_t off[5]; };
    std::vector<TmplAt> btmpl;
    DevBuf bdev;
    size_t bdev_size = 0;
    // radial tables of the batch's sets (k_batch_radial): their buffer and use, and the first
    // block of each set (nsets + 1 entries)
    DevBuf brad;
    size_t brad_size = 0;
    std::vector<int> brblock0;
    const int* brblock0_d = nullptr;
    bool bprep = false;        // the pending batch has sets whose tables the device builds
    bool device_prep = true;   // the batched fit driver stages lattice sets that way
    long long prep_sets = 0, prep_amb = 0;
    const FitSet* bsets_d = nullptr;  // device records of the last uploaded batch
    DevBuf barena_d, bpart_d, bsums_d;
    DevBuf post[2];  // scratch of the post-processing calls (csx_post.hip)
    int tune_force = -1;  // batched-objective variant for every batch of this context (-1: table)
    bool use_graph = true;  // batched rounds as one HIP graph launch
    BatchGraph bgraph;
};

static int fail(csx_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(ctx, CSX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Tuning table of the batched objective: the kernel variant per kind and detector size class.
// Defaults to variant 0; csx_create loads tuning.txt from the library's directory once (the
// autotuner's measured choices), csx_tune_set / csx_tune_load change it.
namespace {
constexpr int kMaxKinds = 16;
std::mutex g_tune_mu;
int g_tune[kMaxKinds][kTuneSizes] = {};
std::once_flag g_tune_once;

int tune_size_class(int ntiles, int nsets) {
    const double m = nsets > 0 ? (double)ntiles / nsets : 0.0;
    return m <= 64 ? 0 : (m <= 256 ? 1 : (m <= 576 ? 2 : 3));
}

int tune_variant(int kind, int size_class) {
    if (kind < 0 || kind >= kMaxKinds || size_class < 0 || size_class >= kTuneSizes) return 0;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    return g_tune[kind][size_class];
}

// Lines "<combination key> <size class> <variant>" (key in hex or decimal); '#' starts a
// comment.  Returns the number of entries taken, or -1 if the file cannot be read.
int tune_load_file(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) return -1;
    char line[512];
    int n = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (char* h = std::strchr(line, '#')) *h = 0;
        unsigned long key = 0;
        int sc = 0, v = 0;
        char* end = nullptr;
        key = std::strtoul(line, &end, 0);
        if (end == line) continue;
        if (std::sscanf(end, "%d %d", &sc, &v) != 2) continue;
        const int k = kind_of((uint32_t)key);
        if (k < 0 || k >= kMaxKinds || sc < 0 || sc >= kTuneSizes || v < 0 || v >= kTuneVariants) continue;
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tune[k][sc] = v;
        ++n;
    }
    std::fclose(f);
    return n;
}

void tune_load_default() {
    Dl_info info;
    if (dladdr((const void*)&tune_load_default, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const size_t slash = p.find_last_of('/');
        p = (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/tuning.txt";
        (void)tune_load_file(p.c_str());
    }
}
}  // namespace

// Device address of `scratch` on each GPU, looked up once per device with that device
// current (csx_create).
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_lookup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = p;
    return e;
}

TileSum* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return static_cast<TileSum*>(g_scratch_ptr[device]);
}
}  // namespace

// Radial pre-pass for kinds whose form factors depend on |q| only, in the midpoint grid.
template <int K, int P>
static void launch_radial(const DevSet& d, const csx_grid& g, Radial* rtab, hipStream_t st) {
    if constexpr (radial_kind(K)) {
        const int m = std::max(std::max(std::abs(g.xmin), std::abs(g.xmax - 1)),
                               std::max(std::abs(g.ymin), std::abs(g.ymax - 1)));
        const int total = (m + 1) * (m + 2) / 2;
        hipLaunchKernelGGL((k_radial<K, P>), dim3((total + 127) / 128), dim3(256), 0, st, d, m, rtab);
    }
}

// Blocks of the fix pass (one wave each; the pass steps through its queue).
constexpr int kFixBlocks = 256;

// ev: null, or two events recorded around the per-pixel kernel (the hot pass when two).
template <int K, int P>
static void launch_image(const DevSet& d, const csx_grid& g, double* out, unsigned long long* stats,
                         const Radial* rtab, const ColdQueue* cq, hipEvent_t* ev, hipStream_t st) {
    dim3 grid((g.xmax - g.xmin + kTile - 1) / kTile, (g.ymax - g.ymin + kTile - 1) / kTile);
    if (ev) (void)hipEventRecord(ev[0], st);
    if constexpr (lattice_kind(K) && P == CSX_PREC_MIXED) {
        if (cq) {
            hipLaunchKernelGGL((k_image_lat_hot<K, P>), grid, dim3(kBlock), 0, st, d, g, out, stats, rtab, *cq);
            if (ev) (void)hipEventRecord(ev[1], st);
            hipLaunchKernelGGL((k_image_lat_fix<K, P>), dim3(kFixBlocks), dim3(64), 0, st, d, g, out, stats, rtab, *cq);
            return;
        }
    }
    if constexpr (lattice_kind(K))
        hipLaunchKernelGGL((k_image_lat<K, P>), grid, dim3(kBlock), 0, st, d, g, out, stats, rtab);
    else
        hipLaunchKernelGGL((k_image<K, P>), grid, dim3(kBlock), 0, st, d, g, out, stats, rtab);
    if (ev) (void)hipEventRecord(ev[1], st);
}

int csx_qgrid(csx_ctx* c, const csx_grid* g, double* d_q3) {
    if (!c || !g || !d_q3) return CSX_E_ARG;
    if (!c->uploaded) return fail(c, CSX_E_STATE, "csx_qgrid before csx_upload");
    if (g->xmax <= g->xmin || g->ymax <= g->ymin) return fail(c, CSX_E_ARG, "empty grid");
    HIPCHK(c, hipSetDevice(c->device));
    dim3 grid((g->xmax - g->xmin + kTile - 1) / kTile, (g->ymax - g->ymin + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_qgrid, grid, dim3(kBlock), 0, c->stream, c->set.s, *g, d_q3);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CSX_OK;
}

template <int K, int P>
static void launch_fit_one(const FitSet& f, int ntiles, double* fqs, TileSum* part, hipStream_t st) {
    hipLaunchKernelGGL((k_fit_one<K, P>), dim3(ntiles), dim3(kBlock), 0, st, f, fqs, part);
}

// Kernel of the batched objective for kind K, precision P, variant v (one workgroup per tile).
template <int K, int P>
static const void* fit_batch_kernel(int variant) {
    if constexpr (lattice_kind(K) && P == CSX_PREC_MIXED) {
        static const void* const tab[kTuneVariants] = {
            (const void*)k_fit_batch_lat<K, P>, (const void*)k_fit_batch_free<K, P>,
            (const void*)k_fit_batch_cap<K, P, 5>, (const void*)k_fit_batch_cap<K, P, 6>};
        return tab[(variant >= 0 && variant < kTuneVariants) ? variant : 0];
    } else if constexpr (lattice_kind(K)) {
        return (const void*)k_fit_batch_lat<K, P>;
    } else {
        return (const void*)k_fit_batch<K, P>;
    }
}
// Kernel of the batch's radial tables for kind K, precision P (kinds without radial tables
// plan no blocks; their launch returns at once).
template <int K, int P>
static const void* batch_radial_kernel() {
    if constexpr (radial_kind(K)) return (const void*)k_batch_radial<K, P>;
    else return (const void*)k_batch_radial<K_SPHERE, P>;
}
template <int K, int P>
static void launch_batch_radial(const FitSet* sets, int nsets, const int* rblock0, int blocks, hipStream_t st) {
    if (blocks <= 0) return;
    void* args[] = {(void*)&sets, (void*)&nsets, (void*)&rblock0};
    (void)hipLaunchKernel(batch_radial_kernel<K, P>(), dim3(blocks), dim3(256), args, 0, st);
}
template <int K, int P>
static void launch_fit_batch(const FitSet* sets, const int2* tiles, int ntiles, TileSum* part, int variant,
                             hipStream_t st) {
    void* args[] = {(void*)&sets, (void*)&tiles, (void*)&ntiles, (void*)&part};
    (void)hipLaunchKernel(fit_batch_kernel<K, P>(variant), dim3(ntiles), dim3(kBlock), args, 0, st);
}

template <int K, int P>
static void launch_image_batch(const FitSet* sets, const int2* tiles, int ntiles, hipStream_t st) {
    if constexpr (lattice_kind(K))
        hipLaunchKernelGGL((k_image_batch_lat<K, P>), dim3(ntiles), dim3(kBlock), 0, st, sets, tiles, ntiles);
    else
        hipLaunchKernelGGL((k_image_batch<K, P>), dim3(ntiles), dim3(kBlock), 0, st, sets, tiles, ntiles);
}

template <class T>
static int put_table(csx_ctx* c, DevBuf& b, const std::vector<T>& v) {
    HIPCHK(c, b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T)));
    if (!v.empty()) HIPCHK(c, hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return CSX_OK;
}

// Kernel kind of a render: the key's kind, or the physics-mode quadrature kernel for
// Gaussian cylinders in mixed precision.
static int render_kind(const csx_ctx* c, int precision) {
    const int k = kind_of(c->set.s.combo_key);
    if (k == K_CYL_GAUSS && c->set.s.fix_orcase && precision == CSX_PREC_MIXED && c->set.cq_tab) return K_CYL_GAUSS_Q;
    return k;
}

// Radial table of a render (midpoint grid, radial kinds), sized for the grid, and its slot;
// null otherwise.
static int radial_table(csx_ctx* c, const csx_grid& g, Radial** out, int* slot) {
    *out = nullptr;
    *slot = -1;
    if (!radial_kind(kind_of(c->set.s.combo_key)) || c->set.s.centerbeam || c->set.s.use1d) return CSX_OK;
    const long long m = std::max(std::max(std::llabs(g.xmin), std::llabs((long long)g.xmax - 1)),
                                 std::max(std::llabs(g.ymin), std::llabs((long long)g.ymax - 1)));
    if (m > 46000) return CSX_OK;  // triangle index would overflow int: evaluate per pixel
    const int sl = c->rslot;
    const size_t bytes = (size_t)((m + 1) * (m + 2) / 2) * sizeof(Radial);
    if (c->radial[sl].n < bytes) {
        // the table is reallocated: its last reader and writer must be done
        if (c->img_recorded[sl]) HIPCHK(c, hipEventSynchronize(c->ev_img[sl]));
        HIPCHK(c, hipStreamSynchronize(c->side));
    }
    HIPCHK(c, c->radial[sl].ensure(bytes));
    c->rslot ^= 1;
    *slot = sl;
    *out = (Radial*)c->radial[sl].p;
    return CSX_OK;
}

#define DISPATCH_KIND(kind, prec, FN, ...)                                          \
    switch (kind) {                                                                 \
    case K_SPHERE: if (prec) FN<K_SPHERE, 1>(__VA_ARGS__); else FN<K_SPHERE, 0>(__VA_ARGS__); break; \
    case K_FCC: if (prec) FN<K_FCC, 1>(__VA_ARGS__); else FN<K_FCC, 0>(__VA_ARGS__); break; \
    case K_CYL_GAUSS: if (prec) FN<K_CYL_GAUSS, 1>(__VA_ARGS__); else FN<K_CYL_GAUSS, 0>(__VA_ARGS__); break; \
    case K_CYL_ZDIR: if (prec) FN<K_CYL_ZDIR, 1>(__VA_ARGS__); else FN<K_CYL_ZDIR, 0>(__VA_ARGS__); break; \
    case K_CYL_GAUSS_Q: FN<K_CYL_GAUSS_Q, 1>(__VA_ARGS__); break; \
    case K_CUBE: if (prec) FN<K_CUBE, 1>(__VA_ARGS__); else FN<K_CUBE, 0>(__VA_ARGS__); break; \
    case K_CYL_ISO: if (prec) FN<K_CYL_ISO, 1>(__VA_ARGS__); else FN<K_CYL_ISO, 0>(__VA_ARGS__); break; \
    case K_CUBE_ZDIR: if (prec) FN<K_CUBE_ZDIR, 1>(__VA_ARGS__); else FN<K_CUBE_ZDIR, 0>(__VA_ARGS__); break; \
    case K_ISO_LAT: if (prec) FN<K_ISO_LAT, 1>(__VA_ARGS__); else FN<K_ISO_LAT, 0>(__VA_ARGS__); break; \
    default: if (prec) FN<K_BCT, 1>(__VA_ARGS__); else FN<K_BCT, 0>(__VA_ARGS__); break; \
    }

// The same dispatch for a function returning a value (out = FN<K, P>(args)).
#define DISPATCH_KIND_FN(kind, prec, FN, out, ...)                                  \
    switch (kind) {                                                                 \
    case K_SPHERE: out = (prec) ? FN<K_SPHERE, 1>(__VA_ARGS__) : FN<K_SPHERE, 0>(__VA_ARGS__); break; \
    case K_FCC: out = (prec) ? FN<K_FCC, 1>(__VA_ARGS__) : FN<K_FCC, 0>(__VA_ARGS__); break; \
    case K_CYL_GAUSS: out = (prec) ? FN<K_CYL_GAUSS, 1>(__VA_ARGS__) : FN<K_CYL_GAUSS, 0>(__VA_ARGS__); break; \
    case K_CYL_ZDIR: out = (prec) ? FN<K_CYL_ZDIR, 1>(__VA_ARGS__) : FN<K_CYL_ZDIR, 0>(__VA_ARGS__); break; \
    case K_CYL_GAUSS_Q: out = FN<K_CYL_GAUSS_Q, 1>(__VA_ARGS__); break; \
    case K_CUBE: out = (prec) ? FN<K_CUBE, 1>(__VA_ARGS__) : FN<K_CUBE, 0>(__VA_ARGS__); break; \
    case K_CYL_ISO: out = (prec) ? FN<K_CYL_ISO, 1>(__VA_ARGS__) : FN<K_CYL_ISO, 0>(__VA_ARGS__); break; \
    case K_CUBE_ZDIR: out = (prec) ? FN<K_CUBE_ZDIR, 1>(__VA_ARGS__) : FN<K_CUBE_ZDIR, 0>(__VA_ARGS__); break; \
    case K_ISO_LAT: out = (prec) ? FN<K_ISO_LAT, 1>(__VA_ARGS__) : FN<K_ISO_LAT, 0>(__VA_ARGS__); break; \
    default: out = (prec) ? FN<K_BCT, 1>(__VA_ARGS__) : FN<K_BCT, 0>(__VA_ARGS__); break; \
    }

// One render: the radial pre-pass (if any) on the side stream, after the image kernel that
// last read the same table; the image kernel on the context stream, after the pre-pass.
// Back-to-back renders thus overlap a pre-pass with the previous image kernel.
static int render_launch(csx_ctx* c, int kind, int precision, const csx_grid& g, double* out,
                         unsigned long long* stats) {
    Radial* rtab = nullptr;
    int sl = -1;
    if (int rc = radial_table(c, g, &rtab, &sl)) return rc;
    if (rtab) {
        if (c->img_recorded[sl]) HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_img[sl], 0));
        DISPATCH_KIND(kind, precision, launch_radial, c->set, g, rtab, c->side);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev_rad[sl], c->side));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_rad[sl], 0));
    }
    // Lattice kinds in mixed precision render in two passes (hot, then fix over a queue of at
    // most one slot per wave), unless unpaired reflections send every pixel to FP64.
    ColdQueue cq{};
    const ColdQueue* cqp = nullptr;
    if (lattice_kind(kind) && precision == CSX_PREC_MIXED && c->set.nsingle == 0) {
        const size_t nwaves = 4 * (size_t)((g.xmax - g.xmin + kTile - 1) / kTile) *
                              (size_t)((g.ymax - g.ymin + kTile - 1) / kTile);
        const size_t need = 256 + nwaves * (sizeof(unsigned long long) + 64 * sizeof(ColdPix));
        if (need > c->coldq.n) {
            HIPCHK(c, c->coldq.ensure(need));
            HIPCHK(c, hipMemsetAsync(c->coldq.p, 0, 256, c->stream));
        }
        int* counts = static_cast<int*>(c->coldq.p);
        cq.count = counts + c->cold_epoch;
        cq.next_count = counts + (c->cold_epoch ^ 1);
        cq.mask = reinterpret_cast<unsigned long long*>(static_cast<char*>(c->coldq.p) + 256);
        cq.ent = reinterpret_cast<ColdPix*>(static_cast<char*>(c->coldq.p) + 256 + nwaves * sizeof(unsigned long long));
        c->cold_epoch ^= 1;
        cqp = &cq;
    }
    hipEvent_t* ev = nullptr;
    if (c->timing) {
        hipEvent_t* e = c->pass_ev[c->pass_n % CSX_PASS_RING];
        for (int k = 0; k < 2; ++k)
            if (!e[k]) HIPCHK(c, hipEventCreate(&e[k]));
        ev = e;
        ++c->pass_n;
    }
    DISPATCH_KIND(kind, precision, launch_image, c->set, g, out, stats, rtab, cqp, ev, c->stream);
    HIPCHK(c, hipGetLastError());
    if (rtab) {
        HIPCHK(c, hipEventRecord(c->ev_img[sl], c->stream));
        c->img_recorded[sl] = true;
    }
    return CSX_OK;
}

extern "C" {

int csx_create(int device, csx_ctx** out) {
    if (!out) return CSX_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CSX_E_HIP;
    if (device < 0 || device >= n) return CSX_E_ARG;
    std::call_once(g_tune_once, tune_load_default);
    csx_ctx* c = new csx_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_rad[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_rad[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_img[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_img[1], hipEventDisableTiming) != hipSuccess ||
        scratch_lookup(device) != hipSuccess) {
        c->own_stream = c->stream != nullptr;
        csx_destroy(c);
        return CSX_E_HIP;
    }
    c->own_stream = true;
    *out = c;
    return CSX_OK;
}

void csx_destroy(csx_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    c->barena.release();
    c->bres.release();
    for (DevBuf* b : {&c->carr, &c->carr11pm, &c->lat1, &c->lat2, &c->lat3, &c->pairs, &c->pairs64, &c->lut, &c->singles,
                      &c->iso, &c->stats, &c->radial[0], &c->radial[1], &c->cqdl, &c->cqtt, &c->cqtab, &c->tiles,
                      &c->fqs_tmp, &c->barena_d, &c->bpart_d, &c->bsums_d, &c->post[0], &c->post[1], &c->coldq,
                      &c->bdev, &c->btiles_d, &c->brad})
        b->release();
    if (c->bgraph.exec) (void)hipGraphExecDestroy(c->bgraph.exec);
    if (c->bgraph.graph) (void)hipGraphDestroy(c->bgraph.graph);
    for (hipEvent_t e : {c->ev0, c->ev1, c->ev_rad[0], c->ev_rad[1], c->ev_img[0], c->ev_img[1]})
        if (e) (void)hipEventDestroy(e);
    for (auto& pe : c->pass_ev)
        for (hipEvent_t e : pe)
            if (e) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int csx_set_stream(csx_ctx* c, void* s) {
    if (!c) return CSX_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    if (s) { c->stream = (hipStream_t)s; c->own_stream = false; }
    else {
        HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    return CSX_OK;
}

int csx_synchronize(csx_ctx* c) {
    if (!c) return CSX_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CSX_OK;
}

int csx_last_error(csx_ctx* c, char* buf, int len) {
    if (!c || !buf || len <= 0) return CSX_E_ARG;
    std::snprintf(buf, (size_t)len, "%s", c->err.c_str());
    return CSX_OK;
}

int csx_tune_set(uint32_t combo_key, int size_class, int variant) {
    const int k = kind_of(combo_key);
    if (k < 0 || k >= kMaxKinds || size_class < -1 || size_class >= kTuneSizes || variant < 0 ||
        variant >= kTuneVariants)
        return CSX_E_ARG;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (int sc = 0; sc < kTuneSizes; ++sc)
        if (size_class < 0 || sc == size_class) g_tune[k][sc] = variant;
    return CSX_OK;
}

int csx_tune_get(uint32_t combo_key, int size_class) {
    const int k = kind_of(combo_key);
    if (k < 0 || k >= kMaxKinds || size_class < 0 || size_class >= kTuneSizes) return CSX_E_ARG;
    return tune_variant(k, size_class);
}

int csx_tune_load(const char* path) {
    if (!path) return CSX_E_ARG;
    const int n = tune_load_file(path);
    return n < 0 ? CSX_E_ARG : n;
}

int csx_set_batch_graph(csx_ctx* c, int enable) {
    if (!c) return CSX_E_ARG;
    if (c->bres.size) return fail(c, CSX_E_STATE, "a batch is in flight");
    c->use_graph = enable != 0;
    return CSX_OK;
}

long long csx_batch_graph_launches(const csx_ctx* c) { return c ? c->bgraph.launches : -1; }

int csx_tune_force(csx_ctx* c, int variant) {
    if (!c || variant < -1 || variant >= kTuneVariants) return CSX_E_ARG;
    c->tune_force = variant;
    return CSX_OK;
}

int csx_set_timing(csx_ctx* c, int enable) {
    if (!c) return CSX_E_ARG;
    c->timing = enable != 0;
    c->pass_n = 0;
    return CSX_OK;
}

int csx_pass_times(csx_ctx* c, int max, double* ms) {
    if (!c || max < 0 || (max > 0 && !ms)) return CSX_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const long long have = std::min<long long>(c->pass_n, CSX_PASS_RING);
    const int n = (int)std::min<long long>(have, max);
    for (int k = 0; k < n; ++k) {
        hipEvent_t* e = c->pass_ev[(c->pass_n - n + k) % CSX_PASS_RING];
        HIPCHK(c, hipEventSynchronize(e[1]));
        float t = 0.f;
        HIPCHK(c, hipEventElapsedTime(&t, e[0], e[1]));
        ms[k] = t;
    }
    return n;
}

double csx_last_kernel_ms(csx_ctx* c) {
    if (!c) return -1.0;
    if (c->timing_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(c->ev1) == hipSuccess && hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess)
            c->last_ms = ms;
        c->timing_pending = false;
    }
    return c->last_ms;
}

static int check_scalars(csx_ctx* c, const csx_scalars* s) {
    if (!s) return fail(c, CSX_E_ARG, "null scalars");
    if (kind_of(s->combo_key) < 0) return fail(c, CSX_E_ARG, "combination key not supported by the device path");
    if (s->peakmax1 < 0 || s->peakmax2 < 0 || s->peakmax1 >= CSX_LATLEN || s->peakmax2 >= CSX_LATLEN)
        return fail(c, CSX_E_ARG, "peakmax out of range");
    if (s->peakmax1 > s->peakmax2 && s->peakmax1 > 0)
        return fail(c, CSX_E_ARG, "isotropic list longer than the anisotropic table");
    if (s->zmax <= 0) return fail(c, CSX_E_ARG, "zmax must be positive");
    return CSX_OK;
}

// The series the device reads (carr_of) from the host's 18 x CSX_NCOEF coefficient block.
static void gather_carr(const double* full, double* out) {
    for (int j = 0; j < 4; ++j) std::memcpy(out + j * kCarrLen, full + kCarrArrays[j] * CSX_NCOEF, kCarrLen * sizeof(double));
}

int csx_upload(csx_ctx* c, const csx_scalars* s, const double* h_carr, const double* h_carr11pm,
               const int32_t* h_l1, const int32_t* h_l2, const float* h_l3) {
    if (!c) return CSX_E_ARG;
    int rc = check_scalars(c, s);
    if (rc) return rc;
    if (!h_carr) return fail(c, CSX_E_ARG, "null coefficient table");
    const int k = kind_of(s->combo_key);
    if (!h_carr11pm && (k == K_CYL_GAUSS || k == K_CUBE_ZDIR))
        return fail(c, CSX_E_ARG, "this combination reads the carr11pm table: pass it");
    HIPCHK(c, hipSetDevice(c->device));
    // a radial pre-pass still running reads the tables about to be replaced
    HIPCHK(c, hipStreamSynchronize(c->side));
    const bool lattice = (s->combo_key >> 8 & 0xff) != 12;
    HIPCHK(c, c->carr.ensure(4 * kCarrLen * sizeof(double)));
    gather_carr(h_carr, c->carr_h);
    HIPCHK(c, hipMemcpyAsync(c->carr.p, c->carr_h, 4 * kCarrLen * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->set = DevSet{};
    c->set.s = *s;
    set_consts(c->set);
    c->set.carr = (const double*)c->carr.p;
    if (h_carr11pm) {
        HIPCHK(c, c->carr11pm.ensure(CSX_NIM2 * CSX_NIM2 * sizeof(double)));
        HIPCHK(c, hipMemcpyAsync(c->carr11pm.p, h_carr11pm, CSX_NIM2 * CSX_NIM2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        c->set.carr11pm = (const double*)c->carr11pm.p;
    }
    FastTables ft;
    if (lattice) {
        if (!h_l1 || !h_l2 || !h_l3) return fail(c, CSX_E_ARG, "null reflection tables");
        const size_t b1 = (size_t)(s->peakmax1 + 1) * 6 * 4, b2 = (size_t)(s->peakmax2 + 1) * 6 * 4,
                     b3 = (size_t)(s->peakmax2 + 1) * 17 * 4;
        HIPCHK(c, c->lat1.ensure(b1));
        HIPCHK(c, c->lat2.ensure(b2));
        HIPCHK(c, c->lat3.ensure(b3));
        HIPCHK(c, hipMemcpyAsync(c->lat1.p, h_l1, b1, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->lat2.p, h_l2, b2, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->lat3.p, h_l3, b3, hipMemcpyHostToDevice, c->stream));
        build_fast_tables(*s, h_l1, h_l2, h_l3, ft);
        c->set.latpar1 = (const int32_t*)c->lat1.p;
        c->set.latpar2 = (const int32_t*)c->lat2.p;
        c->set.latpar3 = (const float*)c->lat3.p;
        rc = put_table(c, c->pairs, ft.pairs);
        if (!rc && !ft.pairs64.empty()) rc = put_table(c, c->pairs64, ft.pairs64);
        if (!rc) rc = put_table(c, c->lut, ft.lut);
        if (!rc) rc = put_table(c, c->singles, ft.singles);
        if (!rc) rc = put_table(c, c->iso, ft.iso);
        if (rc) return rc;
        c->set.pairs = (const PairRec*)c->pairs.p;
        c->set.pairs64 = ft.pairs64.empty() ? nullptr : (const PairRec64*)c->pairs64.p;
        c->set.lut = (const int32_t*)c->lut.p;
        c->set.singles = (const int32_t*)c->singles.p;
        c->set.iso = (const float2*)c->iso.p;
        set_fast(c->set, ft);
    }
    CylQuad cq;
    if (kind_of(s->combo_key) == K_CYL_GAUSS && s->fix_orcase) {
        build_cyl_quad(*s, cq);
        rc = put_table(c, c->cqdl, cq.dl);
        if (!rc) rc = put_table(c, c->cqtt, cq.tt);
        if (!rc) rc = put_table(c, c->cqtab, cq.tab);
        if (rc) return rc;
        c->set.cq_delta = (const float4*)c->cqdl.p;
        c->set.cq_t = (const float*)c->cqtt.p;
        c->set.cq_tab = (const float2*)c->cqtab.p;
        c->set.cq_invh = cq.invh;
        c->set.cq_lz = cq.lz;
    }
    // The host staging memory is reused: wait until the copies have consumed it.
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->uploaded = true;
    return CSX_OK;
}



int csx_render(csx_ctx* c, const csx_grid* g, int precision, double* d_out) {
    if (c) precision = effective_precision(c->set.s.combo_key, precision);
    if (!c || !g || !d_out) return CSX_E_ARG;
    if (!c->uploaded) return fail(c, CSX_E_STATE, "csx_render before csx_upload");
    if (g->xmax <= g->xmin || g->ymax <= g->ymin) return fail(c, CSX_E_ARG, "empty grid");
    if (precision != CSX_PREC_REF64 && precision != CSX_PREC_MIXED) return fail(c, CSX_E_ARG, "bad precision");
    HIPCHK(c, hipSetDevice(c->device));
    const int kind = render_kind(c, precision);
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (int rc = render_launch(c, kind, precision, *g, d_out, nullptr)) return rc;
    if (c->timing) {
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        c->timing_pending = true;  // asynchronous: csx_last_kernel_ms waits for the events
    }
    return CSX_OK;
}

int csx_render_stats(csx_ctx* c, const csx_grid* g, int precision, double* d_out, uint64_t* h_stats) {
    if (c) precision = effective_precision(c->set.s.combo_key, precision);
    if (!c || !g || !d_out || !h_stats) return CSX_E_ARG;
    if (!c->uploaded) return fail(c, CSX_E_STATE, "csx_render_stats before csx_upload");
    if (g->xmax <= g->xmin || g->ymax <= g->ymin) return fail(c, CSX_E_ARG, "empty grid");
    if (precision != CSX_PREC_REF64 && precision != CSX_PREC_MIXED) return fail(c, CSX_E_ARG, "bad precision");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->stats.ensure(CSX_NSTATS * sizeof(uint64_t)));
    const std::vector<uint64_t> zero(CSX_NSTATS, 0);
    HIPCHK(c, hipMemcpyAsync(c->stats.p, zero.data(), CSX_NSTATS * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
    const int kind = render_kind(c, precision);
    if (int rc = render_launch(c, kind, precision, *g, d_out, (unsigned long long*)c->stats.p)) return rc;
    HIPCHK(c, hipMemcpyAsync(h_stats, c->stats.p, CSX_NSTATS * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CSX_OK;
}


int csx_fit_fqs(csx_ctx* c, int w, int h, const csx_mask* mask, int precision_in, const double* d_data,
                double* h_sum, int64_t* h_cnt, int64_t* h_nan, double* d_fqs) {
    if (!c || !mask || !d_data || !h_sum || w <= 0 || h <= 0) return CSX_E_ARG;
    if (!c->uploaded) return fail(c, CSX_E_STATE, "csx_fit_fqs before csx_upload");
    const csx_scalars& s = c->set.s;
    const int precision = effective_precision(s.combo_key, precision_in);
    if ((long long)(s.xmax - s.xmin) * (s.ymax - s.ymin) < (long long)w * h || (s.xmax - s.xmin) != w)
        return fail(c, CSX_E_ARG, "fit array must match the image width and fit inside the image");
    if (s.iimin != s.xmin || s.iimin != s.ymin)
        return fail(c, CSX_E_ARG, "fit needs the full (Q4) or first-quadrant (Q1) image layout");
    HIPCHK(c, hipSetDevice(c->device));
    FitSet f{};
    f.d = c->set;
    f.m = *mask;
    f.data = d_data;
    f.w = w;
    f.h = h;
    f.tiles_x = (w + kTile - 1) / kTile;
    const int ntiles = f.tiles_x * ((h + kTile - 1) / kTile);
    HIPCHK(c, c->tiles.ensure((size_t)ntiles * sizeof(TileSum) + sizeof(SetSum)));
    TileSum* part = (TileSum*)c->tiles.p;
    SetSum* total = (SetSum*)(part + ntiles);
    const int kind = kind_of(s.combo_key);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    DISPATCH_KIND(kind, precision, launch_fit_one, f, ntiles, d_fqs, part, c->stream);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(k_sum_tiles, dim3(1), dim3(kBlock), 0, c->stream, part, ntiles, total);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    SetSum tot{};
    HIPCHK(c, hipMemcpyAsync(&tot, total, sizeof(SetSum), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->timing_pending = false;
    *h_sum = tot.sum;
    if (h_cnt) *h_cnt = tot.cnt;
    if (h_nan) *h_nan = tot.nan;
    return CSX_OK;
}

// The objective of csx_fit_fqs with its tile partials in the static scratch array instead of
// context memory, and their sum taken on the device: nothing is allocated, on the device or
// on the host.
int csx_fit_fqs_noalloc(csx_ctx* c, int w, int h, const csx_mask* mask, int precision_in, const double* d_data,
                        double* h_sum, int64_t* h_cnt, int64_t* h_nan, double* d_fqs) {
    if (!c || !mask || !d_data || !h_sum || w <= 0 || h <= 0) return CSX_E_ARG;
    if (!c->uploaded) return fail(c, CSX_E_STATE, "csx_fit_fqs_noalloc before csx_upload");
    const csx_scalars& s = c->set.s;
    const int precision = effective_precision(s.combo_key, precision_in);
    if ((long long)(s.xmax - s.xmin) * (s.ymax - s.ymin) < (long long)w * h || (s.xmax - s.xmin) != w)
        return fail(c, CSX_E_ARG, "fit array must match the image width and fit inside the image");
    if (s.iimin != s.xmin || s.iimin != s.ymin)
        return fail(c, CSX_E_ARG, "fit needs the full (Q4) or first-quadrant (Q1) image layout");
    const int tiles_x = (w + kTile - 1) / kTile;
    const long long ntiles = (long long)tiles_x * ((h + kTile - 1) / kTile);
    if (ntiles > kScratchTiles)
        return fail(c, CSX_E_ARG, "csx_fit_fqs_noalloc: " + std::to_string(ntiles) + " tiles of 16x16 exceed its " +
                                      std::to_string(kScratchTiles) + " (at most 2048 x 2048 fit pixels)");
    TileSum* part = scratch_address(c->device);
    if (!part) return fail(c, CSX_E_STATE, "csx_fit_fqs_noalloc: no scratch address for this device");
    HIPCHK(c, hipSetDevice(c->device));
    FitSet f{};
    f.d = c->set;
    f.m = *mask;
    f.data = d_data;
    f.w = w;
    f.h = h;
    f.tiles_x = tiles_x;
    const int kind = kind_of(s.combo_key);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    DISPATCH_KIND(kind, precision, launch_fit_one, f, (int)ntiles, d_fqs, part, c->stream);
    HIPCHK(c, hipGetLastError());
    SetSum* d_total = (SetSum*)(part + kScratchTiles);
    hipLaunchKernelGGL(k_sum_tiles, dim3(1), dim3(kBlock), 0, c->stream, part, (int)ntiles, d_total);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    SetSum total{};
    HIPCHK(c, hipMemcpyAsync(&total, d_total, sizeof(SetSum), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->timing_pending = false;
    *h_sum = total.sum;
    if (h_cnt) *h_cnt = total.cnt;
    if (h_nan) *h_nan = total.nan;
    return CSX_OK;
}

// ---- batched objective ------------------------------------------------------
}  // extern "C"

// The whole image of an ExpandImage parameter set from its computed block (ordered on the
// context's stream): d_region holds the block in IDX order, d_full receives the image array.
int csx_expand_image(csx_ctx* c, const double* d_region, double* d_full) {
    if (!c || !d_region || !d_full) return CSX_E_ARG;
    if (!c->uploaded) return fail(c, CSX_E_STATE, "csx_expand_image before csx_upload");
    const csx_scalars& s = c->set.s;
    const int z = s.zmax;
    int quad = 0;
    if (s.xmin == 0 && s.xmax == z && s.ymin == 0 && s.ymax == z) quad = 1;
    if (s.xmin == 0 && s.xmax == z && s.ymin == -z && s.ymax == z) quad = 2;
    if (!quad || s.axmin != -z || s.axmax != z || s.aymin != -z || s.aymax != z)
        return fail(c, CSX_E_ARG, "csx_expand_image: not an ExpandImage parameter set");
    HIPCHK(c, hipSetDevice(c->device));
    const long long n = 4LL * z * z;
    hipLaunchKernelGGL(k_expand, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, d_region,
                       s.xmax - s.xmin, s.xmin, s.ymin, z, quad, d_full);
    HIPCHK(c, hipGetLastError());
    return CSX_OK;
}


// Append bytes at the next 256-byte boundary; returns their offset.
static size_t arena_put(std::vector<char>& a, const void* p, size_t bytes) {
    const size_t off = (a.size() + 255) & ~size_t(255);
    a.resize(off + bytes);
    if (bytes) std::memcpy(a.data() + off, p, bytes);
    return off;
}

struct csx_staged {
    struct Offsets { size_t off_carr, off_l1, off_l2, off_l3, off_pairs, off_pairs64, off_lut, off_singles, off_iso; };
    FitSet f;
    Offsets pd;
    std::vector<char> arena;
    int ntiles = 0;
    int rc = CSX_OK;
    std::string err;
    std::shared_ptr<const RowsTemplate> tmpl;  // tables built on the device (k_prep_sets)
};
csx_staged* csx_staged_new() { return new csx_staged(); }
void csx_staged_free(csx_staged* s) { delete s; }

// The checks and the part of staging common to both kinds of staging: the set record and the
// coefficient tables.
static bool stage_begin(csx_staged& st, const csx_scalars* s, const double* h_carr, int w, int h,
                        const csx_mask* mask, const double* d_data) {
    st.rc = CSX_OK;
    st.err.clear();
    st.arena.clear();
    st.tmpl.reset();
    auto bad = [&](int code, const char* msg) {
        st.rc = code;
        st.err = msg;
        return false;
    };
    if (!s || !h_carr || !mask || !d_data || w <= 0 || h <= 0) return bad(CSX_E_ARG, "null argument or empty fit array");
    if (kind_of(s->combo_key) < 0) return bad(CSX_E_ARG, "combination key not supported by the device path");
    // the batch carries no carr11pm table: kinds that read it render and fit one set at a time
    if (kind_of(s->combo_key) == K_CYL_GAUSS || kind_of(s->combo_key) == K_CUBE_ZDIR)
        return bad(CSX_E_ARG, "this combination needs carr11pm, which batches do not carry");
    if (s->peakmax1 < 0 || s->peakmax2 < 0 || s->peakmax1 >= CSX_LATLEN || s->peakmax2 >= CSX_LATLEN)
        return bad(CSX_E_ARG, "peakmax out of range");
    if (s->peakmax1 > s->peakmax2 && s->peakmax1 > 0)
        return bad(CSX_E_ARG, "isotropic list longer than the anisotropic table");
    if (s->zmax <= 0) return bad(CSX_E_ARG, "zmax must be positive");
    if ((s->xmax - s->xmin) != w || (long long)(s->xmax - s->xmin) * (s->ymax - s->ymin) < (long long)w * h)
        return bad(CSX_E_ARG, "fit array must match the image width and fit inside the image");
    if (s->iimin != s->xmin || s->iimin != s->ymin)
        return bad(CSX_E_ARG, "fit needs the full (Q4) or first-quadrant (Q1) image layout");
    FitSet& f = st.f;
    f = FitSet{};
    f.d.s = *s;
    set_consts(f.d);
    f.m = *mask;
    f.data = d_data;
    f.w = w;
    f.h = h;
    f.tiles_x = (w + kTile - 1) / kTile;
    st.ntiles = f.tiles_x * ((h + kTile - 1) / kTile);
    auto& pd = st.pd;
    pd = csx_staged::Offsets{SIZE_MAX, SIZE_MAX, SIZE_MAX, SIZE_MAX, SIZE_MAX, SIZE_MAX, SIZE_MAX, SIZE_MAX, SIZE_MAX};
    double cc[4 * kCarrLen];
    gather_carr(h_carr, cc);
    pd.off_carr = arena_put(st.arena, cc, sizeof cc);
    return true;
}

// Host part of staging one set: checks, fast tables and the set's bytes, into `out`.
// Touches no context state, so many sets can be built on different threads at once.
void csx_stage_build(csx_staged* out, const csx_scalars* s, const double* h_carr, const int32_t* h_l1,
                     const int32_t* h_l2, const float* h_l3, int w, int h, const csx_mask* mask,
                     const double* d_data) {
    csx_staged& st = *out;
    if (!stage_begin(st, s, h_carr, w, h, mask, d_data)) return;
    const bool lattice = (s->combo_key >> 8 & 0xff) != 12;
    if (!lattice) return;
    if (!h_l1 || !h_l2 || !h_l3) {
        st.rc = CSX_E_ARG;
        st.err = "null reflection tables";
        return;
    }
    auto& pd = st.pd;
    pd.off_l1 = arena_put(st.arena, h_l1, (size_t)(s->peakmax1 + 1) * 6 * 4);
    pd.off_l2 = arena_put(st.arena, h_l2, (size_t)(s->peakmax2 + 1) * 6 * 4);
    pd.off_l3 = arena_put(st.arena, h_l3, (size_t)(s->peakmax2 + 1) * 17 * 4);
    FastTables ft;
    build_fast_tables(*s, h_l1, h_l2, h_l3, ft);
    // kinds with the precise loop read only the FP64 records
    if (ft.pairs64.empty()) pd.off_pairs = arena_put(st.arena, ft.pairs.data(), ft.pairs.size() * sizeof(PairRec));
    else pd.off_pairs64 = arena_put(st.arena, ft.pairs64.data(), ft.pairs64.size() * sizeof(PairRec64));
    pd.off_lut = arena_put(st.arena, ft.lut.data(), ft.lut.size() * sizeof(int32_t));
    pd.off_singles = arena_put(st.arena, ft.singles.data(), ft.singles.size() * sizeof(int32_t));
    pd.off_iso = arena_put(st.arena, ft.iso.data(), ft.iso.size() * sizeof(float2));
    set_fast(st.f.d, ft);
}

// Staging of a lattice set whose tables the device builds (k_prep_sets): the host part is the
// scalars, the coefficient tables and the rotated reciprocal bases; the reflection lists and
// their pairing are shared by every set of the lattice and go over once per batch.
void csx_stage_build_rows(csx_staged* out, const csx::PreparedSet& P, int w, int h, const csx_mask* mask,
                          const double* d_data) {
    csx_staged& st = *out;
    const csx_scalars* s = &P.s;
    if ((s->combo_key >> 8 & 0xff) != 12 && P.lists && P.lists->n2 == 0 && s->peakmax2 == 0 && s->peakmax1 == 0) {
        // no reflection (HKLmax 0, HCP, orthorhombic): the empty tables go over as they are
        static const float l3[17] = {};
        return csx_stage_build(out, s, P.carr.data(), P.lists->l1.data(), P.lists->l2.data(), l3, w, h, mask,
                               d_data);
    }
    if (!stage_begin(st, s, P.carr.data(), w, h, mask, d_data)) return;
    auto bad = [&](const char* msg) {
        st.rc = CSX_E_ARG;
        st.err = msg;
        st.tmpl.reset();
    };
    if ((s->combo_key >> 8 & 0xff) == 12 || !P.lists)
        return bad("device preparation needs a lattice set prepared with rows_on_device");
    if ((s->combo_key >> 16 & 0xff) != 2) return bad("device preparation builds the Gaussian peak's tables only");
    if (P.lists->n2 != s->peakmax2 || P.lists->n1 != s->peakmax1)
        return bad("reflection lists do not match the set's peakmax");
    st.tmpl = rows_template(P.lists, s->peakmax1);
    if ((int)st.tmpl->pairs.size() > kMaxPairs) return bad("too many Friedel pairs for the device preparation");
    FitSet& f = st.f;
    set_fast_consts(f.d);
    f.d.nsingle = (int)st.tmpl->singles.size();
    f.d.niso = (int)st.tmpl->iso.size();
    PrepIn& pin = f.pin;
    std::memcpy(pin.R, P.rot, sizeof pin.R);
    std::memcpy(pin.RT, P.rott, sizeof pin.RT);
    pin.on = 1;
    pin.n2 = s->peakmax2;
    pin.np = (int)st.tmpl->pairs.size();
    pin.niso = (int)st.tmpl->iso.size();
}

// Append built sets to the context's pending batch: offsets in order, then the byte copies
// into the pinned arena on up to `threads` threads.
int csx_stage_commit_many(csx_ctx* c, int n, const csx_staged* const* sts, int threads) {
    if (!c || n < 0 || (n > 0 && !sts)) return CSX_E_ARG;
    std::vector<size_t> base(n);
    size_t off = c->barena.size;
    for (int i = 0; i < n; ++i) {
        const csx_staged& st = *sts[i];
        if (st.rc != CSX_OK) return fail(c, st.rc, st.err);
        const int k0 = c->bsets.empty() ? kind_of(sts[0]->f.d.s.combo_key) : kind_of(c->bsets[0].d.s.combo_key);
        if (kind_of(st.f.d.s.combo_key) != k0) return fail(c, CSX_E_ARG, "all sets of a batch must share the combination");
        off = (off + 255) & ~size_t(255);
        base[i] = off;
        off += st.arena.size();
    }
    HIPCHK(c, c->barena.reserve(off));
    auto copy = [&](int i) {
        if (!sts[i]->arena.empty()) std::memcpy(c->barena.p + base[i], sts[i]->arena.data(), sts[i]->arena.size());
    };
    const int nt = std::max(1, std::min(threads, n / 16));
    if (nt == 1) {
        for (int i = 0; i < n; ++i) copy(i);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                for (int i = t; i < n; i += nt) copy(i);
            });
        for (auto& th : pool) th.join();
    }
    c->barena.size = off;
    // device preparation: each lattice template once per batch in the arena, each set's
    // tables in the device buffer
    auto put = [&](const void* src, size_t bytes) {
        const size_t o = (c->barena.size + 255) & ~size_t(255);
        if (c->barena.reserve(o + bytes) != hipSuccess) return SIZE_MAX;
        if (bytes) std::memcpy(c->barena.p + o, src, bytes);
        c->barena.size = o + bytes;
        return o;
    };
    for (int i = 0; i < n; ++i) {
        const RowsTemplate* t = sts[i]->tmpl.get();
        if (!t) continue;
        bool found = false;
        for (const auto& e : c->btmpl) found = found || e.t == t;
        if (!found) {
            csx_ctx::TmplAt e{t, {}};
            e.off[0] = put(t->lists->l1.data(), t->lists->l1.size() * sizeof(int32_t));
            e.off[1] = put(t->lists->l2.data(), t->lists->l2.size() * sizeof(int32_t));
            e.off[2] = put(t->singles.data(), t->singles.size() * sizeof(int32_t));
            e.off[3] = put(t->pairs.data(), t->pairs.size() * sizeof(int2));
            e.off[4] = put(t->iso.data(), t->iso.size() * sizeof(int2));
            for (size_t o : e.off)
                if (o == SIZE_MAX) return fail(c, CSX_E_HIP, "pinned staging memory");
            c->btmpl.push_back(e);
        }
    }
    // radial table of a set of a radial kind in the midpoint grid: its largest |index| over
    // the set's pixels, its entries in the batch's radial buffer and its blocks
    auto radial_of = [&](FitSet& f, csx_ctx::Pending& pd) {
        pd.dev_rad = SIZE_MAX;
        f.rtab = nullptr;
        f.rad_m = 0;
        const csx_scalars& ss = f.d.s;
        if (c->brblock0.empty()) c->brblock0.push_back(0);
        int blocks = 0;
        if (radial_kind(kind_of(ss.combo_key)) && !ss.centerbeam && !ss.use1d) {
            const long long m = std::max(std::max(std::llabs(ss.xmin), std::llabs((long long)ss.xmin + f.w - 1)),
                                         std::max(std::llabs(ss.ymin), std::llabs((long long)ss.ymin + f.h - 1)));
            if (m <= 46000) {  // the triangle index stays an int
                const size_t entries = (size_t)((m + 1) * (m + 2) / 2);
                pd.dev_rad = (c->brad_size + 255) & ~size_t(255);
                c->brad_size = pd.dev_rad + entries * sizeof(Radial);
                f.rad_m = (int)m;
                blocks = (int)((entries + 127) / 128);
            }
        }
        c->brblock0.push_back(c->brblock0.back() + blocks);
    };
    for (int i = 0; i < n; ++i) {
        const csx_staged& st = *sts[i];
        const size_t b = base[i];
        auto at = [b](size_t o) { return o == SIZE_MAX ? SIZE_MAX : b + o; };
        csx_ctx::Pending pd{};
        pd.off_pairs_t = pd.off_iso_t = SIZE_MAX;
        pd.dev_l3 = pd.dev_pairs = pd.dev_lut = pd.dev_iso = SIZE_MAX;
        if (st.tmpl) {
            const RowsTemplate* t = st.tmpl.get();
            const csx_ctx::TmplAt* e = nullptr;
            for (const auto& x : c->btmpl)
                if (x.t == t) e = &x;
            auto dev = [&](size_t bytes) {
                const size_t o = (c->bdev_size + 255) & ~size_t(255);
                c->bdev_size = o + bytes;
                return o;
            };
            const int np = (int)t->pairs.size();
            const bool precise = precise_kind(kind_of(st.f.d.s.combo_key));
            pd.dev_l3 = dev((size_t)(st.f.pin.n2 + 1) * 17 * sizeof(float));
            pd.dev_pairs = dev((size_t)(3 * np + 4 * kSegments) * (precise ? sizeof(PairRec64) : sizeof(PairRec)));
            pd.dev_lut = dev((size_t)kSegments * (kLutBins + 1) * sizeof(int32_t));
            pd.dev_iso = dev((size_t)t->iso.size() * sizeof(float2));
            pd.off_l1 = e->off[0];
            pd.off_l2 = e->off[1];
            pd.off_singles = e->off[2];
            pd.off_pairs_t = e->off[3];
            pd.off_iso_t = e->off[4];
            pd.off_carr = at(st.pd.off_carr);
            c->bprep = true;
            FitSet f = st.f;
            radial_of(f, pd);
            f.first_tile = c->bntiles_total;
            c->bntiles_total += st.ntiles;
            c->bsets.push_back(f);
            c->bntiles.push_back(st.ntiles);
            c->bpend.push_back(pd);
            continue;
        }
        pd.off_carr = at(st.pd.off_carr);
        pd.off_l1 = at(st.pd.off_l1);
        pd.off_l2 = at(st.pd.off_l2);
        pd.off_l3 = at(st.pd.off_l3);
        pd.off_pairs = at(st.pd.off_pairs);
        pd.off_pairs64 = at(st.pd.off_pairs64);
        pd.off_lut = at(st.pd.off_lut);
        pd.off_singles = at(st.pd.off_singles);
        pd.off_iso = at(st.pd.off_iso);
        FitSet f = st.f;
        radial_of(f, pd);
        f.first_tile = c->bntiles_total;
        c->bntiles_total += st.ntiles;
        c->bsets.push_back(f);
        c->bntiles.push_back(st.ntiles);
        c->bpend.push_back(pd);
    }
    return CSX_OK;
}

int csx_stage_commit(csx_ctx* c, const csx_staged* st) { return csx_stage_commit_many(c, 1, &st, 1); }

extern "C" {
int csx_fit_batch_begin(csx_ctx* c) {
    if (!c) return CSX_E_ARG;
    if (c->bres.size) {  // a launched batch nobody waited for: let its copies finish, drop it
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->bres.size = 0;
    }
    c->bsets.clear();
    c->bntiles.clear();
    c->bntiles_total = 0;
    c->barena.clear();
    c->bpend.clear();
    c->btmpl.clear();
    c->bdev_size = 0;
    c->bprep = false;
    c->brad_size = 0;
    c->brblock0.clear();
    return CSX_OK;
}


int csx_fit_batch_stage(csx_ctx* c, const csx_scalars* s, const double* h_carr, const int32_t* h_l1,
                        const int32_t* h_l2, const float* h_l3, int w, int h, const csx_mask* mask,
                        const double* d_data) {
    if (!c) return CSX_E_ARG;
    csx_staged st;
    csx_stage_build(&st, s, h_carr, h_l1, h_l2, h_l3, w, h, mask, d_data);
    return csx_stage_commit(c, &st);
}

int csx_fit_batch_run(csx_ctx* c, int precision, double* h_sums, int64_t* h_cnts) {
    if (!c || !h_sums) return CSX_E_ARG;
    if (int rc = csx_fit_batch_launch(c, precision)) return rc;
    return csx_fit_batch_wait(c, h_sums, h_cnts);
}

}  // extern "C"

// The staged batch's set records and per-set tile counts appended to the pinned arena, with
// the device addresses of every set's tables filled in; one H2D copy on the context's stream
// (or a node of the round's graph).  Returns the device addresses of the records, of the flat
// tile list (written on the device by k_prep_sets) and of the tile counts.
static int batch_upload(csx_ctx* c, const FitSet** d_sets, const int2** d_tiles, const int** d_ntiles,
                        size_t* bytes = nullptr) {
    const int nsets = (int)c->bsets.size();
    const int ntiles = c->bntiles_total;
    auto al = [](size_t o) { return (o + 255) & ~size_t(255); };
    if (c->brblock0.size() != (size_t)nsets + 1) return fail(c, CSX_E_STATE, "batch radial plan out of step");
    const size_t o_sets = al(c->barena.size);
    const size_t o_ntiles = al(o_sets + nsets * sizeof(FitSet));
    const size_t o_rblock = al(o_ntiles + nsets * sizeof(int));
    const size_t total = o_rblock + (nsets + 1) * sizeof(int);
    HIPCHK(c, c->barena_d.ensure(total));
    if (c->brad_size) HIPCHK(c, c->brad.ensure(c->brad_size));
    HIPCHK(c, c->btiles_d.ensure(std::max<size_t>(1, (size_t)ntiles) * sizeof(int2)));
    HIPCHK(c, c->barena.reserve(total));
    if (c->bprep) HIPCHK(c, c->bdev.ensure(std::max<size_t>(c->bdev_size, 256)));
    char* base = (char*)c->barena_d.p;
    char* dbase = (char*)c->bdev.p;
    for (int i = 0; i < nsets; ++i) {
        FitSet& f = c->bsets[i];
        const auto& pd = c->bpend[i];
        f.d.carr = (const double*)(base + pd.off_carr);
        f.rtab = pd.dev_rad == SIZE_MAX ? nullptr : (const Radial*)((char*)c->brad.p + pd.dev_rad);
        if (pd.dev_l3 != SIZE_MAX) {
            // tables the device builds (k_prep_sets), lists of the lattice's template
            PrepIn& pin = f.pin;
            pin.l2 = (const int32_t*)(base + pd.off_l2);
            pin.pairs_t = (const int2*)(base + pd.off_pairs_t);
            pin.iso_t = (const int2*)(base + pd.off_iso_t);
            pin.l3 = (float*)(dbase + pd.dev_l3);
            const bool precise = precise_kind(kind_of(f.d.s.combo_key));
            pin.pr = precise ? nullptr : (PairRec*)(dbase + pd.dev_pairs);
            pin.pr64 = precise ? (PairRec64*)(dbase + pd.dev_pairs) : nullptr;
            pin.lut = (int32_t*)(dbase + pd.dev_lut);
            pin.iso = (float2*)(dbase + pd.dev_iso);
            f.d.latpar1 = (const int32_t*)(base + pd.off_l1);
            f.d.latpar2 = pin.l2;
            f.d.latpar3 = pin.l3;
            f.d.pairs = pin.pr;
            f.d.pairs64 = pin.pr64;
            f.d.lut = pin.lut;
            f.d.singles = (const int32_t*)(base + pd.off_singles);
            f.d.iso = pin.iso;
        } else if ((f.d.s.combo_key >> 8 & 0xff) != 12) {
            f.d.latpar1 = (const int32_t*)(base + pd.off_l1);
            f.d.latpar2 = (const int32_t*)(base + pd.off_l2);
            f.d.latpar3 = (const float*)(base + pd.off_l3);
            f.d.pairs = pd.off_pairs == SIZE_MAX ? nullptr : (const PairRec*)(base + pd.off_pairs);
            f.d.pairs64 = pd.off_pairs64 == SIZE_MAX ? nullptr : (const PairRec64*)(base + pd.off_pairs64);
            f.d.lut = (const int32_t*)(base + pd.off_lut);
            f.d.singles = (const int32_t*)(base + pd.off_singles);
            f.d.iso = (const float2*)(base + pd.off_iso);
        }
    }
    std::memcpy(c->barena.p + o_sets, c->bsets.data(), nsets * sizeof(FitSet));
    std::memcpy(c->barena.p + o_ntiles, c->bntiles.data(), nsets * sizeof(int));
    std::memcpy(c->barena.p + o_rblock, c->brblock0.data(), (nsets + 1) * sizeof(int));
    c->barena.size = total;
    c->brblock0_d = (const int*)(base + o_rblock);
    if (bytes) *bytes = total;  // the caller copies (a node of the round's graph)
    else HIPCHK(c, hipMemcpyAsync(c->barena_d.p, c->barena.p, total, hipMemcpyHostToDevice, c->stream));
    *d_sets = (const FitSet*)(base + o_sets);
    *d_tiles = (const int2*)c->btiles_d.p;  // filled by k_prep_sets
    *d_ntiles = (const int*)(base + o_ntiles);
    c->bsets_d = *d_sets;
    return CSX_OK;
}

// One round of the batched objective as a HIP graph: the arena copy to HBM, the objective
// kernel over the flat tile list, the per-set reduction and the copy of the sums to pinned
// memory.  The graph is built once per context and kernel; every later round updates its
// four nodes in place (sizes, pointers, grid) and launches it with one call.
static int batch_graph_launch(csx_ctx* c, const void* fit_fn, const void* rad_fn, size_t bytes, const FitSet* d_sets,
                              const int2* d_tiles, const int* d_ntiles, int nsets, int ntiles) {
    BatchGraph& G = c->bgraph;
    TileSum* part = (TileSum*)c->bpart_d.p;
    SetSum* sums = (SetSum*)c->bsums_d.p;
    FitSet* d_sets_w = const_cast<FitSet*>(d_sets);
    int2* d_tiles_w = const_cast<int2*>(d_tiles);
    void* prep_args[] = {(void*)&d_sets_w, (void*)&nsets, (void*)&d_ntiles, (void*)&d_tiles_w};
    const int* d_rblock0 = c->brblock0_d;
    void* rad_args[] = {(void*)&d_sets, (void*)&nsets, (void*)&d_rblock0};
    // the sets' radial tables (a grid of at least one block: the kernel returns past the plan)
    hipKernelNodeParams kq{};
    kq.func = const_cast<void*>(rad_fn);
    kq.gridDim = dim3(std::max(1, c->brblock0.back()));
    kq.blockDim = dim3(256);
    kq.kernelParams = rad_args;
    void* fit_args[] = {(void*)&d_sets, (void*)&d_tiles, (void*)&ntiles, (void*)&part};
    void* red_args[] = {(void*)&d_sets, (void*)&nsets, (void*)&part, (void*)&d_ntiles, (void*)&sums};
    // per set: its tile-list entries and, for sets staged with device preparation, its tables
    hipKernelNodeParams kp{};
    kp.func = (void*)k_prep_sets;
    kp.gridDim = dim3(nsets);
    kp.blockDim = dim3(kPrepBlock);
    kp.kernelParams = prep_args;
    hipKernelNodeParams kf{};
    kf.func = const_cast<void*>(fit_fn);
    kf.gridDim = dim3(ntiles);
    kf.blockDim = dim3(kBlock);
    kf.kernelParams = fit_args;
    hipKernelNodeParams kr{};
    kr.func = (void*)k_reduce_sets;
    kr.gridDim = dim3(nsets);
    kr.blockDim = dim3(kBlock);
    kr.kernelParams = red_args;
    const size_t sbytes = (size_t)nsets * sizeof(SetSum);
    if (!G.exec || G.fit_fn != fit_fn) {
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
        if (G.graph) (void)hipGraphDestroy(G.graph);
        G.exec = nullptr;
        G.graph = nullptr;
        HIPCHK(c, hipGraphCreate(&G.graph, 0));
        HIPCHK(c, hipGraphAddMemcpyNode1D(&G.h2d, G.graph, nullptr, 0, c->barena_d.p, c->barena.p, bytes,
                                          hipMemcpyHostToDevice));
        HIPCHK(c, hipGraphAddKernelNode(&G.prep, G.graph, &G.h2d, 1, &kp));
        HIPCHK(c, hipGraphAddKernelNode(&G.rad, G.graph, &G.prep, 1, &kq));
        HIPCHK(c, hipGraphAddKernelNode(&G.fit, G.graph, &G.rad, 1, &kf));
        HIPCHK(c, hipGraphAddKernelNode(&G.red, G.graph, &G.fit, 1, &kr));
        HIPCHK(c, hipGraphAddMemcpyNode1D(&G.d2h, G.graph, &G.red, 1, c->bres.p, sums, sbytes, hipMemcpyDeviceToHost));
        HIPCHK(c, hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
        G.fit_fn = fit_fn;
    } else {
        HIPCHK(c, hipGraphExecMemcpyNodeSetParams1D(G.exec, G.h2d, c->barena_d.p, c->barena.p, bytes,
                                                    hipMemcpyHostToDevice));
        HIPCHK(c, hipGraphExecKernelNodeSetParams(G.exec, G.prep, &kp));
        HIPCHK(c, hipGraphExecKernelNodeSetParams(G.exec, G.rad, &kq));
        HIPCHK(c, hipGraphExecKernelNodeSetParams(G.exec, G.fit, &kf));
        HIPCHK(c, hipGraphExecKernelNodeSetParams(G.exec, G.red, &kr));
        HIPCHK(c, hipGraphExecMemcpyNodeSetParams1D(G.exec, G.d2h, c->bres.p, sums, sbytes, hipMemcpyDeviceToHost));
    }
    HIPCHK(c, hipGraphLaunch(G.exec, c->stream));
    ++G.launches;
    return CSX_OK;
}

int csx_fit_batch_launch(csx_ctx* c, int precision) {
    if (!c) return CSX_E_ARG;
    const int nsets = (int)c->bsets.size();
    if (nsets == 0) return fail(c, CSX_E_STATE, "empty batch");
    precision = effective_precision(c->bsets[0].d.s.combo_key, precision);
    HIPCHK(c, hipSetDevice(c->device));
    const int ntiles = c->bntiles_total;
    const FitSet* d_sets;
    const int2* d_tiles;
    const int* d_ntiles;
    size_t bytes = 0;
    if (int rc = batch_upload(c, &d_sets, &d_tiles, &d_ntiles, c->use_graph ? &bytes : nullptr)) return rc;
    HIPCHK(c, c->bpart_d.ensure(ntiles * sizeof(TileSum)));
    HIPCHK(c, c->bsums_d.ensure(nsets * sizeof(SetSum)));
    HIPCHK(c, c->bres.reserve(nsets * sizeof(SetSum)));
    const int kind = kind_of(c->bsets[0].d.s.combo_key);
    const int variant = c->tune_force >= 0 ? c->tune_force : tune_variant(kind, tune_size_class(ntiles, nsets));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (c->use_graph) {
        const void* fn = nullptr;
        DISPATCH_KIND_FN(kind, precision, fit_batch_kernel, fn, variant);
        const void* rfn = nullptr;
        DISPATCH_KIND_FN(kind, precision, batch_radial_kernel, rfn);
        if (int rc = batch_graph_launch(c, fn, rfn, bytes, d_sets, d_tiles, d_ntiles, nsets, ntiles)) return rc;
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    } else {
        hipLaunchKernelGGL(k_prep_sets, dim3(nsets), dim3(kPrepBlock), 0, c->stream, const_cast<FitSet*>(d_sets), nsets,
                           d_ntiles, const_cast<int2*>(d_tiles));
        HIPCHK(c, hipGetLastError());
        DISPATCH_KIND(kind, precision, launch_batch_radial, d_sets, nsets, c->brblock0_d, c->brblock0.back(), c->stream);
        HIPCHK(c, hipGetLastError());
        DISPATCH_KIND(kind, precision, launch_fit_batch, d_sets, d_tiles, ntiles, (TileSum*)c->bpart_d.p, variant,
                      c->stream);
        HIPCHK(c, hipGetLastError());
        SetSum* d_sums = (SetSum*)c->bsums_d.p;
        hipLaunchKernelGGL(k_reduce_sets, dim3(nsets), dim3(kBlock), 0, c->stream, d_sets, nsets,
                           (const TileSum*)c->bpart_d.p, d_ntiles, d_sums);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->bres.p, d_sums, nsets * sizeof(SetSum), hipMemcpyDeviceToHost, c->stream));
    }
    c->bres.size = (size_t)nsets;
    return CSX_OK;
}

// Wait for the batch enqueued by csx_fit_batch_launch and return its sums and counts.
int csx_fit_batch_wait(csx_ctx* c, double* h_sums, int64_t* h_cnts) {
    if (!c || !h_sums) return CSX_E_ARG;
    const int nsets = (int)c->bres.size;
    if (nsets == 0) return fail(c, CSX_E_STATE, "no batch in flight");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const SetSum* r = (const SetSum*)c->bres.p;
    for (int i = 0; i < nsets; ++i) h_sums[i] = r[i].sum;
    if (h_cnts) for (int i = 0; i < nsets; ++i) h_cnts[i] = r[i].cnt;
    c->bres.size = 0;
    bool mismatch = false;
    for (int i = 0; i < nsets; ++i)
        if (c->bsets[i].pin.on) {
            ++c->prep_sets;
            c->prep_amb += r[i].flags >> 8;
            mismatch = mismatch || (r[i].flags & 1);
        }
    if (mismatch) return fail(c, CSX_E_STATE, "device preparation: a Friedel pair's members differ");
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->timing_pending = false;
    return CSX_OK;
}

extern "C" {

int csx_set_device_prep(csx_ctx* c, int on) {
    if (!c || (on != 0 && on != 1)) return CSX_E_ARG;
    c->device_prep = on != 0;
    return CSX_OK;
}

int csx_device_prep_counts(const csx_ctx* c, long long* sets, long long* ambiguous) {
    if (!c) return CSX_E_ARG;
    if (sets) *sets = c->prep_sets;
    if (ambiguous) *ambiguous = c->prep_amb;
    return CSX_OK;
}

// A table of set `set` of the last uploaded batch, from device memory (after its run).
int csx_fit_batch_table(csx_ctx* c, int set, int which, void* out, size_t bytes, size_t* needed) {
    if (!c || set < 0 || set >= (int)c->bsets.size() || !c->bsets_d || which < 0 || which > 5) return CSX_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    FitSet f;
    HIPCHK(c, hipMemcpy(&f, c->bsets_d + set, sizeof f, hipMemcpyDeviceToHost));
    const DevSet& d = f.d;
    if ((d.s.combo_key >> 8 & 0xff) == 12) return fail(c, CSX_E_ARG, "not a lattice set");
    const void* src = nullptr;
    size_t n = 0;
    double scal[5];
    switch (which) {
    case 0: src = d.latpar3; n = (size_t)(d.s.peakmax2 + 1) * 17 * sizeof(float); break;
    case 1: src = d.pairs; n = d.pairs ? (size_t)d.npairs * sizeof(PairRec) : 0; break;
    case 2: src = d.pairs64; n = d.pairs64 ? (size_t)d.npairs * sizeof(PairRec64) : 0; break;
    case 3: src = d.lut; n = (size_t)kSegments * (d.nlut + 1) * sizeof(int32_t); break;
    case 4: src = d.iso; n = (size_t)d.niso * sizeof(float2); break;
    case 5:
        scal[0] = d.npairs; scal[1] = d.lut_inv_dq; scal[2] = d.qh_min; scal[3] = d.qh_max; scal[4] = d.asum2;
        n = sizeof scal;
        break;
    }
    if (needed) *needed = n;
    if (!out) return CSX_OK;
    if (bytes < n) return fail(c, CSX_E_ARG, "output buffer too small");
    if (which == 5) std::memcpy(out, scal, n);
    else if (n) HIPCHK(c, hipMemcpy(out, src, n, hipMemcpyDeviceToHost));
    return CSX_OK;
}

}  // extern "C"

int csx_ctx_device_prep(const csx_ctx* c) { return c && c->device_prep ? 1 : 0; }

extern "C" int csx_render_batch_run(csx_ctx* c, int precision) {
    if (!c) return CSX_E_ARG;
    const int nsets = (int)c->bsets.size();
    if (nsets == 0) return fail(c, CSX_E_STATE, "empty batch");
    precision = effective_precision(c->bsets[0].d.s.combo_key, precision);
    HIPCHK(c, hipSetDevice(c->device));
    for (const FitSet& f : c->bsets) {
        if (f.h != f.d.s.ymax - f.d.s.ymin) return fail(c, CSX_E_ARG, "a batched image must cover its whole grid");
    }
    const int ntiles = c->bntiles_total;
    const FitSet* d_sets;
    const int2* d_tiles;
    const int* d_ntiles;
    if (int rc = batch_upload(c, &d_sets, &d_tiles, &d_ntiles)) return rc;
    const int kind = kind_of(c->bsets[0].d.s.combo_key);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_prep_sets, dim3(nsets), dim3(kPrepBlock), 0, c->stream, const_cast<FitSet*>(d_sets), nsets,
                       d_ntiles, const_cast<int2*>(d_tiles));
    HIPCHK(c, hipGetLastError());
    DISPATCH_KIND(kind, precision, launch_batch_radial, d_sets, nsets, c->brblock0_d, c->brblock0.back(), c->stream);
    HIPCHK(c, hipGetLastError());
    DISPATCH_KIND(kind, precision, launch_image_batch, d_sets, d_tiles, ntiles, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->bprep) {
        std::vector<FitSet> back(nsets);
        HIPCHK(c, hipMemcpy(back.data(), d_sets, nsets * sizeof(FitSet), hipMemcpyDeviceToHost));
        bool mismatch = false;
        for (const FitSet& f : back)
            if (f.pin.on) {
                ++c->prep_sets;
                c->prep_amb += f.pin.flags >> 8;
                mismatch = mismatch || (f.pin.flags & 1);
            }
        if (mismatch) return fail(c, CSX_E_STATE, "device preparation: a Friedel pair's members differ");
    }
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->timing_pending = false;
    return CSX_OK;
}

int csx_ctx_device(const csx_ctx* c) { return c ? c->device : -1; }
void* csx_ctx_stream(const csx_ctx* c) { return c ? (void*)c->stream : nullptr; }
int csx_ctx_fail(csx_ctx* c, int code, const char* msg) { return fail(c, code, msg ? msg : ""); }
void* csx_ctx_scratch(csx_ctx* c, int which, size_t bytes) {
    if (!c || which < 0 || which > 1) return nullptr;
    return c->post[which].ensure(bytes) == hipSuccess ? c->post[which].p : nullptr;
}

// A second context on the same device (its own stream and staging memory).
int csx_create_like(const csx_ctx* c, csx_ctx** out) {
    if (!c || !out) return CSX_E_ARG;
    const int rc = csx_create(c->device, out);
    if (rc == CSX_OK) {  // the same batch settings
        (*out)->tune_force = c->tune_force;
        (*out)->use_graph = c->use_graph;
        (*out)->device_prep = c->device_prep;
    }
    return rc;
}
