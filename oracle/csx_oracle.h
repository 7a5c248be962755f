/*
 * CPU oracle for the CrystalScatter per-pixel 2D scattering path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (crystalscatter_amd/, libcsx.so) never links, loads or calls it.
 *
 * It restates, in FP64 C (x87 long double where the reference uses it), the
 * reference engine's preparation and per-pixel evaluation for the
 * specialised kernels the hot path uses:
 *   0x00000C07 sphere, no lattice, isotropic   Source/sc_gpu_pSph_lNon_oIso.h:15
 *   0x00020500 sphere, Gaussian peaks, FCC      Source/sc_gpu_pSph_pGau_lFCC_oGau.h:15
 *   0x01000C00 cylinder, no lattice, Gaussian   Source/sc_gpu_pCyl_lNon_oGau.h:15
 *   0x01000C06 cylinder, no lattice, z-dir      Source/sc_gpu_pCyl_lNon_oZDi.h:17
 *   0x01020806 cylinder, Gaussian peaks, BCT    Source/sc_gpu_pCyl_pGau_lBCT_oZDi.h:16
 * Every function cites the reference lines it follows in csx_oracle.c.
 */
#ifndef CSX_ORACLE_H
#define CSX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csxo_engine csxo_engine;

/* Engine with the reference key-table defaults (Source/sc_calc_generic.cpp:36-138). */
csxo_engine* csxo_new(void);
void csxo_free(csxo_engine* e);
/* Set one parameter by its reference key name; returns 0, or -1 for an unknown key. */
int csxo_set(csxo_engine* e, const char* key, double value);
double csxo_get(const csxo_engine* e, const char* key);

/* prepareCalculation (Source/sc_calc_generic_cpu.cpp:319-988); returns 0 on success. */
int csxo_prepare(csxo_engine* e);
/* combo key as in kernel_selection (Source/sc_calc_generic_gpu.cu:113-136). */
unsigned csxo_combo_key(const csxo_engine* e);

/* Image bounds after prepare: xmin,xmax (ihex), ymin,ymax (i). */
void csxo_bounds(const csxo_engine* e, int* b4);
/* q of pixel (ihex, i): convert_idx2q (Source/sc_calc_generic_gpu.cu:249-287). */
void csxo_idx2q(const csxo_engine* e, int ihex, int i, double* q3);
/* One pixel value through the combo dispatch (calcCPU_selection, .cu:168-242).
 * Returns 0 and the value, or -1 if the combination is not restated. */
int csxo_pixel(const csxo_engine* e, double qx, double qy, double qz, double* out);
/* Whole image in IDX order (doCalculation, Source/sc_calc_generic_cpu.cpp:1151-1275). */
int csxo_image(const csxo_engine* e, double* out, int nthreads);
/* Pixels at explicit (ihex,i) indices (used for sampled parity at large sizes). */
int csxo_image_points(const csxo_engine* e, const int* ihex, const int* ii, long n, double* out, int nthreads);

/* Fit residual (doFitCalculation, Source/sc_calc_generic_cpu.cpp:1334-1443 with the
 * per-pixel masks of checkFitRanges, Source/sc_calc_generic_gpu.cu:293-347).
 * rects: 4 x (x0,y0,x1,y1), x0<0 = unused.  fqs_out may be NULL. */
double csxo_fit_fqs(const csxo_engine* e, const double* data, int w, int h, int bstop, int border,
                    const int* rects, double* fqs_out, long* cnt, long* nancnt, int nthreads);

/* Prepared tables by name ("carr1p","carr4p","carr1f","carr4f","carr2p","carr2f","carr11pm",
 * "limq"(18),"latpar1","latpar2","latpar3") and scalars ("norm","order","orcase","ucvol",
 * "peakmax1","peakmax2","dwfactor","phiwidth","sinphic","cosphic","polPhi","sphno",...). */
const void* csxo_table(const csxo_engine* e, const char* name, long* n, int* elem_kind);
double csxo_scalar(const csxo_engine* e, const char* name);

/* Reference Gamma function (Source/sc_calc_generic_gpu.cu:613-673) and gaussnorm3 (.cu:838-854). */
double csxo_gamma(double z);
double csxo_gaussnorm3(double a);
double csxo_lorentznorm3(double a);
double csxo_pearsonnorm3(double a, double b);

/* Nelder-Mead over an objective callback, restating doSimplexFit2D
 * (Source/sc_simplexfit2d.cpp:254-611) with amotry (:221-251) and the bound rule of
 * fehlerquadratsumme (:25-44).  x: in start values (mmax), out best vertex. */
typedef double (*csxo_objective)(void* user, const double* x);
int csxo_simplex(csxo_objective f, void* user, int numfit, int mmax, double* x,
                 const double* minv, const double* maxv, double stp, int maxit, double ftol,
                 int* iters_out, double* fbest_out, long* nevals_out);

/* Post-processing of the AI-data loop.
 * scaleIntensity (Source/sc_calc_generic_cpu.cpp:11195-11236): p holds n values (the engine
 * scans one element past the image); linlog 1 linear, 0 log10, to [0,1] with its running
 * min/else-if-max rule.
 * scaleAndClipData (Source/sc_postproc.cpp:764-830) with its four flags.
 * generateRphi (Source/sc_postproc.cpp:53-360): r,phi resampling of the image with bounds
 * x0..x1, y0..y1 around (xc, yc) into out (s*s values, cells nothing writes are left as they
 * are), then scaleAndClipData(scale, clip, clip40, no log). */
void csxo_scale_intensity(double* p, long n, int linlog);
void csxo_scale_clip(double* data, long len, int scale, int clip, int clip40, int genlog);
void csxo_rphi(const double* d, int x0, int x1, int y0, int y1, int xc, int yc, int s, double* out,
               int scale, int clip, int clip40);
/* calculateIFFT (Source/sc_postproc.cpp:364-560, the build without FFTW: the radix-2 FFT2D/FFT
 * of :833-1000) for an s x s input the size of the output: input scaled by its maximum, 2D
 * FFT (forward 1 / backward 0), output type 0 real, 1 imag, 2 abs, 3 re^2-im^2, the swap to
 * centre 0, scaleAndClipData, and the centred cut-out when 0 < cutx, cuty < s.  out: s*s
 * values (cutx*cuty with the cut-out).  Returns the output length, 0 when the input maximum
 * is 0 or s is not a power of two >= 4. */
long csxo_ifft(const double* d, int s, int forward, int out_type, int scaled, int clip, int clip40, int swap,
               int cutx, int cuty, double* out);

#ifdef __cplusplus
}
#endif
#endif
