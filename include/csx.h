/*
 * csx.h -- C ABI of the MI355X-native CrystalScatter per-pixel scattering path.
 *
 * Plain C types only: host pointers are marked h_, device (HBM) pointers d_.
 * Every function returns 0 on success or a negative CSX_E* code; nothing
 * throws across this boundary.  One csx_ctx per host thread; all device work
 * of a context is ordered on its HIP stream.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   csx_create / csx_destroy   SasCalc_GENERIC_calculation::initMemory, createMemory,
 *                              memcleanup   Source/sc_calc_generic_gpu.cu:374, :495, :589
 *   csx_upload                 the pinned zero-copy tables params.CR->carr*, latpar1/2/3
 *                              Source/sc_calc_generic_gpu.h:412-415, :443-485
 *   csx_render                 kernel_selection(Nx,Ny,dofit=false)   Source/sc_calc_generic_gpu.cu:87-156
 *                              with the image store setXYIntensity   Source/sc_calc_generic_gpu.cu:353-370
 *   csx_fit_fqs                kernel_selection(_fitWidth,_fitHeight,dofit=true) plus the host
 *                              sum of doFitCalculation  Source/sc_calc_generic_cpu.cpp:1352-1434
 *   csx_fit_batch_*            (new) every simplex vertex of many independent fits in one launch;
 *                              replaces the sequential fehlerquadratsumme calls of
 *                              SasCalc_SimplexFit2D   Source/sc_simplexfit2d.cpp:19-194
 *   csx_engine_*               the SasCalc_GENERIC_calculation facade used by sc_calcCons and
 *                              sc_simplexfit2d: key=value setters (Source/sc_calc_generic.cpp:142-238),
 *                              doCalculation (Source/sc_calc_generic_cpu.cpp:1151),
 *                              doFitCalculation (:1334), setFitData (:8220), setNoFitRect (.h:365)
 */
#ifndef CSX_H
#define CSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CSX_OK = 0,
    CSX_E_ARG = -1,        /* bad argument / unsupported combination */
    CSX_E_HIP = -2,        /* HIP runtime error */
    CSX_E_NOMEM = -3,      /* device allocation failed */
    CSX_E_STATE = -4,      /* call out of order (e.g. render before upload) */
    CSX_E_PARAM = -5       /* parameter check failed (checkAllParameters) */
};

/* Arithmetic of the per-pixel evaluation. */
enum {
    CSX_PREC_REF64 = 0,    /* FP64 throughout, the reference formulas and order */
    CSX_PREC_MIXED = 1     /* FP32 peak/quadrature loops with FP64 series, phases and sums */
};

/* Combination keys of kernel_selection (Source/sc_calc_generic_gpu.cu:113-136):
 * (particle<<24) | (ltype==12 ? 0 : shp<<16) | (ltype<<8) | ordis. */
#define CSX_KEY_SPHERE_NONE_ISO   0x00000C07u
#define CSX_KEY_SPHERE_FCC_GAUSS  0x00020500u
#define CSX_KEY_CYL_NONE_GAUSS    0x01000C00u
#define CSX_KEY_CYL_NONE_ZDIR     0x01000C06u
#define CSX_KEY_CYL_BCT_ZDIR      0x01020806u
#define CSX_KEY_CUBE_NONE_ISO     0x04000C07u
#define CSX_KEY_CYL_NONE_ISO      0x01000C07u  /* generic kernel; L/R >= 2 */
#define CSX_KEY_CUBE_NONE_ZDIR    0x04000C06u  /* generic kernel; homogeneous */
/* Spheres with Gaussian peaks on other cubic lattices and with isotropic orientation (the
 * reference's generic kernel, Source/sc_gpu_generic.h; same sums as the FCC kernel). */
#define CSX_KEY_SPHERE_FCC_ISO    0x00020507u
#define CSX_KEY_SPHERE_BCC_GAUSS  0x00020400u
#define CSX_KEY_SPHERE_BCC_ISO    0x00020407u
#define CSX_KEY_SPHERE_SC_GAUSS   0x00020700u
#define CSX_KEY_SPHERE_SC_ISO     0x00020707u
#define CSX_KEY_SPHERE_BCT_GAUSS  0x00020800u
#define CSX_KEY_SPHERE_BCT_ISO    0x00020807u
#define CSX_KEY_SPHERE_NONE_GAUSS 0x00000C00u  /* no lattice: the isotropic sphere's value */
#define CSX_KEY_SPHERE_NONE_ZDIR  0x00000C06u
/* The other lattice keys served (the generic kernel), without names: on the 2D lattices
 * (LType 1-3), the cubic ones (4, 5, 7, 9, 10, 11, 17), HCP (6), BCT (8) and orthorhombic (18)
 *   spheres, peak shapes 1-6, any orientation distribution: (shp<<16) | (LType<<8) | ordis;
 *   perfectly oriented cylinders, peak shapes 1-6:          0x01000006 | (shp<<16) | (LType<<8);
 *   isotropic cylinders (L/R >= 2), peak shapes 1-6:        0x01000007 | (shp<<16) | (LType<<8);
 *   isotropic homogeneous cubes, peak shapes 1-6:           0x04000007 | (shp<<16) | (LType<<8).
 * Peak shapes other than the Gaussian (2) run the FP64 sums in either precision.
 * Any other key is refused by csx_upload / preparation with CSX_E_ARG. */

#define CSX_NCOEF 151   /* coeffarray_len */
#define CSX_NIM2 131    /* imax2d_len */
#define CSX_LATLEN 2500 /* latparlen */

/* Per-parameter-set scalars the device path reads (the by-value engine copy of
 * the reference, reduced to what the six kernels use).  Filled by the host
 * preparation (csx_engine_* or the caller's own prepareCalculation). */
typedef struct csx_scalars {
    uint32_t combo_key;
    int32_t zmax;                       /* GridPoints */
    int32_t centerbeam;                 /* q from detector geometry (useBeamStop) */
    int32_t orcase;                     /* 0 general, 1 x, 2 y, 3 z (reference enum) */
    int32_t fix_orcase;                 /* 1: compare orientation cases 1-based (physics mode) */
    int32_t peakmax1, peakmax2;
    int32_t dsch, para, twin;           /* Debye-Scherrer width, paracrystal width, twinning */
    int32_t ordis;
    double qmax, pixnox, pixnoy, pixx_m, pixy_m, det, wave, beamx, beamy;
    double radius, sigma, length, sigmal;
    double dbeta;                       /* degrees, as the key table */
    double polTheta, polPhi;            /* degrees, after the prep's phi quirk */
    double cosphi, sinphi, costheta, sintheta, cosphic, sinphic;
    double limq[18];                    /* limq1..9, limq1f..9f */
    double norm;
    double base, i0, iso, ifluc, rfluc;
    double dwfactor, uca, ucb, domainsize, width_zuf, reldis, dist, phiwidth, ceff;
    double por;                         /* Porod constant of the isotropic cube */
    double sphno_cubevol;               /* latpar[1]*latpar[2] */
    int32_t xmin, xmax, ymin, ymax;     /* image bounds _xmin.._ymax (ihex, i) */
    int32_t iimin;                      /* offset of fit coordinates (checkFitRanges) */
    int32_t cs;                         /* particle interior (0 homogeneous) */
    int32_t use1d;                      /* 1D q mode: qx = (qmax-qmin)/zmax*ihex, one row */
    double qmin;                        /* EditQmin (1D mode) */
    int32_t axmin, axmax, aymin, aymax; /* bounds of the image array (data()): the computed
                                           block, or the whole image when ExpandImage fills it
                                           from a computed quadrant or half */
    /* peak shapes other than the Gaussian (setPeakPar, prepareCalculation cpu:585-590):
     * eta = EditPeakPar/100 (Pseudo-Voigt), beta = EditPeakPar (Pearson VII), the width
     * constants c0..c4, and Gamma(beta), Gamma(beta - 1/2) of the reference's Gamma */
    double peak_eta, peak_beta;
    double peak_c[5];
    double peak_gamma_b, peak_gamma_bh;
} csx_scalars;

/* Block of pixels to evaluate: ihex in [xmin,xmax), i in [ymin,ymax); output in the
 * reference IDX order (ihex - xmin) + (xmax - xmin) * (i - ymin). */
typedef struct csx_grid {
    int32_t xmin, xmax, ymin, ymax;
} csx_grid;

/* Fit masks of checkFitRanges (Source/sc_calc_generic_gpu.cu:293-347): up to four
 * inclusive no-fit rectangles (x0<0 = unused), border and beam-stop pixels, with the
 * reference's coordinate conventions. */
typedef struct csx_mask {
    int32_t rects[16];                  /* x0,y0,x1,y1 per rectangle */
    int32_t bstop, border;
} csx_mask;

typedef struct csx_ctx csx_ctx;

int csx_create(int device, csx_ctx** out);
void csx_destroy(csx_ctx* ctx);
/* Use an existing HIP stream (hipStream_t) for all work of this context; NULL = own stream. */
int csx_set_stream(csx_ctx* ctx, void* hip_stream);
int csx_synchronize(csx_ctx* ctx);

/* Copy one parameter set's tables into HBM (once per parameter set).
 * carr: 18 x CSX_NCOEF doubles in the order carr1p..carr9p, carr1f..carr9f;
 * carr11pm: CSX_NIM2 x CSX_NIM2 doubles (may be NULL when the key does not use it);
 * latpar1/latpar2: (peakmax+1) x 6 ints; latpar3: (peakmax2+1) x 17 floats (row 0 unused). */
int csx_upload(csx_ctx* ctx, const csx_scalars* s, const double* h_carr, const double* h_carr11pm,
               const int32_t* h_latpar1, const int32_t* h_latpar2, const float* h_latpar3);

/* Evaluate the pixels of `grid` into d_out (device, (xmax-xmin)*(ymax-ymin) doubles).
 * Asynchronous on the context's stream. */
int csx_render(csx_ctx* ctx, const csx_grid* grid, int precision, double* d_out);

/* csx_render plus work counters (synchronous; for measurement, no reference counterpart):
 * h_stats[0] pair evaluations (lanes x pairs iterated, mixed path), [1] pixels whose peak sum
 * the conditioning guard re-took in FP64 over the window, [2] pixels with FP64 terms for
 * |q^.u| near 1, [3] waves that evaluated both Friedel terms, [4] pixels re-summed in FP64
 * over the full reflection list. */
#define CSX_NSTATS 5
int csx_render_stats(csx_ctx* ctx, const csx_grid* grid, int precision, double* d_out, uint64_t* h_stats);

/* q vectors of the pixels of `grid` as the kernels generate them (convert_idx2q,
 * Source/sc_calc_generic_gpu.cu:249-287): d_q3 receives 3 doubles per pixel, IDX order. */
int csx_qgrid(csx_ctx* ctx, const csx_grid* grid, double* d_q3);

/* Fit objective of one parameter set: sum over fit pixels of
 * (log10 I_calc - log10 I_data)^2 for I_calc>0 and I_data>0 (FQSVERGL, Source/sc_globalConfig.h:32),
 * NaN terms counted and skipped, nonzero terms counted.  d_data: w*h doubles (device).
 * d_fqs (optional, device): per-pixel terms. */
int csx_fit_fqs(csx_ctx* ctx, int w, int h, const csx_mask* mask, int precision, const double* d_data,
                double* h_sum, int64_t* h_cnt, int64_t* h_nancnt, double* d_fqs);

/* csx_fit_fqs without allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving or fitting loops, harnesses
 * that forbid hipMalloc).  Same arguments and results as csx_fit_fqs (bit for bit), but the
 * per-tile partial sums live in a static device array of the library (one per GPU, 256 KiB,
 * held until the process ends) and are summed on the device; nothing is allocated on the
 * device or the host.  Limit: (w/16 rounded up) x (h/16 rounded up) <= 16384 tiles, i.e. at
 * most 2048 x 2048 fit pixels; larger arrays return CSX_E_ARG.  Every call on one GPU uses
 * the same array, so calls that overlap on one GPU (other threads, other contexts) race;
 * the caller serialises them.  Synchronous, like csx_fit_fqs. */
int csx_fit_fqs_noalloc(csx_ctx* ctx, int w, int h, const csx_mask* mask, int precision, const double* d_data,
                        double* h_sum, int64_t* h_cnt, int64_t* h_nancnt, double* d_fqs);

/* Batched objective: many parameter sets, each with its own tables, detector size, data and
 * mask, evaluated in one launch (one workgroup per 16x16 tile of the flat tile list of all
 * staged sets).  Sets are staged with csx_fit_batch_stage after
 * csx_fit_batch_begin; csx_fit_batch_run evaluates all staged sets and writes h_sums[nsets]. */
int csx_fit_batch_begin(csx_ctx* ctx);
int csx_fit_batch_stage(csx_ctx* ctx, const csx_scalars* s, const double* h_carr,
                        const int32_t* h_latpar1, const int32_t* h_latpar2, const float* h_latpar3,
                        int w, int h, const csx_mask* mask, const double* d_data);
int csx_fit_batch_run(csx_ctx* ctx, int precision, double* h_sums, int64_t* h_cnts);
/* Batched images (no reference counterpart; replaces one doCalculation per parameter set in
 * the AI-data loop, Source/sc_mainCons.cpp:916-1440): renders every set staged since
 * csx_fit_batch_begin in one launch.  Each set is staged with w, h = its whole grid
 * (xmax-xmin, ymax-ymin; Q4 or Q1 layout), no mask, and d_data = its device output of w*h
 * doubles, written in IDX order.  All sets share one combination key.  Synchronous. */
int csx_render_batch_run(csx_ctx* ctx, int precision);

/* Kernel variants of the batched objective (no reference counterpart).  For the lattice
 * kinds in mixed precision, csx_fit_batch_run has CSX_TUNE_VARIANTS kernels, the same code
 * with the register budget cut for 4 waves per SIMD (variant 0), uncapped (1), cut for 5 (2)
 * or 6 waves (3).  They compute the same sums; a table picks the fastest per
 * combination and detector size class (mean tiles per set: <= 64, <= 256, <= 576, more), filled
 * from tuning.txt beside the library at the first csx_create (lines "key class variant", as
 * crystalscatter_amd/autotune.py writes them).  size_class -1 sets every class.
 * csx_tune_force makes one context use a variant for every batch (-1: the table), for
 * measurement.  csx_tune_load returns the number of entries read. */
#define CSX_TUNE_VARIANTS 4
#define CSX_TUNE_SIZES 4
int csx_tune_set(uint32_t combo_key, int size_class, int variant);
int csx_tune_get(uint32_t combo_key, int size_class);
int csx_tune_load(const char* path);
int csx_tune_force(csx_ctx* ctx, int variant);
/* Batched rounds as HIP graphs (default on): csx_fit_batch_run and the batched fits launch
 * each round (arena copy, objective kernel, per-set reduction, copy of the sums) as one graph,
 * built once per context and kernel and updated in place between rounds.  Off: the same four
 * operations launched one by one (same results).  csx_batch_graph_launches counts the graph
 * launches of a context. */
int csx_set_batch_graph(csx_ctx* ctx, int enable);
long long csx_batch_graph_launches(const csx_ctx* ctx);

/* Device preparation of batched lattice sets (no reference counterpart; replaces the table
 * fill of prepareCalculation, Source/sc_calc_generic_cpu.cpp:838-954, per simplex vertex).
 * With it on (the default) the batched fit driver (csx_nm_batch, csx_nm_eval) prepares the
 * lattice sets' scalars, coefficient tables and rotated reciprocal bases on the host, and
 * one workgroup per set builds the reflection table (latpar3), the Friedel-pair records,
 * their |q| bin table and the isotropic list on the device, as the first step of the round
 * (a node of its graph).  The tables equal the host's bit for bit; the one value that could
 * differ, the peak norm's exponential, is correctly rounded on the device and every entry
 * where the C library's exp could round to a different float is counted
 * (csx_device_prep_counts: sets built, entries counted).  Off: the host builds every table.
 * csx_fit_batch_table copies a table of set `set` of the last batch run back to the host
 * (which: 0 latpar3, 1 pair records, 2 FP64 pair records, 3 bin table, 4 isotropic list,
 * 5 five doubles: pair-record count, bin scale, |q| min and max, weight bound), for either
 * kind of staging; out = NULL returns the size in *needed. */
int csx_set_device_prep(csx_ctx* ctx, int on);
int csx_device_prep_counts(const csx_ctx* ctx, long long* sets, long long* ambiguous);
int csx_fit_batch_table(csx_ctx* ctx, int set, int which, void* out, size_t bytes, size_t* needed);

/* ---- Batched Nelder-Mead fits ------------------------------------------------------
 * The simplex of SasCalc_SimplexFit2D::doSimplexFit2D (Source/sc_simplexfit2d.cpp:254-611,
 * amotry :221-251) with the bounds rule of fehlerquadratsumme (:25-44: a point outside
 * [lo, hi] in a fitted entry scores 1e20 without evaluation, except entries with bounds
 * exactly 0..360), run as a resumable state machine.  csx_nm_batch advances many independent
 * fits in lockstep: every point they need in a round is prepared on host threads and
 * evaluated with the batched objective in one launch.  It replaces the sequential
 * fehlerquadratsumme calls.
 *
 * A parameter vector has npar entries (the reference's mmax: every entry of the fit's
 * parameter hash); the first nfree are fitted (numFit), the others are the hash's fixed
 * entries.  The reference orders the hash by name, fitted names first (:280-313); callers that
 * want its path pass the entries in that order.
 * CSX_NM_REFERENCE follows doSimplexFit2D exactly, including what it does beyond the textbook
 * simplex:
 *   - the clearing loop of :384-389 zeroes entry 0 of vertex 0 after the vertex was set, while
 *     y[0] is evaluated at the start values; the other vertices are built from that vertex;
 *   - the steps (:343-346) and the trial factor (:231) use npar, not nfree;
 *   - trial points carry the fixed entries at their simplex column sums (:226), and the
 *     objective sets them (:45-52);
 *   - vertex nfree+1 is evaluated when npar > nfree, its value copied to the later ones (:420-428).
 * CSX_NM_CORRECTED is the textbook simplex over the nfree fitted entries with the reference's
 * step rule: vertex 0 kept, steps and trial factor over nfree, fixed entries at their start
 * values, nfree+1 initial evaluations. */
#define CSX_NM_MAXPAR 30   /* map (Source/sc_simplexfit2d.h:114) */
enum { CSX_NM_REFERENCE = 0, CSX_NM_CORRECTED = 1 };
typedef double (*csx_objective)(void* user, const double* x);
typedef struct csx_nm_result {
    double x[CSX_NM_MAXPAR];  /* best vertex, p[ilo] as the reference returns it (all npar entries) */
    double f;                 /* its objective */
    int32_t iters, evals;     /* simplex iterations (repetitions), objective evaluations (numImgCalc) */
} csx_nm_result;
struct csx_engine;
typedef struct csx_nm_fit {
    const struct csx_engine* base;  /* parameters of this fit (host-only engine is enough) */
    const double* d_data;           /* device: w*h doubles in the image's IDX layout */
    int32_t w, h;
    csx_mask mask;
    double x0[CSX_NM_MAXPAR], lo[CSX_NM_MAXPAR], hi[CSX_NM_MAXPAR];
} csx_nm_fit;
/* One fit on a host objective (the same state machine; test and scripting use).  The
 * objective receives all npar entries. */
int csx_nm_run_cb(int npar, int nfree, int mode, csx_objective f, void* user, const double* x0, const double* lo,
                  const double* hi, double stp, int maxit, double ftol, csx_nm_result* out);
/* nfits fits over the npar parameters named in keys (a key "uca,ucb" sets both names to the
 * parameter's value); threads <= 0: all host cores. */
typedef struct csx_nm_stats {
    int64_t rounds;           /* batched launches (from 32 fits on, two groups of fits alternate) */
    int64_t evals;            /* objective evaluations over all fits */
    double prep_ms;           /* host preparation of the parameter sets and their peak tables (threads) */
    double stage_ms;          /* appending the built sets to the batch (serial) */
    double run_ms;            /* batched launch, reductions and the copy of the sums */
    int64_t pruned;           /* trial points screened out of the exact evaluation (csx_nm_batch_screened) */
    double screen_ms;         /* time in the screen callback */
    double launch_ms;         /* enqueueing the rounds: set records, graph update and launch */
} csx_nm_stats;
int csx_nm_batch(csx_ctx* ctx, int nfits, const csx_nm_fit* fits, int npar, int nfree, int mode,
                 const char* const* keys, double stp, int maxit, double ftol, int precision, int threads,
                 csx_nm_result* out, csx_nm_stats* stats);
/* csx_nm_batch with a screen of the simplex's trial points (no reference counterpart; the
 * learned part of the fits).  Each round, fn scores every point the fits posted (npts points
 * of fits fit[k], parameter vectors x[k*npar ...]; score[k]: lower is better, e.g. a surrogate
 * network's objective) and returns 0.  A trial point (reflection, expansion, contraction)
 * scored above (1 + margin) times the worst score of its simplex's vertices is not evaluated:
 * it takes the value of a rejected point (1e20), so the simplex never accepts it.  Initial and
 * shrink vertices, and every point a simplex accepts, have exact values.  fn NULL: csx_nm_batch. */
typedef int (*csx_nm_screen)(void* user, int npts, const int32_t* fit, const double* x, double* score);
typedef struct csx_nm_screen_opts {
    csx_nm_screen fn;
    void* user;
    double margin;
} csx_nm_screen_opts;
int csx_nm_batch_screened(csx_ctx* ctx, int nfits, const csx_nm_fit* fits, int npar, int nfree, int mode,
                          const char* const* keys, double stp, int maxit, double ftol, int precision, int threads,
                          const csx_nm_screen_opts* screen, csx_nm_result* out, csx_nm_stats* stats);
/* The batched objective at explicit points: out[i*nv + v] = fehlerquadratsumme of fit i at
 * x[(i*nv + v)*npar ...] (1e20 outside the bounds of any entry), all points of all fits in
 * one batched launch.  No reference counterpart: used to score proposed starting points. */
int csx_nm_eval(csx_ctx* ctx, int nfits, const csx_nm_fit* fits, int npar, const char* const* keys, int nv,
                const double* x, int precision, int threads, double* out);

int csx_last_error(csx_ctx* ctx, char* buf, int len);
/* Kernel time of the last csx_fit_fqs / csx_fit_batch_run, and of the last csx_render when
 * timing is on, measured with HIP events on the context's stream (milliseconds); waits for
 * a pending render.  Stands in for higResTimerElapsedCalc (Source/sc_calc_generic_gpu.h:376). */
double csx_last_kernel_ms(csx_ctx* ctx);
/* Bracket every csx_render launch with HIP events (default off: the event pair keeps
 * back-to-back launches from overlapping, about 10 us per 1024^2 render). */
int csx_set_timing(csx_ctx* ctx, int enable);
/* Durations (ms) of the image pass of the last renders made with timing on, oldest first:
 * the per-pixel kernel alone (for the lattice kinds in mixed precision, the hot pass; its
 * radial pre-pass and fix pass are not included), from HIP events on the context's stream
 * recorded around that launch, up to the last CSX_PASS_RING renders; waits for them.
 * Returns the number written (at most max).  No reference counterpart: the per-kernel time
 * that the roofline accounting divides by. */
#define CSX_PASS_RING 64
int csx_pass_times(csx_ctx* ctx, int max, double* ms);

/* ---- SasCalc_GENERIC_calculation facade ------------------------------------------- */
typedef struct csx_engine csx_engine;
/* device < 0 creates a host-only engine: preparation and tables, no GPU work. */
csx_engine* csx_engine_new(int device);
void csx_engine_free(csx_engine* e);
/* Set a parameter by its reference key (Source/sc_calc_generic.cpp:36-138). */
int csx_engine_set(csx_engine* e, const char* key, double value);
double csx_engine_get(csx_engine* e, const char* key);
/* prepareCalculation on the host + csx_upload; returns 0 or CSX_E_PARAM (message in last_error). */
int csx_engine_prepare(csx_engine* e);
/* doCalculation: image into h_out ((xmax-xmin)*(ymax-ymin) doubles, IDX order), bounds in b4. */
int csx_engine_calc(csx_engine* e, int precision, double* h_out, int32_t* b4);
/* Same, into device memory. */
int csx_engine_calc_device(csx_engine* e, int precision, double* d_out, int32_t* b4);
/* setFitData + doFitCalculation: returns the objective in *h_sum (1e20 on failure). */
int csx_engine_set_fit_data(csx_engine* e, int w, int h, const double* h_data);
int csx_engine_set_nofit_rect(csx_engine* e, int id, int x0, int y0, int x1, int y1);
int csx_engine_fit(csx_engine* e, int precision, int bstop, int border, double* h_sum,
                   int64_t* h_cnt, int64_t* h_nancnt);
/* Host-side prepared tables (for checks): name as in the reference ("carr4p", "latpar3", ...). */
const void* csx_engine_table(csx_engine* e, const char* name, int64_t* n, int* kind);
const csx_scalars* csx_engine_scalars(csx_engine* e);
csx_ctx* csx_engine_ctx(csx_engine* e);
int csx_engine_last_error(csx_engine* e, char* buf, int len);
/* Stage the engine's image (checks and host preparation included) into ctx's pending batch
 * for csx_render_batch_run, output d_out (room for cap doubles); the engine can be changed
 * right after. */
int csx_engine_stage_image(csx_engine* e, csx_ctx* ctx, double* d_out, int64_t cap);
/* n images staged at once (no reference counterpart; the training patterns of the surrogate):
 * image k has the engine's parameters with keys[j] set to x[k*npar + j] (a key "a,b" sets
 * both names), is checked, prepared on `threads` host threads (lattice tables on the device
 * unless csx_set_device_prep(ctx, 0)) and staged into ctx's pending batch with its output at
 * d_out + k*stride.  Follow with csx_render_batch_run.  The engine itself is not changed. */
int csx_engine_stage_points(csx_engine* e, csx_ctx* ctx, int n, int npar, const char* const* keys, const double* x,
                            double* d_out, int64_t stride, int threads);

/* ---- Random variation for training data (tpvPerformRandom) ------------------------------
 * csx_rand: the C library's rand() sequence (srand(seed), then rand() values) in a private
 * state; the reference draws from the process-wide one, seeded 1 unless srand is called.
 * csx_rand_zuf fills n values of zuf() = (rand()/RAND_MAX*200 - 100)/100 in [-1, 1]
 * (Source/sc_mainCons.cpp:850-853), e.g. the noise of the AI-data loop.
 * csx_rand_skip advances the state by n values without drawing them (a jump of O(log n));
 * csx_rand_zuf_device writes the next n zuf() values to device memory (d_out, on ctx's
 * stream; equal to csx_rand_zuf's) and advances the state by n (no reference counterpart:
 * the device continuation of the sequence). */
typedef struct csx_rand csx_rand;
csx_rand* csx_rand_new(uint32_t seed);
void csx_rand_free(csx_rand* r);
int32_t csx_rand_next(csx_rand* r);
int csx_rand_zuf(csx_rand* r, double* h_out, int64_t n);
int csx_rand_skip(csx_rand* r, int64_t n);
int csx_rand_zuf_device(csx_ctx* ctx, csx_rand* r, int64_t n, double* d_out);
/* tpvPerformRandom (Source/sc_calc_generic_cpu.cpp:1028-1143): ids holds "key=factor" lines
 * ('\n'-separated); each listed variable gets its value at the first call after a clear times
 * (1 + factor*zuf()) (ucb/ucc: the current uca; phi: 2 + 43*(zuf()+1); EditAzi: draws but has
 * no effect; EditDebyeWaller: the Debye-Waller factor), then the cylinder/disk length checks.
 * ids NULL or empty clears.  The values apply to the next preparation; csx_engine_get keeps
 * returning the parameters as set, csx_engine_tpv_value the varied ones.  report (optional)
 * receives the reference's "key = value" lines. */
int csx_engine_tpv(csx_engine* e, const char* ids, csx_rand* rng, char* report, int cap);
double csx_engine_tpv_value(csx_engine* e, const char* key);
/* n images of one AI-data parameter line staged at once (no reference counterpart; the loop
 * of Source/sc_mainCons.cpp:1236-1300 for one line): per image, in the loop's rand() order,
 * csx_engine_tpv(ids), npx noise values (if d_noise: written by the device to
 * d_noise + k*npx, on ctx's stream), nextra values into h_extra + k*nextra (the caller's beam
 * stop and lines); then preparation and staging of all images on `threads` host threads into
 * ctx's pending batch, image k's output at d_out + k*stride.  Follow with
 * csx_render_batch_run. */
int csx_engine_tpv_batch(csx_engine* e, csx_ctx* ctx, const char* ids, csx_rand* rng, int n, double* d_out,
                         int64_t stride, double* d_noise, int nextra, double* h_extra, int threads);

/* ---- Post-processing of image batches in HBM (the AI-data loop after each calculation,
 * Source/sc_mainCons.cpp:1254-1400) -------------------------------------------------------
 * d_img holds nimg consecutive images of npx doubles; work is ordered on ctx's stream.
 * The scans keep the reference's running rule (a value that is a new minimum is not tested
 * against the maximum), exactly.
 * noise: d += z*sqrt(d) (z/100 when scaled) for n values, d_z one zuf() value per pixel in the
 *   reference's draw order (:1275-1290);
 * zero: d[idx[i]] = 0 (beam stop and lines, :1298-1348, indices from the host);
 * scale_intensity: scaleIntensity (Source/sc_calc_generic_cpu.cpp:11195-11236), linlog 1
 *   linear, 0 log10; the engine's scan covers one zero past each image, so does this one;
 * log_scale: the AI loop's log scaling of an image (:1366-1384; skip_axes: ignore the pixels
 *   on x = 0 or y = 0 of an image whose x range xmin.. (width w) is below 0);
 * scale_clip: scaleAndClipData (Source/sc_postproc.cpp:764-830);
 * rphi: generateRphi (Source/sc_postproc.cpp:53-360) of every image (bounds x0..x1, y0..y1,
 *   centre xc, yc) into d_out (nimg x s*s; cells the reference never writes are 0). */
int csx_post_noise(csx_ctx* ctx, double* d_img, const double* d_z, int64_t n, int scaled);
int csx_post_zero(csx_ctx* ctx, double* d_img, const int64_t* d_idx, int64_t n);
/* zero_rects: the same zeroing given as nrect rectangles per image (h_rects: nimg x nrect x
 *   {xa, xb, ya, yb}, inclusive, host memory) over an image with bounds x0..x1, y0..y1: cell
 *   (x, y) zeroes pixel (x - x0) + (x1 - x0)(y - y0) when that lies in the image, so the beam
 *   stop is its box cut to the image and a line is x = xp..xp+width-1 (or y = ...) over the
 *   full other range (:1298-1348). */
int csx_post_zero_rects(csx_ctx* ctx, double* d_img, int64_t npx, int nimg, const int32_t* h_rects, int nrect,
                        int x0, int x1, int y0, int y1);
int csx_post_scale_intensity(csx_ctx* ctx, double* d_img, int64_t npx, int nimg, int linlog);
int csx_post_log_scale(csx_ctx* ctx, double* d_img, int64_t npx, int nimg, int w, int xmin, int ymin,
                       int skip_axes);
int csx_post_scale_clip(csx_ctx* ctx, double* d_img, int64_t npx, int nimg, int scale, int clip,
                        int clip40, int genlog);
int csx_post_rphi(csx_ctx* ctx, const double* d_img, int x0, int x1, int y0, int y1, int xc, int yc,
                  int s, int nimg, double* d_out, int scale, int clip, int clip40);
/* rphi_many: the same with one centre per image (h_centres: nimg x {xc, yc}, host memory), in
 *   one launch; images with the same rounded maximum radius share one cell map. */
int csx_post_rphi_many(csx_ctx* ctx, const double* d_img, int x0, int x1, int y0, int y1, const int32_t* h_centres,
                       int s, int nimg, double* d_out, int scale, int clip, int clip40);
/* FFT image: calculateIFFT of builds without FFTW (Source/sc_postproc.cpp:364-560, its radix-2
 * FFT2D/FFT :833-1000) for inputs of the output's size (s x s, s a power of two 4..2048, as the
 * r,phi image the AI loop passes): each input scaled by its maximum, 2D transform (forward 1 /
 * backward 0), output type 0 real, 1 imaginary, 2 modulus, 3 re^2-im^2, optional swap (0 at the
 * centre; the row and column it never writes are 0), scaleAndClipData(scaled, clip, clip40),
 * then the centred cutx x cuty cut-out when 0 < cutx, cuty < s.  d_out: nimg outputs.  An
 * input whose maximum is 0 has no FFT in the reference (here its output is not finite). */
int csx_post_ifft(csx_ctx* ctx, const double* d_in, int s, int nimg, int forward, int out_type, int scaled,
                  int clip, int clip40, int swap, int cutx, int cuty, double* d_out);

/* ---- Surrogate network (no counterpart in the reference) --------------------------------
 * The learned part of the fits: an MLP from a normalised parameter set (CSX_MLP_NPAR values)
 * and a pixel's (qx, qy) to the normalised log-intensity of that pixel.  Inputs: the
 * parameters and sin/cos of CSX_MLP_NFREQ phases a_f = 2*pi*(F_f0*qx + F_f1*qy) +
 * F_f2*|q|*v_f, where v_f = offset_j + scale_j * (normalised parameter j) for j = pidx_f, or 1
 * for pidx_f = -1 (CSX_MLP_IN values, zero-padded to CSX_MLP_IN_PAD); CSX_MLP_LAYERS hidden
 * layers of CSX_MLP_HIDDEN SiLU units; one linear output.  BF16 weights and activations on
 * the matrix cores with FP32 accumulation; FP32 master weights and Adam state.  Samples are
 * rows: row i uses parameter set set[i] of d_par (nsets x CSX_MLP_NPAR floats; an index
 * outside [0, nsets) is clamped into it) and d_q[2i], d_q[2i+1].  All pointers named d_* are
 * device memory; calls are ordered on the model's stream. */
#define CSX_MLP_NPAR 6
#define CSX_MLP_NFREQ 32
#define CSX_MLP_IN (CSX_MLP_NPAR + 2 * CSX_MLP_NFREQ)
#define CSX_MLP_IN_PAD 128
#define CSX_MLP_HIDDEN 256
#define CSX_MLP_LAYERS 4
typedef struct csx_mlp csx_mlp;
/* Parameter count of the packed FP32 vector: per hidden layer W [HIDDEN][in] (out x in; the
 * first layer's in = IN_PAD, its padding columns zero) then b [HIDDEN]; then w_out [HIDDEN]
 * and b_out. */
int64_t csx_mlp_nparams(void);
/* max_batch: rows per call (rounded up to a multiple of 128); seed: initialisation. */
int csx_mlp_create(int device, int max_batch, uint64_t seed, csx_mlp** out);
void csx_mlp_destroy(csx_mlp* m);
int csx_mlp_set_stream(csx_mlp* m, void* stream);
int csx_mlp_last_error(const csx_mlp* m, char* buf, int len);
/* Host copies of the FP32 parameters (csx_mlp_nparams floats); set_params also resets the
 * Adam state. */
int csx_mlp_get_params(csx_mlp* m, float* h_params);
int csx_mlp_set_params(csx_mlp* m, const float* h_params);
/* Input phases: F (NFREQ x 3, row-major), pidx (NFREQ), affine (NPAR x {offset, scale}).
 * The default after create: F_f2 = 0, pidx = -1 (plain Fourier features of (qx, qy)). */
int csx_mlp_get_features(const csx_mlp* m, float* h_freq, int32_t* h_pidx, float* h_affine);
int csx_mlp_set_features(csx_mlp* m, const float* h_freq, const int32_t* h_pidx, const float* h_affine);
/* d_y[i] = network(row i), n rows (asynchronous). */
int csx_mlp_forward(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q,
                    float* d_y);
/* The same d_y as csx_mlp_forward, bit for bit, from one fused kernel that keeps each block
 * of rows' activations on chip across the layers (inference only; any n >= 1, no max_batch
 * limit; asynchronous). */
int csx_mlp_predict(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q,
                    float* d_y);
/* Gradient of the loss 1/n sum_i (y_i - d_t[i])^2 into the model's gradient buffer;
 * *h_loss (optional, synchronous) = the loss. */
int csx_mlp_gradient(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q,
                     const float* d_t, float* h_loss);
int csx_mlp_get_grads(csx_mlp* m, float* h_grads);
/* One Adam step with the current gradient (bias-corrected). */
int csx_mlp_adam(csx_mlp* m, float lr, float b1, float b2, float eps);
/* csx_mlp_gradient + csx_mlp_adam(lr, 0.9, 0.999, 1e-8). */
int csx_mlp_train_step(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q,
                       const float* d_t, float lr, float* h_loss);
int csx_mlp_synchronize(csx_mlp* m);

#ifdef __cplusplus
}
#endif
#endif
