// Post-processing of batches of images already in HBM, for training data: the steps of the
// console's AI-data loop after each calculation (Source/sc_mainCons.cpp:1254-1400) -- noise,
// zeroed pixels (beam stop, lines), intensity scaling (scaleIntensity,
// Source/sc_calc_generic_cpu.cpp:11195-11236; scaleAndClipData, Source/sc_postproc.cpp:764-830)
// and the r,phi resampling (generateRphi, Source/sc_postproc.cpp:53-360).
//
// Images are n consecutive arrays of npx doubles.  The reference's scans keep a running
// minimum and test the maximum only when a value is not a new minimum ("if (v < lo) lo = v;
// else if (v > hi) hi = v"), so hi is the maximum over the values that are not strict prefix
// minima.  One workgroup per image reproduces that exactly: per-lane chunk minima, a
// workgroup prefix-min scan, then each lane walks its chunk in order from its prefix minimum.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "csx.h"
#include "csx_prep.h"
#include "csx_rand.h"

namespace csp {

constexpr int kStat = 1024;  // lanes of the per-image statistics workgroup

// Which values a statistics pass reads and its start values.
enum StatMode {
    ST_LIN = 0,      // every value; lo 1e20, hi 1e-20 (scaleIntensity, linear)
    ST_LOG = 1,      // log10 of positive values; lo 1e20, hi 1e-20 (scaleIntensity, log)
    ST_GENLOG = 2,   // log10 of positive values; lo 1e26, hi 0 (scaleAndClipData, genlog)
    ST_GENLOG_AXES = 3  // as ST_GENLOG, skipping the pixels on x = 0 or y = 0 (AI loop, :1366-1378)
};

struct StatArgs {
    int64_t npx;
    int mode;
    int extra_zero;   // the scan covers one more value, 0 (scaleIntensity reads one past the image)
    int w, xmin, ymin;  // image geometry for ST_GENLOG_AXES
};

__device__ __forceinline__ bool stat_value(const StatArgs& a, const double* img, int64_t i, double& v) {
    const double p = (i < a.npx) ? img[i] : 0.0;
    if (a.mode == ST_LIN) { v = p; return true; }
    if (!(p > 0)) return false;
    if (a.mode == ST_GENLOG_AXES && a.xmin < 0) {
        const int x = (int)(i % a.w) + a.xmin, y = (int)(i / a.w) + a.ymin;
        if (x == 0 || y == 0) return false;
    }
    v = log10(p);
    return true;
}

// lohi[2*image] = lo, [2*image+1] = hi of each image (blockIdx.x = image).
__global__ void __launch_bounds__(kStat) k_stats(const double* __restrict__ imgs, StatArgs a, double* __restrict__ lohi) {
    __shared__ double s_min[kStat];
    __shared__ double s_max[kStat];
    const double* img = imgs + (int64_t)blockIdx.x * a.npx;
    const int64_t n = a.npx + (a.extra_zero ? 1 : 0);
    const int64_t chunk = (n + kStat - 1) / kStat;
    const int64_t b = (int64_t)threadIdx.x * chunk, e = min(n, b + chunk);
    const bool sil = a.mode == ST_LIN || a.mode == ST_LOG;
    const double lo0 = sil ? 1e20 : 1e26, hi0 = sil ? 1e-20 : 0.0;
    double m = INFINITY;
    for (int64_t i = b; i < e; ++i) {
        double v;
        if (stat_value(a, img, i, v) && v < m) m = v;
    }
    s_min[threadIdx.x] = m;
    __syncthreads();
    // inclusive prefix minimum over the lanes (Hillis-Steele)
    for (int off = 1; off < kStat; off <<= 1) {
        const double o = threadIdx.x >= off ? s_min[threadIdx.x - off] : INFINITY;
        __syncthreads();
        if (o < s_min[threadIdx.x]) s_min[threadIdx.x] = o;
        __syncthreads();
    }
    double run = threadIdx.x ? s_min[threadIdx.x - 1] : INFINITY;
    if (lo0 < run) run = lo0;
    double hi = -INFINITY;
    for (int64_t i = b; i < e; ++i) {
        double v;
        if (!stat_value(a, img, i, v)) continue;
        if (v < run) run = v;
        else if (v > hi) hi = v;
    }
    s_max[threadIdx.x] = hi;
    __syncthreads();
    for (int off = kStat / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off && s_max[threadIdx.x + off] > s_max[threadIdx.x]) s_max[threadIdx.x] = s_max[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double lo = s_min[kStat - 1] < lo0 ? s_min[kStat - 1] : lo0;
        double h = s_max[0] > hi0 ? s_max[0] : hi0;
        if (sil && fabs(lo - h) < 1e-6) {  // scaleIntensity: widen an empty range
            lo *= 0.998;
            h *= 1.002;
            if (h == 0) h = 0.2;
        }
        lohi[2 * blockIdx.x] = lo;
        lohi[2 * blockIdx.x + 1] = h;
    }
}

// Map every value with its image's (lo, hi): linear (v - lo)/(hi - lo), or for the log modes
// (log10 v - lo)/(hi - lo) for v > 0, else 0.  Images along blockIdx.y.
__global__ void k_map(double* __restrict__ imgs, int64_t npx, int lin, const double* __restrict__ lohi) {
    double* img = imgs + (int64_t)blockIdx.y * npx;
    const double lo = lohi[2 * blockIdx.y], hi = lohi[2 * blockIdx.y + 1];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npx; i += (int64_t)gridDim.x * blockDim.x) {
        const double p = img[i];
        img[i] = lin ? (p - lo) / (hi - lo) : ((p > 0) ? (log10(p) - lo) / (hi - lo) : 0);
    }
}

// Per-image maximum with start value 0 (scaleAndClipData, scale).
__global__ void __launch_bounds__(kStat) k_max0(const double* __restrict__ imgs, int64_t npx, double* __restrict__ mx) {
    __shared__ double s[kStat];
    const double* img = imgs + (int64_t)blockIdx.x * npx;
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < npx; i += kStat) if (m < img[i]) m = img[i];
    s[threadIdx.x] = m;
    __syncthreads();
    for (int off = kStat / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off && s[threadIdx.x + off] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) mx[blockIdx.x] = s[0];
}

// scaleAndClipData's elementwise steps after its statistics: divide by the maximum (scale),
// the [1e-6, 1] log clip, the 40 % clip.
__global__ void k_scale_clip(double* __restrict__ imgs, int64_t npx, const double* __restrict__ mx, int clip,
                             int clip40) {
    double* img = imgs + (int64_t)blockIdx.y * npx;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npx; i += (int64_t)gridDim.x * blockDim.x) {
        double p = img[i];
        if (mx) p = p / mx[blockIdx.y];
        if (clip) {
            p = ((log(fabs(p) / 1e-20) / log(10.0)) + 6.0) / 6.0;
            if (p < 0) p = 0;
        }
        if (clip40) {
            p = 1.0 - (1.0 - p) / (1.0 - 0.4);
            if (p < 0) p = 0;
        }
        img[i] = p;
    }
}

// Noise d + z sqrt(d) (z/100 for images already scaled to [0, 1]); z has one value per pixel of
// every image, in the order the reference draws them.
__global__ void k_noise(double* __restrict__ img, const double* __restrict__ z, int64_t n, int scaled) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double d = img[i];
        img[i] = d + (scaled ? z[i] / 100.0 : z[i]) * sqrt(d);
    }
}

__global__ void k_zero(double* __restrict__ img, const int64_t* __restrict__ idx, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) img[idx[i]] = 0.0;
}

// r,phi cell: its writer (step k of the float radius loop, angle row i) or -1.
struct RphiArgs {
    int sx, sy, s;
    double phiber;
    int64_t npx;
};

// Per image: the centre and its cell map (writer table and radius steps of its irmax).
struct RphiImg {
    double xc, yc;
    int64_t writer_off;
    int32_t irs_off, pad;
};

__global__ void k_rphi(const double* __restrict__ imgs, RphiArgs a, const RphiImg* __restrict__ per,
                       const int* __restrict__ writers, const float* __restrict__ irss, double* __restrict__ outs) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= a.s * a.s) return;
    const RphiImg im = per[blockIdx.y];
    const double* d = imgs + (int64_t)blockIdx.y * a.npx;
    double* out = outs + (int64_t)blockIdx.y * a.s * a.s;
    const int wr = writers[im.writer_off + cell];
    if (wr < 0) { out[cell] = 0.0; return; }
    const int k = wr / (a.s + 1), i = wr % (a.s + 1);
    const double rad = irss[im.irs_off + k];
    double twt = 0.0, avg = 0.0;
    int umf = 1024 - (int)(1024. - a.phiber * rad);
    if (umf < 1) umf = 1;
    for (int j = 1; j <= umf; ++j) {
        const double phi = (i - 1.5) * a.phiber + (j - 0.5) * a.phiber / umf;
        const double xs = im.xc + rad * cos(phi), ys = im.yc + rad * sin(phi);
        const int ix1 = (int)xs, iy1 = (int)ys, ix2 = ix1 + 1, iy2 = iy1 + 1;
        const double rex = xs - ix1, rey = ys - iy1;
        const bool x1in = ix1 >= 0 && ix1 < a.sx, x2in = ix2 >= 0 && ix2 < a.sx;
        const bool y1in = iy1 >= 0 && iy1 < a.sy, y2in = iy2 >= 0 && iy2 < a.sy;
        if (x1in && y1in) { const double wt = (1.0 - rex) * (1.0 - rey); twt = twt + wt; avg = avg + wt * d[ix1 + (int64_t)a.sx * iy1]; }
        if (x2in && y1in) { const double wt = (rex) * (1.0 - rey); twt = twt + wt; avg = avg + wt * d[ix2 + (int64_t)a.sx * iy1]; }
        if (x1in && y2in) { const double wt = (1.0 - rex) * (rey); twt = twt + wt; avg = avg + wt * d[ix1 + (int64_t)a.sx * iy2]; }
        if (x2in && y2in) { const double wt = (rex) * (rey); twt = twt + wt; avg = avg + wt * d[ix2 + (int64_t)a.sx * iy2]; }
    }
    if (twt > 0.0) avg = avg / twt;
    out[cell] = (avg < 0) ? 0 : avg;
}

// ---- FFT image (calculateIFFT without FFTW, Source/sc_postproc.cpp:364-560, 833-1000) ----
// One workgroup per line of the 2D transform (rows of one pass, columns of the other), the
// line in LDS: bit-reversal permutation on load, then the butterfly levels.  The twiddles are
// the reference's recurrence, evaluated on the host in the same order and passed in a table
// (level l, l1 = 2^l entries from index l1 - 1), so every butterfly rounds as there.
__global__ void k_fft_lines(double2* __restrict__ buf, int64_t img_stride, int64_t line_stride, int64_t elem_stride,
                            int m, const double2* __restrict__ tw, int forward) {
    extern __shared__ double lds[];
    const int nn = 1 << m;
    double* xr = lds;
    double* xi = lds + nn;
    double2* line = buf + (int64_t)blockIdx.y * img_stride + (int64_t)blockIdx.x * line_stride;
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
        const int r = (int)(__brev((unsigned)i) >> (32 - m));
        const double2 v = line[(int64_t)r * elem_stride];
        xr[i] = v.x;
        xi[i] = v.y;
    }
    __syncthreads();
    for (int l = 0; l < m; ++l) {
        const int l1 = 1 << l, l2 = l1 << 1;
        for (int b = threadIdx.x; b < nn / 2; b += blockDim.x) {
            const int jj = b & (l1 - 1);
            const int i = (b >> l) * l2 + jj, i1 = i + l1;
            const double2 u = tw[l1 - 1 + jj];
            const double t1 = u.x * xr[i1] - u.y * xi[i1];
            const double t2 = u.x * xi[i1] + u.y * xr[i1];
            xr[i1] = xr[i] - t1;
            xi[i1] = xi[i] - t2;
            xr[i] += t1;
            xi[i] += t2;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
        double a = xr[i], b = xi[i];
        if (forward) {
            a /= (double)nn;
            b /= (double)nn;
        }
        line[(int64_t)i * elem_stride] = make_double2(a, b);
    }
}

// Input of the transform: c[y][x] = d[x + s y] / max, imaginary part 0.
__global__ void k_fft_load(const double* __restrict__ d, int64_t len, const double* __restrict__ mx,
                           double2* __restrict__ buf) {
    const double* src = d + (int64_t)blockIdx.y * len;
    double2* dst = buf + (int64_t)blockIdx.y * len;
    const double m = mx[blockIdx.y];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = make_double2(src[i] / m, 0.0);
}

// Output cell (x, y) takes c[x][y] (the reference reads the array transposed) at
// calculateIndex(swap, x, y); with the swap, the sources x = 0 and y = 0 are overwritten by
// x = s-1 and y = s-1 (same target), and row and column 0 are never written: 0 here.
__device__ __forceinline__ int64_t fft_index(int swap, int x, int y, int dim) {
    if (swap) {
        const int h = dim / 2;
        const int64_t cx = x < h ? (h - x) : (dim - 1 - x + h);
        const int64_t cy = y < h ? (h - y) : (dim - 1 - y + h);
        return cx + (int64_t)dim * cy;
    }
    return x + (int64_t)dim * y;
}

__global__ void k_fft_out(const double2* __restrict__ buf, int s, int out_type, int swap, double* __restrict__ out) {
    const int64_t len = (int64_t)s * s;
    const double2* c = buf + (int64_t)blockIdx.y * len;
    double* o = out + (int64_t)blockIdx.y * len;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(k / s), x = (int)(k % s);
        if (swap && (x == 0 || y == 0)) {
            o[k] = 0.0;  // cell (0, k) or (k, 0) of the output: never written
            continue;
        }
        const double2 v = c[(int64_t)x * s + y];
        double r;
        switch (out_type) {
        case 0: r = v.x; break;
        case 1: r = v.y; break;
        case 2: r = sqrt(v.x * v.x + v.y * v.y); break;
        default: r = (v.x + v.y) * (v.x - v.y); break;
        }
        o[fft_index(swap, x, y, s)] = r;
    }
}

// Centred cut-out of each s x s image into cx x cy.
__global__ void k_cut(const double* __restrict__ full, int s, int cx, int cy, double* __restrict__ out) {
    const int64_t n = (int64_t)cx * cy;
    const double* f = full + (int64_t)blockIdx.y * s * s;
    double* o = out + (int64_t)blockIdx.y * n;
    const int x0 = s / 2 - cx / 2, y0 = s / 2 - cy / 2;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int yo = (int)(k / cx), xo = (int)(k % cx);
        o[k] = f[(x0 + xo) + (int64_t)s * (y0 + yo)];
    }
}

// Zeroed rectangles (beam stop, lines): one workgroup per (rectangle, image); a cell (x, y) of
// rectangle xa..xb, ya..yb (inclusive) is pixel ii = (x - x0) + sx (y - y0), zeroed when
// 0 <= ii < npx (a line past the right edge continues on the next row).
__global__ void k_zero_rects(double* __restrict__ img, int64_t npx, const int4* __restrict__ rects, int nrect, int x0,
                             int sx, int y0) {
    const int4 r = rects[(int64_t)blockIdx.y * nrect + blockIdx.x];
    const int64_t wx = (int64_t)r.y - r.x + 1, wy = (int64_t)r.w - r.z + 1;
    if (wx <= 0 || wy <= 0) return;
    double* im = img + (int64_t)blockIdx.y * npx;
    for (int64_t c = threadIdx.x; c < wx * wy; c += blockDim.x) {
        const int64_t x = r.x + c % wx, y = r.z + c / wx;
        const int64_t ii = (x - x0) + (int64_t)sx * (y - y0);
        if (ii >= 0 && ii < npx) im[ii] = 0.0;
    }
}

int blocks_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

// ---- zuf() values of the rand() sequence on the device ------------------------------------
// One lane per (run, chunk of kRandChunk values).  The run's start state (its last 31 terms,
// oldest first) is read, moved to the chunk's start with the chunk's jump rows (the same for
// the whole workgroup, so they are read through the scalar path), and the chunk's terms follow
// from r[i] = r[i-31] + r[i-3] with the 31 terms in registers: in a block of 31 steps the term
// three back is the one written three steps before, or for the first three steps one of the
// block's last three old terms.
constexpr int kRandChunk = 31 * 16;

__device__ inline double zuf_of(int32_t v) {
#pragma clang fp contract(off)
    return ((double)v / 2147483647.0 * 200.0 - 100.0) / 100.0;
}

__global__ void __launch_bounds__(64) k_rand_zuf(const uint32_t* __restrict__ win, int64_t nseq,
                                                 const uint32_t* __restrict__ jump, int64_t len,
                                                 double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int c = blockIdx.y;
    if (k >= nseq) return;
    uint32_t w[31];
    {
        uint32_t u[31];
#pragma unroll
        for (int j = 0; j < 31; ++j) u[j] = win[k * 31 + j];
        if (c == 0) {
#pragma unroll
            for (int j = 0; j < 31; ++j) w[j] = u[j];
        } else {
            const uint32_t* A = jump + (size_t)(c - 1) * 961;
#pragma unroll
            for (int m = 0; m < 31; ++m) {
                uint32_t v = 0;
#pragma unroll
                for (int j = 0; j < 31; ++j) v += A[m * 31 + j] * u[j];
                w[m] = v;
            }
        }
    }
    const int64_t t0 = (int64_t)c * kRandChunk;
    const int n = (int)min((int64_t)kRandChunk, len - t0);
    double* o = out + k * len + t0;
    for (int b = 0; b < n; b += 31) {
#pragma unroll
        for (int j = 0; j < 31; ++j) {
            const uint32_t v = w[j] + w[(j + 28) % 31];
            w[j] = v;
            if (b + j < n) o[b + j] = zuf_of((int32_t)(v >> 1));
        }
    }
}

}  // namespace csp

using namespace csp;

#define PCHK(c, call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return csx_ctx_fail(c, CSX_E_HIP, (std::string(#call) + ": " + hipGetErrorString(e_)).c_str()); \
    } while (0)

int csx_post_rand_zuf(csx_ctx* c, const uint32_t* h_win, int64_t nseq, int64_t len, double* d_out) {
    if (!c) return CSX_E_ARG;
    if (nseq == 0 || len == 0) return CSX_OK;
    const int64_t nch = (len + kRandChunk - 1) / kRandChunk;
    if (!h_win || !d_out || nseq < 0 || len < 0 || nch > 65535 || (nseq + 63) / 64 > INT32_MAX)
        return csx_ctx_fail(c, CSX_E_ARG, "bad random-value runs");
    if (hipSetDevice(csx_ctx_device(c)) != hipSuccess) return csx_ctx_fail(c, CSX_E_HIP, "hipSetDevice");
    // jump rows of c * kRandChunk terms, c = 1 .. nch-1
    std::vector<uint32_t> rows((size_t)std::max<int64_t>(nch - 1, 1) * csr::kRows);
    uint32_t step[csr::kDeg], a[csr::kDeg], t[csr::kDeg];
    csr::xpow(kRandChunk, step);
    for (int j = 0; j < csr::kDeg; ++j) a[j] = step[j];
    for (int64_t ci = 1; ci < nch; ++ci) {
        csr::poly_rows(a, &rows[(size_t)(ci - 1) * csr::kRows]);
        csr::polymulmod(a, step, t);
        for (int j = 0; j < csr::kDeg; ++j) a[j] = t[j];
    }
    const size_t wbytes = (size_t)nseq * csr::kDeg * sizeof(uint32_t), rbytes = rows.size() * sizeof(uint32_t);
    char* scr = (char*)csx_ctx_scratch(c, 0, ((wbytes + 255) & ~size_t(255)) + rbytes);
    if (!scr) return csx_ctx_fail(c, CSX_E_NOMEM, "random-value scratch");
    uint32_t* d_win = (uint32_t*)scr;
    uint32_t* d_rows = (uint32_t*)(scr + ((wbytes + 255) & ~size_t(255)));
    hipStream_t st = (hipStream_t)csx_ctx_stream(c);
    PCHK(c, hipMemcpyAsync(d_win, h_win, wbytes, hipMemcpyHostToDevice, st));
    PCHK(c, hipMemcpyAsync(d_rows, rows.data(), rbytes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_rand_zuf, dim3((unsigned)((nseq + 63) / 64), (unsigned)nch), dim3(64), 0, st, d_win, nseq,
                       d_rows, len, d_out);
    PCHK(c, hipGetLastError());
    // the host arrays must outlive the copies
    PCHK(c, hipStreamSynchronize(st));
    return CSX_OK;
}

namespace {

int stats_and_map(csx_ctx* c, double* d, int64_t npx, int nimg, const StatArgs& a, int lin) {
    hipStream_t st = (hipStream_t)csx_ctx_stream(c);
    double* lohi = (double*)csx_ctx_scratch(c, 0, (size_t)nimg * 2 * sizeof(double));
    if (!lohi) return csx_ctx_fail(c, CSX_E_NOMEM, "post-processing scratch");
    hipLaunchKernelGGL(k_stats, dim3(nimg), dim3(kStat), 0, st, d, a, lohi);
    PCHK(c, hipGetLastError());
    hipLaunchKernelGGL(k_map, dim3(blocks_for(npx), nimg), dim3(256), 0, st, d, npx, lin, lohi);
    PCHK(c, hipGetLastError());
    return CSX_OK;
}

int check(csx_ctx* c, const void* d, int64_t npx, int nimg) {
    if (!c) return CSX_E_ARG;
    if (!d || npx <= 0 || nimg <= 0 || nimg > 65535) return csx_ctx_fail(c, CSX_E_ARG, "bad image batch");
    if (hipSetDevice(csx_ctx_device(c)) != hipSuccess) return csx_ctx_fail(c, CSX_E_HIP, "hipSetDevice");
    return CSX_OK;
}

}  // namespace

extern "C" {

int csx_rand_zuf_device(csx_ctx* c, csx_rand* r, int64_t n, double* d_out) {
    if (!c) return CSX_E_ARG;
    if (!r || n < 0 || (!d_out && n > 0)) return csx_ctx_fail(c, CSX_E_ARG, "bad random-value request");
    // runs of kRun values from states kRun apart (one jump each), then the rest as one run
    constexpr int64_t kRun = (int64_t)kRandChunk * 64;
    const int64_t nfull = n / kRun, rest = n - nfull * kRun;
    std::vector<uint32_t> wins((size_t)(nfull + 1) * csr::kDeg), rows(csr::kRows);
    if (nfull) csr::jump_rows((uint64_t)kRun, rows.data());
    for (int64_t k = 0; k < nfull; ++k) {
        csr::window(r->s, &wins[(size_t)k * csr::kDeg]);
        csr::advance(r->s, rows.data());
    }
    int rc = csx_post_rand_zuf(c, wins.data(), nfull, kRun, d_out);
    if (rc == CSX_OK && rest) {
        csr::window(r->s, &wins[0]);
        csr::jump_rows((uint64_t)rest, rows.data());
        csr::advance(r->s, rows.data());
        rc = csx_post_rand_zuf(c, wins.data(), 1, rest, d_out + nfull * kRun);
    }
    return rc;
}

int csx_post_noise(csx_ctx* c, double* d_img, const double* d_z, int64_t n, int scaled) {
    if (int rc = check(c, d_img, n, 1)) return rc;
    if (!d_z) return csx_ctx_fail(c, CSX_E_ARG, "no noise values");
    hipLaunchKernelGGL(k_noise, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)csx_ctx_stream(c), d_img, d_z, n, scaled);
    PCHK(c, hipGetLastError());
    return CSX_OK;
}

int csx_post_zero(csx_ctx* c, double* d_img, const int64_t* d_idx, int64_t n) {
    if (!c) return CSX_E_ARG;
    if (n == 0) return CSX_OK;
    if (int rc = check(c, d_img, n, 1)) return rc;
    if (!d_idx) return csx_ctx_fail(c, CSX_E_ARG, "no indices");
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)csx_ctx_stream(c), d_img,
                       d_idx, n);
    PCHK(c, hipGetLastError());
    return CSX_OK;
}

int csx_post_zero_rects(csx_ctx* c, double* d_img, int64_t npx, int nimg, const int32_t* h_rects, int nrect, int x0,
                        int x1, int y0, int y1) {
    if (int rc = check(c, d_img, npx, nimg)) return rc;
    if (nrect == 0) return CSX_OK;
    if (!h_rects || nrect < 0 || nrect > 65535 || x1 <= x0 || y1 <= y0 || (int64_t)(x1 - x0) * (y1 - y0) != npx)
        return csx_ctx_fail(c, CSX_E_ARG, "bad rectangles or image bounds");
    const size_t count = (size_t)nimg * nrect;
    for (size_t i = 0; i < count; ++i) {
        const int32_t* r = h_rects + 4 * i;
        const int64_t wx = (int64_t)r[1] - r[0] + 1, wy = (int64_t)r[3] - r[2] + 1;
        if (wx > 0 && wy > 0 && wx * wy > 64 * npx) return csx_ctx_fail(c, CSX_E_ARG, "rectangle far larger than the image");
    }
    const size_t bytes = count * 4 * sizeof(int32_t);
    int4* d_r = (int4*)csx_ctx_scratch(c, 0, bytes);
    if (!d_r) return csx_ctx_fail(c, CSX_E_NOMEM, "post-processing scratch");
    hipStream_t st = (hipStream_t)csx_ctx_stream(c);
    PCHK(c, hipMemcpyAsync(d_r, h_rects, bytes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_rects, dim3(nrect, nimg), dim3(256), 0, st, d_img, npx, d_r, nrect, x0, x1 - x0, y0);
    PCHK(c, hipGetLastError());
    // the host array must outlive the copy
    PCHK(c, hipStreamSynchronize(st));
    return CSX_OK;
}

int csx_post_scale_intensity(csx_ctx* c, double* d_img, int64_t npx, int nimg, int linlog) {
    if (int rc = check(c, d_img, npx, nimg)) return rc;
    StatArgs a{npx, linlog ? ST_LIN : ST_LOG, 1, 1, 0, 0};
    return stats_and_map(c, d_img, npx, nimg, a, linlog);
}

int csx_post_log_scale(csx_ctx* c, double* d_img, int64_t npx, int nimg, int w, int xmin, int ymin, int skip_axes) {
    if (int rc = check(c, d_img, npx, nimg)) return rc;
    if (skip_axes && (w <= 0 || npx % w)) return csx_ctx_fail(c, CSX_E_ARG, "image width does not divide the pixel count");
    StatArgs a{npx, skip_axes ? ST_GENLOG_AXES : ST_GENLOG, 0, w > 0 ? w : 1, xmin, ymin};
    return stats_and_map(c, d_img, npx, nimg, a, 0);
}

int csx_post_scale_clip(csx_ctx* c, double* d_img, int64_t npx, int nimg, int scale, int clip, int clip40, int genlog) {
    if (int rc = check(c, d_img, npx, nimg)) return rc;
    hipStream_t st = (hipStream_t)csx_ctx_stream(c);
    if (genlog) {
        StatArgs a{npx, ST_GENLOG, 0, 1, 0, 0};
        if (int rc = stats_and_map(c, d_img, npx, nimg, a, 0)) return rc;
    }
    if (!scale && !clip && !clip40) return CSX_OK;
    double* mx = nullptr;
    if (scale) {
        mx = (double*)csx_ctx_scratch(c, 1, (size_t)nimg * sizeof(double));
        if (!mx) return csx_ctx_fail(c, CSX_E_NOMEM, "post-processing scratch");
        hipLaunchKernelGGL(k_max0, dim3(nimg), dim3(kStat), 0, st, d_img, npx, mx);
        PCHK(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_scale_clip, dim3(blocks_for(npx), nimg), dim3(256), 0, st, d_img, npx, mx, clip, clip40);
    PCHK(c, hipGetLastError());
    return CSX_OK;
}

int csx_post_rphi_many(csx_ctx* c, const double* d_img, int x0, int x1, int y0, int y1, const int32_t* h_centres,
                       int s, int nimg, double* d_out, int scale, int clip, int clip40) {
    const int64_t npx = (int64_t)(x1 - x0) * (y1 - y0);
    if (int rc = check(c, d_img, npx, nimg)) return rc;
    if (!d_out || !h_centres || s < 2 || s > 8192) return csx_ctx_fail(c, CSX_E_ARG, "r,phi size must be 2..8192");
    // image coordinates from 0 (the reference shifts the range and the centre alike)
    const int dx = x0 < 0 ? -x0 : 0, dy = y0 < 0 ? -y0 : 0;
    x0 += dx; x1 += dx; y0 += dy; y1 += dy;
    const int sx = x1 - x0, sy = y1 - y0;
    const double phiber = 2. * M_PI / s;
    const int64_t cells = (int64_t)s * s;
    auto sq = [](double v) { return v * v; };
    // the float radius loop and the cell each (step, row) writes, as the reference computes
    // them (a later write of a cell replaces an earlier one); it depends on irmax only, so one
    // map serves every image with the same irmax (found by hash).  Images go in chunks whose
    // maps stay within kRphiMapBytes (a chunk takes at least one map), so many distinct beam
    // centres cost launches, not memory.
    constexpr size_t kRphiMapBytes = size_t(256) << 20;
    hipStream_t st = (hipStream_t)csx_ctx_stream(c);
    std::vector<int> writers;
    std::vector<float> irss;
    std::vector<RphiImg> per;
    std::unordered_map<int, std::pair<int64_t, int32_t>> maps;  // irmax -> (writer, radius offsets)
    for (int k0 = 0; k0 < nimg;) {
        writers.clear();
        irss.clear();
        per.clear();
        maps.clear();
        int k = k0;
        for (; k < nimg; ++k) {
            const int xc = h_centres[2 * k] + dx, yc = h_centres[2 * k + 1] + dy;
            double rmax = std::sqrt(sq(1. - xc) + sq(1. - yc));
            rmax = std::max(rmax, std::sqrt(sq((sx - 1.) - xc) + sq(1. - yc)));
            rmax = std::max(rmax, std::sqrt(sq((sx - 1.) - xc) + sq((sy - 1.) - yc)));
            rmax = std::max(rmax, std::sqrt(sq(1. - xc) + sq((sy - 1.) - yc)));
            const int irmax = (int)(rmax + 0.5);
            auto it = maps.find(irmax);
            if (it == maps.end()) {
                if (!maps.empty() && (writers.size() + (size_t)cells) * sizeof(int) > kRphiMapBytes) break;
                const float irstp = (float)irmax / (float)(s - 1);
                const int64_t w0 = (int64_t)writers.size();
                const int32_t i0 = (int32_t)irss.size();
                writers.resize(writers.size() + (size_t)cells, -1);
                int steps = 0;
                for (float ir = 0; ir <= irmax; ir += irstp) {
                    irss.push_back(ir);
                    for (int i = 0; i <= s; ++i) {
                        const int idx = (int)((float)(i * s) + ir / irstp);
                        if (idx < cells && idx >= 0) writers[(size_t)(w0 + idx)] = steps * (s + 1) + i;
                    }
                    if (++steps > 4 * s + 16) return csx_ctx_fail(c, CSX_E_ARG, "r,phi radius loop does not end");
                }
                it = maps.emplace(irmax, std::make_pair(w0, i0)).first;
            }
            per.push_back(RphiImg{(double)xc, (double)yc, it->second.first, it->second.second, 0});
        }
        const int nk = k - k0;
        const size_t wbytes = writers.size() * sizeof(int), ibytes = std::max<size_t>(irss.size(), 1) * sizeof(float),
                     pbytes = per.size() * sizeof(RphiImg);
        auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
        char* scr = (char*)csx_ctx_scratch(c, 0, up(wbytes) + up(ibytes) + pbytes);
        if (!scr) return csx_ctx_fail(c, CSX_E_NOMEM, "post-processing scratch");
        int* d_writers = (int*)scr;
        float* d_irs = (float*)(scr + up(wbytes));
        RphiImg* d_per = (RphiImg*)(scr + up(wbytes) + up(ibytes));
        PCHK(c, hipMemcpyAsync(d_writers, writers.data(), wbytes, hipMemcpyHostToDevice, st));
        if (!irss.empty())
            PCHK(c, hipMemcpyAsync(d_irs, irss.data(), irss.size() * sizeof(float), hipMemcpyHostToDevice, st));
        PCHK(c, hipMemcpyAsync(d_per, per.data(), pbytes, hipMemcpyHostToDevice, st));
        RphiArgs a{sx, sy, s, phiber, npx};
        hipLaunchKernelGGL(k_rphi, dim3((unsigned)((cells + 255) / 256), nk), dim3(256), 0, st, d_img + (int64_t)k0 * npx,
                           a, d_per, d_writers, d_irs, d_out + (int64_t)k0 * cells);
        PCHK(c, hipGetLastError());
        // the host vectors must outlive the copies, and the scratch is reused by the next chunk
        PCHK(c, hipStreamSynchronize(st));
        k0 = k;
    }
    return csx_post_scale_clip(c, d_out, cells, nimg, scale, clip, clip40, 0);
}

int csx_post_rphi(csx_ctx* c, const double* d_img, int x0, int x1, int y0, int y1, int xc, int yc, int s, int nimg,
                  double* d_out, int scale, int clip, int clip40) {
    if (!c) return CSX_E_ARG;
    if (nimg <= 0 || nimg > 65535) return csx_ctx_fail(c, CSX_E_ARG, "bad image batch");
    std::vector<int32_t> centres((size_t)nimg * 2);
    for (int k = 0; k < nimg; ++k) {
        centres[2 * (size_t)k] = xc;
        centres[2 * (size_t)k + 1] = yc;
    }
    return csx_post_rphi_many(c, d_img, x0, x1, y0, y1, centres.data(), s, nimg, d_out, scale, clip, clip40);
}

int csx_post_ifft(csx_ctx* c, const double* d_in, int s, int nimg, int forward, int out_type, int scaled, int clip,
                  int clip40, int swap, int cutx, int cuty, double* d_out) {
    const int64_t len = (int64_t)s * s;
    if (int rc = check(c, d_in, len, nimg)) return rc;
    int m = 0;
    while ((1 << m) < s) ++m;
    if (!d_out || s < 4 || s > 2048 || (1 << m) != s) return csx_ctx_fail(c, CSX_E_ARG, "FFT size must be a power of two, 4..2048");
    if (out_type < 0 || out_type > 3) return csx_ctx_fail(c, CSX_E_ARG, "FFT output type 0..3");
    const bool cut = cutx > 0 && cuty > 0 && cutx < s && cuty < s;
    // twiddles of every level, by the reference's recurrence
    std::vector<double2> tw((size_t)s - 1);
    double c1 = -1.0, c2 = 0.0;
    for (int l = 0, l1 = 1; l < m; ++l, l1 <<= 1) {
        double u1 = 1.0, u2 = 0.0;
        for (int jj = 0; jj < l1; ++jj) {
            tw[(size_t)(l1 - 1 + jj)] = make_double2(u1, u2);
            const double z = u1 * c1 - u2 * c2;
            u2 = u1 * c2 + u2 * c1;
            u1 = z;
        }
        c2 = std::sqrt((1.0 - c1) / 2.0);
        if (forward) c2 = -c2;
        c1 = std::sqrt((1.0 + c1) / 2.0);
    }
    // scratch 0: complex lines; scratch 1: maxima, twiddles, and the full image before a cut-out
    const size_t cbytes = (size_t)len * nimg * sizeof(double2);
    const size_t o_tw = ((size_t)nimg * sizeof(double) + 255) & ~size_t(255);
    const size_t o_full = (o_tw + tw.size() * sizeof(double2) + 255) & ~size_t(255);
    const size_t sbytes = o_full + (cut ? (size_t)len * nimg * sizeof(double) : 0);
    double2* buf = (double2*)csx_ctx_scratch(c, 0, cbytes);
    char* s1 = (char*)csx_ctx_scratch(c, 1, sbytes);
    if (!buf || !s1) return csx_ctx_fail(c, CSX_E_NOMEM, "FFT scratch");
    double* mx = (double*)s1;
    double2* d_tw = (double2*)(s1 + o_tw);
    double* full = cut ? (double*)(s1 + o_full) : d_out;
    hipStream_t st = (hipStream_t)csx_ctx_stream(c);
    PCHK(c, hipMemcpyAsync(d_tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_max0, dim3(nimg), dim3(kStat), 0, st, d_in, len, mx);
    hipLaunchKernelGGL(k_fft_load, dim3(blocks_for(len), nimg), dim3(256), 0, st, d_in, len, mx, buf);
    const int threads = std::min(1024, std::max(64, s / 2));
    const size_t shm = 2 * (size_t)s * sizeof(double);
    // along the first index (stride s) for every second index, then along the second
    hipLaunchKernelGGL(k_fft_lines, dim3(s, nimg), dim3(threads), shm, st, buf, len, (int64_t)1, (int64_t)s, m, d_tw,
                       forward);
    hipLaunchKernelGGL(k_fft_lines, dim3(s, nimg), dim3(threads), shm, st, buf, len, (int64_t)s, (int64_t)1, m, d_tw,
                       forward);
    hipLaunchKernelGGL(k_fft_out, dim3(blocks_for(len), nimg), dim3(256), 0, st, buf, s, out_type, swap, full);
    PCHK(c, hipGetLastError());
    // (its maxima go to the start of scratch 1, which mx no longer needs; the twiddles and
    // the full image lie past them)
    if (int rc = csx_post_scale_clip(c, full, len, nimg, scaled, clip, clip40, 0)) return rc;
    if (cut) {
        hipLaunchKernelGGL(k_cut, dim3(blocks_for((int64_t)cutx * cuty), nimg), dim3(256), 0, st, full, s, cutx, cuty,
                           d_out);
        PCHK(c, hipGetLastError());
    }
    // the twiddle table on the host must outlive its copy
    PCHK(c, hipStreamSynchronize(st));
    return CSX_OK;
}

}  // extern "C"
