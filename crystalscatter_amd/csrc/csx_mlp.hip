// Surrogate network of the fits: a small MLP that maps a parameter set and a pixel's (qx, qy)
// to the normalised log-intensity of that pixel.  It has no counterpart in the reference; the
// fit driver uses it to propose starting simplices and to pre-screen vertices, and the exact
// kernels re-score every point it accepts.
//
//   input   6 normalised parameters + 64 Fourier features sin/cos(a_f), 32 phases
//           a_f = 2*pi*(F_f0*qx + F_f1*qy) + F_f2*|q|*v_f, where v_f is the physical value of
//           parameter pidx_f (offset + scale * normalised value) or 1 (pidx_f < 0)
//           (70 values, zero-padded to 128 columns)
//   hidden  4 x 256, SiLU
//   output  1 (linear)
//
// Precision: BF16 weights and activations on the matrix cores (v_mfma_f32_16x16x32_bf16),
// FP32 accumulation, FP32 master weights and Adam state.  Every matrix product is one
// "NT" kernel, C[m][n] = sum_k A[m][k] * B[n][k] with both operands k-contiguous:
//   forward   Z_l  = H_{l-1} . W_l^T      A = H_{l-1} [M][K], B = W_l   [N][K] (out x in)
//   backward  dH   = dZ_l . W_l          A = dZ_l [M][N],    B = W_l^T [K][N]
//   weights   dW_l = dZ_l^T . H_{l-1}    A = dZ_l^T [N][M],  B = H_{l-1}^T [K][M], split over M
// so the forward pass also stores H^T and the backward pass dZ^T (bf16), and each weight is
// kept in bf16 in both orientations.  Per-sample buffers are sized for max_batch rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "csx.h"

namespace csm {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = CSX_MLP_IN_PAD;    // padded input width
constexpr int kUsed = CSX_MLP_IN;      // used inputs: 6 parameters + 64 Fourier features
constexpr int kH = CSX_MLP_HIDDEN;     // hidden width
constexpr int kL = CSX_MLP_LAYERS;     // hidden layers
constexpr int kNF = CSX_MLP_NFREQ;     // Fourier frequencies
constexpr int kNP = CSX_MLP_NPAR;      // parameters per set
constexpr int kBM = 128, kBN = 128, kBK = 64, kLd = kBK + 8;  // GEMM tile, LDS row stride (bf16)
constexpr int kTp = kBN + 8;            // row stride (bf16) of the epilogue's output tile in LDS
constexpr int kHeadRows = 64;           // rows per block of the output-layer kernels

// Offsets of the FP32 master parameters (one buffer): hidden layer l has W_l [kH][in_l]
// (out x in) then b_l [kH]; then the output weights w [kH] and bias.
__host__ __device__ constexpr int in_of(int l) { return l == 0 ? kIn : kH; }
__host__ __device__ constexpr int64_t w_off(int l) {
    return l == 0 ? 0 : (int64_t)kH * kIn + kH + (int64_t)(l - 1) * ((int64_t)kH * kH + kH);
}
__host__ __device__ constexpr int64_t b_off(int l) { return w_off(l) + (int64_t)kH * in_of(l); }
constexpr int64_t kOutW = w_off(kL);
constexpr int64_t kOutB = kOutW + kH;
constexpr int64_t kNParams = kOutB + 1;

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

// ---- input features ----------------------------------------------------------------
// Feature phases (see the header comment): frequencies F [kNF][3], parameter index per
// phase, and the affine map normalised -> physical value per parameter.
struct Features {
    float f[kNF][3];
    int32_t pidx[kNF];
    float off[kNP], sc[kNP];
};

// Row i < n: parameters of set[i] (already normalised by the caller; an index outside
// [0, nsets) is clamped, the host checks them), then sin and cos of the kNF phases; rows
// n..mpad-1 are zero.  Writes X [mpad][kIn] and X^T [kIn][mpad].
__global__ void __launch_bounds__(256) k_embed(int n, int mpad, int nsets, const float* __restrict__ par,
                                               const int32_t* __restrict__ set, const float* __restrict__ q,
                                               const Features ft, bf16* __restrict__ X, bf16* __restrict__ XT) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= mpad) return;
    float v[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) v[k] = 0.0f;
    if (i < n) {
        const float* p = par + (size_t)min(max(set[i], 0), nsets - 1) * kNP;  // clamped: never out of bounds
#pragma unroll
        for (int k = 0; k < kNP; ++k) v[k] = p[k];
        const float qx = q[2 * i], qy = q[2 * i + 1];
        const float qa = sqrtf(qx * qx + qy * qy);
#pragma unroll
        for (int f = 0; f < kNF; ++f) {
            const int j = ft.pidx[f];
            float pv = 1.0f;  // constant indices only: v stays in registers
#pragma unroll
            for (int k = 0; k < kNP; ++k) pv = (k == j) ? ft.off[k] + ft.sc[k] * v[k] : pv;
            float s, c;
            sincosf(6.283185307f * (ft.f[f][0] * qx + ft.f[f][1] * qy) + ft.f[f][2] * qa * pv, &s, &c);
            v[kNP + f] = s;
            v[kNP + kNF + f] = c;
        }
    }
    bf16* row = X + (size_t)i * kIn;
#pragma unroll
    for (int k = 0; k < kIn; k += 8) {
        bf16x8 pk;
#pragma unroll
        for (int j = 0; j < 8; ++j) pk[j] = (bf16)v[k + j];
        *(bf16x8*)(row + k) = pk;
    }
#pragma unroll
    for (int k = 0; k < kIn; ++k) XT[(size_t)k * mpad + i] = (bf16)v[k];
}

// ---- the NT matrix product with its epilogues ---------------------------------------
enum Epi { EPI_FWD = 0, EPI_BWD = 1, EPI_PART = 2 };
struct EpiArgs {
    const float* bias;   // FWD: b_l
    bf16* z;             // FWD: Z_l [M][N] out
    bf16* h;             // FWD: H_l [M][N] out
    bf16* ht;            // FWD: H_l^T [N][M] out
    const bf16* zprev;   // BWD: Z_{l-1} [M][N]
    bf16* dz;            // BWD: dZ_{l-1} [M][N] out
    bf16* dzt;           // BWD: dZ_{l-1}^T [N][M] out
    float* part;         // PART: partials [split][M][N]
    int ldo;             // leading dimension of [M][N] outputs (N)
    int ldt;             // leading dimension of the transposed outputs (M)
};

// Grid (M/128, N/128, splits); 256 threads, 4 waves in 2x2, each wave a 64x64 block of
// 4x4 MFMA tiles.  Requires M, N multiples of 128, K and kchunk multiples of 64, lda/ldb
// multiples of 8 (checked by the host).
// Row phase of the epilogue: the 128x128 bf16 tile T (row stride kTp) to out[m][n] rows with
// 16-byte stores (out row stride ld).
__device__ __forceinline__ void tile_store_rows(const bf16* T, bf16* __restrict__ out, int ld, int tm, int tn) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int id = threadIdx.x + 256 * it;
        const int row = id >> 4, c8 = (id & 15) * 8;
        *(uint4*)&out[(size_t)(tm + row) * ld + tn + c8] = *(const uint4*)&T[row * kTp + c8];
    }
}
// Column phase: the transpose of T to outT[n][m] (row stride ldt), 8 rows per 16-byte store;
// consecutive threads take consecutive columns (conflict-free LDS reads).
__device__ __forceinline__ void tile_store_cols(const bf16* T, bf16* __restrict__ outT, int ldt, int tm, int tn) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int id = threadIdx.x + 256 * it;
        const int col = id & 127, m8 = (id >> 7) * 8;
        bf16x8 pk;
#pragma unroll
        for (int k = 0; k < 8; ++k) pk[k] = T[(m8 + k) * kTp + col];
        *(bf16x8*)&outT[(size_t)(tn + col) * ldt + tm + m8] = pk;
    }
}

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_nt(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B,
                                                 int ldb, int K, int kchunk, EpiArgs ep) {
    // operand tiles during the main loop, the output tile (128 x kTp bf16) in the epilogue
    __shared__ __attribute__((aligned(16))) bf16 smem[2 * kBM * kLd];
    static_assert(kBM * kTp <= 2 * kBM * kLd, "epilogue tile must fit in the operand tiles");
    bf16* sA = smem;
    bf16* sB = smem + kBM * kLd;
    bf16* T = smem;
    const int tm = blockIdx.x * kBM, tn = blockIdx.y * kBN;
    const int k0 = blockIdx.z * kchunk, k1 = min(K, k0 + kchunk);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    const int fr = lane & 15, fk = (lane >> 4) * 8;  // fragment row, k offset
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kk = k0; kk < k1; kk += kBK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int idx = t + r * 256;
            const int row = idx >> 3, c8 = (idx & 7) * 8;
            *(uint4*)&sA[row * kLd + c8] = *(const uint4*)&A[(size_t)(tm + row) * lda + kk + c8];
            *(uint4*)&sB[row * kLd + c8] = *(const uint4*)&B[(size_t)(tn + row) * ldb + kk + c8];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 32) {
            bf16x8 af[4], bfg[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = *(const bf16x8*)&sA[(wm + i * 16 + fr) * kLd + ks + fk];
#pragma unroll
            for (int j = 0; j < 4; ++j) bfg[j] = *(const bf16x8*)&sB[(wn + j * 16 + fr) * kLd + ks + fk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfg[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator (i, j)[r] is C[tm + wm + 16i + 4(lane/16) + r][tn + wn + 16j + lane%16]
    if constexpr (EPI == EPI_PART) {
        float* out = ep.part + ((size_t)blockIdx.z * gridDim.x * kBM) * ep.ldo;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m0 = tm + wm + i * 16 + (lane >> 4) * 4;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = tn + wn + j * 16 + fr;
#pragma unroll
                for (int r = 0; r < 4; ++r) out[(size_t)(m0 + r) * ep.ldo + n] = acc[i][j][r];
            }
        }
    } else if constexpr (EPI == EPI_FWD) {
        // Z = acc + b, then H = silu(Z), each staged through LDS for row and column stores
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float b = ep.bias[tn + wn + j * 16 + fr];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    T[(wm + i * 16 + (lane >> 4) * 4 + r) * kTp + wn + j * 16 + fr] = (bf16)(acc[i][j][r] + b);
        }
        __syncthreads();
        tile_store_rows(T, ep.z, ep.ldo, tm, tn);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float b = ep.bias[tn + wn + j * 16 + fr];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float z = acc[i][j][r] + b;
                    T[(wm + i * 16 + (lane >> 4) * 4 + r) * kTp + wn + j * 16 + fr] = (bf16)(z * sigm(z));
                }
        }
        __syncthreads();
        tile_store_rows(T, ep.h, ep.ldo, tm, tn);
        tile_store_cols(T, ep.ht, ep.ldt, tm, tn);
    } else {
        // dZ = dH * silu'(Z_prev): Z_prev tile into LDS (row loads), combined in place
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int id = threadIdx.x + 256 * it;
            const int row = id >> 4, c8 = (id & 15) * 8;
            *(uint4*)&T[row * kTp + c8] = *(const uint4*)&ep.zprev[(size_t)(tm + row) * ep.ldo + tn + c8];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    bf16& e = T[(wm + i * 16 + (lane >> 4) * 4 + r) * kTp + wn + j * 16 + fr];
                    e = (bf16)(acc[i][j][r] * silu_grad((float)e));
                }
        __syncthreads();
        tile_store_rows(T, ep.dz, ep.ldo, tm, tn);
        tile_store_cols(T, ep.dzt, ep.ldt, tm, tn);
    }
}

// ---- output layer --------------------------------------------------------------------
// y[m] = H4[m] . w + b for rows m < mpad; one wave per row.
__global__ void __launch_bounds__(256) k_head_fwd(int mpad, const bf16* __restrict__ H, const float* __restrict__ w,
                                                  const float* __restrict__ b, float* __restrict__ y) {
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (m >= mpad) return;
    const bf16* row = H + (size_t)m * kH + lane * 4;
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) s = fmaf((float)row[c], w[lane * 4 + c], s);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[m] = s + b[0];
}

// ---- inference in one pass -------------------------------------------------------------
// Each block takes kPM rows through the embedding, the kL hidden layers and the output layer
// with the rows' activations in one LDS tile (kPM x kH bf16, row stride kPLd): a layer's
// products are accumulated in registers, the block synchronises, and the epilogue writes the
// layer's output over its input.  The weights are MFMA operands read straight from the bf16
// [out][in] copies (L2-resident).  Fragment mapping, k order (32 at a time), epilogue rounding
// and the output layer's reduction are those of k_embed, k_gemm_nt<EPI_FWD> and k_head_fwd,
// so y is bitwise the forward pass's; no per-row buffers, so any row count.
constexpr int kPM = 64;        // rows per block
constexpr int kPLd = kH + 8;   // LDS row stride (bf16): 16-byte aligned rows
__global__ void __launch_bounds__(256) k_mlp_infer(int n, int nsets, const float* __restrict__ par,
                                                   const int32_t* __restrict__ set, const float* __restrict__ q,
                                                   const Features ft, const bf16* __restrict__ wb,
                                                   const float* __restrict__ p, float* __restrict__ y) {
    __shared__ __attribute__((aligned(16))) bf16 act[kPM * kPLd];
    const int r0 = blockIdx.x * kPM;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    {
        // embedding, four threads per row (k_embed's values, written to LDS one by one)
        const int row = t >> 2, h = t & 3;
        const int i = r0 + row;
        bf16* dst = act + row * kPLd;
        float v[kNP];
#pragma unroll
        for (int k = 0; k < kNP; ++k) v[k] = 0.0f;
        const bool live = i < n;
        float qx = 0.0f, qy = 0.0f;
        if (live) {
            const float* pp = par + (size_t)min(max(set[i], 0), nsets - 1) * kNP;  // clamped like k_embed
#pragma unroll
            for (int k = 0; k < kNP; ++k) v[k] = pp[k];
            qx = q[2 * (size_t)i];
            qy = q[2 * (size_t)i + 1];
        }
        const float qa = sqrtf(qx * qx + qy * qy);
        if (h == 0)
#pragma unroll
            for (int k = 0; k < kNP; ++k) dst[k] = (bf16)v[k];
#pragma unroll
        for (int f0 = 0; f0 < kNF / 4; ++f0) {
            const int f = h * (kNF / 4) + f0;
            float sn = 0.0f, cs = 0.0f;
            if (live) {
                const int j = ft.pidx[f];
                float pv = 1.0f;
#pragma unroll
                for (int k = 0; k < kNP; ++k) pv = (k == j) ? ft.off[k] + ft.sc[k] * v[k] : pv;
                sincosf(6.283185307f * (ft.f[f][0] * qx + ft.f[f][1] * qy) + ft.f[f][2] * qa * pv, &sn, &cs);
            }
            dst[kNP + f] = (bf16)sn;
            dst[kNP + kNF + f] = (bf16)cs;
        }
        for (int k = kUsed + h; k < kIn; k += 4) dst[k] = (bf16)0.0f;
    }
    __syncthreads();
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    const int wn = wave * 64;  // this wave's 64 output columns
    for (int l = 0; l < kL; ++l) {
        const int K = in_of(l);
        const bf16* W = wb + (l == 0 ? 0 : (int64_t)kH * kIn + (int64_t)(l - 1) * kH * kH);
        f32x4 acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < K; ks += 32) {
            bf16x8 af[4], bfg[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) bfg[j] = *(const bf16x8*)&W[(size_t)(wn + j * 16 + fr) * K + ks + fk];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = *(const bf16x8*)&act[(i * 16 + fr) * kPLd + ks + fk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfg[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();  // every wave has read the layer's input
        const float* b = p + b_off(l);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float bj = b[wn + j * 16 + fr];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float z = acc[i][j][r] + bj;
                    act[(i * 16 + (lane >> 4) * 4 + r) * kPLd + wn + j * 16 + fr] = (bf16)(z * sigm(z));
                }
        }
        __syncthreads();
    }
    const float* w = p + kOutW;
    for (int rr = wave; rr < kPM; rr += 4) {
        const bf16* row = act + rr * kPLd + lane * 4;
        float sum = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) sum = fmaf((float)row[c], w[lane * 4 + c], sum);
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
        if (lane == 0 && r0 + rr < n) y[r0 + rr] = sum + p[kOutB];
    }
}

// Loss 1/n sum (y - t)^2 and the output layer's backward for a block of kHeadRows rows:
// dy = 2 (y - t) / n (0 for rows >= n) into dy_out, dZ4[m][c] = dy * w[c] * silu'(Z4[m][c])
// written in both layouts, and per-block partials of db = sum dy and of the loss.
__global__ void __launch_bounds__(256) k_head_bwd(int n, int mpad, const float* __restrict__ y,
                                                  const float* __restrict__ tgt, const bf16* __restrict__ Z,
                                                  const float* __restrict__ w, float* __restrict__ dy_out,
                                                  bf16* __restrict__ dZ, bf16* __restrict__ dZT,
                                                  float* __restrict__ part_b) {
    __shared__ bf16 tile[kHeadRows][kH + 8];
    __shared__ float s_dy[kHeadRows];
    const int m0 = blockIdx.x * kHeadRows, t = threadIdx.x;
    if (t < kHeadRows) {
        const int m = m0 + t;
        const float dy = (m < n) ? 2.0f * (y[m] - tgt[m]) / (float)n : 0.0f;
        s_dy[t] = dy;
        dy_out[m] = dy;
    }
    __syncthreads();
    // thread t owns column c = t
    const int c = t;
    const float wc = w[c];
    for (int r = 0; r < kHeadRows; ++r) {
        const size_t o = (size_t)(m0 + r) * kH + c;
        const bf16 dz = (bf16)(s_dy[r] * wc * silu_grad((float)Z[o]));
        tile[r][c] = dz;
        dZ[o] = dz;
    }
    // column c of the tile is 64 consecutive rows of dZ^T
    bf16* dst = dZT + (size_t)c * mpad + m0;
#pragma unroll
    for (int r = 0; r < kHeadRows; r += 8) {
        bf16x8 pk;
#pragma unroll
        for (int j = 0; j < 8; ++j) pk[j] = tile[r + j][c];
        *(bf16x8*)(dst + r) = pk;
    }
    if (t == 0) {
        float db = 0.0f, loss = 0.0f;
        for (int r = 0; r < kHeadRows; ++r) {
            db += s_dy[r];
            const int m = m0 + r;
            if (m < n) loss += (y[m] - tgt[m]) * (y[m] - tgt[m]);
        }
        part_b[2 * blockIdx.x] = db;
        part_b[2 * blockIdx.x + 1] = loss;
    }
}

__device__ __forceinline__ float block_sum(float v, float* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const float s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

// Row reductions over the batch for rows of a [rows][mpad] bf16 matrix, in segments of
// kSeg columns: part[r * S + s] = sum over the segment of A[r][m] * (w ? w[m] : 1).
// Grid (rows, S = ceil(mpad / kSeg)); each thread takes one 8-column chunk.
constexpr int kSeg = 256 * 8;
__global__ void __launch_bounds__(256) k_rowdot(int mpad, const bf16* __restrict__ A, const float* __restrict__ w,
                                                float* __restrict__ part) {
    __shared__ float red[4];
    const int r = blockIdx.x, sg = blockIdx.y;
    const int m = sg * kSeg + threadIdx.x * 8;
    float acc = 0.0f;
    if (m < mpad) {  // mpad is a multiple of 128: chunks lie wholly inside or outside
        const bf16x8 v = *(const bf16x8*)(A + (size_t)r * mpad + m);
        if (w) {
            const float4 w0 = *(const float4*)(w + m), w1 = *(const float4*)(w + m + 4);
            acc = (float)v[0] * w0.x + (float)v[1] * w0.y + (float)v[2] * w0.z + (float)v[3] * w0.w +
                  (float)v[4] * w1.x + (float)v[5] * w1.y + (float)v[6] * w1.z + (float)v[7] * w1.w;
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) acc += (float)v[j];
        }
    }
    const float s = block_sum(acc, red);
    if (threadIdx.x == 0) part[(size_t)r * gridDim.y + sg] = s;
}
// out[r] = sum over s < S of part[r * S + s] (fixed order)
__global__ void k_rowdot_finish(int rows, int S, const float* __restrict__ part, float* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    float s = 0.0f;
    for (int k = 0; k < S; ++k) s += part[(size_t)r * S + k];
    out[r] = s;
}

// Sums of the output layer's block partials: g_b = sum db, loss_sum = sum loss (one block).
__global__ void __launch_bounds__(256) k_head_finish(int nb, const float* __restrict__ part_b, float* __restrict__ g_b,
                                                     float* __restrict__ loss_sum) {
    __shared__ float red[4];
    float db = 0.0f, ls = 0.0f;
    for (int k = threadIdx.x; k < nb; k += 256) {
        db += part_b[2 * k];
        ls += part_b[2 * k + 1];
    }
    db = block_sum(db, red);
    ls = block_sum(ls, red);
    if (threadIdx.x == 0) {
        *g_b = db;
        *loss_sum = ls;
    }
}

// g[c] = sum over s < ns of part[s*ld + c] for the weight-gradient splits, 4 columns per
// thread (fixed order: deterministic).
__global__ void __launch_bounds__(256) k_sum_parts(int n4, int ns, int ld4, const float4* __restrict__ part,
                                                   float4* __restrict__ g) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n4) return;
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 8
    for (int k = 0; k < ns; ++k) {
        const float4 v = part[(size_t)k * ld4 + c];
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
    }
    g[c] = s;
}

// Adam on the FP32 master parameters, then the bf16 copies of the hidden weights in both
// orientations (W [out][in] and W^T [in][out]).
__global__ void __launch_bounds__(256) k_adam(int64_t np, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m1, float* __restrict__ m2, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr * (a * c1) / (sqrtf(v * c2) + eps);
}
__global__ void __launch_bounds__(256) k_cast_weights(const float* __restrict__ p, bf16* __restrict__ wb,
                                                      bf16* __restrict__ wtb) {
    // flat index over all hidden weights, layer by layer
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (int l = 0; l < kL; ++l) {
        const int64_t cnt = (int64_t)kH * in_of(l);
        if (i < cnt) {
            const int o = (int)(i / in_of(l)), k = (int)(i % in_of(l));
            const bf16 v = (bf16)p[w_off(l) + i];
            const int64_t base = l == 0 ? 0 : (int64_t)kH * kIn + (int64_t)(l - 1) * kH * kH;
            wb[base + i] = v;
            wtb[base + (int64_t)k * kH + o] = v;
            return;
        }
        i -= cnt;
    }
}

}  // namespace csm

// ===========================================================================
// C ABI
// ===========================================================================
using namespace csm;

struct csx_mlp {
    int device = 0;
    int max_batch = 0;  // rows, a multiple of 128
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    Features feat;
    // parameters
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    bf16 *wb = nullptr, *wtb = nullptr;  // hidden weights, [out][in] and [in][out]
    int64_t adam_t = 0;
    // per-row buffers
    bf16 *X = nullptr, *XT = nullptr;
    bf16 *Z[kL] = {}, *H[kL] = {}, *HT[kL] = {};
    bf16 *dZ[2] = {}, *dZT[2] = {};
    float *y = nullptr, *dy = nullptr, *part = nullptr, *rpart = nullptr, *hpart_b = nullptr, *loss = nullptr;
    int splits = 0;
};

namespace {

int fail(csx_mlp* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    return code;
}
#define MCHK(m, call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(m, CSX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

const bf16* wb_layer(const csx_mlp* m, int l) {
    return m->wb + (l == 0 ? 0 : (int64_t)kH * kIn + (int64_t)(l - 1) * kH * kH);
}
const bf16* wtb_layer(const csx_mlp* m, int l) {
    return m->wtb + (l == 0 ? 0 : (int64_t)kH * kIn + (int64_t)(l - 1) * kH * kH);
}

// splitmix64 -> uniform doubles, Box-Muller normals (host-side initialisation only)
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * 0x1.0p-53; }
};

int cast_weights(csx_mlp* m) {
    const int64_t nw = (int64_t)kH * kIn + (int64_t)(kL - 1) * kH * kH;
    hipLaunchKernelGGL(k_cast_weights, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->stream, m->p, m->wb, m->wtb);
    MCHK(m, hipGetLastError());
    return CSX_OK;
}

int check_rows(csx_mlp* m, int n, int nsets) {
    if (n < 1 || n > m->max_batch) return fail(m, CSX_E_ARG, "row count must be in [1, max_batch]");
    if (nsets < 1) return fail(m, CSX_E_ARG, "need at least one parameter set");
    return CSX_OK;
}

// Forward pass over n rows (padded to a multiple of 128) into m->y.
int forward(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q) {
    const int mpad = (n + kBM - 1) / kBM * kBM;
    hipLaunchKernelGGL(k_embed, dim3((mpad + 255) / 256), dim3(256), 0, m->stream, n, mpad, nsets, d_par, d_set, d_q,
                       m->feat, m->X, m->XT);
    MCHK(m, hipGetLastError());
    for (int l = 0; l < kL; ++l) {
        EpiArgs ep{};
        ep.bias = m->p + b_off(l);
        ep.z = m->Z[l];
        ep.h = m->H[l];
        ep.ht = m->HT[l];
        ep.ldo = kH;
        ep.ldt = mpad;
        const bf16* A = l == 0 ? m->X : m->H[l - 1];
        const int K = in_of(l);
        hipLaunchKernelGGL(k_gemm_nt<EPI_FWD>, dim3(mpad / kBM, kH / kBN, 1), dim3(256), 0, m->stream, A, K,
                           wb_layer(m, l), K, K, K, ep);
        MCHK(m, hipGetLastError());
    }
    hipLaunchKernelGGL(k_head_fwd, dim3(mpad / 4), dim3(256), 0, m->stream, mpad, m->H[kL - 1], m->p + kOutW,
                       m->p + kOutB, m->y);
    MCHK(m, hipGetLastError());
    return CSX_OK;
}

}  // namespace

extern "C" {

int64_t csx_mlp_nparams(void) { return kNParams; }

int csx_mlp_create(int device, int max_batch, uint64_t seed, csx_mlp** out) {
    if (!out || max_batch < 1 || max_batch > (1 << 24)) return CSX_E_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return CSX_E_HIP;
    csx_mlp* m = new csx_mlp();
    m->device = device;
    m->max_batch = (max_batch + kBM - 1) / kBM * kBM;
    const size_t M = (size_t)m->max_batch;
    // weight-gradient splits over the rows: enough blocks to fill the chip, >= 512 rows each
    m->splits = (int)std::max<size_t>(1, std::min<size_t>(64, M / 512));
    auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, bytes) == hipSuccess; };
    bool ok = hipStreamCreate(&m->stream) == hipSuccess;
    m->own_stream = ok;
    const int64_t nw = (int64_t)kH * kIn + (int64_t)(kL - 1) * kH * kH;
    ok = ok && alloc((void**)&m->p, kNParams * 4) && alloc((void**)&m->g, kNParams * 4) &&
         alloc((void**)&m->m1, kNParams * 4) && alloc((void**)&m->m2, kNParams * 4) &&
         alloc((void**)&m->wb, nw * 2) && alloc((void**)&m->wtb, nw * 2) &&
         alloc((void**)&m->X, M * kIn * 2) && alloc((void**)&m->XT, M * kIn * 2) && alloc((void**)&m->y, M * 4) &&
         alloc((void**)&m->dy, M * 4) && alloc((void**)&m->part, (size_t)m->splits * kH * kH * 4) &&
         alloc((void**)&m->rpart, (size_t)kH * ((M + kSeg - 1) / kSeg) * 4) &&
         alloc((void**)&m->hpart_b, (M / kHeadRows) * 2 * 4) && alloc((void**)&m->loss, 4);
    for (int l = 0; l < kL && ok; ++l)
        ok = alloc((void**)&m->Z[l], M * kH * 2) && alloc((void**)&m->H[l], M * kH * 2) &&
             alloc((void**)&m->HT[l], M * kH * 2);
    for (int k = 0; k < 2 && ok; ++k) ok = alloc((void**)&m->dZ[k], M * kH * 2) && alloc((void**)&m->dZT[k], M * kH * 2);
    if (!ok) {
        csx_mlp_destroy(m);
        return CSX_E_NOMEM;
    }
    // Initialisation: He-uniform hidden weights (fan-in = used inputs), zero biases, a
    // small output layer; Fourier frequencies with log-uniform magnitude in [0.5, 16]
    // cycles per nm^-1 and uniform directions.
    std::vector<float> hp(kNParams, 0.0f);
    Rng r{seed};
    for (int l = 0; l < kL; ++l) {
        const int fan = l == 0 ? kUsed : kH;
        const double bound = std::sqrt(6.0 / fan);
        for (int o = 0; o < kH; ++o)
            for (int k = 0; k < in_of(l); ++k)
                hp[w_off(l) + (int64_t)o * in_of(l) + k] = (l == 0 && k >= kUsed) ? 0.0f : (float)((2 * r.uni() - 1) * bound);
    }
    for (int c = 0; c < kH; ++c) hp[kOutW + c] = (float)((2 * r.uni() - 1) * std::sqrt(3.0 / kH));
    // default phases: 2*pi*F.q with log-uniform magnitudes in [0.5, 16] and uniform directions
    std::memset(&m->feat, 0, sizeof m->feat);
    for (int f = 0; f < kNF; ++f) {
        const double mag = 0.5 * std::pow(32.0, r.uni()), ang = 2 * M_PI * r.uni();
        m->feat.f[f][0] = (float)(mag * std::cos(ang));
        m->feat.f[f][1] = (float)(mag * std::sin(ang));
        m->feat.pidx[f] = -1;
    }
    for (int j = 0; j < kNP; ++j) m->feat.sc[j] = 1.0f;
    ok = hipMemcpyAsync(m->p, hp.data(), kNParams * 4, hipMemcpyHostToDevice, m->stream) == hipSuccess &&
         hipMemsetAsync(m->m1, 0, kNParams * 4, m->stream) == hipSuccess &&
         hipMemsetAsync(m->m2, 0, kNParams * 4, m->stream) == hipSuccess &&
         hipMemsetAsync(m->g, 0, kNParams * 4, m->stream) == hipSuccess && cast_weights(m) == CSX_OK &&
         hipStreamSynchronize(m->stream) == hipSuccess;
    if (!ok) {
        csx_mlp_destroy(m);
        return CSX_E_HIP;
    }
    *out = m;
    return CSX_OK;
}

void csx_mlp_destroy(csx_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* q : {(void*)m->p, (void*)m->g, (void*)m->m1, (void*)m->m2, (void*)m->wb, (void*)m->wtb,
                    (void*)m->X, (void*)m->XT, (void*)m->y, (void*)m->dy, (void*)m->part, (void*)m->rpart,
                    (void*)m->hpart_b, (void*)m->loss})
        if (q) (void)hipFree(q);
    for (int l = 0; l < kL; ++l)
        for (void* q : {(void*)m->Z[l], (void*)m->H[l], (void*)m->HT[l]})
            if (q) (void)hipFree(q);
    for (int k = 0; k < 2; ++k)
        for (void* q : {(void*)m->dZ[k], (void*)m->dZT[k]})
            if (q) (void)hipFree(q);
    if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int csx_mlp_last_error(const csx_mlp* m, char* buf, int len) {
    if (!m || !buf || len <= 0) return CSX_E_ARG;
    std::snprintf(buf, (size_t)len, "%s", m->err.c_str());
    return CSX_OK;
}

int csx_mlp_set_stream(csx_mlp* m, void* stream) {
    if (!m) return CSX_E_ARG;
    MCHK(m, hipStreamSynchronize(m->stream));
    if (m->own_stream) (void)hipStreamDestroy(m->stream);
    m->own_stream = false;
    m->stream = (hipStream_t)stream;
    return CSX_OK;
}

int csx_mlp_get_params(csx_mlp* m, float* h_params) {
    if (!m || !h_params) return CSX_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipMemcpyAsync(h_params, m->p, kNParams * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return CSX_OK;
}

int csx_mlp_get_features(const csx_mlp* m, float* h_freq, int32_t* h_pidx, float* h_affine) {
    if (!m) return CSX_E_ARG;
    for (int f = 0; f < kNF; ++f) {
        if (h_freq) for (int c = 0; c < 3; ++c) h_freq[3 * f + c] = m->feat.f[f][c];
        if (h_pidx) h_pidx[f] = m->feat.pidx[f];
    }
    if (h_affine)
        for (int j = 0; j < kNP; ++j) {
            h_affine[2 * j] = m->feat.off[j];
            h_affine[2 * j + 1] = m->feat.sc[j];
        }
    return CSX_OK;
}

int csx_mlp_set_features(csx_mlp* m, const float* h_freq, const int32_t* h_pidx, const float* h_affine) {
    if (!m || !h_freq || !h_pidx || !h_affine) return CSX_E_ARG;
    for (int f = 0; f < kNF; ++f)
        if (h_pidx[f] < -1 || h_pidx[f] >= kNP) return fail(m, CSX_E_ARG, "parameter index of a phase out of range");
    MCHK(m, hipStreamSynchronize(m->stream));  // launches already queued keep their copy
    for (int f = 0; f < kNF; ++f) {
        for (int c = 0; c < 3; ++c) m->feat.f[f][c] = h_freq[3 * f + c];
        m->feat.pidx[f] = h_pidx[f];
    }
    for (int j = 0; j < kNP; ++j) {
        m->feat.off[j] = h_affine[2 * j];
        m->feat.sc[j] = h_affine[2 * j + 1];
    }
    return CSX_OK;
}

int csx_mlp_set_params(csx_mlp* m, const float* h_params) {
    if (!m) return CSX_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    if (h_params) {
        // the padded input columns of the first layer stay zero
        std::vector<float> hp(h_params, h_params + kNParams);
        for (int o = 0; o < kH; ++o)
            for (int k = kUsed; k < kIn; ++k) hp[w_off(0) + (int64_t)o * kIn + k] = 0.0f;
        MCHK(m, hipMemcpyAsync(m->p, hp.data(), kNParams * 4, hipMemcpyHostToDevice, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
        int rc = cast_weights(m);
        if (rc) return rc;
    }
    m->adam_t = 0;
    MCHK(m, hipMemsetAsync(m->m1, 0, kNParams * 4, m->stream));
    MCHK(m, hipMemsetAsync(m->m2, 0, kNParams * 4, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return CSX_OK;
}

int csx_mlp_forward(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q,
                    float* d_y) {
    if (!m || !d_par || !d_set || !d_q || !d_y) return CSX_E_ARG;
    int rc = check_rows(m, n, nsets);
    if (rc) return rc;
    MCHK(m, hipSetDevice(m->device));
    rc = forward(m, n, nsets, d_par, d_set, d_q);
    if (rc) return rc;
    MCHK(m, hipMemcpyAsync(d_y, m->y, (size_t)n * 4, hipMemcpyDeviceToDevice, m->stream));
    return CSX_OK;
}

int csx_mlp_predict(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q,
                    float* d_y) {
    if (!m || !d_par || !d_set || !d_q || !d_y || n < 1 || nsets < 1) return CSX_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + kPM - 1) / kPM)), dim3(256), 0, m->stream, n, nsets, d_par,
                       d_set, d_q, m->feat, m->wb, m->p, d_y);
    MCHK(m, hipGetLastError());
    return CSX_OK;
}

int csx_mlp_gradient(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q,
                     const float* d_t, float* h_loss) {
    if (!m || !d_par || !d_set || !d_q || !d_t) return CSX_E_ARG;
    int rc = check_rows(m, n, nsets);
    if (rc) return rc;
    MCHK(m, hipSetDevice(m->device));
    rc = forward(m, n, nsets, d_par, d_set, d_q);
    if (rc) return rc;
    const int mpad = (n + kBM - 1) / kBM * kBM;
    const int nhb = mpad / kHeadRows;
    const int nseg = (mpad + kSeg - 1) / kSeg;
    // output layer: dy, dZ of the last hidden layer, db and the loss; dw = H4^T . dy
    hipLaunchKernelGGL(k_head_bwd, dim3(nhb), dim3(256), 0, m->stream, n, mpad, m->y, d_t, m->Z[kL - 1],
                       m->p + kOutW, m->dy, m->dZ[0], m->dZT[0], m->hpart_b);
    MCHK(m, hipGetLastError());
    hipLaunchKernelGGL(k_head_finish, dim3(1), dim3(256), 0, m->stream, nhb, m->hpart_b, m->g + kOutB, m->loss);
    hipLaunchKernelGGL(k_rowdot, dim3(kH, nseg), dim3(256), 0, m->stream, mpad, m->HT[kL - 1], m->dy, m->rpart);
    hipLaunchKernelGGL(k_rowdot_finish, dim3(1), dim3(kH), 0, m->stream, kH, nseg, m->rpart, m->g + kOutW);
    MCHK(m, hipGetLastError());
    // hidden layers, top down; dZ of layer l is in buffer cur
    int cur = 0;
    for (int l = kL - 1; l >= 0; --l) {
        const int K = in_of(l);
        // weight gradient dW_l [kH][K] = dZ_l^T . H_{l-1}, split over the rows
        const int kchunk = ((mpad / m->splits + kBK - 1) / kBK) * kBK;
        const int ns = (mpad + kchunk - 1) / kchunk;
        EpiArgs pe{};
        pe.part = m->part;
        pe.ldo = K;
        const bf16* BT = l == 0 ? m->XT : m->HT[l - 1];
        hipLaunchKernelGGL(k_gemm_nt<EPI_PART>, dim3(kH / kBM, K / kBN, ns), dim3(256), 0, m->stream, m->dZT[cur], mpad,
                           BT, mpad, mpad, kchunk, pe);
        MCHK(m, hipGetLastError());
        const int n4 = kH * K / 4;
        hipLaunchKernelGGL(k_sum_parts, dim3((n4 + 255) / 256), dim3(256), 0, m->stream, n4, ns, n4,
                           (const float4*)m->part, (float4*)(m->g + w_off(l)));
        // bias gradient: row sums of dZ^T
        hipLaunchKernelGGL(k_rowdot, dim3(kH, nseg), dim3(256), 0, m->stream, mpad, m->dZT[cur], nullptr, m->rpart);
        hipLaunchKernelGGL(k_rowdot_finish, dim3(1), dim3(kH), 0, m->stream, kH, nseg, m->rpart, m->g + b_off(l));
        MCHK(m, hipGetLastError());
        if (l == 0) break;
        // dH_{l-1} = dZ_l . W_l, then dZ_{l-1} = dH_{l-1} * silu'(Z_{l-1})
        EpiArgs be{};
        be.zprev = m->Z[l - 1];
        be.dz = m->dZ[cur ^ 1];
        be.dzt = m->dZT[cur ^ 1];
        be.ldo = kH;
        be.ldt = mpad;
        hipLaunchKernelGGL(k_gemm_nt<EPI_BWD>, dim3(mpad / kBM, kH / kBN, 1), dim3(256), 0, m->stream, m->dZ[cur], kH,
                           wtb_layer(m, l), kH, kH, kH, be);
        MCHK(m, hipGetLastError());
        cur ^= 1;
    }
    if (h_loss) {
        float ls = 0.0f;
        MCHK(m, hipMemcpyAsync(&ls, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
        *h_loss = ls / (float)n;
    }
    return CSX_OK;
}

int csx_mlp_get_grads(csx_mlp* m, float* h_grads) {
    if (!m || !h_grads) return CSX_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipMemcpyAsync(h_grads, m->g, kNParams * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return CSX_OK;
}

int csx_mlp_adam(csx_mlp* m, float lr, float b1, float b2, float eps) {
    if (!m || !(lr >= 0) || !(b1 >= 0 && b1 < 1) || !(b2 >= 0 && b2 < 1)) return CSX_E_ARG;
    MCHK(m, hipSetDevice(m->device));
    ++m->adam_t;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)m->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)m->adam_t)));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((kNParams + 255) / 256)), dim3(256), 0, m->stream, kNParams, m->p, m->g,
                       m->m1, m->m2, lr, b1, b2, eps, c1, c2);
    MCHK(m, hipGetLastError());
    return cast_weights(m);
}

int csx_mlp_train_step(csx_mlp* m, int n, int nsets, const float* d_par, const int32_t* d_set, const float* d_q,
                       const float* d_t, float lr, float* h_loss) {
    int rc = csx_mlp_gradient(m, n, nsets, d_par, d_set, d_q, d_t, h_loss);
    if (rc) return rc;
    return csx_mlp_adam(m, lr, 0.9f, 0.999f, 1e-8f);
}

int csx_mlp_synchronize(csx_mlp* m) {
    if (!m) return CSX_E_ARG;
    MCHK(m, hipStreamSynchronize(m->stream));
    return CSX_OK;
}

}  // extern "C"
