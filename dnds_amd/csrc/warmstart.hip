// Learned warm start of the VR reconstruction (gfx950).  No counterpart in
// the reference: a small MLP maps a cell's stencil of cell means and its
// geometry moments to an initial guess of its 9 x 4 rec coefficients; plain
// Jacobi sweeps then iterate to the same tolerance, so the method decides the
// result and remains the fallback.
//
// Network (58 -> 128 -> 128 -> 36, GELU): inputs are the mean differences
// u_j - u_i of 12 stencil cells (per variable divided by s_v, the largest of
// them but at least 1e-3 max_v |u_i,v|), the 9 base moments of the cell and
// 0.5 log(volume); outputs are the
// rec coefficients divided by s_v (the reconstruction is linear in the means,
// so this scaling makes the map scale-free).  The stencil is the first 12
// cells met by a breadth-first walk over faces (cell2face order); missing
// cells repeat the cell itself.
//
// Training and inference use hand-written kernels on gfx950's matrix cores
// (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation): a GEMM with
// fused epilogues (bias, GELU, GELU backward) for training, column sums for
// the bias gradients, the MSE loss and Adam, all FP32 otherwise; and one
// fused kernel for inference that keeps the three layers' BF16 weights in
// LDS and the hidden activations on chip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "ctx.hpp"

using namespace dndsmi_dev;
using namespace dndsmi_host;

namespace {

constexpr int kIn = 58, kHid = 128, kOut = 36, kSten = 12;
// parameter block: W1 [58][128], b1 [128], W2 [128][128], b2 [128], W3 [128][36], b3 [36]
constexpr int64_t oW1 = 0, ob1 = oW1 + kIn * kHid, oW2 = ob1 + kHid, ob2 = oW2 + kHid * kHid, oW3 = ob2 + kHid,
                  ob3 = oW3 + kHid * kOut, kNParam = ob3 + kOut;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ inline unsigned short to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_GRAD = 3 };

// C[M][N] = op(A) op(B) with BF16 operands and FP32 accumulation.
// A(m, k) = transA ? A[k lda + m] : A[m lda + k]; B(k, n) likewise.
// 256 threads; a 64 x 64 tile per workgroup, a 32 x 32 quarter per wave as
// 2 x 2 MFMA 16x16x32 blocks; K in steps of 32 staged through LDS as BF16.
// Operand layout of v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][k] and
// B[k][l % 16] for the eight k = 8 (l / 16) + 0..7 (a product is the same
// under any assignment of k that A and B share); it receives
// D[4 (l / 16) + 0..3][l % 16].
//
// Split K (EPI_STORE only): block z sums k in [z kchunk, (z + 1) kchunk) and
// writes its partial product to C + z M ldcm; k_splitk_sum adds the partials
// in a fixed order.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int32_t M, int32_t N, int32_t K, const float *__restrict__ A,
                                              int64_t lda, int32_t transA, const float *__restrict__ B, int64_t ldb,
                                              int32_t transB, float *__restrict__ C, int64_t ldcm,
                                              const float *__restrict__ bias, float *__restrict__ C2,
                                              const float *__restrict__ Zin, int32_t kchunk) {
    // As: rows m, k contiguous; Bs: columns n, k contiguous (so that a lane's
    // eight k are one 16-byte LDS read); rows padded by 8 against bank conflicts
    __shared__ __attribute__((aligned(16))) unsigned short As[64][32 + 8];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[64][32 + 8];
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int64_t m0 = int64_t(blockIdx.x) * 64, n0 = int64_t(blockIdx.y) * 64;
    const int wm = 32 * (wave / 2), wn = 32 * (wave % 2);
    const int32_t kbeg = int32_t(blockIdx.z) * kchunk, kend = min(K, kbeg + kchunk);
    if (EPI == EPI_STORE)
        C += int64_t(blockIdx.z) * M * ldcm;
    floatx4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
            acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int32_t k0 = kbeg; k0 < kend; k0 += 32) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int e = tid + 256 * q; // 0 .. 2047
            {
                // transA reads A[k lda + m]: consecutive threads take consecutive m
                const int r = transA ? e % 64 : e / 32, kk = transA ? e / 64 : e % 32;
                const int64_t m = m0 + r, k = k0 + kk;
                float v = 0.f;
                if (m < M && k < kend)
                    v = transA ? A[k * lda + m] : A[m * lda + k];
                As[r][kk] = to_bf16(v);
            }
            {
                const int c = transB ? e / 32 : e % 64, kk = transB ? e % 32 : e / 64;
                const int64_t n = n0 + c, k = k0 + kk;
                float v = 0.f;
                if (n < N && k < kend)
                    v = transB ? B[n * ldb + k] : B[k * ldb + n];
                Bs[c][kk] = to_bf16(v);
            }
        }
        __syncthreads();
        bf16x8 af[2], bfr[2];
        const int kb = 8 * (lane / 16);
#pragma unroll
        for (int a = 0; a < 2; a++)
            af[a] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const shortx8 *>(&As[wm + 16 * a + lane % 16][kb]));
#pragma unroll
        for (int b = 0; b < 2; b++)
            bfr[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const shortx8 *>(&Bs[wn + 16 * b + lane % 16][kb]));
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t m = m0 + wm + 16 * a + 4 * (lane / 16) + j;
                const int64_t n = n0 + wn + 16 * b + lane % 16;
                if (m >= M || n >= N)
                    continue;
                float v = acc[a][b][j];
                const int64_t o = m * ldcm + n;
                if (EPI == EPI_STORE) {
                    C[o] = v;
                } else if (EPI == EPI_BIAS) {
                    C[o] = v + bias[n];
                } else if (EPI == EPI_BIAS_GELU) {
                    const float z = v + bias[n];
                    C[o] = z;
                    C2[o] = gelu(z);
                } else {
                    C[o] = v * gelu_grad(Zin[o]);
                }
            }
}

// Inference of the whole network, Y = W3^T g(W2^T g(W1^T x + b1) + b2) + b3
// per row (g = GELU), in one kernel.  Each workgroup (8 waves) stages the
// three weight matrices once, transposed to [n][k] and rounded to BF16, in
// LDS (about 66 KB), then walks its rows 16 at a time per wave: layer 1 reads
// the row's features from memory, every hidden layer's output (bias and GELU
// applied in FP32, then rounded to BF16, as the operand of the next GEMM
// would be) goes through the wave's own 16 x 128 LDS tile, and only the 36
// outputs reach memory.  The arithmetic is that of the three k_gemm launches
// (BF16 operands, FP32 accumulation, FP32 epilogues) without the hidden
// activations' round trips through HBM.
constexpr int kInPad = 64, kOutPad = 48, kLdW1 = kInPad + 8, kLdH = kHid + 8, kInferWaves = 16;

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(int32_t rows, const float *__restrict__ X,
                                                                  const float *__restrict__ P,
                                                                  float *__restrict__ Y) {
    __shared__ __attribute__((aligned(16))) unsigned short W1t[kHid][kLdW1];
    __shared__ __attribute__((aligned(16))) unsigned short W2t[kHid][kLdH];
    __shared__ __attribute__((aligned(16))) unsigned short W3t[kOutPad][kLdH];
    __shared__ float bs[2 * kHid + kOutPad];
    __shared__ __attribute__((aligned(16))) unsigned short H[kInferWaves][16][kLdH];
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    for (int e = tid; e < kHid * kInPad; e += 64 * kInferWaves) {
        const int n = e % kHid, k = e / kHid; // consecutive threads: consecutive n of one row of W1
        W1t[n][k] = to_bf16(k < kIn ? P[oW1 + k * kHid + n] : 0.f);
    }
    for (int e = tid; e < kHid * kHid; e += 64 * kInferWaves) {
        const int n = e % kHid, k = e / kHid;
        W2t[n][k] = to_bf16(P[oW2 + k * kHid + n]);
    }
    for (int e = tid; e < kOutPad * kHid; e += 64 * kInferWaves) {
        const int n = e % kOutPad, k = e / kOutPad;
        W3t[n][k] = to_bf16(n < kOut ? P[oW3 + k * kOut + n] : 0.f);
    }
    for (int e = tid; e < 2 * kHid + kOutPad; e += 64 * kInferWaves)
        bs[e] = e < kHid ? P[ob1 + e] : (e < 2 * kHid ? P[ob2 + e - kHid] : (e - 2 * kHid < kOut ? P[ob3 + e - 2 * kHid] : 0.f));
    __syncthreads();
    const int kb = 8 * (lane / 16), lr = lane % 16;
    const int64_t ntile = (int64_t(rows) + 15) / 16;
    // the loop bound depends on the workgroup only, so every wave reaches
    // every __syncthreads (a wave past the last tile computes on zeros and
    // stores nothing)
    float xc[kInPad / 32][8];
    {
        const int64_t row = (int64_t(blockIdx.x) * kInferWaves + wave) * 16 + lr;
#pragma unroll
        for (int s2 = 0; s2 < kInPad / 32; s2++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int k = 32 * s2 + kb + e;
                xc[s2][e] = row < rows && k < kIn ? X[row * kIn + k] : 0.f;
            }
    }
    for (int64_t t0 = int64_t(blockIdx.x) * kInferWaves; t0 < ntile; t0 += int64_t(gridDim.x) * kInferWaves) {
        const int64_t r0 = (t0 + wave) * 16;
        // layer 1: A operand straight from the features (58 per row, zero
        // padded); the next tile's features are loaded while this one runs
        floatx4 acc[8];
#pragma unroll
        for (int b = 0; b < 8; b++)
            acc[b] = floatx4{0.f, 0.f, 0.f, 0.f};
        float xn[kInPad / 32][8];
        {
            const int64_t nrow = r0 + int64_t(gridDim.x) * kInferWaves * 16 + lr;
#pragma unroll
            for (int s2 = 0; s2 < kInPad / 32; s2++)
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int k = 32 * s2 + kb + e;
                    xn[s2][e] = nrow < rows && k < kIn ? X[nrow * kIn + k] : 0.f;
                }
        }
#pragma unroll
        for (int s2 = 0; s2 < kInPad / 32; s2++) {
            shortx8 av;
#pragma unroll
            for (int e = 0; e < 8; e++)
                av[e] = short(to_bf16(xc[s2][e]));
            const bf16x8 a = __builtin_bit_cast(bf16x8, av);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const bf16x8 w = __builtin_bit_cast(bf16x8, *reinterpret_cast<const shortx8 *>(&W1t[16 * b + lr][32 * s2 + kb]));
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[b], 0, 0, 0);
            }
        }
#pragma unroll
        for (int b = 0; b < 8; b++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                H[wave][4 * (lane / 16) + j][16 * b + lr] = to_bf16(gelu(acc[b][j] + bs[16 * b + lr]));
        __syncthreads();
        // layer 2
#pragma unroll
        for (int b = 0; b < 8; b++)
            acc[b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s2 = 0; s2 < kHid / 32; s2++) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const shortx8 *>(&H[wave][lr][32 * s2 + kb]));
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const bf16x8 w = __builtin_bit_cast(bf16x8, *reinterpret_cast<const shortx8 *>(&W2t[16 * b + lr][32 * s2 + kb]));
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[b], 0, 0, 0);
            }
        }
        __syncthreads(); // every wave has read its layer-1 tile
#pragma unroll
        for (int b = 0; b < 8; b++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                H[wave][4 * (lane / 16) + j][16 * b + lr] = to_bf16(gelu(acc[b][j] + bs[kHid + 16 * b + lr]));
        __syncthreads();
        // layer 3
        floatx4 out[3];
#pragma unroll
        for (int b = 0; b < 3; b++)
            out[b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s2 = 0; s2 < kHid / 32; s2++) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const shortx8 *>(&H[wave][lr][32 * s2 + kb]));
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const bf16x8 w = __builtin_bit_cast(bf16x8, *reinterpret_cast<const shortx8 *>(&W3t[16 * b + lr][32 * s2 + kb]));
                out[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, out[b], 0, 0, 0);
            }
        }
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t m = r0 + 4 * (lane / 16) + j;
                const int n = 16 * b + lr;
                if (m < rows && n < kOut)
                    Y[m * kOut + n] = out[b][j] + bs[2 * kHid + n];
            }
        __syncthreads(); // the tile is free for the next rows
#pragma unroll
        for (int s2 = 0; s2 < kInPad / 32; s2++)
#pragma unroll
            for (int e = 0; e < 8; e++)
                xc[s2][e] = xn[s2][e];
    }
}

// C[i] = sum over z of part[z][i], in order
__global__ __launch_bounds__(256) void k_splitk_sum(int64_t n, int32_t nz, const float *__restrict__ part,
                                                    float *__restrict__ C) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float acc = 0.f;
    for (int32_t z = 0; z < nz; z++)
        acc += part[int64_t(z) * n + i];
    C[i] = acc;
}

// partial column sums of X [rows][n] (bias gradients, n <= 256): block z
// sums rows [z chunk, (z + 1) chunk) in order, one thread per column
__global__ __launch_bounds__(256) void k_colsum(int32_t rows, int32_t n, int32_t chunk, const float *__restrict__ X,
                                                float *__restrict__ part) {
    const int c = threadIdx.x;
    if (c >= n)
        return;
    const int32_t r0 = int32_t(blockIdx.x) * chunk, r1 = min(rows, r0 + chunk);
    float s = 0.f;
    for (int32_t r = r0; r < r1; r++)
        s += X[int64_t(r) * n + c];
    part[int64_t(blockIdx.x) * n + c] = s;
}

// MSE loss and its gradient: dY = 2 (Y - T) / (rows * 36); per-block sums of (Y - T)^2
__global__ __launch_bounds__(256) void k_mse(int64_t n, const float *__restrict__ Y, const float *__restrict__ T,
                                             float scale, float *__restrict__ dY, float *__restrict__ part) {
    __shared__ float red[256];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        const float d = Y[i] - T[i];
        dY[i] = scale * d;
        e = d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size
__global__ __launch_bounds__(256) void k_adam(int64_t n, float *__restrict__ P, const float *__restrict__ G,
                                              float *__restrict__ Mm, float *__restrict__ Vv, float lr_t, float b1,
                                              float b2, float eps_t) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float g = G[i];
    const float m = b1 * Mm[i] + (1.f - b1) * g;
    const float v = b2 * Vv[i] + (1.f - b2) * g * g;
    Mm[i] = m, Vv[i] = v;
    P[i] -= lr_t * m / (sqrtf(v) + eps_t);
}

// features and scales of cell i: the row goes to `row` (58 floats), the
// scales to S
__device__ inline void features_row(int32_t i, int64_t ldc, const int32_t *__restrict__ sten,
                                    const double *__restrict__ u, const float *__restrict__ mom,
                                    const float *__restrict__ logsz, float *row, float *__restrict__ S) {
    double ui[4], d[kSten][4];
#pragma unroll
    for (int v = 0; v < 4; v++)
        ui[v] = u[4 * int64_t(i) + v];
    double s[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < kSten; k++) {
        const int32_t j = sten[k * ldc + i];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            d[k][v] = u[4 * int64_t(j) + v] - ui[v];
            s[v] = fmax(s[v], fabs(d[k][v]));
        }
    }
    // floor: a thousandth of the cell's largest |u_v|, so that nearly uniform
    // stencils (whose coefficients are set by distant cells) stay bounded
    double umax = 0;
#pragma unroll
    for (int v = 0; v < 4; v++)
        umax = fmax(umax, fabs(ui[v]));
#pragma unroll
    for (int v = 0; v < 4; v++) {
        s[v] = fmax(s[v], 1e-3 * umax);
        s[v] = s[v] > 0 ? s[v] : 1.0;
    }
#pragma unroll
    for (int k = 0; k < kSten; k++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            row[4 * k + v] = float(d[k][v] / s[v]);
#pragma unroll
    for (int k = 0; k < 9; k++)
        row[48 + k] = mom[k * ldc + i];
    row[57] = logsz[i];
#pragma unroll
    for (int v = 0; v < 4; v++)
        S[4 * int64_t(i) + v] = float(s[v]);
}

// features and scales of every cell (see the file comment).  The block's
// rows are formed in LDS and written out as one contiguous run: a row is 58
// floats, so stores of whole rows by single threads would not coalesce.
__global__ __launch_bounds__(256) void k_features(int32_t nc, int64_t ldc, const int32_t *__restrict__ sten,
                                                  const double *__restrict__ u, const float *__restrict__ mom,
                                                  const float *__restrict__ logsz, float *__restrict__ X,
                                                  float *__restrict__ S) {
    __shared__ float tile[256][kIn + 1];
    const int64_t i0 = int64_t(blockIdx.x) * 256;
    const int32_t i = int32_t(i0) + int32_t(threadIdx.x);
    if (i < nc)
        features_row(i, ldc, sten, u, mom, logsz, tile[threadIdx.x], S);
    __syncthreads();
    const int64_t nrow = min(int64_t(256), int64_t(nc) - i0);
    for (int64_t e = threadIdx.x; e < nrow * kIn; e += 256)
        X[i0 * kIn + e] = tile[e / kIn][e % kIn];
}

// targets: converged rec coefficients x [k][ldc] (k = r + 9 v) divided by
// s_v, rows of 36 formed in LDS and written contiguously
__global__ __launch_bounds__(256) void k_targets(int32_t nc, int64_t ldc, const double *__restrict__ x,
                                                 const float *__restrict__ S, float *__restrict__ T) {
    __shared__ float tile[256][kOut + 1];
    const int64_t i0 = int64_t(blockIdx.x) * 256;
    const int32_t i = int32_t(i0) + int32_t(threadIdx.x);
    if (i < nc) {
#pragma unroll
        for (int k = 0; k < kOut; k++)
            tile[threadIdx.x][k] = float(x[k * ldc + i] / double(S[4 * int64_t(i) + k / 9]));
    }
    __syncthreads();
    const int64_t nrow = min(int64_t(256), int64_t(nc) - i0);
    for (int64_t e = threadIdx.x; e < nrow * kOut; e += 256)
        T[i0 * kOut + e] = tile[e / kOut][e % kOut];
}

// prediction back to rec coefficients, x = Y s_v (ADD: x += Y s_v, the stage
// correction of the learned warm start inside the mixed-precision step).
// The block's rows of Y are read contiguously into LDS first.
template <bool ADD>
__global__ __launch_bounds__(256) void k_unscale(int32_t nc, int64_t ldc, const float *__restrict__ Y,
                                                 const float *__restrict__ S, double *__restrict__ x) {
    __shared__ float tile[256][kOut + 1];
    const int64_t i0 = int64_t(blockIdx.x) * 256;
    const int64_t nrow = min(int64_t(256), int64_t(nc) - i0);
    for (int64_t e = threadIdx.x; e < nrow * kOut; e += 256)
        tile[e / kOut][e % kOut] = Y[i0 * kOut + e];
    __syncthreads();
    const int32_t i = int32_t(i0) + int32_t(threadIdx.x);
    if (i >= nc)
        return;
#pragma unroll
    for (int k = 0; k < kOut; k++) {
        const double v = double(tile[threadIdx.x][k]) * double(S[4 * int64_t(i) + k / 9]);
        if (ADD)
            x[k * ldc + i] += v;
        else
            x[k * ldc + i] = v;
    }
}

// batch rows by index
__global__ __launch_bounds__(256) void k_gather_rows(int32_t rows, int32_t width, const int32_t *__restrict__ idx,
                                                     const float *__restrict__ src, float *__restrict__ dst) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(rows) * width)
        return;
    const int64_t r = e / width, c = e % width;
    dst[e] = src[int64_t(idx[r]) * width + c];
}

// du = u - u_prev, then u_prev = u (cell means, [cell][4])
__global__ __launch_bounds__(256) void k_mean_change(int32_t nc, const double *__restrict__ u,
                                                     double *__restrict__ uprev, double *__restrict__ du) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * nc)
        return;
    const double v = u[i];
    du[i] = v - uprev[i];
    uprev[i] = v;
}

// per block: sum of (a - b)^2 and of b^2 over n doubles (the corrected state
// against the plain one at the end of a check window)
__global__ __launch_bounds__(256) void k_diff2(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                               double *__restrict__ part) {
    __shared__ double red[2][256];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double d2 = 0.0, b2 = 0.0;
    if (i < n) {
        const double d = a[i] - b[i];
        d2 = d * d, b2 = b[i] * b[i];
    }
    red[0][threadIdx.x] = d2, red[1][threadIdx.x] = b2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            red[0][threadIdx.x] += red[0][threadIdx.x + st], red[1][threadIdx.x] += red[1][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[2 * blockIdx.x] = red[0][0], part[2 * blockIdx.x + 1] = red[1][0];
}

// sum over all cells of |x_new - x_old|^2 (per block), for the rec residual
__global__ __launch_bounds__(256) void k_rec_diff2(int32_t nc, int64_t ldc, const double *__restrict__ a,
                                                   const double *__restrict__ b, double *__restrict__ part) {
    __shared__ double red[256];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < nc)
        for (int k = 0; k < 36; k++) {
            const double d = a[k * ldc + i] - b[k * ldc + i];
            s += d * d;
        }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = red[0];
}

} // namespace

struct dndsmi_warmstart {
    dndsmi_ctx *ctx = nullptr;
    int64_t nc = 0, ldc = 0;
    std::vector<void *> allocs;
    int32_t *sten = nullptr; // [12][ldc]
    float *mom = nullptr;    // [9][ldc]
    float *logsz = nullptr;  // [ldc]
    float *P = nullptr, *G = nullptr, *Mm = nullptr, *Vv = nullptr; // [kNParam]
    int64_t adam_t = 0;
    // training set
    float *X = nullptr, *T = nullptr;
    int64_t n_samples = 0, capacity = 0;
    // per-call work (rows up to max_rows)
    int64_t max_rows = 0;
    float *bX = nullptr, *bT = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Y = nullptr,
          *dY = nullptr, *dZ2 = nullptr, *dZ1 = nullptr, *part = nullptr;
    int32_t *idx = nullptr;
    float *fX = nullptr, *fS = nullptr; // features / scales of all cells
    double *dpart = nullptr;
    float *kpart = nullptr; // split-K partials [64][128][128]
    std::mt19937_64 rng;

    template <class T>
    int alloc(T *&p, int64_t n) {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(T) * size_t(std::max<int64_t>(n, 1))) != hipSuccess) {
            ctx->err = "hipMalloc failed (warm start)";
            return DNDSMI_ERR_OUT_OF_MEMORY;
        }
        allocs.push_back(q);
        p = static_cast<T *>(q);
        return DNDSMI_OK;
    }
    ~dndsmi_warmstart() {
        for (void *q : allocs)
            (void)hipFree(q);
    }
};

namespace {

int ensure_rows(dndsmi_warmstart *w, int64_t rows) {
    if (rows <= w->max_rows)
        return DNDSMI_OK;
    // growth frees nothing: old buffers stay in the allocation list until destroy
    const int64_t r = std::max<int64_t>(rows, 256);
    DNDSMI_TRY(w->alloc(w->bX, r * kIn));
    DNDSMI_TRY(w->alloc(w->bT, r * kOut));
    DNDSMI_TRY(w->alloc(w->Z1, r * kHid));
    DNDSMI_TRY(w->alloc(w->H1, r * kHid));
    DNDSMI_TRY(w->alloc(w->Z2, r * kHid));
    DNDSMI_TRY(w->alloc(w->H2, r * kHid));
    DNDSMI_TRY(w->alloc(w->Y, r * kOut));
    DNDSMI_TRY(w->alloc(w->dY, r * kOut));
    DNDSMI_TRY(w->alloc(w->dZ2, r * kHid));
    DNDSMI_TRY(w->alloc(w->dZ1, r * kHid));
    DNDSMI_TRY(w->alloc(w->idx, r));
    DNDSMI_TRY(w->alloc(w->part, (r * kOut + 255) / 256 + 1));
    w->max_rows = r;
    return DNDSMI_OK;
}

template <int EPI>
void gemm(hipStream_t s, int32_t M, int32_t N, int32_t K, const float *A, int64_t lda, int ta, const float *B,
          int64_t ldb, int tb, float *C, int64_t ldc, const float *bias = nullptr, float *C2 = nullptr,
          const float *Zin = nullptr) {
    dim3 grid((M + 63) / 64, (N + 63) / 64);
    hipLaunchKernelGGL(k_gemm<EPI>, grid, dim3(256), 0, s, M, N, K, A, lda, ta, B, ldb, tb, C, ldc, bias, C2, Zin,
                       K);
}

constexpr int kSplitChunk = 256, kMaxSplit = 64;

// weight gradients C[M][N] = A^T B over a long K (the batch): split K over
// up to 64 blocks per tile, then a fixed-order sum (deterministic)
void gemm_wgrad(hipStream_t s, int32_t M, int32_t N, int32_t K, const float *A, int64_t lda, const float *B,
                int64_t ldb, float *C, float *part) {
    const int32_t chunk = std::max<int32_t>(kSplitChunk, (K + kMaxSplit - 1) / kMaxSplit / 16 * 16 + 16);
    const int32_t nz = (K + chunk - 1) / chunk;
    dim3 grid((M + 63) / 64, (N + 63) / 64, nz);
    hipLaunchKernelGGL(k_gemm<EPI_STORE>, grid, dim3(256), 0, s, M, N, K, A, lda, 1, B, ldb, 0, part, int64_t(N),
                       nullptr, nullptr, nullptr, chunk);
    const int64_t n = int64_t(M) * N;
    hipLaunchKernelGGL(k_splitk_sum, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, nz, part, C);
}

// column sums of X [rows][n] into out (n <= 256, rows >= 1): the rows split
// into at most 1024 chunks, partial sums per chunk, then added in chunk order
// (deterministic); part holds the chunks' sums
void colsum(hipStream_t s, int32_t rows, int32_t n, const float *X, float *out, float *part) {
    static_assert(1024 * kHid <= kMaxSplit * kHid * kHid, "the chunk sums fit in the split-K partials");
    const int32_t chunk = std::max<int32_t>(64, (rows + 1023) / 1024);
    const int32_t nz = (rows + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_colsum, dim3(unsigned(nz)), dim3(256), 0, s, rows, n, chunk, X, part);
    hipLaunchKernelGGL(k_splitk_sum, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, int64_t(n), nz, part, out);
}

// forward pass of `rows` rows of X into Z1, H1, Z2, H2, Y
void forward(dndsmi_warmstart *w, const float *X, int32_t rows) {
    hipStream_t s = w->ctx->stream;
    const float *P = w->P;
    gemm<EPI_BIAS_GELU>(s, rows, kHid, kIn, X, kIn, 0, P + oW1, kHid, 0, w->Z1, kHid, P + ob1, w->H1);
    gemm<EPI_BIAS_GELU>(s, rows, kHid, kHid, w->H1, kHid, 0, P + oW2, kHid, 0, w->Z2, kHid, P + ob2, w->H2);
    gemm<EPI_BIAS>(s, rows, kOut, kHid, w->H2, kHid, 0, P + oW3, kOut, 0, w->Y, kOut, P + ob3);
}

// the network's output for `rows` rows of X into Y (fused kernel, inference
// only: the hidden activations are not kept)
int infer(dndsmi_warmstart *w, const float *X, int32_t rows) {
    const int64_t ntile = (int64_t(rows) + 15) / 16;
    const unsigned grid = unsigned(std::max<int64_t>(1, std::min<int64_t>((ntile + kInferWaves - 1) / kInferWaves, 512)));
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(64 * kInferWaves), 0, w->ctx->stream, rows, X, w->P, w->Y);
    return w->ctx->launch_check("k_mlp_infer");
}

// loss and gradients into G for rows of (X, T); returns the mean squared error
int loss_grad(dndsmi_warmstart *w, const float *X, const float *T, int32_t rows, double *loss) {
    hipStream_t s = w->ctx->stream;
    const float *P = w->P;
    float *G = w->G;
    forward(w, X, rows);
    const int64_t ny = int64_t(rows) * kOut;
    const unsigned nb = unsigned((ny + 255) / 256);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, ny, w->Y, T, float(2.0 / double(ny)), w->dY, w->part);
    // layer 3
    gemm_wgrad(s, kHid, kOut, rows, w->H2, kHid, w->dY, kOut, G + oW3, w->kpart);
    colsum(s, rows, kOut, w->dY, G + ob3, w->kpart);
    gemm<EPI_GELU_GRAD>(s, rows, kHid, kOut, w->dY, kOut, 0, P + oW3, kOut, 1, w->dZ2, kHid, nullptr, nullptr, w->Z2);
    // layer 2
    gemm_wgrad(s, kHid, kHid, rows, w->H1, kHid, w->dZ2, kHid, G + oW2, w->kpart);
    colsum(s, rows, kHid, w->dZ2, G + ob2, w->kpart);
    gemm<EPI_GELU_GRAD>(s, rows, kHid, kHid, w->dZ2, kHid, 0, P + oW2, kHid, 1, w->dZ1, kHid, nullptr, nullptr, w->Z1);
    // layer 1
    gemm_wgrad(s, kIn, kHid, rows, X, kIn, w->dZ1, kHid, G + oW1, w->kpart);
    colsum(s, rows, kHid, w->dZ1, G + ob1, w->kpart);
    DNDSMI_TRY(w->ctx->launch_check("warm-start backward"));
    if (loss) {
        std::vector<float> h(nb);
        DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(h.data(), w->part, sizeof(float) * nb, hipMemcpyDeviceToHost, s),
                               "hipMemcpyAsync"));
        DNDSMI_TRY(w->ctx->hip(hipStreamSynchronize(s), "hipStreamSynchronize"));
        double acc = 0;
        for (float v : h)
            acc += v;
        *loss = acc / double(ny);
    }
    return DNDSMI_OK;
}

void adam(dndsmi_warmstart *w, double lr) {
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    w->adam_t++;
    const double c1 = 1 - std::pow(b1, double(w->adam_t)), c2 = 1 - std::pow(b2, double(w->adam_t));
    // step lr sqrt(c2) / c1 on m / (sqrt(v) + eps sqrt(c2)), the bias-corrected update
    hipLaunchKernelGGL(k_adam, dim3(unsigned((kNParam + 255) / 256)), dim3(256), 0, w->ctx->stream, kNParam, w->P,
                       w->G, w->Mm, w->Vv, float(lr * std::sqrt(c2) / c1), float(b1), float(b2),
                       float(eps * std::sqrt(c2)));
}

int features(dndsmi_warmstart *w, const double *u) {
    if (!w->fX) {
        DNDSMI_TRY(w->alloc(w->fX, w->nc * kIn));
        DNDSMI_TRY(w->alloc(w->fS, w->nc * 4));
    }
    hipLaunchKernelGGL(k_features, dim3(grid_for(w->nc)), dim3(256), 0, w->ctx->stream, int32_t(w->nc), w->ldc,
                       w->sten, u, w->mom, w->logsz, w->fX, w->fS);
    return w->ctx->launch_check("k_features");
}

} // namespace

extern "C" int dndsmi_warmstart_create(dndsmi_ctx *c, const dndsmi_mesh_view *mesh, const dndsmi_vr_view *vr,
                                       uint64_t seed, int64_t capacity, dndsmi_warmstart **out) {
    if (!c || !mesh || !vr || !out || mesh->n_cell != c->nc || vr->n_cell != c->nc || capacity < 0)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    {
        std::string why;
        if (check_mesh_view(mesh, why) != DNDSMI_OK)
            return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    auto *w = new (std::nothrow) dndsmi_warmstart();
    if (!w)
        return DNDSMI_ERR_OUT_OF_MEMORY;
    w->ctx = c, w->nc = c->nc, w->ldc = c->ldc;
    w->rng.seed(seed);
    auto fail = [&](int rc) {
        delete w;
        return rc;
    };
    const int64_t nc = c->nc, ldc = c->ldc;
    // stencil: breadth-first over faces, cell2face order, first 12 cells
    std::vector<int32_t> sten(size_t(kSten) * ldc, 0);
    std::vector<float> mom(size_t(9) * ldc, 0.f), lsz(size_t(ldc), 0.f);
    std::vector<int64_t> ring, next;
    std::vector<int64_t> seen;
    for (int64_t i = 0; i < nc; i++) {
        seen.assign(1, i);
        ring.assign(1, i);
        std::vector<int64_t> got;
        while (int(got.size()) < kSten && !ring.empty()) {
            next.clear();
            for (int64_t a : ring)
                for (int64_t q = mesh->cell_ptr[a]; q < mesh->cell_ptr[a + 1]; q++) {
                    const int64_t f = mesh->cell2face[q];
                    const int64_t L = mesh->face2cell[2 * f], R = mesh->face2cell[2 * f + 1];
                    const int64_t b = L == a ? R : L;
                    if (b == DNDSMI_NO_CELL || std::find(seen.begin(), seen.end(), b) != seen.end())
                        continue;
                    seen.push_back(b);
                    next.push_back(b);
                    if (int(got.size()) < kSten)
                        got.push_back(b);
                }
            ring = next;
        }
        const int64_t ci = c->c2d[i]; // device numbering
        for (int k = 0; k < kSten; k++)
            sten[k * ldc + ci] = c->c2d[k < int(got.size()) ? got[k] : i];
        for (int k = 0; k < 9; k++)
            mom[k * ldc + ci] = float(vr->base_moment[10 * i + 1 + k]);
        lsz[ci] = float(0.5 * std::log(vr->volume[i]));
    }
    try {
        if (w->alloc(w->sten, kSten * ldc) || w->alloc(w->mom, 9 * ldc) || w->alloc(w->logsz, ldc) ||
            w->alloc(w->P, kNParam) || w->alloc(w->G, kNParam) || w->alloc(w->Mm, kNParam) ||
            w->alloc(w->Vv, kNParam) || w->alloc(w->X, capacity * kIn) || w->alloc(w->T, capacity * kOut) ||
            w->alloc(w->dpart, int64_t(grid_for(nc)) + 1) || w->alloc(w->kpart, int64_t(kMaxSplit) * kHid * kHid))
            return fail(DNDSMI_ERR_OUT_OF_MEMORY);
    } catch (...) {
        return fail(DNDSMI_ERR_OUT_OF_MEMORY);
    }
    w->capacity = capacity;
    // He-style initialisation, seeded: weights U(-sqrt(6/fan_in), ...), zero biases
    std::vector<float> P(kNParam, 0.f);
    auto init = [&](int64_t o, int fin, int fout) {
        std::uniform_real_distribution<float> d(-std::sqrt(6.f / fin), std::sqrt(6.f / fin));
        for (int64_t k = 0; k < int64_t(fin) * fout; k++)
            P[o + k] = d(w->rng);
    };
    init(oW1, kIn, kHid), init(oW2, kHid, kHid), init(oW3, kHid, kOut);
    for (int64_t k = oW3; k < oW3 + kHid * kOut; k++)
        P[k] *= 0.1f; // small initial output
    std::vector<float> zero(kNParam, 0.f);
    int rc = DNDSMI_OK;
    rc = rc ? rc : c->upload(w->sten, sten);
    rc = rc ? rc : c->upload(w->mom, mom);
    rc = rc ? rc : c->upload(w->logsz, lsz);
    rc = rc ? rc : c->upload(w->P, P);
    rc = rc ? rc : c->upload(w->Mm, zero);
    rc = rc ? rc : c->upload(w->Vv, zero);
    if (rc)
        return fail(rc);
    c->warmstarts.insert(w);
    *out = w;
    return DNDSMI_OK;
}

void dndsmi_warmstart_detach(dndsmi_warmstart *w) { w->ctx = nullptr; } // keeps its own allocations

extern "C" void dndsmi_warmstart_destroy(dndsmi_warmstart *w) {
    if (w && w->ctx) {
        (void)hipStreamSynchronize(w->ctx->stream);
        w->ctx->warmstarts.erase(w);
        if (w->ctx->step_ws == w)
            w->ctx->step_ws = nullptr;
    }
    delete w;
}

// ---- the learned warm start inside the mixed-precision step ------------
//
// Every stage of dndsmi_ssprk3_step_mixed starts its sweeps from the uRec the
// previous stage left (the reference's own warm start).  With a network
// attached, the stage first adds the network's prediction on the change of
// the cell means since the previous stage.  The plain method judges it: in a
// check window every step runs twice, plain on shadow copies of u and uRec
// and then with the correction, with the plain run's time step; at the
// window's end the corrected state must lie within the tolerance of the
// plain one (relative L2 over all cells), else u and uRec become the plain
// run's and the correction stays off until the next attach.  A corrected
// run that meets a non-positive state inside a window is replaced the same
// way.

namespace dndsmi_host {

int ws_stage_begin(dndsmi_ctx *c, const double *u, double *x, bool use) {
    dndsmi_warmstart *w = c->step_ws;
    if (!c->ws_uprev) {
        DNDSMI_TRY(c->alloc(c->ws_uprev, 4 * c->ldc));
        DNDSMI_TRY(c->alloc(c->ws_du, 4 * c->ldc));
    }
    if (!c->ws_have_prev) {
        DNDSMI_TRY(c->hip(hipMemcpyAsync(c->ws_uprev, u, sizeof(double) * 4 * c->nc, hipMemcpyDeviceToDevice,
                                         c->stream),
                          "copy u"));
        c->ws_have_prev = true;
        return DNDSMI_OK;
    }
    hipLaunchKernelGGL(k_mean_change, dim3(grid_for(4 * c->nc)), dim3(256), 0, c->stream, int32_t(c->nc), u,
                       c->ws_uprev, c->ws_du);
    DNDSMI_TRY(c->launch_check("k_mean_change"));
    if (!use)
        return DNDSMI_OK;
    DNDSMI_TRY(features(w, c->ws_du));
    DNDSMI_TRY(ensure_rows(w, w->nc));
    DNDSMI_TRY(infer(w, w->fX, int32_t(w->nc)));
    hipLaunchKernelGGL(k_unscale<true>, dim3(grid_for(w->nc)), dim3(256), 0, c->stream, int32_t(w->nc), w->ldc, w->Y,
                       w->fS, x);
    DNDSMI_TRY(c->launch_check("warm-start stage correction"));
    c->ws_stats[0]++;
    return DNDSMI_OK;
}

int ws_window_begin(dndsmi_ctx *c, const double *u, const double *x) {
    if (!c->ws_sh_u) {
        DNDSMI_TRY(c->alloc(c->ws_sh_u, 4 * c->ldc));
        DNDSMI_TRY(c->alloc(c->ws_sh_x, 36 * c->ldc));
        DNDSMI_TRY(c->alloc(c->ws_part, 2 * int64_t(grid_for(36 * c->ldc)) + 2));
    }
    DNDSMI_TRY(c->hip(hipMemcpyAsync(c->ws_sh_u, u, sizeof(double) * 4 * c->ldc, hipMemcpyDeviceToDevice, c->stream),
                      "copy u to the shadow"));
    return c->hip(hipMemcpyAsync(c->ws_sh_x, x, sizeof(double) * 36 * c->ldc, hipMemcpyDeviceToDevice, c->stream),
                  "copy uRec to the shadow");
}

int ws_window_end(dndsmi_ctx *c, double *u, double *x, bool failed) {
    bool reject = failed;
    if (!reject) {
        const int64_t n = 4 * c->nc;
        const unsigned g = grid_for(n, 256);
        hipLaunchKernelGGL(k_diff2, dim3(g), dim3(256), 0, c->stream, n, u, c->ws_sh_u, c->ws_part);
        DNDSMI_TRY(c->launch_check("k_diff2"));
        std::vector<double> part(2 * size_t(g));
        DNDSMI_TRY(c->hip(hipMemcpyAsync(part.data(), c->ws_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                         c->stream),
                          "hipMemcpyAsync"));
        DNDSMI_TRY(c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
        double d2 = 0, b2 = 0;
        for (unsigned k = 0; k < g; k++)
            d2 += part[2 * k], b2 += part[2 * k + 1];
        c->ws_last_dev = std::sqrt(d2 / (b2 > 0 ? b2 : 1.0));
        reject = !(c->ws_last_dev <= c->ws_tol);
    }
    if (reject) {
        DNDSMI_TRY(c->hip(hipMemcpyAsync(u, c->ws_sh_u, sizeof(double) * 4 * c->ldc, hipMemcpyDeviceToDevice,
                                         c->stream),
                          "restore u"));
        DNDSMI_TRY(c->hip(hipMemcpyAsync(x, c->ws_sh_x, sizeof(double) * 36 * c->ldc, hipMemcpyDeviceToDevice,
                                         c->stream),
                          "restore uRec"));
        c->ws_rejected = true;
        c->ws_stats[2]++;
    } else {
        c->ws_stats[1]++;
    }
    return DNDSMI_OK;
}

} // namespace dndsmi_host

extern "C" int dndsmi_ssprk3_set_warm_start(dndsmi_ctx *c, dndsmi_warmstart *ws, int32_t check_steps,
                                            int32_t recheck_interval, double tolerance) {
    if (!c || (ws && ws->ctx != c) || check_steps < 1 || recheck_interval < 0 || !(tolerance >= 0))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    c->step_ws = ws;
    c->ws_check_steps = check_steps;
    c->ws_recheck = recheck_interval;
    c->ws_tol = tolerance;
    c->ws_have_prev = false;
    c->ws_rejected = false;
    c->ws_win_left = 0;
    c->ws_since_check = 0;
    c->ws_first = true;
    c->ws_last_dev = -1.0;
    for (int k = 0; k < 3; k++)
        c->ws_stats[k] = 0;
    return DNDSMI_OK;
}

extern "C" int dndsmi_ssprk3_warm_start_stats(const dndsmi_ctx *c, int64_t stats[3], double *last_deviation) {
    if (!c || !stats)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < 3; k++)
        stats[k] = c->ws_stats[k];
    if (last_deviation)
        *last_deviation = c->ws_last_dev;
    return DNDSMI_OK;
}

extern "C" int64_t dndsmi_warmstart_num_params(void) { return kNParam; }

extern "C" int dndsmi_warmstart_get_params(dndsmi_warmstart *w, float *host) {
    if (!w || !host)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(host, w->P, sizeof(float) * kNParam, hipMemcpyDeviceToHost, w->ctx->stream),
                           "hipMemcpyAsync"));
    return w->ctx->hip(hipStreamSynchronize(w->ctx->stream), "hipStreamSynchronize");
}

extern "C" int dndsmi_warmstart_set_params(dndsmi_warmstart *w, const float *host) {
    if (!w || !host)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(w->P, host, sizeof(float) * kNParam, hipMemcpyHostToDevice, w->ctx->stream),
                           "hipMemcpyAsync"));
    w->adam_t = 0;
    std::vector<float> zero(kNParam, 0.f);
    DNDSMI_TRY(w->ctx->upload(w->Mm, zero));
    return w->ctx->upload(w->Vv, zero);
}

extern "C" int dndsmi_warmstart_loss_grad(dndsmi_warmstart *w, const float *X, const float *T, int32_t rows,
                                          float *grad, double *loss) {
    if (!w || !X || !T || rows < 1)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(ensure_rows(w, rows));
    hipStream_t s = w->ctx->stream;
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(w->bX, X, sizeof(float) * rows * kIn, hipMemcpyHostToDevice, s), "copy"));
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(w->bT, T, sizeof(float) * rows * kOut, hipMemcpyHostToDevice, s), "copy"));
    DNDSMI_TRY(loss_grad(w, w->bX, w->bT, rows, loss));
    if (grad) {
        DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(grad, w->G, sizeof(float) * kNParam, hipMemcpyDeviceToHost, s), "copy"));
        DNDSMI_TRY(w->ctx->hip(hipStreamSynchronize(s), "hipStreamSynchronize"));
    }
    return DNDSMI_OK;
}

extern "C" int dndsmi_warmstart_adam_step(dndsmi_warmstart *w, const float *grad, double lr) {
    if (!w || !grad)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(w->G, grad, sizeof(float) * kNParam, hipMemcpyHostToDevice, w->ctx->stream),
                           "copy"));
    adam(w, lr);
    return w->ctx->launch_check("k_adam");
}

extern "C" int dndsmi_warmstart_features(dndsmi_warmstart *w, const dndsmi_array *u, float *X, float *S) {
    if (!w || !same_ctx(w->ctx, u, DNDSMI_ARRAY_DOF))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(features(w, u->d));
    hipStream_t s = w->ctx->stream;
    std::vector<float> hx(size_t(w->nc) * kIn), hs(size_t(w->nc) * 4);
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(hx.data(), w->fX, sizeof(float) * hx.size(), hipMemcpyDeviceToHost, s),
                           "copy"));
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(hs.data(), w->fS, sizeof(float) * hs.size(), hipMemcpyDeviceToHost, s),
                           "copy"));
    DNDSMI_TRY(w->ctx->hip(hipStreamSynchronize(s), "hipStreamSynchronize"));
    for (int64_t i = 0; i < w->nc; i++) { // rows back in mesh numbering
        const int64_t d = w->ctx->c2d[i];
        if (X)
            std::memcpy(X + i * kIn, hx.data() + d * kIn, sizeof(float) * kIn);
        if (S)
            std::memcpy(S + i * 4, hs.data() + d * 4, sizeof(float) * 4);
    }
    return DNDSMI_OK;
}

extern "C" int dndsmi_warmstart_add_samples(dndsmi_warmstart *w, const dndsmi_array *u, const dndsmi_array *uRec) {
    if (!w || !same_ctx(w->ctx, u, DNDSMI_ARRAY_DOF) || !same_ctx(w->ctx, uRec, DNDSMI_ARRAY_REC))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (w->n_samples + w->nc > w->capacity) {
        w->ctx->err = "warm-start training set is full";
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    DNDSMI_TRY(features(w, u->d));
    hipStream_t s = w->ctx->stream;
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(w->X + w->n_samples * kIn, w->fX, sizeof(float) * w->nc * kIn,
                                          hipMemcpyDeviceToDevice, s),
                           "copy"));
    hipLaunchKernelGGL(k_targets, dim3(grid_for(w->nc)), dim3(256), 0, s, int32_t(w->nc), w->ldc, uRec->d, w->fS,
                       w->T + w->n_samples * kOut);
    DNDSMI_TRY(w->ctx->launch_check("k_targets"));
    w->n_samples += w->nc;
    return DNDSMI_OK;
}

extern "C" int dndsmi_warmstart_add_rows(dndsmi_warmstart *w, const float *X, const float *T, int64_t rows) {
    if (!w || !X || !T || rows < 0)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    if (w->n_samples + rows > w->capacity) {
        w->ctx->err = "warm-start training set is full";
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    hipStream_t s = w->ctx->stream;
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(w->X + w->n_samples * kIn, X, sizeof(float) * rows * kIn,
                                          hipMemcpyHostToDevice, s),
                           "copy"));
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(w->T + w->n_samples * kOut, T, sizeof(float) * rows * kOut,
                                          hipMemcpyHostToDevice, s),
                           "copy"));
    DNDSMI_TRY(w->ctx->hip(hipStreamSynchronize(s), "hipStreamSynchronize"));
    w->n_samples += rows;
    return DNDSMI_OK;
}

extern "C" int dndsmi_warmstart_targets(dndsmi_warmstart *w, const dndsmi_array *u, const dndsmi_array *uRec,
                                        float *T) {
    if (!w || !T || !same_ctx(w->ctx, u, DNDSMI_ARRAY_DOF) || !same_ctx(w->ctx, uRec, DNDSMI_ARRAY_REC))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(features(w, u->d));
    DNDSMI_TRY(ensure_rows(w, w->nc));
    hipStream_t s = w->ctx->stream;
    hipLaunchKernelGGL(k_targets, dim3(grid_for(w->nc)), dim3(256), 0, s, int32_t(w->nc), w->ldc, uRec->d, w->fS,
                       w->Y);
    DNDSMI_TRY(w->ctx->launch_check("k_targets"));
    std::vector<float> ht(size_t(w->nc) * kOut);
    DNDSMI_TRY(w->ctx->hip(hipMemcpyAsync(ht.data(), w->Y, sizeof(float) * ht.size(), hipMemcpyDeviceToHost, s),
                           "copy"));
    DNDSMI_TRY(w->ctx->hip(hipStreamSynchronize(s), "hipStreamSynchronize"));
    for (int64_t i = 0; i < w->nc; i++)
        std::memcpy(T + i * kOut, ht.data() + int64_t(w->ctx->c2d[i]) * kOut, sizeof(float) * kOut);
    return DNDSMI_OK;
}

extern "C" int64_t dndsmi_warmstart_num_samples(const dndsmi_warmstart *w) { return w ? w->n_samples : -1; }

extern "C" int dndsmi_warmstart_train(dndsmi_warmstart *w, int32_t epochs, int32_t batch, double lr, double lr_final,
                                      double *loss_hist) {
    if (!w || epochs < 1 || batch < 1 || w->n_samples < 1)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    const int64_t n = w->n_samples;
    const int32_t b = int32_t(std::min<int64_t>(batch, n));
    DNDSMI_TRY(ensure_rows(w, b));
    int32_t *perm = nullptr;
    DNDSMI_TRY(w->alloc(perm, n)); // freed with the warm start
    std::vector<int32_t> hp(static_cast<size_t>(n));
    std::iota(hp.begin(), hp.end(), 0);
    hipStream_t s = w->ctx->stream;
    const int64_t steps_per_epoch = n / b, total = steps_per_epoch * epochs;
    int64_t step = 0;
    for (int32_t e = 0; e < epochs; e++) {
        std::shuffle(hp.begin(), hp.end(), w->rng);
        DNDSMI_TRY(w->ctx->upload(perm, hp));
        double acc = 0;
        for (int64_t q = 0; q < steps_per_epoch; q++, step++) {
            const int32_t *ix = perm + q * b;
            hipLaunchKernelGGL(k_gather_rows, dim3(unsigned((int64_t(b) * kIn + 255) / 256)), dim3(256), 0, s, b, kIn,
                               ix, w->X, w->bX);
            hipLaunchKernelGGL(k_gather_rows, dim3(unsigned((int64_t(b) * kOut + 255) / 256)), dim3(256), 0, s, b,
                               kOut, ix, w->T, w->bT);
            double l = 0;
            const bool last = q == steps_per_epoch - 1;
            DNDSMI_TRY(loss_grad(w, w->bX, w->bT, b, last ? &l : nullptr));
            // cosine decay from lr to lr_final
            const double f = 0.5 * (1 + std::cos(M_PI * double(step) / double(std::max<int64_t>(total, 1))));
            adam(w, lr_final + (lr - lr_final) * f);
            if (last)
                acc = l;
        }
        DNDSMI_TRY(w->ctx->launch_check("warm-start training"));
        if (loss_hist)
            loss_hist[e] = acc;
    }
    return w->ctx->hip(hipStreamSynchronize(s), "hipStreamSynchronize");
}

extern "C" int dndsmi_warmstart_predict(dndsmi_warmstart *w, const dndsmi_array *u, dndsmi_array *uRec) {
    if (!w || !same_ctx(w->ctx, u, DNDSMI_ARRAY_DOF) || !same_ctx(w->ctx, uRec, DNDSMI_ARRAY_REC))
        return DNDSMI_ERR_INVALID_ARGUMENT;
    DNDSMI_TRY(features(w, u->d));
    DNDSMI_TRY(ensure_rows(w, w->nc));
    DNDSMI_TRY(infer(w, w->fX, int32_t(w->nc)));
    hipLaunchKernelGGL(k_unscale<false>, dim3(grid_for(w->nc)), dim3(256), 0, w->ctx->stream, int32_t(w->nc), w->ldc, w->Y,
                       w->fS, uRec->d);
    return w->ctx->launch_check("warm-start predict");
}

extern "C" int dndsmi_reconstruct_to_tolerance(dndsmi_ctx *c, const dndsmi_array *u, dndsmi_array *uRec,
                                               dndsmi_array *uRecNewBuf, double tol, int32_t max_sweeps, double t,
                                               int32_t *sweeps, double *res_out) {
    if (!c || !same_ctx(c, u, DNDSMI_ARRAY_DOF) || !same_ctx(c, uRec, DNDSMI_ARRAY_REC) ||
        !same_ctx(c, uRecNewBuf, DNDSMI_ARRAY_REC) || uRec == uRecNewBuf || max_sweeps < 1)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    const unsigned g = grid_for(c->nc);
    std::vector<double> part(g);
    double *dpart = nullptr;
    if (!c->res_part) { // the implicit path's scratch holds 4 x grid doubles
        DNDSMI_TRY(c->alloc(c->res_part, 4 * int64_t(g)));
        DNDSMI_TRY(c->alloc(c->res, 4));
    }
    dpart = c->res_part;
    double res = 0;
    int32_t k = 0;
    while (k < max_sweeps) {
        DNDSMI_TRY(launch_jacobi(c, u->d, uRec->d, uRecNewBuf->d, t));
        hipLaunchKernelGGL(k_rec_diff2, dim3(g), dim3(256), 0, c->stream, int32_t(c->nc), c->ldc, uRec->d,
                           uRecNewBuf->d, dpart);
        DNDSMI_TRY(c->launch_check("k_rec_diff2"));
        std::swap(uRec->d, uRecNewBuf->d);
        k++;
        DNDSMI_TRY(c->hip(hipMemcpyAsync(part.data(), dpart, sizeof(double) * g, hipMemcpyDeviceToHost, c->stream),
                          "hipMemcpyAsync"));
        DNDSMI_TRY(check_errw(c));
        res = 0;
        for (double v : part)
            res += v;
        res /= double(c->nc);
        if (res < tol)
            break;
    }
    if (sweeps)
        *sweeps = k;
    if (res_out)
        *res_out = res;
    return DNDSMI_OK;
}
