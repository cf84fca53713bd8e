// The learned surrogate for an expensive body's signed distance: a Fourier-feature MLP
// (3 -> 2F sin/cos features -> depth x width SiLU -> 1), BF16 operands on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, FP32 master weights, trained with Adam by the
// kernels in this file.  No counterpart in the reference (whose bodies are SDFLibrary closures).
//
// Kernels:
//   k_mlp_forward   -- fused forward of a 64-point tile through every layer, activations kept in LDS
//                      (ping-pong), weights streamed from L2; optionally saves the pre-activations z
//                      (FP32) and activations a (BF16) of every layer for the backward pass.
//   k_gemm          -- generic 64x64-tiled BF16 GEMM, C = op(A) op(B), for the backward pass:
//                      dA = dZ W^T (with the SiLU' epilogue) and dW = A^T dZ (split over the batch,
//                      FP32 atomics).
//   k_out_*, k_colsum, k_adam -- the output layer, bias gradients and the optimiser step.
//
// MFMA 16x16x32 BF16 operand layout (wave64): lane l holds, for A (16 x 32, M x K), row l%16 and the
// 8 consecutive k = 8*(l/16) .. +7; for B (32 x 16, K x N), column l%16 and the same 8 k; the FP32
// result (16 x 16) gives lane l column l%16 and rows 4*(l/16) .. +3.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "cc_internal.h"


namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

__device__ __forceinline__ u16 f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load8(const u16 *p) {  // 16-byte aligned
    return *reinterpret_cast<const bf16x8 *>(p);
}

constexpr int TILE = 64;  // points per forward tile / GEMM tile edge
constexpr int KS = 32;    // K step of one MFMA

struct MlpDev {
    const float *ffB;    // [F][3] Fourier frequencies (multiplied by 2 pi in the features)
    const u16 *Wt[8];    // per layer, transposed weights [N][K] in BF16 (k contiguous per output)
    const float *bias[8];
    const float *wout;   // [width]
    const float *bout;   // [1], read on the device
    float in_lo, in_scale;  // input map u = (x - in_lo) * in_scale applied before the features
    int F, width, depth;
};

constexpr int WIDTH = 256;        // hidden width the kernels are written for
constexpr int LDW = WIDTH + 8;    // LDS row stride (BF16 elements) of the activation tile

// ---- fused forward ------------------------------------------------------------------------------
// One workgroup (256 threads, 4 waves) runs TILE points through every layer.  The activation tile
// lives in one LDS buffer, [TILE][LDW] BF16 (33 KB): each wave accumulates its 64 output columns
// (4 x 4 MFMA tiles, 64 accumulator registers) from the whole tile, the workgroup synchronises,
// and the new activations overwrite the old in place.  B fragments come straight from the
// transposed weights (L2-resident: 230 k parameters = 460 KB in BF16).
// Points are either explicit (x: n x 3) or, for LAT, the nodes of a lattice: point p is node
// (p % lat.n[0], p / lat.n[0] % lat.n[1], p / (lat.n[0] lat.n[1])) at lat.x0 + index * lat.h.
template <bool SAVE, bool LAT>
__global__ __launch_bounds__(256) void k_mlp_forward(MlpDev m, const float *__restrict__ x, int n,
                                                    float *__restrict__ phi, u16 *__restrict__ feat,
                                                    float *__restrict__ zs, u16 *__restrict__ as,
                                                    cc::Lattice lat) {
    __shared__ __attribute__((aligned(16))) u16 buf[TILE * LDW];
    const int W = WIDTH, F2 = 2 * m.F, LDF = F2 + 8;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int p0 = blockIdx.x * TILE;
    // Fourier features: [sin(2 pi B x), cos(2 pi B x)] for each of the TILE points
    for (int e = t; e < TILE * m.F; e += 256) {
        const int p = e / m.F, f = e - p * m.F;
        float s = 0.0f, c = 1.0f;
        if (p0 + p < n) {
            float xq[3];
            if (LAT) {
                const int q = p0 + p, i = q % lat.n[0], r = q / lat.n[0];
                const int j = r % lat.n[1], k = r / lat.n[1];
                xq[0] = lat.x0[0] + (float)i * lat.h[0];
                xq[1] = lat.x0[1] + (float)j * lat.h[1];
                xq[2] = lat.x0[2] + (float)k * lat.h[2];
            } else {
                xq[0] = x[3 * (size_t)(p0 + p)];
                xq[1] = x[3 * (size_t)(p0 + p) + 1];
                xq[2] = x[3 * (size_t)(p0 + p) + 2];
            }
            const float u0 = (xq[0] - m.in_lo) * m.in_scale, u1 = (xq[1] - m.in_lo) * m.in_scale,
                        u2 = (xq[2] - m.in_lo) * m.in_scale;
            const float *b = m.ffB + 3 * f;
            // sin and cos of 2 pi (B . u): the hardware sine and cosine take revolutions
            const float rev = (b[0] * u0 + b[1] * u1) + b[2] * u2;
            s = __builtin_amdgcn_sinf(rev);
            c = __builtin_amdgcn_cosf(rev);
        }
        buf[p * LDF + f] = f2bf(s);
        buf[p * LDF + m.F + f] = f2bf(c);
        if (SAVE && p0 + p < n) {
            feat[(size_t)(p0 + p) * F2 + f] = f2bf(s);
            feat[(size_t)(p0 + p) * F2 + m.F + f] = f2bf(c);
        }
    }
    __syncthreads();
    int K = F2, LDin = LDF;
    for (int l = 0; l < m.depth; ++l) {
        // this wave's columns [64*wave, 64*wave + 64) of the layer
        f32x4 acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        const u16 *Wt = m.Wt[l];
        for (int k0 = 0; k0 < K; k0 += KS) {
            const int kk = k0 + 8 * (lane >> 4);
            bf16x8 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = load8(buf + (16 * i + (lane & 15)) * LDin + kk);
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = load8(Wt + (size_t)(64 * wave + 16 * j + (lane & 15)) * K + kk);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();  // every wave has read the whole input tile
        const float *bias = m.bias[l];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = 64 * wave + 16 * j + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * i + 4 * (lane >> 4) + r;
                    const float z = acc[i][j][r] + bias[col];
                    const u16 a16 = f2bf(silu(z));
                    buf[row * LDW + col] = a16;
                    if (SAVE && p0 + row < n) {
                        zs[((size_t)l * n + p0 + row) * W + col] = z;
                        as[((size_t)l * n + p0 + row) * W + col] = a16;
                    }
                }
            }
        __syncthreads();
        K = W;
        LDin = LDW;
    }
    // output layer: phi = sum_k a[k] w[k] + b, 4 lanes per point
    const float bo = *m.bout;
    for (int p = t >> 2; p < TILE; p += 64) {
        float s = 0.0f;
        for (int k = (t & 3); k < W; k += 4) s += bf2f(buf[p * LDW + k]) * bf2f(f2bf(m.wout[k]));
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if ((t & 3) == 0 && p0 + p < n) phi[p0 + p] = s + bo;
    }
}

// ---- inference forward --------------------------------------------------------------------------
// The same network for evaluation only, on 128-point tiles, with the matrix-core operands swapped:
// A = the transposed weights (rows = output features, 8 consecutive k per lane, read from L2 with
// the next K step's fragments loaded while the current step's MFMAs run), B = the activation tile
// (columns = points).  Each wave owns 64 output features for all 128 points (4 x 8 MFMA tiles), and
// its accumulators hold 4 consecutive features of one point per lane, so the epilogue -- bias,
// SiLU with the hardware reciprocal, one packed BF16 conversion -- ends in one 8-byte LDS store per
// four values.  The activation tile is [128][LDW] BF16 (67.6 KB), updated in place between layers.
constexpr int ITILE = 128;
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

template <bool LAT>
__global__ __launch_bounds__(256, 2) void k_mlp_infer(MlpDev m, const float *__restrict__ x, int n,
                                                     float *__restrict__ phi, cc::Lattice lat) {
    extern __shared__ __attribute__((aligned(16))) u16 ibuf[];  // [ITILE][LDW], point-major
    __shared__ float wo[WIDTH];                                  // BF16-rounded output weights
    const int F2 = 2 * m.F, LDF = F2 + 8;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int p0 = blockIdx.x * ITILE;
    for (int e = t; e < ITILE * m.F; e += 256) {
        const int p = e / m.F, f = e - p * m.F;
        float sv = 0.0f, cv = 1.0f;
        if (p0 + p < n) {
            float xq[3];
            if (LAT) {
                const int q = p0 + p, i = q % lat.n[0], r = q / lat.n[0];
                const int j = r % lat.n[1], k = r / lat.n[1];
                xq[0] = lat.x0[0] + (float)i * lat.h[0];
                xq[1] = lat.x0[1] + (float)j * lat.h[1];
                xq[2] = lat.x0[2] + (float)k * lat.h[2];
            } else {
                xq[0] = x[3 * (size_t)(p0 + p)];
                xq[1] = x[3 * (size_t)(p0 + p) + 1];
                xq[2] = x[3 * (size_t)(p0 + p) + 2];
            }
            const float u0 = (xq[0] - m.in_lo) * m.in_scale, u1 = (xq[1] - m.in_lo) * m.in_scale,
                        u2 = (xq[2] - m.in_lo) * m.in_scale;
            const float *b = m.ffB + 3 * f;
            const float rev = (b[0] * u0 + b[1] * u1) + b[2] * u2;  // revolutions
            sv = __builtin_amdgcn_sinf(rev);
            cv = __builtin_amdgcn_cosf(rev);
        }
        ibuf[p * LDF + f] = f2bf(sv);
        ibuf[p * LDF + m.F + f] = f2bf(cv);
    }
    wo[t] = bf2f(f2bf(m.wout[t]));
    __syncthreads();
    int K = F2, LDin = LDF;
    const int arow = 64 * wave + (lane & 15), koff = 8 * (lane >> 4);
    for (int l = 0; l < m.depth; ++l) {
        f32x4 acc[4][8];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        const u16 *Wt = m.Wt[l];
        bf16x8 a[4], an[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = load8(Wt + (size_t)(arow + 16 * i) * K + koff);
        for (int k0 = 0; k0 < K; k0 += KS) {
            const int kn = k0 + KS < K ? k0 + KS : k0;  // next step's fragments (the last reloads its own)
#pragma unroll
            for (int i = 0; i < 4; ++i) an[i] = load8(Wt + (size_t)(arow + 16 * i) * K + kn + koff);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const bf16x8 b = load8(ibuf + (16 * j + (lane & 15)) * LDin + k0 + koff);
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[i][j] = mfma(a[i], b, acc[i][j]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = an[i];
        }
        __syncthreads();  // every wave has read the whole input tile
        const float *bias = m.bias[l];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int f0 = 64 * wave + 16 * i + 4 * (lane >> 4);
            const f32x4 bv = *reinterpret_cast<const f32x4 *>(bias + f0);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int p = 16 * j + (lane & 15);
                const bf16x4 pk = {(__bf16)silu_fast(acc[i][j][0] + bv[0]), (__bf16)silu_fast(acc[i][j][1] + bv[1]),
                                   (__bf16)silu_fast(acc[i][j][2] + bv[2]), (__bf16)silu_fast(acc[i][j][3] + bv[3])};
                *reinterpret_cast<bf16x4 *>(ibuf + p * LDW + f0) = pk;
            }
        }
        __syncthreads();
        K = WIDTH;
        LDin = LDW;
    }
    // output layer: two lanes per point, 128 of the 256 k each
    const float bo = *m.bout;
    {
        const int p = t >> 1, h = t & 1;
        float sacc = 0.0f;
        for (int k = h; k < WIDTH; k += 2) sacc += bf2f(ibuf[p * LDW + k]) * wo[k];
        sacc += __shfl_xor(sacc, 1);
        if (h == 0 && p0 + p < n) phi[p0 + p] = sacc + bo;
    }
}

constexpr size_t INFER_LDS = (size_t)ITILE * LDW * sizeof(u16);

// ---- generic GEMM for the backward pass ---------------------------------------------------------
// C[M x N] = op(A)[M x K] op(B)[K x N]; op(A) = A (A stored M x K, row stride lda) or A^T (A stored
// K x M); op(B) = B (stored K x N) or B^T (stored N x K).  M, N multiples of 64, K of 32.
// EPI 0: C32 (+)= result (atomic when the K range is split over gridDim.z)
// EPI 1: C16 = bf16(result * silu'(aux))      (backward through a SiLU layer)
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const u16 *__restrict__ A, int lda, const u16 *__restrict__ B,
                                              int ldb, float *__restrict__ C32, u16 *__restrict__ C16,
                                              int ldc, int M, int N, int K, const float *__restrict__ aux,
                                              int ldaux) {
    __shared__ __attribute__((aligned(16))) u16 As[TILE][KS + 8];  // [m][k]
    __shared__ __attribute__((aligned(16))) u16 Bs[TILE][KS + 8];  // [n][k]
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int m0 = blockIdx.y * TILE, n0 = blockIdx.x * TILE;
    const int kchunk = K / gridDim.z;
    const int kb = blockIdx.z * kchunk, ke = kb + kchunk;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += KS) {
        // stage a 64 x 32 tile of op(A) and of op(B)^T: 2048 elements each, 8 per thread
        {
            const int r = t >> 2, c8 = (t & 3) * 8;  // non-transposed: row r, k c8..c8+7
            if (!TA) {
                *reinterpret_cast<bf16x8 *>(&As[r][c8]) = load8(A + (size_t)(m0 + r) * lda + k0 + c8);
            } else {  // A stored K x M: 8 consecutive m for one k
                const int kk = t >> 3, mm = (t & 7) * 8;
                const bf16x8 v = load8(A + (size_t)(k0 + kk) * lda + m0 + mm);
                const u16 *vs = reinterpret_cast<const u16 *>(&v);
#pragma unroll
                for (int q = 0; q < 8; ++q) As[mm + q][kk] = vs[q];
            }
            if (TB) {  // B stored N x K
                *reinterpret_cast<bf16x8 *>(&Bs[r][c8]) = load8(B + (size_t)(n0 + r) * ldb + k0 + c8);
            } else {  // B stored K x N: 8 consecutive n for one k
                const int kk = t >> 3, nn = (t & 7) * 8;
                const bf16x8 v = load8(B + (size_t)(k0 + kk) * ldb + n0 + nn);
                const u16 *vs = reinterpret_cast<const u16 *>(&v);
#pragma unroll
                for (int q = 0; q < 8; ++q) Bs[nn + q][kk] = vs[q];
            }
        }
        __syncthreads();
        const int kk = 8 * (lane >> 4);
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = load8(&As[wm + 16 * i + (lane & 15)][kk]);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = load8(&Bs[wn + 16 * j + (lane & 15)][kk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + 16 * j + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const float v = acc[i][j][r];
                if (EPI == 0) {
                    if (gridDim.z > 1) atomicAdd(&C32[(size_t)row * ldc + col], v);
                    else C32[(size_t)row * ldc + col] = v;
                } else {
                    C16[(size_t)row * ldc + col] = f2bf(v * dsilu(aux[(size_t)row * ldaux + col]));
                }
            }
        }
}

// Output layer backward and loss, in three coalesced passes:
//   k_out_dphi: dphi = 2 (phi - y) / n per point; loss += sum (phi - y)^2 / n and db_out += sum dphi,
//               one atomic per workgroup;
//   k_out_dz:   dZ_last = bf16((dphi w_out) silu'(z_last)) elementwise;
//   k_colsum (weighted by dphi): dw_out = sum_p dphi a_last.
__global__ __launch_bounds__(256) void k_out_dphi(const float *__restrict__ phi, const float *__restrict__ y,
                                                  int n, float *__restrict__ dphi, float *__restrict__ g_bout,
                                                  float *__restrict__ loss) {
    __shared__ float sl[4], sd[4];
    const int p = blockIdx.x * 256 + threadIdx.x;
    float r2 = 0.0f, dp = 0.0f;
    if (p < n) {
        const float r = phi[p] - y[p];
        dp = 2.0f * r / (float)n;
        dphi[p] = dp;
        r2 = r * r / (float)n;
    }
    for (int o = 32; o > 0; o >>= 1) {
        r2 += __shfl_xor(r2, o);
        dp += __shfl_xor(dp, o);
    }
    if ((threadIdx.x & 63) == 0) {
        sl[threadIdx.x >> 6] = r2;
        sd[threadIdx.x >> 6] = dp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(loss, (sl[0] + sl[1]) + (sl[2] + sl[3]));
        atomicAdd(g_bout, (sd[0] + sd[1]) + (sd[2] + sd[3]));
    }
}

__global__ __launch_bounds__(256) void k_out_dz(const float *__restrict__ dphi, int n, int W,
                                                const float *__restrict__ wout, const float *__restrict__ zlast,
                                                u16 *__restrict__ dz) {
    const size_t o = blockIdx.x * (size_t)256 + threadIdx.x;
    if (o >= (size_t)n * W) return;
    const int p = (int)(o / W), k = (int)(o - (size_t)p * W);
    dz[o] = f2bf(dphi[p] * bf2f(f2bf(wout[k])) * dsilu(zlast[o]));
}

// db[col] += sum over rows of w[row] * dz[row][col] (w == nullptr: weight 1), W == WIDTH.  A
// workgroup takes COLSUM_ROWS rows: each thread two adjacent columns (one 32-bit load) and every
// fourth row, eight rows in flight; the four row phases meet in LDS; one atomic per column.
constexpr int COLSUM_ROWS = 64;
__global__ __launch_bounds__(512) void k_colsum(const u16 *__restrict__ dz, const float *__restrict__ w, int n,
                                                float *__restrict__ db) {
    __shared__ float part[4][WIDTH];
    const int t = threadIdx.x, c2 = (t & 127) * 2, ph = t >> 7;
    const int r0 = blockIdx.x * COLSUM_ROWS;
    float s0 = 0.0f, s1 = 0.0f;
#pragma unroll 8
    for (int q = 0; q < COLSUM_ROWS / 4; ++q) {
        const int r = r0 + ph + 4 * q;
        if (r < n) {
            const uint32_t v = *reinterpret_cast<const uint32_t *>(dz + (size_t)r * WIDTH + c2);
            const float wr = w ? w[r] : 1.0f;
            s0 += wr * bf2f((u16)(v & 0xffffu));
            s1 += wr * bf2f((u16)(v >> 16));
        }
    }
    part[ph][c2] = s0;
    part[ph][c2 + 1] = s1;
    __syncthreads();
    if (t < WIDTH) atomicAdd(&db[t], (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]));
}

// Adam on the FP32 master parameters; refreshes the BF16 transposed copy of a weight matrix.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m1, float *__restrict__ m2, int count,
                                              float lr, float b1, float b2, float eps, float c1, float c2,
                                              u16 *__restrict__ wt, int K, int N) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
    if (wt) {  // p is W[K][N] row-major; wt is [N][K]
        const int k = i / N, nn = i - k * N;
        wt[(size_t)nn * K + k] = f2bf(v);
    }
}

__global__ void k_to_bf16_t(const float *__restrict__ w, u16 *__restrict__ wt, int K, int N) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= K * N) return;
    const int k = i / N, nn = i - k * N;
    wt[(size_t)nn * K + k] = f2bf(w[i]);
}

// Exact samples for training: phi of the body at given points (FP64 evaluation, stored FP32).
__global__ void k_exact_points(cc::Geo geo, const float *__restrict__ x, int n, float *__restrict__ y) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    y[i] = (float)cc::sdf_value<double, 3>(geo, p);
}

}  // namespace

// ---- the model object -----------------------------------------------------------------------------
struct cc_mlp {
    int F = 64, width = 256, depth = 4;
    int device = 0;  // the GPU the model lives on (current when it was created)
    float sigma = 8.0f;
    float in_lo = 0.0f, in_scale = 1.0f;  // set by training: [lo, hi]^3 onto the unit box
    // parameter blocks (FP32 master): W[l] (K_l x width), b[l] (width), wout (width), bout (1)
    std::vector<int> K;
    std::vector<size_t> off_w, off_b;
    size_t off_wout = 0, off_bout = 0, nparam = 0;
    float *ffB = nullptr;     // device [F][3]
    float *params = nullptr;  // device, nparam
    float *grads = nullptr;   // device, nparam
    float *m1 = nullptr, *m2 = nullptr;
    u16 *wt[8] = {};          // device BF16 transposed weights per layer
    int step = 0;
    std::vector<float> ffB_host;
    // training scratch, sized for the batch
    int cap = 0;
    float *x = nullptr, *y = nullptr, *phi = nullptr, *zs = nullptr, *loss = nullptr, *dphi = nullptr;
    u16 *feat = nullptr, *as = nullptr, *dz0 = nullptr, *dz1 = nullptr;
    int ecap = 0;  // inference scratch
    float *ex = nullptr, *ephi = nullptr;
    hipStream_t stream = nullptr;
};

namespace {

int nfail(int code, const std::string &msg) {
    cc_internal_set_error(msg.c_str());
    return code;
}

MlpDev dev_view(const cc_mlp *m) {
    MlpDev d;
    d.ffB = m->ffB;
    for (int l = 0; l < 8; ++l) {
        d.Wt[l] = l < m->depth ? m->wt[l] : nullptr;
        d.bias[l] = l < m->depth ? m->params + m->off_b[l] : nullptr;
    }
    d.wout = m->params + m->off_wout;
    d.bout = m->params + m->off_bout;
    d.in_lo = m->in_lo;
    d.in_scale = m->in_scale;
    d.F = m->F;
    d.width = m->width;
    d.depth = m->depth;
    return d;
}

template <typename P> void release(P *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int ensure_cap(cc_mlp *m, int n) {
    if (n <= m->cap) return CC_OK;
    release(m->x); release(m->y); release(m->phi); release(m->zs); release(m->feat); release(m->as);
    release(m->dz0); release(m->dz1); release(m->loss); release(m->dphi);
    m->cap = 0;
    const size_t W = m->width, D = m->depth, F2 = 2 * m->F;
    if (hipMalloc(&m->x, 3 * sizeof(float) * n) || hipMalloc(&m->y, sizeof(float) * n) ||
        hipMalloc(&m->phi, sizeof(float) * n) || hipMalloc(&m->zs, sizeof(float) * D * n * W) ||
        hipMalloc(&m->feat, sizeof(u16) * n * F2) || hipMalloc(&m->as, sizeof(u16) * D * n * W) ||
        hipMalloc(&m->dz0, sizeof(u16) * n * W) || hipMalloc(&m->dz1, sizeof(u16) * n * W) ||
        hipMalloc(&m->loss, 64) || hipMalloc(&m->dphi, sizeof(float) * n))
        return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate training buffers");
    m->cap = n;
    return CC_OK;
}

// Forward + backward over n points already in m->x / m->y (n a multiple of 64), enqueued on the
// model's stream: gradients into m->grads (zeroed here), the mean squared error added to *m->loss.
void enqueue_loss_grad(cc_mlp *m, int n) {
    hipStream_t s = m->stream;
    const int W = m->width, D = m->depth, F2 = 2 * m->F;
    (void)hipMemsetAsync(m->grads, 0, sizeof(float) * m->nparam, s);
    hipLaunchKernelGGL((k_mlp_forward<true, false>), dim3(n / TILE), dim3(256), 0, s, dev_view(m), m->x, n,
                       m->phi, m->feat, m->zs, m->as, cc::Lattice{});
    const size_t L = (size_t)n * W;
    hipLaunchKernelGGL(k_out_dphi, dim3((n + 255) / 256), dim3(256), 0, s, m->phi, m->y, n, m->dphi,
                       m->grads + m->off_bout, m->loss);
    hipLaunchKernelGGL(k_out_dz, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, s, m->dphi, n, W,
                       m->params + m->off_wout, m->zs + (D - 1) * L, m->dz0);
    const dim3 gcs((n + COLSUM_ROWS - 1) / COLSUM_ROWS);
    hipLaunchKernelGGL(k_colsum, gcs, dim3(512), 0, s, m->as + (D - 1) * L, (const float *)m->dphi, n,
                       m->grads + m->off_wout);
    u16 *dz = m->dz0, *dzp = m->dz1;
    // split the batch of the dW products into equal chunks of whole K steps: split divides n / KS
    int split = std::max(1, std::min(64, n / 256));
    while ((n / KS) % split) --split;
    for (int l = D - 1; l >= 0; --l) {
        const int K = m->K[l];
        const u16 *ain = l == 0 ? m->feat : m->as + (l - 1) * L;
        const int lda = l == 0 ? F2 : W;
        // dW[l] (K x W) = a_in^T (K x n) . dz (n x W)
        hipLaunchKernelGGL((k_gemm<true, false, 0>), dim3(W / TILE, K / TILE, split), dim3(256), 0, s, ain,
                           lda, dz, W, m->grads + m->off_w[l], (u16 *)nullptr, W, K, W, n,
                           (const float *)nullptr, 0);
        hipLaunchKernelGGL(k_colsum, gcs, dim3(512), 0, s, dz, (const float *)nullptr, n,
                           m->grads + m->off_b[l]);
        if (l > 0) {
            // dz_prev (n x W) = (dz (n x W) . W[l]^T (W x W)) * silu'(z_prev); W[l]^T via wt[l] = [W][K]
            hipLaunchKernelGGL((k_gemm<false, false, 1>), dim3(W / TILE, n / TILE, 1), dim3(256), 0, s, dz,
                               W, m->wt[l], K, (float *)nullptr, dzp, W, n, K, W, m->zs + (l - 1) * L, W);
            std::swap(dz, dzp);
        }
    }
}

// One forward + backward with its own loss, synchronised.
int loss_grad(cc_mlp *m, int n, float *loss_out) {
    (void)hipMemsetAsync(m->loss, 0, 4, m->stream);
    enqueue_loss_grad(m, n);
    float lh = 0.0f;
    (void)hipMemcpyAsync(&lh, m->loss, 4, hipMemcpyDeviceToHost, m->stream);
    if (hipStreamSynchronize(m->stream) != hipSuccess)
        return nfail(CC_ERR_HIP, std::string("surrogate step: ") + hipGetErrorString(hipGetLastError()));
    if (loss_out) *loss_out = lh;
    return CC_OK;
}

// Gathers minibatch rows idx[0..n) of the device-resident samples into the step buffers.
__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ Y, const uint32_t *__restrict__ idx,
                         int n, float *__restrict__ x, float *__restrict__ y) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = idx[i];
    x[3 * i] = X[3 * (size_t)s];
    x[3 * i + 1] = X[3 * (size_t)s + 1];
    x[3 * i + 2] = X[3 * (size_t)s + 2];
    y[i] = Y[s];
}

void adam_step(cc_mlp *m, float lr) {
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    for (int l = 0; l < m->depth; ++l) {
        const int cnt = m->K[l] * m->width;
        hipLaunchKernelGGL(k_adam, dim3((cnt + 255) / 256), dim3(256), 0, m->stream, m->params + m->off_w[l],
                           m->grads + m->off_w[l], m->m1 + m->off_w[l], m->m2 + m->off_w[l], cnt, lr, b1,
                           b2, eps, c1, c2, m->wt[l], m->K[l], m->width);
    }
    const int rest = (int)(m->nparam - m->off_b[0]);  // biases, wout, bout are contiguous after W
    hipLaunchKernelGGL(k_adam, dim3((rest + 255) / 256), dim3(256), 0, m->stream, m->params + m->off_b[0],
                       m->grads + m->off_b[0], m->m1 + m->off_b[0], m->m2 + m->off_b[0], rest, lr, b1, b2,
                       eps, c1, c2, (u16 *)nullptr, 0, 0);
}

}  // namespace

namespace cc {

int mlp_device(const cc_mlp *m) { return m->device; }

int launch_mlp_lattice(const cc_mlp *m, const Lattice &lat, float *out, hipStream_t s) {
    const long long n = (long long)lat.n[0] * lat.n[1] * lat.n[2];
    if (n < 1 || n > (1ll << 31) - TILE) return CC_ERR_ARGUMENT;
    hipLaunchKernelGGL((k_mlp_infer<true>), dim3((unsigned)((n + ITILE - 1) / ITILE)), dim3(256), INFER_LDS, s,
                       dev_view(m), (const float *)nullptr, (int)n, out, lat);
    return CC_OK;
}

}  // namespace cc

// ---- C ABI ---------------------------------------------------------------------------------------
extern "C" {

int cc_mlp_create(int nfreq, int width, int depth, float sigma, uint64_t seed, cc_mlp **out) {
    if (!out) return nfail(CC_ERR_ARGUMENT, "NULL output");
    *out = nullptr;
    if (width != WIDTH || depth < 1 || depth > 8 || nfreq < 32 || nfreq > 128 || nfreq % 32 != 0)
        return nfail(CC_ERR_UNSUPPORTED,
                     "the surrogate needs width 256, depth 1..8 and nfreq in {32, 64, 96, 128}");
    cc_mlp *m = new cc_mlp();
    m->F = nfreq;
    m->width = width;
    m->depth = depth;
    m->sigma = sigma;
    size_t off = 0;
    for (int l = 0; l < depth; ++l) {
        m->K.push_back(l == 0 ? 2 * nfreq : width);
        m->off_w.push_back(off);
        off += (size_t)m->K[l] * width;
    }
    for (int l = 0; l < depth; ++l) {
        m->off_b.push_back(off);
        off += width;
    }
    m->off_wout = off;
    off += width;
    m->off_bout = off;
    off += 1;
    m->nparam = off;
    // initialisation: Fourier frequencies N(0, sigma^2); weights uniform(+-sqrt(6/(fan_in+fan_out)))
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, sigma);
    m->ffB_host.resize(3 * nfreq);
    for (auto &v : m->ffB_host) v = nd(rng);
    std::vector<float> p(m->nparam, 0.0f);
    for (int l = 0; l < depth; ++l) {
        const float a = std::sqrt(6.0f / (m->K[l] + width));
        std::uniform_real_distribution<float> ud(-a, a);
        for (size_t i = 0; i < (size_t)m->K[l] * width; ++i) p[m->off_w[l] + i] = ud(rng);
    }
    {
        const float a = std::sqrt(6.0f / (width + 1));
        std::uniform_real_distribution<float> ud(-a, a);
        for (int i = 0; i < width; ++i) p[m->off_wout + i] = ud(rng);
    }
    // the inference kernel's activation tile exceeds the default 64 KB of dynamic LDS
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp_infer<true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)INFER_LDS) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp_infer<false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)INFER_LDS) != hipSuccess ||
        hipGetDevice(&m->device) != hipSuccess) {
        delete m;
        return nfail(CC_ERR_HIP, std::string("surrogate kernel attributes: ") + hipGetErrorString(hipGetLastError()));
    }
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return nfail(CC_ERR_HIP, "surrogate stream");
    }
    bool ok = hipMalloc(&m->ffB, 12 * nfreq) == hipSuccess && hipMalloc(&m->params, 4 * m->nparam) == hipSuccess &&
              hipMalloc(&m->grads, 4 * m->nparam) == hipSuccess && hipMalloc(&m->m1, 4 * m->nparam) == hipSuccess &&
              hipMalloc(&m->m2, 4 * m->nparam) == hipSuccess;
    for (int l = 0; ok && l < depth; ++l) ok = hipMalloc(&m->wt[l], 2 * (size_t)m->K[l] * width) == hipSuccess;
    if (!ok) {
        cc_mlp_destroy(m);
        return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate parameters");
    }
    // everything ordered on the model's (non-blocking) stream
    ok = hipMemcpyAsync(m->ffB, m->ffB_host.data(), 12 * nfreq, hipMemcpyHostToDevice, m->stream) == hipSuccess &&
         hipMemcpyAsync(m->params, p.data(), 4 * m->nparam, hipMemcpyHostToDevice, m->stream) == hipSuccess &&
         hipMemsetAsync(m->m1, 0, 4 * m->nparam, m->stream) == hipSuccess &&
         hipMemsetAsync(m->m2, 0, 4 * m->nparam, m->stream) == hipSuccess;
    for (int l = 0; ok && l < depth; ++l) {
        const int cnt = m->K[l] * width;
        hipLaunchKernelGGL(k_to_bf16_t, dim3((cnt + 255) / 256), dim3(256), 0, m->stream,
                           m->params + m->off_w[l], m->wt[l], m->K[l], width);
    }
    if (!ok || hipStreamSynchronize(m->stream) != hipSuccess) {
        const std::string msg = std::string("surrogate initialisation: ") + hipGetErrorString(hipGetLastError());
        cc_mlp_destroy(m);
        return nfail(CC_ERR_HIP, msg);
    }
    *out = m;
    return CC_OK;
}

int cc_mlp_destroy(cc_mlp *m) {
    if (!m) return CC_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (auto p : {(void *)m->ffB, (void *)m->params, (void *)m->grads, (void *)m->m1, (void *)m->m2,
                   (void *)m->x, (void *)m->y, (void *)m->phi, (void *)m->zs, (void *)m->feat,
                   (void *)m->as, (void *)m->dz0, (void *)m->dz1, (void *)m->loss, (void *)m->ex,
                   (void *)m->ephi, (void *)m->dphi})
        if (p) (void)hipFree(p);
    for (auto w : m->wt)
        if (w) (void)hipFree(w);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return CC_OK;
}

int cc_mlp_num_params(cc_mlp *m, int64_t *n, int64_t *nfreq3) {
    if (!m) return nfail(CC_ERR_ARGUMENT, "NULL model");
    if (n) *n = (int64_t)m->nparam;
    if (nfreq3) *nfreq3 = 3 * m->F;
    return CC_OK;
}

int cc_mlp_get_params(cc_mlp *m, float *params, float *ffB) {
    if (!m) return nfail(CC_ERR_ARGUMENT, "NULL model");
    (void)hipSetDevice(m->device);
    if (params) {
        (void)hipMemcpyAsync(params, m->params, 4 * m->nparam, hipMemcpyDeviceToHost, m->stream);
        if (hipStreamSynchronize(m->stream) != hipSuccess)
            return nfail(CC_ERR_HIP, hipGetErrorString(hipGetLastError()));
    }
    if (ffB) std::memcpy(ffB, m->ffB_host.data(), 12 * m->F);
    return CC_OK;
}

int cc_mlp_set_params(cc_mlp *m, const float *params) {
    if (!m || !params) return nfail(CC_ERR_ARGUMENT, "NULL argument");
    (void)hipSetDevice(m->device);
    (void)hipMemcpyAsync(m->params, params, 4 * m->nparam, hipMemcpyHostToDevice, m->stream);
    for (int l = 0; l < m->depth; ++l) {
        const int cnt = m->K[l] * m->width;
        hipLaunchKernelGGL(k_to_bf16_t, dim3((cnt + 255) / 256), dim3(256), 0, m->stream,
                           m->params + m->off_w[l], m->wt[l], m->K[l], m->width);
    }
    (void)hipStreamSynchronize(m->stream);
    return CC_OK;
}

// phi_nn at n host points (x: n x 3 floats) into y (host).
int cc_mlp_eval(cc_mlp *m, const float *x, int64_t n, float *y) {
    if (!m || !x || !y || n < 1 || n > (int64_t(1) << 30)) return nfail(CC_ERR_ARGUMENT, "bad argument");
    (void)hipSetDevice(m->device);
    const int np = (int)((n + TILE - 1) / TILE * TILE);
    if (np > m->ecap) {
        release(m->ex);
        release(m->ephi);
        m->ecap = 0;
        if (hipMalloc(&m->ex, 12 * (size_t)np) || hipMalloc(&m->ephi, 4 * (size_t)np))
            return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate evaluation buffers");
        m->ecap = np;
    }
    (void)hipMemcpyAsync(m->ex, x, 12 * n, hipMemcpyHostToDevice, m->stream);
    hipLaunchKernelGGL((k_mlp_infer<false>), dim3((unsigned)((n + ITILE - 1) / ITILE)), dim3(256), INFER_LDS,
                       m->stream, dev_view(m), m->ex, (int)n, m->ephi, cc::Lattice{});
    (void)hipMemcpyAsync(y, m->ephi, 4 * n, hipMemcpyDeviceToHost, m->stream);
    if (hipStreamSynchronize(m->stream) != hipSuccess)
        return nfail(CC_ERR_HIP, hipGetErrorString(hipGetLastError()));
    return CC_OK;
}

// Mean squared error and its gradient w.r.t. every parameter at n host samples (x, y); n a multiple
// of 64.  For checking the backward kernels against an independent implementation.
int cc_mlp_loss_grad(cc_mlp *m, const float *x, const float *y, int64_t n, float *loss, float *grad) {
    if (!m || !x || !y || n < 64 || n % 64 || n > (int64_t(1) << 24))
        return nfail(CC_ERR_ARGUMENT, "n must be a multiple of 64 in [64, 2^24]");
    (void)hipSetDevice(m->device);
    int rc = ensure_cap(m, (int)n);
    if (rc) return rc;
    (void)hipMemcpyAsync(m->x, x, 12 * n, hipMemcpyHostToDevice, m->stream);
    (void)hipMemcpyAsync(m->y, y, 4 * n, hipMemcpyHostToDevice, m->stream);
    rc = loss_grad(m, (int)n, loss);
    if (rc) return rc;
    if (grad) {
        (void)hipMemcpyAsync(grad, m->grads, 4 * m->nparam, hipMemcpyDeviceToHost, m->stream);
        if (hipStreamSynchronize(m->stream) != hipSuccess)
            return nfail(CC_ERR_HIP, hipGetErrorString(hipGetLastError()));
    }
    return CC_OK;
}

// Train on exact samples of `geo`: `nsamples` points (half uniform in [lo, hi]^3, half within
// `band` of the surface), minibatches of `batch`, `epochs` passes, Adam with a cosine-decayed
// learning rate.  All work on the device; `history` (optional, epochs floats) receives the mean
// training loss of each epoch.
int cc_mlp_train(cc_mlp *m, const cc_geometry *geo, int64_t nsamples, float lo, float hi, float band,
                 int epochs, int batch, float lr, uint64_t seed, float *history) {
    if (!m || !geo || nsamples < batch || batch < 64 || batch % 64 || batch > (1 << 20) || epochs < 1 ||
        nsamples > (int64_t(1) << 30) || !(band > 0.0f))
        return nfail(CC_ERR_ARGUMENT, "bad training arguments");
    (void)hipSetDevice(m->device);
    if (geo->kind < CC_GEO_EMPTY || geo->kind > CC_GEO_SPHERE_UNION)
        return nfail(CC_ERR_UNSUPPORTED, "the surrogate trains on an unbounded body (sphere, torus, box, plane, "
                                         "sphere union)");
    if (geo->kind == CC_GEO_SPHERE_UNION && (!geo->prims || geo->n_prims < 1))
        return nfail(CC_ERR_ARGUMENT, "sphere union without primitives");
    if (!(hi > lo)) return nfail(CC_ERR_ARGUMENT, "the training box needs hi > lo");
    // the Fourier features see the training box as the unit box (sigma is quoted for unit inputs)
    m->in_lo = lo;
    m->in_scale = 1.0f / (hi - lo);
    cc::Geo g{};
    g.kind = geo->kind;
    std::memcpy(g.p, geo->p, sizeof g.p);
    double *prims = nullptr;
    if (geo->kind == CC_GEO_SPHERE_UNION) {
        if (hipMalloc(&prims, 32 * (size_t)geo->n_prims) != hipSuccess)
            return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate training: primitive table");
        if (hipMemcpy(prims, geo->prims, 32 * (size_t)geo->n_prims, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(prims);
            return nfail(CC_ERR_HIP, "surrogate training: primitive table copy");
        }
        g.nprims = geo->n_prims;
        g.prims = prims;
    }
    // samples: uniform half directly; near-surface half by rejection from uniform candidates
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> ud(lo, hi);
    const size_t N = (size_t)nsamples;
    std::vector<float> X(3 * N), Y(N);
    const size_t chunk = 1 << 20;
    std::vector<float> cx(3 * chunk), cy(chunk);
    float *dcx = nullptr, *dcy = nullptr;
    if (hipMalloc(&dcx, 12 * chunk) != hipSuccess || hipMalloc(&dcy, 4 * chunk) != hipSuccess) {
        (void)hipFree(dcx);
        if (prims) (void)hipFree(prims);
        return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate training: sample buffers");
    }
    size_t nu = 0, nn = 0, half = N / 2;
    int guard = 0;
    while ((nu < half || nn < N - half) && guard < 4096) {
        ++guard;
        for (auto &v : cx) v = ud(rng);
        // every copy is ordered on the model's (non-blocking) stream with the kernel it feeds
        (void)hipMemcpyAsync(dcx, cx.data(), 12 * chunk, hipMemcpyHostToDevice, m->stream);
        hipLaunchKernelGGL(k_exact_points, dim3(chunk / 256), dim3(256), 0, m->stream, g, dcx, (int)chunk, dcy);
        (void)hipMemcpyAsync(cy.data(), dcy, 4 * chunk, hipMemcpyDeviceToHost, m->stream);
        if (hipStreamSynchronize(m->stream) != hipSuccess) {
            (void)hipFree(dcx);
            (void)hipFree(dcy);
            if (prims) (void)hipFree(prims);
            return nfail(CC_ERR_HIP, std::string("surrogate sampling: ") + hipGetErrorString(hipGetLastError()));
        }
        for (size_t i = 0; i < chunk; ++i) {
            size_t slot;
            if (nu < half) slot = nu++;
            else if (std::fabs(cy[i]) < band && nn < N - half) slot = half + nn++;
            else continue;
            X[3 * slot] = cx[3 * i];
            X[3 * slot + 1] = cx[3 * i + 1];
            X[3 * slot + 2] = cx[3 * i + 2];
            Y[slot] = cy[i];
        }
    }
    (void)hipFree(dcx);
    (void)hipFree(dcy);
    if (prims) (void)hipFree(prims);
    if (nn < N - half) return nfail(CC_ERR_ARGUMENT, "could not find enough near-surface samples");
    int rc = ensure_cap(m, batch);
    if (rc) return rc;
    // the samples stay on the device; each epoch uploads one permutation and syncs once
    float *dX = nullptr, *dY = nullptr;
    uint32_t *dperm = nullptr;
    if (hipMalloc(&dX, 12 * N) || hipMalloc(&dY, 4 * N) || hipMalloc(&dperm, 4 * N)) {
        (void)hipFree(dX);
        (void)hipFree(dY);
        return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate training samples");
    }
    (void)hipMemcpyAsync(dX, X.data(), 12 * N, hipMemcpyHostToDevice, m->stream);
    (void)hipMemcpyAsync(dY, Y.data(), 4 * N, hipMemcpyHostToDevice, m->stream);
    std::vector<uint32_t> perm(N);
    for (size_t i = 0; i < N; ++i) perm[i] = (uint32_t)i;
    const size_t nbatch = N / batch;
    const size_t total_steps = nbatch * epochs;
    size_t gstep = 0;
    for (int e = 0; e < epochs && rc == CC_OK; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        (void)hipMemcpyAsync(dperm, perm.data(), 4 * N, hipMemcpyHostToDevice, m->stream);
        (void)hipMemsetAsync(m->loss, 0, 4, m->stream);
        for (size_t b = 0; b < nbatch; ++b, ++gstep) {
            hipLaunchKernelGGL(k_gather, dim3((batch + 255) / 256), dim3(256), 0, m->stream, dX, dY,
                               dperm + b * batch, batch, m->x, m->y);
            enqueue_loss_grad(m, batch);
            const float lrt = 0.5f * lr * (1.0f + std::cos(3.14159265f * (float)gstep / (float)total_steps));
            adam_step(m, lrt);
        }
        float l = 0.0f;
        (void)hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, m->stream);
        if (hipStreamSynchronize(m->stream) != hipSuccess)
            rc = nfail(CC_ERR_HIP, std::string("surrogate training: ") + hipGetErrorString(hipGetLastError()));
        else if (history)
            history[e] = l / (float)nbatch;
    }
    (void)hipFree(dX);
    (void)hipFree(dY);
    (void)hipFree(dperm);
    if (rc) return rc;
    if (hipStreamSynchronize(m->stream) != hipSuccess)
        return nfail(CC_ERR_HIP, hipGetErrorString(hipGetLastError()));
    return CC_OK;
}

}  // extern "C"
