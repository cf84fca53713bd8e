// The learned surrogate for an expensive body's signed distance: a Fourier-feature MLP
// (3 -> 2F sin/cos features -> depth x width SiLU -> 1), BF16 operands on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, FP32 master weights, trained with Adam by the
// kernels in this file.  No counterpart in the reference (whose bodies are SDFLibrary closures).
//
// Kernels:
//   k_mlp_infer     -- evaluation: a 128-point tile through every layer, activations kept in LDS.
//   k_train_fwd     -- the training forward of a 64-point minibatch tile, gathered through the
//                      epoch's permutation; saves, feature-major, what the backward needs and ends
//                      the pass on the device (loss, dphi, the last layer's dz and column sums).
//   k_train_bwd     -- one layer of the backward pass in one launch: the weight gradient as
//                      partial 64x64 tiles over point chunks (no atomics) beside dz of the layer
//                      below with its bias column sums.
//   k_adam_all      -- Adam over the whole parameter vector in one launch.
// A training step is 2 + depth launches, captured once in a graph and replayed; the step index
// (minibatch offset, learning rate, bias corrections) is read on the device.
//
// MFMA 16x16x32 BF16 operand layout (wave64): lane l holds, for A (16 x 32, M x K), row l%16 and the
// 8 consecutive k = 8*(l/16) .. +7; for B (32 x 16, K x N), column l%16 and the same 8 k; the FP32
// result (16 x 16) gives lane l column l%16 and rows 4*(l/16) .. +3.
#include <hip/hip_bf16.h>
#include <hipcub/hipcub.hpp>

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

// ---- inference forward --------------------------------------------------------------------------
// The network for evaluation only, on 128-point tiles, with the matrix-core operands swapped:
// A = the transposed weights (rows = output features, 8 consecutive k per lane, read from L2 with
// the next K step's fragments loaded while the current step's MFMAs run), B = the activation tile
// (columns = points).  Each wave owns 64 output features for all 128 points (4 x 8 MFMA tiles), and
// its accumulators hold 4 consecutive features of one point per lane, so the epilogue -- bias,
// SiLU with the hardware reciprocal, one packed BF16 conversion -- ends in one 8-byte LDS store per
// four values.  The activation tile is [ITILE][LDW] BF16 (67.6 KB), updated in place between
// layers.  Points are either explicit (x: n x 3) or, for LAT, the nodes of a lattice: point p is node
// (p % lat.n[0], p / lat.n[0] % lat.n[1], p / (lat.n[0] lat.n[1])) at lat.x0 + index * lat.h.
// 64-point tiles (three blocks per CU instead of two) were slower: 0.33 against 0.28 ms on the 66^3
// lattice, each weight fragment then feeding half as many MFMAs.
constexpr int ITILE = 128;
constexpr int INJ = ITILE / 16;  // 16-point column tiles of a wave's accumulators
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

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
        f32x4 acc[4][INJ];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < INJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        const u16 *Wt = m.Wt[l];
        // weight fragments two K steps ahead: one step's 32 MFMAs do not cover an L2 round trip
        bf16x8 a[4], an[4], an2[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = load8(Wt + (size_t)(arow + 16 * i) * K + koff);
            an[i] = load8(Wt + (size_t)(arow + 16 * i) * K + (KS < K ? KS : 0) + koff);
        }
        for (int k0 = 0; k0 < K; k0 += KS) {
            const int kn2 = k0 + 2 * KS < K ? k0 + 2 * KS : k0;  // (past the end: reload this step's own)
#pragma unroll
            for (int i = 0; i < 4; ++i) an2[i] = load8(Wt + (size_t)(arow + 16 * i) * K + kn2 + koff);
#pragma unroll
            for (int j = 0; j < INJ; ++j) {
                const bf16x8 b = load8(ibuf + (16 * j + (lane & 15)) * LDin + k0 + koff);
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[i][j] = mfma(a[i], b, acc[i][j]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = an[i];
                an[i] = an2[i];
            }
        }
        __syncthreads();  // every wave has read the whole input tile
        const float *bias = m.bias[l];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int f0 = 64 * wave + 16 * i + 4 * (lane >> 4);
            const f32x4 bv = *reinterpret_cast<const f32x4 *>(bias + f0);
#pragma unroll
            for (int j = 0; j < INJ; ++j) {
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
    // output layer: two lanes per point, 128 of the 256 k each (the first 2 ITILE threads; a whole
    // number of waves)
    const float bo = *m.bout;
    if (t < 2 * ITILE) {
        const int p = t >> 1, h = t & 1;
        float sacc = 0.0f;
        for (int k = h; k < WIDTH; k += 2) sacc += bf2f(ibuf[p * LDW + k]) * wo[k];
        sacc += __shfl_xor(sacc, 1);
        if (h == 0 && p0 + p < n) phi[p0 + p] = sacc + bo;
    }
}

constexpr size_t INFER_LDS = (size_t)ITILE * LDW * sizeof(u16);

// ---- training -----------------------------------------------------------------------------------
// Minibatch order: epoch e visits the samples in the order of a keyed pseudo-random permutation of
// [0, N) (N <= 2^30), a six-round balanced Feistel network on 2 hb bits (2^(2 hb) >= N) with cycle
// walking.  The walk ends: it follows x's cycle of a bijection, on which x itself is below N.  The
// round function is the "lowbias32" integer hash (C. Wellons, hash-prospector, public domain).
__device__ __forceinline__ uint32_t hash32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

__device__ __forceinline__ uint32_t shuffle_index(uint32_t x, uint32_t N, int hb, const uint32_t *__restrict__ key) {
    const uint32_t mask = (1u << hb) - 1u;
    do {
        uint32_t L = x >> hb, R = x & mask;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const uint32_t f = hash32(R ^ key[r]) & mask;
            const uint32_t nl = R;
            R = L ^ f;
            L = nl;
        }
        x = (L << hb) | R;
    } while (x >= N);
    return x;
}

// Saved state of a minibatch of n points (n a multiple of 64), laid out for the matrix cores: the
// weight gradient dW_l = a_{l-1}^T dz_l reduces over points, so its operands are kept feature-major
// ([feature][point], 8 consecutive points per 16-byte fragment load); dz_{l-1} = dz_l W_l^T reduces
// over features, so it reads dz point-major and W_l row-major.
struct TrainBufs {
    u16 *featT;           // [2F][n]        Fourier features (layer 0's input)
    u16 *aT;              // [depth-1][W][n] hidden activations a_l, l < depth - 1
    u16 *ds;              // [depth-1][W][n] silu'(z_l), l < depth - 1
    u16 *dz, *dzT;        // dz of the last layer: [n][W] and [W][n]
    float *gpart;         // [rows][n / 64]: every workgroup's sums, one row per bias, w_out and b_out
                          // entry and one for the loss (no atomics: the optimiser sums the rows)
    int *ctr;             // [0]: step of the next forward, [1]: step of the next Adam
    int rblast, rwout, rbout, rloss;  // rows of the last layer's bias, w_out, b_out, the loss
};

__device__ __forceinline__ void store4(u16 *p, const u16 v[4]) {  // 8-byte aligned
    uint2 w;
    w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
    w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
    *reinterpret_cast<uint2 *>(p) = w;
}

// Two floats to packed BF16 (round to nearest even) with the hardware conversion; the first in the
// low half.
__device__ __forceinline__ uint32_t pk2(float a, float b) {
    const bf16x2 h = __builtin_convertvector(f32x2{a, b}, bf16x2);
    return __builtin_bit_cast(uint32_t, h);
}
__device__ __forceinline__ float lo_bf(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi_bf(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// One workgroup (1024 threads, 16 waves) runs a 64-point tile through every layer.  The activation
// tile lives in two LDS buffers, [TILE][LDW] BF16 each (dynamic LDS, FWD_LDS bytes): layer l reads
// one and writes the other, so a layer needs one barrier; each wave owns FWD_FPW = 16 output
// features (4 x 1 MFMA tiles: lane = one feature, 4 consecutive points), whose weight fragments for
// the whole layer sit in registers, loaded in one burst while the previous layer's epilogue runs.
// Sixteen waves on 64 points keep four waves on every SIMD, so one wave's SiLU epilogue runs beside
// another's matrix products (16 waves measured 5 % faster than 8 on the 16 384-point minibatch).
// Points: X/Y rows shuffle_index(b n + p) under epoch e's keys for step s = ctr[0] = e nb + b when
// keys ([epoch][6]) are given, else rows p.
constexpr int FWD_T = 1024;
constexpr int FWD_NJ = WIDTH / 16 / (FWD_T / 64);  // 16-feature MFMA column tiles per wave
constexpr int FWD_FPW = 16 * FWD_NJ;               // features per wave
constexpr int FWD_LPP = FWD_T / TILE;              // lanes per point in the output layer
constexpr size_t FWD_LDS = 2 * (size_t)TILE * LDW * sizeof(u16);
__global__ __launch_bounds__(FWD_T) void k_train_fwd(MlpDev m, TrainBufs tb, const float *__restrict__ X,
                                                     const float *__restrict__ Y, const uint32_t *__restrict__ keys,
                                                     int hb, uint32_t N, int nb, int n) {
    extern __shared__ __attribute__((aligned(16))) u16 fbuf[];  // 2 x [TILE][LDW]
    __shared__ float pu[TILE][4];  // unit-box coordinates and target of each point
    __shared__ float wo[WIDTH];    // BF16-rounded output weights
    __shared__ float sdp[TILE];    // dphi of each point
    __shared__ float red[2][FWD_T / 64];
    u16 *buf = fbuf, *nxt = fbuf + TILE * LDW;
    const int W = WIDTH, F = m.F, F2 = 2 * F, LDF = F2 + 8;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane & 15;
    const int koff = 8 * (lane >> 4), g4 = 4 * (lane >> 4);
    const int p0 = blockIdx.x * TILE;
    if (t < TILE) {
        size_t q = (size_t)p0 + t;
        if (keys) {
            const int s = tb.ctr[0];
            const int e = s / nb, b = s - e * nb;
            q = shuffle_index((uint32_t)((size_t)b * n + q), N, hb, keys + 6 * e);
            if (blockIdx.x == 0 && t == 0) tb.ctr[1] = s;
        }
        pu[t][0] = (X[3 * q] - m.in_lo) * m.in_scale;
        pu[t][1] = (X[3 * q + 1] - m.in_lo) * m.in_scale;
        pu[t][2] = (X[3 * q + 2] - m.in_lo) * m.in_scale;
        pu[t][3] = Y[q];
    }
    if (t < WIDTH) wo[t] = bf2f(f2bf(m.wout[t]));
    // this wave's weight fragments of a whole layer (FWD_FPW features x K <= 256), loaded in one burst:
    // layer 0's while the features are computed, layer l + 1's while layer l's epilogue runs
    bf16x8 b[8][FWD_NJ];
    auto load_layer = [&](int l, int K) {
        const u16 *wrow = m.Wt[l] + (size_t)(FWD_FPW * wave + lr) * K + koff;
#pragma unroll
        for (int ks = 0; ks < 8; ++ks)
            if (ks * KS < K) {
#pragma unroll
                for (int j = 0; j < FWD_NJ; ++j) b[ks][j] = load8(wrow + 16 * j * (size_t)K + ks * KS);
            }
    };
    load_layer(0, F2);
    __syncthreads();
    // Fourier features [sin(2 pi B u), cos(2 pi B u)]: the hardware sine and cosine take revolutions
    for (int e = t; e < TILE * F; e += FWD_T) {
        const int p = e & (TILE - 1), f = e >> 6;
        const float *b = m.ffB + 3 * f;
        const float rev = (b[0] * pu[p][0] + b[1] * pu[p][1]) + b[2] * pu[p][2];
        const uint32_t sc = pk2(__builtin_amdgcn_sinf(rev), __builtin_amdgcn_cosf(rev));
        const u16 s = (u16)(sc & 0xffffu), c = (u16)(sc >> 16);
        buf[p * LDF + f] = s;
        buf[p * LDF + F + f] = c;
        tb.featT[(size_t)f * n + p0 + p] = s;
        tb.featT[(size_t)(F + f) * n + p0 + p] = c;
    }
    __syncthreads();
    int K = F2, LDin = LDF;
    f32x4 acc[4][FWD_NJ];
    for (int l = 0; l < m.depth; ++l) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < FWD_NJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 8; ++ks)
            if (ks * KS < K) {
                bf16x8 a[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = load8(buf + (16 * i + lr) * LDin + ks * KS + koff);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < FWD_NJ; ++j) acc[i][j] = mfma(a[i], b[ks][j], acc[i][j]);
            }
        if (l + 1 < m.depth) load_layer(l + 1, W);
        const bool last = l == m.depth - 1;
        u16 *aTl = tb.aT + (size_t)l * W * n, *dsl = tb.ds + (size_t)l * W * n;
#pragma unroll
        for (int j = 0; j < FWD_NJ; ++j) {
            const int col = FWD_FPW * wave + 16 * j + lr;
            const float bj = m.bias[l][col];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float av[4], dv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float z = acc[i][j][r] + bj;
                    const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
                    av[r] = z * sg;
                    dv[r] = sg * __builtin_fmaf(z, 1.0f - sg, 1.0f);  // silu'(z)
                    acc[i][j][r] = dv[r];
                }
                const uint2 ap = {pk2(av[0], av[1]), pk2(av[2], av[3])};
                u16 *bp = nxt + (16 * i + g4) * LDW + col;
                bp[0] = (u16)(ap.x & 0xffffu);
                bp[LDW] = (u16)(ap.x >> 16);
                bp[2 * LDW] = (u16)(ap.y & 0xffffu);
                bp[3 * LDW] = (u16)(ap.y >> 16);
                if (!last) {
                    const size_t o = (size_t)col * n + p0 + 16 * i + g4;
                    *reinterpret_cast<uint2 *>(aTl + o) = ap;
                    *reinterpret_cast<uint2 *>(dsl + o) = uint2{pk2(dv[0], dv[1]), pk2(dv[2], dv[3])};
                }
            }
        }
        __syncthreads();
        u16 *const o = buf;
        buf = nxt;
        nxt = o;
        K = W;
        LDin = LDW;
    }
    // output layer, FWD_LPP lanes per point: phi, the residual, dphi = 2 (phi - y) / n
    {
        const int p = t / FWD_LPP;
        float s = 0.0f;
        for (int k = t % FWD_LPP; k < W; k += FWD_LPP) s += bf2f(buf[p * LDW + k]) * wo[k];
        for (int o = 1; o < FWD_LPP; o <<= 1) s += __shfl_xor(s, o);
        const float res = (s + *m.bout) - pu[p][3];
        const float dp = 2.0f * res / (float)n;
        float r2 = 0.0f, dq = 0.0f;
        if (t % FWD_LPP == 0) {
            sdp[p] = dp;
            r2 = res * res / (float)n;
            dq = dp;
        }
        for (int o = 32; o > 0; o >>= 1) {
            r2 += __shfl_xor(r2, o);
            dq += __shfl_xor(dq, o);
        }
        if (lane == 0) {
            red[0][wave] = r2;
            red[1][wave] = dq;
        }
    }
    __syncthreads();
    const size_t nblk = gridDim.x;
    if (t == 0) {
        float a = 0.0f, b = 0.0f;
        for (int w = 0; w < FWD_T / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
        }
        tb.gpart[tb.rloss * nblk + blockIdx.x] = a;
        tb.gpart[tb.rbout * nblk + blockIdx.x] = b;
    }
    // dz of the last layer = bf16(dphi w_out silu'(z)); column sums for its bias and for w_out.
    // Each (point, feature) of the tile belongs to one lane, which replaces a by dz in place.
#pragma unroll
    for (int j = 0; j < FWD_NJ; ++j) {
        const int col = FWD_FPW * wave + 16 * j + lr;
        const float wj = wo[col];
        float sb = 0.0f, sw = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float dv[4];
            u16 *bp = buf + (16 * i + g4) * LDW + col;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float dp = sdp[16 * i + g4 + r];
                dv[r] = dp * wj * acc[i][j][r];
                sw += bf2f(bp[r * LDW]) * dp;
            }
            const uint2 dk = {pk2(dv[0], dv[1]), pk2(dv[2], dv[3])};
            sb += (lo_bf(dk.x) + hi_bf(dk.x)) + (lo_bf(dk.y) + hi_bf(dk.y));
            bp[0] = (u16)(dk.x & 0xffffu);
            bp[LDW] = (u16)(dk.x >> 16);
            bp[2 * LDW] = (u16)(dk.y & 0xffffu);
            bp[3 * LDW] = (u16)(dk.y >> 16);
            *reinterpret_cast<uint2 *>(tb.dzT + (size_t)col * n + p0 + 16 * i + g4) = dk;
        }
        sb += __shfl_xor(sb, 16);
        sb += __shfl_xor(sb, 32);
        sw += __shfl_xor(sw, 16);
        sw += __shfl_xor(sw, 32);
        if (lane < 16) {
            tb.gpart[(tb.rblast + col) * nblk + blockIdx.x] = sb;
            tb.gpart[(tb.rwout + col) * nblk + blockIdx.x] = sw;
        }
    }
    __syncthreads();
    for (int c = t; c < TILE * W / 8; c += FWD_T) {  // the point-major copy, 16 bytes a lane
        const int row = c >> 5, c8 = (c & 31) * 8;
        *reinterpret_cast<bf16x8 *>(tb.dz + (size_t)(p0 + row) * W + c8) = load8(buf + row * LDW + c8);
    }
}

// One layer l of the backward pass.  Workgroups [0, nA): the weight gradient, one 64x64 tile of
// dW_l (features i of the layer's input x features j of its output) over one of S equal point
// chunks; the four waves take interleaved K steps of the chunk and meet in LDS, and the tile is
// written to its chunk's slice of part (the optimiser sums the slices in a fixed order).
// Workgroups [nA, nA + n / 64) when l > 0: dz_{l-1} = bf16((dz_l W_l^T) silu'(z_{l-1})) for 64
// points (each wave 64 features), written in both layouts, with the bias column sums of layer l-1.
struct BwdLayer {
    const u16 *ainT;   // [K][n] the layer's input, feature-major
    const u16 *dz;     // [n][W] dz_l
    const u16 *dzT;    // [W][n] dz_l
    const u16 *w16;    // [K][W] W_l row-major (l > 0)
    const u16 *ds;     // [W][n] silu'(z_{l-1}) (l > 0)
    u16 *dzp, *dzpT;   // dz_{l-1} (l > 0)
    float *part;       // [S][K][W]
    float *gb;         // rows of layer l-1's bias in the workgroup sums (l > 0), n / 64 per row
    int *ctr;          // non-null once per step: advances the forward's step counter
    int n, K, S, nA;
};

// acc[i][j] += A_i(k) B_j(k) over the K steps k = k, k + ks, ... < kend, where A_i(k) = load8(Ar +
// i ast + k) and B_j(k) = load8(Br + j bst + k) are MFMA fragments, loaded DEPTH steps ahead into a
// ring of register sets.
template <int NI, int NJ, int DEPTH>
__device__ __forceinline__ void mma_steps(f32x4 (&acc)[NI][NJ], const u16 *Ar, size_t ast, const u16 *Br, size_t bst,
                                          int k, int kend, int ks) {
    if (k >= kend) return;
    const int nsteps = (kend - 1 - k) / ks + 1;
    bf16x8 x[DEPTH][NI], y[DEPTH][NJ];
    auto ld = [&](int d, int step) {
        const int kk = k + step * ks;
#pragma unroll
        for (int i = 0; i < NI; ++i) x[d][i] = load8(Ar + i * ast + kk);
#pragma unroll
        for (int j = 0; j < NJ; ++j) y[d][j] = load8(Br + j * bst + kk);
    };
#pragma unroll
    for (int d = 0; d < DEPTH; ++d)
        if (d < nsteps) ld(d, d);
    for (int s0 = 0; s0 < nsteps; s0 += DEPTH) {
#pragma unroll
        for (int d = 0; d < DEPTH; ++d) {
            if (s0 + d < nsteps) {
#pragma unroll
                for (int i = 0; i < NI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) acc[i][j] = mfma(x[d][i], y[d][j], acc[i][j]);
                if (s0 + d + DEPTH < nsteps) ld(d, s0 + d + DEPTH);
            }
        }
    }
}

__global__ __launch_bounds__(256, 2) void k_train_bwd(BwdLayer a) {
    // two partial weight-gradient tiles (32 KB), or the point-major dz tile ([TILE][LDW] BF16)
    __shared__ __attribute__((aligned(16))) float smem[std::max(2 * TILE * TILE, TILE * LDW / 2)];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane & 15;
    const int koff = 8 * (lane >> 4), g4 = 4 * (lane >> 4);
    if (a.ctr && blockIdx.x == 0 && t == 0) a.ctr[0] += 1;  // no forward reads it meanwhile
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    if ((int)blockIdx.x < a.nA) {
        constexpr int NT = WIDTH / TILE;
        const int ntile = (a.K / TILE) * NT;
        int tile, s;
        if (a.S % 8 == 0) {
            // workgroups go to the 8 XCDs round-robin: every tile of one chunk on one XCD, whose L2
            // then serves the chunk's activation and dz slices to all of its tiles
            const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
            tile = k % ntile;
            s = 8 * (k / ntile) + x;
        } else {
            tile = blockIdx.x % ntile;
            s = blockIdx.x / ntile;
        }
        const int i0 = (tile / NT) * TILE, j0 = (tile % NT) * TILE;
        const int chunk = a.n / a.S, kb = s * chunk;
        const size_t st = (size_t)16 * a.n;
        mma_steps<4, 4, 4>(acc, a.ainT + (size_t)(i0 + lr) * a.n + koff, st, a.dzT + (size_t)(j0 + lr) * a.n + koff,
                           st, kb + KS * wave, kb + chunk, 4 * KS);
        // the four partial tiles meet in LDS in two rounds, (w0 + w2) + (w1 + w3)
        auto put = [&](float *slot) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) slot[(16 * i + g4 + r) * TILE + 16 * j + lr] = acc[i][j][r];
        };
        auto add = [&](const float *slot) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[i][j][r] += slot[(16 * i + g4 + r) * TILE + 16 * j + lr];
        };
        if (wave >= 2) put(smem + (wave - 2) * TILE * TILE);
        __syncthreads();
        if (wave < 2) add(smem + wave * TILE * TILE);
        __syncthreads();
        if (wave == 1) put(smem);
        __syncthreads();
        if (wave == 0) {
            add(smem);
            float *out = a.part + (size_t)s * a.K * WIDTH + (size_t)i0 * WIDTH + j0;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) out[(size_t)(16 * i + g4 + r) * WIDTH + 16 * j + lr] = acc[i][j][r];
        }
        return;
    }
    const int p0 = (blockIdx.x - a.nA) * TILE;
    u16 *stage = reinterpret_cast<u16 *>(smem);  // [TILE][LDW] BF16: dz_l rows, then dz_{l-1} rows
    for (int c = t; c < TILE * WIDTH / 8; c += 256) {  // the block's dz_l rows in one burst
        const int row = c >> 5, c8 = (c & 31) * 8;
        *reinterpret_cast<bf16x8 *>(stage + row * LDW + c8) = load8(a.dz + (size_t)(p0 + row) * WIDTH + c8);
    }
    uint2 dsv[4][4];  // silu'(z_{l-1}) of this lane's outputs, loaded before the products
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            dsv[j][i] = *reinterpret_cast<const uint2 *>(a.ds + (size_t)(64 * wave + 16 * j + lr) * a.n + p0 + 16 * i + g4);
    __syncthreads();
    // rows: the block's points (from LDS); columns: this wave's 64 features of layer l-1
    mma_steps<4, 4, 2>(acc, stage + lr * LDW + koff, (size_t)16 * LDW, a.w16 + (size_t)(64 * wave + lr) * WIDTH + koff,
                       (size_t)16 * WIDTH, 0, WIDTH, KS);
    __syncthreads();  // every wave has read the dz_l rows
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = 64 * wave + 16 * j + lr;
        float sb = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const size_t o = (size_t)col * a.n + p0 + 16 * i + g4;
            const uint2 d = dsv[j][i];
            const uint2 v = {pk2(acc[i][j][0] * lo_bf(d.x), acc[i][j][1] * hi_bf(d.x)),
                             pk2(acc[i][j][2] * lo_bf(d.y), acc[i][j][3] * hi_bf(d.y))};
            sb += (lo_bf(v.x) + hi_bf(v.x)) + (lo_bf(v.y) + hi_bf(v.y));
            u16 *bp = stage + (16 * i + g4) * LDW + col;
            bp[0] = (u16)(v.x & 0xffffu);
            bp[LDW] = (u16)(v.x >> 16);
            bp[2 * LDW] = (u16)(v.y & 0xffffu);
            bp[3 * LDW] = (u16)(v.y >> 16);
            *reinterpret_cast<uint2 *>(a.dzpT + o) = v;
        }
        sb += __shfl_xor(sb, 16);
        sb += __shfl_xor(sb, 32);
        if (lane < 16) a.gb[(size_t)col * (a.n / TILE) + blockIdx.x - a.nA] = sb;
    }
    __syncthreads();
    for (int c = t; c < TILE * WIDTH / 8; c += 256) {
        const int row = c >> 5, c8 = (c & 31) * 8;
        *reinterpret_cast<bf16x8 *>(a.dzp + (size_t)(p0 + row) * WIDTH + c8) = load8(stage + row * LDW + c8);
    }
}

// Adam over the whole FP32 parameter vector.  Workgroups [0, nW / 256): one thread per weight, whose
// gradient is the sum of the S partial slices of its layer, in slice order.  The workgroups after
// them: one wave per other parameter (biases, w_out, b_out) and one for the loss, each summing its
// row of workgroup sums in a fixed order.  A weight's new value refreshes both BF16 copies: wt
// ([W][K], the forward's operand) and w16 ([K][W], the backward's).  update == 0 only writes the
// gradient into g (the gradient entry point).  The loss is added to *loss.
struct AdamArgs {
    float *p, *g, *m1, *m2;
    const float *part;   // layer l's slices start at part + 32 off_w[l]
    const float *gpart;  // [nq][nblk] workgroup sums of the other parameters, then the loss
    float *loss;
    const float *sched;  // [step][3]: learning rate, 1 - beta1^t, 1 - beta2^t
    const int *ctr;      // ctr[1]: this step's index into sched
    size_t nparam, off_w[9];
    int K[8], S[8];
    u16 *wt[8], *w16[8];
    int depth, update, nblk, nq;
};

__global__ __launch_bounds__(256) void k_adam_all(AdamArgs a) {
    const size_t nW = a.off_w[a.depth];
    float gi;
    int l = -1;
    size_t i, loc = 0;
    if (blockIdx.x < nW / 256) {
        i = blockIdx.x * (size_t)256 + threadIdx.x;
        l = 0;
        while (l + 1 < a.depth && i >= a.off_w[l + 1]) ++l;
        loc = i - a.off_w[l];
        const size_t sz = (size_t)a.K[l] * WIDTH;
        const float *q = a.part + 32 * a.off_w[l] + loc;
        gi = 0.0f;
        int s = 0;
        for (; s + 8 <= a.S[l]; s += 8) {  // eight loads in flight, summed in slice order
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = q[(s + u) * sz];
#pragma unroll
            for (int u = 0; u < 8; ++u) gi += v[u];
        }
        for (; s < a.S[l]; ++s) gi += q[s * sz];
    } else {
        const int q = (int)(blockIdx.x - nW / 256) * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
        if (q >= a.nq) return;
        const float *row = a.gpart + (size_t)q * a.nblk;
        float sum = 0.0f;
        for (int b = lane; b < a.nblk; b += 64) sum += row[b];
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (lane != 0) return;
        i = nW + q;
        if (i >= a.nparam) {  // the loss row
            *a.loss += sum;
            return;
        }
        gi = sum;
    }
    if (!a.update) {
        a.g[i] = gi;
        return;
    }
    const int st = a.ctr[1];
    const float lr = a.sched[3 * st], c1 = a.sched[3 * st + 1], c2 = a.sched[3 * st + 2];
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mo = b1 * a.m1[i] + (1.0f - b1) * gi;
    const float ve = b2 * a.m2[i] + (1.0f - b2) * gi * gi;
    a.m1[i] = mo;
    a.m2[i] = ve;
    const float v = a.p[i] - lr * (mo / c1) / (sqrtf(ve / c2) + eps);
    a.p[i] = v;
    if (l >= 0) {
        const int ii = (int)(loc / WIDTH), jj = (int)(loc - (size_t)ii * WIDTH);
        const u16 h = f2bf(v);
        a.w16[l][loc] = h;
        a.wt[l][(size_t)jj * a.K[l] + ii] = h;
    }
}

// Both BF16 copies of a weight matrix W[K][N] (row-major FP32): wt = W^T, w16 = W.
__global__ void k_to_bf16_t(const float *__restrict__ w, u16 *__restrict__ wt, u16 *__restrict__ w16, int K, int N) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= K * N) return;
    const int k = i / N, nn = i - k * N;
    const u16 h = f2bf(w[i]);
    wt[(size_t)nn * K + k] = h;
    w16[i] = h;
}

// Training samples, drawn on the device: candidate c has coordinates lo + (hi - lo) u_d, with u_d
// the top 24 bits of splitmix64(key + (3 c + d) golden) (the SplitMix64 generator of Steele, Lea and
// Flood, 2014), and the body's phi at it (FP64 evaluation, stored FP32).  ok (when given) flags the
// candidates within band of the surface.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_candidates(cc::Geo geo, uint64_t key, uint64_t c0, int n, float lo, float hi, float band,
                             float *__restrict__ x, float *__restrict__ y, uint8_t *__restrict__ ok) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t c = c0 + (uint64_t)i;
    float u[3];
    double p[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const uint64_t h = splitmix64(key + (3 * c + d) * 0x9e3779b97f4a7c15ull);
        u[d] = lo + (hi - lo) * ((float)(h >> 40) * 0x1p-24f);
        p[d] = u[d];
        x[3 * (size_t)i + d] = u[d];
    }
    const float v = (float)cc::sdf_value<double, 3>(geo, p);
    y[i] = v;
    if (ok) ok[i] = fabsf(v) < band;
}

// Appends the selected candidates sel[0..n) (in candidate order) to the samples.
__global__ void k_place(const float *__restrict__ cx, const float *__restrict__ cy, const int *__restrict__ sel, int n,
                        float *__restrict__ X, float *__restrict__ Y) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int c = sel[i];
    X[3 * (size_t)i] = cx[3 * (size_t)c];
    X[3 * (size_t)i + 1] = cx[3 * (size_t)c + 1];
    X[3 * (size_t)i + 2] = cx[3 * (size_t)c + 2];
    Y[i] = cy[c];
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
    u16 *wt[8] = {};          // device BF16 weights per layer, transposed ([N][K]: the forward's operand)
    u16 *w16[8] = {};         // device BF16 weights per layer, row-major ([K][N]: the backward's operand)
    int step = 0;
    std::vector<float> ffB_host;
    // training state, sized for the batch
    int cap = 0;
    float *x = nullptr, *y = nullptr;  // the gradient entry point's samples
    u16 *featT = nullptr, *aT = nullptr, *ds = nullptr, *dz[2] = {}, *dzT[2] = {};
    float *part = nullptr;   // weight-gradient partials: up to 32 slices of every weight matrix
    float *gpart = nullptr;  // workgroup sums of the other gradients and the loss: [rows][batch / 64]
    float *small = nullptr;  // [0]: loss accumulator; ints at [16], [17]: step counters
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

int *counters(cc_mlp *m) { return reinterpret_cast<int *>(m->small + 16); }

// Rows of the workgroup sums: one per parameter after the weights, then the loss.
int sum_rows(const cc_mlp *m) { return (int)(m->nparam - m->off_b[0]) + 1; }

int ensure_cap(cc_mlp *m, int n) {
    if (n <= m->cap) return CC_OK;
    release(m->x); release(m->y); release(m->featT); release(m->aT); release(m->ds);
    release(m->dz[0]); release(m->dz[1]); release(m->dzT[0]); release(m->dzT[1]);
    release(m->part); release(m->gpart); release(m->small);
    m->cap = 0;
    const size_t W = m->width, H = std::max(m->depth - 1, 1), F2 = 2 * m->F, N = n;
    bool ok = hipMalloc(&m->x, 12 * N) == hipSuccess && hipMalloc(&m->y, 4 * N) == hipSuccess &&
              hipMalloc(&m->featT, 2 * F2 * N) == hipSuccess && hipMalloc(&m->aT, 2 * H * W * N) == hipSuccess &&
              hipMalloc(&m->ds, 2 * H * W * N) == hipSuccess &&
              hipMalloc(&m->part, 4 * 32 * m->off_b[0]) == hipSuccess &&
              hipMalloc(&m->gpart, 4 * sum_rows(m) * (N / TILE)) == hipSuccess && hipMalloc(&m->small, 256) == hipSuccess;
    for (int c = 0; ok && c < 2; ++c)
        ok = hipMalloc(&m->dz[c], 2 * N * W) == hipSuccess && hipMalloc(&m->dzT[c], 2 * N * W) == hipSuccess;
    if (!ok) return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate training buffers");
    m->cap = n;
    return CC_OK;
}

// Point chunks of a layer's weight gradient (K input features): about 256 workgroups, each chunk
// a whole number of K steps.
int chunks(int n, int K) {
    int S = std::max(1, std::min(32, 256 / ((K / TILE) * (WIDTH / TILE))));
    while ((n / KS) % S) --S;
    return S;
}

// Forward and backward over one minibatch of n points (a multiple of 64), enqueued on the model's
// stream: the rows of X / Y are taken in the epoch's shuffled order when keys are given (N samples,
// nb minibatches an epoch, the step read from the device counters), else rows 0..n-1.  Leaves the
// weight-gradient partials in m->part and the workgroup sums of the other gradients and of the loss
// in m->gpart.
void enqueue_forward_backward(cc_mlp *m, int n, const float *X, const float *Y, const uint32_t *keys, int hb,
                              uint32_t N, int nb) {
    hipStream_t s = m->stream;
    const int D = m->depth, W = m->width;
    TrainBufs tb;
    tb.featT = m->featT;
    tb.aT = m->aT;
    tb.ds = m->ds;
    tb.dz = m->dz[0];
    tb.dzT = m->dzT[0];
    tb.gpart = m->gpart;
    tb.ctr = counters(m);
    const size_t nW = m->off_b[0];
    tb.rblast = (int)(m->off_b[D - 1] - nW);
    tb.rwout = (int)(m->off_wout - nW);
    tb.rbout = (int)(m->off_bout - nW);
    tb.rloss = sum_rows(m) - 1;
    hipLaunchKernelGGL(k_train_fwd, dim3(n / TILE), dim3(FWD_T), FWD_LDS, s, dev_view(m), tb, X, Y, keys, hb, N, nb, n);
    for (int l = D - 1, c = 0; l >= 0; --l, c ^= 1) {
        BwdLayer a;
        a.ainT = l == 0 ? m->featT : m->aT + (size_t)(l - 1) * W * n;
        a.dz = m->dz[c];
        a.dzT = m->dzT[c];
        a.w16 = m->w16[l];
        a.ds = l > 0 ? m->ds + (size_t)(l - 1) * W * n : nullptr;
        a.dzp = m->dz[c ^ 1];
        a.dzpT = m->dzT[c ^ 1];
        a.part = m->part + 32 * m->off_w[l];
        a.gb = l > 0 ? m->gpart + (m->off_b[l - 1] - nW) * (size_t)(n / TILE) : nullptr;
        a.ctr = keys && l == D - 1 ? counters(m) : nullptr;
        a.n = n;
        a.K = m->K[l];
        a.S = chunks(n, a.K);
        a.nA = (a.K / TILE) * (W / TILE) * a.S;
        const int nB = l > 0 ? n / TILE : 0;
        hipLaunchKernelGGL(k_train_bwd, dim3(a.nA + nB), dim3(256), 0, s, a);
    }
}

void enqueue_adam(cc_mlp *m, int n, int update, const float *sched) {
    AdamArgs a;
    a.p = m->params;
    a.g = m->grads;
    a.m1 = m->m1;
    a.m2 = m->m2;
    a.part = m->part;
    a.gpart = m->gpart;
    a.loss = m->small;
    a.nblk = n / TILE;
    a.nq = sum_rows(m);
    a.sched = sched;
    a.ctr = counters(m);
    a.nparam = m->nparam;
    for (int l = 0; l < 8; ++l) {
        const bool in = l < m->depth;
        a.off_w[l] = in ? m->off_w[l] : 0;
        a.K[l] = in ? m->K[l] : 0;
        a.S[l] = in ? chunks(n, m->K[l]) : 0;
        a.wt[l] = in ? m->wt[l] : nullptr;
        a.w16[l] = in ? m->w16[l] : nullptr;
    }
    a.off_w[m->depth] = m->off_b[0];
    a.depth = m->depth;
    a.update = update;
    hipLaunchKernelGGL(k_adam_all, dim3((unsigned)(m->off_b[0] / 256 + (a.nq + 3) / 4)), dim3(256), 0, m->stream, a);
}

// The loss and the full gradient (into m->grads) at the n samples in m->x / m->y, synchronised.
int loss_grad(cc_mlp *m, int n, float *loss_out) {
    (void)hipMemsetAsync(m->small, 0, 256, m->stream);
    enqueue_forward_backward(m, n, m->x, m->y, nullptr, 0, 0, 1);
    enqueue_adam(m, n, 0, nullptr);
    float lh = 0.0f;
    (void)hipMemcpyAsync(&lh, m->small, 4, hipMemcpyDeviceToHost, m->stream);
    if (hipStreamSynchronize(m->stream) != hipSuccess)
        return nfail(CC_ERR_HIP, std::string("surrogate step: ") + hipGetErrorString(hipGetLastError()));
    if (loss_out) *loss_out = lh;
    return CC_OK;
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
    // the inference and training forward kernels' activation tiles exceed the default 64 KB of LDS
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp_infer<true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)INFER_LDS) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp_infer<false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)INFER_LDS) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_train_fwd),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)FWD_LDS) != hipSuccess ||
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
    for (int l = 0; ok && l < depth; ++l)
        ok = hipMalloc(&m->wt[l], 2 * (size_t)m->K[l] * width) == hipSuccess &&
             hipMalloc(&m->w16[l], 2 * (size_t)m->K[l] * width) == hipSuccess;
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
                           m->params + m->off_w[l], m->wt[l], m->w16[l], m->K[l], width);
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
                   (void *)m->x, (void *)m->y, (void *)m->featT, (void *)m->aT, (void *)m->ds,
                   (void *)m->dz[0], (void *)m->dz[1], (void *)m->dzT[0], (void *)m->dzT[1],
                   (void *)m->part, (void *)m->small, (void *)m->ex, (void *)m->ephi})
        if (p) (void)hipFree(p);
    for (int l = 0; l < 8; ++l) {
        if (m->wt[l]) (void)hipFree(m->wt[l]);
        if (m->w16[l]) (void)hipFree(m->w16[l]);
    }
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
                           m->params + m->off_w[l], m->wt[l], m->w16[l], m->K[l], m->width);
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
    // samples, drawn on the device: the uniform half directly, the near-surface half by rejection
    // from further uniform candidates, kept in candidate order
    std::mt19937_64 rng(seed);
    const uint64_t skey = rng();
    const size_t N = (size_t)nsamples, half = N / 2, chunk = 1 << 20;
    float *dX = nullptr, *dY = nullptr, *cx = nullptr, *cy = nullptr, *dsched = nullptr;
    uint8_t *cok = nullptr;
    int *sel = nullptr, *nsel = nullptr;
    uint32_t *dkeys = nullptr;
    void *tmp = nullptr;
    size_t tmpb = 0;
    (void)hipcub::DeviceSelect::Flagged(nullptr, tmpb, hipcub::CountingInputIterator<int>(0), (const uint8_t *)nullptr,
                                        (int *)nullptr, (int *)nullptr, (int)chunk);
    auto free_all = [&]() {
        for (void *q : {(void *)dX, (void *)dY, (void *)cx, (void *)cy, (void *)dsched, (void *)cok, (void *)sel,
                        (void *)nsel, (void *)dkeys, tmp, (void *)prims})
            if (q) (void)hipFree(q);
    };
    if (hipMalloc(&dX, 12 * N) || hipMalloc(&dY, 4 * N) || hipMalloc(&cx, 12 * chunk) || hipMalloc(&cy, 4 * chunk) ||
        hipMalloc(&cok, chunk) || hipMalloc(&sel, 4 * chunk) || hipMalloc(&nsel, 64) || hipMalloc(&tmp, tmpb)) {
        free_all();
        return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate training samples");
    }
    hipLaunchKernelGGL(k_candidates, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, m->stream, g, skey,
                       (uint64_t)0, (int)half, lo, hi, band, dX, dY, (uint8_t *)nullptr);
    size_t nn = 0;
    uint64_t c0 = half;
    for (int guard = 0; nn < N - half && guard < 4096; ++guard, c0 += chunk) {
        int got = 0;
        hipLaunchKernelGGL(k_candidates, dim3(chunk / 256), dim3(256), 0, m->stream, g, skey, c0, (int)chunk, lo, hi,
                           band, cx, cy, cok);
        if (hipcub::DeviceSelect::Flagged(tmp, tmpb, hipcub::CountingInputIterator<int>(0), cok, sel, nsel,
                                          (int)chunk, m->stream) != hipSuccess ||
            hipMemcpyAsync(&got, nsel, 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess) {
            free_all();
            return nfail(CC_ERR_HIP, std::string("surrogate sampling: ") + hipGetErrorString(hipGetLastError()));
        }
        const int take = (int)std::min<size_t>((size_t)got, N - half - nn);
        if (take > 0)
            hipLaunchKernelGGL(k_place, dim3((take + 255) / 256), dim3(256), 0, m->stream, cx, cy, sel, take,
                               dX + 3 * (half + nn), dY + half + nn);
        nn += (size_t)take;
    }
    if (nn < N - half) {
        free_all();
        return nfail(CC_ERR_ARGUMENT, "could not find enough near-surface samples");
    }
    int rc = ensure_cap(m, batch);
    if (rc) {
        free_all();
        return rc;
    }
    const size_t nbatch = N / batch;
    const size_t total_steps = nbatch * epochs;
    if (total_steps > (size_t)INT32_MAX) {
        free_all();
        return nfail(CC_ERR_ARGUMENT, "too many training steps");
    }
    // epoch e visits the samples in the order of a keyed permutation (six round keys an epoch)
    int hb = 0;
    while ((uint64_t(1) << (2 * hb)) < N) ++hb;
    std::vector<uint32_t> keys(6 * (size_t)epochs);
    for (auto &k : keys) k = (uint32_t)rng();
    // the learning rate (cosine decay) and Adam's bias corrections of every step, read on the device
    std::vector<float> sched(3 * total_steps);
    for (size_t gs = 0; gs < total_steps; ++gs) {
        const float t = (float)(m->step + 1 + gs);
        sched[3 * gs] = 0.5f * lr * (1.0f + std::cos(3.14159265f * (float)gs / (float)total_steps));
        sched[3 * gs + 1] = 1.0f - std::pow(0.9f, t);
        sched[3 * gs + 2] = 1.0f - std::pow(0.999f, t);
    }
    if (hipMalloc(&dsched, 12 * total_steps) || hipMalloc(&dkeys, 4 * keys.size())) {
        free_all();
        return nfail(CC_ERR_OUT_OF_MEMORY, "surrogate training schedule");
    }
    (void)hipMemcpyAsync(dsched, sched.data(), 12 * total_steps, hipMemcpyHostToDevice, m->stream);
    (void)hipMemcpyAsync(dkeys, keys.data(), 4 * keys.size(), hipMemcpyHostToDevice, m->stream);
    (void)hipMemsetAsync(m->small, 0, 256, m->stream);  // loss and step counters
    // one step: forward, backward, Adam.  Runs of GSTEPS steps and single steps are captured once as
    // graphs and replayed (the step index lives on the device)
    auto enqueue_step = [&]() {
        enqueue_forward_backward(m, batch, dX, dY, dkeys, hb, (uint32_t)N, (int)nbatch);
        enqueue_adam(m, batch, 1, dsched);
    };
    constexpr int GSTEPS = 8;
    hipGraph_t graph[2] = {};
    hipGraphExec_t exec[2] = {};
    for (int k = 0; k < 2; ++k) {
        if (hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) break;
        for (int r = 0; r < (k == 0 ? GSTEPS : 1); ++r) enqueue_step();
        if (hipStreamEndCapture(m->stream, &graph[k]) != hipSuccess ||
            hipGraphInstantiate(&exec[k], graph[k], nullptr, nullptr, 0) != hipSuccess)
            exec[k] = nullptr;  // the same kernels are then launched one by one
    }
    (void)hipGetLastError();
    float *hloss = nullptr;
    if (hipHostMalloc(&hloss, 64) != hipSuccess) hloss = nullptr;
    float lbuf = 0.0f;
    float *lh = hloss ? hloss : &lbuf;
    for (int e = 0; e < epochs && rc == CC_OK; ++e) {
        (void)hipMemsetAsync(m->small, 0, 4, m->stream);
        for (size_t b = 0; b < nbatch;) {
            if (exec[0] && b + GSTEPS <= nbatch) {
                (void)hipGraphLaunch(exec[0], m->stream);
                b += GSTEPS;
            } else {
                if (exec[1]) (void)hipGraphLaunch(exec[1], m->stream);
                else enqueue_step();
                ++b;
            }
        }
        (void)hipMemcpyAsync(lh, m->small, 4, hipMemcpyDeviceToHost, m->stream);
        if (hipStreamSynchronize(m->stream) != hipSuccess)
            rc = nfail(CC_ERR_HIP, std::string("surrogate training: ") + hipGetErrorString(hipGetLastError()));
        else if (history)
            history[e] = lh[0] / (float)nbatch;
    }
    for (int k = 0; k < 2; ++k) {
        if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
        if (graph[k]) (void)hipGraphDestroy(graph[k]);
    }
    if (hloss) (void)hipHostFree(hloss);
    free_all();
    if (rc) return rc;
    m->step += (int)total_steps;
    return CC_OK;
}

}  // extern "C"
