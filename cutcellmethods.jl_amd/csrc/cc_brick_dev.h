// Device pieces of a sphere union's brick-wise nodal phi, shared by the cull and phi passes
// (cc_phi.hip) and the fused stream pass (cc_mc_dev.h): the culled fold, the surrogate network's
// trilinear value, a sub-brick's sign certificate and the per-node decision between them.  No
// counterpart in the reference, whose sample_nodes (src/marching_cubes/cache.jl:33-40) folds the
// whole body at every node; the folds here are bitwise that fold (cc_device.h union_dist)
// restricted to primitives proven inert.  Every symbol is internal to the including unit.
#pragma once
#include "cc_internal.h"

namespace cc {
namespace {

// Primitives the phi pass stages in LDS per brick; a brick keeping more folds from memory.  A brick
// of the 512-sphere composite body keeps 15 on average, and a smaller stage leaves room for more
// blocks per CU.
constexpr int CULL_LDS = 128;

// The brick pass screens its fold with screen_record / screen_skips (cc_device.h), the records
// about the brick's (or the work item's) first node.

// The sphere-union fold over a brick's kept primitives, in the order and with the arithmetic of
// union_dist (cc_device.h, fold_sphere).  Primitive 0 is always kept.  sph: the first
// nk kept primitives staged as T (null: read every one through L).  scr (double only, with sph):
// their screen records about the origin o; a primitive the screen does not skip is folded as the
// oracle folds every one, smin_poly(acc, |x - c| - r, k).
template <typename T>
__device__ __forceinline__ T culled_fold(const T *sph, uint32_t nk, const uint16_t *L, const Geo &geo,
                                         T x, T y, T z, const float4 *scr = nullptr, const T *o = nullptr) {
    const T k = T(geo.p[0]);
    T acc;
    if (sizeof(T) == 8 && sph && scr) {
        acc = sphere_dist<T, 3>(sph[0], sph[1], sph[2], sph[3], x, y, z);
        const float xf = (float)(x - o[0]), yf = (float)(y - o[1]), zf = (float)(z - o[2]);
        float accf = (float)acc;
        for (uint32_t e = 1; e < nk; ++e) {
            if (screen_skips(scr[e], xf, yf, zf, accf)) continue;
            acc = smin_poly(acc, sphere_dist<T, 3>(sph[4 * e], sph[4 * e + 1], sph[4 * e + 2], sph[4 * e + 3], x, y, z),
                            k);
            accf = (float)acc;
        }
    } else if (sph) {
        acc = sphere_dist<T, 3>(sph[0], sph[1], sph[2], sph[3], x, y, z);
        for (uint32_t e = 1; e < nk; ++e)
            acc = fold_sphere<T, 3>(acc, sph[4 * e], sph[4 * e + 1], sph[4 * e + 2], sph[4 * e + 3], x, y, z, k);
    } else {
        const double *s = geo.prims + 4 * (size_t)L[0];
        acc = sphere_dist<T, 3>(T(s[0]), T(s[1]), T(s[2]), T(s[3]), x, y, z);
        for (uint32_t e = 1; e < nk; ++e) {
            s = geo.prims + 4 * (size_t)L[e];
            acc = fold_sphere<T, 3>(acc, T(s[0]), T(s[1]), T(s[2]), T(s[3]), x, y, z, k);
        }
    }
    return acc;
}

// The brick-major sign bitmap the phi pass leaves beside phi for the stream pass, which needs only
// each node's side (a cut cell's exact corners are read from phi by the cut pass): node (gi, gj, gk)
// is bit li + 16 (lj % 4) of word 64 b + lj / 4 + 4 lk, with b its brick and (li, lj, lk) its place
// in it, so the 64 nodes one wave of the phi pass holds (consecutive local indices) fill one word.
// The bit is set when the node's value, zeros nudged to +eps as marching_cubes.jl:240-243 does,
// is >= 0 (NaN: clear).
template <typename T> __device__ __forceinline__ bool node_outside(T v) {
    return fabs(v) < Eps<T>::v || v >= T(0);
}

__device__ __forceinline__ bool sign_bit(const uint64_t *__restrict__ signs, int nbx, int nby, uint32_t gi,
                                         uint32_t gj, uint32_t gk) {
    const uint32_t b = (gi / PHI_BRICK) + (uint32_t)nbx * ((gj / PHI_BRICK) + (uint32_t)nby * (gk / PHI_BRICK));
    const uint32_t li = gi % PHI_BRICK, lj = gj % PHI_BRICK, lk = gk % PHI_BRICK;
    const uint64_t w = signs[64 * (size_t)b + (lj >> 2) + 4 * lk];
    return (w >> (li + 16 * (lj & 3))) & 1ull;
}

// The network lattice's cell and fraction of node index g along one axis: g / f and
// (g - f (g / f)) / f -- with a shift, a mask and the exact reciprocal when f is a power of two, the
// same values.
__device__ __forceinline__ int nn_axis(const SurrogateArgs &sa, int g, float *fr) {
    if (sa.fshift >= 0) {
        *fr = (float)(g & (sa.f - 1)) * sa.finv;
        return g >> sa.fshift;
    }
    const int c = g / sa.f;
    *fr = (float)(g - sa.f * c) / (float)sa.f;
    return c;
}

// The bilinear value in lattice plane K at cell (I, J) and fractions (fx, fy).
__device__ __forceinline__ float nn_plane(const SurrogateArgs &sa, int I, int J, int K, float fx, float fy) {
    const size_t sy = (size_t)sa.nc[0];
    const float *p = sa.nn + I + J * sy + (size_t)K * sy * (size_t)sa.nc[1];
    const float c00 = p[0] + fx * (p[1] - p[0]);
    const float c10 = p[sy] + fx * (p[sy + 1] - p[sy]);
    return c00 + fy * (c10 - c00);
}

// The network's trilinear value at node (gi, gj, gk): the bilinear values of planes K and K + 1,
// then along z.
__device__ __forceinline__ float nn_trilinear(const SurrogateArgs &sa, int gi, int gj, int gk) {
    float fx, fy, fz;
    const int I = nn_axis(sa, gi, &fx), J = nn_axis(sa, gj, &fy), K = nn_axis(sa, gk, &fz);
    const float c0 = nn_plane(sa, I, J, K, fx, fy), c1 = nn_plane(sa, I, J, K + 1, fx, fy);
    return c0 + fz * (c1 - c0);
}

// The sign certificate of sub-brick s (0..63, PHI_SUB^3 nodes, x fastest) of the brick whose first
// node is (i0, j0, k0): the exact phi at the sub-brick's centre c, with every node of it within rho
// of c and the cell diagonal diag, gives every node x of it |phi(x)| > diag and the sign of phi(c)
// when |phi(c)| - rho - diag exceeds the rounding margin delta_c (the body is 1-Lipschitz).  Returns
// +1, -1 or 0 (uncertified); *valid is false for a sub-brick wholly outside the node lattice.
template <typename T>
__device__ __forceinline__ int8_t sub_brick_sign(const GridT<T> &g, const Geo &geo, const T *sph, uint32_t nk,
                                                 const uint16_t *L, int i0, int j0, int k0, int s, T delta_c,
                                                 bool *valid, const float4 *scr = nullptr,
                                                 const T *o = nullptr) {
    const int s3[3] = {s & 3, (s >> 2) & 3, s >> 4};
    const int base[3] = {i0, j0, k0};
    bool ok = true;
    T c[3], e2 = T(0), dg = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int slo = base[a] + PHI_SUB * s3[a];
        const int shi = min(slo + PHI_SUB - 1, g.n[a]);
        ok = ok && slo <= g.n[a];
        const T xl = g.x0[a] + T(slo) * g.d[a], xh = g.x0[a] + T(shi) * g.d[a];
        c[a] = (xl + xh) / T(2);
        const T ea = fabs(c[a] - xl) > fabs(xh - c[a]) ? fabs(c[a] - xl) : fabs(xh - c[a]);
        e2 = e2 + ea * ea;
        dg = dg + g.d[a] * g.d[a];
    }
    *valid = ok;
    if (!ok) return 0;
    const T pc = culled_fold<T>(sph, nk, L, geo, c[0], c[1], c[2], scr, o);
    if (fabs(pc) - sqrt(e2) - sqrt(dg) > delta_c) return pc > T(0) ? 1 : -1;
    return 0;
}

// How a node's phi was obtained (the phi pass's counters).
enum NodeHow { NODE_NET = 0, NODE_EXACT = 1, NODE_BAND = 2, NODE_CONTRADICTED = 3 };

// The surrogate's decision at a node of a certified sub-brick of sign sg (!= 0): keep the network's
// value v unless |v| is under the band or its sign contradicts the certificate.
__device__ __forceinline__ int keep_network(const SurrogateArgs &sa, int8_t sg, float v) {
    if (fabsf(v) < sa.band) return NODE_BAND;
    if ((v > 0.0f) != (sg > 0)) return NODE_CONTRADICTED;
    return NODE_NET;
}

// phi at node (gi, gj, gk) (coordinates x) for the fused stream pass: the phi pass's value there,
// bit for bit.  The certificate is the node's own brick's; the exact fold runs over the kept list
// (nk entries L, all staged in sph unless it is null) of a brick whose extended box holds the node -- the
// calling tile's brick -- which gives the same value as the node's own brick's list.  *how receives
// the NodeHow.
template <typename T>
__device__ __forceinline__ T brick_node_phi(const BrickPhi &bp, const Geo &geo, uint32_t gi, uint32_t gj,
                                            uint32_t gk, const T *x, int *how, const T *sph, uint32_t nk,
                                            const uint16_t *L) {
    *how = NODE_EXACT;
    if (bp.cert) {
        const uint32_t b = (gi / PHI_BRICK) + (uint32_t)bp.nbx * ((gj / PHI_BRICK) + (uint32_t)bp.nby * (gk / PHI_BRICK));
        const int sub = (int)((gi / PHI_SUB) % 4 + 4 * ((gj / PHI_SUB) % 4) + 16 * ((gk / PHI_SUB) % 4));
        const int8_t sg = bp.cert[64 * (size_t)b + sub];
        if (sg != 0) {
            const float v = nn_trilinear(bp.sa, (int)gi, (int)gj, (int)gk);
            *how = keep_network(bp.sa, sg, v);
            if (*how == NODE_NET) return T(v);
        }
    }
    return culled_fold<T>(sph, nk, L, geo, x[0], x[1], x[2]);
}

}  // namespace
}  // namespace cc
