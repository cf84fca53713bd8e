// Nodal phi of a sphere union, brick by brick: the cull pass and the phi pass (culled exact, or
// surrogate with certified signs and exact fallback).  No counterpart in the reference, whose
// sample_nodes (src/marching_cubes/cache.jl:33-40) folds the whole body at every node; the folds
// here are bitwise that fold (cc_device.h union_dist) restricted to primitives proven inert.
//
// Certification (surrogate mode).  Each brick is split into 4^3-node sub-bricks.  The exact phi at
// a sub-brick's centre c, with every node of the sub-brick within rho of c, the cell diagonal diag
// and the body 1-Lipschitz (spheres are, and the polynomial smooth minimum's gradient is a convex
// combination of its arguments'), gives |phi(x)| > diag with the sign of phi(c) at every node x of
// the sub-brick whenever |phi(c)| - rho - diag exceeds a rounding margin.  No cell with a corner in
// such a sub-brick can be cut (its other corners lie within diag of x), so every corner of every
// cut cell sits in an uncertified sub-brick, and all of those nodes are evaluated exactly.  Inside a
// certified sub-brick a node keeps the network's value unless |phi_nn| is under the band or its
// sign contradicts the certificate; then it is evaluated exactly too.  Every node's sign is thus
// the exact one, `kind` is identical everywhere and every cut cell sees its exact corners.
#include "cc_brick_dev.h"

namespace cc {
namespace {

constexpr int BRICK_NODES = PHI_BRICK * PHI_BRICK * PHI_BRICK;

__device__ __forceinline__ void brick_of(int b, int nbx, int nby, int *bi) {
    bi[0] = b % nbx;
    bi[1] = (b / nbx) % nby;
    bi[2] = b / (nbx * nby);
}

template <typename T>
__global__ __launch_bounds__(256) void k_cull(GridT<T> g, Geo geo, T delta, int nbx, int nby,
                                              uint16_t *__restrict__ list, uint32_t *__restrict__ cnt,
                                              int8_t *__restrict__ cert, T delta_c, uint32_t *__restrict__ bstat) {
    __shared__ T wmin[4];
    __shared__ uint32_t wcnt[4];
    __shared__ uint32_t ncert;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int bi[3];
    brick_of(b, nbx, nby, bi);
    // the brick's box extended by one node layer on its upper sides: the kept list is then exact for
    // every node a brick-shaped stream tile touches, its halo included
    T lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int i0 = bi[a] * PHI_BRICK, i1 = min(i0 + PHI_BRICK, g.n[a]);
        lo[a] = g.x0[a] + T(i0) * g.d[a];
        hi[a] = g.x0[a] + T(i1) * g.d[a];
    }
    const T k = T(geo.p[0]);
    const T inf = T(INFINITY);
    T run = inf;
    uint32_t total = 0;
    uint16_t *out = list + (size_t)b * geo.nprims;
    for (int base = 0; base < geo.nprims; base += 256) {
        const int i = base + t;
        T lb = inf, ub = inf;
        if (i < geo.nprims) {
            const double *s = geo.prims + 4 * (size_t)i;
            T sl = T(0), sh = T(0);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const T c = T(s[a]);
                T dl = lo[a] - c, dr = c - hi[a];
                dl = dl > dr ? dl : dr;
                dl = dl > T(0) ? dl : T(0);
                const T el = fabs(c - lo[a]), er = fabs(c - hi[a]);
                const T e = el > er ? el : er;
                sl = sl + dl * dl;
                sh = sh + e * e;
            }
            lb = sqrt(sl) - T(s[3]);
            ub = sqrt(sh) - T(s[3]);
        }
        // inclusive prefix minimum of the upper bounds within the wave
        T v = ub;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const T o = __shfl_up(v, off);
            if (lane >= off) v = o < v ? o : v;
        }
        if (lane == 63) wmin[wave] = v;
        T excl = __shfl_up(v, 1);
        if (lane == 0) excl = inf;
        __syncthreads();
        T before = run;
        for (int w = 0; w < wave; ++w) before = wmin[w] < before ? wmin[w] : before;
        const T U = excl < before ? excl : before;  // min of the upper bounds of primitives < i
        const bool keep = i < geo.nprims && (i == 0 || !(lb - U >= k + delta));
        const unsigned long long m = __ballot(keep);
        if (lane == 0) wcnt[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = total;
        for (int w = 0; w < wave; ++w) off += wcnt[w];
        if (keep) out[off + __popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)i;
        total += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        for (int w = 0; w < 4; ++w) run = wmin[w] < run ? wmin[w] : run;
        __syncthreads();
    }
    if (t == 0) cnt[b] = total;
    if (!bstat) return;
    // the sign certificates of the brick's 64 sub-bricks (the kept list is complete: every thread
    // of the workgroup wrote its entries before the last barrier)
    if (t == 0) ncert = 0;
    __syncthreads();
    if (cert && t < 64) {
        bool valid;
        cert[64 * (size_t)b + t] =
            sub_brick_sign<T>(g, geo, (const T *)nullptr, total, out, bi[0] * PHI_BRICK, bi[1] * PHI_BRICK,
                              bi[2] * PHI_BRICK, t, delta_c, &valid);
        const unsigned long long m = __ballot(valid);
        if (t == 0) ncert = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (t < PHI_BSTAT) bstat[(size_t)b * PHI_BSTAT + t] = t == 2 ? ncert : 0u;
}

template <typename T, bool SURR>
__global__ __launch_bounds__(256) void k_phi_bricks(T *__restrict__ phi, GridT<T> g, Geo geo,
                                                    const uint16_t *__restrict__ list,
                                                    const uint32_t *__restrict__ cnt, SurrogateArgs sa,
                                                    T delta_c, int nbx, int nby,
                                                    uint32_t *__restrict__ bstat, uint64_t *__restrict__ signs) {
    __shared__ T sph[4 * CULL_LDS];
    __shared__ float4 scr[sizeof(T) == 8 ? CULL_LDS : 1];  // screen records (double fold only)
    __shared__ uint64_t sbits[BRICK_NODES / 64];  // the brick's sign bitmap words (node_outside)
    __shared__ uint16_t queue[BRICK_NODES];
    __shared__ int8_t ssign[64];
    __shared__ uint32_t wsum[4], st_uncert, st_band, st_sign, st_nodes, st_cert;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63;
    int bi[3];
    brick_of(b, nbx, nby, bi);
    const int i0 = bi[0] * PHI_BRICK, j0 = bi[1] * PHI_BRICK, k0 = bi[2] * PHI_BRICK;
    const uint32_t nk = cnt[b];
    const uint16_t *L = list + (size_t)b * geo.nprims;
    const uint32_t nstage = nk < (uint32_t)CULL_LDS ? nk : (uint32_t)CULL_LDS;
    // the screen's origin is the brick's first node; every point folded here (its nodes, its
    // sub-bricks' centres) is within R of it
    const T o[3] = {g.x0[0] + T(i0) * g.d[0], g.x0[1] + T(j0) * g.d[1], g.x0[2] + T(k0) * g.d[2]};
    const double od[3] = {double(o[0]), double(o[1]), double(o[2])};
    const double R = 16.0 * sqrt((double(g.d[0]) * double(g.d[0]) + double(g.d[1]) * double(g.d[1])) +
                                 double(g.d[2]) * double(g.d[2]));
    for (uint32_t e = t; e < nstage; e += 256) {
        const double *s = geo.prims + 4 * (size_t)L[e];
        sph[4 * e] = T(s[0]);
        sph[4 * e + 1] = T(s[1]);
        sph[4 * e + 2] = T(s[2]);
        sph[4 * e + 3] = T(s[3]);
        if (sizeof(T) == 8) scr[e] = screen_record(s, od, geo.p[0], R);
    }
    const bool staged = nk <= (uint32_t)CULL_LDS;
    const float4 *scrp = sizeof(T) == 8 && staged ? scr : nullptr;
    if (t == 0) st_uncert = st_band = st_sign = st_nodes = st_cert = 0;
    __syncthreads();
    const size_t m0 = (size_t)g.n[0] + 1, m1 = (size_t)g.n[1] + 1;
    if (SURR && t < 64) {
        bool valid;
        ssign[t] = sub_brick_sign<T>(g, geo, staged ? sph : nullptr, nk, L, i0, j0, k0, t, delta_c, &valid, scrp, o);
        const unsigned long long vm = __ballot(valid);
        if (t == 0) st_cert = (uint32_t)__popcll(vm);
    }
    if (SURR) __syncthreads();
    // classify: nodes the surrogate may keep are written now; the rest are marked in a per-thread
    // mask and queued for the fold below (no per-node synchronisation, so the network's loads of
    // consecutive nodes overlap)
    uint32_t n_unc = 0, n_band = 0, n_sign = 0, n_nodes = 0, pushm = 0;
    // a thread's nodes are one column of the brick, (li, lj) = (t % 16, t / 16) and lk = q, so the
    // layer is the same for the whole block: the network's bilinear values in lattice planes K and
    // K + 1 (those of nn_trilinear) serve every node of the column with gk / f == K, K never
    // decreases along it, and K, Kc and the window's updates are uniform
    const int li = t & 15, lj = t >> 4, gi = i0 + li, gj = j0 + lj;
    const bool col_in = gi <= g.n[0] && gj <= g.n[1];
    const size_t colix = (size_t)gi + m0 * (size_t)gj, sz = m0 * m1;
    float nfx = 0.0f, nfy = 0.0f, b0 = 0.0f, b1 = 0.0f;
    int nI = 0, nJ = 0, Kc = -2;
    if (SURR) {
        nI = nn_axis(sa, gi, &nfx);
        nJ = nn_axis(sa, gj, &nfy);
    }
#pragma unroll 4
    for (int q = 0; q < BRICK_NODES / 256; ++q) {
        const int idx = t + 256 * q, lk = q, gk = k0 + lk;
        const bool in = col_in && gk <= g.n[2];
        bool push = in, kept_out = false;
        float fz = 0.0f;
        if (SURR && gk <= g.n[2]) {
            const int K = nn_axis(sa, gk, &fz);
            if (K != Kc) {
                if (col_in) {
                    b0 = K == Kc + 1 ? b1 : nn_plane(sa, nI, nJ, K, nfx, nfy);
                    b1 = nn_plane(sa, nI, nJ, K + 1, nfx, nfy);
                }
                Kc = K;
            }
        }
        if (SURR) {  // keep_network's decision as selects: one masked store, no nested branches
            const int8_t sg = ssign[(li >> 2) + 4 * (lj >> 2) + 16 * (lk >> 2)];
            const float v = b0 + fz * (b1 - b0);
            const bool cert = in && sg != 0, band = fabsf(v) < sa.band, contra = (v > 0.0f) != (sg > 0);
            const bool keep = cert && !band && !contra;
            n_unc += in && sg == 0 ? 1u : 0u;
            n_band += cert && band ? 1u : 0u;
            n_sign += cert && !band && contra ? 1u : 0u;
            if (keep) phi[colix + sz * (size_t)gk] = T(v);
            push = in && !keep;
            kept_out = keep && node_outside(T(v));
        }
        n_nodes += in ? 1u : 0u;
        pushm |= (push ? 1u : 0u) << q;
        // a wave's 64 nodes (local indices 64 w .. 64 w + 63) are one word of the sign bitmap: the
        // kept network values' bits now, the folded nodes' below
        if (signs) {
            const unsigned long long sw = __ballot(kept_out);
            if (lane == 0) sbits[idx >> 6] = sw;
        }
    }
    // the queue: each thread's marked nodes at its offset, an exclusive scan of the counts over the
    // block (within the wave by shuffles, over the four waves through LDS)
    const uint32_t mine = (uint32_t)__popc(pushm);
    uint32_t incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[t >> 6] = incl;
    // the counters summed over the wave by shuffles (two per word: a wave's sums stay below 2^11) and
    // added by its first lane -- an atomic from every lane would be expanded into a loop over the
    // lanes on the scalar unit
    uint32_t c01 = n_nodes | (n_unc << 16), c23 = n_band | (n_sign << 16);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        c01 += __shfl_xor(c01, o);
        if (SURR) c23 += __shfl_xor(c23, o);
    }
    if (lane == 0) {
        atomicAdd(&st_nodes, c01 & 0xffffu);
        if (SURR) {
            atomicAdd(&st_uncert, c01 >> 16);
            atomicAdd(&st_band, c23 & 0xffffu);
            atomicAdd(&st_sign, c23 >> 16);
        }
    }
    __syncthreads();
    uint32_t off = incl - mine;
    for (int w = 0; w < (t >> 6); ++w) off += wsum[w];
    const uint32_t nq = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    for (int q = 0; q < BRICK_NODES / 256; ++q)
        if ((pushm >> q) & 1u) queue[off++] = (uint16_t)(t + 256 * q);
    __syncthreads();
    for (uint32_t e = t; e < nq; e += 256) {
        const int idx = queue[e];
        const int gi = i0 + (idx & 15), gj = j0 + ((idx >> 4) & 15), gk = k0 + (idx >> 8);
        const T x = g.x0[0] + T(gi) * g.d[0], y = g.x0[1] + T(gj) * g.d[1], z = g.x0[2] + T(gk) * g.d[2];
        const T v = culled_fold<T>(staged ? sph : nullptr, nk, L, geo, x, y, z, scrp, o);
        phi[gi + m0 * (gj + m1 * gk)] = v;
        if (signs && node_outside(v))
            atomicOr(reinterpret_cast<unsigned int *>(sbits) + (idx >> 5), 1u << (idx & 31));
    }
    if (signs) {
        __syncthreads();
        if (t < BRICK_NODES / 64) signs[(size_t)b * (BRICK_NODES / 64) + t] = sbits[t];
    }
    if (t < PHI_BSTAT) {  // this brick's counters (summed on the host when asked for)
        const uint32_t v[PHI_BSTAT] = {st_nodes, nq, st_cert, st_uncert, st_band, st_sign, 0u, 0u};
        bstat[(size_t)b * PHI_BSTAT + t] = v[t];
    }
}

}  // namespace

// Rounding margins.  Every quantity the fold or the bounds round (coordinate differences, squared
// norms, square roots, radii, smooth minima) is within a few units in the last place of `scale`;
// the cull margin covers three such errors, the certificate margin the fold's accumulated rounding
// over every primitive as well.
template <typename T> static T cull_margin(double scale) { return T(256.0 * Eps<T>::v * scale); }
template <typename T> static T cert_margin(double scale, int nprims) {
    return T((4.0 * nprims + 256.0) * Eps<T>::v * scale);
}

template <typename T>
void launch_cull(const GridT<T> &g, const Geo &geo, double scale, uint16_t *list, uint32_t *cnt,
                 int8_t *cert, uint32_t *bstat, hipStream_t s) {
    int nb[3];
    phi_brick_counts(g.n, nb);
    hipLaunchKernelGGL((k_cull<T>), dim3(nb[0] * nb[1] * nb[2]), dim3(256), 0, s, g, geo,
                       cull_margin<T>(scale), nb[0], nb[1], list, cnt, cert, cert_margin<T>(scale, geo.nprims),
                       bstat);
}

template <typename T>
void launch_phi_bricks(T *phi, const GridT<T> &g, const Geo &geo, const uint16_t *list,
                       const uint32_t *cnt, const SurrogateArgs &sa, uint32_t *bstat, uint64_t *signs,
                       hipStream_t s) {
    int nb[3];
    phi_brick_counts(g.n, nb);
    const dim3 grid(nb[0] * nb[1] * nb[2]);
    const T dc = cert_margin<T>(sa.scale, geo.nprims);
    if (sa.nn)
        hipLaunchKernelGGL((k_phi_bricks<T, true>), grid, dim3(256), 0, s, phi, g, geo, list, cnt, sa, dc,
                           nb[0], nb[1], bstat, signs);
    else
        hipLaunchKernelGGL((k_phi_bricks<T, false>), grid, dim3(256), 0, s, phi, g, geo, list, cnt, sa, dc,
                           nb[0], nb[1], bstat, signs);
}

template void launch_cull<double>(const GridT<double> &, const Geo &, double, uint16_t *, uint32_t *, int8_t *,
                                  uint32_t *, hipStream_t);
template void launch_cull<float>(const GridT<float> &, const Geo &, double, uint16_t *, uint32_t *, int8_t *,
                                 uint32_t *, hipStream_t);
template void launch_phi_bricks<double>(double *, const GridT<double> &, const Geo &, const uint16_t *,
                                        const uint32_t *, const SurrogateArgs &, uint32_t *, uint64_t *,
                                        hipStream_t);
template void launch_phi_bricks<float>(float *, const GridT<float> &, const Geo &, const uint16_t *,
                                       const uint32_t *, const SurrogateArgs &, uint32_t *, uint64_t *,
                                       hipStream_t);

}  // namespace cc
