// The PLIC surface on the device: generate_mesh(cache::PLICCutCellCache, grid) and
// extract_surface_plic (src/plic/surface.jl:28-267).
//
// Every cell's stored plane n.xi = a is clipped to its unit cell xi in [0,1]^D (CartesianMeshes'
// cell_plane_clip, a dependency absent here, restated from the geometry): the corners lying on the
// plane, then the crossings of the edges whose two ends lie strictly on opposite sides, in world
// coordinates x_lower + xi * d with x_lower = x0 + i d (local_to_global).
//   2D: one segment per cell, its two ends ordered so that the right-hand normal of (p2 - p1) points
//       along n (out of the body); a plane that misses its cell or touches one corner gives a
//       zero-length segment; segments shorter than min_frac * min(d) are dropped (surface.jl:82-88).
//   3D: a polygon of 3 to 6 vertices per crossed cell, ordered by angle about its centroid in the
//       basis u = normalize(ref x n), v = n x u (ref = e_z unless |n_z| >= 0.9, then e_x)
//       (surface.jl:219-266), fan-triangulated from its first vertex; a triangle with
//       |cross|^2 < (2 min_frac min(d)^2)^2 is dropped, every vertex is kept (surface.jl:150-182).
// Two launches over the cells -- counts, then (after prefix sums) the elements at their offsets --
// keep the reference's cell order with no atomics: the surface is the same on every run.
//
// generate_mesh first marks, one bit per cell, the cells whose plane can touch the closed cell at
// all: some corner value <= 0 and some >= 0, the corner values formed by the very expression the
// clip uses, so every cell the clip would draw from is marked (a cell drawing nothing may be marked
// too).  The marked cells, ranked by a popcount scan, are the entries of the count and emit
// launches, in cell order, so the result is the all-cells result bit for bit.  Not used in 2D with a
// zero floor, where every cell draws a (possibly zero-length) segment.
#include "cc_internal.h"

namespace cc {
namespace {

constexpr int PS_THREADS = 256;

template <typename T> struct Poly {
    T p[6][3];
    int n;
};

constexpr int SQ[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
constexpr int CB[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// n.xi - a at the unit cell's corners (MS_NODE_BITS / MC_NODE_BITS order): the clip and the touch
// test both take their corner values from here.
template <typename T> __device__ __forceinline__ void corners2(const T *nv, T a, T (&f)[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) f[k] = nv[0] * T(SQ[k][0]) + nv[1] * T(SQ[k][1]) - a;
}
template <typename T> __device__ __forceinline__ void corners3(const T *nv, T a, T (&f)[8]) {
#pragma unroll
    for (int c = 0; c < 8; ++c) f[c] = (nv[0] * T(CB[c][0]) + nv[1] * T(CB[c][1])) + nv[2] * T(CB[c][2]) - a;
}

// Whether the clip of this plane can give any point: a nonzero normal, some corner value <= 0 and
// some >= 0 (a zero corner, or corners of both strict signs and so an edge between them).
template <typename T, int D> __device__ __forceinline__ bool touches(const T *nv, T a) {
    if (D == 2) {
        if (nv[0] == T(0) && nv[1] == T(0)) return false;
        T f[4];
        corners2(nv, a, f);
        bool le = false, ge = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            le |= f[k] <= T(0);
            ge |= f[k] >= T(0);
        }
        return le && ge;
    }
    if (nv[0] == T(0) && nv[1] == T(0) && nv[2] == T(0)) return false;
    T f[8];
    corners3(nv, a, f);
    bool le = false, ge = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        le |= f[k] <= T(0);
        ge |= f[k] >= T(0);
    }
    return le && ge;
}

// The 2D clip: 0, 1 or 2 points in the unit square (MS_NODE_BITS corner order).
template <typename T> __device__ __forceinline__ int clip2(const T *nv, T a, T (&xi)[2][2]) {
    xi[0][0] = xi[0][1] = xi[1][0] = xi[1][1] = T(0);
    if (nv[0] == T(0) && nv[1] == T(0)) return 0;
    const int (&C)[4][2] = SQ;
    T f[4];
    corners2(nv, a, f);
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (cnt < 2 && f[k] == T(0)) {
            xi[cnt][0] = T(C[k][0]);
            xi[cnt][1] = T(C[k][1]);
            ++cnt;
        }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int k2 = (k + 1) & 3;
        if (cnt < 2 && ((f[k] < T(0) && f[k2] > T(0)) || (f[k] > T(0) && f[k2] < T(0)))) {
            const T t = f[k] / (f[k] - f[k2]);
            xi[cnt][0] = T(C[k][0]) + t * T(C[k2][0] - C[k][0]);
            xi[cnt][1] = T(C[k][1]) + t * T(C[k2][1] - C[k][1]);
            ++cnt;
        }
    }
    return cnt;
}

// One cell's segment ends in world coordinates, oriented; its length.
template <typename T>
__device__ __forceinline__ T segment(const GridT<T> &g, const T *nv, T a, uint32_t i, uint32_t j, T (&p)[2][2]) {
    T xi[2][2];
    const int cnt = clip2(nv, a, xi);
    if (cnt == 1) {
        xi[1][0] = xi[0][0];
        xi[1][1] = xi[0][1];
    }
    if (cnt == 2) {
        const T v0 = xi[1][0] - xi[0][0], v1 = xi[1][1] - xi[0][1];
        if (v1 * nv[0] - v0 * nv[1] < T(0)) {
            const T t0 = xi[0][0], t1 = xi[0][1];
            xi[0][0] = xi[1][0];
            xi[0][1] = xi[1][1];
            xi[1][0] = t0;
            xi[1][1] = t1;
        }
    }
    const T lo[2] = {g.x0[0] + T(i) * g.d[0], g.x0[1] + T(j) * g.d[1]};
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int x = 0; x < 2; ++x) p[s][x] = lo[x] + xi[s][x] * g.d[x];
    const T l0 = p[1][0] - p[0][0], l1 = p[1][1] - p[0][1];
    return sqrt(l0 * l0 + l1 * l1);
}

// The 3D clip, ordered by angle, in world coordinates.
template <typename T>
__device__ __forceinline__ void polygon(const GridT<T> &g, const T *nv, T a, uint32_t i, uint32_t j, uint32_t k,
                                        Poly<T> &P) {
    P.n = 0;
    if (nv[0] == T(0) && nv[1] == T(0) && nv[2] == T(0)) return;
    const int (&B)[8][3] = CB;
    constexpr int E[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                              {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    T f[8];
    corners3(nv, a, f);
    T xi[6][3];
    int cnt = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (cnt < 6 && f[c] == T(0)) {
            xi[cnt][0] = T(B[c][0]);
            xi[cnt][1] = T(B[c][1]);
            xi[cnt][2] = T(B[c][2]);
            ++cnt;
        }
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const int p = E[e][0], q = E[e][1];
        if (cnt < 6 && ((f[p] < T(0) && f[q] > T(0)) || (f[p] > T(0) && f[q] < T(0)))) {
            const T t = f[p] / (f[p] - f[q]);
#pragma unroll
            for (int x = 0; x < 3; ++x) xi[cnt][x] = T(B[p][x]) + t * T(B[q][x] - B[p][x]);
            ++cnt;
        }
    }
    P.n = cnt;
    if (cnt < 3) return;
    const uint32_t ijk[3] = {i, j, k};
    T cen[3] = {T(0), T(0), T(0)};
    for (int m = 0; m < cnt; ++m)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            P.p[m][x] = (g.x0[x] + T(ijk[x]) * g.d[x]) + xi[m][x] * g.d[x];
            cen[x] += P.p[m][x];
        }
#pragma unroll
    for (int x = 0; x < 3; ++x) cen[x] = cen[x] / T(cnt);
    T ref[3] = {T(0), T(0), T(0)};
    if (fabs(nv[2]) < T(0.9)) ref[2] = T(1);
    else ref[0] = T(1);
    T u[3] = {ref[1] * nv[2] - ref[2] * nv[1], ref[2] * nv[0] - ref[0] * nv[2], ref[0] * nv[1] - ref[1] * nv[0]};
    const T un = sqrt((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
#pragma unroll
    for (int x = 0; x < 3; ++x) u[x] = u[x] / un;
    const T v[3] = {nv[1] * u[2] - nv[2] * u[1], nv[2] * u[0] - nv[0] * u[2], nv[0] * u[1] - nv[1] * u[0]};
    T ang[6];
    for (int m = 0; m < cnt; ++m) {
        const T d0 = P.p[m][0] - cen[0], d1 = P.p[m][1] - cen[1], d2 = P.p[m][2] - cen[2];
        ang[m] = atan2((d0 * v[0] + d1 * v[1]) + d2 * v[2], (d0 * u[0] + d1 * u[1]) + d2 * u[2]);
    }
    for (int m = 1; m < cnt; ++m) {  // insertion sort by angle (surface.jl:251-263)
        const T k0 = P.p[m][0], k1 = P.p[m][1], k2 = P.p[m][2], ka = ang[m];
        int q = m - 1;
        while (q >= 0 && ang[q] > ka) {
            P.p[q + 1][0] = P.p[q][0];
            P.p[q + 1][1] = P.p[q][1];
            P.p[q + 1][2] = P.p[q][2];
            ang[q + 1] = ang[q];
            --q;
        }
        P.p[q + 1][0] = k0;
        P.p[q + 1][1] = k1;
        P.p[q + 1][2] = k2;
        ang[q + 1] = ka;
    }
}

// Whether fan triangle m (vertices 0, m, m+1) is kept.
template <typename T> __device__ __forceinline__ bool keep_tri(const Poly<T> &P, int m, double min_cross_sq) {
    T e1[3], e2[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        e1[x] = P.p[m][x] - P.p[0][x];
        e2[x] = P.p[m + 1][x] - P.p[0][x];
    }
    const T c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2], c2 = e1[0] * e2[1] - e1[1] * e2[0];
    const T s2 = (c0 * c0 + c1 * c1) + c2 * c2;
    return (double)s2 >= min_cross_sq;
}

// Entry c draws, in cell cells[c] (a 0-based linear index; cell c when cells is null), the plane
// stored at index c (extract_surface_plic: the caller's planes, one per entry) or, with by_cell, the
// plane stored at the cell's own index (generate_mesh over the marked cells).
template <typename T, int D>
__global__ __launch_bounds__(PS_THREADS) void k_plic_surf(GridT<T> g, const T *__restrict__ normals,
                                                          const T *__restrict__ intercepts,
                                                          const int64_t *__restrict__ cells, bool by_cell,
                                                          size_t ncell,
                                                          double min_len, double min_cross_sq, bool emit,
                                                          uint32_t *__restrict__ vcnt, uint32_t *__restrict__ ecnt,
                                                          const int64_t *__restrict__ vpre,
                                                          const int64_t *__restrict__ epre, T *__restrict__ nodes,
                                                          int64_t *__restrict__ elems) {
    const size_t c = (size_t)blockIdx.x * PS_THREADS + threadIdx.x;
    if (c >= ncell) return;
    if (emit && vcnt[c] == 0) return;  // nothing of this entry is drawn: skip its plane
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    const size_t lin = cells ? (size_t)cells[c] : c;
    const uint32_t i = (uint32_t)(lin % n0), j = (uint32_t)((lin / n0) % n1), k = (uint32_t)(lin / ((size_t)n0 * n1));
    const size_t pl = by_cell ? lin : c;
    const T *nv = normals + D * pl;
    const T a = intercepts[pl];
    if (D == 2) {
        T p[2][2];
        const T len = segment(g, nv, a, i, j, p);
        const bool kept = (double)len >= min_len;
        if (!emit) {
            vcnt[c] = kept ? 2u : 0u;
            ecnt[c] = kept ? 1u : 0u;
        } else if (kept) {
            const int64_t v0 = vpre[c], e0 = epre[c];
            nodes[2 * v0] = p[0][0];
            nodes[2 * v0 + 1] = p[0][1];
            nodes[2 * v0 + 2] = p[1][0];
            nodes[2 * v0 + 3] = p[1][1];
            elems[2 * e0] = v0 + 1;
            elems[2 * e0 + 1] = v0 + 2;
        }
    } else {
        Poly<T> P;
        polygon(g, nv, a, i, j, k, P);
        if (!emit) {
            uint32_t nt = 0;
            if (P.n >= 3)
                for (int m = 1; m < P.n - 1; ++m) nt += keep_tri(P, m, min_cross_sq) ? 1u : 0u;
            vcnt[c] = P.n >= 3 ? (uint32_t)P.n : 0u;
            ecnt[c] = nt;
        } else if (P.n >= 3) {
            const int64_t v0 = vpre[c];
            int64_t e = epre[c];
            for (int m = 0; m < P.n; ++m) {
                nodes[3 * (v0 + m)] = P.p[m][0];
                nodes[3 * (v0 + m) + 1] = P.p[m][1];
                nodes[3 * (v0 + m) + 2] = P.p[m][2];
            }
            for (int m = 1; m < P.n - 1; ++m)
                if (keep_tri(P, m, min_cross_sq)) {
                    elems[3 * e] = v0 + 1;
                    elems[3 * e + 1] = v0 + m + 1;
                    elems[3 * e + 2] = v0 + m + 2;
                    ++e;
                }
        }
    }
}

// One bit per cell (bit c % 64 of word c / 64): whether cell c's plane can touch it.
template <typename T, int D>
__global__ __launch_bounds__(PS_THREADS) void k_plic_touch(const T *__restrict__ normals,
                                                           const T *__restrict__ intercepts, size_t ncell,
                                                           size_t nwords, unsigned long long *__restrict__ bits) {
    const size_t c = (size_t)blockIdx.x * PS_THREADS + threadIdx.x;
    bool t = false;
    if (c < ncell) t = touches<T, D>(normals + D * c, intercepts[c]);
    const unsigned long long m = __ballot(t);
    const size_t w = c / 64;  // lane 0 holds a multiple of 64: the wave's word
    if ((threadIdx.x & 63) == 0 && w < nwords) bits[w] = m;
}

// The marked cells in cell order: word w's set bits at wrank[w] onwards.
__global__ __launch_bounds__(PS_THREADS) void k_plic_marked(const unsigned long long *__restrict__ bits,
                                                            const int64_t *__restrict__ wrank, size_t nwords,
                                                            int64_t *__restrict__ cells) {
    const size_t w = (size_t)blockIdx.x * PS_THREADS + threadIdx.x;
    if (w >= nwords) return;
    unsigned long long m = bits[w];
    int64_t r = wrank[w];
    while (m) {
        cells[r++] = (int64_t)(w * 64 + (size_t)(__ffsll((long long)m) - 1));
        m &= m - 1;
    }
}

}  // namespace

template <typename T>
void launch_plic_touch(const T *normals, const T *intercepts, size_t ncell, int dim, unsigned long long *bits,
                       int64_t *wrank, unsigned long long *part, size_t part_cap, int64_t *totals, hipStream_t s) {
    const size_t nwords = (ncell + 63) / 64;
    const dim3 grid((unsigned)((ncell + PS_THREADS - 1) / PS_THREADS));
    if (dim == 3)
        hipLaunchKernelGGL((k_plic_touch<T, 3>), grid, dim3(PS_THREADS), 0, s, normals, intercepts, ncell, nwords, bits);
    else
        hipLaunchKernelGGL((k_plic_touch<T, 2>), grid, dim3(PS_THREADS), 0, s, normals, intercepts, ncell, nwords, bits);
    launch_scan_pair({nullptr, bits, wrank, nullptr, nwords}, {nullptr, nullptr, nullptr, nullptr, 0}, part, part_cap,
                     totals, s);
}

void launch_plic_marked(const unsigned long long *bits, const int64_t *wrank, size_t ncell, int64_t *cells,
                        hipStream_t s) {
    const size_t nwords = (ncell + 63) / 64;
    hipLaunchKernelGGL(k_plic_marked, dim3((unsigned)((nwords + PS_THREADS - 1) / PS_THREADS)), dim3(PS_THREADS), 0, s,
                       bits, wrank, nwords, cells);
}

template void launch_plic_touch<double>(const double *, const double *, size_t, int, unsigned long long *, int64_t *,
                                        unsigned long long *, size_t, int64_t *, hipStream_t);
template void launch_plic_touch<float>(const float *, const float *, size_t, int, unsigned long long *, int64_t *,
                                       unsigned long long *, size_t, int64_t *, hipStream_t);

template <typename T>
void launch_plic_surface(const GridT<T> &g, int dim, const T *normals, const T *intercepts, const int64_t *cells,
                         bool by_cell, size_t ncell, double min_frac, const PlicSurfBuffers &b, bool emit, T *nodes,
                         int64_t *elems, hipStream_t s) {
    if (ncell == 0) {
        if (!emit) launch_scan_pair({b.vcnt, nullptr, b.vpre, nullptr, 0}, {b.ecnt, nullptr, b.epre, nullptr, 0},
                                    b.part, b.part_cap, b.totals, s);
        return;
    }
    double dmin = (double)g.d[0] < (double)g.d[1] ? (double)g.d[0] : (double)g.d[1];
    if (dim == 3 && (double)g.d[2] < dmin) dmin = (double)g.d[2];
    const double min_len = min_frac * dmin;
    const double min_area = min_frac * dmin * dmin;
    const double min_cross_sq = (2.0 * min_area) * (2.0 * min_area);
    const dim3 grid((unsigned)((ncell + PS_THREADS - 1) / PS_THREADS));
    if (dim == 3)
        hipLaunchKernelGGL((k_plic_surf<T, 3>), grid, dim3(PS_THREADS), 0, s, g, normals, intercepts, cells, by_cell, ncell, min_len,
                           min_cross_sq, emit, b.vcnt, b.ecnt, (const int64_t *)b.vpre, (const int64_t *)b.epre,
                           nodes, elems);
    else
        hipLaunchKernelGGL((k_plic_surf<T, 2>), grid, dim3(PS_THREADS), 0, s, g, normals, intercepts, cells, by_cell, ncell, min_len,
                           min_cross_sq, emit, b.vcnt, b.ecnt, (const int64_t *)b.vpre, (const int64_t *)b.epre,
                           nodes, elems);
    if (!emit)
        launch_scan_pair({b.vcnt, nullptr, b.vpre, nullptr, ncell}, {b.ecnt, nullptr, b.epre, nullptr, ncell},
                         b.part, b.part_cap, b.totals, s);
}

template void launch_plic_surface<double>(const GridT<double> &, int, const double *, const double *,
                                          const int64_t *, bool, size_t, double, const PlicSurfBuffers &, bool,
                                          double *, int64_t *, hipStream_t);
template void launch_plic_surface<float>(const GridT<float> &, int, const float *, const float *, const int64_t *,
                                         bool, size_t, double, const PlicSurfBuffers &, bool, float *, int64_t *,
                                         hipStream_t);

}  // namespace cc
