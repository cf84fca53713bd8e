// Marching-squares cell pass for gfx950 -- _nodal_cells_kernel! on a 2D grid
// (src/marching_squares/cache.jl:75-78) with cut_cell_moments (src/marching_squares/moments.jl:66-163).
//
// One thread per cell, the whole construction inline: a 2D grid is a few hundred thousand cells at
// most in the workloads this serves (40 000 in the 200x200 circle), so the pass is launch-latency
// bound and a single kernel is the right shape.  Arithmetic follows the reference expression by
// expression (no contraction), so the results are the oracle's bit for bit.
#include "cc_ms_dev.h"

namespace cc {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_ms_cells(const T *__restrict__ phi, GridT<T> g, Cols<T> cols,
                                                  uint32_t *__restrict__ counter) {
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    bool cut = false;
    if (c < n0 * n1) {
        const uint32_t i = c % n0, j = c / n0;
        const size_t sy = n0 + 1, base = i + sy * j;
        // MS_NODE_BITS order: (0,0), (1,0), (1,1), (0,1)
        const T ph[4] = {phi[base], phi[base + 1], phi[base + sy + 1], phi[base + sy]};
        const T xl = g.x0[0] + T(i) * g.d[0], xh = g.x0[0] + T(i + 1) * g.d[0];
        const T yl = g.x0[1] + T(j) * g.d[1], yh = g.x0[1] + T(j + 1) * g.d[1];
        const T nx[4] = {xl, xh, xh, xl}, ny[4] = {yl, yl, yh, yh};
        cut = ms_cell(nx, ny, ph, g.d[0], g.d[1], cols, c);
    }
    add_block_cuts(cut ? 1u : 0u, counter);
}

// The same pass sampling an analytic body itself (_sample_nodes_kernel! and _nodal_cells_kernel!,
// src/marching_squares/cache.jl:67-78, in one launch): each cell evaluates its four corners at the
// lattice points the sampling pass uses, x0 + T(i) d, so its values are phi's bit for bit, and
// writes the nodes it owns -- its lower-left corner, and on the grid's far edges the corners
// beyond -- so phi is written once.  A 200 x 200 update is launch-bound; one launch fewer.
template <typename T>
__global__ __launch_bounds__(256) void k_ms_cells_fused(Geo geo, T *__restrict__ phi, GridT<T> g, Cols<T> cols,
                                                        uint32_t *__restrict__ counter) {
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    bool cut = false;
    if (c < n0 * n1) {
        const uint32_t i = c % n0, j = c / n0;
        const size_t sy = n0 + 1, base = i + sy * j;
        const T xl = g.x0[0] + T(i) * g.d[0], xh = g.x0[0] + T(i + 1) * g.d[0];
        const T yl = g.x0[1] + T(j) * g.d[1], yh = g.x0[1] + T(j + 1) * g.d[1];
        const T nx[4] = {xl, xh, xh, xl}, ny[4] = {yl, yl, yh, yh};
        T ph[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T x[3] = {nx[q], ny[q], T(0)};
            ph[q] = sdf_value<T, 2>(geo, x);
        }
        phi[base] = ph[0];
        if (i + 1 == n0) phi[base + 1] = ph[1];
        if (j + 1 == n1) phi[base + sy] = ph[3];
        if (i + 1 == n0 && j + 1 == n1) phi[base + sy + 1] = ph[2];
        cut = ms_cell(nx, ny, ph, g.d[0], g.d[1], cols, c);
    }
    add_block_cuts(cut ? 1u : 0u, counter);
}

// _leaf_cells_kernel! (src/marching_squares/cache.jl:129-137): a thread per quadtree leaf (a
// grid-stride loop past 2048 blocks, so the cut count takes at most 2048 atomics).  Its corners
// come off the integer lattice, x0 + T(coord + bit) h with h = d / 2^level
// (cell_nodes(mesh, c), marching_squares.jl:113-121): a corner two leaves share across a level jump
// is then one product, (2a)(h/2) = a h exactly, so both sample the body at the same point.  The
// values are sampled there (cell_values(field, mesh, c), :152-155) and kept, four per leaf -- or,
// with SAMPLE false, were put there by the caller (a field the caller samples).
template <typename T, bool SAMPLE>
__global__ __launch_bounds__(256) void k_leaf_cells(Geo geo, const int32_t *__restrict__ level,
                                                    const int32_t *__restrict__ coord, uint32_t n, T x0x, T x0y,
                                                    T dx, T dy, T *__restrict__ phi4, Cols<T> cols,
                                                    uint32_t *__restrict__ counter) {
    uint32_t cuts = 0;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const int lv = level[c];
        const T hx = ldexp(dx, -lv), hy = ldexp(dy, -lv);  // exact: a power-of-two scaling
        const int cx = coord[2 * (size_t)c], cy = coord[2 * (size_t)c + 1];
        const T xl = x0x + T(cx) * hx, xh = x0x + T(cx + 1) * hx;
        const T yl = x0y + T(cy) * hy, yh = x0y + T(cy + 1) * hy;
        const T nx[4] = {xl, xh, xh, xl}, ny[4] = {yl, yl, yh, yh};
        T ph[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (SAMPLE) {
                const T x[2] = {nx[k], ny[k]};
                ph[k] = sdf_value<T, 2>(geo, x);
                phi4[4 * (size_t)c + k] = ph[k];
            } else {
                ph[k] = phi4[4 * (size_t)c + k];
            }
        }
        cuts += ms_cell(nx, ny, ph, hx, hy, cols, c) ? 1u : 0u;
    }
    add_block_cuts(cuts, counter);
}

// generate_mesh over a grid or a tree (src/marching_squares/surface.jl:59-110) draws only the cut
// cells: one bit per cell or leaf, ranked by a popcount scan, then each cut one's index and four
// corner values (and, for a tree held on the device, its level and lattice corner) gathered in
// order, so only they cross to the host contour.  A cell is cut when its four corner values do not
// all share a sign (phi >= 0 is outside) -- _ms_screen_cell (src/marching_squares/surface.jl:131-135),
// read from the field itself, so a field changed through the column's pointer is screened afresh.
// n0 > 0: a grid n0 cells wide over the nodal field; n0 == 0: leaves, four values each.
template <typename T>
__global__ __launch_bounds__(256) void k_sign_bits(const T *__restrict__ phi, size_t n, size_t n0, size_t nwords,
                                                   unsigned long long *__restrict__ bits) {
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    bool cut = false;
    if (c < n) {
        T p[4];
        if (n0) {
            const size_t i = c % n0, j = c / n0, sy = n0 + 1, base = i + sy * j;
            p[0] = phi[base];
            p[1] = phi[base + 1];
            p[2] = phi[base + sy + 1];
            p[3] = phi[base + sy];
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) p[q] = phi[4 * c + q];
        }
        const bool s0 = p[0] >= T(0), s1 = p[1] >= T(0), s2 = p[2] >= T(0), s3 = p[3] >= T(0);
        cut = !(s0 == s1 && s1 == s2 && s2 == s3);
    }
    const unsigned long long m = __ballot(cut);
    const size_t w = c / 64;  // lane 0 holds a multiple of 64: the wave's word
    if ((threadIdx.x & 63) == 0 && w < nwords) bits[w] = m;
}

template <typename T>
__global__ __launch_bounds__(256) void k_leaf_gather(const unsigned long long *__restrict__ bits,
                                                     const int64_t *__restrict__ wrank, size_t nwords,
                                                     const T *__restrict__ phi4, const int32_t *__restrict__ level,
                                                     const int32_t *__restrict__ coord, int64_t *__restrict__ idx,
                                                     T *__restrict__ phi_out, int32_t *__restrict__ lc_out) {
    const size_t w = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (w >= nwords) return;
    unsigned long long m = bits[w];
    int64_t r = wrank[w];
    while (m) {
        const size_t c = w * 64 + (size_t)(__ffsll((long long)m) - 1);
        idx[r] = (int64_t)c;
#pragma unroll
        for (int q = 0; q < 4; ++q) phi_out[4 * r + q] = phi4[4 * c + q];
        if (level) {
            lc_out[3 * r] = level[c];
            lc_out[3 * r + 1] = coord[2 * c];
            lc_out[3 * r + 2] = coord[2 * c + 1];
        }
        ++r;
        m &= m - 1;
    }
}

// The grid contour's gather: cut cell c's index and its four corner values from the nodal field
// (MS_NODE_BITS order), in cell order.
template <typename T>
__global__ __launch_bounds__(256) void k_grid_gather(const unsigned long long *__restrict__ bits,
                                                     const int64_t *__restrict__ wrank, size_t nwords,
                                                     const T *__restrict__ phi, size_t n0, int64_t *__restrict__ idx,
                                                     T *__restrict__ phi_out) {
    const size_t w = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (w >= nwords) return;
    unsigned long long m = bits[w];
    int64_t r = wrank[w];
    const size_t sy = n0 + 1;
    while (m) {
        const size_t c = w * 64 + (size_t)(__ffsll((long long)m) - 1);
        const size_t i = c % n0, j = c / n0, base = i + sy * j;
        idx[r] = (int64_t)c;
        phi_out[4 * r] = phi[base];
        phi_out[4 * r + 1] = phi[base + 1];
        phi_out[4 * r + 2] = phi[base + sy + 1];
        phi_out[4 * r + 3] = phi[base + sy];
        ++r;
        m &= m - 1;
    }
}

}  // namespace

template <typename T>
void launch_grid_gather(const unsigned long long *bits, const int64_t *wrank, size_t n, const T *phi, size_t n0,
                        int64_t *idx, T *phi_out, hipStream_t s) {
    const size_t nwords = (n + 63) / 64;
    hipLaunchKernelGGL((k_grid_gather<T>), dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, bits, wrank,
                       nwords, phi, n0, idx, phi_out);
}

template void launch_grid_gather<double>(const unsigned long long *, const int64_t *, size_t, const double *, size_t,
                                         int64_t *, double *, hipStream_t);
template void launch_grid_gather<float>(const unsigned long long *, const int64_t *, size_t, const float *, size_t,
                                        int64_t *, float *, hipStream_t);

template <typename T>
void launch_cut_marks(const T *phi, size_t n, size_t n0, unsigned long long *bits, int64_t *wrank,
                      unsigned long long *part, size_t part_cap, int64_t *totals, hipStream_t s) {
    const size_t nwords = (n + 63) / 64;
    hipLaunchKernelGGL((k_sign_bits<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, phi, n, n0, nwords, bits);
    launch_scan_pair({nullptr, bits, wrank, nullptr, nwords}, {nullptr, nullptr, nullptr, nullptr, 0}, part, part_cap,
                     totals, s);
}

template void launch_cut_marks<double>(const double *, size_t, size_t, unsigned long long *, int64_t *,
                                       unsigned long long *, size_t, int64_t *, hipStream_t);
template void launch_cut_marks<float>(const float *, size_t, size_t, unsigned long long *, int64_t *,
                                      unsigned long long *, size_t, int64_t *, hipStream_t);

template <typename T>
void launch_leaf_gather(const unsigned long long *bits, const int64_t *wrank, size_t n, const T *phi4,
                        const int32_t *level, const int32_t *coord, int64_t *idx, T *phi_out, int32_t *lc_out,
                        hipStream_t s) {
    const size_t nwords = (n + 63) / 64;
    hipLaunchKernelGGL((k_leaf_gather<T>), dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s, bits, wrank,
                       nwords, phi4, level, coord, idx, phi_out, lc_out);
}

template void launch_leaf_gather<double>(const unsigned long long *, const int64_t *, size_t, const double *,
                                         const int32_t *, const int32_t *, int64_t *, double *, int32_t *,
                                         hipStream_t);
template void launch_leaf_gather<float>(const unsigned long long *, const int64_t *, size_t, const float *,
                                        const int32_t *, const int32_t *, int64_t *, float *, int32_t *, hipStream_t);

template <typename T>
void launch_ms_cells(const T *phi, const GridT<T> &g, const Cols<T> &cols, uint32_t *counter,
                     hipStream_t s) {
    const uint32_t ncell = (uint32_t)g.n[0] * g.n[1];
    hipLaunchKernelGGL((k_ms_cells<T>), dim3((ncell + 255) / 256), dim3(256), 0, s, phi, g, cols,
                       counter);
}

template <typename T>
void launch_ms_cells_fused(const Geo &geo, T *phi, const GridT<T> &g, const Cols<T> &cols, uint32_t *counter,
                           hipStream_t s) {
    const uint32_t ncell = (uint32_t)g.n[0] * g.n[1];
    hipLaunchKernelGGL((k_ms_cells_fused<T>), dim3((ncell + 255) / 256), dim3(256), 0, s, geo, phi, g, cols,
                       counter);
}

template <typename T>
void launch_leaf_cells(const Geo *geo, const int32_t *level, const int32_t *coord, uint32_t n, const double *x0,
                       const double *d, T *phi4, const Cols<T> &cols, uint32_t *counter, hipStream_t s) {
    if (n == 0) return;
    const uint32_t blocks = (n + 255) / 256;
    const dim3 grid(blocks < 2048 ? blocks : 2048);  // leaves beyond one per thread: a grid-stride loop
    if (geo)
        hipLaunchKernelGGL((k_leaf_cells<T, true>), grid, dim3(256), 0, s, *geo, level, coord, n, T(x0[0]),
                           T(x0[1]), T(d[0]), T(d[1]), phi4, cols, counter);
    else
        hipLaunchKernelGGL((k_leaf_cells<T, false>), grid, dim3(256), 0, s, Geo{}, level, coord, n, T(x0[0]),
                           T(x0[1]), T(d[0]), T(d[1]), phi4, cols, counter);
}

template void launch_leaf_cells<double>(const Geo *, const int32_t *, const int32_t *, uint32_t, const double *,
                                        const double *, double *, const Cols<double> &, uint32_t *, hipStream_t);
template void launch_leaf_cells<float>(const Geo *, const int32_t *, const int32_t *, uint32_t, const double *,
                                       const double *, float *, const Cols<float> &, uint32_t *, hipStream_t);
template void launch_ms_cells_fused<double>(const Geo &, double *, const GridT<double> &, const Cols<double> &,
                                            uint32_t *, hipStream_t);
template void launch_ms_cells_fused<float>(const Geo &, float *, const GridT<float> &, const Cols<float> &,
                                           uint32_t *, hipStream_t);
template void launch_ms_cells<double>(const double *, const GridT<double> &, const Cols<double> &,
                                      uint32_t *, hipStream_t);
template void launch_ms_cells<float>(const float *, const GridT<float> &, const Cols<float> &,
                                     uint32_t *, hipStream_t);

}  // namespace cc
