// Marching-squares cell pass for gfx950 -- _nodal_cells_kernel! on a 2D grid
// (src/marching_squares/cache.jl:75-78) with cut_cell_moments (src/marching_squares/moments.jl:66-163).
//
// One thread per cell, the whole construction inline: a 2D grid is a few hundred thousand cells at
// most in the workloads this serves (40 000 in the 200x200 circle), so the pass is launch-latency
// bound and a single kernel is the right shape.  Arithmetic follows the reference expression by
// expression (no contraction), so the results are the oracle's bit for bit.
#include "cc_internal.h"

namespace cc {
namespace {

__device__ __forceinline__ int nx4(int k) { return (k + 1) & 3; }

// MS_DIRECTION_EDGE (src/marching_squares/marching_squares.jl:82), 0-based: direction -> edge.
__device__ __forceinline__ int dir_edge(int dir) { return dir == 0 ? 3 : (dir == 1 ? 1 : (dir == 2 ? 0 : 2)); }

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
        const T cx = xl + T(0.5) * g.d[0], cy = yl + T(0.5) * g.d[1];
        int nout = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) nout += ph[k] >= T(0) ? 1 : 0;

        // _full_face_centre + _face_offset (moments.jl:39-51): the in-plane offset of a face's centre
        T fcl[4], ff[4];
#pragma unroll
        for (int dir = 0; dir < 4; ++dir) {
            const int k = dir_edge(dir), k2 = nx4(k);
            fcl[dir] = (dir >> 1) == 0 ? T(0.5) * (ny[k] + ny[k2]) - yl : T(0.5) * (nx[k] + nx[k2]) - xl;
        }
        T vf, cenx = cx, ceny = cy, icx = cx, icy = cy;
        int8_t kind;
        bool amb = false;
        if (nout == 4 || nout == 0) {  // moments.jl:84-93
            const bool out = nout == 4;
            kind = out ? CELL_OUTSIDE : CELL_INSIDE;
            vf = out ? T(1) : T(0);
#pragma unroll
            for (int dir = 0; dir < 4; ++dir) ff[dir] = vf;
        } else {
            cut = true;
            kind = CELL_CUT;
            // cell_boundary_walk (marching_squares.jl:240-251)
            T fr[4], sax[4], say[4], sbx[4], sby[4];
            bool pres[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int j2 = nx4(k);
                const T fa = ph[k], fb = ph[j2];
                fr[k] = open_fraction(fa, fb);
                pres[k] = fa >= T(0) || fb >= T(0);
                if (fa >= T(0) && fb >= T(0)) {
                    sax[k] = nx[k]; say[k] = ny[k]; sbx[k] = nx[j2]; sby[k] = ny[j2];
                } else if (fa >= T(0)) {
                    sax[k] = nx[k]; say[k] = ny[k];
                    sbx[k] = nx[k] + fr[k] * (nx[j2] - nx[k]);
                    sby[k] = ny[k] + fr[k] * (ny[j2] - ny[k]);
                } else {
                    sax[k] = nx[j2] + fr[k] * (nx[k] - nx[j2]);
                    say[k] = ny[j2] + fr[k] * (ny[k] - ny[j2]);
                    sbx[k] = nx[j2]; sby[k] = ny[j2];
                }
            }
#pragma unroll
            for (int dir = 0; dir < 4; ++dir) {  // moments.jl:105-110
                const int k = dir_edge(dir);
                ff[dir] = fr[k];
                if (pres[k])
                    fcl[dir] = (dir >> 1) == 0 ? T(0.5) * (say[k] + sby[k]) - yl
                                               : T(0.5) * (sax[k] + sbx[k]) - xl;
            }
            T area2 = T(0), cm0 = T(0), cm1 = T(0);  // moments.jl:117-127
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!pres[k]) continue;
                const T la0 = sax[k] - xl, la1 = say[k] - yl, lb0 = sbx[k] - xl, lb1 = sby[k] - yl;
                const T cr = la0 * lb1 - la1 * lb0;
                area2 += cr;
                cm0 += (la0 + lb0) * cr;
                cm1 += (la1 + lb1) * cr;
            }
            // cell_gap_segments (marching_squares.jl:291-317) with the interface sums (moments.jl:129-150)
            int nif = 0;
            T ilen = T(0), icm0 = T(0), icm1 = T(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!pres[k]) continue;
                int jn = k;  // _next_present
                if (pres[nx4(k)]) jn = nx4(k);
                else if (pres[nx4(nx4(k))]) jn = nx4(nx4(k));
                else if (pres[nx4(nx4(nx4(k)))]) jn = nx4(nx4(nx4(k)));
                const T ax = sbx[k], ay = sby[k];
                T bx = sax[0], by = say[0];
                bx = jn == 1 ? sax[1] : bx; by = jn == 1 ? say[1] : by;
                bx = jn == 2 ? sax[2] : bx; by = jn == 2 ? say[2] : by;
                bx = jn == 3 ? sax[3] : bx; by = jn == 3 ? say[3] : by;
                if (ax != bx || ay != by) {
                    nif += 1;
                    if (nif <= 2) {
                        const T la0 = ax - xl, la1 = ay - yl, lb0 = bx - xl, lb1 = by - yl;
                        const T cr = la0 * lb1 - la1 * lb0;
                        area2 += cr;
                        cm0 += (la0 + lb0) * cr;
                        cm1 += (la1 + lb1) * cr;
                        const T vx = bx - ax, vy = by - ay;
                        const T len = sqrt(vx * vx + vy * vy);
                        ilen += len;
                        const T hl = T(0.5) * len;
                        icm0 += hl * (ax + bx);
                        icm1 += hl * (ay + by);
                    }
                }
            }
            amb = nif > 1;
            const T area = T(0.5) * area2;
            if (area2 != T(0)) {
                const T den = T(3) * area2;
                cenx = xl + cm0 / den;
                ceny = yl + cm1 / den;
            }
            if (ilen > T(0)) {
                icx = icm0 / ilen;
                icy = icm1 / ilen;
            }
            vf = area / (g.d[0] * g.d[1]);
        }
        cols.kind[c] = kind;
        cols.amb[c] = amb ? 1 : 0;
        cols.vf[c] = vf;
        cols.centroid[2 * (size_t)c] = cenx;
        cols.centroid[2 * (size_t)c + 1] = ceny;
        cols.ic[2 * (size_t)c] = icx;
        cols.ic[2 * (size_t)c + 1] = icy;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cols.ff[4 * (size_t)c + k] = ff[k];
            cols.fcl[4 * (size_t)c + k] = fcl[k];
        }
    }
    const uint64_t m = __ballot(cut);
    if (m && __lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(counter, (uint32_t)__popcll(m));
}

}  // namespace

template <typename T>
void launch_ms_cells(const T *phi, const GridT<T> &g, const Cols<T> &cols, uint32_t *counter,
                     hipStream_t s) {
    const uint32_t ncell = (uint32_t)g.n[0] * g.n[1];
    hipLaunchKernelGGL((k_ms_cells<T>), dim3((ncell + 255) / 256), dim3(256), 0, s, phi, g, cols,
                       counter);
}

template void launch_ms_cells<double>(const double *, const GridT<double> &, const Cols<double> &,
                                      uint32_t *, hipStream_t);
template void launch_ms_cells<float>(const float *, const GridT<float> &, const Cols<float> &,
                                     uint32_t *, hipStream_t);

}  // namespace cc
