// The marching-squares per-cell construction (cut_cell_moments, src/marching_squares/moments.jl:66-163)
// as one device function: the corners (MS_NODE_BITS order), their values and the cell's size in,
// one CutCellData record written to the columns.  Shared by the grid pass and the quadtree pass
// (src/marching_squares/cache.jl:75-78 and :129-137 build every cell with the same primitive).
// Arithmetic follows the reference expression by expression (no contraction).
#pragma once
#include "cc_internal.h"

namespace cc {

__device__ __forceinline__ int nx4(int k) { return (k + 1) & 3; }

// MS_DIRECTION_EDGE (src/marching_squares/marching_squares.jl:82), 0-based: direction -> edge.
__device__ __forceinline__ int dir_edge(int dir) { return dir == 0 ? 3 : (dir == 1 ? 1 : (dir == 2 ? 0 : 2)); }

// Writes cell c's record; returns whether the cell is cut.  xl, yl: the lower corner (nodes[1]);
// hx, hy: the cell's size (the centre is xl + h/2, the volume hx hy).
template <typename T>
__device__ __forceinline__ bool ms_cell(const T (&nx)[4], const T (&ny)[4], const T (&ph)[4], T hx, T hy,
                                        const Cols<T> &cols, size_t c) {
    const T xl = nx[0], yl = ny[0];
    const T cx = xl + T(0.5) * hx, cy = yl + T(0.5) * hy;
    bool cut = false;
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
        vf = area / (hx * hy);
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
    return cut;
}

// Adds the block's cut cells to the cache's counter with one atomic per block (256 threads, every
// thread arriving): a counter taking one atomic per wave serialises them at one L2 address.
__device__ __forceinline__ void add_block_cuts(uint32_t v, uint32_t *counter) {
    __shared__ uint32_t per_wave[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) per_wave[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = (per_wave[0] + per_wave[1]) + (per_wave[2] + per_wave[3]);
        if (t) atomicAdd(counter, t);
    }
}

}  // namespace cc
