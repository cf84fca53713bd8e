// The marching-squares contour on the device: generate_mesh over a grid or a quadtree's leaves
// (src/marching_squares/surface.jl:59-110, 153-177) for the cut cells the screen found.
//
// The reference walks the cut cells serially in cell (or leaf) order, pushing each cell's
// cell_interface segments (marching_squares.jl:343-351) and welding their ends through a Dict keyed
// by the exact, +0-normalised coordinates, so a vertex's number is the order in which its coordinate
// is first reached.  Here the same numbering comes from data-parallel steps:
//   1. each cut cell counts its segments; an exclusive scan places them in cell order;
//   2. each cell writes its segments' two ends in push order (the end `cell_interface` returns first,
//      then the other): end e of segment g is emission 2g + e;
//   3. the emissions are sorted by coordinate, stably (two 64-bit radix passes, y then x), so each
//      group of equal coordinates starts with its first emission;
//   4. the group heads, flagged in emission order and scanned, number the vertices in first-reach
//      order; every emission takes its group head's number.
// The result is the reference's nodes and lines exactly, whatever the order of the parallel work.
#include <hipcub/hipcub.hpp>

#include "cc_internal.h"

namespace cc {
namespace {

constexpr int CT = 256;

template <typename T> struct Bits;
template <> struct Bits<double> {
    static __device__ __forceinline__ uint64_t of(double v) { return (uint64_t)__double_as_longlong(v); }
};
template <> struct Bits<float> {
    static __device__ __forceinline__ uint64_t of(float v) { return (uint64_t)(uint32_t)__float_as_uint(v); }
};

// Cut cell r's four corners (MS_NODE_BITS order) and values.
template <typename T>
__device__ __forceinline__ void entry(const ContourCells<T> &in, int64_t r, T (&nx)[4], T (&ny)[4], T (&ph)[4]) {
    T xl, xh, yl, yh;
    if (in.lc) {  // a leaf: level, lattice corner; h = d / 2^level (exact)
        const int lv = in.lc[3 * r], cx = in.lc[3 * r + 1], cy = in.lc[3 * r + 2];
        const T hx = ldexp(in.dx, -lv), hy = ldexp(in.dy, -lv);
        xl = in.x0 + T(cx) * hx;
        xh = in.x0 + T(cx + 1) * hx;
        yl = in.y0 + T(cy) * hy;
        yh = in.y0 + T(cy + 1) * hy;
    } else {  // a grid cell: lattice nodes x0 + i d
        const int64_t c = in.idx[r], i = c % in.n0, j = c / in.n0;
        xl = in.x0 + T(i) * in.dx;
        xh = in.x0 + T(i + 1) * in.dx;
        yl = in.y0 + T(j) * in.dy;
        yh = in.y0 + T(j + 1) * in.dy;
    }
    nx[0] = xl; nx[1] = xh; nx[2] = xh; nx[3] = xl;
    ny[0] = yl; ny[1] = yl; ny[2] = yh; ny[3] = yh;
#pragma unroll
    for (int k = 0; k < 4; ++k) ph[k] = in.phi[4 * r + k];
}

// cell_boundary_walk then cell_gap_segments (marching_squares.jl:240-317); segment q runs from
// (bx[q], by[q]) to (ax[q], ay[q]) -- cell_interface's swapped ends, the order _push_segments!
// numbers them in.  Returns the segment count (at most 2).
template <typename T>
__device__ __forceinline__ int cell_segments(const T (&nx)[4], const T (&ny)[4], const T (&ph)[4], T (&bx)[2],
                                             T (&by)[2], T (&ax)[2], T (&ay)[2]) {
    T sax[4], say[4], sbx[4], sby[4];
    bool pres[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int k2 = (k + 1) & 3;
        const T fa = ph[k], fb = ph[k2];
        const bool oa = fa >= T(0), ob = fb >= T(0);
        const T fr = (oa && ob) ? T(1) : ((!oa && !ob) ? T(0) : (oa ? fa / (fa - fb) : fb / (fb - fa)));
        pres[k] = oa || ob;
        if (oa && ob) {
            sax[k] = nx[k]; say[k] = ny[k]; sbx[k] = nx[k2]; sby[k] = ny[k2];
        } else if (oa) {
            sax[k] = nx[k]; say[k] = ny[k];
            sbx[k] = nx[k] + fr * (nx[k2] - nx[k]);
            sby[k] = ny[k] + fr * (ny[k2] - ny[k]);
        } else {
            sax[k] = nx[k2] + fr * (nx[k] - nx[k2]);
            say[k] = ny[k2] + fr * (ny[k] - ny[k2]);
            sbx[k] = nx[k2]; sby[k] = ny[k2];
        }
    }
    int nseg = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (nseg == 2 || !pres[k]) continue;
        int jn = k;
        for (int s = 1; s <= 3; ++s)
            if (pres[(k + s) & 3]) { jn = (k + s) & 3; break; }
        const T gx = sbx[k], gy = sby[k], hx = sax[jn], hy = say[jn];
        if (gx != hx || gy != hy) {
            bx[nseg] = hx; by[nseg] = hy;
            ax[nseg] = gx; ay[nseg] = gy;
            nseg += 1;
        }
    }
    return nseg;
}

template <typename T>
__global__ __launch_bounds__(CT) void k_seg_count(ContourCells<T> in, int64_t ncut, int *__restrict__ cnt) {
    const int64_t r = blockIdx.x * (int64_t)CT + threadIdx.x;
    if (r >= ncut) return;
    T nx[4], ny[4], ph[4], bx[2], by[2], ax[2], ay[2];
    entry(in, r, nx, ny, ph);
    cnt[r] = cell_segments(nx, ny, ph, bx, by, ax, ay);
}

template <typename T>
__global__ __launch_bounds__(CT) void k_seg_emit(ContourCells<T> in, int64_t ncut, const int *__restrict__ off,
                                                 T *__restrict__ pts, uint64_t *__restrict__ kx,
                                                 uint64_t *__restrict__ ky, uint32_t *__restrict__ ord) {
    const int64_t r = blockIdx.x * (int64_t)CT + threadIdx.x;
    if (r >= ncut) return;
    T nx[4], ny[4], ph[4], bx[2], by[2], ax[2], ay[2];
    entry(in, r, nx, ny, ph);
    const int ns = cell_segments(nx, ny, ph, bx, by, ax, ay);
    for (int q = 0; q < ns; ++q) {
        const int64_t e0 = 2 * ((int64_t)off[r] + q);
        // the weld's key: x .+ zero(T), so -0 and +0 are one vertex
        const T x0 = bx[q] + T(0), y0 = by[q] + T(0), x1 = ax[q] + T(0), y1 = ay[q] + T(0);
        pts[2 * e0] = x0; pts[2 * e0 + 1] = y0;
        pts[2 * e0 + 2] = x1; pts[2 * e0 + 3] = y1;
        kx[e0] = Bits<T>::of(x0); ky[e0] = Bits<T>::of(y0);
        kx[e0 + 1] = Bits<T>::of(x1); ky[e0 + 1] = Bits<T>::of(y1);
        ord[e0] = (uint32_t)e0;
        ord[e0 + 1] = (uint32_t)(e0 + 1);
    }
}

// kx of each emission in the order of the first (y) sort, for the second (x) sort.
__global__ __launch_bounds__(CT) void k_gather_keys(const uint32_t *__restrict__ ord, const uint64_t *__restrict__ kx,
                                                    int64_t ne, uint64_t *__restrict__ out) {
    const int64_t p = blockIdx.x * (int64_t)CT + threadIdx.x;
    if (p < ne) out[p] = kx[ord[p]];
}

// Group heads in sorted order (head[p]) and, scattered to emission order, first-reach flags (first[e]).
__global__ __launch_bounds__(CT) void k_heads(const uint32_t *__restrict__ ord, const uint64_t *__restrict__ kx,
                                              const uint64_t *__restrict__ ky, int64_t ne, int *__restrict__ head,
                                              int *__restrict__ first) {
    const int64_t p = blockIdx.x * (int64_t)CT + threadIdx.x;
    if (p >= ne) return;
    const uint32_t e = ord[p];
    bool h = p == 0;
    if (!h) {
        const uint32_t q = ord[p - 1];
        h = kx[e] != kx[q] || ky[e] != ky[q];
    }
    head[p] = h ? 1 : 0;
    first[e] = h ? 1 : 0;
}

// The emission that heads each group (grp: inclusive scan of head, 1-based group numbers).
__global__ __launch_bounds__(CT) void k_head_emission(const uint32_t *__restrict__ ord, const int *__restrict__ head,
                                                      const int *__restrict__ grp, int64_t ne,
                                                      uint32_t *__restrict__ head_e) {
    const int64_t p = blockIdx.x * (int64_t)CT + threadIdx.x;
    if (p < ne && head[p]) head_e[grp[p] - 1] = ord[p];
}

// Every emission's vertex number (1-based, first-reach order: vid0 is the exclusive scan of first)
// into lines; each vertex's coordinates, once, into nodes (as Float64).
template <typename T>
__global__ __launch_bounds__(CT) void k_assign(const uint32_t *__restrict__ ord, const int *__restrict__ grp,
                                               const uint32_t *__restrict__ head_e, const int *__restrict__ vid0,
                                               const T *__restrict__ pts, int64_t ne, int64_t *__restrict__ lines,
                                               double *__restrict__ nodes) {
    const int64_t p = blockIdx.x * (int64_t)CT + threadIdx.x;
    if (p >= ne) return;
    const uint32_t e = ord[p], h = head_e[grp[p] - 1];
    const int v = vid0[h];
    lines[e] = (int64_t)v + 1;
    if (e == h) {
        nodes[2 * (int64_t)v] = (double)pts[2 * (int64_t)e];
        nodes[2 * (int64_t)v + 1] = (double)pts[2 * (int64_t)e + 1];
    }
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + CT - 1) / CT); }

}  // namespace

// Bytes of scratch contour_weld needs for ncut cut cells (its arrays and the library scans' and
// sorts' temporary storage).
template <typename T> size_t contour_weld_bytes(int64_t ncut) {
    const int64_t ne = 4 * (ncut > 0 ? ncut : 1);
    size_t sort_b = 0, scan_b = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_b, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)ne, 0, 64);
    size_t s1 = 0, s2 = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, s1, (const int *)nullptr, (int *)nullptr, (int)ne);
    (void)hipcub::DeviceScan::InclusiveSum(nullptr, s2, (const int *)nullptr, (int *)nullptr, (int)ne);
    scan_b = s1 > s2 ? s1 : s2;
    const size_t tmp = (sort_b > scan_b ? sort_b : scan_b) + 256;
    const size_t n = (size_t)ne, m = (size_t)(ncut > 0 ? ncut : 1);
    return 256 * 20 + 2 * m * 4 + 4 * n * 4 + 4 * n * 8 + 4 * n * 4 + 2 * n * sizeof(T) + tmp;
}

template <typename T>
int contour_weld(const ContourCells<T> &in, int64_t ncut, void *scratch, size_t scratch_bytes, int64_t *lines_dev,
                 double *nodes_dev, int64_t *nseg, int64_t *nvert, hipStream_t s) {
    *nseg = *nvert = 0;
    if (ncut <= 0) return 0;
    if (4 * ncut >= (int64_t(1) << 31)) return -1;
    const int64_t cap = 4 * ncut;
    // carve the scratch, 256-byte aligned pieces
    unsigned char *p = (unsigned char *)scratch, *end = p + scratch_bytes;
    auto take = [&](size_t bytes) -> void * {
        p = (unsigned char *)(((uintptr_t)p + 255) & ~(uintptr_t)255);
        void *r = p;
        p += bytes;
        return r;
    };
    int *cnt = (int *)take(4 * ncut), *off = (int *)take(4 * ncut);
    int *head = (int *)take(4 * cap), *first = (int *)take(4 * cap), *grp = (int *)take(4 * cap),
        *vid0 = (int *)take(4 * cap);
    uint64_t *kx = (uint64_t *)take(8 * cap), *ky = (uint64_t *)take(8 * cap), *kxs = (uint64_t *)take(8 * cap),
             *ks = (uint64_t *)take(8 * cap);
    uint32_t *ord = (uint32_t *)take(4 * cap), *ord1 = (uint32_t *)take(4 * cap), *ord2 = (uint32_t *)take(4 * cap),
             *head_e = (uint32_t *)take(4 * cap);
    T *pts = (T *)take(2 * sizeof(T) * cap);
    void *tmp = take(0);
    if (p > end) return -1;
    size_t tmp_bytes = (size_t)(end - (unsigned char *)tmp);

    hipLaunchKernelGGL((k_seg_count<T>), dim3(blocks(ncut)), dim3(CT), 0, s, in, ncut, cnt);
    size_t tb = tmp_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, off, (int)ncut, s) != hipSuccess) return -2;
    int last[2] = {0, 0};
    if (hipMemcpyAsync(&last[0], off + ncut - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&last[1], cnt + ncut - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return -2;
    const int64_t S = (int64_t)last[0] + last[1], ne = 2 * S;
    *nseg = S;
    if (S == 0) return 0;
    hipLaunchKernelGGL((k_seg_emit<T>), dim3(blocks(ncut)), dim3(CT), 0, s, in, ncut, off, pts, kx, ky, ord);
    const int kbits = 8 * (int)sizeof(T);
    tb = tmp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(tmp, tb, ky, ks, ord, ord1, (int)ne, 0, kbits, s) != hipSuccess)
        return -2;
    hipLaunchKernelGGL(k_gather_keys, dim3(blocks(ne)), dim3(CT), 0, s, ord1, kx, ne, kxs);
    tb = tmp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(tmp, tb, kxs, ks, ord1, ord2, (int)ne, 0, kbits, s) != hipSuccess)
        return -2;
    hipLaunchKernelGGL(k_heads, dim3(blocks(ne)), dim3(CT), 0, s, ord2, kx, ky, ne, head, first);
    tb = tmp_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(tmp, tb, first, vid0, (int)ne, s) != hipSuccess) return -2;
    tb = tmp_bytes;
    if (hipcub::DeviceScan::InclusiveSum(tmp, tb, head, grp, (int)ne, s) != hipSuccess) return -2;
    hipLaunchKernelGGL(k_head_emission, dim3(blocks(ne)), dim3(CT), 0, s, ord2, head, grp, ne, head_e);
    hipLaunchKernelGGL((k_assign<T>), dim3(blocks(ne)), dim3(CT), 0, s, ord2, grp, head_e, vid0, pts, ne, lines_dev,
                       nodes_dev);
    int nv = 0;
    if (hipMemcpyAsync(&nv, grp + ne - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return -2;
    *nvert = nv;
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

template size_t contour_weld_bytes<double>(int64_t);
template size_t contour_weld_bytes<float>(int64_t);
template int contour_weld<double>(const ContourCells<double> &, int64_t, void *, size_t, int64_t *, double *,
                                  int64_t *, int64_t *, hipStream_t);
template int contour_weld<float>(const ContourCells<float> &, int64_t, void *, size_t, int64_t *, double *,
                                 int64_t *, int64_t *, hipStream_t);

}  // namespace cc
