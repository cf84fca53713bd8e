// The whole-grid marching-cubes surface on the device: generate_mesh(cache, grid) for
// MarchingCubesCutCell (src/marching_cubes/surface.jl:82-112) and cell_indices (:128-141).
//
// The reference copies phi to the host and runs MarchingCubes.jl's serial march over every node
// and cube (a 3 (N+1)^3 vertex-index table, about half a gigabyte at N = 200).  Here the march
// starts from the cut-cell list the update already built: every edge whose ends differ in sign is
// an edge of a cut cell, so the work follows the surface, not the grid.  Ranks in two bitmaps turn
// that unordered list into the march's order without a sort:
//
//   k_mesh_mark   one thread per listed cut cell: sets the bits of its crossing edges in a
//                 per-node-row edge bitmap (row j + (n1+1) k, bit 3 i + axis), each edge by the one
//                 cell whose corner 0 is the edge's low node (or, on the grid's far faces where that
//                 cell does not exist, by every cell that has the edge), and its own bit in a
//                 per-cell-row cut bitmap.
//   k_scan_*      exclusive prefix sums (reduce, scan of block sums, down-sweep), two at a time:
//                 first of the two bitmaps' word popcounts, giving every word the rank of its first
//                 bit, later of the ordered cells' triangle and interior-vertex counts.
//   k_mesh_order  one thread per listed cut cell: its rank in cube order (its word's rank + the
//                 popcount below its bit) -- a counting sort -- and, at that rank, its triangle
//                 count and whether its tiling uses the interior vertex.
//   k_mesh_tris   one thread per ordered cut cell: the Lewiner dispatch, the vertices of the edges
//                 it owns (as k_mesh_mark assigns them), its interior vertex and its triangles, a
//                 vertex number being the rank of its edge's bit.
//
// Numbering is that of Lewiner's reference implementation: edge vertices by node (column-major)
// and then axis x, y, z; interior vertices after them in cube order; triangles in cube order, each
// cube's in tiling order, 1-based.  Every rank and offset is an exact integer sum, so the mesh is the
// same on every run.  Crossings are anchored on the edge's low node, as the march anchors them
// (surface.jl:4-7): x_low + u (x_high - x_low), u = phi_low / (phi_low - phi_high), values nudged as
// the cells nudge them; an interior vertex is the mean of its cube's edge vertices, edge order 0..11.
#include "cc_mc_dev.h"

namespace cc {
namespace {

typedef unsigned long long u64;

// Edge e of a cube: its low node's offset (dx, dy, dz) and its axis, packed 5 bits per edge
// (dx | dy << 1 | dz << 2 | axis << 3), from the corner table edge_node and MC_NODE_BITS.
constexpr u64 pack_edges() {
    u64 v = 0;
    for (int e = 0; e < 12; ++e) {
        const int lo = e == 0 ? 0 : e == 1 ? 1 : e == 2 ? 3 : e == 3 ? 0 : e == 4 ? 4 : e == 5 ? 5
                       : e == 6 ? 7 : e == 7 ? 4 : e == 8 ? 0 : e == 9 ? 1 : e == 10 ? 2 : 3;
        const int dx = (lo ^ (lo >> 1)) & 1, dy = (lo >> 1) & 1, dz = (lo >> 2) & 1;
        const int ax = e < 8 ? (e & 1) : 2;
        v |= (u64)(dx | dy << 1 | dz << 2 | ax << 3) << (5 * e);
    }
    return v;
}
constexpr u64 EDGE_PACK = pack_edges();
__device__ __forceinline__ int edge_bits(int e) { return (int)((EDGE_PACK >> (5 * e)) & 31u); }

struct MeshDims {
    uint32_t n0, n1, n2;  // cells per axis
    uint32_t N0, N1;      // nodes per axis (x, y)
    uint32_t W;           // 64-bit words per node row of the edge bitmap
    uint32_t Wc;          // 64-bit words per cell row of the cut bitmap
};

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ u64 wave_incl_scan64(u64 v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u64 t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

template <typename T> __device__ __forceinline__ T lat(const GridT<T> &g, int a, uint32_t i) {
    return g.x0[a] + T(i) * g.d[a];  // get_node, bitwise the cells' own corners
}

// The vertex on the axis-a edge from node (i, j, k) to its +a neighbour, values fl, fh (nudged).
template <typename T>
__device__ __forceinline__ void edge_vertex(const GridT<T> &g, uint32_t i, uint32_t j, uint32_t k, int a,
                                            T fl, T fh, T *p) {
    const T u = fl / (fl - fh);
    const uint32_t ijk[3] = {i, j, k};
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const T xl = lat(g, b, ijk[b]);
        if (b == a) {
            const T xh = lat(g, b, ijk[b] + 1);
            p[b] = xl + u * (xh - xl);
        } else {
            p[b] = xl;
        }
    }
}

// ---- 1. mark ------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_mesh_mark(const T *__restrict__ phi, GridT<T> g, MeshDims md,
                                                   const uint32_t *__restrict__ cutlist, uint32_t ncut,
                                                   u64 *__restrict__ ebits, u64 *__restrict__ cbits) {
    const size_t sy = md.N0, sz = (size_t)md.N0 * md.N1;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ncut; t += gridDim.x * blockDim.x) {
        const uint32_t c = cutlist[t];
        const uint32_t i = c % md.n0, r = c / md.n0, j = r % md.n1, k = r / md.n1;
        T ph[8];
        (void)load_corners(phi, i + sy * j + sz * k, sy, sz, ph);
#pragma unroll
        for (int e = 0; e < 12; ++e) {
            const bool crossed = (ph[edge_node(e, 0)] >= T(0)) != (ph[edge_node(e, 1)] >= T(0));
            const int eb = edge_bits(e);
            const uint32_t ni = i + (eb & 1), nj = j + ((eb >> 1) & 1), nk = k + ((eb >> 2) & 1);
            const bool own = (eb & 7) == 0;
            const bool owner_exists = ni < md.n0 && nj < md.n1 && nk < md.n2;
            if (crossed && (own || !owner_exists)) {
                const uint32_t bit = 3 * ni + (eb >> 3);
                const size_t row = nj + (size_t)md.N1 * nk;
                atomicOr(&ebits[row * md.W + (bit >> 6)], 1ull << (bit & 63));
            }
        }
        const size_t rc = j + (size_t)md.n1 * k;
        atomicOr(&cbits[rc * md.Wc + (i >> 6)], 1ull << (i & 63));
    }
}

// ---- 2. exclusive prefix sums of counts or of bitmap-word popcounts, two at a time --------------------
constexpr int SCAN_THREADS = 256, SCAN_ITEMS = 16, SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ u64 block_excl_scan(u64 v, u64 *total) {
    __shared__ u64 wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u64 inc = wave_incl_scan64(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    u64 before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// A segment's value q: a count, or the popcount of a bitmap word (the bitmap's rank).
__device__ __forceinline__ uint32_t seg_value(const MeshScan::Seg &sg, size_t q) {
    return sg.bits ? (uint32_t)__popcll(sg.bits[q]) : sg.in[q];
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(MeshScan s) {
    const MeshScan::Seg &sg = s.seg[blockIdx.y];
    const size_t b0 = (size_t)blockIdx.x * SCAN_TILE;
    if (b0 >= sg.n) return;
    u64 v = 0;
    for (size_t q = b0 + threadIdx.x; q < b0 + SCAN_TILE && q < sg.n; q += SCAN_THREADS) v += seg_value(sg, q);
    u64 total;
    block_excl_scan(v, &total);
    if (threadIdx.x == 0) s.part[blockIdx.y * s.maxblocks + blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_parts(MeshScan s) {
    const MeshScan::Seg &sg = s.seg[blockIdx.y];
    const size_t nb = (sg.n + SCAN_TILE - 1) / SCAN_TILE;
    u64 *part = s.part + blockIdx.y * s.maxblocks;
    u64 carry = 0;
    for (size_t b0 = 0; b0 < nb; b0 += SCAN_THREADS) {
        const size_t b = b0 + threadIdx.x;
        const u64 v = b < nb ? part[b] : 0;
        u64 total;
        const u64 ex = block_excl_scan(v, &total);
        if (b < nb) part[b] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) s.totals[blockIdx.y] = (int64_t)carry;
}

// The tile moves through LDS so that global loads and stores are lane-contiguous; each thread then
// scans 16 consecutive values (one pad slot per 16 keeps those LDS reads free of bank conflicts).
__device__ __forceinline__ int pad16(int x) { return x + (x >> 4); }

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(MeshScan s) {
    __shared__ uint32_t tile[SCAN_TILE + SCAN_TILE / 16];
    const MeshScan::Seg &sg = s.seg[blockIdx.y];
    const size_t b0 = (size_t)blockIdx.x * SCAN_TILE;
    if (b0 >= sg.n) return;
    const int t = threadIdx.x;
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; ++u) {
        const int x = u * SCAN_THREADS + t;
        tile[pad16(x)] = b0 + x < sg.n ? seg_value(sg, b0 + x) : 0u;
    }
    __syncthreads();
    uint32_t v[SCAN_ITEMS];
    u64 sum = 0;
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; ++u) {
        v[u] = tile[pad16(t * SCAN_ITEMS + u)];
        sum += v[u];
    }
    u64 total;
    const u64 base = s.part[blockIdx.y * s.maxblocks + blockIdx.x];
    // offsets within the tile fit 32 bits: at most 4096 values of at most 64 (a word's popcount)
    uint32_t run = (uint32_t)block_excl_scan(sum, &total);
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; ++u) {
        tile[pad16(t * SCAN_ITEMS + u)] = run;
        run += v[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; ++u) {
        const int x = u * SCAN_THREADS + t;
        if (b0 + x < sg.n) {
            const u64 r = base + tile[pad16(x)];
            if (sg.out32) sg.out32[b0 + x] = (uint32_t)r;
            else sg.out[b0 + x] = (int64_t)r;
        }
    }
}

// ---- 3. the cut cells in cube order ---------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_mesh_order(const T *__restrict__ phi, GridT<T> g, MeshDims md,
                                                    const uint32_t *__restrict__ cutlist, uint32_t ncut,
                                                    const u64 *__restrict__ cbits, const uint32_t *__restrict__ crank,
                                                    uint32_t *__restrict__ order,
                                                    uint32_t *__restrict__ tn, uint32_t *__restrict__ cn) {
    const size_t sy = md.N0, sz = (size_t)md.N0 * md.N1;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ncut; t += gridDim.x * blockDim.x) {
        const uint32_t c = cutlist[t];
        const uint32_t i = c % md.n0, r = c / md.n0, j = r % md.n1, k = r / md.n1;
        const size_t w = (j + (size_t)md.n1 * k) * md.Wc + (i >> 6);
        const uint32_t q = crank[w] + (uint32_t)__popcll(cbits[w] & ((1ull << (i & 63)) - 1ull));
        T ph[8];
        const int lut = load_corners(phi, i + sy * j + sz * k, sy, sz, ph);
        const FaceTests ft = face_tests(ph);
        const signed char *tile;
        int nt = cell_triangles(ph, ft, lut, &tile);
        if (nt < 0) nt = 0;
        bool inner = false;
        for (int x = 0; x < 3 * nt; ++x) inner |= tile[x] == 12;
        order[q] = c;
        tn[q] = (uint32_t)nt;
        cn[q] = inner ? 1u : 0u;
    }
}

// ---- 4. vertices and triangles ---------------------------------------------------------
// The vertex number (0-based) of the crossing on edge e of cell (i, j, k): its bit's rank.
__device__ __forceinline__ int64_t edge_id(const MeshDims &md, const u64 *__restrict__ ebits,
                                           const uint32_t *__restrict__ erank, uint32_t i, uint32_t j, uint32_t k,
                                           int e) {
    const int eb = edge_bits(e);
    const uint32_t ni = i + (eb & 1), nj = j + ((eb >> 1) & 1), nk = k + ((eb >> 2) & 1);
    const uint32_t bit = 3 * ni + (eb >> 3);
    const size_t w = (nj + (size_t)md.N1 * nk) * md.W + (bit >> 6);
    const u64 below = ebits[w] & ((1ull << (bit & 63)) - 1ull);
    return (int64_t)erank[w] + __popcll(below);
}

template <typename T>
__global__ __launch_bounds__(256) void k_mesh_tris(const T *__restrict__ phi, GridT<T> g, MeshDims md,
                                                   uint32_t ncut, const uint32_t *__restrict__ order,
                                                   const int64_t *__restrict__ tpre, const int64_t *__restrict__ cpre,
                                                   const int64_t *__restrict__ totals,
                                                   const u64 *__restrict__ ebits, const uint32_t *__restrict__ erank,
                                                   T *__restrict__ verts, int64_t *__restrict__ tris) {
    __shared__ int64_t ids[13 * 256];  // per thread: the vertex numbers (1-based) of edges 0..11 and 12
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncut) return;
    const uint32_t c = order[q];
    const uint32_t i = c % md.n0, r = c / md.n0, j = r % md.n1, k = r / md.n1;
    const size_t sy = md.N0, sz = (size_t)md.N0 * md.N1;
    T ph[8];
    const int lut = load_corners(phi, i + sy * j + sz * k, sy, sz, ph);
    const FaceTests ft = face_tests(ph);
    const signed char *tile;
    int nt = cell_triangles(ph, ft, lut, &tile);
    if (nt < 0) nt = 0;
    bool inner = false;
    for (int x = 0; x < 3 * nt; ++x) inner |= tile[x] == 12;
    int64_t *my = ids + threadIdx.x;
    // Every crossing's vertex number (independent loads; the tiling reads them below) and, for the
    // edges this cell owns as k_mesh_mark does, the vertex itself; with an interior vertex, the mean
    // of all the cube's edge vertices in edge order.  An edge on a far face of the grid is owned by
    // every cell that has it, which all write the same bits.
    T acc[3] = {T(0), T(0), T(0)};
    int cnt = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const T fl = ph[edge_node(e, 0)], fh = ph[edge_node(e, 1)];
        if ((fl >= T(0)) == (fh >= T(0))) continue;
        const int eb = edge_bits(e);
        const uint32_t ni = i + (eb & 1), nj = j + ((eb >> 1) & 1), nk = k + ((eb >> 2) & 1);
        const int64_t id = edge_id(md, ebits, erank, i, j, k, e);
        my[256 * e] = id + 1;
        const bool own = (eb & 7) == 0 || !(ni < md.n0 && nj < md.n1 && nk < md.n2);
        if (own || inner) {
            T x[3];
            edge_vertex(g, ni, nj, nk, eb >> 3, fl, fh, x);
            if (own) {
                verts[3 * id] = x[0];
                verts[3 * id + 1] = x[1];
                verts[3 * id + 2] = x[2];
            }
            acc[0] += x[0];
            acc[1] += x[1];
            acc[2] += x[2];
            ++cnt;
        }
    }
    if (inner) {
        const int64_t v12 = totals[0] + cpre[q];  // interior vertices follow every edge vertex
        verts[3 * v12] = acc[0] / T(cnt);
        verts[3 * v12 + 1] = acc[1] / T(cnt);
        verts[3 * v12 + 2] = acc[2] / T(cnt);
        my[256 * 12] = v12 + 1;
    }
    int64_t *dst = tris + 3 * tpre[q];
    for (int t = 0; t < nt; ++t) {
        dst[3 * t] = my[256 * tile[3 * t]];
        dst[3 * t + 1] = my[256 * tile[3 * t + 1]];
        dst[3 * t + 2] = my[256 * tile[3 * t + 2]];
    }
}

// ---- cell_indices (surface.jl:128-141) ------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_cell_indices(const T *__restrict__ verts, const int64_t *__restrict__ tris,
                                                      int64_t nt, MeshGrid g, int64_t *__restrict__ cells) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int64_t a = tris[3 * t] - 1, b = tris[3 * t + 1] - 1, c = tris[3 * t + 2] - 1;
    int64_t idx[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const T cen = ((verts[3 * a + x] + verts[3 * b + x]) + verts[3 * c + x]) / T(3);
        const double f = floor(((double)cen - g.x0[x]) / g.d[x]) + 1.0;
        idx[x] = (int64_t)fmin(fmax(f, 1.0), (double)g.n[x]);  // clamp(floor(...) + 1, 1, n)
    }
    cells[t] = idx[0] + g.n[0] * ((idx[1] - 1) + g.n[1] * (idx[2] - 1));
}

template <typename U> __global__ void k_zero(U *p, size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] = 0;
}

MeshDims dims_of(const int *n) {
    MeshDims md;
    md.n0 = n[0];
    md.n1 = n[1];
    md.n2 = n[2];
    md.N0 = n[0] + 1;
    md.N1 = n[1] + 1;
    md.W = (3 * md.N0 + 63) / 64;
    md.Wc = (md.n0 + 63) / 64;
    return md;
}

}  // namespace

MeshSizes mesh_sizes(const int *n) {
    const MeshDims md = dims_of(n);
    MeshSizes s;
    s.node_rows = (size_t)md.N1 * (n[2] + 1);
    s.cell_rows = (size_t)md.n1 * md.n2;
    s.ewords = s.node_rows * md.W;
    s.cwords = s.cell_rows * md.Wc;
    return s;
}

size_t mesh_scan_blocks(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void launch_scan_pair(MeshScan::Seg s0, MeshScan::Seg s1, unsigned long long *part, size_t part_cap,
                      int64_t *totals, hipStream_t s) {
    MeshScan sc;
    sc.seg[0] = s0;
    sc.seg[1] = s1;
    sc.seg[2] = {nullptr, nullptr, nullptr, nullptr, 0};
    sc.part = part;
    sc.maxblocks = part_cap;
    sc.totals = totals;
    const size_t nb = mesh_scan_blocks(s0.n > s1.n ? s0.n : s1.n);
    if (nb == 0) {
        hipLaunchKernelGGL(k_zero<int64_t>, dim3(1), dim3(64), 0, s, totals, (size_t)2);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb, 2), dim3(SCAN_THREADS), 0, s, sc);
    hipLaunchKernelGGL(k_scan_parts, dim3(1, 2), dim3(SCAN_THREADS), 0, s, sc);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb, 2), dim3(SCAN_THREADS), 0, s, sc);
}

namespace {
void scan2(const MeshBuffers &b, MeshScan::Seg s0, MeshScan::Seg s1, int64_t *totals, hipStream_t s) {
    launch_scan_pair(s0, s1, b.part, b.part_cap, totals, s);
}

unsigned grid_for(uint32_t n) {
    const unsigned b = (n + 255) / 256;
    return b < 1 ? 1 : (b > 4096 ? 4096 : b);
}
}  // namespace

template <typename T>
void launch_mesh_count(const T *phi, const GridT<T> &g, const uint32_t *cutlist, uint32_t ncut,
                       const MeshBuffers &b, hipStream_t s) {
    const MeshDims md = dims_of(g.n);
    const MeshSizes sz = mesh_sizes(g.n);
    // cleared by kernels, not hipMemsetAsync, so that a captured graph replays correctly
    hipLaunchKernelGGL(k_zero<u64>, dim3(1024), dim3(256), 0, s, (u64 *)b.ebits, sz.ewords);
    hipLaunchKernelGGL(k_zero<u64>, dim3(1024), dim3(256), 0, s, (u64 *)b.cbits, sz.cwords);
    hipLaunchKernelGGL(k_mesh_mark<T>, dim3(grid_for(ncut)), dim3(256), 0, s, phi, g, md, cutlist, ncut,
                       (u64 *)b.ebits, (u64 *)b.cbits);
    // ranks of both bitmaps: totals[0] edge vertices, totals[1] cut cells
    scan2(b, {nullptr, (const u64 *)b.ebits, nullptr, b.erank, sz.ewords},
          {nullptr, (const u64 *)b.cbits, nullptr, b.crank, sz.cwords}, b.totals, s);
    hipLaunchKernelGGL(k_mesh_order<T>, dim3(grid_for(ncut)), dim3(256), 0, s, phi, g, md, cutlist, ncut,
                       (const u64 *)b.cbits, (const uint32_t *)b.crank, b.order, b.tn, b.cn);
    // triangles and interior vertices in cube order: totals[2], totals[3]
    scan2(b, {b.tn, nullptr, b.tpre, nullptr, ncut}, {b.cn, nullptr, b.cpre, nullptr, ncut}, b.totals + 2, s);
}

template <typename T>
void launch_mesh_emit(const T *phi, const GridT<T> &g, uint32_t ncut, const MeshBuffers &b, T *verts,
                      int64_t *tris, hipStream_t s) {
    const MeshDims md = dims_of(g.n);
    if (ncut == 0) return;
    hipLaunchKernelGGL(k_mesh_tris<T>, dim3((ncut + 255) / 256), dim3(256), 0, s, phi, g, md, ncut,
                       (const uint32_t *)b.order, (const int64_t *)b.tpre, (const int64_t *)b.cpre,
                       (const int64_t *)b.totals, (const u64 *)b.ebits, (const uint32_t *)b.erank, verts, tris);
}

template <typename T>
void launch_cell_indices(const T *verts, const int64_t *tris, int64_t nt, const MeshGrid &g, int64_t *cells,
                         hipStream_t s) {
    if (nt <= 0) return;
    hipLaunchKernelGGL(k_cell_indices<T>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, verts, tris, nt,
                       g, cells);
}

template void launch_mesh_count<double>(const double *, const GridT<double> &, const uint32_t *, uint32_t,
                                        const MeshBuffers &, hipStream_t);
template void launch_mesh_count<float>(const float *, const GridT<float> &, const uint32_t *, uint32_t,
                                       const MeshBuffers &, hipStream_t);
template void launch_mesh_emit<double>(const double *, const GridT<double> &, uint32_t, const MeshBuffers &,
                                       double *, int64_t *, hipStream_t);
template void launch_mesh_emit<float>(const float *, const GridT<float> &, uint32_t, const MeshBuffers &, float *,
                                      int64_t *, hipStream_t);
template void launch_cell_indices<double>(const double *, const int64_t *, int64_t, const MeshGrid &, int64_t *,
                                          hipStream_t);
template void launch_cell_indices<float>(const float *, const int64_t *, int64_t, const MeshGrid &, int64_t *,
                                         hipStream_t);

}  // namespace cc
