// The whole-grid marching-cubes surface on the device: generate_mesh(cache, grid) for
// MarchingCubesCutCell (src/marching_cubes/surface.jl:82-112) and cell_indices (:128-141).
//
// The reference copies phi to the host and runs MarchingCubes.jl's serial march over every node
// and cube (a 3 (N+1)^3 vertex-index table, about half a gigabyte at N = 200).  Here the march
// starts from the cut-cell list the update already built: every edge whose ends differ in sign is
// an edge of a cut cell, so the work follows the surface, not the grid.
//
//   k_mesh_mark   one thread per cut cell: sets the bits of its crossing edges in a per-node-row
//                 edge bitmap (row j + (n1+1) k, bit 3 i + axis), each edge by the one cell whose
//                 corner 0 is the edge's low node (or, on the grid's far faces where that cell does
//                 not exist, by every cell that has the edge); sets its own bit in a per-cell-row
//                 bitmap; adds its triangle count and interior-vertex flag to its row's counters.
//   k_mesh_rows   per node row (one wave): word popcounts -> in-row word offsets and row total.
//   k_scan_*      exclusive prefix sums over the rows (reduce, scan of block sums, down-sweep).
//   k_mesh_verts  per node row (one wave): each crossing edge's vertex at its rank.
//   k_mesh_tris   per cell row (one wave, one lane per cell of a 64-cell bitmap word): the Lewiner
//                 dispatch, the interior vertex and the triangles; a vertex number is the rank of
//                 its edge's bit (row offset + word offset + popcount below the bit).
//
// Numbering is that of Lewiner's reference implementation: edge vertices by node (column-major)
// and then axis x, y, z; interior vertices after them in cube order; triangles in cube order, each
// cube's in tiling order, 1-based.  The counters are integer sums, so the mesh is the same on every
// run.  Crossings are anchored on the edge's low node, as the march anchors them (surface.jl:4-7):
// x_low + u (x_high - x_low), u = phi_low / (phi_low - phi_high), values nudged as the cells nudge
// them; an interior vertex is the mean of its cube's edge vertices in edge order 0..11.
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
                                                   const uint32_t *__restrict__ cutlist,
                                                   const uint32_t *__restrict__ counter,
                                                   u64 *__restrict__ ebits, u64 *__restrict__ cbits,
                                                   uint32_t *__restrict__ tcnt, uint32_t *__restrict__ ccnt) {
    const uint32_t m = *counter;
    const size_t sy = md.N0, sz = (size_t)md.N0 * md.N1;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x) {
        const uint32_t c = cutlist[t];
        const uint32_t i = c % md.n0, r = c / md.n0, j = r % md.n1, k = r / md.n1;
        T ph[8];
        const int lut = load_corners(phi, i + sy * j + sz * k, sy, sz, ph);
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
        const FaceTests ft = face_tests(ph);
        const signed char *tile;
        int nt = cell_triangles(ph, ft, lut, &tile);
        if (nt < 0) nt = 0;
        bool inner = false;
        for (int q = 0; q < 3 * nt; ++q) inner |= tile[q] == 12;
        if (nt) atomicAdd(&tcnt[rc], (uint32_t)nt);
        if (inner) atomicAdd(&ccnt[rc], 1u);
    }
}

// ---- 2. per-row word offsets --------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mesh_rows(const u64 *__restrict__ ebits, MeshDims md, size_t nrows,
                                                   uint32_t *__restrict__ wpre, uint32_t *__restrict__ ecnt) {
    const size_t r = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= nrows) return;
    const int lane = threadIdx.x & 63;
    uint32_t carry = 0;
    for (uint32_t w0 = 0; w0 < md.W; w0 += 64) {
        const uint32_t w = w0 + lane;
        const uint32_t c = w < md.W ? (uint32_t)__popcll(ebits[r * md.W + w]) : 0u;
        const uint32_t inc = wave_incl_scan(c);
        if (w < md.W) wpre[r * md.W + w] = carry + inc - c;
        carry += __shfl(inc, 63, 64);
    }
    if (lane == 0) ecnt[r] = carry;
}

// ---- 3. exclusive prefix sums of up to three uint32 arrays ------------------------------------------
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

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(MeshScan s) {
    const MeshScan::Seg &sg = s.seg[blockIdx.y];
    const size_t b0 = (size_t)blockIdx.x * SCAN_TILE;
    if (b0 >= sg.n) return;
    u64 v = 0;
    for (size_t q = b0 + threadIdx.x; q < b0 + SCAN_TILE && q < sg.n; q += SCAN_THREADS) v += sg.in[q];
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

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(MeshScan s) {
    const MeshScan::Seg &sg = s.seg[blockIdx.y];
    const size_t b0 = (size_t)blockIdx.x * SCAN_TILE;
    if (b0 >= sg.n) return;
    const size_t q0 = b0 + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    u64 sum = 0;
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; ++u) {
        v[u] = q0 + u < sg.n ? sg.in[q0 + u] : 0u;
        sum += v[u];
    }
    u64 total;
    u64 run = block_excl_scan(sum, &total) + s.part[blockIdx.y * s.maxblocks + blockIdx.x];
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; ++u) {
        if (q0 + u < sg.n) sg.out[q0 + u] = (int64_t)run;
        run += v[u];
    }
}

// ---- 4. edge vertices ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_mesh_verts(const T *__restrict__ phi, GridT<T> g, MeshDims md,
                                                    size_t nrows, const u64 *__restrict__ ebits,
                                                    const uint32_t *__restrict__ wpre,
                                                    const uint32_t *__restrict__ ecnt,
                                                    const int64_t *__restrict__ epre, T *__restrict__ verts) {
    const size_t r = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= nrows || ecnt[r] == 0) return;
    const int lane = threadIdx.x & 63;
    const uint32_t j = (uint32_t)(r % md.N1), k = (uint32_t)(r / md.N1);
    const size_t sy = md.N0, sz = (size_t)md.N0 * md.N1;
    const int64_t rowbase = epre[r];
    for (uint32_t w = lane; w < md.W; w += 64) {
        u64 word = ebits[r * md.W + w];
        int64_t id = rowbase + wpre[r * md.W + w];
        while (word) {
            const uint32_t bit = 64 * w + (uint32_t)__builtin_ctzll(word);
            word &= word - 1;
            const uint32_t i = bit / 3;
            const int a = (int)(bit - 3 * i);
            const size_t p = i + sy * j + sz * k;
            const T fl = nudge(phi[p]), fh = nudge(phi[p + (a == 0 ? 1 : (a == 1 ? sy : sz))]);
            T x[3];
            edge_vertex(g, i, j, k, a, fl, fh, x);
            verts[3 * id] = x[0];
            verts[3 * id + 1] = x[1];
            verts[3 * id + 2] = x[2];
            ++id;
        }
    }
}

// ---- 5. interior vertices and triangles ---------------------------------------------------------
// The vertex number (0-based) of the crossing on edge e of cell (i, j, k): its bit's rank.
__device__ __forceinline__ int64_t edge_id(const MeshDims &md, const u64 *__restrict__ ebits,
                                           const uint32_t *__restrict__ wpre, const int64_t *__restrict__ epre,
                                           uint32_t i, uint32_t j, uint32_t k, int e) {
    const int eb = edge_bits(e);
    const uint32_t ni = i + (eb & 1), nj = j + ((eb >> 1) & 1), nk = k + ((eb >> 2) & 1);
    const uint32_t bit = 3 * ni + (eb >> 3);
    const size_t row = nj + (size_t)md.N1 * nk;
    const size_t w = row * md.W + (bit >> 6);
    const u64 below = ebits[w] & ((1ull << (bit & 63)) - 1ull);
    return epre[row] + wpre[w] + __popcll(below);
}

template <typename T>
__global__ __launch_bounds__(256) void k_mesh_tris(const T *__restrict__ phi, GridT<T> g, MeshDims md,
                                                   size_t ncrows, const u64 *__restrict__ cbits,
                                                   const uint32_t *__restrict__ tcnt,
                                                   const uint32_t *__restrict__ ccnt,
                                                   const int64_t *__restrict__ tpre,
                                                   const int64_t *__restrict__ cpre, const int64_t *__restrict__ totals,
                                                   const u64 *__restrict__ ebits, const uint32_t *__restrict__ wpre,
                                                   const int64_t *__restrict__ epre, T *__restrict__ verts,
                                                   int64_t *__restrict__ tris) {
    const size_t rc = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (rc >= ncrows || (tcnt[rc] == 0 && ccnt[rc] == 0)) return;
    const int lane = threadIdx.x & 63;
    const uint32_t j = (uint32_t)(rc % md.n1), k = (uint32_t)(rc / md.n1);
    const size_t sy = md.N0, sz = (size_t)md.N0 * md.N1;
    int64_t tnext = tpre[rc];
    int64_t cnext = totals[0] + cpre[rc];  // interior vertices follow every edge vertex
    for (uint32_t w = 0; w < md.Wc; ++w) {
        const u64 word = cbits[rc * md.Wc + w];
        if (word == 0) continue;
        const uint32_t i = 64 * w + lane;
        const bool cut = (word >> lane) & 1ull;
        T ph[8];
        int nt = 0, lut = 0;
        bool inner = false;
        const signed char *tile = nullptr;
        if (cut) {
            lut = load_corners(phi, i + sy * j + sz * k, sy, sz, ph);
            const FaceTests ft = face_tests(ph);
            nt = cell_triangles(ph, ft, lut, &tile);
            if (nt < 0) nt = 0;
            for (int q = 0; q < 3 * nt; ++q) inner |= tile[q] == 12;
        }
        const uint32_t tinc = wave_incl_scan((uint32_t)nt);
        const uint32_t cinc = wave_incl_scan(inner ? 1u : 0u);
        const int64_t t0 = tnext + tinc - (uint32_t)nt;
        const int64_t v12 = cnext + cinc - (inner ? 1u : 0u);
        tnext += __shfl(tinc, 63, 64);
        cnext += __shfl(cinc, 63, 64);
        if (!cut) continue;
        if (inner) {
            // the mean of the cube's edge vertices, in edge order (each computed as k_mesh_verts does)
            T acc[3] = {T(0), T(0), T(0)};
            int cnt = 0;
#pragma unroll
            for (int e = 0; e < 12; ++e) {
                const T fl = ph[edge_node(e, 0)], fh = ph[edge_node(e, 1)];
                if ((fl >= T(0)) == (fh >= T(0))) continue;
                const int eb = edge_bits(e);
                T x[3];
                edge_vertex(g, i + (eb & 1), j + ((eb >> 1) & 1), k + ((eb >> 2) & 1), eb >> 3, fl, fh, x);
                acc[0] += x[0];
                acc[1] += x[1];
                acc[2] += x[2];
                ++cnt;
            }
            verts[3 * v12] = acc[0] / T(cnt);
            verts[3 * v12 + 1] = acc[1] / T(cnt);
            verts[3 * v12 + 2] = acc[2] / T(cnt);
        }
        for (int t = 0; t < nt; ++t)
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const int code = tile[3 * t + v];
                const int64_t id = code == 12 ? v12 : edge_id(md, ebits, wpre, epre, i, j, k, code);
                tris[3 * (t0 + t) + v] = id + 1;
            }
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
    const size_t rows = s.node_rows > s.cell_rows ? s.node_rows : s.cell_rows;
    s.scan_blocks = (rows + SCAN_TILE - 1) / SCAN_TILE;
    return s;
}

template <typename T>
void launch_mesh_count(const T *phi, const GridT<T> &g, const uint32_t *cutlist, const uint32_t *counter,
                       const MeshBuffers &b, hipStream_t s) {
    const MeshDims md = dims_of(g.n);
    const MeshSizes sz = mesh_sizes(g.n);
    // cleared by kernels, not hipMemsetAsync, so that a captured graph replays correctly
    hipLaunchKernelGGL(k_zero<u64>, dim3(1024), dim3(256), 0, s, (u64 *)b.ebits, sz.ewords);
    hipLaunchKernelGGL(k_zero<u64>, dim3(1024), dim3(256), 0, s, (u64 *)b.cbits, sz.cwords);
    hipLaunchKernelGGL(k_zero<uint32_t>, dim3(256), dim3(256), 0, s, b.tcnt, sz.cell_rows);
    hipLaunchKernelGGL(k_zero<uint32_t>, dim3(256), dim3(256), 0, s, b.ccnt, sz.cell_rows);
    hipLaunchKernelGGL(k_mesh_mark<T>, dim3(2048), dim3(256), 0, s, phi, g, md, cutlist, counter,
                       (u64 *)b.ebits, (u64 *)b.cbits, b.tcnt, b.ccnt);
    hipLaunchKernelGGL(k_mesh_rows, dim3((unsigned)((sz.node_rows + 3) / 4)), dim3(256), 0, s,
                       (const u64 *)b.ebits, md, sz.node_rows, b.wpre, b.ecnt);
    MeshScan sc;
    sc.seg[0] = {b.ecnt, b.epre, sz.node_rows};
    sc.seg[1] = {b.tcnt, b.tpre, sz.cell_rows};
    sc.seg[2] = {b.ccnt, b.cpre, sz.cell_rows};
    sc.part = b.part;
    sc.maxblocks = sz.scan_blocks;
    sc.totals = b.totals;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)sz.scan_blocks, 3), dim3(SCAN_THREADS), 0, s, sc);
    hipLaunchKernelGGL(k_scan_parts, dim3(1, 3), dim3(SCAN_THREADS), 0, s, sc);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)sz.scan_blocks, 3), dim3(SCAN_THREADS), 0, s, sc);
}

template <typename T>
void launch_mesh_emit(const T *phi, const GridT<T> &g, const MeshBuffers &b, T *verts, int64_t *tris,
                      hipStream_t s) {
    const MeshDims md = dims_of(g.n);
    const MeshSizes sz = mesh_sizes(g.n);
    hipLaunchKernelGGL(k_mesh_verts<T>, dim3((unsigned)((sz.node_rows + 3) / 4)), dim3(256), 0, s, phi, g, md,
                       sz.node_rows, (const u64 *)b.ebits, (const uint32_t *)b.wpre, (const uint32_t *)b.ecnt,
                       (const int64_t *)b.epre, verts);
    hipLaunchKernelGGL(k_mesh_tris<T>, dim3((unsigned)((sz.cell_rows + 3) / 4)), dim3(256), 0, s, phi, g, md,
                       sz.cell_rows, (const u64 *)b.cbits, (const uint32_t *)b.tcnt, (const uint32_t *)b.ccnt,
                       (const int64_t *)b.tpre, (const int64_t *)b.cpre, (const int64_t *)b.totals,
                       (const u64 *)b.ebits, (const uint32_t *)b.wpre, (const int64_t *)b.epre, verts, tris);
}

template <typename T>
void launch_cell_indices(const T *verts, const int64_t *tris, int64_t nt, const MeshGrid &g, int64_t *cells,
                         hipStream_t s) {
    if (nt <= 0) return;
    hipLaunchKernelGGL(k_cell_indices<T>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, verts, tris, nt,
                       g, cells);
}

template void launch_mesh_count<double>(const double *, const GridT<double> &, const uint32_t *,
                                        const uint32_t *, const MeshBuffers &, hipStream_t);
template void launch_mesh_count<float>(const float *, const GridT<float> &, const uint32_t *, const uint32_t *,
                                       const MeshBuffers &, hipStream_t);
template void launch_mesh_emit<double>(const double *, const GridT<double> &, const MeshBuffers &, double *,
                                       int64_t *, hipStream_t);
template void launch_mesh_emit<float>(const float *, const GridT<float> &, const MeshBuffers &, float *,
                                      int64_t *, hipStream_t);
template void launch_cell_indices<double>(const double *, const int64_t *, int64_t, const MeshGrid &, int64_t *,
                                          hipStream_t);
template void launch_cell_indices<float>(const float *, const int64_t *, int64_t, const MeshGrid &, int64_t *,
                                         hipStream_t);

}  // namespace cc
