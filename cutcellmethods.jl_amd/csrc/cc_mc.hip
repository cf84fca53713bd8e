// Marching-cubes cell pass for gfx950 -- the reference's _nodal_cells_kernel! on a 3D grid
// (src/marching_squares/cache.jl:75-78) with cut_cell_moments (src/marching_cubes/moments.jl:275-346).
//
// Two kernels replace the reference's one-thread-per-cell launch:
//
//  * k_mc_stream: a workgroup owns MC_TILE consecutive cells (one contiguous byte range in every
//    column).  Each thread classifies one cell from its 8 nudged corner signs; uncut cells -- about
//    99.8% of a 512^3 sphere grid -- have closed-form moments, so the workgroup then streams every
//    column of the tile with 16-byte stores, lane-contiguous (1 KiB per wave instruction), from
//    LDS.  Cut cells are appended to a work list with one atomic per wave.
//  * k_mc_cut: one thread per listed cut cell runs the full construction: Lewiner dispatch, six face
//    clips, divergence-theorem sums.  Dense work, no divergence from the 0.2% sparsity.
//
// The arithmetic follows the reference expression by expression (left folds, outside-anchored
// interpolation, cell-local coordinates) and is compiled without contraction, so the results are
// the oracle's bit for bit.
#include "cc_mc_dev.h"

namespace cc {
namespace {

// Flat variant: a workgroup owns MC_TILE consecutive cells of the whole grid (any shape), reading
// the 8 corners of each from phi.  Used for narrow grids and fields handed in by the caller.
template <typename T, bool NT>
__global__ __launch_bounds__(MC_TILE) void k_mc_stream(const T *__restrict__ phi, GridT<T> g,
                                                       Cols<T> cols, uint32_t *__restrict__ cutlist,
                                                       uint32_t *__restrict__ counter,
                                                       uint32_t ncell, uint32_t ntiles) {
    // XCD-aware order: the hardware deals workgroups round-robin over the 8 XCDs, so give each
    // XCD one contiguous run of tiles; neighbouring tiles then share node rows in that XCD's L2.
    const uint32_t per = gridDim.x >> 3;
    const uint32_t tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (tile >= ntiles) return;
    const uint32_t c0 = tile * MC_TILE;
    const int t = threadIdx.x;
    const uint32_t c = c0 + t;
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    __shared__ TileLDS<T> s;
    bool cut = false;
    if (c < ncell) {
        const uint32_t i = c % n0, r = c / n0, j = r % n1, k = r / n1;
        const size_t sy = n0 + 1, sz = sy * (n1 + 1);
        T ph[8];
        const int lut = load_corners(phi, i + sy * j + sz * k, sy, sz, ph);
        cut = stage_cell(s, t, lut, i, j, k, g);
    }
    append_cut(cut, c, cutlist, counter);
    __syncthreads();
    stream_tile<T, NT>(cols, c0, (int)min((uint32_t)MC_TILE, ncell - c0), s);
}

template <typename T, bool FUSED, bool NT, bool SIGNS = false>
__global__ __launch_bounds__(MC_TILE) void k_mc_march(McBody<T> B, uint32_t nwork) {
    // XCD-aware order: the hardware deals workgroups round-robin over the 8 XCDs, so give each XCD
    // one contiguous run of work items; neighbouring rows then share node rows in that XCD's L2.
    const uint32_t per = gridDim.x >> 3;
    const uint32_t w = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (w >= nwork) return;
    __shared__ MarchLDS<T> L;
    mc_march_item<T, FUSED, NT, false, SIGNS>(B, w, L);
}

// The fused pass for a sphere union's culled or surrogate phi: brick-aligned items (MC_BRICK_W cells
// x MC_TILE / MC_BRICK_W rows, up to 16 layers) that stage their bricks' kept primitives.
template <typename T, bool NT>
__global__ __launch_bounds__(MC_TILE) void k_mc_march_bricks(McBody<T> B, uint32_t nwork) {
    const uint32_t per = gridDim.x >> 3;
    const uint32_t w = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (w >= nwork) return;
    __shared__ MarchLDS<T> L;
    __shared__ T sph[4 * MARCH_SPH];
    __shared__ BrickLists<T> bl;
    mc_march_item<T, true, NT, true>(B, w, L, sph, &bl);
}

template <typename T>
__global__ __launch_bounds__(MC_CUT_THREADS, 3) void k_mc_cut(const T *__restrict__ phi, GridT<T> g, Cols<T> cols,
                                                           const uint32_t *__restrict__ cutlist,
                                                           const uint32_t *__restrict__ counter) {
    __shared__ T pts[MC_CUT_SLOTS * MC_CUT_THREADS];
    const uint32_t count = *counter;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < count; w += gridDim.x * blockDim.x)
        mc_cut_item(phi, g, cols, cutlist[w], pts + threadIdx.x);
}

// _sample_nodes_kernel! (src/marching_squares/cache.jl:67-70): one thread per node, phi at the
// lattice node x0 + i*d.  A workgroup covers one 256-node segment of one node row.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_sample(T *__restrict__ phi, GridT<T> g, Geo geo,
                                                uint32_t m0, uint32_t m1, uint32_t nseg) {
    const uint32_t row = blockIdx.x / nseg, seg = blockIdx.x - row * nseg;
    const uint32_t i = seg * 256 + threadIdx.x;
    if (i >= m0) return;
    const uint32_t j = row % m1, k = row / m1;
    T x[3];
    x[0] = g.x0[0] + T(i) * g.d[0];
    x[1] = g.x0[1] + T(j) * g.d[1];
    x[2] = D == 3 ? g.x0[2] + T(k) * g.d[2] : T(0);
    phi[i + (size_t)m0 * (j + (size_t)m1 * k)] = sdf_value<T, D>(geo, x);
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0; }

__global__ __launch_bounds__(256) void k_screen_records(const double *__restrict__ prims, int n, double k,
                                                        float4 *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const double o[3] = {0.0, 0.0, 0.0};
    if (i < n) out[i] = screen_record(prims + 4 * (size_t)i, o, k, 0.0);
}

}  // namespace

template <typename T>
void launch_sample(T *phi, const GridT<T> &g, const Geo &geo, int dim, hipStream_t s) {
    const uint32_t m0 = g.n[0] + 1, m1 = g.n[1] + 1, m2 = dim == 3 ? g.n[2] + 1 : 1;
    const uint32_t nseg = (m0 + 255) / 256;
    const uint32_t blocks = nseg * m1 * m2;
    if (dim == 3)
        hipLaunchKernelGGL((k_sample<T, 3>), dim3(blocks), dim3(256), 0, s, phi, g, geo, m0, m1, nseg);
    else
        hipLaunchKernelGGL((k_sample<T, 2>), dim3(blocks), dim3(256), 0, s, phi, g, geo, m0, m1, nseg);
}

template <typename T>
void launch_mc_stream(T *phi, const GridT<T> &g, const Geo *geo, const Cols<T> &cols,
                      uint32_t *cutlist, uint32_t *counter, const McTuning &tune, hipStream_t s,
                      const BrickPhi *bp, const uint64_t *signs) {
    const uint32_t ncell = (uint32_t)((uint64_t)g.n[0] * g.n[1] * g.n[2]);
    if (tune.variant == MC_FLAT || (geo == nullptr && tune.variant == MC_MARCH_FUSED)) {
        const uint32_t ntiles = (ncell + MC_TILE - 1) / MC_TILE;
        const uint32_t blocks = ((ntiles + 7) / 8) * 8;
        if (tune.nontemporal)
            hipLaunchKernelGGL((k_mc_stream<T, true>), dim3(blocks), dim3(MC_TILE), 0, s, phi, g, cols,
                               cutlist, counter, ncell, ntiles);
        else
            hipLaunchKernelGGL((k_mc_stream<T, false>), dim3(blocks), dim3(MC_TILE), 0, s, phi, g,
                               cols, cutlist, counter, ncell, ntiles);
        return;
    }
    McBody<T> B;
    B.g = g;
    B.geo = geo ? *geo : Geo{};
    B.phi = phi;
    B.cols = cols;
    B.cutlist = cutlist;
    B.counter = counter;
    B.fused = tune.variant == MC_MARCH_FUSED && geo != nullptr;
    B.nontemporal = tune.nontemporal;
    B.bp = bp ? *bp : BrickPhi{};
    B.signs = B.fused ? nullptr : signs;
    if (B.signs) {
        int nb[3];
        phi_brick_counts(g.n, nb);
        B.sbx = nb[0];
        B.sby = nb[1];
    }
    if (B.fused && bp) {
        mc_brick_tile_shape(g.n, tune.kz > 0 ? tune.kz : 16, &B.W, &B.R, &B.kz, &B.nseg, &B.nrg, &B.nkb);
        const uint32_t nwork = B.nseg * B.nrg * B.nkb;
        const uint32_t blocks = ((nwork + 7) / 8) * 8;
        if (tune.nontemporal)
            hipLaunchKernelGGL((k_mc_march_bricks<T, true>), dim3(blocks), dim3(MC_TILE), 0, s, B, nwork);
        else
            hipLaunchKernelGGL((k_mc_march_bricks<T, false>), dim3(blocks), dim3(MC_TILE), 0, s, B, nwork);
        return;
    }
    mc_tile_shape(g.n, tune.kz > 0 ? tune.kz : 16, &B.W, &B.R, &B.kz, &B.nseg, &B.nrg, &B.nkb);
    const uint32_t nwork = B.nseg * B.nrg * B.nkb;
    const uint32_t blocks = ((nwork + 7) / 8) * 8;
#define CC_LAUNCH_MARCH(F, N, S) \
    hipLaunchKernelGGL((k_mc_march<T, F, N, S>), dim3(blocks), dim3(MC_TILE), 0, s, B, nwork)
    if (B.fused && tune.nontemporal) CC_LAUNCH_MARCH(true, true, false);
    else if (B.fused) CC_LAUNCH_MARCH(true, false, false);
    else if (B.signs && tune.nontemporal) CC_LAUNCH_MARCH(false, true, true);
    else if (B.signs) CC_LAUNCH_MARCH(false, false, true);
    else if (tune.nontemporal) CC_LAUNCH_MARCH(false, true, false);
    else CC_LAUNCH_MARCH(false, false, false);
#undef CC_LAUNCH_MARCH
}

template <typename T>
void launch_mc_cut(const T *phi, const GridT<T> &g, const Cols<T> &cols, const uint32_t *cutlist,
                   const uint32_t *counter, hipStream_t s) {
    hipLaunchKernelGGL((k_mc_cut<T>), dim3(MC_CUT_BLOCKS), dim3(MC_CUT_THREADS), 0, s, phi, g, cols, cutlist,
                       counter);
}

void mc_tile_shape(const int *n, uint32_t kz, uint32_t *W, uint32_t *R, uint32_t *KZ, uint32_t *nseg,
                   uint32_t *nrg, uint32_t *nkb) {
    const uint32_t n0 = n[0], n1 = n[1], n2 = n[2];
    *W = n0 < (uint32_t)MC_TILE ? n0 : (uint32_t)MC_TILE;
    *R = n0 < (uint32_t)MC_TILE ? (uint32_t)MC_TILE / n0 : 1u;
    if (*R > n1) *R = n1;
    *KZ = kz;
    *nseg = (n0 + *W - 1) / *W;
    *nrg = (n1 + *R - 1) / *R;
    *nkb = (n2 + kz - 1) / kz;
}

void mc_brick_tile_shape(const int *n, uint32_t kz, uint32_t *W, uint32_t *R, uint32_t *KZ, uint32_t *nseg,
                         uint32_t *nrg, uint32_t *nkb) {
    const uint32_t n0 = n[0], n1 = n[1], n2 = n[2];
    *W = n0 < (uint32_t)MC_BRICK_W ? n0 : (uint32_t)MC_BRICK_W;
    uint32_t r = PHI_BRICK;  // rows: a power of two dividing the brick edge, W R <= MC_TILE
    while (r > 1 && r * *W > (uint32_t)MC_TILE) r >>= 1;
    *R = n1 < r ? n1 : r;
    uint32_t k = PHI_BRICK;  // the largest divisor of the brick edge not above kz
    while (k > 1 && k > kz) k >>= 1;
    *KZ = k;
    *nseg = (n0 + *W - 1) / *W;
    *nrg = (n1 + *R - 1) / *R;
    *nkb = (n2 + k - 1) / k;
}

void launch_zero_u32(uint32_t *p, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, p);
}

void launch_screen_records(const double *prims, int n, double k, float4 *out, hipStream_t s) {
    hipLaunchKernelGGL(k_screen_records, dim3((n + 255) / 256), dim3(256), 0, s, prims, n, k, out);
}

template void launch_sample<double>(double *, const GridT<double> &, const Geo &, int, hipStream_t);
template void launch_sample<float>(float *, const GridT<float> &, const Geo &, int, hipStream_t);
template void launch_mc_stream<double>(double *, const GridT<double> &, const Geo *,
                                       const Cols<double> &, uint32_t *, uint32_t *,
                                       const McTuning &, hipStream_t, const BrickPhi *, const uint64_t *);
template void launch_mc_stream<float>(float *, const GridT<float> &, const Geo *, const Cols<float> &,
                                      uint32_t *, uint32_t *, const McTuning &, hipStream_t, const BrickPhi *,
                                      const uint64_t *);
template void launch_mc_cut<double>(const double *, const GridT<double> &, const Cols<double> &,
                                    const uint32_t *, const uint32_t *, hipStream_t);
template void launch_mc_cut<float>(const float *, const GridT<float> &, const Cols<float> &,
                                   const uint32_t *, const uint32_t *, hipStream_t);

}  // namespace cc
