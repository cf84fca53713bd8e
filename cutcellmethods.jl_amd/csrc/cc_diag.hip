// Diagnostics over a cache's cells, reduced on the device: cut_cell_report (src/cut_cell.jl:302-317),
// closure_residual (:340-347), interface_normal_area (:224-226) and interface_area (:190-191) over every cell,
// and the shared-face check the reference's tests run (a face two cells share must be single-valued,
// bit for bit: test/test_mc_moments.jl:445-462, test/test_moments.jl:318-340).
//
// One fused pass reads each cell's kind, ambiguous flag, volume fraction, face fractions and face
// centroids once.  A block owns a band of 64 cells along x by 4 x DG_TJ rows by DG_KZ layers, each
// wave a strip of DG_TJ of its rows, and the waves march the band layer by layer together, each
// strip row by row: the +x neighbour's low face comes from the next lane, the +y neighbour's from the
// row before (in registers) or, for a strip's last row, from the next wave's first row through LDS,
// the +z neighbour's from the layer before (in LDS), so only the band's far edges read a neighbour
// again.  Integer counts, the minimum volume fraction and the largest residual are combined
// with integer atomics (order-free); the floating sums (interface area and vector area) are folded per
// lane in a fixed order, per block in a fixed tree and then over the blocks in a fixed tree, so every
// result is the same on every run.
#include "cc_internal.h"

namespace cc {
namespace {

constexpr int DG_THREADS = 256;  // four waves, one strip of the block's band each
constexpr int DG_BLOCKS = 2048;
constexpr int DG_TJ = 4;         // rows per strip
constexpr int DG_KZ = 64;        // layers per strip (3D)
constexpr int DG_BAND = DG_THREADS / 64 * DG_TJ;  // rows per band: a block's four strips, one above another

typedef unsigned long long u64;

template <typename T> __device__ __forceinline__ T block_fold(T v, T *s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = DG_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    const T r = s[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ u64 bits_of(double v) { return (u64)__double_as_longlong(v); }
__device__ __forceinline__ bool differ(double a, double b) { return __double_as_longlong(a) != __double_as_longlong(b); }
__device__ __forceinline__ bool differ(float a, float b) { return __float_as_uint(a) != __float_as_uint(b); }

template <typename T> struct PairOf;
template <> struct PairOf<double> { typedef double2 type; };
template <> struct PairOf<float> { typedef float2 type; };

// N consecutive values from p (N even, p aligned to two values) as N/2 paired loads.
template <typename T, int N> __device__ __forceinline__ void load_pairs(const T *__restrict__ p, T *out) {
    typedef typename PairOf<T>::type P;
#pragma unroll
    for (int q = 0; q < N / 2; ++q) {
        const P v = reinterpret_cast<const P *>(p)[q];
        out[2 * q] = v.x;
        out[2 * q + 1] = v.y;
    }
}

// full_face_area (src/cut_cell.jl:246-253): the product of the sizes along every other axis.
template <typename T, int D> __host__ __device__ __forceinline__ T full_face(const T *cs, int dir) {
    const int ax = dir >> 1;
    if (D == 2) return cs[1 - ax];
    T p = T(1);
#pragma unroll
    for (int a = 0; a < D; ++a)
        if (a != ax) p = p * cs[a];
    return p;
}

// One lane's running terms.  The counts are 32-bit per lane (a lane visits far fewer than 2^32 cells)
// and widen to 64 bits when the block folds them.
// full_face_area of each axis's two faces, computed once on the host (the same IEEE products in the
// same order) and passed by value, so the kernel holds them in scalar registers.
template <typename T> struct FaceAreas {
    T a[3];
};

template <typename T> struct Acc {
    uint32_t cnt[7];  // cut, inside, outside, ambiguous, closure nonzero, face mismatches, centroid mismatches
    uint32_t nonfinite;  // cells whose volume fraction or a face fraction is NaN or infinite
    u64 minvf, maxres;
    T area, vec[3];
};

// One cell's report terms, interface_normal_area and closure_residual (the reference's expressions
// in its order), folded into the lane's accumulators.
template <typename T, int D>
__device__ __forceinline__ void cell_terms(Acc<T> &a, int8_t kd, uint8_t amb, T vf, const T *ff, const T *fa) {
    a.cnt[0] += kd == CELL_CUT;
    a.cnt[1] += kd == CELL_INSIDE;
    a.cnt[2] += kd == CELL_OUTSIDE;
    a.cnt[3] += amb != 0;
    if (kd == CELL_CUT && vf > T(0)) {
        const u64 b = bits_of((double)vf);
        a.minvf = b < a.minvf ? b : a.minvf;  // positive doubles order as their bit patterns
    }
    bool finite = isfinite(vf);
#pragma unroll
    for (int k = 0; k < 2 * D; ++k) finite &= isfinite(ff[k]);
    a.nonfinite += !finite;
    T ina[D], acc[D];
#pragma unroll
    for (int x = 0; x < D; ++x) {
        ina[x] = ff[2 * x + 1] * fa[x] - ff[2 * x] * fa[x];
        acc[x] = T(0);
    }
#pragma unroll
    for (int dir = 0; dir < 2 * D; ++dir) {
        const T s = (dir & 1 ? T(1) : T(-1)) * (ff[dir] * fa[dir >> 1]);
#pragma unroll
        for (int x = 0; x < D; ++x) acc[x] = acc[x] + (x == (dir >> 1) ? s : T(0));
    }
    bool nonzero = false;
    T sq = T(0);
#pragma unroll
    for (int x = 0; x < D; ++x) {
        const T r = acc[x] - ina[x];
        nonzero |= r != T(0);
        const u64 b = bits_of(fabs((double)r));
        a.maxres = b > a.maxres ? b : a.maxres;
        sq = sq + ina[x] * ina[x];
        a.vec[x] = a.vec[x] + ina[x];
    }
    a.cnt[4] += nonzero;
    a.area = a.area + sqrt(sq);
}

// A face two cells share: this cell's fraction and centroid on it against the neighbour's.
template <typename T, int D>
__device__ __forceinline__ void shared_face(Acc<T> &a, T f, const T *fc, T g, const T *gc) {
    a.cnt[5] += differ(f, g);
    bool m = false;
#pragma unroll
    for (int q = 0; q < D - 1; ++q) m |= differ(fc[q], gc[q]);
    a.cnt[6] += m;
}

template <typename T, int D>
__global__ __launch_bounds__(DG_THREADS) void k_diag(Cols<T> cols, GridT<T> g, FaceAreas<T> fa, int sx, int sy,
                                                     int nitems, unsigned long long *__restrict__ counts,
                                                     T *__restrict__ part) {
    constexpr int NF = 2 * D, NC = 2 * D * (D - 1), ZQ = D == 3 ? 3 : 1, NW = DG_THREADS / 64;
    // the layer before: each row's +z face fraction and centroid, per wave, row and lane (lane
    // fastest, so a wave's access is one contiguous run); a lane reads back only what it wrote
    __shared__ T zq[NW][DG_TJ][ZQ][64];
    // each wave's first row's -y face, for the previous wave (the rows just before it); two buffers
    // used in turn layer by layer, so one barrier per layer separates a layer's writes from its reads
    // and from the writes of the layer after next
    __shared__ T yq[2][NW][D][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: strip bounds in scalars
    const int n0 = g.n[0], n1 = g.n[1], n2 = D == 3 ? g.n[2] : 1;
    const size_t row = (size_t)n0, plane = (size_t)n0 * n1;
    // blocks are dealt to the 8 XCDs in turn: renumber so that consecutive bands (neighbouring rows
    // of one layer range) run on one XCD and share its L2 for the rows a band reads again
    const int nb = gridDim.x;
    const int lb = nb % 8 == 0 ? (int)(blockIdx.x % 8) * (nb / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
    Acc<T> a;
#pragma unroll
    for (int q = 0; q < 7; ++q) a.cnt[q] = 0;
    a.minvf = bits_of(1.0);
    a.maxres = 0;
    a.nonfinite = 0;
    a.area = a.vec[0] = a.vec[1] = a.vec[2] = T(0);
    // one band (64 cells x NW * DG_TJ rows x DG_KZ layers) per block and loop turn, the same for every
    // wave of the block, so every wave meets every barrier; wave w owns rows [j0, j1) of the band.  The
    // exchange buffer alternates with every layer the block marches, across bands too.
    int turn = 0;
    for (int item = lb; item < nitems; item += nb) {
        const int tx = item % sx, r = item / sx, ty = r % sy, tz = r / sy;
        const int i = tx * 64 + lane, jb = ty * NW * DG_TJ, j0 = jb + wave * DG_TJ;
        const int j1 = j0 + DG_TJ < n1 ? j0 + DG_TJ : n1;  // j1 <= j0: a wave past the grid's rows
        const int k0 = D == 3 ? tz * DG_KZ : 0;
        const int k1 = D == 3 ? (k0 + DG_KZ < n2 ? k0 + DG_KZ : n2) : 1;
        const bool vi = i < n0, xn = i + 1 < n0;
        // the next wave in the band starts at j1 and has rows: its first -y face comes through yq
        const bool from_next = wave + 1 < NW && j1 < n1;
        // lanes past the grid read its last cell and count nothing; a lane's offset within its row
        // of the strip is small, the row's start is the same for the whole wave
        const uint32_t lo = (uint32_t)(vi ? lane : n0 - 1 - tx * 64);
        for (int k = k0; k < k1; ++k, ++turn) {
            const int par = turn & 1;
            T yf = T(0), yc[D - 1];  // the row before: its +y face
#pragma unroll
            for (int q = 0; q < D - 1; ++q) yc[q] = T(0);
#pragma unroll 1
            for (int j = j0; j < j1; ++j) {
                const int jj = j - j0;
                const size_t c0 = (size_t)tx * 64 + row * j + plane * k;  // the row's first cell
                const T *rff = cols.ff + NF * c0, *rfc = cols.fcl + NC * c0;
                T ff[NF], fc[NC];
                load_pairs<T, NF>(rff + NF * lo, ff);
                load_pairs<T, NC>(rfc + NC * lo, fc);
                const int8_t kd = cols.kind[c0 + lo];
                const uint8_t am = cols.amb[c0 + lo];
                const T vf = cols.vf[c0 + lo];
                if (vi) cell_terms<T, D>(a, kd, am, vf, ff, fa.a);
                // +x: the next lane's -x face (the next strip's first cell for the last lane)
                T nf = __shfl_down(ff[0], 1), nc[D - 1];
#pragma unroll
                for (int q = 0; q < D - 1; ++q) nc[q] = __shfl_down(fc[q], 1);
                if (lane == 63 && xn) {
                    nf = rff[NF * 64];
#pragma unroll
                    for (int q = 0; q < D - 1; ++q) nc[q] = rfc[NC * 64 + q];
                }
                if (vi && xn) shared_face<T, D>(a, ff[1], fc + (D - 1), nf, nc);
                // +y: the row before's +y face against this row's -y face
                if (vi && j > j0) shared_face<T, D>(a, yf, yc, ff[2], fc + 2 * (D - 1));
                if (j == j0 && wave > 0) {  // the previous wave compares its last row against this row
                    yq[par][wave][0][lane] = ff[2];
#pragma unroll
                    for (int q = 0; q < D - 1; ++q) yq[par][wave][1 + q][lane] = fc[2 * (D - 1) + q];
                }
                yf = ff[3];
#pragma unroll
                for (int q = 0; q < D - 1; ++q) yc[q] = fc[3 * (D - 1) + q];
                if (D == 3) {  // +z: the layer before's +z face against this layer's -z face
                    if (vi && k > k0) {
                        T zc[D - 1];
#pragma unroll
                        for (int q = 0; q < D - 1; ++q) zc[q] = zq[wave][jj][1 + q][lane];
                        shared_face<T, D>(a, zq[wave][jj][0][lane], zc, ff[4], fc + 4 * (D - 1));
                    }
                    zq[wave][jj][0][lane] = ff[5];
#pragma unroll
                    for (int q = 0; q < D - 1; ++q) zq[wave][jj][1 + q][lane] = fc[5 * (D - 1) + q];
                }
            }
            __syncthreads();
            if (vi && j0 < j1 && j1 < n1) {  // the strip's last row against the first row past it
                T g0, gc[D - 1];
                if (from_next) {
                    g0 = yq[par][wave + 1][0][lane];
#pragma unroll
                    for (int q = 0; q < D - 1; ++q) gc[q] = yq[par][wave + 1][1 + q][lane];
                } else {  // past the band: read from memory
                    const size_t c = (size_t)i + row * j1 + plane * k;
                    g0 = cols.ff[NF * c + 2];
#pragma unroll
                    for (int q = 0; q < D - 1; ++q) gc[q] = cols.fcl[NC * c + 2 * (D - 1) + q];
                }
                shared_face<T, D>(a, yf, yc, g0, gc);
            }
        }
        if (D == 3 && vi && k1 < n2) {  // the strip's last layer against the first layer past it
#pragma unroll 1
            for (int j = j0; j < j1; ++j) {
                const int jj = j - j0;
                const size_t c = (size_t)i + row * j + plane * k1;
                T gc[D - 1], zc[D - 1];
#pragma unroll
                for (int q = 0; q < D - 1; ++q) {
                    gc[q] = cols.fcl[NC * c + 4 * (D - 1) + q];
                    zc[q] = zq[wave][jj][1 + q][lane];
                }
                shared_face<T, D>(a, zq[wave][jj][0][lane], zc, cols.ff[NF * c + 4], gc);
            }
        }
    }
    // the counts: folded per block, then one atomic per block
    __shared__ u64 ired[DG_THREADS];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const u64 s = block_fold((u64)a.cnt[q], ired);
        if (threadIdx.x == 0 && s) atomicAdd(&counts[q], s);
    }
    {
        const u64 s = block_fold((u64)a.nonfinite, ired);
        if (threadIdx.x == 0 && s) atomicAdd(&counts[9], s);
    }
    // the minimum volume fraction and the largest residual (bit patterns of non-negative doubles)
    ired[threadIdx.x] = a.minvf;
    __syncthreads();
    for (int w = DG_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && ired[threadIdx.x + w] < ired[threadIdx.x]) ired[threadIdx.x] = ired[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(&counts[7], ired[0]);
    __syncthreads();
    ired[threadIdx.x] = a.maxres;
    __syncthreads();
    for (int w = DG_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && ired[threadIdx.x + w] > ired[threadIdx.x]) ired[threadIdx.x] = ired[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&counts[8], ired[0]);
    // the floating sums, per block in a fixed tree
    __shared__ T red[DG_THREADS];
    const T a0 = block_fold(a.area, red);
    const T v0 = block_fold(a.vec[0], red);
    const T v1 = block_fold(a.vec[1], red);
    const T v2 = D == 3 ? block_fold(a.vec[2], red) : T(0);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = a0;
        part[4 * blockIdx.x + 1] = v0;
        part[4 * blockIdx.x + 2] = v1;
        part[4 * blockIdx.x + 3] = v2;
    }
}

template <typename T>
__global__ __launch_bounds__(DG_THREADS) void k_diag_sums(const T *__restrict__ part, int nb, T *__restrict__ out) {
    __shared__ T red[DG_THREADS];
    for (int q = 0; q < 4; ++q) {
        T acc = T(0);
        for (int b = threadIdx.x; b < nb; b += DG_THREADS) acc = acc + part[4 * b + q];
        const T s = block_fold(acc, red);
        if (threadIdx.x == 0) out[q] = s;
    }
}

__global__ void k_diag_init(unsigned long long *counts) {
    const int t = threadIdx.x;
    if (t < 7) counts[t] = 0;
    if (t == 7) counts[7] = bits_of(1.0);
    if (t == 8) counts[8] = 0;
    if (t == 9) counts[9] = 0;
}

}  // namespace

int diag_max_blocks() { return DG_BLOCKS; }

// Bands of 64 x DG_BAND x DG_KZ cells, one per block, at most DG_BLOCKS blocks (each loops over
// bands), at least one.
void diag_shape(const int *n, int dim, int &sx, int &sy, int64_t &items, int &blocks) {
    const int n2 = dim == 3 ? n[2] : 1;
    sx = (n[0] + 63) / 64;
    sy = (n[1] + DG_BAND - 1) / DG_BAND;
    const int64_t sz = dim == 3 ? (n2 + DG_KZ - 1) / DG_KZ : 1;
    items = (int64_t)sx * sy * sz;
    blocks = (int)(items < 1 ? 1 : (items < DG_BLOCKS ? items : DG_BLOCKS));
}

template <typename T>
bool launch_diagnostics(const Cols<T> &cols, const GridT<T> &g, int dim, unsigned long long *counts, T *part,
                        size_t part_blocks, T *sums, hipStream_t s) {
    int sx, sy, nb;
    int64_t items;
    diag_shape(g.n, dim, sx, sy, items, nb);
    if ((size_t)nb > part_blocks || items > INT32_MAX) return false;  // part holds 4 values per block
    const T cs[3] = {g.d[0], g.d[1], dim == 3 ? g.d[2] : T(1)};
    FaceAreas<T> fa;
    for (int ax = 0; ax < 3; ++ax)
        fa.a[ax] = ax >= dim ? T(0) : (dim == 3 ? full_face<T, 3>(cs, 2 * ax) : full_face<T, 2>(cs, 2 * ax));
    hipLaunchKernelGGL(k_diag_init, dim3(1), dim3(64), 0, s, counts);
    if (dim == 3)
        hipLaunchKernelGGL((k_diag<T, 3>), dim3(nb), dim3(DG_THREADS), 0, s, cols, g, fa, sx, sy, (int)items,
                           counts, part);
    else
        hipLaunchKernelGGL((k_diag<T, 2>), dim3(nb), dim3(DG_THREADS), 0, s, cols, g, fa, sx, sy, (int)items,
                           counts, part);
    hipLaunchKernelGGL(k_diag_sums<T>, dim3(1), dim3(DG_THREADS), 0, s, (const T *)part, nb, sums);
    return true;
}

template bool launch_diagnostics<double>(const Cols<double> &, const GridT<double> &, int, unsigned long long *,
                                         double *, size_t, double *, hipStream_t);
template bool launch_diagnostics<float>(const Cols<float> &, const GridT<float> &, int, unsigned long long *,
                                        float *, size_t, float *, hipStream_t);

}  // namespace cc
