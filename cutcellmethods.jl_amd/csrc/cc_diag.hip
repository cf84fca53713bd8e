// Diagnostics over a cache's cells, reduced on the device: cut_cell_report (src/cut_cell.jl:294-316),
// closure_residual (:318-347) and interface_normal_area / interface_area (:176-220) over every cell,
// and the shared-face check the reference's tests run (a face two cells share must be single-valued,
// bit for bit: test/test_mc_moments.jl:445-462, test/test_moments.jl:318-340).
//
// One fused pass reads each cell's kind, ambiguous flag, volume fraction, face fractions and face
// centroids (and its +x/+y/+z neighbours' for the shared faces).  Integer counts, the minimum volume
// fraction and the largest residual are combined with integer atomics (order-free); the floating
// sums (interface area and vector area) are folded per block over a fixed range of cells and then
// over the blocks in a fixed tree, so every result is the same on every run.
#include "cc_internal.h"

namespace cc {
namespace {

constexpr int DG_THREADS = 256;
constexpr int DG_BLOCKS = 2048;

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

// full_face_area (src/cut_cell.jl:246-253): the product of the sizes along every other axis.
template <typename T, int D> __device__ __forceinline__ T full_face(const T *cs, int dir) {
    const int ax = dir >> 1;
    if (D == 2) return cs[1 - ax];
    T p = T(1);
#pragma unroll
    for (int a = 0; a < D; ++a)
        if (a != ax) p = p * cs[a];
    return p;
}

template <typename T, int D>
__global__ __launch_bounds__(DG_THREADS) void k_diag(Cols<T> cols, GridT<T> g, size_t ncell, size_t per,
                                                     unsigned long long *__restrict__ counts,
                                                     T *__restrict__ part) {
    __shared__ T red[DG_THREADS];
    const size_t c0 = blockIdx.x * per, c1 = c0 + per < ncell ? c0 + per : ncell;
    const uint32_t n0 = g.n[0], n1 = g.n[1], n2 = D == 3 ? g.n[2] : 1;
    const size_t st[3] = {1, n0, (size_t)n0 * n1};
    const T cs[3] = {g.d[0], g.d[1], D == 3 ? g.d[2] : T(1)};
    u64 ncut = 0, nin = 0, nout = 0, namb = 0, nclo = 0, nff = 0, nfc = 0;
    u64 minvf = bits_of(1.0), maxres = 0;
    T area = T(0), vec[3] = {T(0), T(0), T(0)};
    for (size_t c = c0 + threadIdx.x; c < c1; c += DG_THREADS) {
        const int8_t kd = cols.kind[c];
        ncut += kd == CELL_CUT;
        nin += kd == CELL_INSIDE;
        nout += kd == CELL_OUTSIDE;
        namb += cols.amb[c] != 0;
        const T vf = cols.vf[c];
        if (kd == CELL_CUT && vf > T(0)) {
            const u64 b = bits_of((double)vf);
            minvf = b < minvf ? b : minvf;  // positive doubles order as their bit patterns
        }
        T ff[2 * D];
#pragma unroll
        for (int k = 0; k < 2 * D; ++k) ff[k] = cols.ff[2 * D * c + k];
        // interface_normal_area and closure_residual, the reference's expressions in its order
        T ina[D], acc[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            ina[a] = ff[2 * a + 1] * full_face<T, D>(cs, 2 * a + 1) - ff[2 * a] * full_face<T, D>(cs, 2 * a);
            acc[a] = T(0);
        }
#pragma unroll
        for (int dir = 0; dir < 2 * D; ++dir) {
            const T s = (dir & 1 ? T(1) : T(-1)) * (ff[dir] * full_face<T, D>(cs, dir));
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] = acc[a] + (a == (dir >> 1) ? s : T(0));
        }
        bool nonzero = false;
        T sq = T(0);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const T r = acc[a] - ina[a];
            nonzero |= r != T(0);
            const u64 b = bits_of(fabs((double)r));
            maxres = b > maxres ? b : maxres;
            sq = sq + ina[a] * ina[a];
            vec[a] = vec[a] + ina[a];
        }
        nclo += nonzero;
        area = area + sqrt(sq);
        // the shared faces towards +x, +y, +z: the neighbour's low slot against this cell's high slot
        const uint32_t ijk[3] = {(uint32_t)(c % n0), (uint32_t)((c / n0) % n1), (uint32_t)(c / st[2])};
        const uint32_t nn[3] = {n0, n1, n2};
#pragma unroll
        for (int a = 0; a < D; ++a) {
            if (ijk[a] + 1 >= nn[a]) continue;
            const size_t o = c + st[a];
            nff += differ(ff[2 * a + 1], cols.ff[2 * D * o + 2 * a]);
            bool fc = false;
#pragma unroll
            for (int q = 0; q < D - 1; ++q)
                fc |= differ(cols.fcl[2 * D * (D - 1) * c + (2 * a + 1) * (D - 1) + q],
                             cols.fcl[2 * D * (D - 1) * o + (2 * a) * (D - 1) + q]);
            nfc += fc;
        }
    }
    // the counts: folded per block, then one atomic per block
    __shared__ u64 ired[DG_THREADS];
    const u64 vals[7] = {ncut, nin, nout, namb, nclo, nff, nfc};
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const u64 s = block_fold(vals[q], ired);
        if (threadIdx.x == 0 && s) atomicAdd(&counts[q], s);
    }
    // the minimum volume fraction and the largest residual (bit patterns of non-negative doubles)
    ired[threadIdx.x] = minvf;
    __syncthreads();
    for (int w = DG_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && ired[threadIdx.x + w] < ired[threadIdx.x]) ired[threadIdx.x] = ired[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(&counts[7], ired[0]);
    __syncthreads();
    ired[threadIdx.x] = maxres;
    __syncthreads();
    for (int w = DG_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && ired[threadIdx.x + w] > ired[threadIdx.x]) ired[threadIdx.x] = ired[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&counts[8], ired[0]);
    // the floating sums, per block in a fixed tree
    const T a0 = block_fold(area, red);
    const T v0 = block_fold(vec[0], red);
    const T v1 = block_fold(vec[1], red);
    const T v2 = D == 3 ? block_fold(vec[2], red) : T(0);
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
}

}  // namespace

int diag_max_blocks() { return DG_BLOCKS; }

int diag_blocks(size_t ncell) {
    const size_t nb = ncell / DG_THREADS + (ncell % DG_THREADS != 0);  // no overflow near SIZE_MAX
    return (int)(nb == 0 ? 1 : (nb < (size_t)DG_BLOCKS ? nb : DG_BLOCKS));
}

template <typename T>
bool launch_diagnostics(const Cols<T> &cols, const GridT<T> &g, int dim, unsigned long long *counts, T *part,
                        size_t part_blocks, T *sums, hipStream_t s) {
    const size_t ncell = (size_t)g.n[0] * g.n[1] * (dim == 3 ? g.n[2] : 1);
    const int nb = diag_blocks(ncell);
    if ((size_t)nb > part_blocks) return false;  // part holds 4 values per block
    const size_t per = (ncell + nb - 1) / nb;
    hipLaunchKernelGGL(k_diag_init, dim3(1), dim3(64), 0, s, counts);
    if (dim == 3)
        hipLaunchKernelGGL((k_diag<T, 3>), dim3(nb), dim3(DG_THREADS), 0, s, cols, g, ncell, per, counts, part);
    else
        hipLaunchKernelGGL((k_diag<T, 2>), dim3(nb), dim3(DG_THREADS), 0, s, cols, g, ncell, per, counts, part);
    hipLaunchKernelGGL(k_diag_sums<T>, dim3(1), dim3(DG_THREADS), 0, s, (const T *)part, nb, sums);
    return true;
}

template bool launch_diagnostics<double>(const Cols<double> &, const GridT<double> &, int, unsigned long long *,
                                         double *, size_t, double *, hipStream_t);
template bool launch_diagnostics<float>(const Cols<float> &, const GridT<float> &, int, unsigned long long *,
                                        float *, size_t, float *, hipStream_t);

}  // namespace cc
