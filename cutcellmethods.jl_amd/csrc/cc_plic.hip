// PLIC passes for gfx950 -- _plic_fit_kernel! and _plic_face_fraction_kernel!
// (src/plic/cache.jl:96-162) with cell_plane (src/plic/plic.jl:72-91) and _face_area
// (src/plic/apertures.jl:73-125).
//
// The fit pass writes each cell's plane; the face pass evaluates, per cell and per face slot, the
// cell's own one-sided wet measure and its neighbour's through the one compiled `face_area` below,
// then the mean `(neg + pos)/2`, negative side first.  Both cells of an interior face form the same
// sum of the same two numbers from one code path, so the resolved face is bitwise single-valued.
#include <mutex>

#include "cc_plic_dev.h"

namespace cc {
namespace {

template <typename T, int D>
__global__ __launch_bounds__(256) void k_plic_fit(GridT<T> g, Geo geo, Cols<T> cols, T *normals,
                                                  T *intercepts, uint8_t *valid, uint32_t ncell) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell) plic_fit_cell<T, D>(g, geo, cols, normals, intercepts, valid, c);
}

template <typename T, int D>
__global__ __launch_bounds__(256) void k_plic_faces(GridT<T> g, Cols<T> cols,
                                                    const T *__restrict__ normals,
                                                    const T *__restrict__ intercepts,
                                                    const uint8_t *__restrict__ valid, uint32_t ncell) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell) plic_faces_cell<T, D>(g, cols, normals, intercepts, valid, c);
}

template <typename T>
__global__ __launch_bounds__(256) void k_plic_fused(GridT<T> g, Geo geo, Cols<T> cols, T *normals,
                                                    T *intercepts, uint8_t *valid, uint32_t nx,
                                                    uint32_t ny, uint32_t kz, uint32_t nwork) {
    const uint32_t per = gridDim.x >> 3;
    const uint32_t w = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (w >= nwork) return;
    __shared__ PlicLDS<T> L;
    plic_fused_item<T>(g, geo, cols, normals, intercepts, valid, w, nx, ny, kz, L);
}

template <typename T>
__global__ __launch_bounds__(PW_W * PW_R, 4) void k_plic_wide(GridT<T> g, Geo geo, Cols<T> cols, T *normals,
                                                          T *intercepts, uint8_t *valid, uint32_t nx,
                                                          uint32_t ny, uint32_t kz, uint32_t nwork) {
    const uint32_t per = gridDim.x >> 3;
    const uint32_t w = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (w >= nwork) return;
    __shared__ PlicWideLDS<T, PW_R> L;
    plic_wide_item<T, PW_R>(g, geo, cols, normals, intercepts, valid, w, nx, ny, kz, L);
}

// calc_volume (src/volume.jl:29-40): every cell's solid fraction from its plane (cell_plane, the fit
// the PLIC cache stores), summed.  Block b sums a fixed contiguous range of cells, lane t the cells
// t, t + 256, ... of it, then the block folds its 256 lane sums in a fixed tree; k_sum_parts folds
// the block sums the same way.  The order is fixed by the launch, so the result is the same on every
// run; it is a pairwise rather than the reference's left fold, so it agrees with that to rounding.
constexpr int VOL_THREADS = 256;

template <typename T> __device__ __forceinline__ T block_sum(T v) {
    __shared__ T s[VOL_THREADS];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = VOL_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

template <typename T, int D>
__global__ __launch_bounds__(VOL_THREADS) void k_plic_volume(GridT<T> g, Geo geo, size_t ncell, size_t per,
                                                             T *__restrict__ part) {
    const size_t c0 = blockIdx.x * per, c1 = c0 + per < ncell ? c0 + per : ncell;
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    T acc = T(0);
    for (size_t c = c0 + threadIdx.x; c < c1; c += VOL_THREADS) {
        const uint32_t i = (uint32_t)(c % n0), j = (uint32_t)((c / n0) % n1), k = (uint32_t)(c / ((size_t)n0 * n1));
        acc = acc + plic_plane<T, D>(g, geo, i, j, k).solid;
    }
    const T sum = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

template <typename T>
__global__ __launch_bounds__(VOL_THREADS) void k_sum_parts(const T *__restrict__ part, int nb, T *__restrict__ out) {
    T acc = T(0);
    for (int b = threadIdx.x; b < nb; b += VOL_THREADS) acc = acc + part[b];
    const T sum = block_sum(acc);
    if (threadIdx.x == 0) out[0] = sum;
}

}  // namespace

template <typename T>
int plic_volume_blocks(const GridT<T> &g, int dim) {
    size_t ncell = (size_t)g.n[0] * g.n[1] * (dim == 3 ? g.n[2] : 1);
    size_t nb = (ncell + VOL_THREADS - 1) / VOL_THREADS;
    return (int)(nb < VOL_SCRATCH_PARTS ? nb : VOL_SCRATCH_PARTS);
}

namespace {
// The allocation-free calc_volume's work arrays (block sums, result, primitive table), one copy per
// GPU for the life of the process.  Host code never passes this name to a kernel: it stands for a
// host-side shadow, not the device address, which volume_scratch looks up.
static __device__ __align__(256) unsigned char scratch[VOL_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};
}  // namespace

int volume_scratch(void **ptr) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= 64) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return (int)hipSuccess;
}

template <typename T>
void launch_plic_volume(const GridT<T> &g, const Geo &geo, int dim, T *part, T *out, hipStream_t s) {
    const size_t ncell = (size_t)g.n[0] * g.n[1] * (dim == 3 ? g.n[2] : 1);
    const int nb = plic_volume_blocks(g, dim);
    const size_t per = (ncell + nb - 1) / nb;
    if (dim == 3)
        hipLaunchKernelGGL((k_plic_volume<T, 3>), dim3(nb), dim3(VOL_THREADS), 0, s, g, geo, ncell, per, part);
    else
        hipLaunchKernelGGL((k_plic_volume<T, 2>), dim3(nb), dim3(VOL_THREADS), 0, s, g, geo, ncell, per, part);
    hipLaunchKernelGGL(k_sum_parts<T>, dim3(1), dim3(VOL_THREADS), 0, s, (const T *)part, nb, out);
}

template int plic_volume_blocks<double>(const GridT<double> &, int);
template int plic_volume_blocks<float>(const GridT<float> &, int);
template void launch_plic_volume<double>(const GridT<double> &, const Geo &, int, double *, double *, hipStream_t);
template void launch_plic_volume<float>(const GridT<float> &, const Geo &, int, float *, float *, hipStream_t);

template <typename T>
void launch_plic_fused(const GridT<T> &g, const Geo &geo, const Cols<T> &cols, T *normals,
                       T *intercepts, uint8_t *valid, uint32_t kz, bool wide, hipStream_t s) {
    const uint32_t tw = wide ? PW_W : PF_W, tr = wide ? PW_R : PF_R;
    const uint32_t nx = (g.n[0] + tw - 1) / tw, ny = (g.n[1] + tr - 1) / tr;
    const uint32_t nkb = (g.n[2] + kz - 1) / kz;
    const uint32_t nwork = nx * ny * nkb;
    const uint32_t blocks = ((nwork + 7) / 8) * 8;
    if (wide)
        hipLaunchKernelGGL((k_plic_wide<T>), dim3(blocks), dim3(PW_W * PW_R), 0, s, g, geo, cols, normals,
                           intercepts, valid, nx, ny, kz, nwork);
    else
        hipLaunchKernelGGL((k_plic_fused<T>), dim3(blocks), dim3(256), 0, s, g, geo, cols, normals,
                           intercepts, valid, nx, ny, kz, nwork);
}

template <typename T>
void launch_plic_fit(const GridT<T> &g, const Geo &geo, int dim, const Cols<T> &cols, T *normals,
                     T *intercepts, uint8_t *valid, hipStream_t s) {
    const uint32_t ncell = (uint32_t)g.n[0] * g.n[1] * (dim == 3 ? g.n[2] : 1);
    const dim3 grid((ncell + 255) / 256);
    if (dim == 3)
        hipLaunchKernelGGL((k_plic_fit<T, 3>), grid, dim3(256), 0, s, g, geo, cols, normals,
                           intercepts, valid, ncell);
    else
        hipLaunchKernelGGL((k_plic_fit<T, 2>), grid, dim3(256), 0, s, g, geo, cols, normals,
                           intercepts, valid, ncell);
}

template <typename T>
void launch_plic_faces(const GridT<T> &g, int dim, const Cols<T> &cols, const T *normals,
                       const T *intercepts, const uint8_t *valid, hipStream_t s) {
    const uint32_t ncell = (uint32_t)g.n[0] * g.n[1] * (dim == 3 ? g.n[2] : 1);
    const dim3 grid((ncell + 255) / 256);
    if (dim == 3)
        hipLaunchKernelGGL((k_plic_faces<T, 3>), grid, dim3(256), 0, s, g, cols, normals,
                           intercepts, valid, ncell);
    else
        hipLaunchKernelGGL((k_plic_faces<T, 2>), grid, dim3(256), 0, s, g, cols, normals,
                           intercepts, valid, ncell);
}

template void launch_plic_fit<double>(const GridT<double> &, const Geo &, int, const Cols<double> &,
                                      double *, double *, uint8_t *, hipStream_t);
template void launch_plic_fit<float>(const GridT<float> &, const Geo &, int, const Cols<float> &,
                                     float *, float *, uint8_t *, hipStream_t);
template void launch_plic_fused<double>(const GridT<double> &, const Geo &, const Cols<double> &,
                                        double *, double *, uint8_t *, uint32_t, bool, hipStream_t);
template void launch_plic_fused<float>(const GridT<float> &, const Geo &, const Cols<float> &, float *,
                                       float *, uint8_t *, uint32_t, bool, hipStream_t);
template void launch_plic_faces<double>(const GridT<double> &, int, const Cols<double> &,
                                        const double *, const double *, const uint8_t *, hipStream_t);
template void launch_plic_faces<float>(const GridT<float> &, int, const Cols<float> &, const float *,
                                       const float *, const uint8_t *, hipStream_t);

}  // namespace cc
