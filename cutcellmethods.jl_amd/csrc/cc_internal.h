// Host-side declarations shared by the kernel translation units and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cc_device.h"

namespace cc {

// One device pointer per CutCellData column (src/cut_cell.jl:112-120).
template <typename T> struct Cols {
    int8_t *kind;
    uint8_t *amb;
    T *vf;
    T *centroid;
    T *ff;
    T *fcl;
    T *ic;
};

// Tile width of the marching-cubes stream pass: cells per workgroup, one per thread.
constexpr int MC_TILE = 256;
// Workgroups of the cut-cell pass (a grid-stride loop over the device-side count).
constexpr int MC_CUT_BLOCKS = 2048;

template <typename T>
void launch_sample(T *phi, const GridT<T> &g, const Geo &geo, int dim, hipStream_t s);

// Variants of the marching-cubes stream pass (picked per cache; the autotuner sweeps them).
enum McVariant { MC_FLAT = 0, MC_MARCH = 1, MC_MARCH_FUSED = 2 };
struct McTuning {
    int variant = MC_MARCH_FUSED;  // fused needs an analytic body; falls back to MC_FLAT otherwise
    int kz = 16;                   // z-march depth (layers per workgroup)
    int nontemporal = 0;           // streaming (non-temporal) stores for the column writes
};

// The stream pass.  `geo` is the body when the variant samples it in-kernel (MC_MARCH_FUSED, phi is
// then written by this pass); otherwise phi must already hold the nodal field.
template <typename T>
void launch_mc_stream(T *phi, const GridT<T> &g, const Geo *geo, const Cols<T> &cols,
                      uint32_t *cutlist, uint32_t *counter, const McTuning &tune, hipStream_t s);
template <typename T>
void launch_mc_cut(const T *phi, const GridT<T> &g, const Cols<T> &cols, const uint32_t *cutlist,
                   const uint32_t *counter, hipStream_t s);

template <typename T>
void launch_ms_cells(const T *phi, const GridT<T> &g, const Cols<T> &cols, uint32_t *counter,
                     hipStream_t s);

template <typename T>
void launch_plic_fit(const GridT<T> &g, const Geo &geo, int dim, const Cols<T> &cols, T *normals,
                     T *intercepts, uint8_t *valid, hipStream_t s);
template <typename T>
void launch_plic_faces(const GridT<T> &g, int dim, const Cols<T> &cols, const T *normals,
                       const T *intercepts, const uint8_t *valid, hipStream_t s);

// The fused 3D PLIC pass (fit + resolved faces in one sweep, no read-back).
template <typename T>
void launch_plic_fused(const GridT<T> &g, const Geo &geo, const Cols<T> &cols, T *normals,
                       T *intercepts, uint8_t *valid, uint32_t kz, hipStream_t s);

void launch_zero_u32(uint32_t *p, hipStream_t s);

// Tile shape of the z-marching stream pass for a grid of n cells: W cells along x (a 256-cell run,
// or the whole row when narrower), R whole rows per tile when rows are short, kz layers per item.
void mc_tile_shape(const int *n, uint32_t kz, uint32_t *W, uint32_t *R, uint32_t *KZ, uint32_t *nseg,
                   uint32_t *nrg, uint32_t *nkb);

}  // namespace cc
