// Host-side declarations shared by the kernel translation units and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cc_device.h"

// Hooks between the library's own translation units (the cache is defined in cc_api.hip).  Hidden:
// they link inside libcutcell_hip.so and are not part of its exported C ABI.
#define CC_HIDDEN __attribute__((visibility("hidden")))
struct cc_cache;
extern "C" {
CC_HIDDEN void cc_internal_set_error(const char *msg);
CC_HIDDEN int cc_internal_body(cc_cache *c, int *method, int *dtype, int64_t *n, void **cols, void **cutlist,
                               void **counter, int *kz);
CC_HIDDEN int cc_internal_leaves(const cc_cache *c);
// A body of a batched step was updated: a surface marched before describes an older field.
CC_HIDDEN void cc_internal_touch(cc_cache *c);
// Makes stream s wait for the work cc_update_async left on the cache's stream (nothing if none).
CC_HIDDEN int cc_internal_order_after(cc_cache *c, hipStream_t s);
}

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
// Workgroups of the cut-cell pass (a grid-stride loop over the device-side count) and their size.
// Each lane stages its cell's 13 points in MC_CUT_SLOTS values of LDS: one per edge crossing (its
// coordinate along the edge; the other two are the edge's corner coordinates) and three for the
// interior vertex.
constexpr int MC_CUT_BLOCKS = 4096;
constexpr int MC_CUT_THREADS = 128;
constexpr int MC_CUT_SLOTS = 15;

template <typename T>
void launch_sample(T *phi, const GridT<T> &g, const Geo &geo, int dim, hipStream_t s);
// A 3D sphere union's screen records about the origin (screen_record with R = 0): out[i] for the n
// primitives at prims, smoothing k.
void launch_screen_records(const double *prims, int n, double k, float4 *out, hipStream_t s);

// Variants of the marching-cubes stream pass (picked per cache; the autotuner sweeps them).
enum McVariant { MC_FLAT = 0, MC_MARCH = 1, MC_MARCH_FUSED = 2 };
struct McTuning {
    int variant = MC_MARCH_FUSED;  // fused needs an analytic body; falls back to MC_FLAT otherwise
    int kz = 4;                    // z-march depth (layers per workgroup)
    int nontemporal = 0;           // streaming (non-temporal) stores for the column writes
};

struct BrickPhi;
// The stream pass.  `geo` is the body when the variant samples it in-kernel (MC_MARCH_FUSED, phi is
// then written by this pass); otherwise phi must already hold the nodal field.  With `bp` (and the
// fused variant) a sphere union's culled or surrogate phi is evaluated in-kernel from the bricks.
// With `signs` (the z-march reading a field, MC_MARCH) each node's side comes from the phi pass's
// sign bitmap instead of phi.
template <typename T>
void launch_mc_stream(T *phi, const GridT<T> &g, const Geo *geo, const Cols<T> &cols,
                      uint32_t *cutlist, uint32_t *counter, const McTuning &tune, hipStream_t s,
                      const BrickPhi *bp = nullptr, const uint64_t *signs = nullptr);
template <typename T>
void launch_mc_cut(const T *phi, const GridT<T> &g, const Cols<T> &cols, const uint32_t *cutlist,
                   const uint32_t *counter, hipStream_t s);

template <typename T>
void launch_ms_cells(const T *phi, const GridT<T> &g, const Cols<T> &cols, uint32_t *counter,
                     hipStream_t s);
// The 2D cell pass sampling an analytic body itself (phi written once, bitwise the sampling pass's).
template <typename T>
void launch_ms_cells_fused(const Geo &geo, T *phi, const GridT<T> &g, const Cols<T> &cols, uint32_t *counter,
                           hipStream_t s);

// The quadtree pass: leaf c at `level[c]`, lower corner `coord[2c..2c+1]` on its level's lattice
// (spacing d / 2^level from x0); samples `geo` at each leaf's corners into phi4 (four per leaf) --
// or, with geo == nullptr, reads them from phi4 -- and writes one record per leaf.
template <typename T>
void launch_leaf_cells(const Geo *geo, const int32_t *level, const int32_t *coord, uint32_t n, const double *x0,
                       const double *d, T *phi4, const Cols<T> &cols, uint32_t *counter, hipStream_t s);

template <typename T>
void launch_plic_fit(const GridT<T> &g, const Geo &geo, int dim, const Cols<T> &cols, T *normals,
                     T *intercepts, uint8_t *valid, hipStream_t s);
template <typename T>
void launch_plic_faces(const GridT<T> &g, int dim, const Cols<T> &cols, const T *normals,
                       const T *intercepts, const uint8_t *valid, hipStream_t s);

// calc_volume: the sum of every cell's solid fraction (out[0]) through plic_volume_blocks(g, dim)
// block sums in `part`; the caller multiplies by the cell volume.
template <typename T> int plic_volume_blocks(const GridT<T> &g, int dim);
template <typename T>
void launch_plic_volume(const GridT<T> &g, const Geo &geo, int dim, T *part, T *out, hipStream_t s);

// calc_volume's work arrays without allocation: a file-scope device array of VOL_SCRATCH_BYTES per
// GPU, laid out as the block sums (at most VOL_SCRATCH_PARTS of them, which every grid's
// plic_volume_blocks fits), the result at VOL_SCRATCH_OUT, and a body's primitive table from
// VOL_SCRATCH_PRIMS on (at most VOL_SCRATCH_MAX_PRIM_DOUBLES doubles).  volume_scratch gives the
// current device's address of that array, looked up once per device; it returns a hipError_t.
constexpr size_t VOL_SCRATCH_BYTES = 256 * 1024;
constexpr size_t VOL_SCRATCH_PARTS = 2048;
constexpr size_t VOL_SCRATCH_OUT = VOL_SCRATCH_PARTS * 8;
constexpr size_t VOL_SCRATCH_PRIMS = VOL_SCRATCH_OUT + 256;
constexpr size_t VOL_SCRATCH_MAX_PRIM_DOUBLES = (VOL_SCRATCH_BYTES - VOL_SCRATCH_PRIMS) / 8;
int volume_scratch(void **ptr);

// The fused 3D PLIC pass (fit + resolved faces in one sweep, no read-back): `wide` picks 64 x 8
// tiles with one wave per row, else 32 x 8 tiles.
template <typename T>
void launch_plic_fused(const GridT<T> &g, const Geo &geo, const Cols<T> &cols, T *normals,
                       T *intercepts, uint8_t *valid, uint32_t kz, bool wide, hipStream_t s);

void launch_zero_u32(uint32_t *p, hipStream_t s);

// Tile shape of the z-marching stream pass for a grid of n cells: W cells along x (a 256-cell run,
// or the whole row when narrower), R whole rows per tile when rows are short, kz layers per item.
void mc_tile_shape(const int *n, uint32_t kz, uint32_t *W, uint32_t *R, uint32_t *KZ, uint32_t *nseg,
                   uint32_t *nrg, uint32_t *nkb);
// The brick-aligned tiles of the fused pass over a sphere union's culled or surrogate phi:
// MC_BRICK_W cells along x (a multiple of the brick edge; fewer on a narrower grid) by a power of two
// of rows dividing the brick edge, kz rounded down to a power of two dividing it.  One brick's
// footprint (16 x 16 cells) measured fastest: the exact nodes cluster in 4^3-node sub-bricks, which
// a compact tile meets in fewer waves than a strip (64-, 128- and 256-cell strips were 1.3-1.7x
// slower on the composite body at 512^3).
constexpr int MC_BRICK_W = 16;
void mc_brick_tile_shape(const int *n, uint32_t kz, uint32_t *W, uint32_t *R, uint32_t *KZ, uint32_t *nseg,
                         uint32_t *nrg, uint32_t *nkb);

// ---- nodal phi of a sphere union, brick by brick ------------------------------------------------
// The node lattice is cut into bricks of PHI_BRICK^3 nodes.  For each brick a cull pass keeps, in
// fold order, the primitives that a bound over the brick's box cannot prove inert: primitive i is
// dropped when its lower bound over the box exceeds min_{j<i} (upper bound of primitive j) by more
// than the smoothing k (the polynomial smooth minimum of a and b with b - a >= k returns a exactly,
// and the fold so far never exceeds any earlier primitive), so the culled fold is bitwise the full
// one.  The phi pass then folds only the kept primitives, at every node (culled exact) or, with a
// surrogate, only where the surrogate cannot be trusted.
constexpr int PHI_BRICK = 16;  // nodes per brick edge
constexpr int PHI_SUB = 4;     // nodes per certification sub-brick edge

// Per-brick counters of the phi pass, PHI_BSTAT uint32 per brick: nodes written, nodes evaluated
// with the exact (culled) fold, exact evaluations at sub-brick centres (sign certificates), nodes
// in sub-bricks whose sign could not be certified, nodes with |phi_nn| under the band, nodes whose
// network sign contradicts the certified sign.
constexpr int PHI_BSTAT = 8;

// A lattice of points for the surrogate network: node (i, j, k) at x0 + (i, j, k) * h.
struct Lattice {
    float x0[3];
    float h[3];
    int n[3];
};

// The surrogate's inputs to the phi pass: network values on a coarse lattice of spacing f cells
// whose node (I, J, K) sits on fine node (f I, f J, f K); nn == nullptr means exact everywhere.
struct SurrogateArgs {
    const float *nn;
    int nc[3];
    int f;
    int fshift;    // log2 f when f is a power of two, else -1
    float finv;    // 1 / f (exact when f is a power of two)
    float band;    // nodes with |phi_nn| < band are evaluated exactly
    double scale;  // bound on the magnitudes (coordinates, radii, distances) the fold rounds
};

inline void phi_brick_counts(const int *n, int *nb) {
    for (int a = 0; a < 3; ++a) nb[a] = (n[a] + 1 + PHI_BRICK - 1) / PHI_BRICK;
}

// The culled or surrogate phi as the fused stream pass evaluates it at a node: the node's brick
// (nbx x nby bricks per layer) folds its kept primitives list[b * nprims ..][0, cnt[b]), unless the
// sub-brick's sign certificate cert[64 b + sub] (null: culled exact) and the network value allow
// keeping the latter -- the phi pass's decisions exactly.  wstat receives PHI_BSTAT counters per
// work item of the stream pass, laid out as the phi pass's per-brick counters (no certificates).
struct BrickPhi {
    const uint16_t *list;
    const uint32_t *cnt;
    const int8_t *cert;
    SurrogateArgs sa;
    int nbx, nby, nprims;
    uint32_t *wstat;
};

// The cull pass; with cert non-null it also certifies the signs of each brick's 64 sub-bricks
// (int8 per sub-brick: +1, -1, or 0 when uncertified) and writes each brick's PHI_BSTAT counters
// into bstat (certificate evaluations only).
template <typename T>
void launch_cull(const GridT<T> &g, const Geo &geo, double scale, uint16_t *list, uint32_t *cnt,
                 int8_t *cert, uint32_t *bstat, hipStream_t s);
// The phi pass; with signs non-null it also leaves the brick-major sign bitmap of phi there (64
// words per brick; cc_brick_dev.h sign_bit) for the stream pass to read instead of phi.
template <typename T>
void launch_phi_bricks(T *phi, const GridT<T> &g, const Geo &geo, const uint16_t *list,
                       const uint32_t *cnt, const SurrogateArgs &sa, uint32_t *bstat, uint64_t *signs,
                       hipStream_t s);

// ---- the whole-grid marching-cubes surface (generate_mesh) ---------------------------------------
// Sizes of the march's per-grid work arrays for a grid of n cells: node rows (j, k) of (n0+1)
// nodes and cell rows of n0 cells, an edge bitmap of 3 bits per node, a cut bitmap of 1 bit per cell.
struct MeshSizes {
    size_t node_rows, cell_rows;
    size_t ewords, cwords;  // 64-bit words of the edge and cut bitmaps
};
MeshSizes mesh_sizes(const int *n);
// Blocks of a prefix sum over n values (the block-sum array holds this many per sum).
size_t mesh_scan_blocks(size_t n);

// The march's device work arrays (owned by the cache).
struct MeshBuffers {
    void *ebits;              // [node_rows][W] edge bits, bit 3 i + axis
    void *cbits;              // [cell_rows][Wc] cut-cell bits
    uint32_t *erank, *crank;  // per bitmap word: the set bits in all words before it
    unsigned long long *part; // [2][part_cap] block sums of a pair of prefix sums
    size_t part_cap;
    int64_t *totals;          // [4] edge vertices, cut cells, triangles, interior vertices
    // per cut cell in cube order, capacity cut_cap
    uint32_t *order;          // the cell's linear index
    uint32_t *tn, *cn;        // its triangle count, 1 if its tiling uses the interior vertex
    int64_t *tpre, *cpre;     // their exclusive prefix sums
    size_t cut_cap;
};

// Up to three exclusive prefix sums in one set of launches.  A segment sums `in` (uint32 counts of
// at most 64 each) or, with `bits`, the popcounts of bitmap words; it writes int64 `out` or uint32
// `out32`.
struct MeshScan {
    struct Seg {
        const uint32_t *in;
        const unsigned long long *bits;
        int64_t *out;
        uint32_t *out32;
        size_t n;
    } seg[3];
    unsigned long long *part;
    size_t maxblocks;
    int64_t *totals;
};

// Two exclusive prefix sums in one set of launches (block sums in part, part_cap per sum; the two
// totals into totals[0..1]).
void launch_scan_pair(MeshScan::Seg s0, MeshScan::Seg s1, unsigned long long *part, size_t part_cap,
                      int64_t *totals, hipStream_t s);

// The PLIC surface's per-cell work arrays: vertex and element counts, their prefix sums.
struct PlicSurfBuffers {
    uint32_t *vcnt, *ecnt;
    int64_t *vpre, *epre;
    unsigned long long *part;
    size_t part_cap;
    int64_t *totals;  // [2] vertices, elements
};
// Over ncell entries, entry c being plane c in cell cells[c] (0-based linear index; cells == nullptr:
// cell c).  emit == false: counts every entry's vertices and elements and their prefix sums (totals
// on the device); emit == true: writes nodes (D T per vertex) and elements (D int64, 1-based).
template <typename T>
void launch_plic_surface(const GridT<T> &g, int dim, const T *normals, const T *intercepts, const int64_t *cells,
                         bool by_cell, size_t ncell, double min_frac, const PlicSurfBuffers &b, bool emit, T *nodes,
                         int64_t *elems, hipStream_t s);

// generate_mesh's filter: bits (one per cell, (ncell + 63) / 64 words) marks each cell whose plane can
// touch it; wrank receives each word's exclusive rank and totals[0] the number of marked cells (part
// holds the scan's block sums, part_cap per sum).  launch_plic_marked then lists the marked cells in
// cell order.
template <typename T>
void launch_plic_touch(const T *normals, const T *intercepts, size_t ncell, int dim, unsigned long long *bits,
                       int64_t *wrank, unsigned long long *part, size_t part_cap, int64_t *totals, hipStream_t s);
void launch_plic_marked(const unsigned long long *bits, const int64_t *wrank, size_t ncell, int64_t *cells,
                        hipStream_t s);

// generate_mesh over a grid or a tree: bits marks the cut cells (corner values not all of one sign;
// n0 > 0 a grid n0 cells wide over the nodal field, n0 == 0 leaves with four values each), wrank
// ranks each word and totals[0] counts them (part: the scan's block sums, part_cap per sum).
// launch_leaf_gather then writes, in leaf order, each cut leaf's index, its four corner values and
// -- when level is not null -- its level and lattice corner (three int32 each).
template <typename T>
void launch_cut_marks(const T *phi, size_t n, size_t n0, unsigned long long *bits, int64_t *wrank,
                      unsigned long long *part, size_t part_cap, int64_t *totals, hipStream_t s);
// The grid's cut cells in cell order: each one's index and its four corner values from the nodal
// field of an n0-wide grid (n cells).
template <typename T>
void launch_grid_gather(const unsigned long long *bits, const int64_t *wrank, size_t n, const T *phi, size_t n0,
                        int64_t *idx, T *phi_out, hipStream_t s);
template <typename T>
void launch_leaf_gather(const unsigned long long *bits, const int64_t *wrank, size_t n, const T *phi4,
                        const int32_t *level, const int32_t *coord, int64_t *idx, T *phi_out, int32_t *lc_out,
                        hipStream_t s);

// The marching-squares contour welded on the device (cc_contour.hip): the cut cells as the gathers
// above leave them, entry r being grid cell idx[r] (a grid n0 cells wide) or the leaf with level and
// lattice corner lc[3r..3r+2] (lc not null), with corner values phi[4r..4r+3].  contour_weld writes
// the reference's lines (two 1-based vertex numbers per segment, *nseg segments) and nodes (two
// Float64 values per vertex, *nvert vertices) to device memory sized for 4 ncut ends; it returns 0,
// -1 when the scratch (contour_weld_bytes) is too small, -2 on a HIP failure.  Blocks on stream s.
template <typename T> struct ContourCells {
    const int64_t *idx;
    const int32_t *lc;
    const T *phi;
    int64_t n0;
    T x0, y0, dx, dy;
};
template <typename T> size_t contour_weld_bytes(int64_t ncut);
template <typename T>
int contour_weld(const ContourCells<T> &in, int64_t ncut, void *scratch, size_t scratch_bytes, int64_t *lines_dev,
                 double *nodes_dev, int64_t *nseg, int64_t *nvert, hipStream_t s);

// Diagnostics over a cache's cells: counts[0..6] = cut, inside, outside, ambiguous, cells with a
// nonzero closure residual, shared faces whose fractions differ, shared faces whose centroids differ;
// counts[7] = the bits of the smallest positive cut volume fraction (1.0 if none); counts[8] = the bits
// of the largest |closure residual| component; counts[9] = cells with a NaN or infinite volume
// fraction or face fraction.  sums[0..3] = total interface area and vector area.
// part holds part_blocks x 4 values; a grid needing more than part_blocks blocks is refused, with
// nothing launched (false).  No grid needs more than diag_max_blocks().
int diag_max_blocks();
template <typename T>
bool launch_diagnostics(const Cols<T> &cols, const GridT<T> &g, int dim, unsigned long long *counts, T *part,
                        size_t part_blocks, T *sums, hipStream_t s);

// The grid cell_indices floors into (Float64, as the reference converts it).
struct MeshGrid {
    double x0[3], d[3];
    int64_t n[3];
};

// Marks, ranks and orders the ncut listed cut cells; leaves the four totals in b.totals.
template <typename T>
void launch_mesh_count(const T *phi, const GridT<T> &g, const uint32_t *cutlist, uint32_t ncut,
                       const MeshBuffers &b, hipStream_t s);
// Writes every vertex (3 T each) and triangle (3 int64, 1-based), into arrays sized from the totals.
template <typename T>
void launch_mesh_emit(const T *phi, const GridT<T> &g, uint32_t ncut, const MeshBuffers &b, T *verts,
                      int64_t *tris, hipStream_t s);
template <typename T>
void launch_cell_indices(const T *verts, const int64_t *tris, int64_t nt, const MeshGrid &g, int64_t *cells,
                         hipStream_t s);

}  // namespace cc

struct cc_mlp;
namespace cc {
// The network at every node of `lat` (n[0] n[1] n[2] values, i fastest) into `out`, on stream s.
int launch_mlp_lattice(const cc_mlp *net, const Lattice &lat, float *out, hipStream_t s);
// The GPU the network lives on.
int mlp_device(const cc_mlp *net);
}  // namespace cc
