/* libcutcell_hip -- the C ABI of an MI355X (gfx950) implementation of CutCellMethods.jl's
 * update_cache! hot path: nodal marching squares (2D), nodal marching cubes (3D) and PLIC.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns a status code and never
 * throws across the ABI.  A failing call leaves a message for cc_last_error().  Device buffers are
 * owned by the library and live until cc_cache_destroy.  Calls on different caches are thread-safe;
 * calls on one cache must be serialised by the caller (as in the reference, which has no locking).
 *
 * Layout (src/cut_cell.jl:112-120, src/cache.jl:101-104): one device column per CutCellData field,
 * column-major over the grid (i fastest), vector fields interleaved per cell:
 *   kind int8[N] | ambiguous uint8[N] | volume_fraction T[N] | centroid T[N][D] |
 *   face_fraction T[N][2D] | face_centroid_local T[N][2D][D-1] | interface_centroid T[N][D]
 * plus phi T[(n0+1)(n1+1)(n2+1)] for the nodal methods, and normals T[N][D], intercepts T[N],
 * is_valid uint8[N] for PLIC.  Directions are 0 = -x, 1 = +x, 2 = -y, ... (the reference's 1..2D).
 * A marching-squares cache on a quadtree's leaves (cc_leaf_cache_create) has N = leaves, in leaf
 * order, and phi T[N][4]: each leaf's four corner values. 
 */
#ifndef CUTCELL_HIP_H
#define CUTCELL_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------------------------- */
#define CC_OK 0
#define CC_ERR_DIMENSION_MISMATCH 1 /* the reference's DimensionMismatch (src/cache.jl:107-121) */
#define CC_ERR_ARGUMENT 2           /* the reference's ArgumentError (src/plic/surface.jl:41-45) */
#define CC_ERR_HIP 3                /* a HIP runtime failure                                    */
#define CC_ERR_UNSUPPORTED 4        /* method/geometry/precision combination not offered         */
#define CC_ERR_OUT_OF_MEMORY 5

/* ---- method tags (src/CutCellMethods.jl:4-5) and element types ------------------------------ */
#define CC_MARCHING_SQUARES 1 /* MarchingSquaresCutCell, 2D  (src/marching_squares/marching_squares.jl:57) */
#define CC_MARCHING_CUBES 2   /* MarchingCubesCutCell, 3D    (src/marching_cubes/marching_cubes.jl:54)    */
#define CC_PLIC 3             /* PLICCutCell, 2D or 3D       (src/plic/plic.jl:30)                        */
#define CC_FLOAT64 0
#define CC_FLOAT32 1

/* ---- cell kinds (src/CutCellMethods.jl:56-58) ------------------------------------------------ */
#define CC_CELL_INSIDE 0
#define CC_CELL_OUTSIDE 1
#define CC_CELL_CUT 2

/* ---- columns --------------------------------------------------------------------------------- */
#define CC_FIELD_KIND 0
#define CC_FIELD_AMBIGUOUS 1
#define CC_FIELD_VOLUME_FRACTION 2
#define CC_FIELD_CENTROID 3
#define CC_FIELD_FACE_FRACTION 4
#define CC_FIELD_FACE_CENTROID_LOCAL 5
#define CC_FIELD_INTERFACE_CENTROID 6
#define CC_FIELD_PHI 7        /* nodal methods (MarchingSquaresCutCellCache.phi, src/marching_squares/cache.jl:32-35) */
#define CC_FIELD_NORMALS 8    /* PLIC (src/plic/cache.jl:48-53) */
#define CC_FIELD_INTERCEPTS 9 /* PLIC */
#define CC_FIELD_IS_VALID 10  /* PLIC */
#define CC_NUM_FIELDS 11

/* ---- geometry (POD form of an SDFLibrary body; the library's own bodies) --------------------- */
#define CC_GEO_EMPTY 0        /* p[0] constant value; EmptyGeo(one(T)) */
#define CC_GEO_SPHERE 1       /* p[0..2] centre, p[3] radius (SDFSphere; SDFCircle in 2D) */
#define CC_GEO_TORUS 2        /* p[0..2] centre, p[3] major R, p[4] minor r, p[5..13] rotation (row-major) */
#define CC_GEO_BOX 3          /* p[0..2] centre, p[3..5] half extents, p[6..14] rotation */
#define CC_GEO_PLANE 4        /* p[0..2] unit normal, p[3..5] point on the plane */
#define CC_GEO_SPHERE_UNION 5 /* p[0] smoothing k; prims = n_prims x (cx, cy, cz, r), host memory */
/* SDFUnion(children...; smoothing) (SDFLibrary, whose use the reference tests at test/test_plic.jl:96-112):
 * p[0] smoothing k >= 0 (0 = the plain minimum); prims = n_prims records of CC_UNION_RECORD doubles in
 * host memory, each the child's kind (CC_GEO_SPHERE, _TORUS, _BOX or _PLANE, as a double) followed by
 * that child's own p[0..14].  The children fold left with the polynomial smooth minimum
 * min(a,b) - h*h*k/4, h = max(k - |a-b|, 0)/k; the normal is the fold's gradient (a blended one scaled
 * to unit length), the winning child's own normal wherever no step blended. */
#define CC_GEO_UNION 6
#define CC_UNION_RECORD 16
/* A flag on any of the bounded kinds above: SDFBounded(child; pad_distance, fill_distance) (SDFLibrary,
 * whose use the reference documents at src/marching_cubes/surface.jl:70-71 and src/volume.jl:26-27).
 * The child's own parameters stay in p[0..15]; p[16..18] / p[19..21] are a box holding the child,
 * p[22] the pad and p[23] the fill (> 0).  Points farther than the pad from the box evaluate to the
 * fill, with no normal; every other point to the child. */
#define CC_GEO_BOUNDED 0x100

typedef struct cc_geometry {
    int32_t kind;
    int32_t n_prims;
    double p[24];
    const double *prims;
} cc_geometry;

/* A CartesianGrid{D}: node i (0-based) of axis a sits at x0[a] + i*d[a] (CartesianMeshes.get_node). */
typedef struct cc_grid {
    int32_t dim;
    int32_t reserved;
    int64_t n[3];
    double x0[3];
    double d[3];
} cc_grid;

typedef struct cc_cache cc_cache;

/* allocate_cache(grid, method; backend) -- src/marching_squares/cache.jl:41-47,
 * src/marching_cubes/cache.jl:30-36, src/plic/cache.jl:60-75.  Allocates every column on the current
 * device and seeds the cache as "no body" with one update against EmptyGeo(1) (src/cache.jl:22-25);
 * PLIC additionally fills the centroid columns with NaN once.  Returns CC_ERR_ARGUMENT for an
 * anisotropic PLIC grid (src/plic/surface.jl:41-45) or a method/dimension mismatch. */
int cc_cache_create(int method, int dtype, const cc_grid *grid, cc_cache **out);

/* Releases every device buffer of the cache (the Julia finalizer's job). */
int cc_cache_destroy(cc_cache *cache);

/* update_cache!(cache, geo, grid) -- src/marching_squares/cache.jl:49-64, src/marching_cubes/cache.jl:38-39,
 * src/plic/cache.jl:77-92.  `grid` must have the cache's n (CC_ERR_DIMENSION_MISMATCH otherwise,
 * src/cache.jl:107-112); its origin may move.  Blocks until the device has finished (src/cache.jl:40). */
int cc_update(cc_cache *cache, const cc_geometry *geo, const cc_grid *grid);

/* update_cache! with a field the caller sampled: the reference's "any callable x -> phi"
 * (src/marching_squares/marching_squares.jl:138-141) for the nodal methods.  `phi` holds one value
 * per node in the cache's element type, column-major; `phi_on_device` != 0 for a device pointer.
 * Checks the node count (src/cache.jl:115-121).  Blocks until done. */
int cc_update_from_phi(cc_cache *cache, const void *phi, int phi_on_device, const cc_grid *grid);

/* The same two updates without the final synchronize, for callers that order work on the cache's
 * stream themselves (benchmarks, graph capture).  Not a reference entry point. */
int cc_update_async(cc_cache *cache, const cc_geometry *geo, const cc_grid *grid);
int cc_synchronize(cc_cache *cache);

/* The cache's own storage (volume_fractions/kinds/face_fractions return storage, src/cache.jl:47-82):
 * device pointer and byte size of one column. */
int cc_column(cc_cache *cache, int field, void **device_ptr, int64_t *nbytes);

/* Copies bytes [offset, offset + nbytes) of one column to host memory (Adapt.adapt(Array, ...)). */
int cc_copy_to_host(cc_cache *cache, int field, int64_t offset, int64_t nbytes, void *dst);

/* Number of cut cells found by the last update (the length of its cut-cell work list). */
int cc_cut_count(cc_cache *cache, int64_t *count);

/* The cache's HIP stream (hipStream_t), for callers that time or order work on it. */
int cc_stream(cc_cache *cache, void **stream);

/* Launch shape of the marching-cubes stream pass, chosen per cache (the autotuner's knob; no
 * counterpart in the reference, whose KernelAbstractions launch is fixed at workgroup 64):
 *   variant 0 = flat tiles of 256 consecutive cells reading phi,
 *           1 = z-march of 256-cell row runs over kz layers reading phi (for a sphere union's
 *               culled or surrogate phi, the sign bitmap its phi pass leaves beside it: the cells of
 *               this pass need only each node's side, and the cut cells read phi),
 *           2 = z-march sampling an analytic body in-kernel (phi written once, never re-read);
 *               for a sphere union with a culled or surrogate phi source it evaluates that phi
 *               in-kernel from the bricks (16 x 16-cell tiles, kz rounded down to a power of two
 *               dividing 16) instead of running the separate brick pass;
 *   nontemporal != 0 streams the column stores past the caches.  For a 3D PLIC cache, variant 0
 *   is the reference's two passes (fit, then faces), 1 the fused single sweep (fit + resolved
 *   faces, kz layers per work item) on 32 x 8-cell tiles and 2 the same sweep on 64 x 8-cell
 *   tiles, one wave per row; nontemporal has no effect there.  Results are identical for every
 *   setting; only the speed differs. */
#define CC_MC_FLAT 0
#define CC_MC_MARCH 1
#define CC_MC_MARCH_FUSED 2
int cc_set_tuning(cc_cache *cache, int variant, int kz, int nontemporal);
int cc_get_tuning(cc_cache *cache, int *variant, int *kz, int *nontemporal);

/* Per-kernel device times (milliseconds) of the last update, in launch order, recorded with HIP
 * events on the cache's stream when profiling is on.  `names` receives static strings. */
int cc_set_profiling(cc_cache *cache, int on);
int cc_kernel_times(cc_cache *cache, int max, float *ms, const char **names, int *count);

/* ---- diagnostics over a cache's cells, reduced on the device ------------------------------------ */
/* cut_cell_report(moments) -- src/cut_cell.jl:302-317: counts by kind, ambiguous cells, and the
 * smallest nonzero volume fraction of a cut cell (1 when there is none). */
typedef struct cc_report {
    int64_t n, cut, inside, outside, ambiguous;
    double min_volume_fraction;
} cc_report;
int cc_cut_cell_report(cc_cache *cache, cc_report *out);

/* The report plus, over every cell of a grid cache: closure_residual (src/cut_cell.jl:340-347) --
 * how many cells have a nonzero one and its largest component magnitude (both zero by construction);
 * the shared faces (interior faces, each once) whose two cells store different bits for the face
 * fraction or the face centroid (zero for a watertight cache; test/test_moments.jl:318-340); and the
 * totals of interface_area and interface_normal_area (src/cut_cell.jl:190-191, :224-226).  Counts and extremes
 * are exact; the totals are summed in a fixed order (the same on every run). */
typedef struct cc_diagnostics {
    cc_report report;
    int64_t closure_nonzero;
    double closure_max;
    int64_t face_mismatches, centroid_mismatches;
    double interface_area;
    double interface_normal_area[3];
    int64_t nonfinite; /* cells whose volume fraction or a face fraction is NaN or infinite */
} cc_diagnostics;
int cc_cache_diagnostics(cc_cache *cache, const cc_grid *grid, cc_diagnostics *out);

/* ---- files ------------------------------------------------------------------------------------- */
/* For ParaView: `prefix_cells.vti` (VTK XML ImageData over the grid: volume_fraction, kind, ambiguous,
 * face_fraction_-x ... +z, and centroid / interface_centroid with phi as point data, or for PLIC
 * normal / intercept / is_valid) and, with surface != 0, `prefix_surface.vtu` (the generate_mesh
 * result), as write_cache_vtk writes them for the reference's triangle-clipping cache
 * (src/tri_clipping/surface.jl:214-252).  Raw little-endian appended data.  Not for leaf caches. */
int cc_write_cache_vtk(cc_cache *cache, const cc_grid *grid, const char *prefix, int surface);

/* One column as a .npy file in the reference's layout: fortran_order, components first, then the
 * cells (or nodes, for phi): e.g. face_fraction of an n0 x n1 x n2 cache is (6, n0, n1, n2) float64. */
int cc_write_npy(cc_cache *cache, int field, const char *path);

/* generate_mesh(cache, grid) for a marching-squares cache -- src/marching_squares/surface.jl:59-86:
 * the contour of the cache's phi, welded by exact coordinate equality.  Two passes: call with
 * nodes/lines NULL to get the counts, then again with buffers of that size; the size call draws the
 * contour (the cut cells gathered on the device, contoured on the host) and the fill call right after
 * it copies that contour out unless an update or another grid frame came between.  nodes: 2*nnodes
 * doubles; lines: 2*nlines int64 (1-based node ids, in order of first appearance, as MeshLibrary
 * stores them). */
int cc_generate_mesh_ms(cc_cache *cache, const cc_grid *grid, double *nodes, int64_t *nnodes,
                        int64_t *lines, int64_t *nlines);

/* ---- marching squares on a quadtree (CartesianMeshes.AdaptiveMesh{2}) ----------------------------
 * The tree crosses the ABI as its leaf list, in the tree's leaf order: leaf c has refinement level
 * level[c] (0 = the base grid's cells) and its lower corner at integer lattice coordinates
 * coord[2c], coord[2c+1] of its level, so its corners are x0 + (coord + bit) * d / 2^level
 * (cell_nodes(mesh, c), src/marching_squares/marching_squares.jl:113-121). */
typedef struct cc_leaves {
    int64_t n;             /* leaves */
    const int32_t *level;  /* [n] */
    const int32_t *coord;  /* [n][2] */
    double x0[2];          /* the base grid's origin */
    double d[2];           /* the base grid's cell size */
    int32_t on_device;     /* 1: level and coord are device pointers (the tree adapted to the device) */
    int32_t version;       /* host arrays: 0, or a tag the caller changes whenever the leaf arrays
                              change; with the same nonzero tag, arrays and n as the cache's last update
                              the library keeps its device copy instead of copying the leaves again */
} cc_leaves;

/* Where generate_mesh welds a marching-squares contour (grid or tree): on the device from 4096 cut
 * cells (CC_WELD_AUTO), always on the host, or always on the device.  No counterpart in the
 * reference, whose weld is a serial host Dict; every mode gives the same nodes and lines. */
#define CC_WELD_AUTO 0
#define CC_WELD_HOST 1
#define CC_WELD_DEVICE 2
int cc_set_contour_weld(cc_cache *cache, int mode);

/* allocate_cache(mesh::AdaptiveMesh{2}, MarchingSquaresCutCell()) -- src/marching_squares/cache.jl:100-106:
 * a marching-squares cache laid out on the leaves (one CutCellData record per leaf, in leaf order;
 * the phi column holds each leaf's four corner values in MS_NODE_BITS order), seeded as no body. */
int cc_leaf_cache_create(int dtype, const cc_leaves *tree, cc_cache **out);

/* update_cache!(cache, geo, mesh::AdaptiveMesh{2}) -- src/marching_squares/cache.jl:115-137: every
 * leaf's corners sampled from geo and its moments built from exactly those values, one thread per
 * leaf; blocks until done.  CC_ERR_DIMENSION_MISMATCH when the tree has another leaf count
 * (_check_leaf_size, :140-147); CC_ERR_UNSUPPORTED on a grid cache. */
int cc_update_leaves(cc_cache *cache, const cc_geometry *geo, const cc_leaves *tree);

/* The same from corner values the caller sampled (cell_values(field, mesh, c) with a callable field,
 * src/marching_squares/marching_squares.jl:152-155): phi4 holds four values per leaf in
 * MS_NODE_BITS order, in the cache's element type, in host memory or (on_device != 0) device memory. */
int cc_update_leaves_from_phi(cc_cache *cache, const void *phi4, int on_device, const cc_leaves *tree);

/* generate_mesh(cache, mesh::AdaptiveMesh{2}) -- src/marching_squares/surface.jl:92-110: the contour
 * of each cut leaf's stored corner values, in leaf order, welded by exact coordinate equality; the
 * same two-call protocol and buffers as cc_generate_mesh_ms (the fill call reuses the size call's
 * contour only for a tree with a nonzero version tag, the same arrays and count). */
int cc_generate_mesh_leaves(cc_cache *cache, const cc_leaves *tree, double *nodes, int64_t *nnodes,
                            int64_t *lines, int64_t *nlines);

/* generate_mesh(cache::PLICCutCellCache, grid; min_tri_area_fraction) -- src/plic/surface.jl:28-267:
 * every cell's stored plane clipped to its cell on the device, in cell order: 2D segments (two nodes
 * each, never shared) or 3D triangles (each crossed cell's polygon ordered about its centroid and
 * fanned from its first vertex).  Elements shorter (2D) or smaller (3D) than min_frac of a cell
 * edge / face are dropped.  nodes: D values per node in the cache's element type; elems: D int64 per
 * element, 1-based.  Two-call protocol as cc_generate_mesh_mc; the surface stays in the cache
 * (cc_mesh_device) until the next update. */
int cc_generate_mesh_plic(cc_cache *cache, const cc_grid *grid, double min_frac, void *nodes, int64_t *nnodes,
                          int64_t *elems, int64_t *nelems);

/* extract_surface_plic(normals, intercepts, cell_indices, grid; min_tri_area_fraction) --
 * src/plic/surface.jl:44-101, :126-182: the same surface for planes the caller holds: entry q is the
 * plane (normals[D q .. D q + D - 1], intercepts[q]) in the unit-cell frame of cell cells[q] (0-based
 * linear index).  Host arrays in; nodes / elems as cc_generate_mesh_plic, entries in the given order.
 * Two-call protocol (counts with nodes/elems NULL; every call recomputes). */
int cc_extract_surface_plic(int dtype, const cc_grid *grid, int64_t count, const int64_t *cells, const void *normals,
                            const void *intercepts, double min_frac, void *nodes, int64_t *nnodes, int64_t *elems,
                            int64_t *nelems);

/* calc_volume(geo, grid) -- src/volume.jl:29-40: the measure of the region where geo's per-cell
 * planes (cell_plane, the PLIC fit) are solid, over the grid's own bounds, in the element type dtype:
 * the sum of every cell's solid fraction times the cell volume.  CC_ERR_ARGUMENT for an anisotropic
 * grid.  The cells' fractions are summed on the device in a fixed order (the same result on every
 * run); the reference's serial left fold can differ from it in the last bits. */
int cc_calc_volume(int dtype, const cc_geometry *geo, const cc_grid *grid, double *volume);

/* cc_calc_volume without allocating device memory.  No counterpart in the reference: this entry point
 * is for callers that cannot call hipMalloc (code under graph capture, tight serving loops, harnesses
 * that forbid allocation).  It makes no hipMalloc/hipFree and creates no stream: its block sums, its
 * result and the body's primitive table live in a 256 KiB device array that each GPU holds for the
 * life of the process (at most 30 688 doubles of primitive table: 7672 spheres of a sphere union or
 * 1918 children of a union; a larger table is CC_ERR_ARGUMENT).  It runs on `stream` (a hipStream_t;
 * NULL for the null stream) and blocks until the volume is on the host.  Every call on one GPU shares
 * that array: calls that overlap on one GPU race, so the caller serialises them. */
int cc_calc_volume_noalloc(int dtype, const cc_geometry *geo, const cc_grid *grid, void *stream, double *volume);

/* generate_mesh(cache, grid) for a marching-cubes cache -- src/marching_cubes/surface.jl:82-112:
 * the cache's phi marched on the device (Lewiner cases, shared vertices, wound so each triangle's
 * normal points out of the body) into a mesh the cache keeps until its next update.  Call with
 * nodes/tris NULL to march and get the counts, then with buffers of that size to copy: nodes is
 * 3*nnodes values of the cache's element type (double or float); tris is 3*ntris int64, 1-based
 * node ids (MeshLibrary's Tri connectivity).  A call after an update, or with a moved grid, marches
 * again.  An empty surface is not an error (counts 0); the reference's warning is the caller's. */
int cc_generate_mesh_mc(cc_cache *cache, const cc_grid *grid, void *nodes, int64_t *nnodes, int64_t *tris,
                        int64_t *ntris);

/* The last marched surface in the cache's device memory (valid until the next march or destroy),
 * and a serial number that increases with every march.  No counterpart in the reference, whose
 * march is host-only; for device consumers of the mesh. */
int cc_mesh_device(cc_cache *cache, const void **nodes, const int64_t **tris, int64_t *nnodes,
                   int64_t *ntris, int64_t *serial);

/* cell_indices(mesh, grid) -- src/marching_cubes/surface.jl:128-141 -- for the cache's last marched
 * surface: for each triangle the 1-based linear index of the cell its centroid floors into, clamped
 * to the grid.  cells: ntris int64, host memory, or device memory when on_device != 0. */
int cc_mesh_cell_indices(cc_cache *cache, const cc_grid *grid, int64_t *cells, int on_device);

/* ---- batched bodies (no counterpart in the reference, which updates one cache per call) --------
 * A batch references B existing 3D caches (marching cubes and PLIC mixed, one element type) and
 * updates all of them with ONE hipGraph launch per step: a counter-reset kernel, one stream/fit pass
 * over every body's work items, one cut-cell/face pass.  Bodies must be analytic (sphere, torus,
 * box, plane).  Each step's grids (origins may move) and bodies are read from `grids[b]`/`geos[b]`;
 * every body's results land in its own cache's columns, identical to cc_update on that cache.
 * The caches must outlive the batch and must not be updated on their own while it runs.  A step
 * waits for any cc_update_async still running on a member cache; other work a caller queues on a
 * member's stream (cc_stream) must be synchronised by the caller before the step. */
typedef struct cc_batch cc_batch;
int cc_batch_create(int nbodies, cc_cache *const *caches, cc_batch **out);
int cc_batch_destroy(cc_batch *batch);
/* One step: every body updated; blocks.  `device_ms` (optional) receives the graph's device time. */
int cc_batch_step(cc_batch *batch, const cc_geometry *geos, const cc_grid *grids, float *device_ms);
/* Tile depth (z layers per work item) of one marching-cubes body: the autotuner's per-body knob. */
int cc_batch_set_kz(cc_batch *batch, int body, int kz);
/* Work items of the two passes of the captured graph. */
int cc_batch_work_items(cc_batch *batch, int64_t *pass_a, int64_t *pass_b);

/* ---- the learned surrogate (no counterpart in the reference, whose bodies are SDFLibrary
 * closures evaluated exactly at every node) ---------------------------------------------------------
 * A Fourier-feature MLP for an expensive body's signed distance: 3 -> 2*nfreq features
 * [sin(2 pi B x), cos(2 pi B x)] (B: nfreq x 3, N(0, sigma^2)) -> depth x width SiLU -> 1.  FP32
 * master parameters; forward and backward run in BF16 on the matrix cores with FP32 accumulation.
 * Parameters are packed W_0 (2 nfreq x width, row-major), W_1..W_{depth-1} (width x width), b_0..,
 * w_out (width), b_out.  width must be 256, depth 1..8, nfreq 32/64/96/128. */
typedef struct cc_mlp cc_mlp;
int cc_mlp_create(int nfreq, int width, int depth, float sigma, uint64_t seed, cc_mlp **out);
int cc_mlp_destroy(cc_mlp *net);
/* Parameter count and 3*nfreq (the size of B). */
int cc_mlp_num_params(cc_mlp *net, int64_t *nparams, int64_t *nfreq3);
/* Copies the packed parameters and B to host memory (either pointer may be NULL). */
int cc_mlp_get_params(cc_mlp *net, float *params, float *ffB);
int cc_mlp_set_params(cc_mlp *net, const float *params);
/* The network at n host points (x: n x 3) into y (host, n floats). */
int cc_mlp_eval(cc_mlp *net, const float *x, int64_t n, float *y);
/* Mean squared error over n host samples (n a multiple of 64) and its gradient (nparams floats). */
int cc_mlp_loss_grad(cc_mlp *net, const float *x, const float *y, int64_t n, float *loss, float *grad);
/* Fits the network to exact samples of `geo` (half uniform in [lo, hi]^3, half within `band` of
 * the surface), Adam with a cosine-decayed rate; `history` (optional) gets each epoch's mean loss.
 * From then on the features see u = (x - lo) / (hi - lo) (the training box as the unit box, where
 * sigma is quoted); a network never trained sees u = x. */
int cc_mlp_train(cc_mlp *net, const cc_geometry *geo, int64_t nsamples, float lo, float hi, float band,
                 int epochs, int batch, float lr, uint64_t seed, float *history);

/* How cc_update produces a sphere union's nodal phi in a marching-cubes cache (sample_nodes,
 * src/marching_cubes/cache.jl:33-40, folds the whole body at every node):
 *   CC_PHI_NAIVE     -- every node folds every primitive (the plain exact path; the default);
 *   CC_PHI_CULLED    -- each 16^3-node brick folds only the primitives a bound over its box
 *                       (extended by one node on its upper sides) cannot prove inert: bitwise
 *                       the same phi;
 *   CC_PHI_SURROGATE -- the network `net` on a lattice of spacing `coarse` cells, interpolated;
 *                       exact (culled) wherever a Lipschitz sign certificate is missing, where
 *                       |phi_nn| < band_cells * h, or where the network's sign contradicts the
 *                       certificate.  kind and every moment are bitwise the exact path's; phi
 *                       differs only at nodes of certified sub-bricks, none of which is a corner
 *                       of a cut cell.
 * Other bodies always take the plain path.  The network must outlive its use by the cache. */
#define CC_PHI_NAIVE 0
#define CC_PHI_CULLED 1
#define CC_PHI_SURROGATE 2
int cc_set_phi_source(cc_cache *cache, int source, cc_mlp *net, int coarse, double band_cells);
/* Counters of the last culled/surrogate update (blocks until it is done), CC_PHI_STATS entries:
 * nodes, exact node evaluations, certificate evaluations, sum over bricks of kept primitives,
 * nodes in uncertified sub-bricks, nodes under the band, nodes with a contradicted sign, bricks. */
#define CC_PHI_STATS 8
int cc_phi_stats(cc_cache *cache, int64_t *stats);

/* Measurement aid (no counterpart in the reference): the best time over `iters` of a write-only
 * fill of `nbytes` and of a copy of `nbytes` (read + write) on the current device, with 16-byte
 * lane-contiguous accesses -- the achievable-bandwidth yardstick the cut-cell passes are read
 * against, and the plain copy kernel that scales rocprof's FETCH_SIZE/WRITE_SIZE counters. */
int cc_probe_bandwidth(int64_t nbytes, int iters, float *fill_ms, float *copy_ms);
/* The same fill kernel over one of a cache's own columns (its first nbytes/16*16 bytes), best time of
 * `iters`: the write rate that allocation reaches, the yardstick for the passes writing it, since
 * where an allocation lands moves it.  Overwrites the column: update the cache before reading it. */
int cc_probe_fill_column(cc_cache *cache, int field, int iters, float *ms);
/* Write-pattern study for the same yardstick: best time of a fill of `nbytes` with `blocks`
 * workgroups of 256 and store pattern `variant` (0 grid-stride, 1 non-temporal, 2 one contiguous
 * 64 KiB run per workgroup iteration, 3 four stores in flight per lane). */
int cc_probe_write(int64_t nbytes, int iters, int variant, int blocks, float *ms);

/* The last error message of this thread ("" if none). */
const char *cc_last_error(void);

/* Library version string. */
const char *cc_version(void);

#ifdef __cplusplus
}
#endif
#endif
