/* CPU oracle for the cut-cell update hot path -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C restatement of CutCellMethods.jl's Float64 `update_cache!` for the nodal methods
 * (MarchingSquaresCutCell, MarchingCubesCutCell) and PLICCutCell, expression by expression, so that
 * the GPU path can be compared against it bit for bit.  Only tests, the smoke check and the
 * benchmark's CPU-baseline leg may load it.
 *
 * Conventions (all 0-based here; the reference is 1-based):
 *   - grid: x0[D], d[D], n[D] cells; nodes are n+1 per axis; node (i,j,k) sits at x0 + i*d per axis
 *     (CartesianMeshes.get_node, pinned by test/test_moments.jl:604-620).
 *   - every array is column-major over the grid (i fastest), one column per CutCellData field
 *     (src/cut_cell.jl:112-120, src/cache.jl:101-104); vector fields are interleaved per cell.
 *
 * Pinning: tests/test_oracle_{ms,mc,plic,geometry,mesh}.py check every function here against the
 * reference tests' golden values (hand cells, the flush cube, the PLIC count 84, the documented
 * sphere area of the marched surface), closed forms, invariants (closure, watertightness, the 256
 * sign patterns, the marched surface closing the face clips' apertures) and brute force.  Parity
 * unpinned, because no reference output exists to compare with: the exact bits of SDFLibrary's
 * bodies (the geometry here is the library's own, pinned by mathematics), of CartesianMeshes'
 * get_volume_fraction (orc_volume_fraction: pinned to tolerance against exact rationals), of
 * get_elem_centroid (restated as x0 + (i + 0.5) d), and MarchingCubes.jl's own vertex numbering and
 * index-to-world bits for the whole-grid surface (orc_mc_mesh: pinned by closure to the ulp-level
 * agreement the reference states, by topology and by the documented area).
 */
#ifndef CUTCELL_ORACLE_H
#define CUTCELL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One column pointer per CutCellData field (src/cut_cell.jl:112-120). */
typedef struct orc_cols {
    int8_t  *kind;                /* [N]                      */
    uint8_t *ambiguous;           /* [N]                      */
    double  *volume_fraction;     /* [N]                      */
    double  *centroid;            /* [N][D]                   */
    double  *face_fraction;       /* [N][2D]                  */
    double  *face_centroid_local; /* [N][2D][D-1]             */
    double  *interface_centroid;  /* [N][D]                   */
} orc_cols;

/* Geometry descriptor kinds (the build's own geometry; SDFLibrary is un-vendored). */
enum {
    ORC_GEO_EMPTY = 0,        /* p[0] = constant value (EmptyGeo(one(T)) seeds with 1)          */
    ORC_GEO_SPHERE = 1,       /* p[0..2] centre, p[3] radius (2D: p[0..1] centre, p[3] radius)   */
    ORC_GEO_TORUS = 2,        /* p[0..2] centre, p[3] R, p[4] r, p[5..13] rotation (row-major)   */
    ORC_GEO_BOX = 3,          /* p[0..2] centre, p[3..5] half extents, p[6..14] rotation        */
    ORC_GEO_PLANE = 4,        /* p[0..2] unit normal, p[3..5] point on the plane                */
    ORC_GEO_SPHERE_UNION = 5, /* p[0] smoothing k; prims = n_prims x (cx,cy,cz,r)               */
    ORC_GEO_UNION = 6,        /* p[0] smoothing k >= 0; prims = n_prims x ORC_UNION_RECORD: the
                                 child's kind (sphere, torus, box, plane), then its p[0..14]      */
    ORC_UNION_RECORD = 16,
    ORC_GEO_BOUNDED = 0x100   /* flag on a kind: SDFBounded(child; pad, fill) -- the child's own
                                 parameters in p[0..15], its bounding box p[16..18] (low corner),
                                 p[19..21] (high), p[22] pad, p[23] fill                          */
};

typedef struct orc_geometry {
    int32_t kind;
    int32_t n_prims;
    double p[24];
    const double *prims;
} orc_geometry;

/* ---- geometry ---------------------------------------------------------------------------- */
double orc_sdf_value(const orc_geometry *g, int dim, const double *x);
/* (distance, normal); returns the distance, writes dim normal components. */
double orc_get_sdf(const orc_geometry *g, int dim, const double *x, double *normal);

/* ---- nodal sampling: src/marching_squares/cache.jl:67-70 ----------------------------------- */
void orc_sample_nodes(const orc_geometry *g, int dim, const double *x0, const double *d,
                      const int64_t *n, double *phi);

/* ---- marching squares: src/marching_squares/moments.jl:66-163 (one whole grid) ------------- */
void orc_ms_cells(const double *phi, const double *x0, const double *d, const int64_t *n,
                  const orc_cols *out);

/* ---- marching squares on a quadtree's leaves: src/marching_squares/cache.jl:129-137 ----------
 * Leaf c: level[c], lower corner coord[2c..2c+1] on its level's lattice (spacing d / 2^level from
 * the base origin x0).  phi4: the four corner values per leaf; out: one entry per leaf. */
/* generate_mesh over a 2D grid (src/marching_squares/surface.jl:59-86, 123-177): the contour of phi
 * with exactly welded vertices; counts only when nodes/lines are NULL.  0, or -1 out of memory. */
int orc_ms_contour(const double *phi, const double *x0, const double *d, const int64_t *n,
                   int64_t *nnodes, int64_t *nlines, double *nodes, int64_t *lines);
void orc_ms_leaves(const orc_geometry *g, const double *x0, const double *d, int64_t n,
                   const int32_t *level, const int32_t *coord, double *phi4, const orc_cols *out);

/* ---- marching cubes: src/marching_cubes/moments.jl:275-346, cells with k in [k0, k1) -------
 * `out` columns are indexed relative to cell (0,0,k0), so a slab can be checked on its own. */
void orc_mc_cells(const double *phi, const double *x0, const double *d, const int64_t *n,
                  int64_t k0, int64_t k1, const orc_cols *out);

/* Per-cell pieces exported for the non-circular tests. */
int  orc_mc_cell_case(const double *phi_raw8);                     /* 0-based Lewiner case   */
int  orc_mc_cell_triangles(const double *phi_raw8, int8_t *codes36); /* n triangles, 0..12 codes */
void orc_mc_edge_crossings(const double *local_nodes24, const double *phi_raw8, double *pts39);
void orc_mc_cell(const double *nodes24, const double *phi_raw8, const double *cellsize3,
                 int8_t *kind, uint8_t *amb, double *vf, double *centroid3, double *ff6,
                 double *fcl12, double *ic3);

/* ---- the whole-grid marching-cubes surface: src/marching_cubes/surface.jl:82-112 ------------
 * Vertices (3 per node, column-major node order, then interior vertices in cube order) and
 * 1-based triangles in cube order; owner = each triangle's 0-based cell.  nodes == NULL counts.
 * Returns 0, or -1 when the scratch allocation fails. */
int  orc_mc_mesh(const double *phi, const double *x0, const double *d, const int64_t *n,
                 int64_t *nverts, int64_t *ntris, double *nodes, int64_t *tris, int64_t *owner);

/* ---- PLIC: src/plic/cache.jl:77-162, src/plic/plic.jl:72-91, src/plic/apertures.jl --------- */
double orc_volume_fraction(int dim, const double *normal, double a);
void orc_plic_fit(const orc_geometry *g, int dim, const double *x0, const double *d,
                  const int64_t *n, double *vf, int8_t *kind, double *normals,
                  double *intercepts, uint8_t *is_valid);
void orc_plic_faces(int dim, const double *d, const int64_t *n, const double *normals,
                    const double *intercepts, const uint8_t *is_valid, const double *vf,
                    double *face_fraction);

/* generate_mesh(cache::PLICCutCellCache, grid) (src/plic/surface.jl:28-267): every cell's plane
 * clipped to its cell -- 2D segments (2 nodes each) or 3D fan triangles of the ordered polygon --
 * in cell order; elements are 1-based node ids, D per element.  nodes == NULL counts. */
int orc_plic_surface(int dim, const double *x0, const double *d, const int64_t *n, const double *normals,
                     const double *intercepts, double min_frac, int64_t *nverts, int64_t *nelems,
                     double *nodes, int64_t *elems);

/* calc_volume(geo, grid) (src/volume.jl:29-40): the measure where the body's planes are solid. */
double orc_calc_volume(const orc_geometry *g, int dim, const double *x0, const double *d, const int64_t *n);

/* ---- whole-grid drivers (OpenMP over k) ----------------------------------------------------- */
int  orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
