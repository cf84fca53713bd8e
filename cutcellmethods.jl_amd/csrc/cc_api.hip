// The C ABI of libcutcell_hip: cache lifetime, the update entry points, column access and the
// marching-squares contour.  Mirrors the reference's cache API (src/cache.jl:15-121,
// src/marching_squares/cache.jl, src/marching_cubes/cache.jl, src/plic/cache.jl) with status codes
// in place of exceptions.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "cc_internal.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPC(expr)                                                                            \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(CC_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));           \
    } while (0)

template <typename T> __global__ void k_fill(T *p, size_t n, T v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

constexpr int MAX_TIMED = 8;

// Calibration kernels for the bandwidth probe: a write-only fill and a read+write copy, both one-shot
// -- one 16-byte element per thread, a workgroup per 4 KiB, as many workgroups as the buffer needs.
// Of the write patterns measured on MI355X this one reached the highest rate (6.9 TB/s on 2 and
// 12 GiB buffers, where grid-stride loops of 2048-32768 workgroups reached 4.3-5.9 TB/s and two or
// four stores per thread 3.7-6.2 TB/s), so the probe's rate is the yardstick a write stream is
// read against.
typedef double __attribute__((ext_vector_type(2))) f64x2;
// Buffers beyond 32 GiB (2^31 elements) take a capped grid and loop, which keeps a launch's
// work-items under 2^32.
constexpr size_t PROBE_MAX_BLOCKS = size_t(1) << 23;
__global__ __launch_bounds__(256) void k_probe_fill(f64x2 *dst, size_t n, double v) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = f64x2{v, v};
}
__global__ __launch_bounds__(256) void k_probe_copy(f64x2 *dst, const f64x2 *src, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
// Workgroups of the probe kernels for n 16-byte elements: one element per thread up to the cap.
inline unsigned probe_blocks(size_t n) { return (unsigned)std::min((n + 255) / 256, PROBE_MAX_BLOCKS); }

// Write-pattern variants of the fill probe: 0 grid-stride, 1 grid-stride non-temporal,
// 2 one contiguous 64 KiB run per workgroup iteration, 3 grid-stride with 4 stores in flight.
__global__ __launch_bounds__(256) void k_probe_fill_v(f64x2 *dst, size_t n, double v, int variant) {
    const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const f64x2 val = f64x2{v, v};
    if (variant == 1) {
        for (size_t i = tid; i < n; i += stride) __builtin_nontemporal_store(val, dst + i);
    } else if (variant == 2) {
        const size_t run = 4096;  // 64 KiB of 16-byte elements
        for (size_t b = blockIdx.x * run; b < n; b += (size_t)gridDim.x * run)
            for (size_t i = b + threadIdx.x; i < b + run && i < n; i += blockDim.x) dst[i] = val;
    } else if (variant == 3) {
        size_t i = tid;
        for (; i + 3 * stride < n; i += 4 * stride) {
            dst[i] = val;
            dst[i + stride] = val;
            dst[i + 2 * stride] = val;
            dst[i + 3 * stride] = val;
        }
        for (; i < n; i += stride) dst[i] = val;
    } else {
        for (size_t i = tid; i < n; i += stride) dst[i] = val;
    }
}

}  // namespace

struct cc_cache {
    int method = 0, dtype = 0, dim = 0, device = 0;
    int64_t n[3] = {1, 1, 1};
    int64_t ncell = 0, nnode = 0;
    void *col[CC_NUM_FIELDS] = {};
    int64_t colbytes[CC_NUM_FIELDS] = {};
    uint32_t *cutlist = nullptr;
    uint32_t *counter = nullptr;
    double *prims_dev = nullptr;
    int64_t prims_cap = 0;
    float4 *scr_dev = nullptr;  // a sphere union's screen records for the plain exact fold
    int64_t scr_cap = 0;
    std::vector<double> prims_host;
    hipStream_t stream = nullptr;
    bool async_pending = false;    // cc_update_async left work on `stream` not yet waited for
    hipEvent_t order_ev = nullptr;  // orders another stream (a batch's) after that work
    bool profiling = false;
    hipEvent_t ev[MAX_TIMED + 1] = {};
    const char *ev_name[MAX_TIMED] = {};
    int nev = 0;
    bool ev_ready = false;
    cc::McTuning tune;
    // how a sphere union's nodal phi is produced (marching cubes)
    int phi_source = CC_PHI_NAIVE;
    cc_mlp *net = nullptr;
    int coarse = 8;
    double band_cells = 3.0;
    uint16_t *cull_list = nullptr;
    int64_t cull_cap = 0;  // entries
    uint32_t *cull_cnt = nullptr;
    int64_t cnt_cap = 0;
    float *nn = nullptr;
    int64_t nn_cap = 0;
    uint32_t *bstat = nullptr;  // per-brick counters of the last culled/surrogate update (then, when
                                // the stream pass evaluated phi, its per-work-item counters)
    int64_t bstat_cap = 0;
    int8_t *cert = nullptr;     // sub-brick sign certificates for the fused surrogate pass
    int64_t cert_cap = 0;
    uint64_t *signs = nullptr;  // the phi pass's brick-major sign bitmap, for the stream pass reading it
    int64_t signs_cap = 0;
    int64_t phi_nbricks = 0;
    int64_t phi_nwork = 0;  // work items of the stream pass that evaluated phi (their counters follow)
    // the whole-grid surface (marching cubes): work arrays, the last mesh, and what it was made from
    uint64_t gen = 0;  // updates so far
    cc::MeshBuffers mb = {};
    std::vector<void *> mb_allocs;
    int64_t *mesh_totals_host = nullptr;  // pinned, 3 values
    void *mverts = nullptr;
    int64_t mverts_cap = 0;  // bytes
    int64_t *mtris = nullptr;
    int64_t mtris_cap = 0;  // int64 entries
    int64_t *mcells = nullptr;
    int64_t mcells_cap = 0;
    int64_t mesh_nv = -1, mesh_nt = -1, mesh_serial = 0;
    uint64_t mesh_gen = 0;
    double mesh_x0[3] = {}, mesh_d[3] = {};
    double mesh_min_frac = -1.0;  // the PLIC surface's element floor
    cc::PlicSurfBuffers ps = {};
    std::vector<void *> ps_allocs;
    int64_t ps_cap = 0;  // entries the PLIC surface's work arrays hold
    // generate_mesh's cell marks (the PLIC cells a plane can touch, the cut leaves of a tree): one
    // bit per cell, each word's rank, the marked cells in order
    unsigned long long *mark_bits = nullptr;
    int64_t mark_bits_cap = 0;
    int64_t *mark_rank = nullptr;
    int64_t mark_rank_cap = 0;
    int64_t *mark_cells = nullptr;
    int64_t mark_cells_cap = 0;
    // diagnostics: counters, block sums, totals (device), and a pinned copy of the results
    void *dg = nullptr;
    void *dg_host = nullptr;
    // marching squares laid out on a quadtree's leaves (ncell = leaves, phi = 4 values per leaf)
    bool leaves = false;
    int32_t *leaf_level = nullptr, *leaf_coord = nullptr;  // device copies of a host leaf list
    const void *leaf_src[2] = {};  // the host arrays and version tag they were copied from
    int32_t leaf_version = 0;
    // the last contour (grid or tree), for the fill call of generate_mesh's two calls: valid for
    // this update (gen) and this grid frame, or this tagged tree (arrays, count, version, base grid)
    std::vector<double> contour_nodes;
    std::vector<int64_t> contour_lines;
    bool contour_dev = false;  // the kept contour is the device weld's (weld_nodes / weld_lines)
    int64_t contour_dev_nn = 0, contour_dev_nl = 0;
    bool contour_ok = false;
    uint64_t contour_gen = 0;
    const void *contour_src[2] = {};
    int64_t contour_n = -1;
    int32_t contour_version = 0;
    double contour_frame[4] = {};
    int64_t *mark_idx = nullptr;  // the cut leaves gathered for the tree contour
    int64_t mark_idx_cap = 0;
    char *mark_phi = nullptr;
    int64_t mark_phi_cap = 0;
    int32_t *mark_lc = nullptr;
    int64_t mark_lc_cap = 0;
    // where the contour is welded (CC_WELD_*), and the device weld's scratch and outputs
    int weld_mode = CC_WELD_AUTO;
    char *weld_scratch = nullptr;
    int64_t weld_scratch_cap = 0;
    int64_t *weld_lines = nullptr;
    int64_t weld_lines_cap = 0;
    double *weld_nodes = nullptr;
    int64_t weld_nodes_cap = 0;
};

namespace {

size_t elem_size(int dtype) { return dtype == CC_FLOAT32 ? 4 : 8; }

int check_grid(const cc_cache *c, const cc_grid *g) {
    if (c->leaves)
        return fail(CC_ERR_UNSUPPORTED, "this cache is laid out on a tree's leaves: update and contour "
                                        "it over the tree (cc_update_leaves, cc_generate_mesh_leaves)");
    if (!g) return fail(CC_ERR_ARGUMENT, "grid is NULL");
    if (g->dim != c->dim)
        return fail(CC_ERR_DIMENSION_MISMATCH, "the cache is %dD but the grid is %dD", c->dim, g->dim);
    for (int a = 0; a < c->dim; ++a)
        if (g->n[a] != c->n[a])
            return fail(CC_ERR_DIMENSION_MISMATCH,
                        "the cache holds (%lld, %lld, %lld) cells, but `grid` has (%lld, %lld, %lld); a "
                        "cache is allocated for one grid's `n`, and only its origin may move between "
                        "updates",
                        (long long)c->n[0], (long long)c->n[1], (long long)(c->dim == 3 ? c->n[2] : 1),
                        (long long)g->n[0], (long long)g->n[1],
                        (long long)(g->dim == 3 ? g->n[2] : 1));
    for (int a = 0; a < c->dim; ++a)
        if (!(g->d[a] > 0.0) || !std::isfinite(g->d[a]) || !std::isfinite(g->x0[a]))
            return fail(CC_ERR_ARGUMENT, "grid.d must be positive and finite, x0 finite");
    if (c->method == CC_PLIC) {
        for (int a = 1; a < c->dim; ++a)
            if (g->d[a] != g->d[0])
                return fail(CC_ERR_ARGUMENT,
                            "PLICCutCell needs an isotropic grid (equal cell size on every axis), got "
                            "grid.d = (%g, %g%s)",
                            g->d[0], g->d[1], c->dim == 3 ? ", ..." : "");
    }
    return CC_OK;
}

template <typename T> cc::GridT<T> grid_t(const cc_grid *g, int dim) {
    cc::GridT<T> r;
    for (int a = 0; a < 3; ++a) {
        r.x0[a] = a < dim ? T(g->x0[a]) : T(0);
        r.d[a] = a < dim ? T(g->d[a]) : T(1);
        r.n[a] = a < dim ? (int)g->n[a] : 1;
    }
    return r;
}

template <typename T> cc::Cols<T> cols_t(const cc_cache *c) {
    cc::Cols<T> r;
    r.kind = (int8_t *)c->col[CC_FIELD_KIND];
    r.amb = (uint8_t *)c->col[CC_FIELD_AMBIGUOUS];
    r.vf = (T *)c->col[CC_FIELD_VOLUME_FRACTION];
    r.centroid = (T *)c->col[CC_FIELD_CENTROID];
    r.ff = (T *)c->col[CC_FIELD_FACE_FRACTION];
    r.fcl = (T *)c->col[CC_FIELD_FACE_CENTROID_LOCAL];
    r.ic = (T *)c->col[CC_FIELD_INTERFACE_CENTROID];
    return r;
}

// A geometry descriptor's kind and parameters, checked on the host before any launch.
int check_geometry(const cc_geometry *geo, int dim) {
    if (!geo) return fail(CC_ERR_ARGUMENT, "geometry is NULL");
    const int base = geo->kind & ~CC_GEO_BOUNDED;
    if ((geo->kind & ~(CC_GEO_BOUNDED | 0xff)) != 0 || base < CC_GEO_EMPTY || base > CC_GEO_UNION)
        return fail(CC_ERR_UNSUPPORTED, "unknown geometry kind %d", geo->kind);
    if (base == CC_GEO_TORUS && dim != 3) return fail(CC_ERR_UNSUPPORTED, "a torus is a 3D body");
    if (base == CC_GEO_SPHERE_UNION) {
        if (geo->n_prims < 1 || !geo->prims)
            return fail(CC_ERR_ARGUMENT, "a sphere union needs at least one primitive");
        if (!(geo->p[0] > 0.0)) return fail(CC_ERR_ARGUMENT, "the smoothing k must be positive");
    }
    if (base == CC_GEO_UNION) {
        if (geo->n_prims < 1 || !geo->prims)
            return fail(CC_ERR_ARGUMENT, "a union needs at least one child");
        if (!(geo->p[0] >= 0.0) || !std::isfinite(geo->p[0]))
            return fail(CC_ERR_ARGUMENT, "the smoothing k must be finite and >= 0");
        for (int i = 0; i < geo->n_prims; ++i) {
            const double ck = geo->prims[(size_t)CC_UNION_RECORD * i];
            if (!(ck == CC_GEO_SPHERE || ck == CC_GEO_TORUS || ck == CC_GEO_BOX || ck == CC_GEO_PLANE))
                return fail(CC_ERR_UNSUPPORTED, "union child %d: kind %g is not a sphere, torus, box or plane", i, ck);
            if (ck == CC_GEO_TORUS && dim != 3) return fail(CC_ERR_UNSUPPORTED, "a torus is a 3D body");
        }
    }
    if (geo->kind & CC_GEO_BOUNDED) {
        if (base == CC_GEO_EMPTY || base == CC_GEO_PLANE)
            return fail(CC_ERR_ARGUMENT, "a bounded geometry needs a bounded child (no empty body or plane)");
        for (int a = 0; a < dim; ++a)
            if (!(geo->p[16 + a] <= geo->p[19 + a]))
                return fail(CC_ERR_ARGUMENT, "a bounded geometry's box must have low <= high on every axis");
        if (!(geo->p[22] >= 0.0)) return fail(CC_ERR_ARGUMENT, "a bounded geometry's pad_distance must be >= 0");
        if (!(geo->p[23] > 0.0)) return fail(CC_ERR_ARGUMENT, "a bounded geometry's fill_distance must be positive");
    }
    return CC_OK;
}

// Doubles in a checked geometry's primitive table (host memory): four per sphere of a sphere union,
// CC_UNION_RECORD per child of a mixed union, none for any other body.
int64_t prim_doubles(const cc_geometry *geo) {
    const int base = geo->kind & ~CC_GEO_BOUNDED;
    if (base == CC_GEO_SPHERE_UNION) return 4 * (int64_t)geo->n_prims;
    if (base == CC_GEO_UNION) return CC_UNION_RECORD * (int64_t)geo->n_prims;
    return 0;
}

int stage_geometry(cc_cache *c, const cc_geometry *geo, cc::Geo *out) {
    const int rc = check_geometry(geo, c->dim);
    if (rc) return rc;
    out->kind = geo->kind;
    out->nprims = 0;
    std::memcpy(out->p, geo->p, sizeof out->p);
    out->prims = nullptr;
    if (prim_doubles(geo) > 0) {
        const int64_t nd = prim_doubles(geo);
        bool same = (int64_t)c->prims_host.size() == nd &&
                    std::memcmp(c->prims_host.data(), geo->prims, nd * sizeof(double)) == 0;
        if (!same) {
            if (nd > c->prims_cap) {
                if (c->prims_dev) HIPC(hipFree(c->prims_dev));
                c->prims_dev = nullptr;
                HIPC(hipMalloc(&c->prims_dev, nd * sizeof(double)));
                c->prims_cap = nd;
            }
            HIPC(hipStreamSynchronize(c->stream));
            HIPC(hipMemcpy(c->prims_dev, geo->prims, nd * sizeof(double), hipMemcpyHostToDevice));
            c->prims_host.assign(geo->prims, geo->prims + nd);
        }
        out->nprims = geo->n_prims;
        out->prims = c->prims_dev;
    }
    return CC_OK;
}

void mark(cc_cache *c, const char *name) {
    if (!c->profiling || c->nev >= MAX_TIMED) return;
    c->ev_name[c->nev] = name;
    (void)hipEventRecord(c->ev[c->nev + 1], c->stream);
    c->nev += 1;
}

template <typename T> int ensure(T **p, int64_t *cap, int64_t want) {
    if (want <= *cap) return CC_OK;
    if (*p) HIPC(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, sizeof(T) * want) != hipSuccess)
        return fail(CC_ERR_OUT_OF_MEMORY, "could not allocate %lld bytes", (long long)(sizeof(T) * want));
    *cap = want;
    return CC_OK;
}

// A sphere union's nodal phi through the cull pass and the brick pass (culled exact, or the
// surrogate with certified signs).  Bitwise the plain fold at every node the surrogate does not keep.
// With bp (the fused stream pass, `nwork` work items) the phi pass is left to the stream pass: the
// cull pass certifies the sub-bricks' signs and *bp receives what the stream pass evaluates from.
template <typename T>
int phi_bricks(cc_cache *c, const cc::Geo &geo, const cc::GridT<T> &gt, T *phi, cc::BrickPhi *bp = nullptr,
               int64_t nwork = 0, bool want_signs = false) {
    if (geo.nprims > 65535)
        return fail(CC_ERR_UNSUPPORTED, "culled and surrogate phi take at most 65535 primitives");
    int nb[3];
    cc::phi_brick_counts(gt.n, nb);
    const int64_t nbrick = (int64_t)nb[0] * nb[1] * nb[2];
    int rc = ensure(&c->cull_list, &c->cull_cap, nbrick * geo.nprims);
    if (!rc) rc = ensure(&c->cull_cnt, &c->cnt_cap, nbrick);
    if (!rc) rc = ensure(&c->bstat, &c->bstat_cap, (nbrick + (bp ? nwork : 0)) * cc::PHI_BSTAT);
    if (!rc && bp && c->phi_source == CC_PHI_SURROGATE) rc = ensure(&c->cert, &c->cert_cap, nbrick * 64);
    if (!rc && want_signs) rc = ensure(&c->signs, &c->signs_cap, nbrick * (cc::PHI_BRICK * cc::PHI_BRICK * cc::PHI_BRICK / 64));
    if (rc) return rc;
    c->phi_nbricks = nbrick;
    c->phi_nwork = bp ? nwork : 0;
    // magnitude bound for the rounding margins: coordinates of primitives and lattice, radii
    double scale = 1.0, pm = 0.0, gm = 0.0;
    for (int64_t i = 0; i < geo.nprims; ++i) {
        const double *q = c->prims_host.data() + 4 * i;
        const double m = std::fmax(std::fabs(q[0]), std::fmax(std::fabs(q[1]), std::fabs(q[2]))) + std::fabs(q[3]);
        pm = std::fmax(pm, m);
    }
    for (int a = 0; a < 3; ++a)
        gm = std::fmax(gm, std::fmax(std::fabs((double)gt.x0[a]),
                                     std::fabs((double)gt.x0[a] + (double)gt.n[a] * (double)gt.d[a])));
    scale += 2.0 * (pm + gm) + std::fabs(geo.p[0]);
    cc::SurrogateArgs sa{};
    sa.scale = scale;
    if (c->phi_source == CC_PHI_SURROGATE) {
        cc::Lattice lat{};
        const int f = c->coarse;
        for (int a = 0; a < 3; ++a) {
            lat.n[a] = gt.n[a] / f + 2;
            lat.x0[a] = (float)gt.x0[a];
            lat.h[a] = (float)((double)f * (double)gt.d[a]);
            sa.nc[a] = lat.n[a];
        }
        rc = ensure(&c->nn, &c->nn_cap, (int64_t)lat.n[0] * lat.n[1] * lat.n[2]);
        if (rc) return rc;
        if (cc::launch_mlp_lattice(c->net, lat, c->nn, c->stream) != CC_OK)
            return fail(CC_ERR_ARGUMENT, "the surrogate lattice is too large");
        mark(c, "nn_lattice");
        sa.nn = c->nn;
        sa.f = f;
        sa.fshift = (f & (f - 1)) == 0 ? __builtin_ctz((unsigned)f) : -1;
        sa.finv = 1.0f / (float)f;
        const double h = std::fmax((double)gt.d[0], std::fmax((double)gt.d[1], (double)gt.d[2]));
        sa.band = (float)(c->band_cells * h);
    }
    int8_t *cert = bp && sa.nn ? c->cert : nullptr;
    cc::launch_cull<T>(gt, geo, scale, c->cull_list, c->cull_cnt, cert, bp ? c->bstat : nullptr, c->stream);
    mark(c, "cull");
    if (bp) {
        bp->list = c->cull_list;
        bp->cnt = c->cull_cnt;
        bp->cert = cert;
        bp->sa = sa;
        bp->nbx = nb[0];
        bp->nby = nb[1];
        bp->nprims = geo.nprims;
        bp->wstat = c->bstat + nbrick * cc::PHI_BSTAT;
        return CC_OK;
    }
    cc::launch_phi_bricks<T>(phi, gt, geo, c->cull_list, c->cull_cnt, sa, c->bstat, want_signs ? c->signs : nullptr,
                             c->stream);
    mark(c, "phi_bricks");
    return CC_OK;
}

// The screen's margin 1e-6 R for a grid whose nodes all lie within R of the origin (the largest
// corner norm), rounded up to a float.
template <typename T> float screen_margin(const cc::GridT<T> &gt) {
    double R = 0.0;
    for (int corner = 0; corner < 8; ++corner) {
        double s = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double v = (corner >> a) & 1 ? (double)gt.x0[a] + (double)gt.n[a] * (double)gt.d[a] : (double)gt.x0[a];
            s += v * v;
        }
        R = std::fmax(R, std::sqrt(s));
    }
    return std::nextafter((float)(1e-6 * R * (1.0 + 1e-9)), INFINITY);
}

template <typename T> int run_update(cc_cache *c, const cc::Geo *geo, const void *phi_in, int on_dev,
                                     const cc_grid *g) {
    const cc::GridT<T> gt = grid_t<T>(g, c->dim);
    const cc::Cols<T> cols = cols_t<T>(c);
    c->gen += 1;  // any surface marched before now describes an older field
    c->nev = 0;
    c->ev_ready = false;
    if (c->profiling) (void)hipEventRecord(c->ev[0], c->stream);
    T *phi = (T *)c->col[CC_FIELD_PHI];
    if (c->method == CC_MARCHING_SQUARES || c->method == CC_MARCHING_CUBES) {
        // The marching-cubes stream pass samples an analytic body itself (phi written once, never
        // re-read); a sphere union is sampled once by its own pass and then read.
        // A sphere union's culled or surrogate phi is evaluated by the fused pass from its bricks.
        cc::McTuning tn = c->tune;
        const bool bricks = c->method == CC_MARCHING_CUBES && geo && geo->kind == cc::GEO_SPHERE_UNION &&
                            c->phi_source != CC_PHI_NAIVE;
        const bool fuse = c->method == CC_MARCHING_CUBES && geo && tn.variant == cc::MC_MARCH_FUSED &&
                          (geo->kind != cc::GEO_SPHERE_UNION || bricks) &&
                          (geo->kind & cc::GEO_KIND) != cc::GEO_UNION;
        if (!fuse && tn.variant == cc::MC_MARCH_FUSED) tn.variant = cc::MC_MARCH;
        cc::BrickPhi bp{};
        // the z-march reading a brick pass's field reads its sign bitmap instead of phi: it needs only
        // each node's side, and a cut cell's exact corners are read from phi by the cut pass
        const bool signs = bricks && !fuse && tn.variant == cc::MC_MARCH;
        if (bricks) {
            uint32_t W, R, KZ, nseg, nrg, nkb;
            cc::mc_brick_tile_shape(gt.n, tn.kz > 0 ? (uint32_t)tn.kz : 16u, &W, &R, &KZ, &nseg, &nrg, &nkb);
            const int rc = phi_bricks<T>(c, *geo, gt, phi, fuse ? &bp : nullptr, (int64_t)nseg * nrg * nkb, signs);
            if (rc) return rc;
        } else if (geo && !fuse && c->method == CC_MARCHING_CUBES) {
            cc::Geo gs = *geo;
            if (sizeof(T) == 8 && c->dim == 3 && (geo->kind & cc::GEO_KIND) == cc::GEO_SPHERE_UNION &&
                geo->nprims > 1) {  // the double fold screened in single precision
                const int rc = ensure(&c->scr_dev, &c->scr_cap, (int64_t)geo->nprims);
                if (rc) return rc;
                cc::launch_screen_records(geo->prims, geo->nprims, geo->p[0], c->scr_dev, c->stream);
                gs.scr = c->scr_dev;
                gs.scr_m = screen_margin(gt);
            }
            cc::launch_sample<T>(phi, gt, gs, c->dim, c->stream);
            mark(c, "sample_nodes");
        } else if (phi_in) {
            HIPC(hipMemcpyAsync(phi, phi_in, c->colbytes[CC_FIELD_PHI],
                                on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
            mark(c, "copy_phi");
        }
        cc::launch_zero_u32(c->counter, c->stream);
        if (c->method == CC_MARCHING_CUBES) {
            cc::launch_mc_stream<T>(phi, gt, fuse ? geo : nullptr, cols, c->cutlist, c->counter, tn,
                                    c->stream, fuse && bricks ? &bp : nullptr, signs ? c->signs : nullptr);
            mark(c, "mc_stream");
            cc::launch_mc_cut<T>(phi, gt, cols, c->cutlist, c->counter, c->stream);
            mark(c, "mc_cut");
        } else if (geo) {  // the cell pass samples the body itself
            cc::launch_ms_cells_fused<T>(*geo, phi, gt, cols, c->counter, c->stream);
            mark(c, "ms_cells");
        } else {
            cc::launch_ms_cells<T>(phi, gt, cols, c->counter, c->stream);
            mark(c, "ms_cells");
        }
    } else if (c->dim == 3 && c->tune.variant != 0 && (geo->kind & cc::GEO_KIND) != cc::GEO_UNION) {
        // the fused 3D pass: fit and resolved faces in one sweep (a mixed union takes the two
        // passes, whose fit kernel carries its evaluation)
        cc::launch_plic_fused<T>(gt, *geo, cols, (T *)c->col[CC_FIELD_NORMALS],
                                 (T *)c->col[CC_FIELD_INTERCEPTS], (uint8_t *)c->col[CC_FIELD_IS_VALID],
                                 (uint32_t)c->tune.kz, c->tune.variant == cc::MC_MARCH_FUSED, c->stream);
        mark(c, "plic_fused");
    } else {
        cc::launch_plic_fit<T>(gt, *geo, c->dim, cols, (T *)c->col[CC_FIELD_NORMALS],
                               (T *)c->col[CC_FIELD_INTERCEPTS], (uint8_t *)c->col[CC_FIELD_IS_VALID],
                               c->stream);
        mark(c, "plic_fit");
        cc::launch_plic_faces<T>(gt, c->dim, cols, (const T *)c->col[CC_FIELD_NORMALS],
                                 (const T *)c->col[CC_FIELD_INTERCEPTS],
                                 (const uint8_t *)c->col[CC_FIELD_IS_VALID], c->stream);
        mark(c, "plic_faces");
    }
    HIPC(hipGetLastError());
    c->ev_ready = c->profiling;
    return CC_OK;
}

void free_all(cc_cache *c);

// Every column, the stream, counters and events of a cache whose method, dtype, dim, ncell and
// nnode are set (phi holds nnode values).  On failure everything is freed (the caller deletes c).
int allocate_columns(cc_cache *c) {
    const int method = c->method, dtype = c->dtype, dim = c->dim;
    const int64_t ncell = c->ncell, nnode = c->nnode;
    const size_t es = elem_size(dtype);
    const int64_t D = dim;
    int64_t bytes[CC_NUM_FIELDS] = {};
    bytes[CC_FIELD_KIND] = ncell;
    bytes[CC_FIELD_AMBIGUOUS] = ncell;
    bytes[CC_FIELD_VOLUME_FRACTION] = ncell * es;
    bytes[CC_FIELD_CENTROID] = ncell * D * es;
    bytes[CC_FIELD_FACE_FRACTION] = ncell * 2 * D * es;
    bytes[CC_FIELD_FACE_CENTROID_LOCAL] = ncell * 2 * D * (D - 1) * es;
    bytes[CC_FIELD_INTERFACE_CENTROID] = ncell * D * es;
    if (method == CC_PLIC) {
        bytes[CC_FIELD_NORMALS] = ncell * D * es;
        bytes[CC_FIELD_INTERCEPTS] = ncell * es;
        bytes[CC_FIELD_IS_VALID] = ncell;
    } else {
        bytes[CC_FIELD_PHI] = nnode * es;
    }
#define CREATE_HIP(expr)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            int code = e_ == hipErrorOutOfMemory ? CC_ERR_OUT_OF_MEMORY : CC_ERR_HIP;           \
            fail(code, "%s failed: %s", #expr, hipGetErrorString(e_));                          \
            free_all(c);                                                                        \
            return code;                                                                        \
        }                                                                                       \
    } while (0)
    CREATE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int f = 0; f < CC_NUM_FIELDS; ++f) {
        c->colbytes[f] = bytes[f];
        if (bytes[f] > 0) CREATE_HIP(hipMalloc(&c->col[f], bytes[f]));
    }
    CREATE_HIP(hipMalloc(&c->counter, 64));
    if (method == CC_MARCHING_CUBES) CREATE_HIP(hipMalloc(&c->cutlist, ncell * sizeof(uint32_t)));
    for (int e = 0; e <= MAX_TIMED; ++e) CREATE_HIP(hipEventCreate(&c->ev[e]));
    if (method == CC_PLIC) {
        // What a centroid plane fit has no value for, written once (src/plic/cache.jl:65-69).
        const int fields[3] = {CC_FIELD_CENTROID, CC_FIELD_FACE_CENTROID_LOCAL,
                               CC_FIELD_INTERFACE_CENTROID};
        for (int f : fields) {
            const size_t cnt = bytes[f] / es;
            if (dtype == CC_FLOAT32)
                hipLaunchKernelGGL(k_fill<float>, dim3(1024), dim3(256), 0, c->stream,
                                   (float *)c->col[f], cnt, std::nanf(""));
            else
                hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, c->stream,
                                   (double *)c->col[f], cnt, std::nan(""));
        }
        hipLaunchKernelGGL(k_fill<uint8_t>, dim3(1024), dim3(256), 0, c->stream,
                           (uint8_t *)c->col[CC_FIELD_AMBIGUOUS], (size_t)ncell, (uint8_t)0);
        CREATE_HIP(hipGetLastError());
    }
#undef CREATE_HIP
    return CC_OK;
}

int dispatch_update(cc_cache *c, const cc::Geo *geo, const void *phi, int on_dev, const cc_grid *g) {
    HIPC(hipSetDevice(c->device));
    return c->dtype == CC_FLOAT32 ? run_update<float>(c, geo, phi, on_dev, g)
                                  : run_update<double>(c, geo, phi, on_dev, g);
}

int update_impl(cc_cache *c, const cc_geometry *geo, const cc_grid *g, bool sync) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    int rc = check_grid(c, g);
    if (rc) return rc;
    cc::Geo gd;
    rc = stage_geometry(c, geo, &gd);
    if (rc) return rc;
    rc = dispatch_update(c, &gd, nullptr, 0, g);
    if (rc) return rc;
    if (sync) HIPC(hipStreamSynchronize(c->stream));
    c->async_pending = !sync;
    return CC_OK;
}

// ---- the whole-grid marching-cubes surface (src/marching_cubes/surface.jl:82-141) ----------------
// The march's per-grid arrays (once per cache) and its per-cut-cell arrays (grown to ncut).
int mesh_buffers(cc_cache *c, uint32_t ncut) {
    const int n[3] = {(int)c->n[0], (int)c->n[1], (int)c->n[2]};
    const cc::MeshSizes s = cc::mesh_sizes(n);
    cc::MeshBuffers &b = c->mb;
    auto get = [&](void **p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes < 8 ? 8 : bytes) != hipSuccess) return false;
        c->mb_allocs.push_back(*p);
        return true;
    };
    if (!b.ebits) {
        const bool ok = get(&b.ebits, 8 * s.ewords) && get(&b.cbits, 8 * s.cwords) &&
                        get((void **)&b.erank, 4 * s.ewords) && get((void **)&b.crank, 4 * s.cwords) &&
                        get((void **)&b.totals, 8 * 4) &&
                        hipHostMalloc((void **)&c->mesh_totals_host, 8 * 4, hipHostMallocDefault) == hipSuccess;
        if (!ok) {
            for (void *p : c->mb_allocs) (void)hipFree(p);
            c->mb_allocs.clear();
            b = {};
            if (c->mesh_totals_host) (void)hipHostFree(c->mesh_totals_host);
            c->mesh_totals_host = nullptr;
            return fail(CC_ERR_OUT_OF_MEMORY, "could not allocate the surface march's work arrays");
        }
    }
    if (ncut > b.cut_cap || !b.order) {
        for (void *p : {(void *)b.order, (void *)b.tn, (void *)b.cn, (void *)b.tpre, (void *)b.cpre})
            if (p) (void)hipFree(p);
        b.order = b.tn = b.cn = nullptr;
        b.tpre = b.cpre = nullptr;
        b.cut_cap = 0;
        const size_t cap = ncut > 0 ? ncut : 1;
        if (hipMalloc((void **)&b.order, 4 * cap) != hipSuccess || hipMalloc((void **)&b.tn, 4 * cap) != hipSuccess ||
            hipMalloc((void **)&b.cn, 4 * cap) != hipSuccess || hipMalloc((void **)&b.tpre, 8 * cap) != hipSuccess ||
            hipMalloc((void **)&b.cpre, 8 * cap) != hipSuccess)
            return fail(CC_ERR_OUT_OF_MEMORY, "could not allocate the surface march's %u-cell arrays", ncut);
        b.cut_cap = cap;
    }
    size_t most = s.ewords > s.cwords ? s.ewords : s.cwords;
    if (ncut > most) most = ncut;
    const size_t blocks = cc::mesh_scan_blocks(most);
    if (blocks > b.part_cap || !b.part) {
        if (b.part) (void)hipFree(b.part);
        b.part = nullptr;
        b.part_cap = 0;
        if (hipMalloc((void **)&b.part, 8 * 2 * (blocks > 0 ? blocks : 1)) != hipSuccess)
            return fail(CC_ERR_OUT_OF_MEMORY, "could not allocate the surface march's block sums");
        b.part_cap = blocks > 0 ? blocks : 1;
    }
    return CC_OK;
}

bool mesh_current(const cc_cache *c, const cc_grid *g) {
    if (c->mesh_nv < 0 || c->mesh_gen != c->gen) return false;
    for (int a = 0; a < 3; ++a)
        if (std::memcmp(&c->mesh_x0[a], &g->x0[a], 8) != 0 || std::memcmp(&c->mesh_d[a], &g->d[a], 8) != 0)
            return false;
    return true;
}

// Marks, counts and ranks on the device; one host round trip for the three totals (the output
// sizes); then every vertex and triangle.  Leaves the mesh in the cache's own device buffers.
template <typename T> int mesh_march(cc_cache *c, const cc_grid *g) {
    // the cut list of the last update holds every cell the surface passes through
    uint32_t ncut = 0;
    HIPC(hipMemcpyAsync(&ncut, c->counter, sizeof ncut, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    int rc = mesh_buffers(c, ncut);
    if (rc) return rc;
    const cc::GridT<T> gt = grid_t<T>(g, 3);
    const T *phi = (const T *)c->col[CC_FIELD_PHI];
    c->nev = 0;
    c->ev_ready = false;
    if (c->profiling) (void)hipEventRecord(c->ev[0], c->stream);
    cc::launch_mesh_count<T>(phi, gt, c->cutlist, ncut, c->mb, c->stream);
    mark(c, "mesh_count");
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(c->mesh_totals_host, c->mb.totals, 4 * sizeof(int64_t), hipMemcpyDeviceToHost,
                        c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    const int64_t ne = c->mesh_totals_host[0], nt = c->mesh_totals_host[2], ni = c->mesh_totals_host[3];
    if (c->mesh_totals_host[1] != (int64_t)ncut)
        return fail(CC_ERR_HIP, "surface march: %lld cut cells ranked, %u listed",
                    (long long)c->mesh_totals_host[1], ncut);
    const int64_t nv = ne + ni;
    if (nv > INT32_MAX || nt > INT32_MAX)
        return fail(CC_ERR_UNSUPPORTED, "a surface of %lld vertices / %lld triangles is beyond 2^31",
                    (long long)nv, (long long)nt);
    rc = ensure((char **)&c->mverts, &c->mverts_cap, 3 * sizeof(T) * (nv > 0 ? nv : 1));
    if (rc) return rc;
    rc = ensure(&c->mtris, &c->mtris_cap, 3 * (nt > 0 ? nt : 1));
    if (rc) return rc;
    mark(c, "mesh_host_sync");  // the round trip above: host time, not a kernel
    cc::launch_mesh_emit<T>(phi, gt, ncut, c->mb, (T *)c->mverts, c->mtris, c->stream);
    mark(c, "mesh_emit");
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->stream));
    c->ev_ready = c->profiling;
    c->mesh_nv = nv;
    c->mesh_nt = nt;
    c->mesh_gen = c->gen;
    for (int a = 0; a < 3; ++a) {
        c->mesh_x0[a] = g->x0[a];
        c->mesh_d[a] = g->d[a];
    }
    c->mesh_serial += 1;
    return CC_OK;
}

// ---- marching squares on a quadtree's leaves (src/marching_squares/cache.jl:91-147) -------------
int check_leaves(const cc_cache *c, const cc_leaves *t) {
    if (!c->leaves)
        return fail(CC_ERR_UNSUPPORTED, "this cache is laid out on a grid: update it over the grid (cc_update)");
    if (!t || !t->level || !t->coord) return fail(CC_ERR_ARGUMENT, "tree or its leaf arrays are NULL");
    if (t->n != c->ncell)
        return fail(CC_ERR_DIMENSION_MISMATCH,
                    "the cache holds %lld corner values and %lld cells, but `mesh` has %lld leaves; on an "
                    "`AdaptiveMesh` both are one entry per leaf, in leaf order",
                    (long long)c->ncell, (long long)c->ncell, (long long)t->n);
    for (int a = 0; a < 2; ++a)
        if (!(t->d[a] > 0.0) || !std::isfinite(t->d[a]) || !std::isfinite(t->x0[a]))
            return fail(CC_ERR_ARGUMENT, "the tree's base cell size must be positive and finite, its origin finite");
    return CC_OK;
}

template <typename T>
int run_leaves(cc_cache *c, const cc::Geo *geo, const void *phi4, int on_dev, const cc_leaves *t) {
    c->gen += 1;
    c->nev = 0;
    c->ev_ready = false;
    if (c->profiling) (void)hipEventRecord(c->ev[0], c->stream);
    const int32_t *level = t->level, *coord = t->coord;
    if (!t->on_device) {
        // the host leaf list, copied in stream order into the cache's own buffers -- unless the
        // caller's version tag says the copy from the last update is still these leaves
        const bool same = t->version != 0 && t->version == c->leaf_version && c->leaf_src[0] == t->level &&
                          c->leaf_src[1] == t->coord;
        if (!same) {
            HIPC(hipMemcpyAsync(c->leaf_level, t->level, sizeof(int32_t) * t->n, hipMemcpyHostToDevice,
                                c->stream));
            HIPC(hipMemcpyAsync(c->leaf_coord, t->coord, 2 * sizeof(int32_t) * t->n, hipMemcpyHostToDevice,
                                c->stream));
            c->leaf_src[0] = t->level;
            c->leaf_src[1] = t->coord;
            c->leaf_version = t->version;
            mark(c, "copy_leaves");
        }
        level = c->leaf_level;
        coord = c->leaf_coord;
    }
    if (phi4) {
        HIPC(hipMemcpyAsync(c->col[CC_FIELD_PHI], phi4, c->colbytes[CC_FIELD_PHI],
                            on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
        mark(c, "copy_phi");
    }
    cc::launch_zero_u32(c->counter, c->stream);
    cc::launch_leaf_cells<T>(geo, level, coord, (uint32_t)t->n, t->x0, t->d, (T *)c->col[CC_FIELD_PHI],
                             cols_t<T>(c), c->counter, c->stream);
    mark(c, "leaf_cells");
    HIPC(hipGetLastError());
    c->ev_ready = c->profiling;
    return CC_OK;
}

int update_leaves_impl(cc_cache *c, const cc_geometry *geo, const void *phi4, int on_dev, const cc_leaves *t) {
    int rc = check_leaves(c, t);
    if (rc) return rc;
    cc::Geo gd;
    if (geo) {
        rc = stage_geometry(c, geo, &gd);
        if (rc) return rc;
    }
    HIPC(hipSetDevice(c->device));
    const cc::Geo *gp = geo ? &gd : nullptr;
    rc = c->dtype == CC_FLOAT32 ? run_leaves<float>(c, gp, phi4, on_dev, t)
                                : run_leaves<double>(c, gp, phi4, on_dev, t);
    if (rc) return rc;
    HIPC(hipStreamSynchronize(c->stream));
    return CC_OK;
}

// The PLIC surface (src/plic/surface.jl:28-267): count every cell, prefix sums, one round trip for
// the sizes, then every node and element at its offset.
// The surface passes' per-entry work arrays (counts, their prefix sums, the scans' block sums and
// totals) for n entries, grown on demand (contents not kept).
int surface_work(cc_cache *c, int64_t n) {
    if (n < 1) n = 1;
    if (c->ps.vcnt && n <= c->ps_cap) return CC_OK;
    for (void *p : c->ps_allocs) (void)hipFree(p);
    c->ps_allocs.clear();
    c->ps = {};
    c->ps_cap = 0;
    auto get = [&](void **p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes < 8 ? 8 : bytes) != hipSuccess) return false;
        c->ps_allocs.push_back(*p);
        return true;
    };
    const size_t blocks = cc::mesh_scan_blocks((size_t)n);
    cc::PlicSurfBuffers &b = c->ps;
    b.part_cap = blocks > 0 ? blocks : 1;
    if (!(get((void **)&b.vcnt, 4 * n) && get((void **)&b.ecnt, 4 * n) && get((void **)&b.vpre, 8 * n) &&
          get((void **)&b.epre, 8 * n) && get((void **)&b.part, 8 * 2 * b.part_cap) && get((void **)&b.totals, 8 * 2))) {
        for (void *p : c->ps_allocs) (void)hipFree(p);
        c->ps_allocs.clear();
        c->ps = {};
        return fail(CC_ERR_OUT_OF_MEMORY, "could not allocate the surface work arrays");
    }
    c->ps_cap = n;
    return CC_OK;
}

// generate_mesh for a PLIC cache: the cells whose plane can touch them are marked and listed (one
// host round trip for their number), then counted, ranked (a second round trip for the output
// sizes) and emitted in cell order.  In 2D with a zero floor every cell draws, so all are entries.
template <typename T> int plic_surface(cc_cache *c, const cc_grid *g, double min_frac) {
    const int64_t ncell = c->ncell;
    const cc::GridT<T> gt = grid_t<T>(g, c->dim);
    const T *nrm = (const T *)c->col[CC_FIELD_NORMALS], *icp = (const T *)c->col[CC_FIELD_INTERCEPTS];
    double dmin = (double)gt.d[0] < (double)gt.d[1] ? (double)gt.d[0] : (double)gt.d[1];
    if (c->dim == 3 && (double)gt.d[2] < dmin) dmin = (double)gt.d[2];
    const bool filter = c->dim == 3 || min_frac * dmin > 0.0;  // the 2D floor launch_plic_surface applies
    c->nev = 0;
    c->ev_ready = false;
    if (c->profiling) (void)hipEventRecord(c->ev[0], c->stream);
    int64_t n = ncell;
    const int64_t *cells = nullptr;
    int rc;
    if (filter) {
        const int64_t nwords = (ncell + 63) / 64;
        rc = ensure(&c->mark_bits, &c->mark_bits_cap, nwords);
        if (rc == CC_OK) rc = ensure(&c->mark_rank, &c->mark_rank_cap, nwords);
        if (rc == CC_OK) rc = surface_work(c, nwords);
        if (rc) return rc;
        cc::launch_plic_touch<T>(nrm, icp, (size_t)ncell, c->dim, c->mark_bits, c->mark_rank, c->ps.part, c->ps.part_cap,
                                 c->ps.totals, c->stream);
        mark(c, "plic_surface_touch");
        HIPC(hipGetLastError());
        int64_t marked[2] = {0, 0};
        HIPC(hipMemcpyAsync(marked, c->ps.totals, sizeof marked, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        n = marked[0];
        if (n < 0 || n > ncell) return fail(CC_ERR_HIP, "PLIC surface: %lld cells marked of %lld", (long long)n,
                                            (long long)ncell);
        rc = ensure(&c->mark_cells, &c->mark_cells_cap, n > 0 ? n : 1);
        if (rc == CC_OK) rc = surface_work(c, n);
        if (rc) return rc;
        mark(c, "plic_surface_host_sync0");  // the round trip above: host time, not a kernel
        if (n > 0) cc::launch_plic_marked(c->mark_bits, c->mark_rank, (size_t)ncell, c->mark_cells, c->stream);
        mark(c, "plic_surface_marked");
        cells = c->mark_cells;
    } else {
        rc = surface_work(c, ncell);
        if (rc) return rc;
    }
    cc::launch_plic_surface<T>(gt, c->dim, nrm, icp, cells, true, (size_t)n, min_frac, c->ps, false, nullptr, nullptr,
                               c->stream);
    mark(c, "plic_surface_count");
    HIPC(hipGetLastError());
    int64_t tot[2] = {0, 0};
    HIPC(hipMemcpyAsync(tot, c->ps.totals, sizeof tot, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    rc = ensure((char **)&c->mverts, &c->mverts_cap, c->dim * sizeof(T) * (tot[0] > 0 ? tot[0] : 1));
    if (rc) return rc;
    rc = ensure(&c->mtris, &c->mtris_cap, c->dim * (tot[1] > 0 ? tot[1] : 1));
    if (rc) return rc;
    mark(c, "plic_surface_host_sync");
    cc::launch_plic_surface<T>(gt, c->dim, nrm, icp, cells, true, (size_t)n, min_frac, c->ps, true, (T *)c->mverts,
                               c->mtris, c->stream);
    mark(c, "plic_surface_emit");
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->stream));
    c->ev_ready = c->profiling;
    c->mesh_nv = tot[0];
    c->mesh_nt = tot[1];
    c->mesh_gen = c->gen;
    c->mesh_min_frac = min_frac;
    for (int a = 0; a < 3; ++a) {
        c->mesh_x0[a] = g->x0[a];
        c->mesh_d[a] = g->d[a];
    }
    c->mesh_serial += 1;
    return CC_OK;
}

// The diagnostics pass over cache c's columns, viewed with grid g (a pseudo-grid of n x 1 x 1 for a
// leaf cache, whose report is all that is read back).
template <typename T> int run_diagnostics(cc_cache *c, const cc::GridT<T> &g, int dim, unsigned long long *counts,
                                          T *sums) {
    constexpr size_t NCOUNT = 10;  // counters the pass writes (cc_internal.h)
    const size_t most = (size_t)cc::diag_max_blocks();  // the pass's largest block count
    const size_t bytes = NCOUNT * 8 + 4 * 8 + 4 * 8 * most;
    if (!c->dg) {
        if (hipMalloc(&c->dg, bytes) != hipSuccess) { c->dg = nullptr; return fail(CC_ERR_OUT_OF_MEMORY, "diagnostics"); }
        if (hipHostMalloc(&c->dg_host, NCOUNT * 8 + 4 * 8, hipHostMallocDefault) != hipSuccess) {
            c->dg_host = nullptr;
            return fail(CC_ERR_OUT_OF_MEMORY, "diagnostics");
        }
    }
    unsigned long long *dc = (unsigned long long *)c->dg;
    T *dsum = (T *)((char *)c->dg + NCOUNT * 8), *dpart = dsum + 4;
    c->nev = 0;
    c->ev_ready = false;
    if (c->profiling) (void)hipEventRecord(c->ev[0], c->stream);
    if (!cc::launch_diagnostics<T>(cols_t<T>(c), g, dim, dc, dpart, most, dsum, c->stream))
        return fail(CC_ERR_ARGUMENT, "diagnostics: the grid needs more blocks than the partial sums hold");
    mark(c, "diagnostics");
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(c->dg_host, c->dg, NCOUNT * 8 + 4 * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    c->ev_ready = c->profiling;
    std::memcpy(counts, c->dg_host, NCOUNT * 8);
    std::memcpy(sums, (char *)c->dg_host + NCOUNT * 8, 4 * sizeof(T));
    return CC_OK;
}

int diagnostics_impl(cc_cache *c, const cc_grid *grid, cc_diagnostics *out) {
    unsigned long long cnt[10];  // the pass's counters, cc_internal.h
    double sums[4];
    int rc;
    const int dim = c->leaves ? 2 : c->dim;
    cc_grid pg{};
    if (c->leaves) {  // the report reads kind, ambiguous and volume fraction only
        pg.dim = 2;
        pg.n[0] = c->ncell;
        pg.n[1] = pg.n[2] = 1;
        pg.d[0] = pg.d[1] = pg.d[2] = 1.0;
        grid = &pg;
    }
    if (c->dtype == CC_FLOAT32) {
        float fs[4];
        rc = run_diagnostics<float>(c, grid_t<float>(grid, dim), dim, cnt, fs);
        for (int q = 0; q < 4; ++q) sums[q] = fs[q];
    } else {
        rc = run_diagnostics<double>(c, grid_t<double>(grid, dim), dim, cnt, sums);
    }
    if (rc) return rc;
    double minvf, resmax;
    std::memcpy(&minvf, &cnt[7], 8);
    std::memcpy(&resmax, &cnt[8], 8);
    out->report.n = c->ncell;
    out->report.cut = (int64_t)cnt[0];
    out->report.inside = (int64_t)cnt[1];
    out->report.outside = (int64_t)cnt[2];
    out->report.ambiguous = (int64_t)cnt[3];
    out->report.min_volume_fraction = minvf;
    out->closure_nonzero = (int64_t)cnt[4];
    out->closure_max = resmax;
    out->face_mismatches = (int64_t)cnt[5];
    out->centroid_mismatches = (int64_t)cnt[6];
    out->nonfinite = (int64_t)cnt[9];
    out->interface_area = sums[0];
    for (int q = 0; q < 3; ++q) out->interface_normal_area[q] = q < dim ? sums[q + 1] : 0.0;
    return CC_OK;
}

void free_all(cc_cache *c) {
    if (c->order_ev) (void)hipEventDestroy(c->order_ev);
    c->order_ev = nullptr;
    if (c->dg) (void)hipFree(c->dg);
    if (c->dg_host) (void)hipHostFree(c->dg_host);
    for (void *p : c->ps_allocs) (void)hipFree(p);
    c->ps_allocs.clear();
    for (void *p : c->mb_allocs) (void)hipFree(p);
    c->mb_allocs.clear();
    for (void *p : {(void *)c->mb.order, (void *)c->mb.tn, (void *)c->mb.cn, (void *)c->mb.tpre,
                    (void *)c->mb.cpre, (void *)c->mb.part})
        if (p) (void)hipFree(p);
    if (c->mesh_totals_host) (void)hipHostFree(c->mesh_totals_host);
    if (c->mverts) (void)hipFree(c->mverts);
    if (c->mtris) (void)hipFree(c->mtris);
    if (c->mcells) (void)hipFree(c->mcells);
    if (c->mark_bits) (void)hipFree(c->mark_bits);
    if (c->mark_rank) (void)hipFree(c->mark_rank);
    if (c->mark_cells) (void)hipFree(c->mark_cells);
    if (c->mark_idx) (void)hipFree(c->mark_idx);
    if (c->mark_phi) (void)hipFree(c->mark_phi);
    if (c->mark_lc) (void)hipFree(c->mark_lc);
    if (c->weld_scratch) (void)hipFree(c->weld_scratch);
    if (c->weld_lines) (void)hipFree(c->weld_lines);
    if (c->weld_nodes) (void)hipFree(c->weld_nodes);
    if (c->leaf_level) (void)hipFree(c->leaf_level);
    if (c->leaf_coord) (void)hipFree(c->leaf_coord);
    for (auto &p : c->col)
        if (p) { (void)hipFree(p); p = nullptr; }
    if (c->cutlist) (void)hipFree(c->cutlist);
    if (c->counter) (void)hipFree(c->counter);
    if (c->prims_dev) (void)hipFree(c->prims_dev);
    if (c->scr_dev) (void)hipFree(c->scr_dev);
    if (c->cull_list) (void)hipFree(c->cull_list);
    if (c->cull_cnt) (void)hipFree(c->cull_cnt);
    if (c->nn) (void)hipFree(c->nn);
    if (c->bstat) (void)hipFree(c->bstat);
    if (c->cert) (void)hipFree(c->cert);
    if (c->signs) (void)hipFree(c->signs);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// ---- the marching-squares contour (src/marching_squares/surface.jl:59-177), on the host --------
// Welds contour vertices by exact coordinate equality (surface.jl:170-177): `+ 0` folds -0.0 onto
// 0.0; ids are 1-based, in order of first appearance, as MeshLibrary numbers nodes.  The keys are the
// coordinates' bit patterns in a flat open-addressing table (linear probing, at most half full).
struct Welder {
    std::vector<uint64_t> kx, ky;  // slot keys; an empty slot has id 0
    std::vector<int64_t> id;
    uint64_t mask = 0;
    size_t used = 0;
    std::vector<double> &nodes;
    explicit Welder(std::vector<double> &n) : nodes(n) { grow(64); }
    void reserve(size_t keys) {
        size_t cap = 64;
        while (cap < 2 * keys) cap <<= 1;
        if (cap > mask + 1) grow(cap);
    }
    static uint64_t mix(uint64_t a, uint64_t b) {
        uint64_t h = a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull + (a << 6) + (a >> 2));
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        return h ^ (h >> 29);
    }
    void grow(size_t cap) {
        std::vector<uint64_t> ox(std::move(kx)), oy(std::move(ky));
        std::vector<int64_t> oi(std::move(id));
        kx.assign(cap, 0);
        ky.assign(cap, 0);
        id.assign(cap, 0);
        mask = cap - 1;
        for (size_t q = 0; q < oi.size(); ++q)
            if (oi[q]) {
                uint64_t h = mix(ox[q], oy[q]) & mask;
                while (id[h]) h = (h + 1) & mask;
                kx[h] = ox[q];
                ky[h] = oy[q];
                id[h] = oi[q];
            }
    }
    template <typename T> int64_t operator()(T x, T y) {
        const double px = (double)(x + T(0)), py = (double)(y + T(0));
        uint64_t bx, by;
        std::memcpy(&bx, &px, 8);
        std::memcpy(&by, &py, 8);
        uint64_t h = mix(bx, by) & mask;
        while (id[h]) {
            if (kx[h] == bx && ky[h] == by) return id[h];
            h = (h + 1) & mask;
        }
        nodes.push_back(px);
        nodes.push_back(py);
        const int64_t v = (int64_t)nodes.size() / 2;
        kx[h] = bx;
        ky[h] = by;
        id[h] = v;
        if (2 * ++used > mask + 1) grow(2 * (mask + 1));
        return v;
    }
};

// One cut cell's interface segments (cell_boundary_walk, cell_gap_segments and cell_interface's
// orientation, marching_squares.jl:240-351), corners nx/ny and values ph in MS_NODE_BITS order,
// appended as welded lines.
template <typename T>
void cell_segments(const T (&nx)[4], const T (&ny)[4], const T (&ph)[4], Welder &vid, std::vector<int64_t> &lines) {
    T sax[4], say[4], sbx[4], sby[4];
    bool pres[4];
    for (int k = 0; k < 4; ++k) {  // cell_boundary_walk
        const int k2 = (k + 1) & 3;
        const T fa = ph[k], fb = ph[k2];
        const bool oa = fa >= T(0), ob = fb >= T(0);
        const T fr = (oa && ob) ? T(1) : ((!oa && !ob) ? T(0) : (oa ? fa / (fa - fb) : fb / (fb - fa)));
        pres[k] = oa || ob;
        if (oa && ob) { sax[k] = nx[k]; say[k] = ny[k]; sbx[k] = nx[k2]; sby[k] = ny[k2]; }
        else if (oa) {
            sax[k] = nx[k]; say[k] = ny[k];
            sbx[k] = nx[k] + fr * (nx[k2] - nx[k]);
            sby[k] = ny[k] + fr * (ny[k2] - ny[k]);
        } else {
            sax[k] = nx[k2] + fr * (nx[k] - nx[k2]);
            say[k] = ny[k2] + fr * (ny[k] - ny[k2]);
            sbx[k] = nx[k2]; sby[k] = ny[k2];
        }
    }
    int nseg = 0;
    for (int k = 0; k < 4 && nseg < 2; ++k) {  // cell_gap_segments, then cell_interface's swap
        if (!pres[k]) continue;
        int jn = k;
        for (int s = 1; s <= 3; ++s)
            if (pres[(k + s) & 3]) { jn = (k + s) & 3; break; }
        const T ax = sbx[k], ay = sby[k], bx = sax[jn], by = say[jn];
        if (ax != bx || ay != by) {
            nseg += 1;
            const int64_t i1 = vid(bx, by);
            const int64_t i2 = vid(ax, ay);
            lines.push_back(i1);
            lines.push_back(i2);
        }
    }
}

// The grid contour: cut cells in cell order (the screen: corners not all of one sign), entry r being
// cell idx[r] with its corner values phi[4r..4r+3] (surface.jl:59-86).
template <typename T>
void contour(int64_t ncut, const int64_t *idx, const T *phi, const cc_grid *g, std::vector<double> &nodes,
             std::vector<int64_t> &lines) {
    const int64_t n0 = g->n[0];
    Welder vid(nodes);
    vid.reserve((size_t)(2 * ncut));
    nodes.reserve((size_t)(4 * ncut));
    lines.reserve((size_t)(2 * ncut));
    const T x0 = T(g->x0[0]), y0 = T(g->x0[1]), dx = T(g->d[0]), dy = T(g->d[1]);
    for (int64_t r = 0; r < ncut; ++r) {
        const int64_t c = idx[r], i = c % n0, j = c / n0;
        const T ph[4] = {phi[4 * r], phi[4 * r + 1], phi[4 * r + 2], phi[4 * r + 3]};
        const T xl = x0 + T(i) * dx, xh = x0 + T(i + 1) * dx;
        const T yl = y0 + T(j) * dy, yh = y0 + T(j + 1) * dy;
        const T nx[4] = {xl, xh, xh, xl}, ny[4] = {yl, yl, yh, yh};
        cell_segments(nx, ny, ph, vid, lines);
    }
}

// The tree contour: cut leaves in leaf order, each from its own four stored corner values
// (surface.jl:92-110), corners off the lattice as the leaf pass places them.  Entry r is leaf idx[r],
// its corner values phi[4r..4r+3], its level and lattice corner lc[3r..3r+2] (or, lc null, from the
// host leaf list).
template <typename T>
void contour_leaves(int64_t ncut, const int64_t *idx, const T *phi, const int32_t *lc, const cc_leaves *t,
                    std::vector<double> &nodes, std::vector<int64_t> &lines) {
    Welder vid(nodes);
    vid.reserve((size_t)(2 * ncut));
    nodes.reserve((size_t)(4 * ncut));
    lines.reserve((size_t)(2 * ncut));
    const T x0 = T(t->x0[0]), y0 = T(t->x0[1]), dx = T(t->d[0]), dy = T(t->d[1]);
    for (int64_t r = 0; r < ncut; ++r) {
        const int64_t c = idx[r];
        const int lv = lc ? lc[3 * r] : t->level[c];
        const int cx = lc ? lc[3 * r + 1] : t->coord[2 * c], cy = lc ? lc[3 * r + 2] : t->coord[2 * c + 1];
        const T hx = std::ldexp(dx, -lv), hy = std::ldexp(dy, -lv);
        const T xl = x0 + T(cx) * hx, xh = x0 + T(cx + 1) * hx;
        const T yl = y0 + T(cy) * hy, yh = y0 + T(cy + 1) * hy;
        const T nx[4] = {xl, xh, xh, xl}, ny[4] = {yl, yl, yh, yh};
        const T ph[4] = {phi[4 * r], phi[4 * r + 1], phi[4 * r + 2], phi[4 * r + 3]};
        cell_segments(nx, ny, ph, vid, lines);
    }
}

// Marks and ranks the cut cells (or leaves) of the cache on the device from the corner signs of its
// field; *ncut receives their number.  n0: the grid's cells along x, or 0 for a leaf cache.
template <typename T> int mark_cut_cells(cc_cache *c, int64_t n, int64_t n0, int64_t *ncut) {
    const int64_t nwords = (n + 63) / 64;
    int rc = ensure(&c->mark_bits, &c->mark_bits_cap, nwords);
    if (rc == CC_OK) rc = ensure(&c->mark_rank, &c->mark_rank_cap, nwords);
    if (rc == CC_OK) rc = surface_work(c, nwords);
    if (rc) return rc;
    cc::launch_cut_marks<T>((const T *)c->col[CC_FIELD_PHI], (size_t)n, (size_t)n0, c->mark_bits, c->mark_rank,
                            c->ps.part, c->ps.part_cap, c->ps.totals, c->stream);
    HIPC(hipGetLastError());
    int64_t tot[2] = {0, 0};
    HIPC(hipMemcpyAsync(tot, c->ps.totals, sizeof tot, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    if (tot[0] < 0 || tot[0] > n) return fail(CC_ERR_HIP, "contour: %lld cut cells of %lld", (long long)tot[0],
                                             (long long)n);
    *ncut = tot[0];
    return CC_OK;
}

// The device weld (cc_contour.hip) of ncut gathered cut cells into the cache's kept contour.
template <typename T> int device_weld(cc_cache *c, const cc::ContourCells<T> &in, int64_t ncut) {
    const int64_t bytes = (int64_t)cc::contour_weld_bytes<T>(ncut);
    int rc = ensure(&c->weld_scratch, &c->weld_scratch_cap, bytes);
    if (rc == CC_OK) rc = ensure(&c->weld_lines, &c->weld_lines_cap, 4 * ncut);
    if (rc == CC_OK) rc = ensure(&c->weld_nodes, &c->weld_nodes_cap, 8 * ncut);
    if (rc) return rc;
    int64_t nseg = 0, nvert = 0;
    const int w = cc::contour_weld<T>(in, ncut, c->weld_scratch, (size_t)c->weld_scratch_cap, c->weld_lines,
                                      c->weld_nodes, &nseg, &nvert, c->stream);
    if (w == -1) return fail(CC_ERR_HIP, "contour: the weld's scratch is too small");
    if (w != 0) return fail(CC_ERR_HIP, "contour weld: %s", hipGetErrorString(hipGetLastError()));
    // kept on the device: the fill call copies it straight into the caller's buffers
    c->contour_dev = true;
    c->contour_dev_nn = nvert;
    c->contour_dev_nl = nseg;
    return CC_OK;
}

// Whether a contour of ncut cut cells is welded on the device: from 4096 cut cells unless the cache
// says otherwise (below that the host weld's single pass beats the device's dozen launches).
bool weld_on_device(const cc_cache *c, int64_t ncut) {
    if (c->weld_mode == CC_WELD_DEVICE) return true;
    if (c->weld_mode == CC_WELD_HOST) return false;
    return ncut >= 4096;
}

// The grid contour of a marching-squares cache into its kept contour: the cut cells' corners gathered
// on the device, welded there or on the host.
template <typename T> int grid_contour(cc_cache *c, const cc_grid *g) {
    int64_t ncut = 0;
    int rc = mark_cut_cells<T>(c, c->ncell, c->n[0], &ncut);
    if (rc) return rc;
    const int64_t m = ncut > 0 ? ncut : 1;
    rc = ensure(&c->mark_idx, &c->mark_idx_cap, m);
    if (rc == CC_OK) rc = ensure(&c->mark_phi, &c->mark_phi_cap, 4 * (int64_t)sizeof(T) * m);
    if (rc) return rc;
    c->contour_nodes.clear();
    c->contour_lines.clear();
    c->contour_dev = false;
    if (ncut > 0) {
        cc::launch_grid_gather<T>(c->mark_bits, c->mark_rank, (size_t)c->ncell, (const T *)c->col[CC_FIELD_PHI],
                                  (size_t)c->n[0], c->mark_idx, (T *)c->mark_phi, c->stream);
        HIPC(hipGetLastError());
    }
    if (ncut > 0 && weld_on_device(c, ncut)) {
        const cc::ContourCells<T> in{c->mark_idx, nullptr, (const T *)c->mark_phi, c->n[0], T(g->x0[0]),
                                     T(g->x0[1]), T(g->d[0]), T(g->d[1])};
        return device_weld<T>(c, in, ncut);
    }
    std::vector<int64_t> idx(m);
    std::vector<T> phi(4 * m);
    if (ncut > 0) {
        HIPC(hipMemcpyAsync(idx.data(), c->mark_idx, sizeof(int64_t) * ncut, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipMemcpyAsync(phi.data(), c->mark_phi, 4 * sizeof(T) * ncut, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    }
    contour<T>(ncut, idx.data(), phi.data(), g, c->contour_nodes, c->contour_lines);
    return CC_OK;
}

// The device copy of a host leaf list, as run_leaves keeps it: copied again unless the caller's
// version tag says the last copy holds these leaves.
int leaves_on_device(cc_cache *c, const cc_leaves *t, const int32_t **level, const int32_t **coord) {
    if (t->on_device) {
        *level = t->level;
        *coord = t->coord;
        return CC_OK;
    }
    const bool same = t->version != 0 && t->version == c->leaf_version && c->leaf_src[0] == t->level &&
                      c->leaf_src[1] == t->coord;
    if (!same) {
        HIPC(hipMemcpyAsync(c->leaf_level, t->level, sizeof(int32_t) * t->n, hipMemcpyHostToDevice, c->stream));
        HIPC(hipMemcpyAsync(c->leaf_coord, t->coord, 2 * sizeof(int32_t) * t->n, hipMemcpyHostToDevice, c->stream));
        c->leaf_src[0] = t->level;
        c->leaf_src[1] = t->coord;
        c->leaf_version = t->version;
    }
    *level = c->leaf_level;
    *coord = c->leaf_coord;
    return CC_OK;
}

// Marks, ranks and gathers the cut leaves on the device and contours them into the cache's kept
// contour, welded on the device or on the host.
template <typename T> int tree_contour(cc_cache *c, const cc_leaves *tree) {
    const int64_t n = tree->n;
    int64_t ncut = 0;
    int rc = mark_cut_cells<T>(c, n, 0, &ncut);
    if (rc) return rc;
    const int64_t m = ncut > 0 ? ncut : 1;
    const bool dev = ncut > 0 && weld_on_device(c, ncut);
    rc = ensure(&c->mark_idx, &c->mark_idx_cap, m);
    if (rc == CC_OK) rc = ensure(&c->mark_phi, &c->mark_phi_cap, 4 * (int64_t)sizeof(T) * m);
    if (rc == CC_OK && (tree->on_device || dev)) rc = ensure(&c->mark_lc, &c->mark_lc_cap, 3 * m);
    if (rc) return rc;
    c->contour_nodes.clear();
    c->contour_lines.clear();
    c->contour_dev = false;
    if (dev) {
        const int32_t *level = nullptr, *coord = nullptr;
        rc = leaves_on_device(c, tree, &level, &coord);
        if (rc) return rc;
        cc::launch_leaf_gather<T>(c->mark_bits, c->mark_rank, (size_t)n, (const T *)c->col[CC_FIELD_PHI], level,
                                  coord, c->mark_idx, (T *)c->mark_phi, c->mark_lc, c->stream);
        HIPC(hipGetLastError());
        const cc::ContourCells<T> in{c->mark_idx, c->mark_lc, (const T *)c->mark_phi, 0, T(tree->x0[0]),
                                     T(tree->x0[1]), T(tree->d[0]), T(tree->d[1])};
        return device_weld<T>(c, in, ncut);
    }
    std::vector<int64_t> idx(m);
    std::vector<T> phi(4 * m);
    std::vector<int32_t> lc(tree->on_device ? 3 * m : 0);
    if (ncut > 0) {
        cc::launch_leaf_gather<T>(c->mark_bits, c->mark_rank, (size_t)n, (const T *)c->col[CC_FIELD_PHI],
                                  tree->on_device ? tree->level : nullptr, tree->on_device ? tree->coord : nullptr,
                                  c->mark_idx, (T *)c->mark_phi, c->mark_lc, c->stream);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(idx.data(), c->mark_idx, sizeof(int64_t) * ncut, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipMemcpyAsync(phi.data(), c->mark_phi, 4 * sizeof(T) * ncut, hipMemcpyDeviceToHost, c->stream));
        if (tree->on_device)
            HIPC(hipMemcpyAsync(lc.data(), c->mark_lc, 3 * sizeof(int32_t) * ncut, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    }
    contour_leaves<T>(ncut, idx.data(), phi.data(), tree->on_device ? lc.data() : nullptr, tree, c->contour_nodes,
                      c->contour_lines);
    return CC_OK;
}

// The kept contour's counts, and with buffers its nodes and lines (from the host weld's vectors or
// straight from the device weld's arrays).
int contour_out(cc_cache *c, double *nodes, int64_t *nnodes, int64_t *lines, int64_t *nlines) {
    const int64_t nn = c->contour_dev ? c->contour_dev_nn : (int64_t)c->contour_nodes.size() / 2;
    const int64_t nl = c->contour_dev ? c->contour_dev_nl : (int64_t)c->contour_lines.size() / 2;
    if (nodes && lines) {
        if (*nnodes < nn || *nlines < nl)
            return fail(CC_ERR_ARGUMENT, "buffers hold %lld nodes / %lld lines, need %lld / %lld",
                        (long long)*nnodes, (long long)*nlines, (long long)nn, (long long)nl);
        if (c->contour_dev) {
            if (nl > 0) {
                HIPC(hipMemcpyAsync(nodes, c->weld_nodes, sizeof(double) * 2 * nn, hipMemcpyDeviceToHost, c->stream));
                HIPC(hipMemcpyAsync(lines, c->weld_lines, sizeof(int64_t) * 2 * nl, hipMemcpyDeviceToHost, c->stream));
                HIPC(hipStreamSynchronize(c->stream));
            }
        } else {
            std::memcpy(nodes, c->contour_nodes.data(), c->contour_nodes.size() * sizeof(double));
            std::memcpy(lines, c->contour_lines.data(), c->contour_lines.size() * sizeof(int64_t));
        }
    }
    *nnodes = nn;
    *nlines = nl;
    return CC_OK;
}

// calc_volume's host checks (src/volume.jl:29-40): element type, grid, isotropy, body; fills the
// kernel's body descriptor without its primitive table.
int check_volume_args(int dtype, const cc_geometry *geo, const cc_grid *grid, double *volume, cc::Geo *gd) {
    if (!geo || !grid || !volume) return fail(CC_ERR_ARGUMENT, "NULL argument");
    if (dtype != CC_FLOAT64 && dtype != CC_FLOAT32)
        return fail(CC_ERR_UNSUPPORTED, "element type must be CC_FLOAT64 or CC_FLOAT32");
    const int dim = grid->dim;
    if (dim != 2 && dim != 3) return fail(CC_ERR_ARGUMENT, "grid.dim must be 2 or 3");
    int64_t ncell = 1;
    for (int a = 0; a < dim; ++a) {
        if (grid->n[a] < 1) return fail(CC_ERR_ARGUMENT, "every axis needs at least one cell");
        if (!(grid->d[a] > 0.0) || !std::isfinite(grid->d[a]) || !std::isfinite(grid->x0[a]))
            return fail(CC_ERR_ARGUMENT, "grid.d must be positive and finite, x0 finite");
        ncell *= grid->n[a];
    }
    if (ncell >= (int64_t(1) << 32)) return fail(CC_ERR_UNSUPPORTED, "grids of 2^32 cells or more are not supported");
    for (int a = 1; a < dim; ++a)  // src/volume.jl:30-32
        if (grid->d[a] != grid->d[0])
            return fail(CC_ERR_ARGUMENT,
                        "calc_volume needs an isotropic grid (equal cell size on every axis), got grid.d = "
                        "(%g, %g%s)", grid->d[0], grid->d[1], dim == 3 ? ", ..." : "");
    const int grc = check_geometry(geo, dim);
    if (grc) return grc;
    gd->kind = geo->kind;
    gd->nprims = prim_doubles(geo) > 0 ? geo->n_prims : 0;
    std::memcpy(gd->p, geo->p, sizeof gd->p);
    gd->prims = nullptr;
    return CC_OK;
}

// calc_volume on stream st with the given block sums and result arrays: the fold of every cell's solid
// fraction, times prod(grid.d) in the grid's type; blocks until the result is on the host.
int run_volume(int dtype, const cc_grid *grid, const cc::Geo &gd, void *part, void *out, hipStream_t st,
               double *volume) {
    const int dim = grid->dim;
    double sum = 0.0;
    if (dtype == CC_FLOAT32) {
        const cc::GridT<float> g32 = grid_t<float>(grid, dim);
        cc::launch_plic_volume<float>(g32, gd, dim, (float *)part, (float *)out, st);
        float v = 0.0f;
        if (hipMemcpyAsync(&v, out, 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            return fail(CC_ERR_HIP, "calc_volume: %s", hipGetErrorString(hipGetLastError()));
        float c32 = g32.d[0] * g32.d[1];
        if (dim == 3) c32 = c32 * g32.d[2];
        sum = (double)(v * c32);  // vol * prod(grid.d), in the grid's type
    } else {
        const cc::GridT<double> g64 = grid_t<double>(grid, dim);
        cc::launch_plic_volume<double>(g64, gd, dim, (double *)part, (double *)out, st);
        if (hipMemcpyAsync(&sum, out, 8, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            return fail(CC_ERR_HIP, "calc_volume: %s", hipGetErrorString(hipGetLastError()));
        double cv = g64.d[0] * g64.d[1];
        if (dim == 3) cv = cv * g64.d[2];
        sum = sum * cv;
    }
    *volume = sum;
    return CC_OK;
}

}  // namespace

extern "C" {

const char *cc_last_error(void) { return g_err.c_str(); }

CC_HIDDEN void cc_internal_set_error(const char *msg) { g_err = msg ? msg : ""; }

CC_HIDDEN int cc_internal_body(cc_cache *c, int *method, int *dtype, int64_t *n, void **cols, void **cutlist,
                     void **counter, int *kz) {
    if (!c) return CC_ERR_ARGUMENT;
    *method = c->method;
    *dtype = c->dtype;
    for (int a = 0; a < 3; ++a) n[a] = c->dim == 2 && a == 2 ? 0 : c->n[a];
    for (int f = 0; f < CC_NUM_FIELDS; ++f) cols[f] = c->col[f];
    *cutlist = c->cutlist;
    *counter = c->counter;
    *kz = c->tune.kz;
    return CC_OK;
}

// Whether the cache is laid out on a quadtree's leaves.
CC_HIDDEN int cc_internal_leaves(const cc_cache *c) { return c && c->leaves ? 1 : 0; }

CC_HIDDEN void cc_internal_touch(cc_cache *c) {
    if (c) c->gen += 1;
}

CC_HIDDEN int cc_internal_order_after(cc_cache *c, hipStream_t s) {
    if (!c || !c->async_pending) return CC_OK;
    if (!c->order_ev && hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming) != hipSuccess)
        return fail(CC_ERR_HIP, "could not create an ordering event");
    HIPC(hipEventRecord(c->order_ev, c->stream));
    HIPC(hipStreamWaitEvent(s, c->order_ev, 0));
    c->async_pending = false;
    return CC_OK;
}

const char *cc_version(void) { return "libcutcell_hip 0.1.0 (gfx950)"; }

int cc_cache_create(int method, int dtype, const cc_grid *grid, cc_cache **out) {
    if (!out || !grid) return fail(CC_ERR_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (dtype != CC_FLOAT64 && dtype != CC_FLOAT32)
        return fail(CC_ERR_UNSUPPORTED, "element type must be CC_FLOAT64 or CC_FLOAT32");
    const int dim = grid->dim;
    if (method == CC_MARCHING_SQUARES && dim != 2)
        return fail(CC_ERR_ARGUMENT, "MarchingSquaresCutCell needs a 2D grid");
    if (method == CC_MARCHING_CUBES && dim != 3)
        return fail(CC_ERR_ARGUMENT, "MarchingCubesCutCell needs a 3D grid");
    if (method != CC_MARCHING_SQUARES && method != CC_MARCHING_CUBES && method != CC_PLIC)
        return fail(CC_ERR_UNSUPPORTED, "unknown method %d", method);
    if (dim != 2 && dim != 3) return fail(CC_ERR_ARGUMENT, "grid.dim must be 2 or 3");
    int64_t ncell = 1, nnode = 1;
    for (int a = 0; a < dim; ++a) {
        if (grid->n[a] < 1) return fail(CC_ERR_ARGUMENT, "every axis needs at least one cell");
        ncell *= grid->n[a];
        nnode *= grid->n[a] + 1;
    }
    if (nnode >= (int64_t(1) << 32) || ncell >= (int64_t(1) << 32))
        return fail(CC_ERR_UNSUPPORTED, "grids of 2^32 cells or nodes or more are not supported");

    cc_cache *c = new cc_cache();
    c->method = method;
    c->dtype = dtype;
    c->dim = dim;
    for (int a = 0; a < dim; ++a) c->n[a] = grid->n[a];
    c->ncell = ncell;
    c->nnode = nnode;
    // The fused PLIC sweep fits the layer below and above each work item's z-range again, so it
    // marches deeper than the nodal stream pass (measured on the 384^3 torus: kz 1 2.09 ms, 2 1.46,
    // 4 1.20, 8 1.11, 16 1.08, 32 1.13, 64 1.36 ms).
    if (method == CC_PLIC) c->tune.kz = 16;
    (void)hipGetDevice(&c->device);
    int rc = check_grid(c, grid);
    if (rc) { delete c; return rc; }

    rc = allocate_columns(c);
    if (rc) { delete c; return rc; }
    // Seeded as no body: one update against EmptyGeo(one(T)) (src/cache.jl:22-25).
    cc_geometry empty{};
    empty.kind = CC_GEO_EMPTY;
    empty.p[0] = 1.0;
    rc = update_impl(c, &empty, grid, true);
    if (rc) {
        std::string msg = g_err;
        free_all(c);
        delete c;
        g_err = msg;
        return rc;
    }
    *out = c;
    return CC_OK;
}

int cc_cache_destroy(cc_cache *c) {
    if (!c) return CC_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return CC_OK;
}

int cc_update(cc_cache *c, const cc_geometry *geo, const cc_grid *grid) {
    return update_impl(c, geo, grid, true);
}

int cc_update_async(cc_cache *c, const cc_geometry *geo, const cc_grid *grid) {
    return update_impl(c, geo, grid, false);
}

int cc_synchronize(cc_cache *c) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    HIPC(hipStreamSynchronize(c->stream));
    c->async_pending = false;
    return CC_OK;
}

int cc_set_contour_weld(cc_cache *c, int mode) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    if (mode != CC_WELD_AUTO && mode != CC_WELD_HOST && mode != CC_WELD_DEVICE)
        return fail(CC_ERR_ARGUMENT, "unknown weld mode %d", mode);
    c->weld_mode = mode;
    c->contour_ok = false;
    return CC_OK;
}

int cc_set_phi_source(cc_cache *c, int source, cc_mlp *net, int coarse, double band_cells) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    if (source != CC_PHI_NAIVE && source != CC_PHI_CULLED && source != CC_PHI_SURROGATE)
        return fail(CC_ERR_ARGUMENT, "unknown phi source %d", source);
    if (source != CC_PHI_NAIVE && c->method != CC_MARCHING_CUBES)
        return fail(CC_ERR_UNSUPPORTED, "culled and surrogate phi serve the marching-cubes cache");
    if (source == CC_PHI_SURROGATE) {
        if (!net) return fail(CC_ERR_ARGUMENT, "the surrogate source needs a network");
        if (cc::mlp_device(net) != c->device)
            return fail(CC_ERR_ARGUMENT, "the network lives on GPU %d, the cache on GPU %d",
                        cc::mlp_device(net), c->device);
        if (coarse < 1 || coarse > 64) return fail(CC_ERR_ARGUMENT, "coarse spacing must be 1..64 cells");
        if (!(band_cells >= 0.0) || !std::isfinite(band_cells))
            return fail(CC_ERR_ARGUMENT, "band must be a finite number of cells >= 0");
    }
    c->phi_source = source;
    c->net = source == CC_PHI_SURROGATE ? net : nullptr;
    if (source == CC_PHI_SURROGATE) {
        c->coarse = coarse;
        c->band_cells = band_cells;
    }
    return CC_OK;
}

int cc_phi_stats(cc_cache *c, int64_t *stats) {
    if (!c || !stats) return fail(CC_ERR_ARGUMENT, "NULL argument");
    for (int i = 0; i < CC_PHI_STATS; ++i) stats[i] = 0;
    if (!c->phi_nbricks) return CC_OK;
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->stream));
    const int64_t nb = c->phi_nbricks, nr = nb + c->phi_nwork;
    std::vector<uint32_t> cnt(nb), bs(nr * cc::PHI_BSTAT);
    HIPC(hipMemcpy(cnt.data(), c->cull_cnt, 4 * nb, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(bs.data(), c->bstat, 4 * nr * cc::PHI_BSTAT, hipMemcpyDeviceToHost));
    // nodes, exact, certificate evaluations, kept primitives, uncertified, band, contradicted, bricks;
    // the records are the bricks' and then, when the stream pass evaluated phi, its work items'
    for (int64_t b = 0; b < nb; ++b) stats[3] += cnt[b];
    for (int64_t r = 0; r < nr; ++r) {
        const uint32_t *v = bs.data() + r * cc::PHI_BSTAT;
        stats[0] += v[0];
        stats[1] += v[1];
        stats[2] += v[2];
        stats[4] += v[3];
        stats[5] += v[4];
        stats[6] += v[5];
    }
    stats[7] = nb;
    return CC_OK;
}

int cc_update_from_phi(cc_cache *c, const void *phi, int on_dev, const cc_grid *grid) {
    if (!c || !phi) return fail(CC_ERR_ARGUMENT, "NULL argument");
    if (c->method == CC_PLIC)
        return fail(CC_ERR_UNSUPPORTED, "PLIC fits planes from get_sdf, not from nodal values");
    int rc = check_grid(c, grid);
    if (rc) return rc;
    rc = dispatch_update(c, nullptr, phi, on_dev, grid);
    if (rc) return rc;
    HIPC(hipStreamSynchronize(c->stream));
    return CC_OK;
}

int cc_column(cc_cache *c, int field, void **ptr, int64_t *nbytes) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    if (field < 0 || field >= CC_NUM_FIELDS || !c->col[field])
        return fail(CC_ERR_ARGUMENT, "the cache has no column %d", field);
    if (ptr) *ptr = c->col[field];
    if (nbytes) *nbytes = c->colbytes[field];
    return CC_OK;
}

int cc_copy_to_host(cc_cache *c, int field, int64_t offset, int64_t nbytes, void *dst) {
    if (!c || !dst) return fail(CC_ERR_ARGUMENT, "NULL argument");
    if (field < 0 || field >= CC_NUM_FIELDS || !c->col[field])
        return fail(CC_ERR_ARGUMENT, "the cache has no column %d", field);
    if (offset < 0 || nbytes < 0 || offset + nbytes > c->colbytes[field])
        return fail(CC_ERR_ARGUMENT, "byte range [%lld, %lld) is outside the column (%lld bytes)",
                    (long long)offset, (long long)(offset + nbytes), (long long)c->colbytes[field]);
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->stream));
    HIPC(hipMemcpy(dst, (char *)c->col[field] + offset, nbytes, hipMemcpyDeviceToHost));
    return CC_OK;
}

int cc_cut_count(cc_cache *c, int64_t *count) {
    if (!c || !count) return fail(CC_ERR_ARGUMENT, "NULL argument");
    if (c->method == CC_PLIC) return fail(CC_ERR_UNSUPPORTED, "PLIC keeps no cut-cell list");
    uint32_t v = 0;
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->stream));
    HIPC(hipMemcpy(&v, c->counter, sizeof v, hipMemcpyDeviceToHost));
    *count = v;
    return CC_OK;
}

int cc_stream(cc_cache *c, void **stream) {
    if (!c || !stream) return fail(CC_ERR_ARGUMENT, "NULL argument");
    *stream = (void *)c->stream;
    return CC_OK;
}

int cc_set_tuning(cc_cache *c, int variant, int kz, int nontemporal) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    if (variant < cc::MC_FLAT || variant > cc::MC_MARCH_FUSED)
        return fail(CC_ERR_ARGUMENT, "unknown stream-pass variant %d", variant);
    if (kz < 1 || kz > 4096) return fail(CC_ERR_ARGUMENT, "kz must be in [1, 4096]");
    c->tune.variant = variant;
    c->tune.kz = kz;
    c->tune.nontemporal = nontemporal != 0;
    return CC_OK;
}

int cc_get_tuning(cc_cache *c, int *variant, int *kz, int *nontemporal) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    if (variant) *variant = c->tune.variant;
    if (kz) *kz = c->tune.kz;
    if (nontemporal) *nontemporal = c->tune.nontemporal;
    return CC_OK;
}

int cc_set_profiling(cc_cache *c, int on) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    c->profiling = on != 0;
    return CC_OK;
}

int cc_kernel_times(cc_cache *c, int max, float *ms, const char **names, int *count) {
    if (!c || !count) return fail(CC_ERR_ARGUMENT, "NULL argument");
    *count = 0;
    if (!c->ev_ready) return CC_OK;
    HIPC(hipEventSynchronize(c->ev[c->nev]));
    for (int k = 0; k < c->nev && k < max; ++k) {
        float t = 0.0f;
        HIPC(hipEventElapsedTime(&t, c->ev[k], c->ev[k + 1]));
        if (ms) ms[k] = t;
        if (names) names[k] = c->ev_name[k];
        *count = k + 1;
    }
    return CC_OK;
}

int cc_probe_fill_column(cc_cache *c, int field, int iters, float *ms) {
    if (!c || !ms || iters < 1) return fail(CC_ERR_ARGUMENT, "NULL cache or result, or iters < 1");
    if (field < 0 || field >= CC_NUM_FIELDS || !c->col[field])
        return fail(CC_ERR_ARGUMENT, "field %d is not a column of this cache", field);
    const size_t n = (size_t)c->colbytes[field] / 16;
    if (n == 0) return fail(CC_ERR_ARGUMENT, "column %d is shorter than 16 bytes", field);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CC_OK;
    float best = 1e30f;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(CC_ERR_HIP, "probe events: %s", hipGetErrorString(hipGetLastError()));
    } else {
        for (int it = 0; it < iters + 1; ++it) {
            float t = 0.0f;
            (void)hipEventRecord(e0, c->stream);
            hipLaunchKernelGGL(k_probe_fill, dim3(probe_blocks(n)), dim3(256), 0, c->stream, (f64x2 *)c->col[field], n,
                               0.0);
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&t, e0, e1);
            if (it > 0 && t < best) best = t;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(CC_ERR_HIP, "probe failed: %s", hipGetErrorString(e));
    }
    *ms = best;
    c->gen += 1;  // the column no longer holds the last update's values
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

int cc_probe_bandwidth(int64_t nbytes, int iters, float *fill_ms, float *copy_ms) {
    if (nbytes < (1 << 20) || nbytes > (int64_t(1) << 40) || iters < 1)
        return fail(CC_ERR_ARGUMENT, "need 1 MiB <= nbytes <= 1 TiB, iters >= 1");
    nbytes &= ~int64_t(4095);
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CC_OK;
    float best_fill = 1e30f, best_copy = 1e30f;
    const size_t n = (size_t)nbytes / 16;
    const unsigned blocks = probe_blocks(n);
    if (hipMalloc(&a, nbytes) != hipSuccess || hipMalloc(&b, nbytes) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(CC_ERR_HIP, "probe allocation failed: %s", hipGetErrorString(hipGetLastError()));
    } else {
        hipLaunchKernelGGL(k_probe_fill, dim3(blocks), dim3(256), 0, st, (f64x2 *)b, n, 2.0);
        for (int it = 0; it < iters + 1; ++it) {
            float t = 0.0f;
            (void)hipEventRecord(e0, st);
            hipLaunchKernelGGL(k_probe_fill, dim3(blocks), dim3(256), 0, st, (f64x2 *)a, n, 1.0);
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&t, e0, e1);
            if (it > 0 && t < best_fill) best_fill = t;
            (void)hipEventRecord(e0, st);
            hipLaunchKernelGGL(k_probe_copy, dim3(blocks), dim3(256), 0, st, (f64x2 *)a,
                               (const f64x2 *)b, n);
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&t, e0, e1);
            if (it > 0 && t < best_copy) best_copy = t;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(CC_ERR_HIP, "probe failed: %s", hipGetErrorString(e));
    }
    if (fill_ms) *fill_ms = best_fill;
    if (copy_ms) *copy_ms = best_copy;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int cc_probe_write(int64_t nbytes, int iters, int variant, int blocks, float *ms) {
    if (nbytes < (1 << 20) || nbytes > (int64_t(1) << 40) || iters < 1 || blocks < 1 ||
        (size_t)blocks > PROBE_MAX_BLOCKS)
        return fail(CC_ERR_ARGUMENT, "need 1 MiB <= nbytes <= 1 TiB and 1 <= blocks <= 2^23");
    nbytes &= ~int64_t(4095);
    void *a = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t st = nullptr;
    int rc = CC_OK;
    float best = 1e30f;
    if (hipMalloc(&a, nbytes) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(CC_ERR_HIP, "probe allocation failed");
    } else {
        for (int it = 0; it < iters + 1; ++it) {
            float t = 0.0f;
            (void)hipEventRecord(e0, st);
            hipLaunchKernelGGL(k_probe_fill_v, dim3(blocks), dim3(256), 0, st, (f64x2 *)a,
                               (size_t)nbytes / 16, 1.0, variant);
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&t, e0, e1);
            if (it > 0 && t < best) best = t;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(CC_ERR_HIP, "probe failed: %s", hipGetErrorString(e));
    }
    if (ms) *ms = best;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    return rc;
}

int cc_generate_mesh_ms(cc_cache *c, const cc_grid *grid, double *nodes, int64_t *nnodes,
                        int64_t *lines, int64_t *nlines) {
    if (!c || !nnodes || !nlines) return fail(CC_ERR_ARGUMENT, "NULL argument");
    if (c->method != CC_MARCHING_SQUARES)
        return fail(CC_ERR_UNSUPPORTED, "generate_mesh here is the marching-squares contour");
    int rc = check_grid(c, grid);
    if (rc) return rc;
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->stream));
    // the fill call of the two-call protocol takes the contour its size call made, when nothing was
    // updated and the grid's frame is the same; every size call contours afresh
    const double frame[4] = {grid->x0[0], grid->x0[1], grid->d[0], grid->d[1]};
    const bool reuse = nodes && lines && c->contour_ok && c->contour_gen == c->gen &&
                       std::memcmp(frame, c->contour_frame, sizeof frame) == 0;
    if (!reuse) {
        c->contour_ok = false;
        rc = c->dtype == CC_FLOAT32 ? grid_contour<float>(c, grid) : grid_contour<double>(c, grid);
        if (rc) return rc;
        c->contour_ok = true;
        c->contour_gen = c->gen;
        std::memcpy(c->contour_frame, frame, sizeof frame);
    }
    return contour_out(c, nodes, nnodes, lines, nlines);
}

int cc_leaf_cache_create(int dtype, const cc_leaves *tree, cc_cache **out) {
    if (!out || !tree) return fail(CC_ERR_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (dtype != CC_FLOAT64 && dtype != CC_FLOAT32)
        return fail(CC_ERR_UNSUPPORTED, "element type must be CC_FLOAT64 or CC_FLOAT32");
    if (tree->n < 1 || tree->n >= (int64_t(1) << 30))
        return fail(CC_ERR_ARGUMENT, "a tree of %lld leaves is not supported (1 to 2^30)", (long long)tree->n);
    cc_cache *c = new cc_cache();
    c->method = CC_MARCHING_SQUARES;
    c->dtype = dtype;
    c->dim = 2;
    c->leaves = true;
    c->n[0] = tree->n;
    c->ncell = tree->n;
    c->nnode = 4 * tree->n;  // four corner values per leaf
    (void)hipGetDevice(&c->device);
    int rc = check_leaves(c, tree);
    if (rc) { delete c; return rc; }
    rc = allocate_columns(c);
    if (rc) { delete c; return rc; }
    if (hipMalloc((void **)&c->leaf_level, sizeof(int32_t) * tree->n) != hipSuccess ||
        hipMalloc((void **)&c->leaf_coord, 2 * sizeof(int32_t) * tree->n) != hipSuccess) {
        free_all(c);
        delete c;
        return fail(CC_ERR_OUT_OF_MEMORY, "could not allocate the leaf list");
    }
    // seeded as no body: one update against EmptyGeo(1) over the tree (src/marching_squares/cache.jl:105)
    cc_geometry empty{};
    empty.kind = CC_GEO_EMPTY;
    empty.p[0] = 1.0;
    rc = update_leaves_impl(c, &empty, nullptr, 0, tree);
    if (rc) {
        std::string msg = g_err;
        free_all(c);
        delete c;
        g_err = msg;
        return rc;
    }
    *out = c;
    return CC_OK;
}

int cc_update_leaves(cc_cache *c, const cc_geometry *geo, const cc_leaves *tree) {
    if (!c || !geo) return fail(CC_ERR_ARGUMENT, "NULL argument");
    return update_leaves_impl(c, geo, nullptr, 0, tree);
}

int cc_update_leaves_from_phi(cc_cache *c, const void *phi4, int on_device, const cc_leaves *tree) {
    if (!c || !phi4) return fail(CC_ERR_ARGUMENT, "NULL argument");
    return update_leaves_impl(c, nullptr, phi4, on_device, tree);
}

int cc_generate_mesh_leaves(cc_cache *c, const cc_leaves *tree, double *nodes, int64_t *nnodes, int64_t *lines,
                            int64_t *nlines) {
    if (!c || !nnodes || !nlines) return fail(CC_ERR_ARGUMENT, "NULL argument");
    int rc = check_leaves(c, tree);
    if (rc) return rc;
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->stream));
    const int64_t n = tree->n;
    const double frame[4] = {tree->x0[0], tree->x0[1], tree->d[0], tree->d[1]};
    const bool reuse = nodes && lines && c->contour_ok && c->contour_gen == c->gen && tree->version != 0 &&
                       tree->version == c->contour_version && n == c->contour_n &&
                       c->contour_src[0] == tree->level && c->contour_src[1] == tree->coord &&
                       std::memcmp(frame, c->contour_frame, sizeof frame) == 0;
    if (!reuse) {
        c->contour_ok = false;
        rc = c->dtype == CC_FLOAT32 ? tree_contour<float>(c, tree) : tree_contour<double>(c, tree);
        if (rc) return rc;
        c->contour_ok = true;
        c->contour_gen = c->gen;
        c->contour_src[0] = tree->level;
        c->contour_src[1] = tree->coord;
        c->contour_n = n;
        c->contour_version = tree->version;
        std::memcpy(c->contour_frame, frame, sizeof frame);
    }
    return contour_out(c, nodes, nnodes, lines, nlines);
}

int cc_calc_volume(int dtype, const cc_geometry *geo, const cc_grid *grid, double *volume) {
    cc::Geo gd;
    int rc = check_volume_args(dtype, geo, grid, volume, &gd);
    if (rc) return rc;
    const int dim = grid->dim;
    const size_t prim_bytes = 8 * (size_t)prim_doubles(geo);
    const size_t es = elem_size(dtype);
    const int nb = dtype == CC_FLOAT32 ? cc::plic_volume_blocks(grid_t<float>(grid, dim), dim)
                                       : cc::plic_volume_blocks(grid_t<double>(grid, dim), dim);
    hipStream_t st = nullptr;
    void *part = nullptr, *out = nullptr;
    double *prims = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipMalloc(&part, es * nb) != hipSuccess ||
        hipMalloc(&out, es) != hipSuccess ||
        (gd.nprims > 0 && (hipMalloc((void **)&prims, prim_bytes) != hipSuccess ||
                           hipMemcpy(prims, geo->prims, prim_bytes, hipMemcpyHostToDevice) != hipSuccess))) {
        rc = fail(CC_ERR_OUT_OF_MEMORY, "calc_volume: could not set up its work arrays");
    } else {
        gd.prims = prims;
        rc = run_volume(dtype, grid, gd, part, out, st, volume);
    }
    if (prims) (void)hipFree(prims);
    if (out) (void)hipFree(out);
    if (part) (void)hipFree(part);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

int cc_calc_volume_noalloc(int dtype, const cc_geometry *geo, const cc_grid *grid, void *stream, double *volume) {
    cc::Geo gd;
    int rc = check_volume_args(dtype, geo, grid, volume, &gd);
    if (rc) return rc;
    const int64_t nd = prim_doubles(geo);
    if (nd > (int64_t)cc::VOL_SCRATCH_MAX_PRIM_DOUBLES)
        return fail(CC_ERR_ARGUMENT,
                    "calc_volume without allocation: the body's primitive table holds %lld doubles, more than "
                    "the %lld its scratch array keeps (%lld spheres of a sphere union, %lld children of a union)",
                    (long long)nd, (long long)cc::VOL_SCRATCH_MAX_PRIM_DOUBLES,
                    (long long)(cc::VOL_SCRATCH_MAX_PRIM_DOUBLES / 4),
                    (long long)(cc::VOL_SCRATCH_MAX_PRIM_DOUBLES / CC_UNION_RECORD));
    void *base = nullptr;
    const int e = cc::volume_scratch(&base);
    if (e != (int)hipSuccess)
        return fail(CC_ERR_HIP, "calc_volume without allocation: scratch lookup: %s", hipGetErrorString((hipError_t)e));
    unsigned char *ptr = (unsigned char *)base;
    hipStream_t st = (hipStream_t)stream;
    if (nd > 0) {
        HIPC(hipMemcpyAsync(ptr + cc::VOL_SCRATCH_PRIMS, geo->prims, 8 * (size_t)nd, hipMemcpyHostToDevice, st));
        gd.prims = (const double *)(ptr + cc::VOL_SCRATCH_PRIMS);
    }
    return run_volume(dtype, grid, gd, ptr, ptr + cc::VOL_SCRATCH_OUT, st, volume);
}

int cc_generate_mesh_plic(cc_cache *c, const cc_grid *grid, double min_frac, void *nodes, int64_t *nnodes,
                          int64_t *elems, int64_t *nelems) {
    if (!c || !nnodes || !nelems) return fail(CC_ERR_ARGUMENT, "NULL argument");
    if (c->method != CC_PLIC) return fail(CC_ERR_UNSUPPORTED, "this is the PLIC surface");
    if (!(min_frac >= 0.0) || !std::isfinite(min_frac))
        return fail(CC_ERR_ARGUMENT, "min_tri_area_fraction must be finite and non-negative");
    int rc = check_grid(c, grid);
    if (rc) return rc;
    HIPC(hipSetDevice(c->device));
    if (!mesh_current(c, grid) || c->mesh_min_frac != min_frac) {
        rc = c->dtype == CC_FLOAT32 ? plic_surface<float>(c, grid, min_frac) : plic_surface<double>(c, grid, min_frac);
        if (rc) return rc;
    }
    if (nodes && elems) {
        if (*nnodes < c->mesh_nv || *nelems < c->mesh_nt)
            return fail(CC_ERR_ARGUMENT, "buffers hold %lld nodes / %lld elements, need %lld / %lld",
                        (long long)*nnodes, (long long)*nelems, (long long)c->mesh_nv, (long long)c->mesh_nt);
        if (c->mesh_nv > 0)
            HIPC(hipMemcpyAsync(nodes, c->mverts, c->dim * elem_size(c->dtype) * c->mesh_nv, hipMemcpyDeviceToHost,
                                c->stream));
        if (c->mesh_nt > 0)
            HIPC(hipMemcpyAsync(elems, c->mtris, c->dim * sizeof(int64_t) * c->mesh_nt, hipMemcpyDeviceToHost,
                                c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    }
    *nnodes = c->mesh_nv;
    *nelems = c->mesh_nt;
    return CC_OK;
}

int cc_extract_surface_plic(int dtype, const cc_grid *grid, int64_t count, const int64_t *cells, const void *normals,
                            const void *intercepts, double min_frac, void *nodes, int64_t *nnodes, int64_t *elems,
                            int64_t *nelems) {
    if (!grid || !nnodes || !nelems || count < 0 || (count > 0 && (!cells || !normals || !intercepts)))
        return fail(CC_ERR_ARGUMENT, "NULL argument or negative count");
    if (dtype != CC_FLOAT64 && dtype != CC_FLOAT32)
        return fail(CC_ERR_UNSUPPORTED, "element type must be CC_FLOAT64 or CC_FLOAT32");
    if (!(min_frac >= 0.0) || !std::isfinite(min_frac))
        return fail(CC_ERR_ARGUMENT, "min_tri_area_fraction must be finite and non-negative");
    const int dim = grid->dim;
    if (dim != 2 && dim != 3) return fail(CC_ERR_ARGUMENT, "grid.dim must be 2 or 3");
    int64_t ncell = 1;
    for (int a = 0; a < dim; ++a) {
        if (grid->n[a] < 1 || !(grid->d[a] > 0.0) || !std::isfinite(grid->d[a]) || !std::isfinite(grid->x0[a]))
            return fail(CC_ERR_ARGUMENT, "grid: n >= 1, d positive and finite, x0 finite");
        ncell *= grid->n[a];
    }
    for (int64_t q = 0; q < count; ++q)
        if (cells[q] < 0 || cells[q] >= ncell)
            return fail(CC_ERR_ARGUMENT, "cell index %lld is outside the grid", (long long)cells[q]);
    const size_t es = elem_size(dtype);
    std::vector<void *> allocs;
    hipStream_t st = nullptr;
    auto get = [&](void **p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes < 8 ? 8 : bytes) != hipSuccess) return false;
        allocs.push_back(*p);
        return true;
    };
    auto cleanup = [&]() {
        for (void *p : allocs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    };
    cc::PlicSurfBuffers b = {};
    void *dn = nullptr, *di = nullptr, *dc = nullptr, *vout = nullptr, *eout = nullptr;
    const size_t n = (size_t)count;
    b.part_cap = cc::mesh_scan_blocks(n) > 0 ? cc::mesh_scan_blocks(n) : 1;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        !(get((void **)&b.vcnt, 4 * n) && get((void **)&b.ecnt, 4 * n) && get((void **)&b.vpre, 8 * n) &&
          get((void **)&b.epre, 8 * n) && get((void **)&b.part, 16 * b.part_cap) && get((void **)&b.totals, 16) &&
          get(&dn, dim * es * n) && get(&di, es * n) && get(&dc, 8 * n))) {
        cleanup();
        return fail(CC_ERR_OUT_OF_MEMORY, "extract_surface_plic: could not allocate its work arrays");
    }
    int rc = CC_OK;
    int64_t tot[2] = {0, 0};
    auto run = [&](auto tag) -> int {
        using T = decltype(tag);
        const cc::GridT<T> gt = grid_t<T>(grid, dim);
        if (n > 0) {
            HIPC(hipMemcpyAsync(dn, normals, dim * es * n, hipMemcpyHostToDevice, st));
            HIPC(hipMemcpyAsync(di, intercepts, es * n, hipMemcpyHostToDevice, st));
            HIPC(hipMemcpyAsync(dc, cells, 8 * n, hipMemcpyHostToDevice, st));
        }
        cc::launch_plic_surface<T>(gt, dim, (const T *)dn, (const T *)di, (const int64_t *)dc, false, n, min_frac, b,
                                   false, nullptr, nullptr, st);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(tot, b.totals, sizeof tot, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        if (!(nodes && elems)) return CC_OK;
        if (*nnodes < tot[0] || *nelems < tot[1])
            return fail(CC_ERR_ARGUMENT, "buffers hold %lld nodes / %lld elements, need %lld / %lld",
                        (long long)*nnodes, (long long)*nelems, (long long)tot[0], (long long)tot[1]);
        if (!get(&vout, dim * es * (tot[0] > 0 ? tot[0] : 1)) || !get(&eout, dim * 8 * (tot[1] > 0 ? tot[1] : 1)))
            return fail(CC_ERR_OUT_OF_MEMORY, "extract_surface_plic: could not allocate its output");
        cc::launch_plic_surface<T>(gt, dim, (const T *)dn, (const T *)di, (const int64_t *)dc, false, n, min_frac, b,
                                   true, (T *)vout, (int64_t *)eout, st);
        HIPC(hipGetLastError());
        if (tot[0] > 0) HIPC(hipMemcpyAsync(nodes, vout, dim * es * tot[0], hipMemcpyDeviceToHost, st));
        if (tot[1] > 0) HIPC(hipMemcpyAsync(elems, eout, dim * 8 * tot[1], hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        return CC_OK;
    };
    rc = dtype == CC_FLOAT32 ? run(0.0f) : run(0.0);
    cleanup();
    if (rc) return rc;
    *nnodes = tot[0];
    *nelems = tot[1];
    return CC_OK;
}

int cc_cut_cell_report(cc_cache *c, cc_report *out) {
    if (!c || !out) return fail(CC_ERR_ARGUMENT, "NULL argument");
    HIPC(hipSetDevice(c->device));
    cc_diagnostics d{};
    cc_grid g{};
    g.dim = c->dim;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = a < c->dim ? c->n[a] : 1;
        g.d[a] = 1.0;
    }
    const int rc = diagnostics_impl(c, &g, &d);
    if (rc) return rc;
    *out = d.report;
    return CC_OK;
}

int cc_cache_diagnostics(cc_cache *c, const cc_grid *grid, cc_diagnostics *out) {
    if (!c || !out) return fail(CC_ERR_ARGUMENT, "NULL argument");
    const int rc = check_grid(c, grid);  // also refuses a leaf cache: its faces need the tree
    if (rc) return rc;
    HIPC(hipSetDevice(c->device));
    return diagnostics_impl(c, grid, out);
}

int cc_generate_mesh_mc(cc_cache *c, const cc_grid *grid, void *nodes, int64_t *nnodes, int64_t *tris,
                        int64_t *ntris) {
    if (!c || !nnodes || !ntris) return fail(CC_ERR_ARGUMENT, "NULL argument");
    if (c->method != CC_MARCHING_CUBES)
        return fail(CC_ERR_UNSUPPORTED, "generate_mesh here is the marching-cubes surface");
    int rc = check_grid(c, grid);
    if (rc) return rc;
    HIPC(hipSetDevice(c->device));
    if (!mesh_current(c, grid)) {
        rc = c->dtype == CC_FLOAT32 ? mesh_march<float>(c, grid) : mesh_march<double>(c, grid);
        if (rc) return rc;
    }
    if (nodes && tris) {
        if (*nnodes < c->mesh_nv || *ntris < c->mesh_nt)
            return fail(CC_ERR_ARGUMENT, "buffers hold %lld nodes / %lld triangles, need %lld / %lld",
                        (long long)*nnodes, (long long)*ntris, (long long)c->mesh_nv, (long long)c->mesh_nt);
        HIPC(hipMemcpyAsync(nodes, c->mverts, 3 * elem_size(c->dtype) * c->mesh_nv, hipMemcpyDeviceToHost,
                            c->stream));
        HIPC(hipMemcpyAsync(tris, c->mtris, 3 * sizeof(int64_t) * c->mesh_nt, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    }
    *nnodes = c->mesh_nv;
    *ntris = c->mesh_nt;
    return CC_OK;
}

int cc_mesh_device(cc_cache *c, const void **nodes, const int64_t **tris, int64_t *nnodes, int64_t *ntris,
                   int64_t *serial) {
    if (!c) return fail(CC_ERR_ARGUMENT, "cache is NULL");
    if (c->mesh_nv < 0) return fail(CC_ERR_ARGUMENT, "no surface has been generated on this cache");
    if (nodes) *nodes = c->mverts;
    if (tris) *tris = c->mtris;
    if (nnodes) *nnodes = c->mesh_nv;
    if (ntris) *ntris = c->mesh_nt;
    if (serial) *serial = c->mesh_serial;
    return CC_OK;
}

int cc_mesh_cell_indices(cc_cache *c, const cc_grid *grid, int64_t *cells, int on_device) {
    if (!c || !grid || !cells) return fail(CC_ERR_ARGUMENT, "NULL argument");
    if (c->method != CC_MARCHING_CUBES)
        return fail(CC_ERR_UNSUPPORTED, "cell_indices is for a marching-cubes surface");
    if (c->mesh_nv < 0) return fail(CC_ERR_ARGUMENT, "no surface has been generated on this cache");
    if (grid->dim != 3) return fail(CC_ERR_DIMENSION_MISMATCH, "cell_indices takes a 3D grid");
    cc::MeshGrid mg;
    for (int a = 0; a < 3; ++a) {
        if (!(grid->d[a] > 0.0) || !std::isfinite(grid->d[a]) || !std::isfinite(grid->x0[a]) || grid->n[a] < 1)
            return fail(CC_ERR_ARGUMENT, "grid.d must be positive and finite, x0 finite, n >= 1");
        mg.x0[a] = grid->x0[a];
        mg.d[a] = grid->d[a];
        mg.n[a] = grid->n[a];
    }
    HIPC(hipSetDevice(c->device));
    const int64_t nt = c->mesh_nt;
    int64_t *out = cells;
    if (!on_device) {
        const int rc = ensure(&c->mcells, &c->mcells_cap, nt > 0 ? nt : 1);
        if (rc) return rc;
        out = c->mcells;
    }
    if (c->dtype == CC_FLOAT32)
        cc::launch_cell_indices<float>((const float *)c->mverts, c->mtris, nt, mg, out, c->stream);
    else
        cc::launch_cell_indices<double>((const double *)c->mverts, c->mtris, nt, mg, out, c->stream);
    HIPC(hipGetLastError());
    if (!on_device && nt > 0)
        HIPC(hipMemcpyAsync(cells, out, sizeof(int64_t) * nt, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return CC_OK;
}

}  // extern "C"
