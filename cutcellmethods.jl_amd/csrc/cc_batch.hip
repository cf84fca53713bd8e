// Many bodies, one launch per pass: the batched update of independent 3D caches (marching cubes and
// PLIC mixed), captured once in a hipGraph and replayed every step.
//
// The reference updates one cache per call (src/marching_squares/cache.jl:55-64,
// src/plic/cache.jl:77-92), two or three launches each.  Here a step over B bodies is one graph of
// three kernels, whatever B:
//   reset  -- zero every marching-cubes body's cut-cell counter (a kernel, not a memset node);
//   pass A -- every marching-cubes body's fused z-marching stream items and every PLIC body's fused
//             fit-and-faces items, in one grid.  Each body's items are cut into eight contiguous
//             chunks, chunk x going to XCD x (block b runs on XCD b mod 8), so every XCD holds the
//             same mix of memory-bound stream items and compute-heavy PLIC items, and neighbouring
//             rows of a body still share one XCD's L2 (64 mixed bodies: 5.27-5.32 ms per step, where
//             one contiguous range of work items per XCD took 5.73 ms on the same box);
//   pass B -- every marching-cubes body's cut-cell items (a grid-stride share of its work list).
// A PLIC body is done entirely in pass A by the fused sweep (fit and resolved faces together).
// The per-step inputs (grid origin, body parameters) live in a device descriptor array that the
// host refreshes with one copy before each replay; the graph itself never changes.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "cc_mc_dev.h"
#include "cc_plic_dev.h"

namespace cc {
namespace {

enum { BODY_MC = 0, BODY_PLIC = 1 };

template <typename T> struct BatchBody {
    int method;
    uint32_t ncell;
    uint32_t na;         // pass-A work items
    uint32_t cut_items;  // pass-B work items of a marching-cubes body
    uint32_t pnx, pny, pkz;  // fused PLIC tiling: tiles along x and y, layers per item
    McBody<T> mc;        // grid, body, columns, work list and tile shape (PLIC uses g, geo, cols)
    T *normals;
    T *intercepts;
    uint8_t *valid;
};

__device__ __forceinline__ int find_body(const uint32_t *begin, int nb, uint32_t w) {
    int lo = 0, hi = nb;  // begin[lo] <= w < begin[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (begin[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t xcd_order(uint32_t b) {
    const uint32_t per = gridDim.x >> 3;
    return (b & 7) * per + (b >> 3);
}

template <typename T>
__global__ void k_batch_reset(const BatchBody<T> *__restrict__ bodies, int nb) {
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        if (bodies[b].method == BODY_MC) *bodies[b].mc.counter = 0;
}

// First pass-A item of a body's chunk for XCD x: the chunks split its na items at na*x/8.
__host__ __device__ __forceinline__ uint32_t xcd_chunk(uint32_t na, uint32_t x) {
    return (uint32_t)(((uint64_t)na * x) >> 3);
}

// xbegin[x * (nb + 1) + b]: where body b's chunk starts in XCD x's list of items (a prefix over the
// bodies' chunk sizes); the grid holds 8 blocks per entry of the longest list.
template <typename T>
__global__ __launch_bounds__(MC_TILE) void k_batch_a(const BatchBody<T> *__restrict__ bodies,
                                                     const uint32_t *__restrict__ xbegin, int nb) {
    const uint32_t x = blockIdx.x & 7, l = blockIdx.x >> 3;
    const uint32_t *xb = xbegin + x * (uint32_t)(nb + 1);
    if (l >= xb[nb]) return;
    const int b = find_body(xb, nb, l);
    const uint32_t local = xcd_chunk(bodies[b].na, x) + (l - xb[b]);
    constexpr size_t lds_bytes = sizeof(MarchLDS<T>) > sizeof(PlicLDS<T>) ? sizeof(MarchLDS<T>)
                                                                           : sizeof(PlicLDS<T>);
    __shared__ __attribute__((aligned(16))) unsigned char lds[lds_bytes];
    if (bodies[b].method == BODY_MC) {
        const McBody<T> B = bodies[b].mc;
        mc_march_item<T, true, false>(B, local, *reinterpret_cast<MarchLDS<T> *>(lds));
    } else {
        const BatchBody<T> &P = bodies[b];
        plic_fused_item<T>(P.mc.g, P.mc.geo, P.mc.cols, P.normals, P.intercepts, P.valid, local, P.pnx,
                           P.pny, P.pkz, *reinterpret_cast<PlicLDS<T> *>(lds));
    }
}

template <typename T>
__global__ __launch_bounds__(MC_CUT_THREADS, 3) void k_batch_b(const BatchBody<T> *__restrict__ bodies,
                                                            const uint32_t *__restrict__ begin, int nb,
                                                            uint32_t total) {
    __shared__ T pts[MC_CUT_SLOTS * MC_CUT_THREADS];
    const uint32_t w = xcd_order(blockIdx.x);
    if (w >= total) return;
    const int b = find_body(begin, nb, w);
    const uint32_t local = w - begin[b];
    const BatchBody<T> &P = bodies[b];
    if (P.method == BODY_MC) {
        const uint32_t count = *P.mc.counter;
        const uint32_t stride = P.cut_items * (uint32_t)MC_CUT_THREADS;
        for (uint32_t e = local * (uint32_t)MC_CUT_THREADS + threadIdx.x; e < count; e += stride)
            mc_cut_item(P.mc.phi, P.mc.g, P.mc.cols, P.mc.cutlist[e], pts + threadIdx.x);
    }
}

}  // namespace
}  // namespace cc


struct cc_batch {
    int nb = 0, dtype = 0;
    std::vector<cc_cache *> caches;
    std::vector<int> kz;
    hipStream_t stream = nullptr;
    void *desc_host = nullptr;  // pinned
    void *desc_dev = nullptr;
    uint32_t *begin_dev = nullptr;
    size_t desc_bytes = 0;
    uint32_t total_a = 0, total_b = 0;
    uint32_t xlen_a = 0;  // the longest XCD list of pass-A items
    uint32_t *begin_a_dev = nullptr, *begin_b_dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float last_ms = 0.0f;
    std::vector<int> method;
};

namespace {

int bfail(int code, const char *msg) {
    cc_internal_set_error(msg);
    return code;
}

template <typename T> int build_desc(cc_batch *bt, const cc_geometry *geos, const cc_grid *grids,
                                     std::vector<uint32_t> *ia, std::vector<uint32_t> *ib) {
    auto *desc = static_cast<cc::BatchBody<T> *>(bt->desc_host);
    if (ia) { ia->assign(1, 0u); ib->assign(1, 0u); }
    for (int b = 0; b < bt->nb; ++b) {
        cc::BatchBody<T> &B = desc[b];
        int method = 0, dtype = 0, kz = 16;
        int64_t n[3];
        void *cols[11];
        void *cutlist = nullptr, *counter = nullptr;
        cc_internal_body(bt->caches[b], &method, &dtype, n, cols, &cutlist, &counter, &kz);
        const cc_grid &g = grids[b];
        const cc_geometry &geo = geos[b];
        if (g.dim != 3 || g.n[0] != n[0] || g.n[1] != n[1] || g.n[2] != n[2])
            return bfail(CC_ERR_DIMENSION_MISMATCH, "a body's grid does not match its cache's n");
        if (geo.kind == CC_GEO_SPHERE_UNION || geo.kind < CC_GEO_EMPTY || geo.kind > CC_GEO_PLANE)
            return bfail(CC_ERR_UNSUPPORTED, "batched bodies are analytic (sphere, torus, box, plane)");
        if (method == CC_PLIC && !(g.d[1] == g.d[0] && g.d[2] == g.d[0]))
            return bfail(CC_ERR_ARGUMENT, "PLICCutCell needs an isotropic grid");
        std::memset(&B, 0, sizeof B);
        B.method = method == CC_MARCHING_CUBES ? cc::BODY_MC : cc::BODY_PLIC;
        B.ncell = (uint32_t)(n[0] * n[1] * n[2]);
        for (int a = 0; a < 3; ++a) {
            B.mc.g.x0[a] = T(g.x0[a]);
            B.mc.g.d[a] = T(g.d[a]);
            B.mc.g.n[a] = (int)n[a];
        }
        B.mc.geo.kind = geo.kind;
        B.mc.geo.nprims = 0;
        std::memcpy(B.mc.geo.p, geo.p, sizeof geo.p);
        B.mc.geo.prims = nullptr;
        B.mc.cols.kind = (int8_t *)cols[CC_FIELD_KIND];
        B.mc.cols.amb = (uint8_t *)cols[CC_FIELD_AMBIGUOUS];
        B.mc.cols.vf = (T *)cols[CC_FIELD_VOLUME_FRACTION];
        B.mc.cols.centroid = (T *)cols[CC_FIELD_CENTROID];
        B.mc.cols.ff = (T *)cols[CC_FIELD_FACE_FRACTION];
        B.mc.cols.fcl = (T *)cols[CC_FIELD_FACE_CENTROID_LOCAL];
        B.mc.cols.ic = (T *)cols[CC_FIELD_INTERFACE_CENTROID];
        B.mc.phi = (T *)cols[CC_FIELD_PHI];
        B.mc.cutlist = (uint32_t *)cutlist;
        B.mc.counter = (uint32_t *)counter;
        B.normals = (T *)cols[CC_FIELD_NORMALS];
        B.intercepts = (T *)cols[CC_FIELD_INTERCEPTS];
        B.valid = (uint8_t *)cols[CC_FIELD_IS_VALID];
        uint32_t na, nbb;
        if (B.method == cc::BODY_MC) {
            cc::mc_tile_shape(B.mc.g.n, (uint32_t)bt->kz[b], &B.mc.W, &B.mc.R, &B.mc.kz, &B.mc.nseg,
                              &B.mc.nrg, &B.mc.nkb);
            B.mc.fused = 1;
            na = B.mc.nseg * B.mc.nrg * B.mc.nkb;
            // a share of the cut-cell list: about one item per 64k cells, at least one
            B.cut_items = std::max(1u, B.ncell / 65536u);
            nbb = B.cut_items;
        } else {
            // the fused PLIC sweep does the whole update in pass A
            B.pnx = (B.mc.g.n[0] + cc::PF_W - 1) / cc::PF_W;
            B.pny = (B.mc.g.n[1] + cc::PF_R - 1) / cc::PF_R;
            B.pkz = (uint32_t)bt->kz[b];
            na = B.pnx * B.pny * ((B.mc.g.n[2] + B.pkz - 1) / B.pkz);
            nbb = 0;
        }
        B.na = na;
        if (ia) {
            ia->push_back(ia->back() + na);
            ib->push_back(ib->back() + nbb);
        }
    }
    return CC_OK;
}

template <typename T> int capture(cc_batch *bt) {
    const int nb = bt->nb;
    const auto *dd = static_cast<const cc::BatchBody<T> *>(bt->desc_dev);
    if (hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bfail(CC_ERR_HIP, "hipStreamBeginCapture failed");
    hipLaunchKernelGGL((cc::k_batch_reset<T>), dim3(1), dim3(256), 0, bt->stream, dd, nb);
    const uint32_t ga = 8 * bt->xlen_a, gb = ((bt->total_b + 7) / 8) * 8;
    if (ga > 0)
        hipLaunchKernelGGL((cc::k_batch_a<T>), dim3(ga), dim3(cc::MC_TILE), 0, bt->stream, dd,
                           bt->begin_a_dev, nb);
    if (bt->total_b > 0)
        hipLaunchKernelGGL((cc::k_batch_b<T>), dim3(gb), dim3(cc::MC_CUT_THREADS), 0, bt->stream, dd, bt->begin_b_dev,
                           nb, bt->total_b);
    if (hipStreamEndCapture(bt->stream, &bt->graph) != hipSuccess)
        return bfail(CC_ERR_HIP, "hipStreamEndCapture failed");
    if (hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0) != hipSuccess)
        return bfail(CC_ERR_HIP, "hipGraphInstantiate failed");
    return CC_OK;
}

template <typename T> int rebuild(cc_batch *bt, const cc_geometry *geos, const cc_grid *grids) {
    std::vector<uint32_t> ia, ib;
    int rc = build_desc<T>(bt, geos, grids, &ia, &ib);
    if (rc) return rc;
    bt->total_a = ia.back();
    bt->total_b = ib.back();
    // pass A's per-XCD lists: body b contributes items [na*x/8, na*(x+1)/8) to XCD x
    const int nb = bt->nb;
    std::vector<uint32_t> xa(8 * (size_t)(nb + 1), 0u);
    bt->xlen_a = 0;
    for (uint32_t x = 0; x < 8; ++x) {
        uint32_t *xb = xa.data() + x * (size_t)(nb + 1);
        for (int b = 0; b < nb; ++b) {
            const uint32_t na = ia[b + 1] - ia[b];
            xb[b + 1] = xb[b] + (cc::xcd_chunk(na, x + 1) - cc::xcd_chunk(na, x));
        }
        bt->xlen_a = std::max(bt->xlen_a, xb[nb]);
    }
    if (bt->exec) { (void)hipGraphExecDestroy(bt->exec); bt->exec = nullptr; }
    if (bt->graph) { (void)hipGraphDestroy(bt->graph); bt->graph = nullptr; }
    if (hipMemcpy(bt->begin_a_dev, xa.data(), xa.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(bt->begin_b_dev, ib.data(), ib.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bfail(CC_ERR_HIP, "copying the work-item prefixes failed");
    return capture<T>(bt);
}

}  // namespace

extern "C" {

int cc_batch_destroy(cc_batch *bt);

int cc_batch_create(int nbodies, cc_cache *const *caches, cc_batch **out) {
    if (!out || !caches || nbodies < 1) return bfail(CC_ERR_ARGUMENT, "need at least one cache");
    *out = nullptr;
    cc_batch *bt = new cc_batch();
    bt->nb = nbodies;
    bt->caches.assign(caches, caches + nbodies);
    bt->kz.assign(nbodies, 8);
    for (int b = 0; b < nbodies; ++b) {
        int method = 0, dtype = 0, kz = 0;
        int64_t n[3];
        void *cols[11];
        void *cl, *ct;
        if (!caches[b] || cc_internal_body(caches[b], &method, &dtype, n, cols, &cl, &ct, &kz) != CC_OK) {
            delete bt;
            return bfail(CC_ERR_ARGUMENT, "a cache is NULL");
        }
        if (method == CC_MARCHING_SQUARES || n[2] < 1 || cols[CC_FIELD_CENTROID] == nullptr) {
            delete bt;
            return bfail(CC_ERR_UNSUPPORTED, "batched bodies are 3D marching-cubes or PLIC caches");
        }
        if (b == 0) bt->dtype = dtype;
        if (dtype != bt->dtype) {
            delete bt;
            return bfail(CC_ERR_UNSUPPORTED, "every cache of a batch has one element type");
        }
        bt->method.push_back(method);
        // In one launch over every body, short PLIC work items balance the tail better than the
        // deep z-march a PLIC cache takes alone (64 mixed bodies: 4 layers 5.6 ms, 32 layers 6.4 ms
        // per step); cc_batch_set_kz and the batch autotuner override it.
        bt->kz[b] = method == CC_PLIC ? 4 : kz;
    }
    const size_t es = bt->dtype == CC_FLOAT32 ? sizeof(cc::BatchBody<float>) : sizeof(cc::BatchBody<double>);
    bt->desc_bytes = es * nbodies;
    if (hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc(&bt->desc_host, bt->desc_bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&bt->desc_dev, bt->desc_bytes) != hipSuccess ||
        hipMalloc(&bt->begin_a_dev, 8 * (size_t)(nbodies + 1) * 4) != hipSuccess ||
        hipMalloc(&bt->begin_b_dev, (nbodies + 1) * 4) != hipSuccess ||
        hipEventCreate(&bt->e0) != hipSuccess || hipEventCreate(&bt->e1) != hipSuccess) {
        cc_batch_destroy(bt);
        return bfail(CC_ERR_HIP, "batch allocation failed");
    }
    *out = bt;
    return CC_OK;
}

int cc_batch_destroy(cc_batch *bt) {
    if (!bt) return CC_OK;
    if (bt->stream) (void)hipStreamSynchronize(bt->stream);
    if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    if (bt->desc_host) (void)hipHostFree(bt->desc_host);
    if (bt->desc_dev) (void)hipFree(bt->desc_dev);
    if (bt->begin_a_dev) (void)hipFree(bt->begin_a_dev);
    if (bt->begin_b_dev) (void)hipFree(bt->begin_b_dev);
    if (bt->e0) (void)hipEventDestroy(bt->e0);
    if (bt->e1) (void)hipEventDestroy(bt->e1);
    if (bt->stream) (void)hipStreamDestroy(bt->stream);
    delete bt;
    return CC_OK;
}

int cc_batch_set_kz(cc_batch *bt, int body, int kz) {
    if (!bt || body < 0 || body >= bt->nb || kz < 1 || kz > 4096)
        return bfail(CC_ERR_ARGUMENT, "bad batch body or kz");
    bt->kz[body] = kz;
    if (bt->exec) { (void)hipGraphExecDestroy(bt->exec); bt->exec = nullptr; }
    if (bt->graph) { (void)hipGraphDestroy(bt->graph); bt->graph = nullptr; }
    return CC_OK;
}

int cc_batch_step(cc_batch *bt, const cc_geometry *geos, const cc_grid *grids, float *device_ms) {
    if (!bt || !geos || !grids) return bfail(CC_ERR_ARGUMENT, "NULL argument");
    int rc;
    if (!bt->exec) {
        rc = bt->dtype == CC_FLOAT32 ? rebuild<float>(bt, geos, grids) : rebuild<double>(bt, geos, grids);
    } else {
        rc = bt->dtype == CC_FLOAT32 ? build_desc<float>(bt, geos, grids, nullptr, nullptr)
                                     : build_desc<double>(bt, geos, grids, nullptr, nullptr);
    }
    if (rc) return rc;
    for (int b = 0; b < bt->nb; ++b) {
        cc_internal_touch(bt->caches[b]);
        // an update a caller left running on a member cache (cc_update_async) finishes first
        if (cc_internal_order_after(bt->caches[b], bt->stream) != CC_OK) return CC_ERR_HIP;
    }
    if (hipMemcpyAsync(bt->desc_dev, bt->desc_host, bt->desc_bytes, hipMemcpyHostToDevice, bt->stream) !=
            hipSuccess ||
        hipEventRecord(bt->e0, bt->stream) != hipSuccess ||
        hipGraphLaunch(bt->exec, bt->stream) != hipSuccess ||
        hipEventRecord(bt->e1, bt->stream) != hipSuccess ||
        hipStreamSynchronize(bt->stream) != hipSuccess)
        return bfail(CC_ERR_HIP, hipGetErrorString(hipGetLastError()));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, bt->e0, bt->e1);
    bt->last_ms = ms;
    if (device_ms) *device_ms = ms;
    return CC_OK;
}

int cc_batch_work_items(cc_batch *bt, int64_t *pass_a, int64_t *pass_b) {
    if (!bt) return bfail(CC_ERR_ARGUMENT, "NULL batch");
    if (pass_a) *pass_a = bt->total_a;
    if (pass_b) *pass_b = bt->total_b;
    return CC_OK;
}

}  // extern "C"
