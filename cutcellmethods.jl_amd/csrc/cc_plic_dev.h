// Device building blocks of the PLIC passes (src/plic/plic.jl, src/plic/apertures.jl,
// src/plic/cache.jl), shared by the single-cache kernels and the batched launch.  Every symbol is
// internal to the including translation unit.
#pragma once
#include "cc_internal.h"

namespace cc {
namespace {

// CartesianMeshes.get_volume_fraction(n, a) (un-vendored): the measure of {xi in [0,1]^D : n.xi <= a}
// by Scardovelli & Zaleski (2000): negative components reflected, m_i = |n_i| sorted, s = sum m_i.
// Every formula is homogeneous of degree zero in (a, m), so there is no normalisation by s, and
// one division per evaluation; the alpha <= s/2 half by symmetry.  In 3D the m1 factor of
// a^3 - (a - m1)^3 = m1 * P1 is kept explicit so a nearly axis-aligned plane loses no digits.
template <typename T> __device__ __forceinline__ T vfrac2(T n0, T n1, T a) {
    if (n0 < T(0)) a -= n0;
    if (n1 < T(0)) a -= n1;
    const T m0 = fabs(n0), m1 = fabs(n1);
    const T s = m0 + m1;
    if (!(s > T(0))) return a >= T(0) ? T(1) : T(0);
    if (a <= T(0)) return T(0);
    if (a >= s) return T(1);
    const T lo = m0 < m1 ? m0 : m1, hi = m0 < m1 ? m1 : m0;
    const bool flip = a > T(0.5) * s;
    const T al = flip ? s - a : a;
    const T v = al < lo ? (al * al) / (T(2) * lo * hi) : (al - T(0.5) * lo) / hi;
    return flip ? T(1) - v : v;
}

template <typename T> __device__ __forceinline__ T vfrac3(const T *n, T a) {
    T m[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (n[i] < T(0)) a -= n[i];
        m[i] = fabs(n[i]);
    }
    const T s = (m[0] + m[1]) + m[2];
    if (!(s > T(0))) return a >= T(0) ? T(1) : T(0);
    if (a <= T(0)) return T(0);
    if (a >= s) return T(1);
    T t;
    if (m[0] > m[1]) { t = m[0]; m[0] = m[1]; m[1] = t; }
    if (m[1] > m[2]) { t = m[1]; m[1] = m[2]; m[2] = t; }
    if (m[0] > m[1]) { t = m[0]; m[0] = m[1]; m[1] = t; }
    const bool flip = a > T(0.5) * s;
    const T al = flip ? s - a : a;
    const T m1 = m[0], m2 = m[1], m3 = m[2], m12 = m1 + m2;
    T v;
    if (al < m1) {
        v = (al * al * al) / (T(6) * m1 * m2 * m3);
    } else if (al < m2) {
        v = (T(3) * al * (al - m1) + m1 * m1) / (T(6) * m2 * m3);
    } else if (al < m12) {
        const T u2 = al - m2;
        T c = u2 * u2 * u2;
        if (al > m3) {
            const T u3 = al - m3;
            c = c + u3 * u3 * u3;
        }
        v = (m1 * (T(3) * al * (al - m1) + m1 * m1) - c) / (T(6) * m1 * m2 * m3);
    } else {
        v = (T(2) * al - m12) / (T(2) * m3);
    }
    return flip ? T(1) - v : v;
}

// _face_area (apertures.jl:112-125) with _cell_wet (:73-97): one cell's own wet measure of its face
// `k` (0-based direction) from its stored fit.  The reference marks this @noinline because the
// plane-clip it calls is @fastmath and could round differently where inlined; here nothing is
// contracted or reassociated (-ffp-contract=off, no fast-math), so every inlined copy of this one
// expression rounds identically and both sides of a face get the same bits.
template <typename T, int D>
__device__ __forceinline__ T face_area(const T *nv, T a, T frac, bool valid, const T *cs, int k) {
    const T s = (k & 1) ? T(1) : T(0);
    if (D == 2) {
        const int c = k < 2 ? 0 : 1, p = 1 - c;
        const T len = cs[p];
        if (!valid) return frac < T(0.5) ? T(0) : len;
        const T base = nv[c] * s - a;
        return open_fraction(base, base + nv[p]) * len;
    }
    int c, p, q;
    if (k < 2) { c = 0; p = 1; q = 2; }
    else if (k < 4) { c = 1; p = 0; q = 2; }
    else { c = 2; p = 0; q = 1; }
    const T U = cs[p], V = cs[q];
    if (!valid) return frac < T(0.5) ? T(0) : U * V;
    const T base = nv[c] * s - a;
    if (nv[p] == T(0) && nv[q] == T(0)) return base >= T(0) ? U * V : T(0);
    const T solid = vfrac2(nv[p], nv[q], -base);
    return U * V * (T(1) - solid);
}

// Where every face of a cell is known without evaluating it (see plic_plane): not known, all six
// faces wholly wet (one-sided fraction exactly 1), or all wholly dry (exactly 0).
enum : int { FACES_EVAL = 0, FACES_WET = 1, FACES_DRY = 2 };

// One cell's fitted plane: cell_plane (src/plic/plic.jl:72-91) as the fit pass stores it.
template <typename T> struct Plane {
    T n[3];
    T a;
    T vf;     // the fluid fraction, 1 - solid
    T solid;  // cell_plane's fraction itself (calc_volume sums it)
    bool ok;
    int faces;  // FACES_*: whether face_area needs evaluating on this cell's faces
};

template <typename T, int D, bool MIXED = true>
__device__ __forceinline__ Plane<T> plic_plane(const GridT<T> &g, const Geo &geo, uint32_t i,
                                               uint32_t j, uint32_t k) {
    const uint32_t ijk[3] = {i, j, k};
    T x[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = g.x0[a] + (T(ijk[a]) + T(0.5)) * g.d[a];  // get_elem_centroid
    Plane<T> P;
    P.n[0] = P.n[1] = P.n[2] = T(0);
    const T dist = get_sdf<T, D, MIXED>(geo, x, P.n);
    T sum = P.n[0] + P.n[1];
    if (D == 3) sum = sum + P.n[2];
    T a = T(0.5) * sum - dist / g.d[0];
    const bool zero = P.n[0] == T(0) && P.n[1] == T(0) && (D == 2 || P.n[2] == T(0));
    T fraction;
    if (isnan(a) || zero) {
#pragma unroll
        for (int q = 0; q < 3; ++q) P.n[q] = T(0);
        a = T(0);
        fraction = dist <= T(0) ? T(1) : T(0);
        P.ok = false;
        // face_area's invalid branch: frac < 1/2 ? 0 : the full face, whose quotient is exactly 1
        P.faces = T(1) - fraction < T(0.5) ? FACES_DRY : FACES_WET;
    } else {
        fraction = D == 2 ? vfrac2(P.n[0], P.n[1], a) : vfrac3(P.n, a);
        P.ok = true;
        // A plane clear of the cell's closed cube leaves each face's own clip at an early exit of
        // vfrac2.  With a3 = a minus the negative components and S = sum |n_i|, every face's
        // reflected offset A (face_area: -(n_c s - a) minus the negative in-plane components) obeys
        // A <= a3 and A - (|n_p| + |n_q|) >= a3 - S exactly; each side is a handful of roundings of
        // magnitude at most |a| + S.  So a3 below -16 eps (|a| + S) makes every face's solid share
        // exactly 0 (wet area U*V, quotient 1), and a3 - S above the same margin makes it exactly 1
        // (area +0, quotient +0), the parallel-face guard included.
        T a3 = a, S = T(0);
#pragma unroll
        for (int q = 0; q < D; ++q) {
            if (P.n[q] < T(0)) a3 -= P.n[q];
            S = S + fabs(P.n[q]);
        }
        const T margin = T(16) * Eps<T>::v * (fabs(a) + S);
        P.faces = a3 < -margin ? FACES_WET : (a3 - S > margin ? FACES_DRY : FACES_EVAL);
    }
    P.a = a;
    P.vf = T(1) - fraction;
    P.solid = fraction;
    return P;
}

// The fit pass's stores for cell c (src/plic/cache.jl:104-112): kind from the stored fraction.
template <typename T, int D>
__device__ __forceinline__ void plic_store_fit(const Plane<T> &P, const Cols<T> &cols, T *normals,
                                               T *intercepts, uint8_t *valid, size_t c) {
#pragma unroll
    for (int q = 0; q < D; ++q) normals[D * c + q] = P.n[q];
    intercepts[c] = P.a;
    valid[c] = P.ok ? 1 : 0;
    cols.vf[c] = P.vf;
    cols.kind[c] = P.vf == T(0) ? CELL_INSIDE : (P.vf == T(1) ? CELL_OUTSIDE : CELL_CUT);
}

// _plic_fit_kernel! (src/plic/cache.jl:96-113) for cell c.
template <typename T, int D>
__device__ __forceinline__ void plic_fit_cell(const GridT<T> &g, const Geo &geo, const Cols<T> &cols,
                                              T *normals, T *intercepts, uint8_t *valid, uint32_t c) {
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    const Plane<T> P = plic_plane<T, D>(g, geo, c % n0, (c / n0) % n1, c / (n0 * n1));
    plic_store_fit<T, D>(P, cols, normals, intercepts, valid, c);
}

// _plic_face_fraction_kernel! and _resolved_slot (src/plic/cache.jl:118-162) for cell c.
template <typename T, int D>
__device__ __forceinline__ void plic_faces_cell(const GridT<T> &g, const Cols<T> &cols,
                                                const T *__restrict__ normals,
                                                const T *__restrict__ intercepts,
                                                const uint8_t *__restrict__ valid, uint32_t c) {
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    const uint32_t ijk[3] = {c % n0, (c / n0) % n1, c / (n0 * n1)};
    const uint32_t st[3] = {1u, n0, n0 * n1};
    const T *vf = cols.vf;
#pragma unroll
    for (int slot = 0; slot < 2 * D; ++slot) {
        const int ax = slot >> 1;
        T full;
        if (D == 2) full = g.d[1 - ax];
        else {
            full = T(1);
#pragma unroll
            for (int b = 0; b < 3; ++b)
                if (b != ax) full *= g.d[b];
        }
        const T own = face_area<T, D>(normals + D * (size_t)c, intercepts[c], vf[c], valid[c] != 0,
                                      g.d, slot);
        T r;
        if ((slot & 1) == 0) {
            const bool has = ijk[ax] >= 1;
            const uint32_t nb = has ? c - st[ax] : c;
            const T below = face_area<T, D>(normals + D * (size_t)nb, intercepts[nb], vf[nb],
                                            valid[nb] != 0, g.d, slot + 1);
            const T an = below / full, ap = own / full;
            r = has ? (an + ap) / T(2) : ap;
        } else {
            const bool has = ijk[ax] + 1 <= (uint32_t)g.n[ax] - 1;
            const uint32_t nb = has ? c + st[ax] : c;
            const T above = face_area<T, D>(normals + D * (size_t)nb, intercepts[nb], vf[nb],
                                            valid[nb] != 0, g.d, slot - 1);
            const T an = own / full, ap = above / full;
            r = has ? (an + ap) / T(2) : an;
        }
        cols.ff[2 * D * (size_t)c + slot] = r;
    }
}

template <typename T> __device__ __forceinline__ T plane_face(const Plane<T> &P, const T *cs, int k) {
    return face_area<T, 3>(P.n, P.a, P.vf, P.ok, cs, k);
}

// _stored_face_area(k) / full_face_area(k) for a fitted plane: the quotient both cells of a face
// resolve with.  A face known wholly wet or dry is exactly 1 or +0 (U*V equals the full area, both
// the product of the same two cell sizes in increasing axis order); an area of exactly 0 or of the
// full face gives those quotients too, so the division runs only on partly wet faces.
template <typename T>
__device__ __forceinline__ T plane_candidate(const Plane<T> &P, const T *cs, int k, T full) {
    if (P.faces == FACES_WET) return T(1);
    if (P.faces == FACES_DRY) return T(0);
    const T own = plane_face(P, cs, k);
    if (own == T(0)) return T(0);
    if (own == full) return T(1);
    return own / full;
}

// Tile of the fused 3D PLIC pass: PF_W cells along x by PF_R rows, marched in z.
constexpr int PF_W = 32, PF_R = 8;

template <typename T> struct PlicLDS {
    // one-sided wet areas of the tile's cells and of the x/y halo, each already divided by its
    // face's full area (the quotient both cells of a face use, so it is formed once)
    T cand[PF_R + 2][PF_W + 2][6];
};

// One work item of the fused 3D PLIC pass: the fit (cell_plane, src/plic/plic.jl:72-91) and the
// face resolution (src/plic/cache.jl:118-162) of the cells [i0, i0+PF_W) x [j0, j0+PF_R) over the
// layers [k0, k0+kz), with no intermediate read-back.  Each thread owns one (i, j) column and
// carries its cell's planes up the z-march in registers, so the -z and +z neighbours' candidates
// come from its own previous and next layer; the x and y neighbours' candidates come through LDS,
// the ring of cells just outside the tile being fitted again by spare threads.  A cell's plane is
// a pure function of its centroid, so a recomputed plane is bitwise the one its owner stores, and
// every face is resolved from the same two numbers as the two-pass form.
template <typename T>
__device__ void plic_fused_item(const GridT<T> &g, const Geo &geo, const Cols<T> &cols, T *normals,
                                T *intercepts, uint8_t *valid, uint32_t w, uint32_t nx, uint32_t ny,
                                uint32_t kz, PlicLDS<T> &L) {
    const uint32_t n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const uint32_t tx = w % nx, rest = w / nx, ty = rest % ny, kb = rest / ny;
    const uint32_t i0 = tx * PF_W, j0 = ty * PF_R, k0 = kb * kz, k1 = min(k0 + kz, n2);
    const int wc = (int)min((uint32_t)PF_W, n0 - i0), rc = (int)min((uint32_t)PF_R, n1 - j0);
    const int t = threadIdx.x, ii = t % PF_W, r = t / PF_W;
    const bool inside = ii < wc && r < rc;
    const uint32_t i = i0 + ii, j = j0 + r;
    T fullf[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        T f = T(1);
#pragma unroll
        for (int b = 0; b < 3; ++b)
            if (b != ax) f *= g.d[b];
        fullf[ax] = f;
    }
    // the halo cell this thread fits, if any: bottom row, top row, left column, right column
    int hr = -1, hc = -1, hslot = 0;
    uint32_t hi_ = 0, hj = 0;
    {
        const int h = t;
        if (h < wc) {
            if (j0 >= 1) { hr = 0; hc = h + 1; hslot = 3; hi_ = i0 + h; hj = j0 - 1; }
        } else if (h >= PF_W && h < PF_W + wc) {
            if (j0 + rc < n1) { hr = rc + 1; hc = h - PF_W + 1; hslot = 2; hi_ = i0 + h - PF_W; hj = j0 + rc; }
        } else if (h >= 2 * PF_W && h < 2 * PF_W + rc) {
            if (i0 >= 1) { hr = h - 2 * PF_W + 1; hc = 0; hslot = 1; hi_ = i0 - 1; hj = j0 + h - 2 * PF_W; }
        } else if (h >= 2 * PF_W + PF_R && h < 2 * PF_W + PF_R + rc) {
            if (i0 + wc < n0) {
                hr = h - 2 * PF_W - PF_R + 1; hc = wc + 1; hslot = 0; hi_ = i0 + wc; hj = j0 + h - 2 * PF_W - PF_R;
            }
        }
    }
    Plane<T> cur, nxt;
    // per face slot q of this cell's plane: its one-sided wet area / full face area (_stored_face_area)
    T cz_prev = T(0);
    T cc[6], cn[6];
    if (inside) {
        if (k0 >= 1) cz_prev = plane_candidate(plic_plane<T, 3, false>(g, geo, i, j, k0 - 1), g.d, 5, fullf[2]);
        cur = plic_plane<T, 3, false>(g, geo, i, j, k0);
#pragma unroll
        for (int q = 0; q < 6; ++q) cc[q] = plane_candidate(cur, g.d, q, fullf[q >> 1]);
    }
    for (uint32_t k = k0; k < k1; ++k) {
        if (inside && k + 1 < n2) {
            nxt = plic_plane<T, 3, false>(g, geo, i, j, k + 1);
#pragma unroll
            for (int q = 0; q < 6; ++q) cn[q] = plane_candidate(nxt, g.d, q, fullf[q >> 1]);
        }
        if (inside) {
#pragma unroll
            for (int q = 0; q < 6; ++q) L.cand[r + 1][ii + 1][q] = cc[q];
        }
        if (hr >= 0)
            L.cand[hr][hc][hslot] =
                plane_candidate(plic_plane<T, 3, false>(g, geo, hi_, hj, k), g.d, hslot, fullf[hslot >> 1]);
        __syncthreads();
        if (inside) {
            T ff[6];
            // _resolved_slot: low faces take the neighbour below's high candidate first
            ff[0] = i >= 1 ? (L.cand[r + 1][ii][1] + cc[0]) / T(2) : cc[0];
            ff[1] = i + 2 <= n0 ? (cc[1] + L.cand[r + 1][ii + 2][0]) / T(2) : cc[1];
            ff[2] = j >= 1 ? (L.cand[r][ii + 1][3] + cc[2]) / T(2) : cc[2];
            ff[3] = j + 2 <= n1 ? (cc[3] + L.cand[r + 2][ii + 1][2]) / T(2) : cc[3];
            ff[4] = k >= 1 ? (cz_prev + cc[4]) / T(2) : cc[4];
            ff[5] = k + 2 <= n2 ? (cc[5] + cn[4]) / T(2) : cc[5];
            const size_t c = i + (size_t)n0 * (j + (size_t)n1 * k);
            plic_store_fit<T, 3>(cur, cols, normals, intercepts, valid, c);
#pragma unroll
            for (int q = 0; q < 6; ++q) cols.ff[6 * c + q] = ff[q];
        }
        __syncthreads();
        if (inside) {
            cz_prev = cc[5];
            cur = nxt;
#pragma unroll
            for (int q = 0; q < 6; ++q) cc[q] = cn[q];
        }
    }
}

// Tile of the wide fused 3D PLIC pass: PW_W cells along x, one per lane of a wave, by R rows, one
// wave per row, marched in z.  Rows of 64 cells write each column in runs twice as long as the
// 32 x 8 tile's, which the HBM write stream rewards (the same sweep's stores alone, 384^3: 0.88-0.94
// against 0.98-1.13 ms), and the refitted ring is 144 cells per 512 (80 per 256 for 32 x 8).  Eight
// rows at 128 VGPRs hold two blocks per CU, whose waves cover each other's barriers (16 rows, one
// block per CU: 1.08 ms; 8 rows: 1.05 ms; 32 x 8: 1.11 ms, one box, one process).
constexpr int PW_W = 64, PW_R = 8;

template <typename T, int R> struct PlicWideLDS {
    // per layer parity: for the row below the tile, the R rows and the row above, each lane's -y
    // (slot 2) and +y (slot 3) candidates; then the left ring cell's +x and the right ring cell's -x
    // candidate of each row
    static constexpr int Y = (R + 2) * 2 * PW_W;
    static constexpr int N = Y + 2 * R;
    T v[2][N];
};

template <int R> __device__ __forceinline__ int wide_y(int row, int s, int lane) {
    return (row * 2 + s) * PW_W + lane;
}

// One work item of the wide fused sweep: cells [i0, i0+64) x [j0, j0+R) over the layers
// [k0, k0+kz).  The same per-cell functions as plic_fused_item, so every column is bitwise the
// same; only who computes what differs:
//  - wave r owns row j0 + r and lane l column i0 + l, and carries the planes' candidates up the
//    z-march in registers (the -z and +z neighbours);
//  - a layer's fit is stored as soon as it is computed (the next layer's, one step ahead), so only
//    the candidates are carried;
//  - the x neighbours' candidates come from the adjacent lanes; the y neighbours' through LDS,
//    double-buffered by layer parity so one barrier per layer orders them (a wave writing parity p
//    again at layer k + 2 has passed the barrier of layer k + 1, which every wave reaches only after
//    its reads of layer k);
//  - the ring just outside the tile is fitted again: the row below by wave 0, the row above by wave
//    R - 1, the cells left and right of each row by lanes of wave 1, so three waves fit twice per
//    layer and the four SIMDs share that work evenly.
template <typename T, int R>
__device__ void plic_wide_item(const GridT<T> &g, const Geo &geo, const Cols<T> &cols, T *normals,
                               T *intercepts, uint8_t *valid, uint32_t w, uint32_t nx, uint32_t ny,
                               uint32_t kz, PlicWideLDS<T, R> &L) {
    static_assert(R >= 3 && 2 * R <= PW_W, "the ring duties need waves 0, 1 and R - 1 distinct");
    constexpr int Y = PlicWideLDS<T, R>::Y;
    const uint32_t n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const uint32_t tx = w % nx, rest = w / nx, ty = rest % ny, kb = rest / ny;
    const uint32_t i0 = tx * PW_W, j0 = ty * R, k0 = kb * kz, k1 = min(k0 + kz, n2);
    const int wc = (int)min((uint32_t)PW_W, n0 - i0), rc = (int)min((uint32_t)R, n1 - j0);
    const int lane = threadIdx.x % PW_W, r = threadIdx.x / PW_W;
    const bool inside = lane < wc && r < rc;
    const uint32_t i = i0 + lane, j = j0 + r;
    T fullf[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        T f = T(1);
#pragma unroll
        for (int b = 0; b < 3; ++b)
            if (b != ax) f *= g.d[b];
        fullf[ax] = f;
    }
    // this lane's ring cell, if any: its LDS slot, the face slot it supplies and its (i, j)
    int hoff = -1, hslot = 0;
    uint32_t hi_ = 0, hj = 0;
    if (r == 0) {
        if (lane < wc && j0 >= 1) { hoff = wide_y<R>(0, 1, lane); hslot = 3; hi_ = i0 + lane; hj = j0 - 1; }
    } else if (r == R - 1) {
        if (lane < wc && j0 + rc < n1) {
            hoff = wide_y<R>(rc + 1, 0, lane); hslot = 2; hi_ = i0 + lane; hj = j0 + rc;
        }
    } else if (r == 1) {
        if (lane < R) {
            if (i0 >= 1 && lane < rc) { hoff = Y + lane; hslot = 1; hi_ = i0 - 1; hj = j0 + lane; }
        } else if (lane < 2 * R) {
            const int q = lane - R;
            if (i0 + wc < n0 && q < rc) { hoff = Y + R + q; hslot = 0; hi_ = i0 + wc; hj = j0 + q; }
        }
    }
    T cz_prev = T(0), cc[6], cn[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) cc[q] = cn[q] = T(0);
    if (inside) {
        if (k0 >= 1) cz_prev = plane_candidate(plic_plane<T, 3, false>(g, geo, i, j, k0 - 1), g.d, 5, fullf[2]);
        const Plane<T> P = plic_plane<T, 3, false>(g, geo, i, j, k0);
        plic_store_fit<T, 3>(P, cols, normals, intercepts, valid, i + (size_t)n0 * (j + (size_t)n1 * k0));
#pragma unroll
        for (int q = 0; q < 6; ++q) cc[q] = plane_candidate(P, g.d, q, fullf[q >> 1]);
    }
    for (uint32_t k = k0; k < k1; ++k) {
        T *buf = L.v[k & 1];
        if (inside && k + 1 < n2) {
            const Plane<T> P = plic_plane<T, 3, false>(g, geo, i, j, k + 1);
            if (k + 1 < k1)
                plic_store_fit<T, 3>(P, cols, normals, intercepts, valid,
                                     i + (size_t)n0 * (j + (size_t)n1 * (k + 1)));
#pragma unroll
            for (int q = 0; q < 6; ++q) cn[q] = plane_candidate(P, g.d, q, fullf[q >> 1]);
        }
        if (inside) {
            buf[wide_y<R>(r + 1, 0, lane)] = cc[2];
            buf[wide_y<R>(r + 1, 1, lane)] = cc[3];
        }
        if (hoff >= 0)
            buf[hoff] = plane_candidate(plic_plane<T, 3, false>(g, geo, hi_, hj, k), g.d, hslot,
                                        fullf[hslot >> 1]);
        __syncthreads();
        // every lane takes part in the shuffles; lanes past the tile's x edge hold zeros
        const T left = __shfl_up(cc[1], 1, PW_W), right = __shfl_down(cc[0], 1, PW_W);
        if (inside) {
            const T lft = lane == 0 ? buf[Y + r] : left;
            const T rgt = lane + 1 == wc ? buf[Y + R + r] : right;
            T ff[6];
            // _resolved_slot: low faces take the neighbour below's high candidate first
            ff[0] = i >= 1 ? (lft + cc[0]) / T(2) : cc[0];
            ff[1] = i + 2 <= n0 ? (cc[1] + rgt) / T(2) : cc[1];
            ff[2] = j >= 1 ? (buf[wide_y<R>(r, 1, lane)] + cc[2]) / T(2) : cc[2];
            ff[3] = j + 2 <= n1 ? (cc[3] + buf[wide_y<R>(r + 2, 0, lane)]) / T(2) : cc[3];
            ff[4] = k >= 1 ? (cz_prev + cc[4]) / T(2) : cc[4];
            ff[5] = k + 2 <= n2 ? (cc[5] + cn[4]) / T(2) : cc[5];
            const size_t c = i + (size_t)n0 * (j + (size_t)n1 * k);
#pragma unroll
            for (int q = 0; q < 6; ++q) cols.ff[6 * c + q] = ff[q];
            cz_prev = cc[5];
#pragma unroll
            for (int q = 0; q < 6; ++q) cc[q] = cn[q];
        }
    }
}

}  // namespace
}  // namespace cc
