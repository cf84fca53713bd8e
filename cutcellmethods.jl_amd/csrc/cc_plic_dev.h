// Device building blocks of the PLIC passes (src/plic/plic.jl, src/plic/apertures.jl,
// src/plic/cache.jl), shared by the single-cache kernels and the batched launch.  Every symbol is
// internal to the including translation unit.
#pragma once
#include "cc_internal.h"

namespace cc {
namespace {

// CartesianMeshes.get_volume_fraction(n, a) (un-vendored): the measure of {xi in [0,1]^D : n.xi <= a}
// by Scardovelli & Zaleski (2000), negative components reflected, normalised by sum|n|, the
// alpha <= 1/2 half by symmetry.  In 3D the m1 factor of alpha^3 - (alpha - m1)^3 is cancelled by
// hand so a nearly axis-aligned plane loses no digits.
template <typename T> __device__ __forceinline__ T vfrac2(T n0, T n1, T a) {
    if (n0 < T(0)) a -= n0;
    if (n1 < T(0)) a -= n1;
    T m0 = fabs(n0), m1 = fabs(n1);
    const T s = m0 + m1;
    if (!(s > T(0))) return a >= T(0) ? T(1) : T(0);
    m0 /= s;
    m1 /= s;
    a /= s;
    if (a <= T(0)) return T(0);
    if (a >= T(1)) return T(1);
    const T lo = m0 < m1 ? m0 : m1, hi = m0 < m1 ? m1 : m0;
    const bool flip = a > T(0.5);
    const T al = flip ? T(1) - a : a;
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
#pragma unroll
    for (int i = 0; i < 3; ++i) m[i] /= s;
    a /= s;
    if (a <= T(0)) return T(0);
    if (a >= T(1)) return T(1);
    T t;
    if (m[0] > m[1]) { t = m[0]; m[0] = m[1]; m[1] = t; }
    if (m[1] > m[2]) { t = m[1]; m[1] = m[2]; m[2] = t; }
    if (m[0] > m[1]) { t = m[0]; m[0] = m[1]; m[1] = t; }
    const bool flip = a > T(0.5);
    const T al = flip ? T(1) - a : a;
    const T m1 = m[0], m2 = m[1], m3 = m[2], m12 = m1 + m2;
    T v;
    if (al < m1) {
        v = (al * al * al) / (T(6) * m1 * m2 * m3);
    } else {
        const T P1 = T(3) * al * (al - m1) + m1 * m1;
        if (al < m2) {
            v = P1 / (T(6) * m2 * m3);
        } else if (al < m12) {
            const T u2 = al - m2;
            T c = u2 * u2 * (u2 / m1);
            if (al > m3) {
                const T u3 = al - m3;
                c = c + u3 * u3 * (u3 / m1);
            }
            v = (P1 - c) / (T(6) * m2 * m3);
        } else {
            v = (T(2) * al - m12) / (T(2) * m3);
        }
    }
    return flip ? T(1) - v : v;
}

// _face_area (apertures.jl:112-125) with _cell_wet (:73-97): one cell's own wet measure of its face
// `k` (0-based direction) from its stored fit.  noinline: one compiled body for both sides of a face.
template <typename T, int D>
__device__ __noinline__ T face_area(const T *nv, T a, T frac, bool valid, const T *cs, int k) {
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

// cell_plane (src/plic/plic.jl:72-91) and _plic_fit_kernel! (src/plic/cache.jl:96-113) for cell c.
template <typename T, int D>
__device__ __forceinline__ void plic_fit_cell(const GridT<T> &g, const Geo &geo, const Cols<T> &cols,
                                              T *normals, T *intercepts, uint8_t *valid, uint32_t c) {
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    const uint32_t ijk[3] = {c % n0, (c / n0) % n1, c / (n0 * n1)};
    T x[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = g.x0[a] + (T(ijk[a]) + T(0.5)) * g.d[a];  // get_elem_centroid
    T nrm[3] = {T(0), T(0), T(0)};
    const T dist = get_sdf<T, D>(geo, x, nrm);
    T sum = nrm[0] + nrm[1];
    if (D == 3) sum = sum + nrm[2];
    T a = T(0.5) * sum - dist / g.d[0];
    bool zero = nrm[0] == T(0) && nrm[1] == T(0) && (D == 2 || nrm[2] == T(0));
    T fraction;
    bool ok;
    if (isnan(a) || zero) {
#pragma unroll
        for (int q = 0; q < D; ++q) nrm[q] = T(0);
        a = T(0);
        fraction = dist <= T(0) ? T(1) : T(0);
        ok = false;
    } else {
        fraction = D == 2 ? vfrac2(nrm[0], nrm[1], a) : vfrac3(nrm, a);
        ok = true;
    }
    const T frac = T(1) - fraction;
#pragma unroll
    for (int q = 0; q < D; ++q) normals[D * (size_t)c + q] = nrm[q];
    intercepts[c] = a;
    valid[c] = ok ? 1 : 0;
    cols.vf[c] = frac;
    cols.kind[c] = frac == T(0) ? CELL_INSIDE : (frac == T(1) ? CELL_OUTSIDE : CELL_CUT);
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

}  // namespace
}  // namespace cc
