// Device-side building blocks shared by the cut-cell kernels: element-type limits, the body
// descriptor and its signed-distance functions, and the lattice helpers.
//
// Compiled with -ffp-contract=off: the reference's Julia arithmetic fuses nothing (no @fastmath in
// src/), and the watertightness argument (src/marching_squares/marching_squares.jl:17-31) needs the
// same expression, rounded the same way, in both cells that share a face.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cc {

template <typename T> struct Eps;
template <> struct Eps<double> { static constexpr double v = 2.220446049250313e-16; };
template <> struct Eps<float> { static constexpr float v = 1.1920928955078125e-07f; };

enum : int8_t { CELL_INSIDE = 0, CELL_OUTSIDE = 1, CELL_CUT = 2 };  // src/CutCellMethods.jl:56-58

enum { GEO_EMPTY = 0, GEO_SPHERE = 1, GEO_TORUS = 2, GEO_BOX = 3, GEO_PLANE = 4,
       GEO_SPHERE_UNION = 5, GEO_UNION = 6, GEO_BOUNDED = 0x100, GEO_KIND = 0xff };

// Doubles per child of a mixed union: the child's kind, then its own parameters p[0..14].
constexpr int UNION_REC = 16;

// The body as the kernels see it: parameters by value, the primitive table in device memory.
struct Geo {
    int kind;
    int nprims;
    double p[24];
    const double *prims;  // device pointer: nprims x (cx, cy, cz, r), or nprims x UNION_REC
    // a 3D sphere union's float screen records (screen_record about the origin), or null, and the
    // margin 1e-6 R (rounded up) for |x| <= R over the points evaluated
    const float4 *scr = nullptr;
    float scr_m = 0.0f;
};

// The grid as the kernels see it, converted to the cache's element type once on the host
// (update_cache! converts the grid to CartesianGrid{D,T}, src/marching_squares/cache.jl:58).
template <typename T> struct GridT {
    T x0[3];
    T d[3];
    int n[3];
};

template <typename T> __device__ __forceinline__ T sgn(T v) {
    return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0));
}

// _open_fraction, src/marching_squares/marching_squares.jl:192-198: anchored on the outside end.
template <typename T> __device__ __forceinline__ T open_fraction(T a, T b) {
    const bool fa = a >= T(0), fb = b >= T(0);
    if (fa && fb) return T(1);
    if (!fa && !fb) return T(0);
    return fa ? a / (a - b) : b / (b - a);
}

// ||x - c|| - r with the norm accumulated ((dx^2 + dy^2) + dz^2).
template <typename T, int D>
__device__ __forceinline__ T sphere_dist(T cx, T cy, T cz, T r, T x, T y, T z, T *nr_out = nullptr,
                                         T *v = nullptr) {
    T dx = x - cx, dy = y - cy, dz = z - cz;
    T s = dx * dx + dy * dy;
    if (D == 3) s = s + dz * dz;
    T nr = sqrt(s);
    if (nr_out) *nr_out = nr;
    if (v) { v[0] = dx; v[1] = dy; v[2] = dz; }
    return nr - r;
}

// Polynomial smooth minimum: h = max(k - |a-b|, 0)/k; min(a,b) - h*h*k/4.  When k - |a-b| is not
// positive the correction is exactly zero and min(a,b) - 0 is min(a,b) bit for bit (also for -0,
// inf and NaN), so that case returns early and skips the division.
template <typename T> __device__ __forceinline__ T smin_poly(T a, T b, T k) {
    T h = k - fabs(a - b);
    const T m = a < b ? a : b;
    if (!(h > T(0))) return m;
    h = h / k;
    return m - ((h * h) * k) / T(4);
}

// p = R^T (x - c), R row-major 3x3.
template <typename T, int D>
__device__ __forceinline__ void to_local(const double *r, const double *c, const T *x, T *p) {
    T v[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int a = 0; a < D; ++a) v[a] = x[a] - T(c[a]);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        T s = T(r[0 * 3 + i]) * v[0];
#pragma unroll
        for (int j = 1; j < D; ++j) s = s + T(r[j * 3 + i]) * v[j];
        p[i] = s;
    }
}

template <typename T, int D>
__device__ __forceinline__ void from_local(const double *r, const T *gl, T *n) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
        T s = T(r[i * 3 + 0]) * gl[0];
#pragma unroll
        for (int j = 1; j < D; ++j) s = s + T(r[i * 3 + j]) * gl[j];
        n[i] = s;
    }
}

// A single-precision screen for the double fold.  A primitive leaves the fold unchanged when
// fold_sphere's own skip test holds (|x - c| >= acc + k + r up to its margin: the smooth minimum
// then returns acc bit for bit); the screen proves that test true from float data, so a wave whose
// lanes all pass it skips the primitive's double work.  Per primitive (screen_record): the centre
// minus an origin o, rounded to nearest, and w >= (k + r) + 1e-6 (|k + r| + R + |c - o|), where R
// bounds |x - o| over every point folded (R may instead come as the separate margin m >= 1e-6 R).
// For a point x, with xf = fl32(x - o) and u = 2^-24: the float difference vector is within
// 3.7 u (|x - o| + |c - o|) of x - c, so s = |xf - cf|^2 (fused multiply-adds) and
// tp = w + m + accf + 2^-18 |accf| (accf = fl32(acc)) satisfy
// tp > (acc + k + r) + 5.9 u (|acc| + k + r) + 4 u (R + |c - o|) > 0 -- and s >= fl32(tp * tp)
// gives |x - c| > acc + k + r + 3.6 u (|acc| + k + r), far above the skip test's margin (under 40
// double epsilons of the same sum); tp <= 0 gives acc + k + r < 0, where the skip test holds too.
// |tp| >= 1e18 or (positive) below 2^-60 (where tp * tp could underflow), infinities, NaNs, negative
// radii and w outside [1e-18, 1e30] (w = inf) never skip.
__device__ __forceinline__ float4 screen_record(const double *s, const double *o, double k, double R) {
    const double cx = s[0] - o[0], cy = s[1] - o[1], cz = s[2] - o[2];
    const double kr = k + s[3];
    double w = kr + 1e-6 * ((fabs(kr) + R) + sqrt((cx * cx + cy * cy) + cz * cz));
    if (!(s[3] >= 0.0) || !(w < 1e30) || !(w >= 1e-18)) w = INFINITY;
    return make_float4((float)cx, (float)cy, (float)cz, (float)w);
}

__device__ __forceinline__ bool screen_skips(const float4 &q, float xf, float yf, float zf, float accf,
                                             float m = 0.0f) {
    const float dx = xf - q.x, dy = yf - q.y, dz = zf - q.z;
    const float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float tp = fmaf(fabsf(accf), 0x1p-18f, accf + q.w) + m;
    // evaluated whole (no short circuit): three compares and two mask operations, no branches
    return (fabsf(tp) < 1e18f) & ((tp <= 0.0f) | ((tp >= 0x1p-60f) & (s >= tp * tp)));
}

// One step acc <- smin_poly(acc, |x - c| - r, k) of the sphere-union fold.  s is the squared norm
// sphere_dist forms.  When sqrt(s) >= acc + k + r plus a rounding margin (or acc + k + r is below
// minus the margin, so that even the centre is k above acc), the computed distance d satisfies
// d - acc > k, the smooth minimum returns acc bit for bit, and the square root and the smooth
// minimum are skipped.  The margin, 16 eps times the magnitudes involved, covers the rounding of
// t, of t*t, of the square root and of d and acc - d.
template <typename T, int D>
__device__ __forceinline__ T fold_sphere(T acc, T cx, T cy, T cz, T r, T x, T y, T z, T k) {
    const T dx = x - cx, dy = y - cy, dz = z - cz;
    T s = dx * dx + dy * dy;
    if (D == 3) s = s + dz * dz;
    const T t = (acc + k) + r;
    const T tm = t + T(16) * Eps<T>::v * (((fabs(acc) + k) + r) + fabs(t));
    if (tm <= T(0) || s >= tm * tm) return acc;
    return smin_poly(acc, sqrt(s) - r, k);
}

// The sphere union folded left with the polynomial smooth minimum.
template <typename T, int D>
__device__ __forceinline__ T union_dist(const Geo &g, const T *x) {
    const T z = D == 3 ? x[2] : T(0);
    const double *s = g.prims;
    const T k = T(g.p[0]);
    T acc = sphere_dist<T, D>(T(s[0]), T(s[1]), T(s[2]), T(s[3]), x[0], x[1], z);
    if (sizeof(T) == 8 && D == 3 && g.scr) {  // screened: a primitive not skipped folds as the oracle's
        const float xf = (float)x[0], yf = (float)x[1], zf = (float)z;
        float accf = (float)acc;
        for (int i = 1; i < g.nprims; ++i) {
            if (screen_skips(g.scr[i], xf, yf, zf, accf, g.scr_m)) continue;
            const double *si = s + 4 * i;
            acc = smin_poly(acc, sphere_dist<T, D>(T(si[0]), T(si[1]), T(si[2]), T(si[3]), x[0], x[1], z), k);
            accf = (float)acc;
        }
        return acc;
    }
    for (int i = 1; i < g.nprims; ++i) {
        const double *si = s + 4 * i;
        acc = fold_sphere<T, D>(acc, T(si[0]), T(si[1]), T(si[2]), T(si[3]), x[0], x[1], z, k);
    }
    return acc;
}

// SDFBounded (SDFLibrary, un-vendored; its use documented at src/marching_cubes/surface.jl:70-71
// and src/volume.jl:26-27): beyond pad (p[22]) of the child's bounding box (p[16..18] to p[19..21])
// the value is the positive fill (p[23]), with no normal, so those nodes and centroids read as
// outside.  The band is the Euclidean distance to the box, summed over the axes in order and
// compared squared.
template <typename T, int D> __device__ __forceinline__ bool beyond_band(const Geo &g, const T *x) {
    T s = T(0);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const T lo = T(g.p[16 + a]), hi = T(g.p[19 + a]);
        const T e = x[a] < lo ? lo - x[a] : (x[a] > hi ? x[a] - hi : T(0));
        s = s + e * e;
    }
    const T pad = T(g.p[22]);
    return s > pad * pad;
}

// One primitive body's get_sdf -> (distance, normal) from its kind and parameters p (a body's own
// p, or a union child's record after its kind).  The normal is the analytic gradient; it is NaN or
// zero where the distance is not differentiable, which the PLIC fit reports as a degenerate plane
// (src/plic/plic.jl:82-85).  Returns NaN with a zero normal for a kind that is not a primitive.
template <typename T, int D>
__device__ __forceinline__ T primitive_sdf(int kind, const double *pp, const T *x, T *normal) {
    T v[3] = {T(0), T(0), T(0)}, p[3] = {T(0), T(0), T(0)}, gl[3] = {T(0), T(0), T(0)};
    switch (kind) {
    case GEO_SPHERE: {
        T nr;
        T dist = sphere_dist<T, D>(T(pp[0]), T(pp[1]), T(pp[2]), T(pp[3]), x[0], x[1],
                                   D == 3 ? x[2] : T(0), &nr, v);
#pragma unroll
        for (int a = 0; a < D; ++a) normal[a] = v[a] / nr;
        return dist;
    }
    case GEO_TORUS: {  // 3D only
        const double *r = pp + 5;
        to_local<T, 3>(r, pp, x, p);
        T rho = sqrt(p[0] * p[0] + p[1] * p[1]);
        T qx = rho - T(pp[3]), qz = p[2];
        T nq = sqrt(qx * qx + qz * qz);
        // the gradient (qx p_x/rho, qx p_y/rho, qz) / nq over one common denominator: one division
        // where the quotients took four (on the axis or the core circle it is NaN either way)
        const T t = T(1) / (nq * rho);
        gl[0] = (qx * p[0]) * t;
        gl[1] = (qx * p[1]) * t;
        gl[2] = (qz * rho) * t;
        from_local<T, 3>(r, gl, normal);
        return nq - T(pp[4]);
    }
    case GEO_BOX: {
        const double *r = pp + 6;
        T q[3] = {T(0), T(0), T(0)}, mx[3] = {T(0), T(0), T(0)};
        to_local<T, D>(r, pp, x, p);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            q[a] = fabs(p[a]) - T(pp[3 + a]);
            mx[a] = q[a] > T(0) ? q[a] : T(0);
        }
        T o = mx[0] * mx[0];
#pragma unroll
        for (int a = 1; a < D; ++a) o = o + mx[a] * mx[a];
        o = sqrt(o);
        T qm = q[0];
        int am = 0;
#pragma unroll
        for (int a = 1; a < D; ++a)
            if (q[a] > qm) { qm = q[a]; am = a; }
        T inner = qm < T(0) ? qm : T(0);
        if (qm > T(0)) {
#pragma unroll
            for (int a = 0; a < D; ++a) gl[a] = (mx[a] / o) * sgn(p[a]);
        } else {
#pragma unroll
            for (int a = 0; a < D; ++a) gl[a] = a == am ? sgn(p[a]) : T(0);
        }
        from_local<T, D>(r, gl, normal);
        return o + inner;
    }
    case GEO_PLANE: {
        T s = T(pp[0]) * (x[0] - T(pp[3]));
#pragma unroll
        for (int a = 1; a < D; ++a) s = s + T(pp[a]) * (x[a] - T(pp[3 + a]));
#pragma unroll
        for (int a = 0; a < D; ++a) normal[a] = T(pp[a]);
        return s;
    }
    default: break;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) normal[a] = T(0);
    return T(NAN);
}

// The union of mixed primitives (SDFUnion(children...; smoothing), SDFLibrary, un-vendored): the
// children folded left with the polynomial smooth minimum of smoothing k = p[0]; k = 0 is the plain
// minimum (a tie keeps the later child, as min(a, b) = a < b ? a : b does).  The normal is the
// gradient of the fold: where a step takes the plain minimum it is the winner's normal unchanged;
// where it blends (h = (k - |a - b|)/k > 0) it is (1 - h/2) n_min + (h/2) n_max, the derivative of
// min(a, b) - h^2 k/4, and a blended normal is scaled to unit length once at the end.
template <typename T, int D>
__device__ __forceinline__ T mixed_union_sdf(const Geo &g, const T *x, T *normal) {
    const T k = T(g.p[0]);
    T acc = T(0);
    bool blended = false;
    for (int i = 0; i < g.nprims; ++i) {
        const double *ri = g.prims + UNION_REC * i;
        T nb[3] = {T(0), T(0), T(0)};
        const T b = primitive_sdf<T, D>((int)ri[0], ri + 1, x, nb);
        T h = k - fabs(acc - b);
        const bool a_min = acc < b;
        if (i == 0 || !(h > T(0))) {
            if (i == 0 || !a_min) {
                acc = b;
#pragma unroll
                for (int a = 0; a < D; ++a) normal[a] = nb[a];
            }
            continue;
        }
        h = h / k;
        const T wmax = h / T(2), wmin = T(1) - wmax;
#pragma unroll
        for (int a = 0; a < D; ++a)
            normal[a] = a_min ? wmin * normal[a] + wmax * nb[a] : wmax * normal[a] + wmin * nb[a];
        acc = (a_min ? acc : b) - ((h * h) * k) / T(4);
        blended = true;
    }
    if (blended) {
        T s = normal[0] * normal[0];
#pragma unroll
        for (int a = 1; a < D; ++a) s = s + normal[a] * normal[a];
        s = sqrt(s);
#pragma unroll
        for (int a = 0; a < D; ++a) normal[a] = normal[a] / s;
    }
    return acc;
}

// get_sdf(geo, x) -> (distance, normal) (SDFLibrary, un-vendored: the build's own bodies).  MIXED =
// false leaves the mixed union out of the code, for the fused kernels whose register budget it
// would raise; the host never hands them a mixed union (they read NaN for one).
template <typename T, int D, bool MIXED = true>
__device__ __forceinline__ T get_sdf(const Geo &g, const T *x, T *normal) {
    if ((g.kind & GEO_BOUNDED) && beyond_band<T, D>(g, x)) {
#pragma unroll
        for (int a = 0; a < D; ++a) normal[a] = T(0);
        return T(g.p[23]);
    }
    const int kind = g.kind & GEO_KIND;
    if (kind >= GEO_SPHERE && kind <= GEO_PLANE) return primitive_sdf<T, D>(kind, g.p, x, normal);
    if constexpr (MIXED) {
        if (kind == GEO_UNION) return mixed_union_sdf<T, D>(g, x, normal);
    }
#pragma unroll
    for (int a = 0; a < D; ++a) normal[a] = T(0);
    if (kind == GEO_EMPTY) return T(g.p[0]);
    if (kind == GEO_SPHERE_UNION) return union_dist<T, D>(g, x);  // no analytic normal offered
    return T(NAN);
}

// sdf_value(geo, x) (SDFLibrary, un-vendored: the build's own bodies).
template <typename T, int D, bool MIXED = true>
__device__ __forceinline__ T sdf_value(const Geo &g, const T *x) {
    const T z = D == 3 ? x[2] : T(0);
    if ((g.kind & GEO_BOUNDED) && beyond_band<T, D>(g, x)) return T(g.p[23]);
    switch (g.kind & GEO_KIND) {
    case GEO_EMPTY: return T(g.p[0]);
    case GEO_SPHERE:
        return sphere_dist<T, D>(T(g.p[0]), T(g.p[1]), T(g.p[2]), T(g.p[3]), x[0], x[1], z);
    case GEO_SPHERE_UNION: return union_dist<T, D>(g, x);
    default: break;
    }
    T nrm[3];
    return get_sdf<T, D, MIXED>(g, x, nrm);
}

}  // namespace cc
