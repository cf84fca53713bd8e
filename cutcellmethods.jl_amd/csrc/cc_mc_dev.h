// Device building blocks of the marching-cubes passes, shared by the single-cache kernels and the
// batched launch: the Lewiner dispatch and per-cell construction (src/marching_cubes/marching_cubes.jl,
// src/marching_cubes/moments.jl), the tile staging and column streaming, and one work item of the
// z-marching stream pass and of the cut-cell pass.  Included by exactly one place per translation
// unit; every symbol is internal to that unit.
#pragma once
#include "cc_brick_dev.h"
#include "cc_internal.h"

#define CC_LUT_QUAL __constant__
#include "lewiner_tables.h"

namespace cc {
namespace {


__device__ __forceinline__ int nx4(int k) { return (k + 1) & 3; }  // _ms_next
__device__ __forceinline__ int pv4(int k) { return (k + 3) & 3; }  // _ms_prev

// MC_NODE_BITS (src/marching_cubes/marching_cubes.jl:106-107), 0-based corner c.
__device__ __forceinline__ constexpr int bitx(int c) { return (c ^ (c >> 1)) & 1; }
__device__ __forceinline__ constexpr int bity(int c) { return (c >> 1) & 1; }
__device__ __forceinline__ constexpr int bitz(int c) { return (c >> 2) & 1; }

// MC_FACE_NODES (:169-180), MC_FACE_TEST_CODE (:385-388), MC_TEST_FACE_QUADS (:372-373) and
// MC_EDGE_NODES (:135-137), 0-based.  constexpr look-ups: in the unrolled per-face and per-edge
// loops every index folds to a constant, so the corner values stay in registers.
__device__ __forceinline__ constexpr int face_node(int dir, int k) {
    constexpr int8_t t[6][4] = {{0, 3, 7, 4}, {1, 2, 6, 5}, {0, 1, 5, 4},
                                {3, 2, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
    return t[dir][k];
}
__device__ __forceinline__ constexpr int face_test_code(int dir) {
    constexpr int8_t t[6] = {4, 2, 1, 3, 5, 6};
    return t[dir];
}
__device__ __forceinline__ constexpr int test_quad(int f, int k) {
    constexpr int8_t t[6][4] = {{0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3},
                                {3, 7, 4, 0}, {0, 3, 2, 1}, {4, 7, 6, 5}};
    return t[f][k];
}
__device__ __forceinline__ constexpr int edge_node(int e, int end) {
    constexpr int8_t t[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                                 {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    return t[e][end];
}

// _test_face (:401-406), the asymptotic decider on one face's four corner values, for all six
// faces and both signs of the face code at once: bit f-1 of `pos` / `neg` is the result for face
// code +f / -f.  The table-driven tests of the dispatch then read a bit instead of indexing the
// corner values at run time (which would move them out of registers).  Each bit is the reference's
// expression: |AC - BD| < eps ? (code >= 0) : code * A * (AC - BD) >= 0.
struct FaceTests {
    unsigned pos, neg;
};

template <int F, typename T> __device__ __forceinline__ void face_test_bits(const T *c, FaceTests &ft) {
    const T A = c[test_quad(F - 1, 0)], B = c[test_quad(F - 1, 1)];
    const T C = c[test_quad(F - 1, 2)], D = c[test_quad(F - 1, 3)];
    const T det = A * C - B * D;
    const bool small = fabs(det) < Eps<T>::v;
    const bool p = small || T(F) * A * det >= T(0);
    const bool n = !small && T(-F) * A * det >= T(0);
    ft.pos |= (p ? 1u : 0u) << (F - 1);
    ft.neg |= (n ? 1u : 0u) << (F - 1);
}

template <typename T> __device__ __forceinline__ FaceTests face_tests(const T *c) {
    FaceTests ft{0u, 0u};
    face_test_bits<1>(c, ft);
    face_test_bits<2>(c, ft);
    face_test_bits<3>(c, ft);
    face_test_bits<4>(c, ft);
    face_test_bits<5>(c, ft);
    face_test_bits<6>(c, ft);
    return ft;
}

__device__ __forceinline__ bool test_face(const FaceTests &ft, int face) {
    const int f = face < 0 ? -face : face;
    return (((face >= 0 ? ft.pos : ft.neg) >> (f - 1)) & 1u) != 0u;
}

// t on reference edge (E0, E1) and the values at t on its three parallel edges (:466-489).
template <int E0, int E1, int B0, int B1, int C0, int C1, int D0, int D1, typename T>
__device__ __forceinline__ void interior_edge(const T *c, T &t, T &Bt, T &Ct, T &Dt) {
    t = c[E0] / (c[E0] - c[E1]);
    Bt = c[B0] + (c[B1] - c[B0]) * t;
    Ct = c[C0] + (c[C1] - c[C0]) * t;
    Dt = c[D0] + (c[D1] - c[D0]) * t;
}

// _test_interior (:427-504); `case1` is the reference's 1-based case, `subcfg` 1-based.
template <typename T>
__device__ __forceinline__ bool test_interior(int case1, const T *c, int cfg, int subcfg, int s) {
    T At = T(0), Bt = T(0), Ct = T(0), Dt = T(0), t;
    if (case1 == 5 || case1 == 11) {
        T a = (c[4] - c[0]) * (c[6] - c[2]) - (c[7] - c[3]) * (c[5] - c[1]);
        T b = c[2] * (c[4] - c[0]) + c[0] * (c[6] - c[2]) - c[1] * (c[7] - c[3]) -
              c[3] * (c[5] - c[1]);
        t = -b / (T(2) * a);
        if (t < T(0) || t > T(1)) return s > 0;
        At = c[0] + (c[4] - c[0]) * t;
        Bt = c[3] + (c[7] - c[3]) * t;
        Ct = c[2] + (c[6] - c[2]) * t;
        Dt = c[1] + (c[5] - c[1]) * t;
    } else {
        int edge;
        if (case1 == 7) edge = LW_TEST6[cfg][2];
        else if (case1 == 8) edge = LW_TEST7[cfg][4];
        else if (case1 == 13) edge = LW_TEST12[cfg][3];
        else edge = LW_TILING13_5_1[cfg][subcfg - 1][0];
        // Reference edge e: t on that edge, then B/C/D on the three parallel edges.
        switch (edge) {
        case 0:  interior_edge<0, 1, 3, 2, 7, 6, 4, 5>(c, t, Bt, Ct, Dt); break;
        case 1:  interior_edge<1, 2, 0, 3, 4, 7, 5, 6>(c, t, Bt, Ct, Dt); break;
        case 2:  interior_edge<2, 3, 1, 0, 5, 4, 6, 7>(c, t, Bt, Ct, Dt); break;
        case 3:  interior_edge<3, 0, 2, 1, 6, 5, 7, 4>(c, t, Bt, Ct, Dt); break;
        case 4:  interior_edge<4, 5, 7, 6, 3, 2, 0, 1>(c, t, Bt, Ct, Dt); break;
        case 5:  interior_edge<5, 6, 4, 7, 0, 3, 1, 2>(c, t, Bt, Ct, Dt); break;
        case 6:  interior_edge<6, 7, 5, 4, 1, 0, 2, 3>(c, t, Bt, Ct, Dt); break;
        case 7:  interior_edge<7, 4, 6, 5, 2, 1, 3, 0>(c, t, Bt, Ct, Dt); break;
        case 8:  interior_edge<0, 4, 3, 7, 2, 6, 1, 5>(c, t, Bt, Ct, Dt); break;
        case 9:  interior_edge<1, 5, 0, 4, 3, 7, 2, 6>(c, t, Bt, Ct, Dt); break;
        case 10: interior_edge<2, 6, 1, 5, 0, 4, 3, 7>(c, t, Bt, Ct, Dt); break;
        default: interior_edge<3, 7, 2, 6, 1, 5, 0, 4>(c, t, Bt, Ct, Dt); break;
        }
    }
    int test = 0;
    if (At >= T(0)) test += 1;
    if (Bt >= T(0)) test += 2;
    if (Ct >= T(0)) test += 4;
    if (Dt >= T(0)) test += 8;
    if (test == 6 || test == 8 || test == 9 || test == 12 || (test >= 0 && test <= 4)) return s > 0;
    if (test == 5) return (At * Ct - Bt * Dt < Eps<T>::v) ? s > 0 : s < 0;
    if (test == 10) return (At * Ct - Bt * Dt >= Eps<T>::v) ? s > 0 : s < 0;
    return s < 0;
}

// cell_triangles (:538-651): the Lewiner dispatch.  Returns the triangle count (-1 for a
// SUBCONFIG13 entry with no tiling) and points *tile at the flat run of 0-based edge codes.
template <typename T>
__device__ __forceinline__ int cell_triangles(const T *phi, const FaceTests &ft, int lut,
                                              const signed char **tile) {
    const int cs = LW_CASES[lut][0], cfg = LW_CASES[lut][1];
    int subcfg = 1;
    switch (cs) {
    case 0: *tile = nullptr; return 0;
    case 1: *tile = LW_TILING1[cfg]; return 1;
    case 2: *tile = LW_TILING2[cfg]; return 2;
    case 3:
        if (test_face(ft, LW_TEST3[cfg])) { *tile = LW_TILING3_2[cfg]; return 4; }
        *tile = LW_TILING3_1[cfg]; return 2;
    case 4:
        if (test_interior(5, phi, cfg, subcfg, LW_TEST4[cfg])) { *tile = LW_TILING4_1[cfg]; return 2; }
        *tile = LW_TILING4_2[cfg]; return 6;
    case 5: *tile = LW_TILING5[cfg]; return 3;
    case 6:
        if (test_face(ft, LW_TEST6[cfg][0])) { *tile = LW_TILING6_2[cfg]; return 5; }
        if (test_interior(7, phi, cfg, subcfg, LW_TEST6[cfg][1])) { *tile = LW_TILING6_1_1[cfg]; return 3; }
        *tile = LW_TILING6_1_2[cfg]; return 9;
    case 7:
        if (test_face(ft, LW_TEST7[cfg][0])) subcfg += 1;
        if (test_face(ft, LW_TEST7[cfg][1])) subcfg += 2;
        if (test_face(ft, LW_TEST7[cfg][2])) subcfg += 4;
        switch (subcfg) {
        case 1: *tile = LW_TILING7_1[cfg]; return 3;
        case 2: case 3: *tile = LW_TILING7_2[cfg][subcfg - 2]; return 5;
        case 4: *tile = LW_TILING7_3[cfg][0]; return 9;
        case 5: *tile = LW_TILING7_2[cfg][2]; return 5;
        case 6: case 7: *tile = LW_TILING7_3[cfg][subcfg - 5]; return 9;
        default:
            if (test_interior(8, phi, cfg, subcfg, LW_TEST7[cfg][3])) { *tile = LW_TILING7_4_2[cfg]; return 9; }
            *tile = LW_TILING7_4_1[cfg]; return 5;
        }
    case 8: *tile = LW_TILING8[cfg]; return 2;
    case 9: *tile = LW_TILING9[cfg]; return 4;
    case 10:
        if (test_face(ft, LW_TEST10[cfg][0])) {
            if (test_face(ft, LW_TEST10[cfg][1])) { *tile = LW_TILING10_1_1_[cfg]; return 4; }
            *tile = LW_TILING10_2[cfg]; return 8;
        }
        if (test_face(ft, LW_TEST10[cfg][1])) { *tile = LW_TILING10_2_[cfg]; return 8; }
        if (test_interior(11, phi, cfg, subcfg, LW_TEST10[cfg][2])) { *tile = LW_TILING10_1_1[cfg]; return 4; }
        *tile = LW_TILING10_1_2[cfg]; return 8;
    case 11: *tile = LW_TILING11[cfg]; return 4;
    case 12:
        if (test_face(ft, LW_TEST12[cfg][0])) {
            if (test_face(ft, LW_TEST12[cfg][1])) { *tile = LW_TILING12_1_1_[cfg]; return 4; }
            *tile = LW_TILING12_2[cfg]; return 8;
        }
        if (test_face(ft, LW_TEST12[cfg][1])) { *tile = LW_TILING12_2_[cfg]; return 8; }
        if (test_interior(13, phi, cfg, subcfg, LW_TEST12[cfg][2])) { *tile = LW_TILING12_1_1[cfg]; return 4; }
        *tile = LW_TILING12_1_2[cfg]; return 8;
    case 13: {
        for (int f = 0; f < 6; ++f)
            if (test_face(ft, LW_TEST13[cfg][f])) subcfg += 1 << f;
        const int sc = LW_SUBCONFIG13[subcfg - 1];
        if (sc < 0) { *tile = nullptr; return -1; }
        if (sc == 0) { *tile = LW_TILING13_1[cfg]; return 4; }
        if (sc <= 6) { *tile = LW_TILING13_2[cfg][sc - 1]; return 6; }
        if (sc <= 18) { *tile = LW_TILING13_3[cfg][sc - 7]; return 10; }
        if (sc <= 22) { *tile = LW_TILING13_4[cfg][sc - 19]; return 12; }
        if (sc <= 26) {
            const int sub = sc - 22;
            // test13[cfg][6] in the reference's 1-based spelling (:631-637)
            if (test_interior(14, phi, cfg, sub, LW_TEST13[cfg][5])) { *tile = LW_TILING13_5_1[cfg][sub - 1]; return 6; }
            *tile = LW_TILING13_5_2[cfg][sub - 1]; return 10;
        }
        if (sc <= 38) { *tile = LW_TILING13_3_[cfg][sc - 27]; return 10; }
        if (sc <= 44) { *tile = LW_TILING13_2_[cfg][sc - 39]; return 6; }
        *tile = LW_TILING13_1_[cfg]; return 4;
    }
    default: *tile = LW_TILING14[cfg]; return 4;
    }
}

// _face_polygon (src/marching_cubes/moments.jl:111-139) on a face's (p, q) frame: the boundary
// walk (src/marching_squares/marching_squares.jl:240-251) with outside-anchored open segments,
// then the chords (_chord_target, moments.jl:91-98), summed as directed edges.
template <typename T>
__device__ __forceinline__ void face_polygon(const T *fx, const T *fy, const T *fp, bool connect,
                                             T &area2, T &c0, T &c1) {
    T sax[4], say[4], sbx[4], sby[4];
    bool pres[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = nx4(k);
        const T fa = fp[k], fb = fp[j];
        const T frac = open_fraction(fa, fb);
        pres[k] = fa >= T(0) || fb >= T(0);
        if (fa >= T(0) && fb >= T(0)) {
            sax[k] = fx[k]; say[k] = fy[k]; sbx[k] = fx[j]; sby[k] = fy[j];
        } else if (fa >= T(0)) {
            sax[k] = fx[k]; say[k] = fy[k];
            sbx[k] = fx[k] + frac * (fx[j] - fx[k]);
            sby[k] = fy[k] + frac * (fy[j] - fy[k]);
        } else {
            sax[k] = fx[j] + frac * (fx[k] - fx[j]);
            say[k] = fy[j] + frac * (fy[k] - fy[j]);
            sbx[k] = fx[j]; sby[k] = fy[j];
        }
    }
    T a2 = T(0), m0 = T(0), m1 = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!pres[k]) continue;
        const T cr = sax[k] * sby[k] - say[k] * sbx[k];
        a2 += cr;
        m0 += (sax[k] + sbx[k]) * cr;
        m1 += (say[k] + sby[k]) * cr;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!(fp[k] >= T(0) && fp[nx4(k)] < T(0))) continue;
        // the chord target: the first entering edge met walking on (connect) or back from k; both
        // walks are unrolled so that every index is a constant
        bool found = false;
        T lbx = T(0), lby = T(0);
        if (connect) {
#pragma unroll
            for (int it = 1; it <= 4; ++it) {
                const int j = (k + it) & 3;
                if (!found && fp[j] < T(0) && fp[nx4(j)] >= T(0)) { found = true; lbx = sax[j]; lby = say[j]; }
            }
        } else {
#pragma unroll
            for (int it = 1; it <= 4; ++it) {
                const int j = (k + 4 - it) & 3;
                if (!found && fp[j] < T(0) && fp[nx4(j)] >= T(0)) { found = true; lbx = sax[j]; lby = say[j]; }
            }
        }
        if (!found) continue;
        const T lax = sbx[k], lay = sby[k];
        const T cr = lax * lby - lay * lbx;
        a2 += cr;
        m0 += (lax + lbx) * cr;
        m1 += (lay + lby) * cr;
    }
    area2 = a2;
    c0 = m0;
    c1 = m1;
}

template <typename T> __device__ __forceinline__ T full_face_area3(const T *cs, int dir) {
    const int ax = dir >> 1;
    T p = T(1);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (a != ax) p *= cs[a];
    return p;
}

// One face of cut_cell_faces (moments.jl:224-257): its clip in the face's (p, q) frame, the face
// fraction and the centroid offsets, and the saddle flag.  DIR is a template parameter so that
// every corner and table index folds to a constant and the corner values stay in registers.
template <int DIR, typename T>
__device__ __forceinline__ void face_clip(const T *phi, const FaceTests &ft, const T *L, const T *cs, T *ff,
                                         T *fcl, bool &amb_face) {
    constexpr int ax = DIR >> 1;
    constexpr int p = ax == 0 ? 1 : 0, q = ax == 2 ? 1 : 2;
    T fx[4], fy[4], fp[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int nd = face_node(DIR, k);
        const int bp = p == 0 ? bitx(nd) : bity(nd);
        const int bq = q == 1 ? bity(nd) : bitz(nd);
        fx[k] = bp ? L[p] : T(0);
        fy[k] = bq ? L[q] : T(0);
        fp[k] = phi[nd];
    }
    int nout = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) nout += fp[k] >= T(0) ? 1 : 0;
    amb_face |= nout == 2 && ((fp[0] >= T(0)) == (fp[2] >= T(0)));
    const T avg0 = T(0.25) * (((fx[0] + fx[1]) + fx[2]) + fx[3]);
    const T avg1 = T(0.25) * (((fy[0] + fy[1]) + fy[2]) + fy[3]);
    if (nout == 4 || nout == 0) {
        ff[DIR] = nout == 4 ? T(1) : T(0);
        fcl[2 * DIR] = avg0;
        fcl[2 * DIR + 1] = avg1;
    } else {
        T a2, m0, m1;
        face_polygon(fx, fy, fp, test_face(ft, face_test_code(DIR)), a2, m0, m1);
        ff[DIR] = T(0.5) * a2 / full_face_area3(cs, DIR);
        if (a2 != T(0)) {
            const T den = T(3) * a2;
            fcl[2 * DIR] = m0 / den;
            fcl[2 * DIR + 1] = m1 / den;
        } else {
            fcl[2 * DIR] = avg0;
            fcl[2 * DIR + 1] = avg1;
        }
    }
}

// The axis edge E runs along (edges 0-7 alternate x and y, 8-11 run along z).
__device__ __forceinline__ constexpr int edge_axis(int e) { return e < 8 ? (e & 1) : 2; }

// Crossing of edge E (edge_crossings, marching_cubes.jl:306-330), outside-anchored.  Its coordinate
// along the edge goes to the lane's LDS slot E (zero where the edge is not crossed, which sets no
// bit of `cmask`); on the two other axes the point is pa + s (pb - pa) with pa = pb there, which
// cut_point restores.  acc and ncr gather the interior vertex.
template <int E, typename T>
__device__ __forceinline__ void edge_point(const T *phi, const T *L, T *P, T *acc, int &ncr, unsigned &cmask) {
    constexpr int a = edge_node(E, 0), b = edge_node(E, 1);
    const bool crossed = (phi[a] >= T(0)) != (phi[b] >= T(0));
    const T pa[3] = {bitx(a) ? L[0] : T(0), bity(a) ? L[1] : T(0), bitz(a) ? L[2] : T(0)};
    const T pb[3] = {bitx(b) ? L[0] : T(0), bity(b) ? L[1] : T(0), bitz(b) ? L[2] : T(0)};
    T q[3] = {T(0), T(0), T(0)};
    if (crossed) {
        const T s = open_fraction(phi[a], phi[b]);
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            q[x] = phi[a] >= T(0) ? pa[x] + s * (pb[x] - pa[x]) : pb[x] + s * (pa[x] - pb[x]);
            acc[x] += q[x];
        }
        ncr += 1;
        cmask |= 1u << E;
    }
    P[E * MC_CUT_THREADS] = q[edge_axis(E)];
}

// Point e of a cut cell (0-11 an edge crossing, 12 the interior vertex) from the lane's slots as
// edge_point and mc_cut_cell left them.  On a crossed edge's two other axes the point is the edge's
// corner coordinate c (0 or L) plus s * 0 -- c itself for a finite fraction s -- and (q - q) * 0
// gives the same for every q the slot can hold (NaN or infinite q, from a NaN or infinite s, give
// NaN there as s * 0 does); an edge not crossed is the zero point.
template <typename T>
__device__ __forceinline__ void cut_point(int e, const T *P, const T *L, unsigned crossed, T *v) {
    if (e == 12) {
#pragma unroll
        for (int x = 0; x < 3; ++x) v[x] = P[(12 + x) * MC_CUT_THREADS];
        return;
    }
    if (!((crossed >> e) & 1u)) {
        v[0] = v[1] = v[2] = T(0);
        return;
    }
    const T q = P[e * MC_CUT_THREADS];
    const T z = (q - q) * T(0);
    const int a = (int)((0x321047540310ull >> (4 * e)) & 15u);  // edge_node(e, 0), four bits per edge
    const int ax = edge_axis(e);
    v[0] = ax == 0 ? q : (bitx(a) ? L[0] : T(0)) + z;
    v[1] = ax == 1 ? q : (bity(a) ? L[1] : T(0)) + z;
    v[2] = ax == 2 ? q : (bitz(a) ? L[2] : T(0)) + z;
}

// The whole construction for one cut cell: cut_cell_moments (moments.jl:275-346) with
// cut_cell_faces (:224-257) and edge_crossings (marching_cubes.jl:306-330).  `phi` is nudged;
// `L` the cell-local upper corner (node(i+1) - node(i) per axis), `xl` the lower lattice corner.
template <typename T>
__device__ __forceinline__ void mc_cut_cell(const T *phi, int lut, const T *xl, const T *L, const T *cs,
                            const Cols<T> &cols, uint32_t c, T *P) {
    // ---- six face clips (one instantiation per face: every table index is a constant) ----
    const FaceTests ft = face_tests(phi);
    T ff[6], fcl[12];
    bool amb_face = false;
    face_clip<0>(phi, ft, L, cs, ff, fcl, amb_face);
    face_clip<1>(phi, ft, L, cs, ff, fcl, amb_face);
    face_clip<2>(phi, ft, L, cs, ff, fcl, amb_face);
    face_clip<3>(phi, ft, L, cs, ff, fcl, amb_face);
    face_clip<4>(phi, ft, L, cs, ff, fcl, amb_face);
    face_clip<5>(phi, ft, L, cs, ff, fcl, amb_face);

    // ---- edge crossings on cell-local corners ----
    // The tiling codes index the 13 points at run time, so they live in this lane's LDS slots
    // P[slot * MC_CUT_THREADS] (lane-major: consecutive lanes, consecutive banks), not in a private
    // array (which would go to scratch): MC_CUT_SLOTS values, read back by cut_point.
    T acc[3] = {T(0), T(0), T(0)};
    int ncr = 0;
    unsigned crossed = 0;
    edge_point<0>(phi, L, P, acc, ncr, crossed);
    edge_point<1>(phi, L, P, acc, ncr, crossed);
    edge_point<2>(phi, L, P, acc, ncr, crossed);
    edge_point<3>(phi, L, P, acc, ncr, crossed);
    edge_point<4>(phi, L, P, acc, ncr, crossed);
    edge_point<5>(phi, L, P, acc, ncr, crossed);
    edge_point<6>(phi, L, P, acc, ncr, crossed);
    edge_point<7>(phi, L, P, acc, ncr, crossed);
    edge_point<8>(phi, L, P, acc, ncr, crossed);
    edge_point<9>(phi, L, P, acc, ncr, crossed);
    edge_point<10>(phi, L, P, acc, ncr, crossed);
    edge_point<11>(phi, L, P, acc, ncr, crossed);
#pragma unroll
    for (int x = 0; x < 3; ++x) P[(12 + x) * MC_CUT_THREADS] = ncr > 0 ? acc[x] / T(ncr) : T(0);

    // ---- the closed polyhedron, by the divergence theorem ----
    T vol3 = T(0), cm[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
        const int ax = dir >> 1;
        const T s = (dir & 1) ? T(1) : T(-1);
        const T x_ax = (dir & 1) ? L[ax] : T(0);  // local coordinate of the face's first corner
        const T a_open = ff[dir] * full_face_area3(cs, dir);
        vol3 += s * x_ax * a_open;
#pragma unroll
        for (int a = 0; a < 3; ++a) cm[a] += a == ax ? s * x_ax * x_ax * a_open : T(0);
    }
    const signed char *tile;
    int ntri = cell_triangles(phi, ft, lut, &tile);
    if (ntri < 0) ntri = 0;
    T iarea = T(0), icm[3] = {T(0), T(0), T(0)};
    for (int t = 0; t < ntri; ++t) {
        const int ia = tile[3 * t], ib = tile[3 * t + 1], id = tile[3 * t + 2];
        T a[3], b[3], d[3];
        cut_point(ia, P, L, crossed, a);
        cut_point(ib, P, L, crossed, b);
        cut_point(id, P, L, crossed, d);
        T u[3], v[3], w[3], xc[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) { u[x] = b[x] - a[x]; v[x] = d[x] - a[x]; }
        w[0] = T(0.5) * (u[1] * v[2] - u[2] * v[1]);
        w[1] = T(0.5) * (u[2] * v[0] - u[0] * v[2]);
        w[2] = T(0.5) * (u[0] * v[1] - u[1] * v[0]);
#pragma unroll
        for (int x = 0; x < 3; ++x) xc[x] = (a[x] + b[x] + d[x]) / T(3);
        vol3 -= w[0] * xc[0] + w[1] * xc[1] + w[2] * xc[2];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const T m2 = (a[x] * a[x] + b[x] * b[x] + d[x] * d[x] + a[x] * b[x] + b[x] * d[x] +
                          d[x] * a[x]) / T(6);
            cm[x] -= w[x] * m2;
        }
        const T len = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        iarea += len;
#pragma unroll
        for (int x = 0; x < 3; ++x) icm[x] += len * xc[x];
    }
    const T volume = vol3 / T(3);
    const int cs0 = LW_CASES[lut][0];
    const bool amb = amb_face || cs0 == 4 || cs0 == 6 || cs0 == 7 || cs0 == 10 || cs0 == 12 ||
                     cs0 == 13;
    const T vol_cell = (cs[0] * cs[1]) * cs[2];

    cols.kind[c] = CELL_CUT;
    cols.amb[c] = amb ? 1 : 0;
    cols.vf[c] = volume / vol_cell;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T centre = xl[a] + T(0.5) * cs[a];
        cols.centroid[3 * (size_t)c + a] = volume > T(0) ? xl[a] + cm[a] / (T(2) * volume) : centre;
        cols.ic[3 * (size_t)c + a] = iarea > T(0) ? xl[a] + icm[a] / iarea : centre;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) cols.ff[6 * (size_t)c + k] = ff[k];
#pragma unroll
    for (int k = 0; k < 12; ++k) cols.fcl[12 * (size_t)c + k] = fcl[k];
}

template <typename T> __device__ __forceinline__ T nudge(T v) {
    return fabs(v) < Eps<T>::v ? Eps<T>::v : v;  // nudge_zeros, marching_cubes.jl:240-243
}

// The 8 corner values of cell (i, j, k) in MC_NODE_BITS order, nudged, and their sign mask
// (cell_case's lut - 1, :343-349).
template <typename T>
__device__ __forceinline__ int load_corners(const T *__restrict__ phi, size_t base, size_t sy,
                                            size_t sz, T *ph) {
    ph[0] = nudge(phi[base]);
    ph[1] = nudge(phi[base + 1]);
    ph[2] = nudge(phi[base + sy + 1]);
    ph[3] = nudge(phi[base + sy]);
    ph[4] = nudge(phi[base + sz]);
    ph[5] = nudge(phi[base + sz + 1]);
    ph[6] = nudge(phi[base + sz + sy + 1]);
    ph[7] = nudge(phi[base + sz + sy]);
    int lut = 0;
#pragma unroll
    for (int p = 0; p < 8; ++p) lut |= (ph[p] >= T(0) ? 1 : 0) << p;
    return lut;
}

template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double __attribute__((ext_vector_type(2))) type; };
template <> struct Vec16<float> { typedef float __attribute__((ext_vector_type(4))) type; };

typedef int __attribute__((ext_vector_type(4))) i32x4;

template <bool NT, typename V> __device__ __forceinline__ void put(V *p, V v) {
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// One column of the tile, [tc cells x M components], written from `f(q)` (q = flat element index
// within the tile) with 16-byte lane-contiguous stores, then a scalar tail.
template <typename T, int M, bool NT, typename F>
__device__ __forceinline__ void stream_column(T *__restrict__ dst, int tc, F f) {
    constexpr int V = 16 / sizeof(T);
    typedef typename Vec16<T>::type vec;
    const int total = tc * M;
    const int nfull = total / V;
    for (int e = threadIdx.x; e < nfull; e += blockDim.x) {
        vec v;
#pragma unroll
        for (int u = 0; u < V; ++u) v[u] = f(e * V + u);
        put<NT>(reinterpret_cast<vec *>(dst + (size_t)e * V), v);
    }
    for (int q = nfull * V + threadIdx.x; q < total; q += blockDim.x) put<NT>(dst + q, f(q));
}

// Per-tile staging in LDS: each cell's kind, centre, and half cell-local extents.
template <typename T> struct TileLDS {
    __attribute__((aligned(16))) int8_t kind[MC_TILE];
    T cen[MC_TILE * 3];
    T hl[MC_TILE * 3];
};

// Classify one cell from its nudged corner mask and stage its closed-form uncut moments.
template <typename T>
__device__ __forceinline__ bool stage_cell(TileLDS<T> &s, int t, int lut, uint32_t i, uint32_t j,
                                           uint32_t k, const GridT<T> &g) {
    const int8_t kind = lut == 255 ? CELL_OUTSIDE : (lut == 0 ? CELL_INSIDE : CELL_CUT);
    s.kind[t] = kind;
    const uint32_t ijk[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T xl = g.x0[a] + T(ijk[a]) * g.d[a];
        const T xh = g.x0[a] + T(ijk[a] + 1) * g.d[a];
        const T L = xh - xl;
        s.cen[3 * t + a] = xl + T(0.5) * g.d[a];
        // 0.25 * (((0 + L) + L) + 0): the uncut face's corner average in its (p, q) frame
        s.hl[3 * t + a] = T(0.25) * (L + L);
    }
    return kind == CELL_CUT;
}

// Append this wave's cut cells to the work list: one atomic per wave.  Called by every lane.
__device__ __forceinline__ void append_cut(bool cut, uint32_t c, uint32_t *cutlist, uint32_t *counter) {
    const uint64_t m = __ballot(cut);
    if (m) {
        const int lane = __lane_id();
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t b = 0;
        if (lane == leader) b = atomicAdd(counter, (uint32_t)__popcll(m));
        b = __shfl(b, leader);
        if (cut) cutlist[b + __popcll(m & ((1ull << lane) - 1))] = c;
    }
}

// Every column of cells [c0, c0 + tc) from the staged tile.  Cut cells get placeholder values
// here; k_mc_cut overwrites them afterwards on the same stream.
template <typename T, bool NT>
__device__ __forceinline__ void stream_tile(const Cols<T> &cols, uint32_t c0, int tc,
                                            const TileLDS<T> &s) {
    const int t = threadIdx.x;
    {
        int8_t *dk = cols.kind + c0;
        uint8_t *da = cols.amb + c0;
        const int nfull = tc / 16;
        for (int e = t; e < nfull; e += blockDim.x) {
            put<NT>(reinterpret_cast<i32x4 *>(dk + 16 * e), *reinterpret_cast<const i32x4 *>(s.kind + 16 * e));
            put<NT>(reinterpret_cast<i32x4 *>(da + 16 * e), i32x4{0, 0, 0, 0});
        }
        for (int q = nfull * 16 + t; q < tc; q += blockDim.x) {
            dk[q] = s.kind[q];
            da[q] = 0;
        }
    }
    stream_column<T, 1, NT>(cols.vf + c0, tc, [&](int q) { return s.kind[q] == CELL_OUTSIDE ? T(1) : T(0); });
    stream_column<T, 3, NT>(cols.centroid + 3 * (size_t)c0, tc, [&](int q) { return s.cen[q]; });
    stream_column<T, 6, NT>(cols.ff + 6 * (size_t)c0, tc,
                            [&](int q) { return s.kind[q / 6] == CELL_OUTSIDE ? T(1) : T(0); });
    // face_centroid_local: per direction the (p, q) half extents; axes (y,z),(x,z),(x,y).
    stream_column<T, 12, NT>(cols.fcl + 12 * (size_t)c0, tc, [&](int q) {
        const int cell = q / 12, r = q - 12 * cell, dir = r >> 1, w = r & 1;
        const int ax = dir >> 1;
        const int a = w == 0 ? (ax == 0 ? 1 : 0) : (ax == 2 ? 1 : 2);
        return s.hl[3 * cell + a];
    });
    stream_column<T, 3, NT>(cols.ic + 3 * (size_t)c0, tc, [&](int q) { return s.cen[q]; });
}


// One column of a tile of rc rows of wc cells, the rows rowstride cells apart: [rc x wc cells x M
// components] from `f(q)` (q = flat element index within the tile, rows in order), 16-byte
// lane-contiguous stores within each row when a row holds whole vectors.
template <typename T, int M, bool NT, typename F>
__device__ __forceinline__ void stream_column_rows(T *__restrict__ dst, int wc, int rc, size_t rowstride, F f) {
    constexpr int V = 16 / sizeof(T);
    typedef typename Vec16<T>::type vec;
    const int per = wc * M;
    if (per % V == 0) {
        const int vper = per / V, nv = vper * rc;
        for (int e = threadIdx.x; e < nv; e += blockDim.x) {
            const int r = e / vper, ve = e - r * vper;
            vec v;
#pragma unroll
            for (int u = 0; u < V; ++u) v[u] = f(r * per + ve * V + u);
            put<NT>(reinterpret_cast<vec *>(dst + (size_t)r * rowstride * M + (size_t)ve * V), v);
        }
    } else {
        for (int q = threadIdx.x; q < per * rc; q += blockDim.x) {
            const int r = q / per;
            put<NT>(dst + (size_t)r * rowstride * M + (q - r * per), f(q));
        }
    }
}

// Every column of a tile of rc rows of wc cells starting at cell c0, the rows n0 cells apart (a
// brick-shaped tile), from the staged tile.  Cut cells get placeholder values, as in stream_tile.
template <typename T, bool NT>
__device__ __forceinline__ void stream_tile_rows(const Cols<T> &cols, uint32_t c0, int wc, int rc, uint32_t n0,
                                                 const TileLDS<T> &s) {
    const int t = threadIdx.x;
    {
        int8_t *dk = cols.kind + c0;
        uint8_t *da = cols.amb + c0;
        if (wc % 16 == 0) {
            const int vper = wc / 16;
            for (int e = t; e < vper * rc; e += blockDim.x) {
                const int r = e / vper, ve = e - r * vper;
                const size_t o = (size_t)r * n0 + 16 * ve;
                put<NT>(reinterpret_cast<i32x4 *>(dk + o), *reinterpret_cast<const i32x4 *>(s.kind + r * wc + 16 * ve));
                put<NT>(reinterpret_cast<i32x4 *>(da + o), i32x4{0, 0, 0, 0});
            }
        } else {
            for (int q = t; q < wc * rc; q += blockDim.x) {
                const int r = q / wc;
                const size_t o = (size_t)r * n0 + (q - r * wc);
                dk[o] = s.kind[q];
                da[o] = 0;
            }
        }
    }
    stream_column_rows<T, 1, NT>(cols.vf + c0, wc, rc, n0, [&](int q) { return s.kind[q] == CELL_OUTSIDE ? T(1) : T(0); });
    stream_column_rows<T, 3, NT>(cols.centroid + 3 * (size_t)c0, wc, rc, n0, [&](int q) { return s.cen[q]; });
    stream_column_rows<T, 6, NT>(cols.ff + 6 * (size_t)c0, wc, rc, n0,
                                 [&](int q) { return s.kind[q / 6] == CELL_OUTSIDE ? T(1) : T(0); });
    stream_column_rows<T, 12, NT>(cols.fcl + 12 * (size_t)c0, wc, rc, n0, [&](int q) {
        const int cell = q / 12, r = q - 12 * cell, dir = r >> 1, w = r & 1;
        const int ax = dir >> 1;
        const int a = w == 0 ? (ax == 0 ? 1 : 0) : (ax == 2 ? 1 : 2);
        return s.hl[3 * cell + a];
    });
    stream_column_rows<T, 3, NT>(cols.ic + 3 * (size_t)c0, wc, rc, n0, [&](int q) { return s.cen[q]; });
}

// A marching-cubes cache as the kernels see it: one body of a batch, or the one cache of an update.
template <typename T> struct McBody {
    GridT<T> g;
    Geo geo;
    T *phi;
    Cols<T> cols;
    uint32_t *cutlist;
    uint32_t *counter;
    uint32_t W, R, kz;        // tile: W cells along x, R rows along y (R > 1 only when W == n0), kz layers
    uint32_t nseg, nrg, nkb;  // tiles along x, row groups along y, layer blocks along z
    uint32_t fused;           // sample the body in-kernel (phi written) or read phi
    uint32_t nontemporal;
    BrickPhi bp;              // a sphere union's culled or surrogate phi (the BRICKS items)
    const uint64_t *signs;    // reading a field: its nodes' sides from a phi pass's bitmap, not phi
    int sbx, sby;             // that bitmap's bricks along x and y
};

constexpr int MARCH_NODES = 2 * (MC_TILE + 1);  // max (R+1)(W+1) over R*W <= MC_TILE
constexpr int MARCH_SPH = 256;  // kept primitives a brick-aligned item stages in LDS (more: read from memory)
constexpr int MARCH_BX = MC_TILE / PHI_BRICK;  // bricks along x a brick-aligned item may span

// The kept lists of the bricks a brick-shaped item spans along x (in LDS): count, list, and the LDS
// copy of the primitives (null when they did not fit).
template <typename T> struct BrickLists {
    uint32_t nk[MARCH_BX];
    const uint16_t *L[MARCH_BX];
    T *sph[MARCH_BX];
};

// Cut cells of up to CUT_STEPS z-steps of one work item, held as the waves' ballot masks and
// appended to the work list with ONE atomic per flush: a single global counter taking one atomic
// per wave and step serialises in L2 once bodies have millions of cut cells.
constexpr int CUT_STEPS = 32;
struct CutMasks {
    uint64_t m[CUT_STEPS][MC_TILE / 64];   // ballot of each wave at each step
    uint32_t off[CUT_STEPS][MC_TILE / 64]; // exclusive prefix of their popcounts
    uint32_t c0[CUT_STEPS];                // first cell of each step's tile
    uint32_t base;
};

template <typename T> struct MarchLDS {
    TileLDS<T> s;
    T rows[2][MARCH_NODES];  // [layer parity][(R+1) x (W+1) nodes, row-major]
    CutMasks cm;
    uint32_t st[5];          // BRICKS: owned nodes, exact, uncertified, band, contradicted
};

// Appends the cut cells recorded in cm (steps [0, ns)) to the work list.  Called by every thread
// of the workgroup after a barrier; ends with one.  Cell of thread t at step s: c0[s] + (t / wc) n0
// + t % wc.
__device__ __forceinline__ void flush_cuts(CutMasks &cm, int ns, int wc, uint32_t n0, uint32_t *cutlist,
                                           uint32_t *counter) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int NW = MC_TILE / 64;
    if (t == 0) {
        uint32_t acc = 0;
        for (int q = 0; q < ns; ++q)
            for (int w = 0; w < NW; ++w) {
                cm.off[q][w] = acc;
                acc += (uint32_t)__popcll(cm.m[q][w]);
            }
        cm.base = acc ? atomicAdd(counter, acc) : 0u;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t rel = (uint32_t)(t / wc) * n0 + (uint32_t)(t % wc);
    for (int q = 0; q < ns; ++q) {
        const uint64_t m = cm.m[q][wave];
        if ((m >> lane) & 1ull) cutlist[cm.base + cm.off[q][wave] + __popcll(m & below)] = cm.c0[q] + rel;
    }
    __syncthreads();
}

// One work item of the z-marching stream pass: the R x W cells of rows [j0, j0+R) and columns
// [i0, i0+W), marched over the layers [k0, k0+kz).  The node rows of the layer below stay in LDS, so
// each step brings in only the layer above: computed from the body (FUSED: owned nodes written to
// phi exactly once, phi never re-read; BRICKS: a sphere union's culled or surrogate phi from its
// bricks, with the counters of the owned nodes written to B.bp.wstat[w]) or read from phi.  The
// cells of one step are one contiguous range of every column (R > 1 only for whole rows).
template <typename T, bool FUSED, bool NT, bool BRICKS = false, bool SIGNS = false>
__device__ void mc_march_item(const McBody<T> &B, uint32_t w, MarchLDS<T> &L, T *sph = nullptr,
                              BrickLists<T> *bl = nullptr) {
    const GridT<T> &g = B.g;
    const uint32_t n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const uint32_t seg = w % B.nseg, rest = w / B.nseg, rg = rest % B.nrg, kb = rest / B.nrg;
    const uint32_t i0 = seg * B.W, j0 = rg * B.R;
    const int wc = (int)min(B.W, n0 - i0), rc = (int)min(B.R, n1 - j0);
    const int tc = wc * rc, rowlen = wc + 1, nn = (rc + 1) * rowlen;
    const uint32_t k0 = kb * B.kz, k1 = min(k0 + B.kz, n2);
    const int t = threadIdx.x;
    const size_t sy = n0 + 1, sz = sy * (n1 + 1);

    // BRICKS: the item spans BX = ceil(wc / PHI_BRICK) bricks along x and one brick along y and z
    // (i0 a multiple of PHI_BRICK, j0 and k0 multiples of R and kz, which divide it), so the kept
    // list of brick bx, exact over its box extended by one node, serves every node of the item with
    // (gi - i0) / PHI_BRICK == bx (the last brick also the halo node i0 + wc); the lists are staged in
    // LDS one after another while they fit in MARCH_SPH primitives, the rest read from memory
    int nbxi = 1;
    if (BRICKS) {
        nbxi = (wc + PHI_BRICK - 1) / PHI_BRICK;
        const uint32_t brow = (uint32_t)B.bp.nbx * (j0 / PHI_BRICK + (uint32_t)B.bp.nby * (k0 / PHI_BRICK));
        if (t < nbxi) {
            const uint32_t b = i0 / PHI_BRICK + t + brow;
            bl->nk[t] = B.bp.cnt[b];
            bl->L[t] = B.bp.list + (size_t)b * B.bp.nprims;
        }
        __syncthreads();
        if (t == 0) {  // the lists one after another in the LDS budget; those past it stay in memory
            uint32_t off = 0;
            for (int bx = 0; bx < nbxi; ++bx) {
                const bool fits = off + bl->nk[bx] <= (uint32_t)MARCH_SPH;
                bl->sph[bx] = fits ? sph + 4 * (size_t)off : nullptr;
                off += fits ? bl->nk[bx] : 0u;
            }
        }
        __syncthreads();
        for (int bx = 0; bx < nbxi; ++bx) {
            T *dst = bl->sph[bx];
            if (!dst) continue;
            for (uint32_t e = t; e < bl->nk[bx]; e += blockDim.x) {
                const double *q = B.geo.prims + 4 * (size_t)bl->L[bx][e];
                dst[4 * e] = T(q[0]);
                dst[4 * e + 1] = T(q[1]);
                dst[4 * e + 2] = T(q[2]);
                dst[4 * e + 3] = T(q[3]);
            }
        }
        __syncthreads();
    }
    uint32_t st_own = 0, st_exact = 0, st_unc = 0, st_band = 0, st_sign = 0;  // BRICKS, owned nodes
    auto fill = [&](int par, uint32_t k) {
        const bool own_layer = k < k1 || k == n2;
        for (int v = t; v < nn; v += blockDim.x) {
            const int r = v / rowlen, ii = v - r * rowlen;
            const uint32_t gi = i0 + ii, gj = j0 + r;
            const size_t idx = gi + sy * gj + sz * k;
            T val;
            if (FUSED) {
                const T x[3] = {g.x0[0] + T(gi) * g.d[0], g.x0[1] + T(gj) * g.d[1],
                                g.x0[2] + T(k) * g.d[2]};
                int how = NODE_EXACT;
                if (BRICKS) {
                    int bx = ii / PHI_BRICK;
                    bx = bx < nbxi ? bx : nbxi - 1;
                    val = brick_node_phi<T>(B.bp, B.geo, gi, gj, k, x, &how, bl->sph[bx], bl->nk[bx], bl->L[bx]);
                } else {
                    val = sdf_value<T, 3, false>(B.geo, x);
                }
                const bool own = own_layer && (r < rc || gj == n1) && (ii < wc || gi == n0);
                if (own) {
                    B.phi[idx] = val;
                    if (BRICKS) {
                        st_own += 1u;
                        st_exact += how != NODE_NET ? 1u : 0u;
                        st_unc += how == NODE_EXACT && B.bp.cert ? 1u : 0u;
                        st_band += how == NODE_BAND ? 1u : 0u;
                        st_sign += how == NODE_CONTRADICTED ? 1u : 0u;
                    }
                }
            } else if (SIGNS) {  // the side is all the cells of this pass read (+-1 keeps it)
                val = sign_bit(B.signs, B.sbx, B.sby, gi, gj, k) ? T(1) : T(-1);
            } else {
                val = B.phi[idx];
            }
            L.rows[par][v] = nudge(val);
        }
    };
    fill(0, k0);
    int par = 0, ns = 0;
    for (uint32_t k = k0; k < k1; ++k) {
        fill(par ^ 1, k + 1);
        __syncthreads();
        bool cut = false;
        const uint32_t c0 = i0 + n0 * (j0 + n1 * k);
        if (t < tc) {
            const int r = t / wc, ii = t - r * wc, b = r * rowlen + ii;
            const T *lo = L.rows[par], *hi = L.rows[par ^ 1];
            // MC_NODE_BITS order
            const T ph[8] = {lo[b], lo[b + 1], lo[b + rowlen + 1], lo[b + rowlen],
                             hi[b], hi[b + 1], hi[b + rowlen + 1], hi[b + rowlen]};
            int lut = 0;
#pragma unroll
            for (int p = 0; p < 8; ++p) lut |= (ph[p] >= T(0) ? 1 : 0) << p;
            cut = stage_cell(L.s, t, lut, i0 + ii, j0 + r, k, g);
        }
        const uint64_t m = __ballot(cut);
        if ((t & 63) == 0) L.cm.m[ns][t >> 6] = m;
        if (t == 0) L.cm.c0[ns] = c0;
        ++ns;
        __syncthreads();
        if (BRICKS) stream_tile_rows<T, NT>(B.cols, c0, wc, rc, n0, L.s);
        else stream_tile<T, NT>(B.cols, c0, tc, L.s);
        __syncthreads();
        if (ns == CUT_STEPS || k + 1 == k1) {
            flush_cuts(L.cm, ns, wc, n0, B.cutlist, B.counter);
            ns = 0;
        }
        par ^= 1;
    }
    if (BRICKS) {  // the item's counters, laid out as the phi pass's per-brick ones
        if (t < 5) L.st[t] = 0u;
        __syncthreads();
        const uint32_t v5[5] = {st_own, st_exact, st_unc, st_band, st_sign};
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            uint32_t v = v5[q];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if ((t & 63) == 0 && v) atomicAdd(&L.st[q], v);
        }
        __syncthreads();
        if (t < PHI_BSTAT) {
            const uint32_t rec[PHI_BSTAT] = {L.st[0], L.st[1], 0u, L.st[2], L.st[3], L.st[4], 0u, 0u};
            B.bp.wstat[(size_t)w * PHI_BSTAT + t] = rec[t];
        }
    }
}

// One entry of the cut-cell work list (if it exists): the full construction for that cell.  P is
// the calling lane's slot in a workgroup array of MC_CUT_SLOTS * MC_CUT_THREADS T in LDS.
template <typename T>
__device__ __forceinline__ void mc_cut_item(const T *__restrict__ phi, const GridT<T> &g,
                                            const Cols<T> &cols, uint32_t c, T *P) {
    const uint32_t n0 = g.n[0], n1 = g.n[1];
    const size_t sy = n0 + 1, sz = sy * (n1 + 1);
    const uint32_t i = c % n0, r = c / n0, j = r % n1, k = r / n1;
    T ph[8];
    const int lut = load_corners(phi, i + sy * j + sz * k, sy, sz, ph);
    const uint32_t ijk[3] = {i, j, k};
    T xl[3], L[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        xl[a] = g.x0[a] + T(ijk[a]) * g.d[a];
        L[a] = (g.x0[a] + T(ijk[a] + 1) * g.d[a]) - xl[a];
    }
    mc_cut_cell(ph, lut, xl, L, g.d, cols, c, P);
}

}  // namespace
}  // namespace cc
