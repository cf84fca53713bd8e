"""Brute force on tiny inputs: the oracle's marching-cubes moments of random single cubes against a
sub-cell quadrature of the reconstructed polyhedron itself (the Lewiner triangles and the cube's
faces), computed here from nothing but the corner signs, the edge crossings and the triangles.

Face fractions: a point of a face is fluid when its straight path within the face to the face's
u = 1 edge crosses the surface's boundary segments on that face an even number of times and the
edge is fluid where the path meets it (the edge's state follows its corner signs and its crossing).
Volume fraction: vertical rays from the top face's state, toggled at every triangle they cross, each
ray's fluid length integrated over a jittered grid of rays.  Neither uses the oracle's face clips or
its divergence-theorem sums, so a wrong face walk, a dropped or misplaced triangle, or a sign slip
in the moments shows as a difference far above the quadrature's own error.  Cubes with a facet lying
in a face plane (the flush configurations of test/test_mc_moments.jl:480-494) make rays graze a
face and are skipped; they are counted and stay few.
"""
import numpy as np

import oracle as orc
from helpers import MC_BITS, MC_EDGE, cube_nodes

EPS = np.finfo(float).eps


def nudged(phi):
    phi = np.asarray(phi, float)
    return np.where(np.abs(phi) < EPS, EPS, phi)


def surface(phi):
    """Triangles (k, 3, 3) of the cube's reconstruction and its boundary segments (a, b points)."""
    n, codes = orc.mc_cell_triangles(phi)
    pts = orc.mc_edge_crossings(cube_nodes(), phi)
    tris = pts[codes] if n > 0 else np.zeros((0, 3, 3))
    count = {}
    for t in codes:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            key = (min(a, b), max(a, b))
            count[key] = count.get(key, 0) + 1
    bnd = [k for k, v in count.items() if v == 1]
    return tris, codes, pts, bnd


def face_state(phi, pts, bnd, ax, side, u, v):
    """Fluid (True) at the points (u, v) of face (ax, side): u, v the other two axes in order."""
    p, q = [a for a in range(3) if a != ax]
    fl = nudged(phi) > 0
    # the face edge at u = 1: corners with bit[ax] = side and bit[p] = 1, running along q
    c0 = next(c for c in range(8) if MC_BITS[c][ax] == side and MC_BITS[c][p] == 1 and MC_BITS[c][q] == 0)
    c1 = next(c for c in range(8) if MC_BITS[c][ax] == side and MC_BITS[c][p] == 1 and MC_BITS[c][q] == 1)
    e = next(k for k, (a, b) in enumerate(MC_EDGE) if {a, b} == {c0, c1})
    if fl[c0] == fl[c1]:
        edge = np.full(u.shape, fl[c0])
    else:
        edge = np.where(v < pts[e][q], fl[c0], fl[c1])
    on_face = {k for k, (a, b) in enumerate(MC_EDGE) if MC_BITS[a][ax] == side and MC_BITS[b][ax] == side}
    par = np.zeros(u.shape, bool)
    for a, b in bnd:
        if a in on_face and b in on_face:
            (ua, va), (ub, vb) = (pts[a][p], pts[a][q]), (pts[b][p], pts[b][q])
            if va == vb:
                continue
            t = (v - va) / (vb - va)
            hit = (t >= 0) & (t < 1) & (ua + t * (ub - ua) > u)
            par ^= hit
    return edge ^ par


def ray_volume(phi, tris, pts, bnd, N, rng):
    x, y = np.meshgrid((np.arange(N) + rng.uniform(0.1, 0.9, N)[:, None].T) / N,
                       (np.arange(N) + rng.uniform(0.1, 0.9, N)[:, None]) / N)
    x, y = x.ravel(), y.ravel()
    state = face_state(phi, pts, bnd, 2, 1, x, y)  # at z = 1
    hits = []
    for A, B, C in tris:
        det = (B[0] - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (B[1] - A[1])
        if det == 0:
            continue
        l1 = ((x - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (y - A[1])) / det
        l2 = ((B[0] - A[0]) * (y - A[1]) - (x - A[0]) * (B[1] - A[1])) / det
        inside = (l1 >= 0) & (l2 >= 0) & (l1 + l2 <= 1)
        z = A[2] + l1 * (B[2] - A[2]) + l2 * (C[2] - A[2])
        hits.append(np.where(inside, z, np.nan))
    zs = np.sort(np.array(hits), axis=0) if hits else np.full((0, x.size), np.nan)
    length = np.zeros(x.size)
    top = np.ones(x.size)
    for z in zs[::-1]:  # descending: NaN (no hit) sorts last, so it comes first here
        ok = ~np.isnan(z)
        seg = np.where(ok, top - z, 0.0)
        length += np.where(state & ok, seg, 0.0)
        state = np.where(ok, ~state, state)
        top = np.where(ok, z, top)
    length += np.where(state, top, 0.0)
    return length.mean()


def test_random_cubes_against_sub_cell_quadrature():
    rng = np.random.default_rng(20261204)
    N = 64
    checked = skipped = 0
    worst_v = worst_f = 0.0
    while checked < 40:
        phi = rng.standard_normal(8) * rng.choice([1.0, 0.1, 3.0], 8)
        fl = nudged(phi) > 0
        if fl.all() or (~fl).all():
            continue
        tris, codes, pts, bnd = surface(phi)
        flush = any(len({a for a in range(3) if np.all(t[:, a] == t[0, a]) and t[0, a] in (0.0, 1.0)}) for t in tris)
        if flush:
            skipped += 1
            continue
        m = orc.mc_cell(cube_nodes(), phi, (1.0, 1.0, 1.0))
        vf = ray_volume(phi, tris, pts, bnd, N, rng)
        worst_v = max(worst_v, abs(vf - m["volume_fraction"]))
        u, v = np.meshgrid((np.arange(2 * N) + 0.5) / (2 * N), (np.arange(2 * N) + 0.5) / (2 * N))
        u, v = u.ravel(), v.ravel()
        for ax in range(3):
            for side in (0, 1):
                f = face_state(phi, pts, bnd, ax, side, u, v).mean()
                worst_f = max(worst_f, abs(f - m["face_fraction"][2 * ax + side]))
        checked += 1
    assert skipped <= 10
    assert worst_v < 5e-3, worst_v  # the quadrature's own error is about 1e-3 at 64 x 64 rays
    assert worst_f < 5e-3, worst_f


def test_the_quadrature_sees_a_planted_slip():
    # the same quadrature over the surface with one triangle dropped misses the moments by far more
    # than its own error: it is not blind to the triangles it is given
    rng = np.random.default_rng(7)
    while True:
        phi = rng.standard_normal(8)
        fl = nudged(phi) > 0
        if not (fl.all() or (~fl).all()):
            break
    tris, codes, pts, bnd = surface(phi)
    m = orc.mc_cell(cube_nodes(), phi, (1.0, 1.0, 1.0))
    assert abs(ray_volume(phi, tris, pts, bnd, 64, rng) - m["volume_fraction"]) < 5e-3
    assert abs(ray_volume(phi, tris[1:], pts, bnd, 64, rng) - m["volume_fraction"]) > 2e-2
