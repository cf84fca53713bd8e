"""The reference's marching-squares moment tests that go through a cache
(test/test_moments.jl:286-659), on the device through the C ABI: convergence of the apertures and of
the volume against closed forms, the interface facet recomputed independently from the cache's own
corner values, exact closure, the report of a banded tree, children of a refined leaf cut afresh, a
plain field, the contour against the moments and on its own domain only, and face centroids
single-valued on an off-lattice grid.

Bar: the reference's own thresholds (rates, tolerances, counts), and bitwise equality where the
reference asks `===` or `==`.
"""
import math

import numpy as np
import pytest

from cutcell import (AdaptiveMesh, CartesianGrid, CartesianIsoGrid, DimensionMismatch, MarchingSquaresCutCell,
                     SDFSphere, Unsupported, allocate_cache, cut_cell_report, generate_mesh, readers,
                     update_cache)
from helpers import GEO_CENTRE, GEO_RADIUS, banded_tree, circle_sdf

pytestmark = pytest.mark.gpu

GEO = SDFSphere(GEO_CENTRE, GEO_RADIUS)  # test/test_moments.jl:30-32


def pair_rates(e):  # test/test_moments.jl:105-106
    return [math.log2(e[j] / e[j + 1]) for j in range(len(e) - 1)]


def overall_rate(e):
    return math.log2(e[0] / e[-1]) / (len(e) - 1)


def uniform(n):
    """n x n cells over [-2, 2]^2 (test/test_moments.jl:61-64), as a grid: a tree at level 0 over the
    same cells is the same construction (checked in test_gpu_leaves.py)."""
    return CartesianGrid((-2.0, -2.0), (n, n), (4.0 / n, 4.0 / n))


def circle_cache(g, geo=GEO):
    return update_cache(allocate_cache(g, MarchingSquaresCutCell), geo, g)


def test_open_face_apertures_against_the_exact_chord():
    # test/test_moments.jl:286-312: the covered length of each vertical grid line within 0.8 r of
    # the centre against the chord 2 sqrt(r^2 - dx^2)
    def chord_error(n):
        g = uniform(n)
        h = 4.0 / n
        ff = circle_cache(g).read("face_fraction").reshape(n, n, 4)  # [j, i, direction]
        covered = (ff[:, :, 0] * h).sum(axis=0)  # the -x face of every cell of column i
        worst = 0.0
        for ix in range(n):
            dx = (-2.0 + ix * h) - GEO_CENTRE[0]
            if abs(dx) >= 0.8 * GEO_RADIUS:
                continue
            chord = 2 * math.sqrt(GEO_RADIUS ** 2 - dx ** 2)
            worst = max(worst, abs((4.0 - covered[ix]) - chord))
        return worst

    e = [chord_error(n) for n in (64, 128, 256, 512)]
    print("chord errors", e)
    assert e[-1] < 1e-4
    assert overall_rate(e) > 1.8
    assert all(r > 1.4 for r in pair_rates(e))


def test_per_cell_conservation_and_the_facet_recomputed_independently():
    # test/test_moments.jl:468-478: closure exactly zero on the uniform 64 x 64 cells
    g = uniform(64)
    c = circle_cache(g)
    ff = c.read("face_fraction")
    assert np.count_nonzero(readers.closure_residual(ff, g.d)) == 0
    assert int((c.read("kind") == 2).sum()) > 50
    # :481-524: the segment between the edge crossings of the cache's own corner values, computed
    # here, against the normal and area the closure gives, and pointing out of the disc
    g = uniform(128)
    c = circle_cache(g)
    n0, n1 = g.n
    kind, amb = c.read("kind"), c.read("ambiguous")
    ff = c.read("face_fraction")
    ic = c.read("interface_centroid")
    phi = c.phi.reshape(n1 + 1, n0 + 1)  # [j, i]
    sel = np.flatnonzero((kind == 2) & (amb == 0))
    nrm = readers.interface_normal(ff[sel], g.d)
    area = readers.interface_area(ff[sel], g.d)
    checked, worst_angle, worst_area, worst_sense = 0, 0.0, 0.0, math.inf
    for r, cell in enumerate(sel):
        i, j = cell % n0, cell // n0
        corners = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]  # MS_NODE_BITS order
        nodes = [np.array([g.x0[0] + a * g.d[0], g.x0[1] + b * g.d[1]]) for a, b in corners]
        p = [phi[b, a] for a, b in corners]
        pts = []
        for k in range(4):
            k2 = (k + 1) % 4
            if (p[k] >= 0) != (p[k2] >= 0):
                t = p[k] / (p[k] - p[k2])
                pts.append(nodes[k] + t * (nodes[k2] - nodes[k]))
        if len(pts) != 2:
            continue
        v = pts[1] - pts[0]
        nA = np.array([v[1], -v[0]])
        ln = float(np.hypot(*nA))
        if ln <= 1e-13:
            continue
        worst_angle = max(worst_angle, abs(1 - abs(float(nA @ nrm[r]) / ln)))
        worst_area = max(worst_area, abs(ln - float(area[r])) / ln)
        worst_sense = min(worst_sense, float(nrm[r] @ (ic[cell] - np.array(GEO_CENTRE))))
        checked += 1
    assert checked > 100
    assert worst_angle < 1e-12 and worst_area < 1e-11
    assert worst_sense > 0  # every facet normal points out of the body


def test_refinement_consistency():
    # test/test_moments.jl:537-551: the volume from n x n and from 2n x 2n cells converge at second
    # order
    def volume(n):
        g = uniform(n)
        return float((circle_cache(g).read("volume_fraction") * (g.d[0] * g.d[1])).sum())

    gaps = [abs(volume(n) - volume(2 * n)) for n in (32, 64, 128, 256)]
    print("parent-child gaps", gaps)
    assert overall_rate(gaps) > 1.8
    assert all(r > 1.4 for r in pair_rates(gaps))
    assert gaps[-1] < 5e-3


def test_the_report_surfaces_the_smallest_cell_of_a_banded_tree():
    # test/test_moments.jl:587-594
    tree = banded_tree()
    c = update_cache(allocate_cache(tree, MarchingSquaresCutCell), GEO, tree)
    r = cut_cell_report(c)
    assert r["n"] == tree.nleaves
    assert r["cut"] + r["inside"] + r["outside"] == r["n"]
    assert r["ambiguous"] == 0
    assert 0 < r["min_volume_fraction"] < 0.05


def test_a_field_that_is_no_geometry_and_the_contour_against_the_moments():
    # test/test_moments.jl:635-640: a plain field y - 0.25
    g = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    c = update_cache(allocate_cache(g, MarchingSquaresCutCell), lambda x: x[1] - 0.25, g)
    vf = c.read("volume_fraction")
    assert np.all((vf >= 0) & (vf <= 1))
    assert int((c.read("kind") == 2).sum()) == g.n[0]
    # :652-659: one contour segment per cut cell, no ambiguous cell, every vertex on the circle
    c = circle_cache(g)
    nodes, lines = generate_mesh(c, g)
    assert int((c.read("kind") == 2).sum()) == len(lines)
    assert not np.any(c.read("ambiguous"))
    assert np.all(np.abs(circle_sdf(nodes, GEO_CENTRE, GEO_RADIUS)) < 2e-3)


def test_face_centroids_are_single_valued_on_an_off_lattice_grid():
    # test/test_moments.jl:343-364: the global face centroid of every interior face from both of its
    # cells, bitwise, uncut outside cells beside cut ones included
    g = CartesianGrid((-1.3, 0.7), (17, 23), ((2.1 + 1.3) / 17, (3.9 - 0.7) / 23))
    c = update_cache(allocate_cache(g, MarchingSquaresCutCell),
                     lambda x: math.sin(3 * x[0]) * math.cos(2 * x[1]) + 0.1, g)
    n0, n1 = g.n
    fcl = c.read("face_centroid_local")  # (ncell, 4, 1)
    kind = c.read("kind")
    ii, jj = np.meshgrid(np.arange(n0), np.arange(n1))  # [j, i]
    idx = np.stack([ii.ravel(), jj.ravel()], axis=1)  # cell order: i fastest
    x0, d = np.array(g.x0), np.array(g.d)
    lower = x0 + idx * d
    upper = x0 + (idx + 1) * d
    mismatches = outside_by_cut = 0
    for ax in range(2):
        step = 1 if ax == 0 else n0
        inner = (idx[:, ax] + 1 < g.n[ax])
        ci = np.flatnonzero(inner)
        cj = ci + step
        a = readers.face_centroid(fcl[ci], 2 * ax + 1, lower[ci], upper[ci])
        b = readers.face_centroid(fcl[cj], 2 * ax, lower[cj], upper[cj])
        mismatches += int(np.any(a.view(np.uint64) != b.view(np.uint64), axis=1).sum())
        outside_by_cut += int((((kind[ci] == 1) & (kind[cj] == 2)) | ((kind[ci] == 2) & (kind[cj] == 1))).sum())
    assert outside_by_cut > 50  # the case the test exists for is exercised
    assert mismatches == 0


def test_children_are_recomputed_never_inherited():
    # test/test_moments.jl:556-585: one cut leaf of a uniform 32 x 32 tree refined; its four children
    # are cut afresh from the body, not copied from the parent
    mesh = AdaptiveMesh(uniform(32), max_level=1, initial_level=0)
    c = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), GEO, mesh)
    kind, vf = c.read("kind"), c.read("volume_fraction")
    i = int(np.flatnonzero((kind == 2) & (vf < 0.35))[0])
    parent_frac = float(vf[i])
    h = mesh.elem_size(i)
    parent_vol = parent_frac * h[0] * h[1]
    pcoord = mesh.coord[i].copy()
    flags = np.zeros(mesh.nleaves, dtype=bool)
    flags[i] = True
    mesh.refine(flags)
    c = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), GEO, mesh)
    kids = np.flatnonzero((mesh.level == 1) & np.all(mesh.coord // 2 == pcoord, axis=1))
    assert kids.size == 4
    vf = c.read("volume_fraction")
    child = vf[kids]
    assert not np.allclose(child, parent_frac)  # not a copied fraction
    assert np.any(child == 0)                   # at least one wholly inside
    hk = mesh.elem_size(kids[0])
    assert abs(float(child.sum()) * hk[0] * hk[1] - parent_vol) <= 0.2 * parent_vol
    ff = c.read("face_fraction")[kids]
    assert np.count_nonzero(readers.closure_residual(ff, tuple(hk))) == 0


def test_a_caches_contour_is_asked_for_on_its_own_domain():
    # test/test_moments.jl:623-633: another grid, another tree, or a tree for a grid cache is refused
    g = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    c = circle_cache(g)
    with pytest.raises(DimensionMismatch):
        generate_mesh(c, CartesianGrid(g.x0, tuple(n + 1 for n in g.n), g.d))
    tree = banded_tree()
    tc = update_cache(allocate_cache(tree, MarchingSquaresCutCell), GEO, tree)
    with pytest.raises(DimensionMismatch):
        generate_mesh(tc, banded_tree(level=2))
    with pytest.raises(Unsupported):
        generate_mesh(c, tree)
