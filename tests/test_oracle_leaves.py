"""Pins the oracle's marching-squares leaf cache (src/marching_squares/cache.jl:129-137) against what
the reference's tests and the grid path fix: a leaf coinciding with a grid cell is that cell bit for
bit (cache.jl:83-89); a corner two leaves share across a level jump is one sample
(test/test_moments.jl:383-410); same-level neighbours agree bitwise on their shared face
(:318-340); every leaf closes exactly (test/test_cache.jl:137-138)."""
import numpy as np

import oracle as orc
from cutcell import CartesianGrid, SDFCircle, readers
from cutcell.tree import AdaptiveMesh
from helpers import (GEO_CENTRE, GEO_RADIUS, banded_tree, cache_tree, circle_sdf, fortran,
                     same_level_neighbours)

COLS = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
        "face_centroid_local", "interface_centroid")


def leaves_of(mesh, geo):
    return orc.ms_leaves(geo, mesh.base.x0, mesh.base.d, mesh.level, mesh.coord)


def test_a_uniform_tree_is_the_grid():
    base = CartesianGrid((-2.0, -2.0), (8, 8), (0.5, 0.5))
    mesh = AdaptiveMesh(base, max_level=2, initial_level=2)
    geo = SDFCircle(GEO_CENTRE, GEO_RADIUS)
    phi4, cols = leaves_of(mesh, geo)
    n = (32, 32)
    g = CartesianGrid((-2.0, -2.0), n, (0.125, 0.125))
    phi = orc.sample_nodes(geo, 2, g.x0, g.d, g.n)
    ref = orc.ms_cells(phi, g.x0, g.d, g.n)
    lin = mesh.coord[:, 0] + n[0] * mesh.coord[:, 1]
    for name in COLS:
        a, b = cols[name], ref[name][lin]
        assert np.array_equal(a.view(np.uint8), np.ascontiguousarray(b).view(np.uint8)), name
    P = fortran(phi, (33, 33))
    i, j = mesh.coord[:, 0], mesh.coord[:, 1]
    want = np.stack([P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]], axis=1)
    assert np.array_equal(phi4.view(np.uint64), want.view(np.uint64))
    assert (cols["kind"] == 2).sum() > 50


def test_corner_samples_and_moments():
    mesh = banded_tree()
    assert len(set(mesh.level.tolist())) > 1
    geo = SDFCircle(GEO_CENTRE, GEO_RADIUS)
    phi4, cols = leaves_of(mesh, geo)
    # each leaf's values are the body at its lattice corners (cell_values(field, mesh, c))
    assert np.array_equal(phi4, circle_sdf(mesh.cell_nodes(), GEO_CENTRE, GEO_RADIUS))
    # a corner shared across a level jump is one sample
    L = int(mesh.level.max())
    seen, mism, levels = {}, 0, {}
    bits = [(0, 0), (1, 0), (1, 1), (0, 1)]
    for c in range(mesh.nleaves):
        s = 2 ** (L - int(mesh.level[c]))
        for k, b in enumerate(bits):
            key = ((int(mesh.coord[c, 0]) + b[0]) * s, (int(mesh.coord[c, 1]) + b[1]) * s)
            v = phi4[c, k]
            if key in seen:
                mism += seen[key] != v
                levels[key].add(int(mesh.level[c]))
            else:
                seen[key] = v
                levels[key] = {int(mesh.level[c])}
    assert mism == 0
    assert sum(len(v) > 1 for v in levels.values()) > 100
    # every leaf closes exactly with its own size; same-level faces are single-valued
    h = mesh.elem_size()
    clo = np.array([readers.closure_residual(cols["face_fraction"][c:c + 1], tuple(h[c]))[0]
                    for c in range(mesh.nleaves)])
    assert np.count_nonzero(clo) == 0
    pairs = same_level_neighbours(mesh)
    assert len(pairs) > 1000
    ff = cols["face_fraction"]
    assert all(ff[a, 2 * ax + 1] == ff[b, 2 * ax] for a, b, ax in pairs)
    # one contour segment per cut leaf here: nothing is ambiguous
    assert (cols["kind"] == 2).sum() > 100 and not cols["ambiguous"].any()


def test_cache_tree_sizes():
    tree = cache_tree()
    assert len(set(tree.level.tolist())) > 1
    phi4, cols = leaves_of(tree, SDFCircle((0.013, -0.021), 0.83))
    assert (cols["kind"] == 2).sum() > 20
    cut_levels = set(tree.level[cols["kind"] == 2].tolist())
    assert len(cut_levels) >= 1
