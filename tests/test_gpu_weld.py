"""The marching-squares contour welded on the device (cc_contour.hip) against the host weld and the
tests' host restatements of generate_mesh (src/marching_squares/surface.jl:59-110, 153-177): the
same nodes and lines, bit for bit and in the same order, for grids and quadtrees, Float64 and
Float32, fields with exact zeros at nodes (-0 and +0 welded as one vertex), fields with saddles
(two segments in a cell) and a tree of about a million leaves."""
import numpy as np
import pytest

import oracle as orc
from cutcell import (AdaptiveMesh, CartesianGrid, CartesianIsoGrid, MarchingSquaresCutCell, SDFCircle,
                     allocate_cache, generate_mesh, update_cache)
from helpers import banded_tree
from refrng import XorShift, randn_field
from test_gpu_leaves import contour_reference as tree_reference
from test_gpu_parity import contour_reference as grid_reference

pytestmark = pytest.mark.gpu


def both_welds(cache, grid):
    out = []
    for mode in ("host", "device"):
        cache.set_contour_weld(mode)
        out.append(generate_mesh(cache, grid))
    cache.set_contour_weld("auto")
    return out


def test_grid_contours_weld_the_same_on_the_device():
    g = CartesianIsoGrid((-1.0, -1.0), (1.0, 1.0), sz=0.01)
    c = update_cache(allocate_cache(g, MarchingSquaresCutCell), SDFCircle((0.0, 0.0), 0.5), g)
    (hn, hl), (dn, dl) = both_welds(c, g)
    wn, wl = grid_reference(c.phi, g)
    assert np.array_equal(dn, wn) and np.array_equal(dl, wl) and np.array_equal(hn, dn) and np.array_equal(hl, dl)
    ga = CartesianGrid((-1.3, -0.7), (97, 83), (0.027, 0.019))
    rng = XorShift(99)
    m = 98 * 84
    for dtype in (np.float64, np.float32):
        for field in (randn_field(rng, ga.nodegrid_size()),
                      np.where(np.arange(m) % 7 == 0, 0.0, randn_field(rng, ga.nodegrid_size())),
                      np.where(np.arange(m) % 5 == 0, -0.0, randn_field(rng, ga.nodegrid_size()))):
            c = allocate_cache(ga, MarchingSquaresCutCell, dtype)
            update_cache(c, field.astype(dtype), ga)
            (hn, hl), (dn, dl) = both_welds(c, ga)
            wn, wl = grid_reference(c.phi, ga, dtype)
            assert np.array_equal(dn, wn) and np.array_equal(dl, wl), dtype
            assert np.array_equal(hn, dn) and np.array_equal(hl, dl)
            if dtype == np.float64:
                on, ol = orc.ms_contour(field, ga.x0, ga.d, ga.n)
                assert np.array_equal(dn, on) and np.array_equal(dl, ol)
            assert len(dl) > 1000


def test_tree_contours_weld_the_same_on_the_device():
    mesh = banded_tree()
    c = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), SDFCircle((0.0, 0.0), 1.0), mesh)
    (hn, hl), (dn, dl) = both_welds(c, mesh)
    rn, rl = tree_reference(c, mesh)
    assert np.array_equal(dn, rn) and np.array_equal(dl, rl) and np.array_equal(hn, dn) and np.array_equal(hl, dl)
    # about a million leaves, refined round a circle (the quadtree workload): the auto mode welds on
    # the device there
    base = CartesianGrid((-2.0, -2.0), (8, 8), (0.5, 0.5))
    big = AdaptiveMesh(base, max_level=12, initial_level=1)
    centre, radius = (0.13, -0.07), 0.9
    for _ in range(12):
        h = big.elem_size()
        q = big.elem_centroid() - np.asarray(centre)
        dist = np.abs(np.sqrt((q * q).sum(axis=1)) - radius)
        flags = (big.level < 12) & (dist <= 2.5 * np.sqrt((h * h).sum(axis=1)))
        if not flags.any():
            break
        big.refine(flags)
    cb = update_cache(allocate_cache(big, MarchingSquaresCutCell), SDFCircle(centre, radius), big)
    auto = generate_mesh(cb, big)
    (hn, hl), (dn, dl) = both_welds(cb, big)
    assert np.array_equal(hn, dn) and np.array_equal(hl, dl)
    assert np.array_equal(auto[0], dn) and np.array_equal(auto[1], dl)
    assert len(dl) > 50000
    counts = np.bincount(dl.reshape(-1), minlength=dn.shape[0] + 1)[1:]
    assert np.all(counts == 2)  # one closed loop
