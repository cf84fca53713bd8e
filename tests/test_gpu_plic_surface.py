"""GPU parity of the PLIC surface (generate_mesh / extract_surface_plic, src/plic/surface.jl) and of
calc_volume (src/volume.jl) against the oracle, through the C ABI.

Bar: the device surface bit-exact against the oracle's (nodes and 1-based elements, 2D and 3D, at
several element floors), the reference's golden counts (84; 3384 / 1688), generate_mesh equal to
extract_surface_plic of the cut cells' planes, corner-touching planes drawing points; calc_volume
within 1e-12 (relative) of the oracle's serial sum and of the cache's fluid fractions (the
tolerance the reference's own test states), identical on every run; Float32 variants; the errors.
"""
import numpy as np
import pytest

import oracle as orc
from cutcell import (ArgumentError, CartesianGrid, CartesianIsoGrid, DimensionMismatch,
                     MarchingCubesCutCell, PLICCutCell, SDFBox, SDFPlane, SDFSphere, SDFSphereUnion,
                     SDFTorus, Unsupported, allocate_cache, calc_volume, extract_surface_plic,
                     generate_mesh, random_rotation, update_cache)

pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def check_surface(g, geo, fracs=(1e-4, 0.0, 0.1)):
    dim = g.dim
    c = update_cache(allocate_cache(g, PLICCutCell), geo, g)
    cols, nrm, ic, _ = orc.plic(geo, dim, g.x0, g.d, g.n)
    assert same(c.normals, nrm.reshape(c.normals.shape)) and same(c.intercepts, ic)
    out = {}
    for f in fracs:
        nodes, el = generate_mesh(c, g, warn=False, min_tri_area_fraction=f)
        rn, re = orc.plic_surface(dim, g.x0, g.d, g.n, nrm, ic, f)
        assert same(nodes, rn) and same(el, re), (type(geo).__name__, f)
        out[f] = (nodes, el)
    return c, cols, out


def test_2d_circle_golden_and_bitwise():
    g = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)  # test/test_plic.jl:25-26
    c, cols, out = check_surface(g, SDFSphere((0.0, 0.0), 1.0))
    nodes, el = out[1e-4]
    assert len(el) == int((cols["kind"] == 2).sum()) == 84
    assert len(out[0.0][1]) == np.prod(g.n) and len(out[0.1][1]) <= 84
    # the cache's own planes, clipped in cell order: what extract_surface_plic draws from the cut cells
    cut = np.flatnonzero(c.cells.kind == 2)
    rn, re = extract_surface_plic(c.normals[cut], c.intercepts[cut], cut, g)
    assert same(rn, nodes) and same(re, el)
    with pytest.raises(DimensionMismatch):
        generate_mesh(c, CartesianGrid(g.x0, (31, 30), g.d))


def test_plane_through_corners():
    g = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    n0 = np.array([1.0, 1.0]) / np.sqrt(2.0)
    c = update_cache(allocate_cache(g, PLICCutCell), SDFPlane((n0[0], n0[1], 0.0), (0.0, 0.0, 0.0)), g)
    allc = np.arange(g.ncells)
    nodes, el = extract_surface_plic(c.normals, c.intercepts, allc, g, min_tri_area_fraction=0)
    assert len(el) == g.ncells
    seg = nodes.reshape(-1, 2, 2)
    length = np.linalg.norm(seg[:, 1] - seg[:, 0], axis=1)
    cut = c.cells.kind == 2
    assert np.all(cut | (length < 1e-12 * g.d[0])) and np.sum(~cut & (length == 0.0)) > 0
    ref, _ = generate_mesh(c, g)
    assert same(seg[cut].reshape(-1, 2), ref)
    assert np.all(np.abs(ref @ n0) < 1e-12)


def test_3d_sphere_golden_torus_box_union_bitwise():
    g = CartesianIsoGrid((-1.5,) * 3, (1.5,) * 3, sz=0.15)  # test/test_plic.jl:139-148
    c, cols, out = check_surface(g, SDFSphere((0.0, 0.0, 0.0), 1.0))
    nodes, el = out[1e-4]
    assert len(nodes) == 3384 and len(el) == 1688
    rng = np.random.default_rng(7)
    g2 = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=48)
    for geo in (SDFTorus((0.02, -0.03, 0.01), 0.5, 0.15, random_rotation(rng)),
                SDFBox((0.05, 0.0, 0.0), (0.4, 0.3, 0.5), random_rotation(rng)),
                SDFSphereUnion(rng.uniform(-0.5, 0.5, (16, 3)), rng.uniform(0.1, 0.3, 16), 0.05)):
        check_surface(g2, geo, fracs=(1e-4, 0.0))


def test_3d_planes_on_lattice_faces_and_nodes():
    # a plane lying on a lattice face: the cells beside it are not cut (fraction 0 or 1) yet their
    # planes run along a face of theirs and are drawn; a plane through lattice nodes touches cells at
    # corners and edges only.  generate_mesh's cell filter must keep every such cell.
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=16)
    for nrm, pt, on_faces in (((0.0, 0.0, 1.0), (0.0, 0.0, 0.0), True), ((0.0, 1.0, 0.0), (0.0, 0.25, 0.0), True),
                              ((1.0, 1.0, 1.0), (0.0, 0.0, 0.0), False),
                              ((1.0, -1.0, 0.0), (0.125, 0.0, 0.0), False)):
        c, cols, out = check_surface(g, SDFPlane(nrm, pt), fracs=(1e-4, 0.0))
        nodes, el = out[0.0]
        if on_faces:  # cells that are not cut draw too
            assert len(el) > int((cols["kind"] == 2).sum())
    c, cols, out = check_surface(g, SDFPlane((0.0, 0.0, 1.0), (0.0, 0.0, 0.0)), fracs=(1e-4,))
    nodes, el = out[1e-4]
    assert int((cols["kind"] == 2).sum()) == 0 and len(el) == 2 * 2 * 16 * 16 and np.all(nodes[:, 2] == 0.0)


def test_calc_volume_against_the_oracle_and_the_cache():
    rng = np.random.default_rng(11)
    cases = [(CartesianIsoGrid((-1.5,) * 3, (1.5,) * 3, sz=0.15), SDFSphere((0.0, 0.0, 0.0), 1.0)),
             (CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=96), SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15,
                                                                       random_rotation(rng))),
             (CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.01), SDFSphere((0.1, -0.05), 1.0))]
    for g, geo in cases:
        v = calc_volume(geo, g)
        ref = orc.calc_volume(geo, g.dim, g.x0, g.d, g.n)
        assert abs(v - ref) < 1e-12 * ref
        assert calc_volume(geo, g) == v  # a fixed summation order: the same on every run
        c = update_cache(allocate_cache(g, PLICCutCell), geo, g)
        domain = np.prod(g.n) * np.prod(g.d)
        assert abs(domain - c.cells.volume_fraction.sum() * np.prod(g.d) - v) < 1e-12 * v
        v32 = calc_volume(geo, g, np.float32)
        assert v32.dtype == np.float32 and abs(float(v32) - ref) < 1e-4 * ref
    with pytest.raises(ArgumentError):
        calc_volume(SDFSphere((0.0, 0.0, 0.0), 1.0), CartesianGrid((-1.0,) * 3, (10, 10, 10), (0.2, 0.2, 0.3)))


def test_float32_surface_and_errors():
    g = CartesianIsoGrid((-1.5,) * 3, (1.5,) * 3, sz=0.15)
    geo = SDFSphere((0.0, 0.0, 0.0), 1.0)
    c32 = update_cache(allocate_cache(g, PLICCutCell, np.float32), geo, g)
    nodes, el = generate_mesh(c32, g)
    assert nodes.dtype == np.float32
    c64 = update_cache(allocate_cache(g, PLICCutCell), geo, g)
    n64, e64 = generate_mesh(c64, g)
    assert abs(len(el) - len(e64)) <= 8 and abs(len(nodes) - len(n64)) <= 16
    with pytest.raises(Unsupported):
        mc = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=8)
        import cutcell._lib as L
        import ctypes
        c = allocate_cache(mc, MarchingCubesCutCell)
        nn, ne = ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib().cc_generate_mesh_plic(c.handle, ctypes.byref(L.make_grid(mc)), 1e-4, None,
                                               ctypes.byref(nn), None, ctypes.byref(ne)))
    with pytest.warns(UserWarning):
        generate_mesh(update_cache(allocate_cache(g, PLICCutCell), SDFSphere((9.0, 9.0, 9.0), 0.5), g), g)
