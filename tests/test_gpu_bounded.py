"""GPU parity of bounded bodies (SDFLibrary's SDFBounded as test/test_bounded.jl uses it), through the
C ABI.

Bar: the reference test's claims on the device -- calc_volume of the bounded circle equal to the
plain circle's (bit for bit here, where the reference asks 1e-12) and to pi within 1e-3, the
marching-squares contour the circle's (the same nodes; here the same arrays) -- and every column,
phi and PLIC plane of bounded marching-squares, marching-cubes and PLIC caches bit-exact against the
oracle; a bounded plane refused.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc
from cutcell import (ArgumentError, CartesianIsoGrid, MarchingCubesCutCell, MarchingSquaresCutCell,
                     PLICCutCell, SDFBounded, SDFBox, SDFCircle, SDFSphere, SDFTorus, allocate_cache,
                     calc_volume, generate_mesh, random_rotation, update_cache)

pytestmark = pytest.mark.gpu

COLS = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction", "face_centroid_local",
        "interface_centroid")


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_the_reference_bounded_test_on_the_device():
    circle = SDFCircle((0.0, 0.0), 1.0)
    bounded = SDFBounded(circle, pad_distance=0.5)
    g = CartesianIsoGrid((-2.0, -2.0), (2.0, 2.0), n=40)  # test/test_bounded.jl:15-17
    vb, vc = calc_volume(bounded, g), calc_volume(circle, g)
    assert vb == vc and abs(vb - np.pi) < 1e-3 * np.pi
    assert abs(vb - orc.calc_volume(bounded, 2, g.x0, g.d, g.n)) < 1e-12 * vb
    cb = update_cache(allocate_cache(g, MarchingSquaresCutCell), bounded, g)
    cc = update_cache(allocate_cache(g, MarchingSquaresCutCell), circle, g)
    nb, lb = generate_mesh(cb, g)
    nc, lc = generate_mesh(cc, g)
    assert len(nb) == len(nc) and np.array_equal(nb, nc) and np.array_equal(lb, lc)  # :21-25
    for k in COLS:
        assert same(cb.read(k), cc.read(k)), k
    ref = orc.sample_nodes(bounded, 2, g.x0, g.d, g.n)
    assert same(cb.phi, ref) and np.any(cb.phi != cc.phi)
    cols = orc.ms_cells(ref, g.x0, g.d, g.n)
    for k in COLS:
        assert same(cb.read(k), cols[k]), k


def test_bounded_3d_bodies_bitwise_against_the_oracle():
    rng = np.random.default_rng(17)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=48)
    for child, pad in ((SDFSphere((0.05, -0.1, 0.02), 0.45), 0.1),
                       (SDFTorus((0.0, 0.05, 0.0), 0.45, 0.15, random_rotation(rng)), 0.08),
                       (SDFBox((0.1, 0.0, -0.05), (0.3, 0.25, 0.35), random_rotation(rng)), 0.06)):
        geo = SDFBounded(child, pad_distance=pad)
        c = update_cache(allocate_cache(g, MarchingCubesCutCell), geo, g)
        phi = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
        assert same(c.phi, phi) and np.any(phi == pad)
        ref = orc.mc_cells(phi, g.x0, g.d, g.n)
        for k in COLS:
            assert same(c.read(k), ref[k]), (type(child).__name__, k)
        p = update_cache(allocate_cache(g, PLICCutCell), geo, g)
        cols, nrm, ic, valid = orc.plic(geo, 3, g.x0, g.d, g.n)
        assert same(p.normals, nrm.reshape(p.normals.shape)) and same(p.intercepts, ic)
        for k in ("kind", "volume_fraction", "face_fraction"):
            assert same(p.read(k), cols[k]), (type(child).__name__, k)
        assert not np.all(p.is_valid)  # planes beyond the band have no normal


def test_a_bounded_plane_is_refused_by_the_abi():
    import cutcell._lib as L
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=8)
    c = allocate_cache(g, MarchingCubesCutCell)
    gg = L.make_geometry(SDFSphere((0.0, 0.0, 0.0), 0.5))
    gg.kind = 4 | 0x100  # a plane flagged bounded
    for i in range(16, 22):
        gg.p[i] = -1.0 if i < 19 else 1.0
    gg.p[22], gg.p[23] = 0.1, 0.1
    with pytest.raises(ArgumentError):
        L.check(L.lib().cc_update(c.handle, ctypes.byref(gg), ctypes.byref(L.make_grid(g))))
    gg = L.make_geometry(SDFBounded(SDFSphere((0.0, 0.0, 0.0), 0.5), 0.1))
    gg.p[23] = 0.0  # no positive fill
    with pytest.raises(ArgumentError):
        L.check(L.lib().cc_update(c.handle, ctypes.byref(gg), ctypes.byref(L.make_grid(g))))
