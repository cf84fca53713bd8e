"""GPU parity for the union of mixed primitives: the reference's sharp-corner PLIC case
(``test/test_plic.jl:96-112``) and its 3D analogue on the device, bit for bit against the oracle,
with the reference's own assertions on the device output; marching cubes, marching squares and
``calc_volume`` over unions with and without smoothing; the host checks of the descriptor."""
import numpy as np
import pytest

import oracle as orc
from cutcell import (ArgumentError, CartesianIsoGrid, MarchingCubesCutCell, MarchingSquaresCutCell,
                     PLICCutCell, SDFBox, SDFCircle, SDFPlane, SDFRectangle, SDFSphere, SDFTorus,
                     SDFUnion, Unsupported, allocate_cache, calc_volume, closure_residual,
                     random_rotation, update_cache)
from cutcell.geometry import GEO_EMPTY
from helpers import shared_face_mismatches
from test_gpu_parity import assert_cells_equal, same_bits

pytestmark = pytest.mark.gpu


def children_3d():
    rng = np.random.default_rng(41)
    return (SDFSphere((0.1, -0.05, 0.02), 0.35),
            SDFBox((-0.2, 0.15, 0.05), (0.25, 0.15, 0.3), random_rotation(rng)),
            SDFTorus((0.05, 0.1, -0.1), 0.45, 0.12, random_rotation(rng)),
            SDFPlane(tuple(np.array([0.2, -0.3, 0.93]) / np.linalg.norm([0.2, -0.3, 0.93])),
                     (0.0, 0.0, -0.55)))


def check_plic(geo, g):
    c = update_cache(allocate_cache(g, PLICCutCell), geo, g)
    ref, normals, intercepts, valid = orc.plic(geo, g.dim, g.x0, g.d, g.n)
    assert_cells_equal(c, ref)
    assert same_bits(c.normals, normals) and same_bits(c.intercepts, intercepts)
    assert np.array_equal(c.is_valid, valid)
    return c


def test_sharp_corner_plic_2d():
    # test/test_plic.jl:96-112, on the device
    geo = SDFUnion(SDFCircle((0.02, 0.03), 0.31), SDFRectangle((0.25, 0.10), (0.18, 0.18)),
                   smoothing=0.0)
    g = CartesianIsoGrid((-0.6, -0.6), (0.6, 0.6), sz=0.05)
    c = check_plic(geo, g)
    ff = c.cells.face_fraction
    assert int((c.cells.kind == 2).sum()) > 40
    assert shared_face_mismatches(ff, g.n, 2) == 0
    assert np.all((ff >= 0) & (ff <= 1))
    assert np.count_nonzero(closure_residual(ff, g.d)) == 0


@pytest.mark.parametrize("smoothing", [0.0, 0.03])
def test_sharp_corner_plic_3d(smoothing):
    geo = SDFUnion(SDFSphere((0.02, 0.03, -0.01), 0.31), SDFBox((0.25, 0.10, 0.05), (0.09, 0.09, 0.12)),
                   smoothing=smoothing)
    g = CartesianIsoGrid((-0.6,) * 3, (0.6,) * 3, sz=0.04)
    for variant in (0, 2):  # a 3D union takes the two passes whatever the tuning asks
        c = allocate_cache(g, PLICCutCell)
        c.set_tuning(variant, 8, False)
        update_cache(c, geo, g)
        ref, normals, intercepts, valid = orc.plic(geo, 3, g.x0, g.d, g.n)
        assert_cells_equal(c, ref)
        assert same_bits(c.normals, normals) and np.array_equal(c.is_valid, valid)
        ff = c.cells.face_fraction
        assert int((c.cells.kind == 2).sum()) > 500
        assert shared_face_mismatches(ff, g.n, 3) == 0
        assert np.count_nonzero(closure_residual(ff, g.d)) == 0


@pytest.mark.parametrize("smoothing", [0.0, 0.05])
def test_mc_union_bitwise(smoothing):
    geo = SDFUnion(*children_3d(), smoothing=smoothing)
    g = CartesianIsoGrid((-0.8,) * 3, (0.8,) * 3, n=44)
    for variant in (1, 2):  # the fused variant samples a union in a pass of its own
        c = allocate_cache(g, MarchingCubesCutCell)
        c.set_tuning(variant, 4, False)
        update_cache(c, geo, g)
        phi = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
        assert same_bits(c.phi, phi)
        assert_cells_equal(c, orc.mc_cells(phi, g.x0, g.d, g.n))
        assert shared_face_mismatches(c.cells.face_fraction, g.n, 3) == 0


def test_ms_union_bitwise():
    geo = SDFUnion(SDFCircle((0.02, 0.03), 0.31), SDFRectangle((0.25, 0.10), (0.18, 0.18)),
                   SDFBox((-0.3, -0.2), (0.1, 0.25), np.array([[0.8, -0.6], [0.6, 0.8]])),
                   smoothing=0.02)
    g = CartesianIsoGrid((-0.6, -0.6), (0.6, 0.6), n=120)
    c = update_cache(allocate_cache(g, MarchingSquaresCutCell), geo, g)
    phi = orc.sample_nodes(geo, 2, g.x0, g.d, g.n)
    assert same_bits(c.phi, phi)
    assert_cells_equal(c, orc.ms_cells(phi, g.x0, g.d, g.n))


def test_calc_volume_of_a_union():
    geo = SDFUnion(*children_3d()[:3], smoothing=0.02)
    g = CartesianIsoGrid((-0.8,) * 3, (0.8,) * 3, n=40)
    v = calc_volume(geo, g)
    want = orc.calc_volume(geo, 3, g.x0, g.d, g.n)
    assert abs(v - want) <= 1e-12 * abs(want)


def test_union_descriptor_checks():
    g = CartesianIsoGrid((-0.6,) * 3, (0.6,) * 3, n=8)
    c = allocate_cache(g, MarchingCubesCutCell)
    bad = SDFUnion(SDFSphere((0.0, 0.0, 0.0), 0.3))
    bad._prims = bad._prims.copy()
    bad._prims[0, 0] = GEO_EMPTY
    with pytest.raises(Unsupported):
        update_cache(c, bad, g)
    neg = SDFUnion(SDFSphere((0.0, 0.0, 0.0), 0.3))
    neg.smoothing = -1.0
    with pytest.raises(ArgumentError):
        update_cache(c, neg, g)
    c2 = allocate_cache(CartesianIsoGrid((-0.6, -0.6), (0.6, 0.6), n=8), MarchingSquaresCutCell)
    with pytest.raises(Unsupported):
        update_cache(c2, SDFUnion(SDFTorus((0, 0, 0), 0.3, 0.1)), CartesianIsoGrid((-0.6, -0.6), (0.6, 0.6), n=8))
