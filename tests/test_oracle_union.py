"""Pins the oracle's union of mixed primitives (``SDFUnion(children...; smoothing)``; SDFLibrary is not
vendored, so the body is the build's own) and ports the reference's sharp-corner PLIC case
(``test/test_plic.jl:96-112``) and a 3D analogue onto the oracle.

* smoothing 0: the value is the children's plain minimum bit for bit, the normal the winner's own
  (a tie keeps the later child);
* smoothing k > 0: the value is the left fold of the polynomial smooth minimum (closed form, in
  numpy), equal to the plain minimum wherever no step blends, and the normal is the unit gradient
  of the value (central differences);
* the sharp corner: a circle and a rectangle unioned with smoothing 0, where the two cells flanking
  a face fit genuinely different planes, still gives every shared face one value, every fraction in
  [0, 1] and exactly zero closure.
"""
import numpy as np
import pytest

import oracle as orc
from cutcell import (CartesianIsoGrid, SDFBox, SDFCircle, SDFPlane, SDFRectangle, SDFSphere,
                     SDFTorus, SDFUnion, random_rotation)
from cutcell import readers
from helpers import fortran, shared_face_mismatches


def children_3d():
    rng = np.random.default_rng(31)
    return (SDFSphere((0.1, -0.05, 0.02), 0.35),
            SDFBox((-0.2, 0.15, 0.05), (0.25, 0.15, 0.3), random_rotation(rng)),
            SDFTorus((0.05, 0.1, -0.1), 0.45, 0.12, random_rotation(rng)),
            SDFPlane(tuple(np.array([0.2, -0.3, 0.93]) / np.linalg.norm([0.2, -0.3, 0.93])),
                     (0.0, 0.0, -0.55)))


def smin(a, b, k):
    if k == 0.0:
        return a if a < b else b
    h = k - abs(a - b)
    m = a if a < b else b
    if not h > 0.0:
        return m
    h = h / k
    return m - ((h * h) * k) / 4.0


@pytest.mark.parametrize("dim", [2, 3])
def test_plain_union_is_the_minimum_with_the_winners_normal(dim):
    kids = children_3d() if dim == 3 else (SDFCircle((0.02, 0.03), 0.31),
                                           SDFRectangle((0.25, 0.10), (0.18, 0.18)),
                                           SDFBox((-0.2, -0.1), (0.1, 0.3)))
    geo = SDFUnion(*kids, smoothing=0.0)
    rng = np.random.default_rng(32 + dim)
    for x in rng.uniform(-0.8, 0.8, (500, 3)):
        vals, nrms = zip(*(orc.get_sdf(c, dim, x) for c in kids))
        w = 0
        for i in range(1, len(vals)):
            if not vals[w] < vals[i]:
                w = i
        d, n = orc.get_sdf(geo, dim, x)
        assert d == vals[w] and np.array_equal(n, nrms[w])
        assert orc.sdf_value(geo, dim, x) == d
    # a tie keeps the later child (min(a, b) = a < b ? a : b)
    twin = SDFUnion(SDFSphere((0.0, 0.0, 0.0), 0.5), SDFPlane((1.0, 0.0, 0.0), (0.5, 0.0, 0.0)))
    d, n = orc.get_sdf(twin, 3, np.array([1.0, 0.0, 0.0]))
    assert d == 0.5 and np.array_equal(n, [1.0, 0.0, 0.0])
    d, n = orc.get_sdf(twin, 3, np.array([0.0, 0.0, 0.0]))
    assert d == -0.5 and np.array_equal(n, [1.0, 0.0, 0.0])  # the sphere's centre: the plane wins


def test_smooth_union_value_and_normal():
    kids = children_3d()[:3]
    k = 0.08
    geo = SDFUnion(*kids, smoothing=k)
    rng = np.random.default_rng(34)
    blended = 0
    for x in rng.uniform(-0.8, 0.8, (700, 3)):
        vals = [orc.sdf_value(c, 3, x) for c in kids]
        acc = vals[0]
        blends = False
        for v in vals[1:]:
            blends |= k - abs(acc - v) > 0.0
            acc = smin(acc, v, k)
        d, n = orc.get_sdf(geo, 3, x)
        assert d == acc
        if not blends:
            assert d == min(vals)
        # the normal is the value's unit gradient, away from the children's own kinks
        h = 1e-6
        grad = np.array([(orc.sdf_value(geo, 3, x + e) - orc.sdf_value(geo, 3, x - e)) / (2 * h)
                         for e in np.eye(3) * h])
        fd_ok = all(np.linalg.norm([(orc.sdf_value(c, 3, x + e) - orc.sdf_value(c, 3, x - e)) / (2 * h)
                                    for e in np.eye(3) * h]) > 0.999 for c in kids)
        if fd_ok and np.linalg.norm(grad) > 0.2:
            assert abs(np.linalg.norm(n) - 1.0) < 1e-14
            assert np.max(np.abs(n - grad / np.linalg.norm(grad))) < 1e-6
            blended += blends
    assert blended > 20


def test_smoothing_zero_never_divides():
    geo = SDFUnion(SDFSphere((0.0, 0.0, 0.0), 0.5), SDFSphere((0.2, 0.0, 0.0), 0.5), smoothing=0.0)
    for x in np.random.default_rng(35).uniform(-1, 1, (200, 3)):
        d, n = orc.get_sdf(geo, 3, x)
        assert np.isfinite(d) and np.all(np.isfinite(n))


def sharp_corner():
    # test/test_plic.jl:97-101
    geo = SDFUnion(SDFCircle((0.02, 0.03), 0.31), SDFRectangle((0.25, 0.10), (0.18, 0.18)),
                   smoothing=0.0)
    return geo, CartesianIsoGrid((-0.6, -0.6), (0.6, 0.6), sz=0.05)


def flanking_planes_disagree(normals, valid, ff_own, n, D):
    """How many interior faces lie between two valid cells whose fitted normals differ."""
    N = fortran(normals, n)
    V = fortran(valid, n)
    count = 0
    for c in range(D):
        lo = [slice(None)] * D
        hi = [slice(None)] * D
        lo[c] = slice(0, n[c] - 1)
        hi[c] = slice(1, n[c])
        diff = np.any(N[tuple(lo)] != N[tuple(hi)], axis=-1) & (V[tuple(lo)] == 1) & (V[tuple(hi)] == 1)
        count += int(diff.sum())
    return count


def test_sharp_corner_faces_are_single_valued():
    # test/test_plic.jl:96-112
    geo, g = sharp_corner()
    cols, normals, intercepts, valid = orc.plic(geo, 2, g.x0, g.d, g.n)
    ff = cols["face_fraction"]
    assert int((cols["kind"] == 2).sum()) > 40
    assert shared_face_mismatches(ff, g.n, 2) == 0
    assert np.all((ff >= 0) & (ff <= 1))
    assert np.count_nonzero(readers.closure_residual(ff, g.d)) == 0
    # the fixture is real: the rectangle's straight sides and the circle meet, so cells on either
    # side of some faces carry different planes (circle-fitted vs rectangle-fitted normals)
    rect = np.array([orc.get_sdf(SDFRectangle((0.25, 0.10), (0.18, 0.18)), 2, x)[1] for x in
                     [np.array([0.34 + 0.02, 0.1]), np.array([0.25, 0.19 + 0.02])]])
    assert any(np.any(np.all(normals == r, axis=1)) for r in rect)
    assert flanking_planes_disagree(normals, valid, ff, g.n, 2) > 20


def test_sharp_corner_3d_analogue():
    # a sphere and a box meeting at re-entrant edges, smoothing 0, on an offset isotropic grid
    geo = SDFUnion(SDFSphere((0.02, 0.03, -0.01), 0.31), SDFBox((0.25, 0.10, 0.05), (0.09, 0.09, 0.12)),
                   smoothing=0.0)
    g = CartesianIsoGrid((-0.6, -0.6, -0.6), (0.6, 0.6, 0.6), sz=0.04)
    cols, normals, intercepts, valid = orc.plic(geo, 3, g.x0, g.d, g.n)
    ff = cols["face_fraction"]
    assert int((cols["kind"] == 2).sum()) > 500
    assert shared_face_mismatches(ff, g.n, 3) == 0
    assert np.all(np.isfinite(ff) & (ff >= 0) & (ff <= 1))
    assert np.count_nonzero(readers.closure_residual(ff, g.d)) == 0
    assert flanking_planes_disagree(normals, valid, ff, g.n, 3) > 100


def test_nodal_union_samples_the_minimum():
    kids = children_3d()
    geo = SDFUnion(*kids, smoothing=0.0)
    g = CartesianIsoGrid((-0.7,) * 3, (0.7,) * 3, n=12)
    phi = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
    each = np.array([orc.sample_nodes(c, 3, g.x0, g.d, g.n) for c in kids])
    # min over children in fold order, ties to the later child
    want = each[0].copy()
    for v in each[1:]:
        want = np.where(want < v, want, v)
    assert np.array_equal(phi.view(np.uint64), want.view(np.uint64))
