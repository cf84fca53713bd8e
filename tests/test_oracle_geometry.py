"""Pins the oracle's bodies (the library's own geometry: SDFLibrary is not available) against the
mathematics that fixes them: brute-force distances to densely sampled surfaces, the inside test,
closed forms, and the reduction of the polynomial smooth union to the plain minimum.  A wrong sign,
a dropped rotation or a transposed frame fails one of these."""
import numpy as np
import pytest

import oracle as orc
from cutcell import SDFBox, SDFPlane, SDFSphere, SDFSphereUnion, SDFTorus, random_rotation


def sdf(geo, pts):
    return np.array([orc.sdf_value(geo, 3, p) for p in pts])


def test_sphere_closed_form():
    rng = np.random.default_rng(0)
    c, r = np.array([0.1, -0.2, 0.3]), 0.45
    pts = rng.uniform(-1, 1, (500, 3))
    want = np.sqrt(((pts[:, 0] - c[0]) ** 2 + (pts[:, 1] - c[1]) ** 2) + (pts[:, 2] - c[2]) ** 2) - r
    assert np.array_equal(sdf(SDFSphere(tuple(c), r), pts), want)  # the same rounding, bit for bit


def test_torus_against_brute_force_distance():
    rng = np.random.default_rng(1)
    R = random_rotation(rng)
    c, Rmaj, rmin = np.array([0.05, 0.1, -0.05]), 0.5, 0.15
    geo = SDFTorus(tuple(c), Rmaj, rmin, R)
    # the tube's centre circle in world coordinates: c + R (Rmaj cos t, Rmaj sin t, 0)
    t = np.linspace(0, 2 * np.pi, 20000, endpoint=False)
    circle = c + (np.stack([Rmaj * np.cos(t), Rmaj * np.sin(t), 0 * t], 1) @ np.asarray(R).T)
    pts = rng.uniform(-0.9, 0.9, (300, 3))
    d = np.array([np.min(np.linalg.norm(circle - p, axis=1)) for p in pts]) - rmin
    assert np.max(np.abs(sdf(geo, pts) - d)) < 1e-6  # circle sampling error ~ (Rmaj dt)^2 / 8


def test_box_against_brute_force_distance_and_inside_test():
    rng = np.random.default_rng(2)
    R = np.asarray(random_rotation(rng))
    c, h = np.array([-0.1, 0.05, 0.2]), np.array([0.3, 0.2, 0.4])
    geo = SDFBox(tuple(c), tuple(h), R)
    pts = rng.uniform(-0.9, 0.9, (400, 3))
    p_loc = (pts - c) @ R  # R^T (x - c) per row
    inside = np.all(np.abs(p_loc) <= h, axis=1)
    # distance to the box surface by clamping (outside) or to the nearest face (inside)
    q = np.abs(p_loc) - h
    out_d = np.linalg.norm(np.maximum(q, 0), axis=1)
    in_d = np.min(h - np.abs(p_loc), axis=1)
    want = np.where(inside, -in_d, out_d)
    got = sdf(geo, pts)
    assert np.max(np.abs(got - want)) < 1e-14
    assert np.array_equal(got < 0, inside)


def test_plane_closed_form():
    n, p0 = np.array([0.6, 0.0, 0.8]), np.array([0.1, 0.1, 0.1])
    pts = np.random.default_rng(3).uniform(-1, 1, (200, 3))
    want = ((n[0] * (pts[:, 0] - p0[0]) + n[1] * (pts[:, 1] - p0[1])) + n[2] * (pts[:, 2] - p0[2]))
    assert np.array_equal(sdf(SDFPlane(tuple(n), tuple(p0)), pts), want)


def test_smooth_union_fold_and_its_reduction_to_min():
    rng = np.random.default_rng(4)
    cen = rng.uniform(-0.6, 0.6, (24, 3))
    rad = rng.uniform(0.05, 0.2, 24)
    k = 0.02
    geo = SDFSphereUnion(cen, rad, k)
    pts = rng.uniform(-1, 1, (300, 3))
    d = np.sqrt(((pts[:, None, 0] - cen[None, :, 0]) ** 2 + (pts[:, None, 1] - cen[None, :, 1]) ** 2)
                + (pts[:, None, 2] - cen[None, :, 2]) ** 2) - rad[None, :]
    acc = d[:, 0].copy()
    for i in range(1, 24):  # left fold, h = max(k - |a-b|, 0)/k, min(a,b) - h^2 k / 4
        b = d[:, i]
        h = np.maximum(k - np.abs(acc - b), 0.0) / k
        acc = np.minimum(acc, b) - ((h * h) * k) / 4.0
    got = sdf(geo, pts)
    assert np.array_equal(got, acc)
    # the smoothing only ever lowers the minimum, and by at most k/4 per fold step that blends
    m = d.min(axis=1)
    assert np.all(got <= m) and np.all(got >= m - 24 * k / 4)
    # far apart (|a - b| >= k everywhere): exactly the plain minimum
    geo2 = SDFSphereUnion(np.array([[-0.5, 0, 0], [0.5, 0, 0]]), np.array([0.1, 0.1]), k)
    far = np.array([[-0.5, 0.3, 0.0], [0.5, -0.2, 0.1], [-0.45, 0.0, 0.05]])
    dd = np.stack([np.linalg.norm(far - [-0.5, 0, 0], axis=1) - 0.1,
                   np.linalg.norm(far - [0.5, 0, 0], axis=1) - 0.1], 1)
    assert np.all(np.abs(dd[:, 0] - dd[:, 1]) >= k)
    assert np.array_equal(sdf(geo2, far), dd.min(axis=1))


def test_bounded_body_band_and_the_reference_test():
    # SDFBounded (test/test_bounded.jl): the child within the pad of its box, the fill beyond, and a
    # pad wider than a cell diagonal leaves every cut cell -- and so the volume and the contour -- alone
    from cutcell import CartesianIsoGrid, SDFBounded, SDFBox, SDFCircle, SDFPlane, SDFSphere, SDFTorus
    circle = SDFCircle((0.0, 0.0), 1.0)
    b = SDFBounded(circle, pad_distance=0.5)
    rng = np.random.default_rng(9)
    pts = rng.uniform(-3.0, 3.0, (4000, 2))
    for x in pts:
        lo = np.clip(x, -1.0, 1.0)
        far = np.sum((x - lo) ** 2) > 0.25
        v = orc.sdf_value(b, 2, x)
        assert v == (0.5 if far else orc.sdf_value(circle, 2, x))
        val, nrm = orc.get_sdf(b, 2, x)
        assert val == v and (np.all(nrm == 0.0) if far else np.array_equal(nrm, orc.get_sdf(circle, 2, x)[1]))
    # test/test_bounded.jl:15-19: cell 0.1, pad 0.5
    g = CartesianIsoGrid((-2.0, -2.0), (2.0, 2.0), n=40)
    vb = orc.calc_volume(b, 2, g.x0, g.d, g.n)
    assert vb == orc.calc_volume(circle, 2, g.x0, g.d, g.n)
    assert abs(vb - np.pi) < 1e-3 * np.pi
    # :21-25, the nodal reconstruction: every cell the same (only far phi differs)
    pb, pc = orc.sample_nodes(b, 2, g.x0, g.d, g.n), orc.sample_nodes(circle, 2, g.x0, g.d, g.n)
    cb, cc = orc.ms_cells(pb, g.x0, g.d, g.n), orc.ms_cells(pc, g.x0, g.d, g.n)
    for k in cc:
        assert np.array_equal(cb[k], cc[k]), k
    assert np.any(pb != pc) and np.all((pb == pc) | (pb == 0.5))
    # bounding boxes contain their bodies: on the box's boundary every child is >= 0 (outside)
    for child in (SDFSphere((0.1, -0.2, 0.3), 0.7), SDFTorus((0.0, 0.1, 0.0), 0.5, 0.15, np.eye(3)),
                  SDFBox((0.0, 0.0, 0.0), (0.3, 0.2, 0.5), random_rotation(rng))):
        lo, hi = SDFBounded(child, 0.1).bounds()
        face = rng.uniform(lo, hi, (2000, 3))
        ax = rng.integers(0, 3, 2000)
        side = rng.integers(0, 2, 2000)
        face[np.arange(2000), ax] = np.where(side == 1, np.array(hi)[ax], np.array(lo)[ax])
        assert all(orc.sdf_value(child, 3, p) >= -1e-12 for p in face), type(child).__name__
    with pytest.raises(ValueError):
        SDFBounded(SDFPlane((0.0, 0.0, 1.0), (0.0, 0.0, 0.0)), 0.5)
    with pytest.raises(ValueError):
        SDFBounded(circle, 0.5, fill_distance=0.0)
