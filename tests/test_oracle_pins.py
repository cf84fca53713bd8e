"""Pins the oracle functions that the reference's own literals leave open, each against what the
mathematics fixes, so that a plausible slip in the restatement fails here:

* the 3D PLIC one-sided face areas and their resolved mean (``src/plic/apertures.jl:91-125``,
  ``src/plic/cache.jl:143-162``) against the exact area of a plane-clipped square, computed in
  rationals by inclusion-exclusion, on oblique, axis-parallel and axis-aligned planes over an offset
  isotropic grid, boundary faces included -- the 3D form of ``test/test_cache.jl:228-234``;
* the torus and box normals ``get_sdf`` returns (the build's own bodies: SDFLibrary is not
  vendored) against central differences of the distance, which ``test_oracle_geometry`` pins by
  brute force, and unit length, off the medial sets;
* the zero-corner convention of ``nudge_zeros`` (``src/marching_cubes/marching_cubes.jl:225-243``):
  a corner exactly on the surface reads outside;
* the degenerate fit's faces (``test/test_plic.jl:180-206``) in 2D and 3D: the centre cell on the
  body's medial point has every face exactly closed.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as orc
from cutcell import SDFBox, SDFPlane, SDFSphere, SDFTorus, random_rotation
from cutcell.grid import CartesianGrid
from helpers import cube_nodes, fortran


def square_fluid_fraction(A, B, C):
    """Exact share of the unit square where ``A u + B v + C >= 0``, in rationals."""
    A, B, C = Fraction(A), Fraction(B), Fraction(C)
    if A == 0 and B == 0:
        return Fraction(1) if C >= 0 else Fraction(0)
    if A < 0:  # u -> 1 - u
        C, A = C + A, -A
    if B < 0:
        C, B = C + B, -B
    t = -C  # solid: A u + B v < t
    if A == 0 or B == 0:
        s = max(A, B)
        solid = min(max(t / s, Fraction(0)), Fraction(1))
    else:
        solid = Fraction(0)
        for a in (0, 1):
            for b in (0, 1):
                r = t - a * A - b * B
                if r > 0:
                    solid += (-1) ** (a + b) * r * r
        solid /= 2 * A * B
    return 1 - solid


def exact_plane_faces(g, nrm, p0):
    """Every cell's six face fractions of the half-space ``nrm . (x - p0) >= 0`` (outside = fluid)."""
    n, d = g.n, Fraction(g.d[0])
    nF = [Fraction(v) for v in nrm]
    pF = [Fraction(v) for v in p0]
    x0 = [Fraction(v) for v in g.x0]
    out = np.zeros((g.ncells, 6))
    ax = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                c = i + n[0] * (j + n[1] * k)
                lo = [x0[0] + i * d, x0[1] + j * d, x0[2] + k * d]
                for slot in range(6):
                    a = slot >> 1
                    corner = list(lo)
                    corner[a] += d * (slot & 1)
                    phi0 = sum(nF[b] * (corner[b] - pF[b]) for b in range(3))
                    p, q = ax[a]
                    out[c, slot] = float(square_fluid_fraction(nF[p] * d, nF[q] * d, phi0))
    return out


@pytest.mark.parametrize("nrm,p0", [
    ((0.43, -0.71, 0.56), (0.013, -0.021, 0.037)),   # oblique
    ((0.6, 0.0, -0.8), (-0.05, 0.2, 0.011)),          # parallel to y: n_p = 0 on two face pairs
    ((0.0, 0.0, 1.0), (0.0, 0.0, 0.0123)),            # parallel to the z faces (the guard)
    ((0.0, -1.0, 0.0), (0.0, -0.0311, 0.0)),          # parallel to the y faces, pointing down
])
def test_3d_plic_faces_are_exact_plane_square_areas(nrm, p0):
    nrm = np.asarray(nrm, dtype=np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    g = CartesianGrid((-0.61, -0.57, -0.63), (14, 12, 13), (0.09, 0.09, 0.09))
    cols, normals, intercepts, valid = orc.plic(SDFPlane(tuple(nrm), p0), 3, g.x0, g.d, g.n)
    assert valid.all()
    assert int((cols["kind"] == 2).sum()) > 100
    want = exact_plane_faces(g, nrm, p0)
    got = cols["face_fraction"]
    assert np.max(np.abs(got - want)) < 1e-13
    # the boundary slots keep their one side and are exact too
    F = fortran(got, g.n)
    W = fortran(want, g.n)
    for a in range(3):
        lo = [slice(None)] * 3
        lo[a] = 0
        hi = [slice(None)] * 3
        hi[a] = g.n[a] - 1
        assert np.max(np.abs(F[tuple(lo) + (2 * a,)] - W[tuple(lo) + (2 * a,)])) < 1e-13
        assert np.max(np.abs(F[tuple(hi) + (2 * a + 1,)] - W[tuple(hi) + (2 * a + 1,)])) < 1e-13


def fd_gradient(geo, dim, x, h=1e-6):
    g = np.zeros(dim)
    for a in range(dim):
        e = np.zeros(3)
        e[a] = h
        g[a] = (orc.sdf_value(geo, dim, x + e) - orc.sdf_value(geo, dim, x - e)) / (2 * h)
    return g


def test_torus_normals_are_the_distance_gradient():
    rng = np.random.default_rng(11)
    R = np.asarray(random_rotation(rng))
    c, Rmaj, rmin = np.array([0.05, -0.1, 0.02]), 0.5, 0.15
    geo = SDFTorus(tuple(c), Rmaj, rmin, R)
    checked = 0
    for x in rng.uniform(-0.9, 0.9, (600, 3)):
        p = (x - c) @ R  # R^T (x - c)
        rho = np.hypot(p[0], p[1])
        if rho < 0.05 or np.hypot(rho - Rmaj, p[2]) < 0.05:  # the axis and the tube's centre circle
            continue
        dist, nrm = orc.get_sdf(geo, 3, x)
        assert abs(np.linalg.norm(nrm) - 1.0) < 1e-14
        assert np.max(np.abs(nrm - fd_gradient(geo, 3, x))) < 1e-7
        checked += 1
    assert checked > 400


@pytest.mark.parametrize("dim", [2, 3])
def test_box_normals_are_the_distance_gradient(dim):
    rng = np.random.default_rng(12 + dim)
    R = np.asarray(random_rotation(rng))
    if dim == 2:
        t = 0.7
        R = np.array([[np.cos(t), -np.sin(t), 0.0], [np.sin(t), np.cos(t), 0.0], [0.0, 0.0, 1.0]])
    c, h = np.array([-0.1, 0.05, 0.2]), np.array([0.3, 0.2, 0.4])
    geo = SDFBox(tuple(c), tuple(h), R)
    inside = outside = 0
    for x in rng.uniform(-0.9, 0.9, (800, 3)):
        if dim == 2:
            x[2] = 0.0
        p = ((x - c) @ R)[:dim]
        q = np.abs(p) - h[:dim]
        qs = np.sort(q)
        if q.max() > 0:
            if np.linalg.norm(np.maximum(q, 0)) < 1e-3 or np.min(np.abs(p)) < 1e-3:
                continue  # on the surface, or on a plane where sign(p) flips
            outside += 1
        else:
            if qs[-1] - qs[-2] < 1e-3 or q.max() > -1e-3:
                continue  # the medial set, or the surface
            inside += 1
        dist, nrm = orc.get_sdf(geo, dim, x)
        assert abs(np.linalg.norm(nrm) - 1.0) < 1e-14
        assert np.max(np.abs(nrm - fd_gradient(geo, dim, x))) < 1e-7
    assert inside > 20 and outside > 200


def test_a_zero_corner_reads_outside():
    # nudge_zeros (src/marching_cubes/marching_cubes.jl:225-243): |phi| < eps -> +eps, the
    # `phi >= 0 is outside` convention; -0.0 and sub-eps values alike
    for p in range(8):
        for z in (0.0, -0.0, 1e-17, -1e-17):
            phi = np.full(8, 0.3)
            phi[p] = z
            cell = orc.mc_cell(cube_nodes(), phi, (1.0, 1.0, 1.0))
            assert cell["kind"] == 1
            assert cell["volume_fraction"] == 1.0
            assert np.all(cell["face_fraction"] == 1.0)
    # the same corner with negative neighbours is the only outside corner: a cut cell whose fluid
    # corner is eps-sized (its volume, a sum of opposite terms, may round to either sign)
    phi = np.full(8, -0.3)
    phi[6] = 0.0
    cell = orc.mc_cell(cube_nodes(), phi, (1.0, 1.0, 1.0))
    assert cell["kind"] == 2 and abs(cell["volume_fraction"]) < 1e-30


def test_a_plane_through_lattice_nodes():
    # phi = z - z5 with z5 a node layer: layer 5 is exactly zero and nudged outward, so the cells
    # below it are cut with an eps-sized fluid sliver and the cells above it are outside
    g = CartesianGrid((-0.4, -0.35, -0.45), (7, 6, 10), (0.1, 0.1, 0.1))
    z5 = g.x0[2] + 5.0 * g.d[2]
    phi = orc.sample_nodes(SDFPlane((0.0, 0.0, 1.0), (0.0, 0.0, z5)), 3, g.x0, g.d, g.n)
    layer = fortran(phi, tuple(v + 1 for v in g.n))[:, :, 5]
    assert np.all(layer == 0.0)
    cols = orc.mc_cells(phi, g.x0, g.d, g.n)
    kind = fortran(cols["kind"], g.n)
    vf = fortran(cols["volume_fraction"], g.n)
    assert np.all(kind[:, :, :4] == 0) and np.all(vf[:, :, :4] == 0.0)
    assert np.all(kind[:, :, 4] == 2)
    assert np.all((vf[:, :, 4] > 0.0) & (vf[:, :, 4] < 1e-13))
    assert np.all(kind[:, :, 5:] == 1) and np.all(vf[:, :, 5:] == 1.0)
    # the sliver's top face is fully open, its bottom face closed
    ff = fortran(cols["face_fraction"], g.n)
    assert np.all(ff[:, :, 4, 5] == 1.0) and np.all(ff[:, :, 4, 4] == 0.0)


def test_degenerate_centre_faces_are_exactly_closed_2d():
    # test/test_plic.jl:180-206: the centre cell's centroid is the circle's centre (a NaN normal)
    g = CartesianGrid((-1.5, -1.5), (3, 3), (1.0, 1.0))
    cols, normals, intercepts, valid = orc.plic(SDFSphere((0.0, 0.0), 1.0), 2, g.x0, g.d, g.n)
    centre = 1 + 3 * 1
    assert not valid[centre] and int((valid == 0).sum()) == 1
    assert cols["volume_fraction"][centre] == 0.0
    assert np.all(cols["face_fraction"][centre] == 0.0)


def test_degenerate_centre_faces_are_exactly_closed_3d():
    g = CartesianGrid((-1.5, -1.5, -1.5), (3, 3, 3), (1.0, 1.0, 1.0))
    cols, normals, intercepts, valid = orc.plic(SDFSphere((0.0, 0.0, 0.0), 1.0), 3, g.x0, g.d, g.n)
    centre = 1 + 3 * (1 + 3 * 1)
    assert not valid[centre] and int((valid == 0).sum()) == 1
    assert np.all(normals[centre] == 0.0) and intercepts[centre] == 0.0
    assert cols["volume_fraction"][centre] == 0.0
    assert np.all(cols["face_fraction"][centre] == 0.0)
    # a face neighbour's plane touches the shared face along the sphere: its own side is closed too
    assert np.all(np.isfinite(cols["face_fraction"]))
