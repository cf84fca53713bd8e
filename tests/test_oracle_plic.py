"""Pins the oracle's PLIC restatement against ``test/test_plic.jl`` and ``test/test_cache.jl``:
the golden cut-cell count (84), the fit conventions, exact faces on a straight boundary, the
degenerate centre fit, the wholly-inside / wholly-outside extremes, single-valued shared faces and
the exact closure, plus the Scardovelli-Zaleski volume fraction against exact rational arithmetic."""
import itertools
import math
from fractions import Fraction

import numpy as np

import oracle as orc
from cutcell import readers
from cutcell.geometry import SDFBox, SDFPlane, SDFSphere
from cutcell.grid import CartesianIsoGrid
from helpers import fortran, shared_face_mismatches
from refrng import XorShift


def run(geo, g):
    return orc.plic(geo, g.dim, g.x0, g.d, g.n)


def exact_volume_fraction(n, a):
    """Inclusion-exclusion over the cube's vertices, in exact rationals (all components nonzero)."""
    n = [Fraction(x) for x in n]
    a = Fraction(a)
    D = len(n)
    tot = Fraction(0)
    for v in itertools.product([0, 1], repeat=D):
        s = a - sum(ni * vi for ni, vi in zip(n, v))
        if s > 0:
            tot += (-1) ** sum(v) * s ** D
    den = math.factorial(D)
    for ni in n:
        den *= ni
    return float(tot / den)


def test_volume_fraction_against_exact_rationals():
    rng = XorShift(777)
    for D in (2, 3):
        worst = 0.0
        for _ in range(400):
            nv = np.array([rng.randn() for _ in range(D)])
            a = (rng.rand() - 0.5) * 3
            ref = exact_volume_fraction(nv, a)
            worst = max(worst, abs(orc.volume_fraction(nv, a) - ref))
        assert worst < 1e-14
    # axis-aligned and degenerate normals
    assert orc.volume_fraction([0.0, 0.0, 1.0], 0.3) == 0.3
    assert orc.volume_fraction([1.0, 0.0], 0.25) == 0.25
    assert orc.volume_fraction([-1.0, 0.0], -0.25) == 0.75


def test_fit_conventions_and_golden_count():
    # test/test_plic.jl:45-80 and :124 (84 cut cells)
    g = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    geo = SDFSphere((0.0, 0.0), 1.0)
    cols, normals, intercepts, valid = run(geo, g)
    assert g.n == (30, 30)
    kind, vf = cols["kind"], cols["volume_fraction"]
    assert int((kind == 2).sum()) == 84
    assert valid.all()
    I, J = np.unravel_index(np.arange(900), g.n, order="F")
    for c in range(0, 900, 7):
        x = np.array([g.x0[0] + (I[c] + 0.5) * g.d[0], g.x0[1] + (J[c] + 0.5) * g.d[1]])
        dist, nrm = orc.get_sdf(geo, 2, x)
        a = nrm.sum() / 2 - dist / g.d[0]
        assert np.array_equal(normals[c], nrm)
        assert math.isclose(intercepts[c], a, rel_tol=1e-14, abs_tol=1e-14)
        assert abs(vf[c] - (1 - exact_volume_fraction(nrm, a))) <= 1e-14
    expect = np.where(vf == 0, 0, np.where(vf == 1, 1, 2))
    assert np.array_equal(kind, expect)
    assert np.all((vf >= 0) & (vf <= 1))
    assert np.all(np.isnan(cols["centroid"])) and np.all(np.isnan(cols["face_centroid_local"]))
    assert np.all(np.isnan(cols["interface_centroid"])) and not cols["ambiguous"].any()
    ff = cols["face_fraction"]
    assert np.all(np.isfinite(ff) & (ff >= 0) & (ff <= 1))
    assert shared_face_mismatches(ff, g.n, 2) == 0
    assert np.count_nonzero(readers.closure_residual(ff, g.d)) == 0


def test_straight_boundary_exact_faces():
    # test/test_cache.jl:228-234
    from cutcell.grid import CartesianGrid
    g = CartesianGrid((-1.3, -1.25), (48, 44), (0.055, 0.055))
    c0 = np.array([0.013, -0.021])
    nrm = np.array([1.0, 0.37])
    nrm /= np.linalg.norm(nrm)
    cols, *_ = run(SDFPlane(tuple(nrm), tuple(c0)), g)
    assert int((cols["kind"] == 2).sum()) > 40

    def line(p):
        return nrm @ (p - c0)

    def edge_open(p0, p1):
        a, b = line(p0), line(p1)
        if a >= 0 and b >= 0:
            return 1.0
        if a < 0 and b < 0:
            return 0.0
        return a / (a - b) if a >= 0 else b / (b - a)

    worst = 0.0
    ff = cols["face_fraction"]
    for c in range(g.ncells):
        i, j = c % g.n[0], c // g.n[0]
        lo = g.get_node((i, j))
        hi = g.get_node((i + 1, j + 1))
        ex = [edge_open(lo, np.array([lo[0], hi[1]])), edge_open(np.array([hi[0], lo[1]]), hi),
              edge_open(lo, np.array([hi[0], lo[1]])), edge_open(np.array([lo[0], hi[1]]), hi)]
        worst = max(worst, np.max(np.abs(ff[c] - ex)))
    assert worst < 1e-13


def test_degenerate_centre_fit():
    # test/test_plic.jl:180-206
    from cutcell.grid import CartesianGrid
    g = CartesianGrid((-1.5, -1.5), (3, 3), (1.0, 1.0))
    cols, normals, intercepts, valid = run(SDFSphere((0.0, 0.0), 1.0), g)
    centre = 1 + 3 * 1
    assert not valid[centre] and int((valid == 0).sum()) == 1
    assert np.all(np.isfinite(intercepts)) and np.all(np.isfinite(normals))
    assert np.all(normals[centre] == 0) and intercepts[centre] == 0
    assert cols["volume_fraction"][centre] == 0
    assert np.all(np.isfinite(cols["face_fraction"]))
    ff = cols["face_fraction"][centre][None]
    assert all(readers.face_area_of(ff, k, g.d)[0] < g.d[0] for k in range(4))


def test_wholly_inside_and_outside():
    # test/test_plic.jl:208-229
    geo = SDFSphere((0.0, 0.0), 1.0)
    inside, *_ = run(geo, CartesianIsoGrid((-0.2, -0.2), (0.2, 0.2), sz=0.1))
    assert np.all(inside["volume_fraction"] == 0) and np.all(inside["face_fraction"] == 0)
    assert np.all(inside["kind"] == 0)
    out, *_ = run(geo, CartesianIsoGrid((10.0, 10.0), (11.0, 11.0), sz=0.25))
    assert np.all(out["volume_fraction"] == 1) and np.all(out["face_fraction"] == 1)
    assert np.all(out["kind"] == 1)


def test_3d_sphere_and_box():
    # test/test_plic.jl:82-93 and test/test_cache.jl:238-245
    from cutcell.grid import CartesianGrid
    g = CartesianIsoGrid((-1.5,) * 3, (1.5,) * 3, sz=0.3)
    cols, *_ = run(SDFSphere((0.0, 0.0, 0.0), 1.0), g)
    ff = cols["face_fraction"]
    assert np.all(np.isfinite(ff) & (ff >= 0) & (ff <= 1))
    g3 = CartesianGrid((-1.5,) * 3, (24,) * 3, (0.125,) * 3)
    cols, *_ = run(SDFSphere((0.01, 0.02, -0.03), 1.0), g3)
    assert int((cols["kind"] == 2).sum()) > 1000
    assert shared_face_mismatches(cols["face_fraction"], g3.n, 3) == 0
    ff = cols["face_fraction"]
    assert np.all((ff >= 0) & (ff <= 1))
    assert np.count_nonzero(readers.closure_residual(ff, g3.d)) == 0
    # a box has medial planes: degenerate fits are flagged, never NaN
    cols, normals, intercepts, valid = run(SDFBox((0.0, 0.0, 0.0), (0.6, 0.6, 0.6)), g3)
    assert np.all(np.isfinite(cols["face_fraction"])) and np.all(np.isfinite(normals))
    assert shared_face_mismatches(cols["face_fraction"], g3.n, 3) == 0
