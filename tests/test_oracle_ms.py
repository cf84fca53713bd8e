"""Pins the oracle's marching-squares restatement against ``test/test_moments.jl``: hand-computed
cells, exact uncut cells, the saddle, analytic-circle convergence rates, the non-circular normal
check (an independent crossing computation), and bitwise watertightness on an off-lattice grid."""
import math

import numpy as np

import oracle as orc
from cutcell import readers
from helpers import fortran, sample, sample_vec, shared_face_mismatches

UNIT = (1.0, 1.0)


def box_moments(f, lo, hi):
    d = (hi[0] - lo[0], hi[1] - lo[1])
    phi = sample(f, lo, d, (1, 1))
    c = orc.ms_cells(phi, lo, d, (1, 1))
    return {k: v[0] for k, v in c.items()}, d


def ia(m, h):
    return readers.interface_area(m["face_fraction"][None], h)[0]


def inrm(m, h):
    return readers.interface_normal(m["face_fraction"][None], h)[0]


def clo(m, h):
    return readers.closure_residual(m["face_fraction"][None], h)[0]


def test_hand_cells():
    # test/test_moments.jl:111-148
    m, _ = box_moments(lambda x: x[1] - 0.3, (0.0, 0.0), (1.0, 1.0))
    assert m["kind"] == 2 and not m["ambiguous"]
    assert math.isclose(m["volume_fraction"], 0.7, rel_tol=1e-14)
    assert np.allclose(m["centroid"], [0.5, 0.65])
    assert np.allclose(m["face_fraction"], [0.7, 0.7, 0.0, 1.0])
    assert np.allclose(m["face_centroid_local"][:, 0][[0, 3]], [0.65, 0.5])
    assert math.isclose(ia(m, UNIT), 1.0)
    assert np.allclose(inrm(m, UNIT), [0.0, 1.0])
    assert np.allclose(m["interface_centroid"], [0.5, 0.3])

    m2, _ = box_moments(lambda x: x[0] + x[1] - 0.5, (0.0, 0.0), (1.0, 1.0))
    assert math.isclose(m2["volume_fraction"], 1 - 0.125)
    assert math.isclose(ia(m2, UNIT), math.sqrt(0.5))
    assert np.allclose(inrm(m2, UNIT), np.array([1, 1]) / math.sqrt(2))
    assert np.allclose(m2["interface_centroid"], [0.25, 0.25])

    aniso = (1.0, 4.0)
    m3, _ = box_moments(lambda x: x[0] - 0.25, (0.0, 0.0), aniso)
    assert math.isclose(m3["volume_fraction"], 0.75)
    ff = m3["face_fraction"][None]
    assert np.allclose([readers.face_area_of(ff, k, aniso)[0] for k in range(4)], [0, 4, .75, .75])
    assert np.allclose(m3["face_fraction"], [0, 1, .75, .75])
    assert math.isclose(ia(m3, aniso), 4.0)


def test_uncut_cells_exact():
    # test/test_moments.jl:150-169
    small = (0.5, 0.25)
    mo, _ = box_moments(lambda x: 1.0, (1.5, -3.0), (2.0, -2.75))
    assert mo["kind"] == 1 and mo["volume_fraction"] == 1.0
    assert np.all(mo["face_fraction"] == 1.0)
    assert ia(mo, small) == 0.0 and np.all(inrm(mo, small) == 0.0)
    assert np.allclose(mo["centroid"], [1.75, -2.875])
    ff = mo["face_fraction"][None]
    assert readers.face_area_of(ff, 0, small)[0] == 0.25
    assert readers.face_area_of(ff, 2, small)[0] == 0.5
    mi, _ = box_moments(lambda x: -1.0, (1.5, -3.0), (2.0, -2.75))
    assert mi["kind"] == 0 and mi["volume_fraction"] == 0.0 and np.all(mi["face_fraction"] == 0)


def test_sign_convention_complement_and_degenerate():
    # test/test_moments.jl:171-209
    g2 = lambda x: math.hypot(x[0], x[1]) - 0.6
    m4, _ = box_moments(g2, (0.0, 0.0), (1.0, 1.0))
    assert m4["kind"] == 2 and m4["volume_fraction"] > 0.5
    assert np.all(inrm(m4, UNIT) > 0)
    out, _ = box_moments(lambda x: x[1] - 0.3, (0.0, 0.0), (1.0, 1.0))
    inn, _ = box_moments(lambda x: -(x[1] - 0.3), (0.0, 0.0), (1.0, 1.0))
    assert math.isclose(out["volume_fraction"] + inn["volume_fraction"], 1.0)
    assert np.allclose(inrm(inn, UNIT), -inrm(out, UNIT))
    md, _ = box_moments(lambda x: x[1], (0.0, 0.0), (1.0, 1.0))
    for k in ("face_fraction", "centroid"):
        assert np.all(np.isfinite(md[k]))
    assert math.isclose(md["volume_fraction"], 1.0)


def test_saddle_flagged_and_closed():
    # test/test_moments.jl:211-230
    ma, _ = box_moments(lambda x: (x[0] - 0.5) * (x[1] - 0.5), (0.0, 0.0), (1.0, 1.0))
    assert ma["ambiguous"] and ma["kind"] == 2
    assert np.all(clo(ma, UNIT) == 0.0)
    ma2, h = box_moments(lambda x: (x[0] - 0.5) * (x[1] - 2.0), (0.0, 0.0), (1.0, 4.0))
    assert np.all(clo(ma2, h) == 0.0)
    m, _ = box_moments(lambda x: x[0] + 2 * x[1] - 0.9, (0.0, 0.0), (1.0, 1.0))
    ff = m["face_fraction"][None]
    tot = sum(readers.face_vector_area(ff, k, UNIT) for k in range(4))
    assert np.array_equal(tot, readers.interface_normal_area(ff, UNIT))


CENTRE, RADIUS = np.array([0.13, -0.07]), 0.9


def circle(x, y):
    return np.sqrt((x - CENTRE[0]) ** 2 + (y - CENTRE[1]) ** 2) - RADIUS


def circle_errors(n):
    x0, d = (-2.0, -2.0), (4.0 / n, 4.0 / n)
    phi = sample_vec(circle, x0, d, (n, n))
    c = orc.ms_cells(phi, x0, d, (n, n))
    V = d[0] * d[1]
    I, J = np.unravel_index(np.arange(n * n), (n, n), order="F")
    cen = np.stack([x0[0] + (I + 0.5) * d[0], x0[1] + (J + 0.5) * d[1]], axis=1)
    vf = c["volume_fraction"]
    area_inside = np.sum(V - vf * V)
    outside_moment = ((vf * V)[:, None] * c["centroid"]).sum(0)
    domain_moment = (V * cen).sum(0)
    A_int = readers.interface_area(c["face_fraction"], d)
    L_tot = A_int.sum()
    i_mom = (A_int[:, None] * c["interface_centroid"]).sum(0)
    A = math.pi * RADIUS ** 2
    L = 2 * math.pi * RADIUS
    return (abs(area_inside - A) / A,
            np.linalg.norm(outside_moment - (domain_moment - A * CENTRE)) / (A * np.linalg.norm(CENTRE)),
            abs(L_tot - L) / L,
            np.linalg.norm(i_mom - L * CENTRE) / (L * np.linalg.norm(CENTRE)))


def test_analytic_circle_second_order():
    # test/test_moments.jl:234-281 (on the uniform grid the tree reduces to)
    res = (64, 128, 256, 512)
    errs = [circle_errors(n) for n in res]
    for k in range(4):
        e = [errs[j][k] for j in range(len(res))]
        assert e[-1] < 1e-4
        rate = math.log2(e[0] / e[-1]) / (len(e) - 1)
        assert 1.8 < rate < 2.2, (k, e)
        assert all(math.log2(e[j] / e[j + 1]) > 1.4 for j in range(len(e) - 1))


def test_interface_normal_non_circular():
    # test/test_moments.jl:481-524: an independent crossing computation against the
    # closure-derived normal and area.
    n = 128
    x0, d = (-2.0, -2.0), (4.0 / n, 4.0 / n)
    phi = sample_vec(circle, x0, d, (n, n))
    c = orc.ms_cells(phi, x0, d, (n, n))
    P = fortran(phi, (n + 1, n + 1))
    nrm = readers.interface_normal(c["face_fraction"], d)
    area = readers.interface_area(c["face_fraction"], d)
    worst_angle = worst_area = 0.0
    worst_sense = math.inf
    checked = 0
    bits = [(0, 0), (1, 0), (1, 1), (0, 1)]
    for ci in np.flatnonzero((c["kind"] == 2) & (c["ambiguous"] == 0)):
        i, j = ci % n, ci // n
        nodes = [np.array([x0[0] + (i + b[0]) * d[0], x0[1] + (j + b[1]) * d[1]]) for b in bits]
        ph = [P[i + b[0], j + b[1]] for b in bits]
        pts = []
        for k in range(4):
            kk = (k + 1) % 4
            if (ph[k] >= 0) != (ph[kk] >= 0):
                t = ph[k] / (ph[k] - ph[kk])
                pts.append(nodes[k] + t * (nodes[kk] - nodes[k]))
        if len(pts) != 2:
            continue
        v = pts[1] - pts[0]
        nA = np.array([v[1], -v[0]])
        ln = np.linalg.norm(nA)
        if ln <= 1e-13:
            continue
        worst_angle = max(worst_angle, abs(1 - abs(nA / ln @ nrm[ci])))
        worst_area = max(worst_area, abs(ln - area[ci]) / ln)
        worst_sense = min(worst_sense, nrm[ci] @ (c["interface_centroid"][ci] - CENTRE))
        checked += 1
    assert checked > 100
    assert worst_angle < 1e-12
    assert worst_area < 1e-11
    assert worst_sense > 0


def test_watertight_off_lattice():
    # test/test_moments.jl:354-373
    n = (17, 23)
    x0 = (-1.3, 0.7)
    d = ((2.1 - -1.3) / 17, (3.9 - 0.7) / 23)
    phi = sample_vec(lambda x, y: np.sin(3 * x) * np.cos(2 * y) + 0.1, x0, d, n)
    c = orc.ms_cells(phi, x0, d, n)
    assert shared_face_mismatches(c["face_fraction"], n, 2) == 0
    assert shared_face_mismatches(c["face_centroid_local"], n, 2) == 0
    K = fortran(c["kind"], n)
    outside_by_cut = 0
    for ax in range(2):
        a = K[:-1, :] if ax == 0 else K[:, :-1]
        b = K[1:, :] if ax == 0 else K[:, 1:]
        outside_by_cut += int((((a == 1) & (b == 2)) | ((a == 2) & (b == 1))).sum())
    assert outside_by_cut > 50
    assert np.count_nonzero(readers.closure_residual(c["face_fraction"], d)) == 0


def test_contour_against_an_independent_restatement_and_its_shape():
    # generate_mesh over a grid (src/marching_squares/surface.jl:59-177): the oracle's C stitch
    # against the tests' own Python restatement (cell_interface per cut cell in cell order, exact
    # +0-normalised weld), on the 200^2 circle, a random field on an offset anisotropic grid, and a
    # field with exact zeros at nodes (crossings landing on corners, -0 against +0 in the weld)
    from cutcell.grid import CartesianGrid, CartesianIsoGrid
    from refrng import XorShift, randn_field
    from test_gpu_parity import contour_reference
    g = CartesianIsoGrid((-1.0, -1.0), (1.0, 1.0), sz=0.01)
    phi = sample_vec(lambda x, y: np.sqrt(x * x + y * y) - 0.5, g.x0, g.d, g.n)
    nodes, lines = orc.ms_contour(phi, g.x0, g.d, g.n)
    wn, wl = contour_reference(phi, g)
    assert np.array_equal(nodes, wn) and np.array_equal(lines, wl)
    # one closed loop: every vertex ends two segments; wound with the body on the left of b - a
    counts = np.bincount(lines.reshape(-1), minlength=nodes.shape[0] + 1)[1:]
    assert np.all(counts == 2) and nodes.shape[0] == lines.shape[0] > 300
    a, b = nodes[lines[:, 0] - 1], nodes[lines[:, 1] - 1]
    v = b - a
    assert np.all(np.sum(np.stack([v[:, 1], -v[:, 0]], 1) * (0.5 * (a + b)), axis=1) > 0)
    assert np.all(np.abs(np.hypot(nodes[:, 0], nodes[:, 1]) - 0.5) < 1e-3)
    ga = CartesianGrid((-1.3, -0.7), (37, 29), (0.07, 0.05))
    rng = XorShift(4242)
    for field in (randn_field(rng, ga.nodegrid_size()),
                  np.where(np.arange(38 * 30) % 7 == 0, 0.0, randn_field(rng, ga.nodegrid_size())),
                  np.where(np.arange(38 * 30) % 5 == 0, -0.0, randn_field(rng, ga.nodegrid_size()))):
        nodes, lines = orc.ms_contour(field, ga.x0, ga.d, ga.n)
        wn, wl = contour_reference(field, ga)
        assert np.array_equal(nodes, wn) and np.array_equal(lines, wl)
        assert lines.shape[0] > 100
