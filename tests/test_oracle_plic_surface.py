"""Pins the oracle's PLIC surface (generate_mesh / extract_surface_plic, src/plic/surface.jl) and
calc_volume (src/volume.jl) against the reference's own tests: the golden counts of test_plic.jl
(84 segments on the 2D circle, 3384 nodes and 1688 triangles on the 3D sphere), one segment per cell
with nothing dropped, a plane through cell corners drawing points rather than false diagonals, and
calc_volume against the cache's fluid fractions.  CartesianMeshes' cell_plane_clip is absent, so the
clip is restated from the geometry; these counts are what pin it."""
import numpy as np

import oracle as orc
from cutcell import CartesianIsoGrid, SDFBox, SDFPlane, SDFSphere, SDFTorus, random_rotation


def plic_2d():
    # test/test_plic.jl:25-26
    g = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    return g, SDFSphere((0.0, 0.0), 1.0)


def test_2d_golden_counts():
    g, geo = plic_2d()
    cols, nrm, ic, _ = orc.plic(geo, 2, g.x0, g.d, g.n)
    nodes, el = orc.plic_surface(2, g.x0, g.d, g.n, nrm, ic)
    assert len(el) == int((cols["kind"] == 2).sum()) == 84  # test_plic.jl:122-124
    assert len(nodes) == 2 * len(el)  # never shared
    assert np.array_equal(el, np.arange(1, 2 * len(el) + 1).reshape(-1, 2))
    assert len(orc.plic_surface(2, g.x0, g.d, g.n, nrm, ic, 0.0)[1]) == np.prod(g.n)  # :134
    assert len(orc.plic_surface(2, g.x0, g.d, g.n, nrm, ic, 0.1)[1]) <= len(el)  # :135
    # every segment lies on its cell's plane and is wound with the body on its left-hand side
    cut = np.flatnonzero(cols["kind"] == 2)
    p, q = nodes[0::2], nodes[1::2]
    v = q - p
    rn = np.stack([v[:, 1], -v[:, 0]], axis=1)
    assert np.all(np.einsum("ij,ij->i", rn, nrm[cut]) > 0)
    i, j = cut % g.n[0], cut // g.n[0]
    lo = np.stack([g.x0[0] + i * g.d[0], g.x0[1] + j * g.d[1]], axis=1)
    for x in (p, q):
        xi = (x - lo) / g.d[0]
        assert np.all(np.abs(np.einsum("ij,ij->i", xi, nrm[cut]) - ic[cut]) < 1e-12)


def test_a_line_through_corners_draws_points():
    # test/test_plic.jl:151-180
    g, _ = plic_2d()
    nrm0 = np.array([1.0, 1.0]) / np.sqrt(2.0)
    geo = SDFPlane((nrm0[0], nrm0[1], 0.0), (0.0, 0.0, 0.0))
    cols, nrm, ic, _ = orc.plic(geo, 2, g.x0, g.d, g.n)
    m_nodes, m_el = orc.plic_surface(2, g.x0, g.d, g.n, nrm, ic, 0.0)
    assert len(m_el) == np.prod(g.n)
    cut = cols["kind"] == 2
    seg = m_nodes.reshape(-1, 2, 2)
    length = np.linalg.norm(seg[:, 1] - seg[:, 0], axis=1)
    assert cut.sum() > 0
    assert np.all(cut | (length < 1e-12 * g.d[0]))
    assert np.sum(~cut & (length == 0.0)) > 0
    ref_nodes, ref_el = orc.plic_surface(2, g.x0, g.d, g.n, nrm, ic)
    assert len(ref_el) == cut.sum()
    assert np.array_equal(seg[cut].reshape(-1, 2), ref_nodes)
    assert np.all(np.abs(ref_nodes @ nrm0) < 1e-12)


def test_3d_golden_counts_and_volume():
    # test/test_plic.jl:139-148
    g = CartesianIsoGrid((-1.5,) * 3, (1.5,) * 3, sz=0.15)
    geo = SDFSphere((0.0, 0.0, 0.0), 1.0)
    cols, nrm, ic, _ = orc.plic(geo, 3, g.x0, g.d, g.n)
    nodes, el = orc.plic_surface(3, g.x0, g.d, g.n, nrm, ic)
    assert len(nodes) == 3384 and len(el) == 1688
    domain = np.prod(g.n) * np.prod(g.d)
    vol = orc.calc_volume(geo, 3, g.x0, g.d, g.n)
    assert abs(domain - cols["volume_fraction"].sum() * np.prod(g.d) - vol) < 1e-12 * vol
    assert abs(vol - 4 * np.pi / 3) < 0.01 * 4 * np.pi / 3  # a plane reconstruction contains the sphere
    assert vol > 4 * np.pi / 3
    # every triangle on its cell's plane, wound along the cell's normal (out of the body)
    a, b, c = (nodes[el[:, k] - 1] for k in range(3))
    w = np.cross(b - a, c - a)
    cen = (a + b + c) / 3
    cell = np.floor((cen - np.asarray(g.x0)) / g.d[0]).astype(int)
    lin = cell[:, 0] + g.n[0] * (cell[:, 1] + g.n[1] * cell[:, 2])
    assert np.all(np.einsum("ij,ij->i", w, nrm[lin]) > 0)


def test_calc_volume_2d_and_other_bodies():
    # test/test_plic.jl:75-78 and test/test_bounded.jl:18-19: the circle's area to 1e-3
    g, geo = plic_2d()
    cols, *_ = orc.plic(geo, 2, g.x0, g.d, g.n)
    area = orc.calc_volume(geo, 2, g.x0, g.d, g.n)
    assert abs(np.prod(g.n) * np.prod(g.d) - cols["volume_fraction"].sum() * np.prod(g.d) - area) < 1e-12 * area
    g2 = CartesianIsoGrid((-2.0, -2.0), (2.0, 2.0), sz=0.01)
    assert abs(orc.calc_volume(geo, 2, g2.x0, g2.d, g2.n) - np.pi) < 1e-3 * np.pi
    rng = np.random.default_rng(2)
    g3 = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, sz=0.04)
    box = SDFBox((0.0, 0.0, 0.0), (0.3, 0.2, 0.4), random_rotation(rng))
    assert abs(orc.calc_volume(box, 3, g3.x0, g3.d, g3.n) - 8 * 0.3 * 0.2 * 0.4) < 0.02 * 0.192
    tor = SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15, random_rotation(rng))
    assert abs(orc.calc_volume(tor, 3, g3.x0, g3.d, g3.n) - 2 * np.pi ** 2 * 0.5 * 0.15 ** 2) < 0.01 * 0.222
