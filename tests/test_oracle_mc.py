"""Pins the CPU oracle's marching-cubes restatement against what the reference's own tests fix.

Each check here is from ``test/test_mc_moments.jl`` (closed forms, golden literals, invariants,
seeded random fields) or is an independent re-derivation (the numpy "march" below, which places
crossings anchored on the edge's low node as MarchingCubes.jl's march does, and integrates the
closed polyhedron on its own).  A plausible slip in the oracle -- a wrong table offset, a swapped
corner, a dropped term -- fails at least one of them.
"""
import numpy as np
import pytest

import oracle as orc
from cutcell import readers
from helpers import MC_BITS, MC_EDGE, cube_nodes, fortran, sample_vec, shared_face_mismatches
from refrng import XorShift, randn_field

H1 = (1.0, 1.0, 1.0)


def cube(phi, h=H1):
    return orc.mc_cell(cube_nodes(h=h), phi, h)


def marched(phi, m, h=H1):
    """Independent closure of one cell (test/test_mc_moments.jl:103-138): crossings anchored on the
    edge's LOW node, triangles from the Lewiner dispatch, open faces from the stored fractions."""
    phi = np.asarray(phi, dtype=np.float64)
    ph = np.where(np.abs(phi) < np.finfo(float).eps, np.finfo(float).eps, phi)
    ln = cube_nodes(h=h)
    pts = np.zeros((13, 3))
    acc, cnt = np.zeros(3), 0
    for e, (a, b) in enumerate(MC_EDGE):
        if (ph[a] >= 0) != (ph[b] >= 0):
            t = ph[a] / (ph[a] - ph[b])
            pts[e] = ln[a] + t * (ln[b] - ln[a])
            acc += pts[e]
            cnt += 1
    pts[12] = acc / max(cnt, 1)
    n, codes = orc.mc_cell_triangles(phi)
    ff = m["face_fraction"][None, :]
    S = np.zeros(3)
    A = 0.0
    Ac = np.zeros(3)
    vol3 = 0.0
    cm2 = np.zeros(3)
    for ax in range(3):
        a_open = float(readers.face_area_of(ff, 2 * ax + 1, h)[0])
        vol3 += h[ax] * a_open
        cm2[ax] += h[ax] ** 2 * a_open
    for t in range(n):
        a, b, c = pts[codes[t, 0]], pts[codes[t, 1]], pts[codes[t, 2]]
        w = 0.5 * np.cross(b - a, c - a)
        xc = (a + b + c) / 3
        S += w
        ln_ = np.linalg.norm(w)
        A += ln_
        Ac += ln_ * xc
        vol3 -= w @ xc
        cm2 -= w * ((a * a + b * b + c * c + a * b + b * c + c * a) / 6)
    V = vol3 / 3
    return dict(vector_area=S, interface_centroid=Ac / A if A > 0 else None,
                volume_fraction=V / np.prod(h), centroid=cm2 / (2 * V) if V != 0 else None)


def march_errors(phi, m, h=H1):
    r = marched(phi, m, h)
    ina = readers.interface_normal_area(m["face_fraction"][None, :], h)[0]
    out = dict(area=np.max(np.abs(r["vector_area"] - ina)),
               volume=abs(r["volume_fraction"] - m["volume_fraction"]))
    if r["centroid"] is not None:
        out["centroid"] = np.linalg.norm(r["centroid"] - m["centroid"])
    if r["interface_centroid"] is not None:
        out["interface_centroid"] = np.linalg.norm(r["interface_centroid"] - m["interface_centroid"])
    return out


def closure(m, h=H1):
    return readers.closure_residual(m["face_fraction"][None, :], h)[0]


def test_corner_order_matches_lewiner_bits():
    # test/test_mc_moments.jl:206-216
    for p in range(8):
        assert tuple(MC_BITS[p]) == ((p ^ (p >> 1)) & 1, (p >> 1) & 1, (p >> 2) & 1)


def test_every_sign_pattern_closes_against_the_march():
    # test/test_mc_moments.jl:219-244: all 256 sign patterns x 4 magnitudes.
    rng = XorShift(20260919)
    worst = dict(area=0.0, volume=0.0, centroid=0.0, interface_centroid=0.0)
    ncut = 0
    cases = set()
    for mask in range(256):
        for _ in range(4):
            phi = [(1 if (mask >> k) & 1 else -1) * (0.05 + rng.rand()) for k in range(8)]
            m = cube(phi)
            if m["kind"] != 2:
                continue
            ncut += 1
            cases.add(orc.mc_cell_case(phi))
            for k, v in march_errors(phi, m).items():
                worst[k] = max(worst[k], v)
            assert np.all(closure(m) == 0.0)
    assert ncut == 254 * 4
    assert len(cases) == 14
    assert worst["area"] < 1e-14
    assert worst["volume"] < 1e-14
    assert worst["centroid"] < 1e-12
    assert worst["interface_centroid"] < 1e-12


def test_plane_closed_form():
    # test/test_mc_moments.jl:247-260: phi = x - 0.3
    phi = [b[0] - 0.3 for b in MC_BITS]
    m = cube(phi)
    assert m["kind"] == 2
    assert abs(m["volume_fraction"] - 0.7) <= 1e-15
    assert np.allclose(m["centroid"], [0.65, 0.5, 0.5], atol=1e-15, rtol=0)
    ff = m["face_fraction"][None, :]
    assert abs(readers.interface_area(ff, H1)[0] - 1.0) <= 1e-15
    assert np.allclose(readers.interface_normal(ff, H1)[0], [1, 0, 0], atol=1e-15, rtol=0)
    assert np.allclose(m["interface_centroid"], [0.3, 0.5, 0.5], atol=1e-15, rtol=0)
    assert np.allclose(m["face_fraction"], [0, 1, .7, .7, .7, .7], atol=1e-15, rtol=0)
    lower, upper = np.zeros((1, 3)), np.ones((1, 3))
    fc = readers.face_centroid(m["face_centroid_local"][None], 1, lower, upper)[0]
    assert np.allclose(fc, [1.0, 0.5, 0.5], atol=1e-15, rtol=0)


def _plane_sweep(seed, count, h, dfun):
    rng = XorShift(seed)
    worst_vol = worst_nrm = worst_clo = worst_close = 0.0
    ntested = 0
    nodes = cube_nodes(h=h)
    for _ in range(count):
        nv = np.array([rng.randn(), rng.randn(), rng.randn()])
        nv = nv / np.linalg.norm(nv)
        d = dfun(nv, rng)
        phi = [(nv[0] * x[0] + nv[1] * x[1]) + nv[2] * x[2] - d for x in nodes]
        m = orc.mc_cell(nodes, phi, h)
        if m["kind"] != 2:
            continue
        ntested += 1
        scaled = nv * np.array(h)
        ns = np.linalg.norm(scaled)
        ref = 1 - orc.volume_fraction(scaled / ns, d / ns)
        worst_vol = max(worst_vol, abs(m["volume_fraction"] - ref))
        ff = m["face_fraction"][None, :]
        worst_nrm = max(worst_nrm, np.linalg.norm(readers.interface_normal(ff, h)[0] -
                                                  scaled / ns * np.array(h) / np.array(h)))
        worst_clo = max(worst_clo, np.max(np.abs(closure(m, h))))
        worst_close = max(worst_close, march_errors(phi, m, h)["area"])
    return ntested, worst_vol, worst_nrm, worst_clo, worst_close


def test_random_planes_against_the_analytic_volume():
    # test/test_mc_moments.jl:263-290
    ntested, wv, wn, wc, _ = _plane_sweep(4242, 1500, H1, lambda nv, rng: 0.05 + 0.9 * rng.rand())
    assert ntested > 900
    assert wv < 1e-14
    assert wn < 1e-12
    assert wc == 0.0


def test_stretched_cells_and_tilted_planes():
    # test/test_mc_moments.jl:498-552
    h = (2.0, 0.5, 1.5)
    nodes = cube_nodes(h=h)
    m = orc.mc_cell(nodes, [x[0] - 0.5 for x in nodes], h)
    assert abs(m["volume_fraction"] - 0.75) <= 1e-15
    assert np.allclose(m["centroid"], [1.25, 0.25, 0.75], atol=1e-15, rtol=0)
    assert np.allclose(m["face_fraction"], [0, 1, .75, .75, .75, .75], atol=1e-15, rtol=0)
    assert readers.full_face_area(h, 0) == readers.full_face_area(h, 1) == 0.5 * 1.5
    assert readers.full_face_area(h, 2) == 2.0 * 1.5
    assert readers.full_face_area(h, 4) == 2.0 * 0.5
    ff = m["face_fraction"][None, :]
    assert abs(readers.interface_area(ff, h)[0] - 0.75) <= 1e-15
    assert np.all(closure(m, h) == 0.0)
    hv = np.array(h)
    ntested, wv, _, wc, wclose = _plane_sweep(
        3141, 800, h, lambda nv, rng: float(nv @ hv) * (0.1 + 0.8 * rng.rand()))
    assert ntested > 400
    assert wv < 1e-14
    assert wc == 0.0
    assert wclose < 1e-14


def test_uncut_cells_are_exact():
    # test/test_mc_moments.jl:381-399
    out = cube([1.0] * 8)
    inn = cube([-1.0] * 8)
    assert out["kind"] == 1 and not out["ambiguous"]
    assert inn["kind"] == 0 and not inn["ambiguous"]
    assert out["volume_fraction"] == 1.0 and inn["volume_fraction"] == 0.0
    assert np.all(out["face_fraction"] == 1.0) and np.all(inn["face_fraction"] == 0.0)
    assert np.all(closure(out) == 0.0) and np.all(closure(inn) == 0.0)
    fc = readers.face_centroid(out["face_centroid_local"][None], 1, np.zeros((1, 3)),
                               np.ones((1, 3)))[0]
    assert np.allclose(fc, [1.0, 0.5, 0.5], atol=1e-15, rtol=0)


def test_ambiguous_configurations_are_flagged():
    # test/test_mc_moments.jl:402-418
    saddle = cube([1, -1, 1, -1, -1, -1, -1, -1])
    diag = cube([1, -1, 1, -1, -1, 1, -1, 1])
    for m in (saddle, diag):
        assert m["kind"] == 2 and m["ambiguous"]
        assert np.all(closure(m) == 0.0)
        assert 0 < m["volume_fraction"] < 1
    n = np.array([1.0, 2.0, 3.0])
    n /= np.linalg.norm(n)
    plane = cube([n @ b - 0.5 for b in MC_BITS.astype(float)])
    assert not plane["ambiguous"]


def test_flush_cube_golden():
    # test/test_mc_moments.jl:486-494: Lewiner case 10.1.2, reference case index 11 (1-based).
    flush = [-0.6727807769355654, -0.23277645053557436, 1.0026954720119468, 0.22568989049603774,
             1.8172940934743478, 0.02419684567652038, -0.7057160958113278, -2.13709186873607]
    assert orc.mc_cell_case(flush) + 1 == 11
    m = cube(flush)
    e = march_errors(flush, m)
    assert e["area"] < 1e-14 and e["volume"] < 1e-14
    assert np.all(closure(m) == 0.0)


@pytest.fixture(scope="module")
def random_field():
    # test/test_mc_moments.jl:421-495: a 35^3 grid of standard normals from seed 31337.
    rng = XorShift(31337)
    n = (35, 35, 35)
    vals = randn_field(rng, (36, 36, 36))
    cols = orc.mc_cells(vals, (0.0, 0.0, 0.0), H1, n)
    return n, vals, cols


def test_random_field_every_case_closed_single_valued(random_field):
    n, vals, cols = random_field
    kind = cols["kind"]
    assert (kind == 2).sum() > 40_000
    clo = readers.closure_residual(cols["face_fraction"], H1)
    assert np.count_nonzero(clo) == 0
    vf = cols["volume_fraction"]
    assert np.all((vf >= -1e-12) & (vf <= 1 + 1e-12))
    V = fortran(vals, (36, 36, 36))
    cases = set()
    cut = np.flatnonzero(kind == 2)
    I, J, K = np.unravel_index(cut, n, order="F")
    corners = np.stack([V[I + b[0], J + b[1], K + b[2]] for b in MC_BITS], axis=1)
    for row in corners[::5]:
        cases.add(orc.mc_cell_case(row))
    assert len(cases) == 14
    assert shared_face_mismatches(cols["face_fraction"], n, 3) == 0
    assert shared_face_mismatches(cols["face_centroid_local"], n, 3) == 0
    nsaddle = 0
    for ax in range(3):
        p, q = [a for a in range(3) if a != ax]
        sl = [slice(0, 35)] * 3
        sl[ax] = slice(1, 35)
        s = []
        for (u, v) in ((0, 0), (1, 0), (1, 1), (0, 1)):
            off = [0, 0, 0]
            off[p], off[q] = u, v
            s.append(V[tuple(slice(o + (1 if a == ax else 0), o + 35) for a, o in
                             enumerate(off))] >= 0)
        nsaddle += int(((s[0] == s[2]) & (s[1] == s[3]) & (s[0] != s[1])).sum())
    assert nsaddle > 1000
    # Every 7th cut cell against the independent march.
    worst = 0.0
    for c in range(0, cut.size, 7):
        phi = corners[c]
        m = cube(phi)
        assert m["volume_fraction"] == vf[cut[c]]  # a cell's moments are its own corners'
        assert np.array_equal(m["face_fraction"], cols["face_fraction"][cut[c]])
        e = march_errors(phi, m)
        worst = max(worst, e["area"], e["volume"])
    assert worst < 1e-14


def test_watertight_sphere_and_wiggly():
    # test/test_mc_moments.jl:316-343
    N = 24
    x0, d, n = (-1.5,) * 3, (3.0 / N,) * 3, (N, N, N)
    for f in (lambda x, y, z: np.sqrt((x * x + y * y) + z * z) - 1.0,
              lambda x, y, z: np.sin(2.7 * x) * np.sin(3.1 * y) * np.sin(2.3 * z) - 0.12):
        phi = sample_vec(f, x0, d, n)
        cols = orc.mc_cells(phi, x0, d, n)
        assert shared_face_mismatches(cols["face_fraction"], n, 3) == 0
        assert shared_face_mismatches(cols["face_centroid_local"], n, 3) == 0
        assert (cols["kind"] == 2).sum() > 100
        assert np.count_nonzero(readers.closure_residual(cols["face_fraction"], d)) == 0


def test_sphere_volume_converges_and_first_moment():
    # test/test_mc_moments.jl:346-378
    errs = []
    for N in (20, 40, 80):
        x0, d, n = (-1.5,) * 3, (3.0 / N,) * 3, (N, N, N)
        phi = sample_vec(lambda x, y, z: np.sqrt((x * x + y * y) + z * z) - 1.0, x0, d, n)
        cols = orc.mc_cells(phi, x0, d, n)
        outside = cols["volume_fraction"].sum() * np.prod(d)
        errs.append((27.0 - outside - 4 * np.pi / 3) / (4 * np.pi / 3))
    assert all(e < 0 for e in errs)
    assert abs(errs[1]) < abs(errs[0]) / 3.5
    assert abs(errs[2]) < abs(errs[1]) / 3.5
    c0 = np.array([0.13, -0.21, 0.07])
    for N, tol in ((30, 1e-4), (60, 1e-6)):
        x0, d, n = (-1.5,) * 3, (3.0 / N,) * 3, (N, N, N)
        phi = sample_vec(lambda x, y, z: np.sqrt(((x - c0[0]) ** 2 + (y - c0[1]) ** 2) +
                                                 (z - c0[2]) ** 2) - 1.0, x0, d, n)
        cols = orc.mc_cells(phi, x0, d, n)
        v = cols["volume_fraction"] * np.prod(d)
        outv = v.sum()
        outmom = (v[:, None] * cols["centroid"]).sum(axis=0)
        assert np.linalg.norm(-outmom / (27.0 - outv) - c0) < tol
