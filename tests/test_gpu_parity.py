"""GPU parity: the HIP path through the C ABI against the CPU oracle, on the same seeded inputs.

Bar: bit-exact for every column of the nodal methods and of PLIC (the kernels keep the reference's
expression order and are compiled without contraction), plus the reference's own invariants on the
device output: exactly-zero closure, bitwise single-valued shared faces, a moved grid bitwise equal
to a fresh cache, the error types of the host checks.
"""
import numpy as np
import pytest

import oracle as orc
from cutcell import (ArgumentError, CartesianGrid, CartesianIsoGrid, DimensionMismatch, EmptyGeo,
                     MarchingCubesCutCell, MarchingSquaresCutCell, PLICCutCell, SDFBox, SDFPlane,
                     SDFSphere, SDFSphereUnion, SDFTorus, allocate_cache, closure_residual,
                     generate_mesh, random_rotation, update_cache)
from helpers import MC_BITS, DeviceArray, fortran, sample_vec, shared_face_mismatches
from refrng import XorShift, randn_field

pytestmark = pytest.mark.gpu

COLS = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
        "face_centroid_local", "interface_centroid")


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        ua = a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)
        ub = b.view(np.uint64 if b.dtype.itemsize == 8 else np.uint32)
        return bool(np.all((ua == ub) | (np.isnan(a) & np.isnan(b))))
    return bool(np.array_equal(a, b))


def assert_cells_equal(cache, ref, start=0, stop=None):
    for name in COLS:
        got = cache.read(name, start, stop)
        want = ref[name]
        if not same_bits(got, want):
            bad = np.flatnonzero(~np.all((got == want).reshape(got.shape[0], -1), axis=1))
            raise AssertionError(f"{name}: {bad.size} cells differ, first {bad[:5]}: "
                                 f"got {got[bad[:2]]} want {want[bad[:2]]}")


def contour_reference(phi, g, dtype=np.float64):
    """An independent host restatement of generate_mesh on a grid (surface.jl:59-177):
    cell_interface per cut cell in linear order, vertices welded by exact (+0-normalised) keys;
    the arithmetic in ``dtype``, the cache's element type."""
    T = np.dtype(dtype).type
    P = fortran(np.asarray(phi, dtype=dtype), g.nodegrid_size())
    n0, n1 = g.n
    ids, nodes, lines = {}, [], []
    x0, y0, dx, dy = T(g.x0[0]), T(g.x0[1]), T(g.d[0]), T(g.d[1])

    def vid(p):
        key = (float(p[0] + T(0)), float(p[1] + T(0)))
        if key not in ids:
            nodes.append(key)
            ids[key] = len(nodes)
        return ids[key]

    for c in range(n0 * n1):
        i, j = c % n0, c // n0
        corners = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
        ph = [P[a, b] for a, b in corners]
        if len({v >= 0 for v in ph}) == 1:
            continue
        X = [(x0 + T(a) * dx, y0 + T(b) * dy) for a, b in corners]
        pres, sa, sb = [], [], []
        for k in range(4):
            k2 = (k + 1) % 4
            fa, fb = ph[k], ph[k2]
            pres.append(fa >= 0 or fb >= 0)
            if fa >= 0 and fb >= 0:
                sa.append(X[k]); sb.append(X[k2])
            elif fa >= 0:
                fr = fa / (fa - fb)
                sa.append(X[k]); sb.append(tuple(X[k][q] + fr * (X[k2][q] - X[k][q]) for q in range(2)))
            else:
                fr = fb / (fb - fa) if fb >= 0 else 0.0
                sa.append(tuple(X[k2][q] + fr * (X[k][q] - X[k2][q]) for q in range(2))); sb.append(X[k2])
        nseg = 0
        for k in range(4):
            if not pres[k] or nseg == 2:
                continue
            jn = next((k + s) % 4 for s in (1, 2, 3) if pres[(k + s) % 4]) if any(
                pres[(k + s) % 4] for s in (1, 2, 3)) else k
            a, b = sb[k], sa[jn]
            if a != b:
                nseg += 1
                lines.append((vid(b), vid(a)))
    return np.array(nodes, dtype=float).reshape(-1, 2), np.array(lines, dtype=np.int64).reshape(-1, 2)


def test_fresh_cache_reads_as_no_body():
    # test/test_cache.jl:65-74
    g = CartesianGrid((-1.5,) * 3, (24,) * 3, (0.125,) * 3)
    c = allocate_cache(g, MarchingCubesCutCell)
    assert np.all(c.cells.kind == 1)
    assert np.all(c.cells.volume_fraction == 1.0)
    assert np.all(c.cells.face_fraction == 1.0)
    assert np.all(c.phi == 1.0)


def test_mc_sphere_bitwise_and_invariants():
    # test/test_cache.jl:164-191 (CACHE_G3, CACHE_SPHERE)
    g = CartesianGrid((-1.5,) * 3, (24,) * 3, (0.125,) * 3)
    geo = SDFSphere((0.01, 0.02, -0.03), 1.0)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), geo, g)
    phi_ref = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
    assert same_bits(c.phi, phi_ref)  # the sampled field, bit for bit (sqrt correctly rounded)
    ref = orc.mc_cells(phi_ref, g.x0, g.d, g.n)
    assert_cells_equal(c, ref)
    assert (c.cells.kind == 2).sum() > 1000
    assert c.cut_count() == int((ref["kind"] == 2).sum())
    ff = c.cells.face_fraction
    assert shared_face_mismatches(ff, g.n, 3) == 0
    assert shared_face_mismatches(c.cells.face_centroid_local, g.n, 3) == 0
    assert np.count_nonzero(closure_residual(ff, g.d)) == 0
    # a moved grid rewrites every cell: bitwise a fresh cache's
    moved = g.moved((0.013, -0.007, 0.021))
    update_cache(c, geo, moved)
    fresh = update_cache(allocate_cache(moved, MarchingCubesCutCell), geo, moved)
    for name in COLS:
        assert same_bits(c.read(name), fresh.read(name))
    with pytest.raises(DimensionMismatch):
        update_cache(c, geo, CartesianGrid((0.0,) * 3, (10,) * 3, (0.1,) * 3))


def test_mc_random_field_bitwise():
    # test/test_mc_moments.jl:421-495: every Lewiner case, through update_cache! from nodal values
    rng = XorShift(31337)
    vals = randn_field(rng, (36, 36, 36))
    g = CartesianGrid((0.0,) * 3, (35,) * 3, (1.0,) * 3)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), vals, g)
    assert same_bits(c.phi, vals)
    ref = orc.mc_cells(vals, g.x0, g.d, g.n)
    assert_cells_equal(c, ref)
    assert shared_face_mismatches(c.cells.face_fraction, g.n, 3) == 0
    assert shared_face_mismatches(c.cells.face_centroid_local, g.n, 3) == 0
    assert np.count_nonzero(closure_residual(c.cells.face_fraction, g.d)) == 0


def test_update_from_a_device_field():
    # cc_update_from_phi with a device pointer (a caller's field already in HBM): the same cells as
    # the host-array update, and the field kept as phi
    from cutcell import update_cache_from_device_phi
    vals = randn_field(XorShift(4242), (21, 18, 25))
    g = CartesianGrid((-0.5, 0.25, -1.0), (20, 17, 24), (0.1, 0.15, 0.08))
    host = update_cache(allocate_cache(g, MarchingCubesCutCell), vals, g)
    buf = DeviceArray(vals)
    dev = update_cache_from_device_phi(allocate_cache(g, MarchingCubesCutCell), buf.ptr.value, g)
    assert same_bits(dev.phi, vals)
    for name in COLS:
        assert same_bits(dev.read(name), host.read(name)), name
    del dev
    buf.free()


def test_mc_every_sign_pattern_single_cubes():
    # test/test_mc_moments.jl:219-244 on the device: 256 masks x 4 magnitude draws
    rng = XorShift(20260919)
    g = CartesianGrid((0.0,) * 3, (1, 1, 1), (1.0,) * 3)
    c = allocate_cache(g, MarchingCubesCutCell)
    nodes = MC_BITS.astype(float)
    ncut = 0
    for mask in range(256):
        for _ in range(4):
            phi = [(1 if (mask >> k) & 1 else -1) * (0.05 + rng.rand()) for k in range(8)]
            vals = np.zeros((2, 2, 2))
            for k in range(8):
                b = MC_BITS[k]
                vals[b[0], b[1], b[2]] = phi[k]
            update_cache(c, vals, g)
            m = orc.mc_cell(nodes, phi, (1.0, 1.0, 1.0))
            ncut += m["kind"] == 2
            for name in COLS:
                got = c.read(name)[0]
                want = np.asarray(m[name]).reshape(got.shape)
                assert same_bits(got, want), (mask, name, got, want)
    assert ncut == 1016


def test_mc_stretched_and_offset_grid():
    # anisotropic cells off the lattice origin: every face's own full measure
    g = CartesianGrid((-1.3, 0.7, -0.45), (33, 17, 21), (0.11, 0.23, 0.07))
    f = lambda x, y, z: np.sin(2.7 * x) * np.sin(3.1 * y) * np.sin(2.3 * z) - 0.12
    vals = sample_vec(f, g.x0, g.d, g.n)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), vals, g)
    assert_cells_equal(c, orc.mc_cells(vals, g.x0, g.d, g.n))


def test_mc_torus_box_union_geometries():
    rng = np.random.default_rng(5)
    R = random_rotation(rng)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=40)
    centers = rng.uniform(-0.5, 0.5, (24, 3))
    radii = rng.uniform(0.05, 0.2, 24)
    for geo in (SDFTorus((0.02, -0.03, 0.01), 0.5, 0.15, R), SDFBox((0.05, 0, 0), (0.4, 0.3, 0.5), R),
                SDFSphereUnion(centers, radii, 0.02), SDFPlane((0.6, 0.0, 0.8), (0.1, 0.1, 0.1))):
        c = update_cache(allocate_cache(g, MarchingCubesCutCell), geo, g)
        phi_ref = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
        assert same_bits(c.phi, phi_ref), type(geo).__name__
        assert_cells_equal(c, orc.mc_cells(phi_ref, g.x0, g.d, g.n))


def test_ms_circle_config_bitwise_and_contour():
    # MarchingSquaresCutCell on [-1,1]^2, 200 x 200, circle r = 0.5, then generate_mesh
    g = CartesianIsoGrid((-1.0, -1.0), (1.0, 1.0), sz=0.01)
    assert g.n == (200, 200)
    geo = SDFSphere((0.0, 0.0), 0.5)
    c = update_cache(allocate_cache(g, MarchingSquaresCutCell), geo, g)
    phi_ref = orc.sample_nodes(geo, 2, g.x0, g.d, g.n)
    assert same_bits(c.phi, phi_ref)
    ref = orc.ms_cells(phi_ref, g.x0, g.d, g.n)
    assert_cells_equal(c, ref)
    ff = c.cells.face_fraction
    assert shared_face_mismatches(ff, g.n, 2) == 0
    assert np.count_nonzero(closure_residual(ff, g.d)) == 0
    nodes, lines = generate_mesh(c, g)
    want_nodes, want_lines = contour_reference(phi_ref, g)
    assert np.array_equal(nodes, want_nodes) and np.array_equal(lines, want_lines)
    ncut = int((ref["kind"] == 2).sum())
    # a crossing landing exactly on a node leaves a zero-length gap, which draws nothing
    assert ncut - 16 <= lines.shape[0] <= ncut
    counts = np.bincount(lines.reshape(-1), minlength=nodes.shape[0] + 1)[1:]
    assert np.all(counts == 2) and nodes.shape[0] == lines.shape[0]  # one closed loop, bitwise
    r = np.hypot(nodes[:, 0], nodes[:, 1])
    assert np.all(np.abs(r - 0.5) < 2e-3 * 0.5)
    # winding: the right-hand normal points out of the body
    a, b = nodes[lines[:, 0] - 1], nodes[lines[:, 1] - 1]
    v = b - a
    nrm = np.stack([v[:, 1], -v[:, 0]], axis=1)
    assert np.all(np.sum(nrm * (0.5 * (a + b)), axis=1) > 0)
    # the library keeps a contour only from a size call to its fill call: a moved grid, or a field
    # changed through the column's own pointer with no update, is drawn afresh by the next call
    moved = g.moved((0.003, -0.002))
    n2, l2 = generate_mesh(c, moved)
    assert np.array_equal(n2, contour_reference(phi_ref, moved)[0]) and np.array_equal(l2, lines)
    import ctypes
    ptr, _ = c.column_ptr("phi")
    P = fortran(phi_ref, g.nodegrid_size())
    i, j = np.argwhere((P[:-1] > 0) & (P[1:] < 0))[0]  # a node whose +x edge the surface crosses
    node = int(i) + 201 * int(j)
    new = phi_ref.copy()
    new[node] = 0.5 * phi_ref[node]  # the same sign, so the same cut cells
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(ctypes.c_void_p(ptr + 8 * node), new[node:node + 1].ctypes.data, 8, 1) == 0
    assert hip.hipDeviceSynchronize() == 0
    n3, l3 = generate_mesh(c, g)
    w3, wl3 = contour_reference(new, g)
    assert np.array_equal(n3, w3) and np.array_equal(l3, wl3) and not np.array_equal(n3, nodes)


def test_ms_contour_reads_the_field_not_kind():
    # generate_mesh screens cells by their corner signs (_ms_screen_cell, surface.jl:131-135), so a
    # node whose sign is flipped through the column's own pointer -- no update, kind left stale --
    # changes which cells the next contour draws
    import ctypes
    g = CartesianIsoGrid((-1.0, -1.0), (1.0, 1.0), sz=0.01)
    geo = SDFSphere((0.0, 0.0), 0.5)
    c = update_cache(allocate_cache(g, MarchingSquaresCutCell), geo, g)
    phi = c.phi.copy()
    P = fortran(phi, g.nodegrid_size())
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    ptr, _ = c.column_ptr("phi")
    n0, l0 = generate_mesh(c, g)
    # a node 3 cells deep turned positive: its 8 neighbours stay negative (phi is 1-Lipschitz), so
    # the four cells around it are newly cut, one segment each
    i, j = np.argwhere(P < -0.03)[0]
    node = int(i) + 201 * int(j)
    assert np.all(P[i - 1:i + 2, j - 1:j + 2] < 0)
    new = phi.copy()
    new[node] = 1e-3
    assert hip.hipMemcpy(ctypes.c_void_p(ptr + 8 * node), new[node:node + 1].ctypes.data, 8, 1) == 0
    assert hip.hipDeviceSynchronize() == 0
    n1, l1 = generate_mesh(c, g)
    w1, wl1 = contour_reference(new, g)
    assert np.array_equal(n1, w1) and np.array_equal(l1, wl1)
    assert len(l1) == len(l0) + 4  # the stale kind would have drawn none of them


def test_ms_float32_contour():
    # a Float32 cache contours in Float32, as the reference converts into the grid's float type
    g = CartesianIsoGrid((-1.0, -1.0), (1.0, 1.0), sz=0.01)
    c = update_cache(allocate_cache(g, MarchingSquaresCutCell, np.float32), SDFSphere((0.0, 0.0), 0.5), g)
    nodes, lines = generate_mesh(c, g)
    want_nodes, want_lines = contour_reference(c.phi, g, np.float32)
    assert np.array_equal(nodes, want_nodes) and np.array_equal(lines, want_lines) and len(lines) > 300


def test_ms_callable_field_and_errors():
    # test/test_cache.jl:62-107
    g = CartesianGrid((-1.3, -1.25), (48, 44), (0.055, 0.055))
    c = allocate_cache(g, MarchingSquaresCutCell)
    assert np.all(c.cells.kind == 1)
    update_cache(c, lambda x: x[1] - 0.3, g)
    assert int((c.cells.kind == 2).sum()) == g.n[0]
    with pytest.raises(DimensionMismatch):
        update_cache(c, SDFSphere((0, 0), 0.5), CartesianGrid((0.0, 0.0), (10, 10), (0.1, 0.1)))


def test_plic_2d_golden_and_bitwise():
    # test/test_plic.jl:45-124: 84 cut cells
    g = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    geo = SDFSphere((0.0, 0.0), 1.0)
    c = update_cache(allocate_cache(g, PLICCutCell), geo, g)
    ref, normals, intercepts, valid = orc.plic(geo, 2, g.x0, g.d, g.n)
    assert int((c.cells.kind == 2).sum()) == 84
    assert_cells_equal(c, ref)
    assert same_bits(c.normals, normals) and same_bits(c.intercepts, intercepts)
    assert np.array_equal(c.is_valid, valid)
    assert shared_face_mismatches(c.cells.face_fraction, g.n, 2) == 0
    with pytest.raises(ArgumentError):
        allocate_cache(CartesianGrid((-1.5, -1.5), (30, 15), (0.1, 0.2)), PLICCutCell)
    iso = CartesianGrid((-1.5, -1.5), (30, 15), (0.1, 0.1))
    ci = allocate_cache(iso, PLICCutCell)
    assert np.all(ci.cells.kind == 1) and np.all(ci.cells.face_fraction == 1)
    with pytest.raises(ArgumentError):
        update_cache(ci, geo, CartesianGrid((-1.5, -1.5), (30, 15), (0.1, 0.2)))


def test_plic_float32_2d():
    # test/test_plic.jl:238-250: a Float32 grid and body; every column, plane and intercept in
    # Float32, every face fraction finite, every volume fraction in [0, 1]
    f = lambda v: float(np.float32(v))
    g = CartesianGrid((f(-0.6), f(-0.6)), (24, 24), (f(0.05), f(0.05)))
    c = update_cache(allocate_cache(g, PLICCutCell, np.float32), SDFSphere((f(0.02), f(0.03)), f(0.31)), g)
    for name in ("volume_fraction", "face_fraction", "centroid", "face_centroid_local", "interface_centroid"):
        assert c.read(name).dtype == np.float32, name
    assert c.normals.dtype == np.float32 and c.intercepts.dtype == np.float32
    assert np.all(np.isfinite(c.read("face_fraction")))
    vf = c.read("volume_fraction")
    assert np.all((vf >= 0) & (vf <= 1)) and int((c.read("kind") == 2).sum()) > 20


def test_the_uniform_readers_read_the_caches_own_storage():
    # test/test_cache.jl:250-287: every cache answers face_fractions / volume_fractions / kinds from its
    # own storage, is_cell_open agrees with the per-cell reading, and a write into that storage (the
    # device columns) is what the readers then return
    import ctypes
    from cutcell import face_fractions, is_cell_open, kinds, volume_fractions
    g2 = CartesianGrid((-1.3, -1.25), (48, 44), (0.055, 0.055))
    g3 = CartesianGrid((-1.5,) * 3, (24,) * 3, (0.125,) * 3)
    circle, sphere = SDFSphere((0.013, -0.021), 0.83), SDFSphere((0.01, 0.02, -0.03), 1.0)
    ms = update_cache(allocate_cache(g2, MarchingSquaresCutCell), circle, g2)
    mc = update_cache(allocate_cache(g3, MarchingCubesCutCell), sphere, g3)
    pl = update_cache(allocate_cache(g2, PLICCutCell), circle, g2)
    for c, g in ((ms, g2), (mc, g3), (pl, g2)):
        n = int(np.prod(g.n))
        for reader, name, per in ((face_fractions, "face_fraction", 2 * g.dim * 8),
                                  (volume_fractions, "volume_fraction", 8), (kinds, "kind", 1)):
            assert same_bits(reader(c), c.read(name)), name
            assert c.column_ptr(name)[1] == n * per, name  # the column is the cache's storage
        ff = face_fractions(c)
        opened = is_cell_open(c)
        assert np.array_equal(opened, np.array([bool(np.any(row != 0)) for row in ff]))
        assert int((~opened).sum()) > 0
    cell = int(np.flatnonzero(kinds(ms) == 2)[0])
    assert is_cell_open(ms)[cell]
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library links
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    zeros, inside = np.zeros(4), np.array([0], dtype=np.int8)
    ffp, _ = ms.column_ptr("face_fraction")
    kp, _ = ms.column_ptr("kind")
    assert hip.hipMemcpy(ctypes.c_void_p(ffp + 4 * 8 * cell), zeros.ctypes.data, 32, 1) == 0
    assert hip.hipMemcpy(ctypes.c_void_p(kp + cell), inside.ctypes.data, 1, 1) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert np.all(face_fractions(ms)[cell] == 0) and kinds(ms)[cell] == 0
    assert not is_cell_open(ms)[cell]  # it reads the edited fractions


def test_plic_3d_torus_box_degenerate_bitwise():
    rng = np.random.default_rng(11)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=48)
    for geo in (SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15, random_rotation(rng)),
                SDFBox((0.0, 0.0, 0.0), (0.4, 0.4, 0.4)), SDFSphere((0.0, 0.0, 0.0), 0.6)):
        c = update_cache(allocate_cache(g, PLICCutCell), geo, g)
        ref, normals, intercepts, valid = orc.plic(geo, 3, g.x0, g.d, g.n)
        assert_cells_equal(c, ref)
        assert same_bits(c.normals, normals) and np.array_equal(c.is_valid, valid)
        ff = c.cells.face_fraction
        assert shared_face_mismatches(ff, g.n, 3) == 0
        assert np.count_nonzero(closure_residual(ff, g.d)) == 0


def test_float32_variant_against_the_float64_oracle():
    # test/test_cache.jl:290-312 / test_mc_moments.jl:555-568: Float32 end to end
    g = CartesianGrid((-1.5,) * 3, (24,) * 3, (0.125,) * 3)
    geo = SDFSphere((0.01, 0.02, -0.03), 1.0)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell, np.float32), geo, g)
    assert c.cells.volume_fraction.dtype == np.float32
    phi64 = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
    ref = orc.mc_cells(phi64, g.x0, g.d, g.n)
    assert np.array_equal(c.cells.kind, ref["kind"])
    assert np.max(np.abs(c.cells.volume_fraction - ref["volume_fraction"])) < 100 * np.finfo(np.float32).eps
    assert np.max(np.abs(c.cells.face_fraction - ref["face_fraction"])) < 1e-4
    assert shared_face_mismatches(c.cells.face_fraction, g.n, 3) == 0
    assert np.count_nonzero(closure_residual(c.cells.face_fraction, np.float32(g.d))) == 0


@pytest.mark.parametrize("variant,kz,nt", [(0, 16, False), (1, 7, False), (2, 16, False),
                                           (2, 5, True), (1, 64, True), (0, 16, True)])
def test_mc_stream_variants_bitwise(variant, kz, nt):
    # every launch shape of the stream pass gives the same bits, on a grid whose x extent is not a
    # multiple of the 256-cell run and whose z extent is not a multiple of kz
    g = CartesianGrid((-1.1, -0.9, -1.05), (300, 13, 23), (0.0071, 0.13, 0.091))
    geo = SDFSphere((0.05, 0.1, -0.02), 0.7)
    c = allocate_cache(g, MarchingCubesCutCell)
    c.set_tuning(variant, kz, nt)
    assert c.get_tuning() == (variant, kz, nt)
    update_cache(c, geo, g)
    phi_ref = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
    assert same_bits(c.phi, phi_ref)
    assert_cells_equal(c, orc.mc_cells(phi_ref, g.x0, g.d, g.n))
    # and from a caller-sampled field (no body for the fused pass to sample)
    update_cache(c, phi_ref, g)
    assert_cells_equal(c, orc.mc_cells(phi_ref, g.x0, g.d, g.n))


def test_batched_tiny_bodies_match_single_updates():
    # bodies with fewer work items than XCDs (a single cell, a 2 x 3 x 1 PLIC grid, a row of cells)
    # leave some of the eight per-XCD lists empty; each body still gets exactly its own update
    from cutcell.batch import Batch
    specs = [((-0.5, -0.5, -0.5), (1, 1, 1), 1.0, MarchingCubesCutCell, SDFSphere((0.1, 0.0, 0.0), 0.4)),
             ((-0.3, -0.45, -0.15), (2, 3, 1), 0.3, PLICCutCell, SDFSphere((0.05, 0.0, 0.0), 0.35)),
             ((-1.0, -0.1, -0.1), (40, 1, 1), 0.05, MarchingCubesCutCell, SDFSphere((0.2, 0.0, 0.0), 0.5)),
             ((-0.5, -0.5, -0.5), (9, 5, 7), 0.11, PLICCutCell,
              SDFTorus((0.0, 0.0, 0.0), 0.3, 0.1, random_rotation(np.random.default_rng(2))))]
    grids = [CartesianGrid(x0, n, (d,) * 3) for x0, n, d, _, _ in specs]
    caches = [allocate_cache(g, m) for g, (_, _, _, m, _) in zip(grids, specs)]
    geos = [s[4] for s in specs]
    batch = Batch(caches)
    for step in range(2):
        moved = [g.moved((0.01 * step, 0.0, -0.005 * step)) for g in grids]
        batch.step(geos, moved)
        for c, g, geo, (_, _, _, m, _) in zip(caches, moved, geos, specs):
            solo = update_cache(allocate_cache(g, m), geo, g)
            for name in COLS:
                assert same_bits(c.read(name), solo.read(name)), (step, g.n, name)


def test_batched_bodies_match_single_updates_and_oracle():
    # mixed marching-cubes and PLIC bodies on uneven grids, moving: one graph launch per step gives
    # every body exactly what cc_update gives it alone
    from cutcell.batch import Batch
    from cutcell.workloads import MovingBodies
    wl = MovingBodies(seed=7, nbodies=6, nmin=12, nmax=70)
    meth = {"mc": MarchingCubesCutCell, "plic": PLICCutCell}
    caches = [allocate_cache(wl.grid(b, 0), meth[B["method"]]) for b, B in enumerate(wl.bodies)]
    solo = [allocate_cache(wl.grid(b, 0), meth[B["method"]]) for b, B in enumerate(wl.bodies)]
    batch = Batch(caches)
    assert {B["method"] for B in wl.bodies} == {"mc", "plic"}
    for step in range(3):
        geos = [wl.geometry(b, step) for b in range(6)]
        grids = [wl.grid(b, step) for b in range(6)]
        batch.step(geos, grids)
        for b in range(6):
            update_cache(solo[b], geos[b], grids[b])
            for name in COLS:
                assert same_bits(caches[b].read(name), solo[b].read(name)), (step, b, name)
    b = next(i for i, B in enumerate(wl.bodies) if B["method"] == "mc")
    g = wl.grid(b, 2)
    phi_ref = orc.sample_nodes(wl.geometry(b, 2), 3, g.x0, g.d, g.n)
    assert_cells_equal(caches[b], orc.mc_cells(phi_ref, g.x0, g.d, g.n))
    p = next(i for i, B in enumerate(wl.bodies) if B["method"] == "plic")
    g = wl.grid(p, 2)
    ref, *_ = orc.plic(wl.geometry(p, 2), 3, g.x0, g.d, g.n)
    assert_cells_equal(caches[p], ref)
    # a different slab depth rebuilds the graph; results unchanged
    batch.set_kz(b, 3)
    batch.step([wl.geometry(i, 2) for i in range(6)], [wl.grid(i, 2) for i in range(6)])
    assert_cells_equal(caches[b], orc.mc_cells(orc.sample_nodes(wl.geometry(b, 2), 3, wl.grid(b, 2).x0,
                                                                wl.grid(b, 2).d, wl.grid(b, 2).n),
                                               wl.grid(b, 2).x0, wl.grid(b, 2).d, wl.grid(b, 2).n))
    # the batch autotuner gives every body a depth (the PLIC bodies as a class of their own); the
    # tuned step is still what each body gets alone
    from cutcell.autotune import tune_batch
    geos = [wl.geometry(i, 3) for i in range(6)]
    grids = [wl.grid(i, 3) for i in range(6)]
    kz = tune_batch(batch, geos, grids, kz_candidates=(2, 32), classes=2, repeats=1)
    assert sorted(kz) == list(range(6)) and all(k in (2, 4, 8, 32) for k in kz.values())
    batch.step(geos, grids)
    for i in range(6):
        update_cache(solo[i], geos[i], grids[i])
        for name in COLS:
            assert same_bits(caches[i].read(name), solo[i].read(name)), (i, name)


def test_autotuner_picks_a_shape_and_results_stay_identical(tmp_path, monkeypatch):
    from cutcell.autotune import tune_stream
    monkeypatch.setenv("CUTCELL_TUNING", str(tmp_path / "t.json"))
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=96)
    geo = SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15, random_rotation(np.random.default_rng(3)))
    c = allocate_cache(g, MarchingCubesCutCell)
    rec = tune_stream(c, geo, g, repeats=1)
    assert c.get_tuning() == (rec["variant"], rec["kz"], rec["nontemporal"])
    update_cache(c, geo, g)
    phi_ref = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
    assert_cells_equal(c, orc.mc_cells(phi_ref, g.x0, g.d, g.n))
    assert (tmp_path / "t.json").exists()


@pytest.mark.parametrize("variant,kz,n", [(0, 16, (70, 37, 29)), (1, 16, (70, 37, 29)), (1, 5, (70, 37, 29)),
                                         (2, 16, (70, 37, 29)), (2, 5, (70, 37, 29)), (2, 1, (70, 37, 29)),
                                         (2, 7, (131, 17, 20)), (2, 3, (40, 33, 9)), (2, 4, (64, 9, 6)),
                                         (2, 2, (5, 3, 4)), (1, 3, (5, 3, 4)), (2, 1, (1, 1, 1))])
def test_plic_fused_and_two_pass_bitwise(variant, kz, n):
    # the fused single sweeps (variant 1: 32 x 8 tiles, variant 2: 64 x 8 tiles, one wave per row)
    # and the reference's two passes give the oracle's bits, on grids whose extents are not
    # multiples of the tiles or of kz (a last tile row of one row, a last tile column of three
    # cells, a grid narrower than one tile); torus and box (degenerate fits) bodies
    rng = np.random.default_rng(17)
    g = CartesianGrid((-0.93, -1.07, -0.88), n, (0.052,) * 3)
    for geo in (SDFTorus((0.05, -0.02, 0.03), 0.5, 0.15, random_rotation(rng)),
                SDFBox((0.0, 0.0, 0.0), (0.4, 0.3, 0.35), random_rotation(rng)),
                SDFBox((0.026, 0.0, 0.0), (0.4, 0.3, 0.35))):
        c = allocate_cache(g, PLICCutCell)
        c.set_tuning(variant, kz, False)
        update_cache(c, geo, g)
        ref, normals, intercepts, valid = orc.plic(geo, 3, g.x0, g.d, g.n)
        assert_cells_equal(c, ref)
        assert same_bits(c.normals, normals) and same_bits(c.intercepts, intercepts)
        assert np.array_equal(c.is_valid, valid)


def test_plic_float32_3d_every_sweep_agrees():
    # test/test_plic.jl:238-250 in 3D: a Float32 PLIC cache computes in Float32 end to end; the
    # reference's two passes and both fused sweeps are the same expressions in the same order, so
    # they agree bit for bit in Float32 too; faces single-valued, closure exactly zero, and the
    # result close to the Float64 oracle's
    rng = np.random.default_rng(29)
    g = CartesianGrid((-0.93, -1.07, -0.88), (70, 37, 29), (0.052,) * 3)
    names = ("kind", "volume_fraction", "face_fraction", "normals", "intercepts", "is_valid")
    for geo in (SDFTorus((0.05, -0.02, 0.03), 0.5, 0.15, random_rotation(rng)),
                SDFBox((0.0, 0.0, 0.0), (0.4, 0.3, 0.35), random_rotation(rng))):
        runs = []
        for variant, kz in ((0, 16), (1, 16), (2, 16), (2, 5)):
            c = allocate_cache(g, PLICCutCell, np.float32)
            c.set_tuning(variant, kz, False)
            update_cache(c, geo, g)
            runs.append({k: c.read(k) for k in names})
        assert runs[0]["volume_fraction"].dtype == np.float32 and runs[0]["normals"].dtype == np.float32
        for r in runs[1:]:
            for k in names:
                assert same_bits(r[k], runs[0][k]), k
        ff = runs[0]["face_fraction"]
        assert np.all((ff >= 0) & (ff <= 1))
        assert shared_face_mismatches(ff, g.n, 3) == 0
        assert np.count_nonzero(closure_residual(ff, np.float32(g.d))) == 0
        ref, *_ = orc.plic(geo, 3, g.x0, g.d, g.n)
        vf = runs[0]["volume_fraction"].astype(np.float64)
        assert np.max(np.abs(vf - ref["volume_fraction"])) < 1e-4
        assert np.max(np.abs(ff.astype(np.float64) - ref["face_fraction"])) < 1e-3
        assert int(np.count_nonzero(runs[0]["kind"] != ref["kind"])) <= g.ncells // 1000


@pytest.mark.parametrize("nrm", [(1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (1.0, 1.0, 0.0), (1.0, 1.0, 1.0),
                                 (-1.0, 2.0, 2.0), (0.3, -0.4, 1e-17)])
def test_plic_fused_planes_through_lattice_nodes(nrm):
    # planes through a lattice node on a dyadic grid touch many cells exactly at a corner, an edge or
    # a face: where a cell's faces are exactly wet, exactly dry or partly cut is decided at the
    # boundary, and the fused sweep's shortcuts for wholly wet or dry cells must agree with the
    # oracle's full evaluation bit for bit
    nrm = np.asarray(nrm) / np.linalg.norm(nrm)
    g = CartesianGrid((-1.0, -0.75, -0.5), (40, 36, 33), (0.0625,) * 3)
    p0 = (g.x0[0] + 17 * 0.0625, g.x0[1] + 19 * 0.0625, g.x0[2] + 15 * 0.0625)
    geo = SDFPlane(tuple(nrm), p0)
    ref, normals, intercepts, valid = orc.plic(geo, 3, g.x0, g.d, g.n)
    for variant, kz in ((2, 8), (2, 33), (1, 8), (0, 1)):
        c = allocate_cache(g, PLICCutCell)
        c.set_tuning(variant, kz, False)
        update_cache(c, geo, g)
        assert_cells_equal(c, ref)
        assert same_bits(c.normals, normals) and same_bits(c.intercepts, intercepts)
    assert shared_face_mismatches(ref["face_fraction"], g.n, 3) == 0

