"""GPU parity of the marching-squares cache on a quadtree's leaves (src/marching_squares/cache.jl:91-147)
and its contour (src/marching_squares/surface.jl:92-110), through the C ABI.

Bar: every column and each leaf's four corner values bit-exact against the oracle's leaf restatement
(circle, plane and smooth union; Float64), a uniform tree bitwise the grid cache, the reference's
tree checks (seeded as no body, corners off the lattice, exact closure, one contour segment per cut
leaf agreeing with the moments, other trees and other layouts refused), the contour equal to a host
restatement, and Float32 trees against Float64.
"""
import numpy as np
import pytest

import oracle as orc
from cutcell import (CartesianGrid, DimensionMismatch, MarchingCubesCutCell, MarchingSquaresCutCell,
                     SDFCircle, SDFPlane, SDFSphereUnion, Unsupported, allocate_cache, generate_mesh,
                     readers, update_cache)
from cutcell.tree import AdaptiveMesh
from helpers import (CACHE_CENTRE, CACHE_RADIUS, GEO_CENTRE, GEO_RADIUS, DeviceArray, banded_tree, cache_tree,
                     circle_sdf, same_level_neighbours)

pytestmark = pytest.mark.gpu

COLS = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
        "face_centroid_local", "interface_centroid")


def assert_leaves_equal(cache, mesh, geo):
    phi4, ref = orc.ms_leaves(geo, mesh.base.x0, mesh.base.d, mesh.level, mesh.coord)
    assert np.array_equal(cache.phi.view(np.uint64), phi4.view(np.uint64))
    for name in COLS:
        got, want = cache.read(name), ref[name]
        assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(want).view(np.uint8)), name
    return ref


def contour_reference(cache, mesh):
    """Host restatement of the tree contour: cell_interface (boundary walk, gap segments, outward
    orientation) of every leaf whose stored corners do not all share a sign, in leaf order, welded
    by exact keys (src/marching_squares/surface.jl:92-110)."""
    phi4 = cache.phi
    X = mesh.cell_nodes()
    signs = phi4 >= 0
    ids, nodes, lines = {}, [], []

    def vid(p):
        key = (p[0] + 0.0, p[1] + 0.0)
        if key not in ids:
            nodes.append(key)
            ids[key] = len(nodes)
        return ids[key]

    for c in np.flatnonzero(signs.any(axis=1) & ~signs.all(axis=1)):
        ph, nd = phi4[c], X[c]
        sa, sb, pres = [None] * 4, [None] * 4, [False] * 4
        for k in range(4):
            k2 = (k + 1) % 4
            fa, fb = ph[k], ph[k2]
            oa, ob = fa >= 0, fb >= 0
            pres[k] = oa or ob
            if oa and ob:
                sa[k], sb[k] = nd[k], nd[k2]
            elif oa:
                fr = fa / (fa - fb)
                sa[k], sb[k] = nd[k], nd[k] + fr * (nd[k2] - nd[k])
            elif ob:
                fr = fb / (fb - fa)
                sa[k], sb[k] = nd[k2] + fr * (nd[k] - nd[k2]), nd[k2]
        nseg = 0
        for k in range(4):
            if nseg == 2 or not pres[k]:
                continue
            jn = next((k + s) % 4 for s in range(1, 4) if pres[(k + s) % 4]) if any(
                pres[(k + s) % 4] for s in range(1, 4)) else k
            a, b = sb[k], sa[jn]
            if a[0] != b[0] or a[1] != b[1]:
                nseg += 1
                lines.append((vid(b), vid(a)))
    return np.array(nodes).reshape(-1, 2), np.array(lines, dtype=np.int64).reshape(-1, 2)


def test_leaf_cache_bitwise_and_seeded():
    mesh = banded_tree()
    cache = allocate_cache(mesh, MarchingSquaresCutCell)
    assert cache.on_leaves and cache.phi.shape == (mesh.nleaves, 4)
    assert np.all(cache.read("kind") == 1) and np.all(cache.read("volume_fraction") == 1.0)
    rng = np.random.default_rng(3)
    for geo in (SDFCircle(GEO_CENTRE, GEO_RADIUS), SDFPlane((0.6, 0.8), (0.1, -0.2)),
                SDFSphereUnion(np.c_[rng.uniform(-1, 1, (12, 2)), np.zeros(12)], rng.uniform(0.1, 0.4, 12),
                               0.05)):
        update_cache(cache, geo, mesh)
        ref = assert_leaves_equal(cache, mesh, geo)
        assert cache.cut_count() == int((ref["kind"] == 2).sum()) > 20, type(geo).__name__


def test_tree_invariants_and_contour():
    mesh = banded_tree()
    geo = SDFCircle(GEO_CENTRE, GEO_RADIUS)
    cache = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), geo, mesh)
    # corners off the lattice: the body at each leaf's corners (test/test_cache.jl:134-135)
    assert np.array_equal(cache.phi, circle_sdf(mesh.cell_nodes(), GEO_CENTRE, GEO_RADIUS))
    h = mesh.elem_size()
    ff = cache.read("face_fraction")
    clo = np.array([readers.closure_residual(ff[c:c + 1], tuple(h[c]))[0] for c in range(mesh.nleaves)])
    assert np.count_nonzero(clo) == 0
    pairs = same_level_neighbours(mesh)
    assert len(pairs) > 1000 and all(ff[a, 2 * ax + 1] == ff[b, 2 * ax] for a, b, ax in pairs)
    # the contour off the stored corners: one segment per cut leaf, agreeing with the moments
    # (test/test_moments.jl:420-437)
    kind = cache.read("kind")
    nodes, lines = generate_mesh(cache, mesh)
    cut = np.flatnonzero(kind == 2)
    assert len(lines) == cut.size > 100
    p, q = nodes[lines[:, 0] - 1], nodes[lines[:, 1] - 1]
    ic = cache.read("interface_centroid")[cut]
    assert np.allclose(0.5 * (p + q), ic, rtol=0, atol=1e-14)
    ia = 0.0
    for lv in np.unique(mesh.level[cut]):
        at = cut[mesh.level[cut] == lv]
        ia += readers.interface_area(ff[at], tuple(h[at[0]])).sum()
    assert abs(np.linalg.norm(q - p, axis=1).sum() - ia) < 1e-12 * ia
    rn, rl = contour_reference(cache, mesh)
    assert np.array_equal(nodes, rn) and np.array_equal(lines, rl)
    # wound so the right-hand normal points out of the body
    v = q - p
    nrm = np.stack([v[:, 1], -v[:, 0]], axis=1)
    out = 0.5 * (p + q) - np.asarray(GEO_CENTRE)
    assert np.all(np.einsum("ij,ij->i", nrm, out) > 0)
    # the library keeps a contour between the size and fill calls: a later update, or a refined
    # tree, must draw afresh
    update_cache(cache, SDFCircle((0.2, 0.1), 0.7), mesh)
    n2, l2 = generate_mesh(cache, mesh)
    rn, rl = contour_reference(cache, mesh)
    assert np.array_equal(n2, rn) and np.array_equal(l2, rl) and not np.array_equal(n2[:10], nodes[:10])
    mesh.refine(np.abs(circle_sdf(mesh.elem_centroid(), (0.2, 0.1), 0.7)) < 0.05)
    cache3 = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), SDFCircle((0.2, 0.1), 0.7), mesh)
    n3, l3 = generate_mesh(cache3, mesh)
    rn, rl = contour_reference(cache3, mesh)
    assert np.array_equal(n3, rn) and np.array_equal(l3, rl) and len(l3) > len(l2)


def test_tree_contour_reads_the_field_not_kind():
    # the contour screens each leaf by its corner signs, so a corner value changed through the
    # column's own pointer -- with no update, kind left stale -- is drawn by the next generate_mesh
    import ctypes
    mesh = banded_tree()
    geo = SDFCircle(GEO_CENTRE, GEO_RADIUS)
    cache = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), geo, mesh)
    n0, l0 = generate_mesh(cache, mesh)
    phi4 = cache.phi
    kind = cache.read("kind")
    inside = np.flatnonzero((kind == 0) & np.all(phi4 > -0.05, axis=1))[0]  # an inside leaf by the band
    new = phi4[inside].copy()
    new[2] = -new[2]  # one corner now outside: the leaf is cut, its kind still says inside
    ptr, _ = cache.column_ptr("phi")
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(ctypes.c_void_p(ptr + 32 * int(inside)), new.ctypes.data, 32, 1) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert cache.read("kind")[inside] == 0
    n1, l1 = generate_mesh(cache, mesh)
    rn, rl = contour_reference(cache, mesh)
    assert np.array_equal(n1, rn) and np.array_equal(l1, rl)
    assert len(l1) == len(l0) + 1


def test_tree_held_on_the_device():
    # the leaf list as device arrays (on_device = 1, the reference's tree adapted to the device): the
    # update and the contour are the host tree's bit for bit
    import ctypes
    import cutcell._lib as L
    mesh = banded_tree()
    geo = SDFCircle(GEO_CENTRE, GEO_RADIUS)
    host = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), geo, mesh)
    hn, hl = generate_mesh(host, mesh)
    lv, co = DeviceArray(mesh.level.astype(np.int32)), DeviceArray(mesh.coord.astype(np.int32))
    t = mesh.leaves()
    t.level, t.coord, t.on_device, t.version = lv.ptr, co.ptr, 1, 0
    dev = allocate_cache(mesh, MarchingSquaresCutCell)
    L.check(L.lib().cc_update_leaves(dev.handle, ctypes.byref(L.make_geometry(geo)), ctypes.byref(t)))
    for name in ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction", "face_centroid_local",
                 "interface_centroid", "phi"):
        assert np.array_equal(dev.read(name), host.read(name)), name
    nn, nl = ctypes.c_int64(), ctypes.c_int64()
    L.check(L.lib().cc_generate_mesh_leaves(dev.handle, ctypes.byref(t), None, ctypes.byref(nn), None, ctypes.byref(nl)))
    nodes, lines = np.empty((nn.value, 2)), np.empty((nl.value, 2), dtype=np.int64)
    L.check(L.lib().cc_generate_mesh_leaves(dev.handle, ctypes.byref(t), nodes.ctypes.data, ctypes.byref(nn),
                                            lines.ctypes.data, ctypes.byref(nl)))
    assert np.array_equal(nodes, hn) and np.array_equal(lines, hl)
    del dev
    lv.free()
    co.free()


def test_uniform_tree_is_the_grid_cache():
    base = CartesianGrid((-2.0, -2.0), (8, 8), (0.5, 0.5))
    mesh = AdaptiveMesh(base, max_level=2, initial_level=2)
    g = CartesianGrid((-2.0, -2.0), (32, 32), (0.125, 0.125))
    geo = SDFCircle(GEO_CENTRE, GEO_RADIUS)
    tc = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), geo, mesh)
    gc = update_cache(allocate_cache(g, MarchingSquaresCutCell), geo, g)
    lin = mesh.coord[:, 0] + 32 * mesh.coord[:, 1]
    for name in COLS:
        a, b = tc.read(name), gc.read(name)[lin]
        assert np.array_equal(a.view(np.uint8), np.ascontiguousarray(b).view(np.uint8)), name


def test_callable_field_other_trees_and_layouts():
    tree = cache_tree()
    geo = SDFCircle(CACHE_CENTRE, CACHE_RADIUS)
    cache = update_cache(allocate_cache(tree, MarchingSquaresCutCell), geo, tree)
    by_geo = {n: cache.read(n) for n in COLS}
    # a callable field gives the same leaves as the body when it is the same expression
    update_cache(cache, lambda x: circle_sdf(np.asarray(x), CACHE_CENTRE, CACHE_RADIUS), tree)
    for n in COLS:
        assert np.array_equal(cache.read(n).view(np.uint8), by_geo[n].view(np.uint8)), n
    update_cache(cache, lambda x: x[1] - 0.3, tree)
    assert (cache.read("kind") == 2).sum() > 0
    # another tree has other leaves (test/test_cache.jl:146-150); a tree cache is not a grid cache
    other = AdaptiveMesh(tree.base, max_level=3, initial_level=2)
    with pytest.raises(DimensionMismatch):
        update_cache(cache, geo, other)
    with pytest.raises(DimensionMismatch):
        generate_mesh(cache, other)
    g2 = CartesianGrid((-1.3, -1.25), (48, 44), (0.055, 0.055))
    with pytest.raises(Unsupported):
        update_cache(cache, geo, g2)
    with pytest.raises(Unsupported):
        update_cache(allocate_cache(g2, MarchingSquaresCutCell), geo, tree)
    with pytest.raises(Unsupported):
        allocate_cache(tree, MarchingCubesCutCell)


def test_float32_tree():
    mesh = banded_tree()
    geo = SDFCircle(GEO_CENTRE, GEO_RADIUS)
    c32 = update_cache(allocate_cache(mesh, MarchingSquaresCutCell, np.float32), geo, mesh)
    c64 = update_cache(allocate_cache(mesh, MarchingSquaresCutCell), geo, mesh)
    assert c32.read("volume_fraction").dtype == np.float32
    assert np.array_equal(c32.read("kind"), c64.read("kind"))
    assert np.max(np.abs(c32.read("volume_fraction") - c64.read("volume_fraction"))) < 1e-4
    h = mesh.elem_size().astype(np.float32)
    ff = c32.read("face_fraction")
    clo = np.array([readers.closure_residual(ff[c:c + 1], tuple(h[c]))[0] for c in range(mesh.nleaves)])
    assert np.count_nonzero(clo) == 0
    nodes, lines = generate_mesh(c32, mesh)
    assert len(lines) == int((c32.read("kind") == 2).sum())
