"""GPU parity of the device surface march (generate_mesh / cell_indices for marching cubes,
src/marching_cubes/surface.jl:82-141) against the oracle's march, through the C ABI.

Bar: bit-exact -- the same vertex count and order, every coordinate's bits, the same 1-based
triangles in the same order, and cell_indices equal to its host restatement -- on seeded fields that
reach every Lewiner case, bodies of every descriptor kind, ragged and anisotropic grids, moved grids,
and the 512^3 sphere at full size.  Float32 caches are checked by the surface's invariants.
"""
import warnings

import numpy as np
import pytest

import oracle as orc
from cutcell import (ArgumentError, CartesianGrid, CartesianIsoGrid, DimensionMismatch,
                     MarchingCubesCutCell, MarchingSquaresCutCell, PLICCutCell, SDFBox, SDFPlane,
                     SDFSphere, SDFSphereUnion, SDFTorus, Unsupported, allocate_cache, cell_indices,
                     generate_mesh, random_rotation, update_cache)
from cutcell import readers
from helpers import MC_BITS, cell_indices_ref, mesh_topology, sample_vec, tri_vector_areas
from refrng import XorShift, randn_field

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool(np.array_equal(a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32),
                                   b.view(np.uint64 if b.dtype.itemsize == 8 else np.uint32)))
    return bool(np.array_equal(a, b))


def check_against_oracle(cache, grid, phi=None):
    phi = cache.phi if phi is None else phi
    mesh = generate_mesh(cache, grid, warn=False)
    nodes, tris, owner = orc.mc_mesh(phi, grid.x0, grid.d, grid.n)
    assert mesh.nodes.shape == nodes.shape and mesh.elements.shape == tris.shape, \
        (mesh.nodes.shape, nodes.shape, mesh.elements.shape, tris.shape)
    bad = np.flatnonzero(~np.all(mesh.nodes.view(np.uint64) == nodes.view(np.uint64), axis=1))
    assert bad.size == 0, (bad[:5], mesh.nodes[bad[:2]], nodes[bad[:2]])
    assert np.array_equal(mesh.elements, tris)
    cells = cell_indices(mesh, grid)
    assert np.array_equal(cells, cell_indices_ref(nodes, tris, grid.x0, grid.d, grid.n))
    return mesh, owner


def test_sphere_mesh_bitwise_moved_grid_and_determinism():
    # test/test_cache.jl:164-191: the cache's field marched twice gives the same surface; a second
    # cache gives the same surface; a moved grid marches again
    g = CartesianGrid((-1.5,) * 3, (24,) * 3, (0.125,) * 3)
    geo = SDFSphere((0.01, 0.02, -0.03), 1.0)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), geo, g)
    mesh, _ = check_against_oracle(c, g)
    hist, dmax, euler = mesh_topology(mesh.elements)
    assert dmax == 1 and euler == 2 and hist.tolist() == [0, 0, int(hist[2])]
    again = generate_mesh(c, g)
    assert same_bits(again.nodes, mesh.nodes) and np.array_equal(again.elements, mesh.elements)
    other = update_cache(allocate_cache(g, MarchingCubesCutCell), geo, g)
    ref = generate_mesh(other, g)
    assert same_bits(ref.nodes, mesh.nodes) and np.array_equal(ref.elements, mesh.elements)
    moved = g.moved((0.013, -0.007, 0.021))
    update_cache(c, geo, moved)
    check_against_oracle(c, moved)
    with pytest.raises(ArgumentError):  # the device copy of `mesh` has been replaced
        cell_indices(mesh, g)
    with pytest.raises(DimensionMismatch):
        generate_mesh(c, CartesianGrid(g.x0, (25, 24, 24), g.d))


def test_random_field_every_case_bitwise():
    # test/test_mc_moments.jl:421-495: all 14 cases, saddles, interior vertices, flush facets
    rng = XorShift(31337)
    vals = randn_field(rng, (36, 36, 36))
    g = CartesianGrid((0.0,) * 3, (35,) * 3, (1.0,) * 3)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), vals, g)
    mesh, owner = check_against_oracle(c, g, vals)
    assert len(mesh.elements) > 100_000
    # the per-cube triangle order: the oracle's owner of every triangle is nondecreasing
    assert np.all(np.diff(owner) >= 0)


def test_every_sign_pattern_single_cubes():
    rng = XorShift(20260919)
    g = CartesianGrid((0.25, -0.5, 1.0), (1, 1, 1), (0.5, 1.5, 2.0))
    c = allocate_cache(g, MarchingCubesCutCell)
    for mask in range(256):
        for _ in range(2):
            phi = [(1 if (mask >> k) & 1 else -1) * (0.05 + rng.rand()) for k in range(8)]
            vals = np.zeros((2, 2, 2))
            for k in range(8):
                b = MC_BITS[k]
                vals[b[0], b[1], b[2]] = phi[k]
            update_cache(c, vals, g)
            check_against_oracle(c, g)


def test_bodies_ragged_and_anisotropic_grids():
    rng = np.random.default_rng(11)
    R = random_rotation(rng)
    centers = rng.uniform(-0.5, 0.5, (24, 3))
    radii = rng.uniform(0.05, 0.2, 24)
    cases = [
        (CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=40), SDFTorus((0.02, -0.03, 0.01), 0.5, 0.15, R)),
        (CartesianGrid((-1.3, 0.7 - 1.5, -0.45), (33, 17, 21), (0.11, 0.23, 0.07)),
         SDFBox((0.05, 0, 0), (0.4, 0.3, 0.5), R)),
        (CartesianGrid((-1.0, -1.0, -1.0), (70, 3, 65), (2 / 70, 0.4, 2 / 65)),
         SDFSphereUnion(centers, radii, 0.02)),
        # a plane through lattice nodes: exact zeros, nudged to +eps
        (CartesianGrid((-1.0,) * 3, (16, 16, 16), (0.125,) * 3), SDFPlane((0.0, 0.0, 1.0), (0.0, 0.0, 0.0))),
        # x extent above 64 cells per bitmap word, several words per row
        (CartesianGrid((-1.1, -0.9, -1.05), (300, 13, 23), (0.0071, 0.13, 0.091)),
         SDFSphere((0.05, 0.1, -0.02), 0.7)),
    ]
    for g, geo in cases:
        c = update_cache(allocate_cache(g, MarchingCubesCutCell), geo, g)
        phi = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
        assert same_bits(c.phi, phi)
        mesh, _ = check_against_oracle(c, g, phi)
        assert len(mesh.elements) > 0, type(geo).__name__
        _, dmax, _ = mesh_topology(mesh.elements)
        assert dmax == 1


def test_surface_closes_against_the_cells():
    # the device mesh against the device cells: each cut cell's interface vector area is the sum of
    # the vector areas of the triangles cell_indices gives it, except where a facet lies flush in a
    # shared face (test/test_mc_moments.jl:155-186)
    N = 32
    g = CartesianGrid((-1.5,) * 3, (N,) * 3, (3.0 / N,) * 3)
    f = lambda x, y, z: np.sin(2.7 * x) * np.sin(3.1 * y) * np.sin(2.3 * z) - 0.12
    vals = sample_vec(f, g.x0, g.d, g.n)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), vals, g)
    mesh = generate_mesh(c, g)
    cells = cell_indices(mesh, g) - 1
    w, _, _, _ = tri_vector_areas(mesh.nodes, mesh.elements)
    S = np.zeros((g.ncells, 3))
    np.add.at(S, cells, w)
    ina = readers.interface_normal_area(c.cells.face_fraction, g.d)
    cut = np.flatnonzero(c.cells.kind == 2)
    err = np.max(np.abs(S[cut] - ina[cut]), axis=1)
    assert cut.size > 1000
    assert np.mean(err < 1e-14) > 0.99


def test_empty_surface_warns_and_errors():
    g = CartesianGrid((0.0,) * 3, (8, 8, 8), (0.1,) * 3)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), SDFSphere((5.0, 5.0, 5.0), 0.5), g)
    with pytest.warns(UserWarning):
        mesh = generate_mesh(c, g)
    assert mesh.nodes.shape == (0, 3) and mesh.elements.shape == (0, 3)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        generate_mesh(c, g, warn=False)
    assert cell_indices(mesh, g).shape == (0,)
    # a PLIC cache's surface is its planes' shards, not a marching-cubes Surface
    grid = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=8)
    shards = generate_mesh(allocate_cache(grid, PLICCutCell), grid, warn=False)
    with pytest.raises(ArgumentError):
        cell_indices(shards, grid)


def test_float32_surface():
    # a Float32 cache marches in Float32 (surface.jl:84-91 converts into the grid's float type):
    # closed and oriented, each vertex within a few Float32 ulps of the Float64 march where the two
    # fields agree in sign everywhere
    g = CartesianGrid((-1.5,) * 3, (24,) * 3, (0.125,) * 3)
    geo = SDFSphere((0.01, 0.02, -0.03), 1.0)
    c32 = update_cache(allocate_cache(g, MarchingCubesCutCell, np.float32), geo, g)
    m32 = generate_mesh(c32, g)
    assert m32.nodes.dtype == np.float32
    hist, dmax, euler = mesh_topology(m32.elements)
    assert dmax == 1 and euler == 2 and hist.tolist() == [0, 0, int(hist[2])]
    c64 = update_cache(allocate_cache(g, MarchingCubesCutCell), geo, g)
    m64 = generate_mesh(c64, g)
    assert np.array_equal(m32.elements, m64.elements)
    assert np.max(np.abs(m32.nodes.astype(np.float64) - m64.nodes)) < 8 * np.finfo(np.float32).eps * 1.5
    assert np.array_equal(cell_indices(m32, g), cell_indices_ref(m32.nodes, m32.elements, g.x0, g.d, g.n))


def test_sphere_512_full_size_bitwise():
    # the 512^3 marching-cubes workload (sphere r = 0.4, centre U[-0.1, 0.1]^3 from seed 0) at full
    # size: the device surface against the oracle's march of the same field
    rng = np.random.default_rng(0)
    centre = tuple(rng.uniform(-0.1, 0.1, 3))
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=512)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), SDFSphere(centre, 0.4), g)
    mesh, _ = check_against_oracle(c, g)
    hist, dmax, euler = mesh_topology(mesh.elements)
    assert dmax == 1 and euler == 2
    w, a, b, cc = tri_vector_areas(mesh.nodes, mesh.elements)
    area = np.linalg.norm(w, axis=1).sum()
    assert abs(area - 4 * np.pi * 0.16) < 1e-3 * 4 * np.pi * 0.16
