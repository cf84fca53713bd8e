"""The device path on the committed golden fixtures (tests/golden/), through the C ABI.

Bar: every column the device computes from a fixture field bit-exact against the stored oracle
columns (xorshift fields on anisotropic offset grids, a sampled sphere from its descriptor, the 2D
PLIC circle and its surface); the reference tests' literal cells and counts reproduced on the device,
`approx` entries with Julia's default isapprox, `exact` ones bit for bit.
"""
import numpy as np
import pytest

from cutcell import (CartesianGrid, CartesianIsoGrid, MarchingCubesCutCell, MarchingSquaresCutCell,
                     PLICCutCell, SDFSphere, allocate_cache, generate_mesh, readers, update_cache)
from test_golden import check_entry, fields, linear_phi, literals, ms_cell_quantities

pytestmark = pytest.mark.gpu

COLS = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction", "face_centroid_local",
        "interface_centroid")


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("case", ["mc_rand", "ms_rand"])
def test_device_reproduces_the_random_field_fixtures(case):
    z = fields()
    g = CartesianGrid(tuple(z[case + "/x0"]), tuple(int(v) for v in z[case + "/n"]), tuple(z[case + "/d"]))
    method = MarchingCubesCutCell if case.startswith("mc") else MarchingSquaresCutCell
    c = update_cache(allocate_cache(g, method), z[case + "/phi"], g)
    for k in COLS:
        assert same(c.read(k), z[case + "/" + k]), (case, k)


def test_device_reproduces_the_sphere_fixture_from_its_descriptor():
    z = fields()
    b = z["mc_sphere/body"]
    g = CartesianGrid(tuple(z["mc_sphere/x0"]), tuple(int(v) for v in z["mc_sphere/n"]), tuple(z["mc_sphere/d"]))
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), SDFSphere(tuple(b[:3]), b[3]), g)
    assert same(c.read("phi"), z["mc_sphere/phi"])
    for k in COLS:
        assert same(c.read(k), z["mc_sphere/" + k]), k


def test_device_reproduces_the_plic_fixture_and_its_surface():
    z = fields()
    g = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    assert g.x0 == tuple(z["plic2d/x0"]) and g.d == tuple(z["plic2d/d"])
    c = update_cache(allocate_cache(g, PLICCutCell), SDFSphere((0.0, 0.0), 1.0), g)
    for k in ("kind", "volume_fraction", "face_fraction"):
        assert same(c.read(k), z["plic2d/" + k]), k
    assert same(c.normals, z["plic2d/normals"]) and same(c.intercepts, z["plic2d/intercepts"])
    assert same(c.is_valid, z["plic2d/is_valid"])
    nodes, el = generate_mesh(c, g)
    assert same(nodes, z["plic2d/surface_nodes"]) and same(el, z["plic2d/surface_elements"])


def test_literal_hand_cells_on_the_device():
    L = literals()
    for e in L["ms_hand_cells"]:
        g = CartesianGrid(tuple(e["lo"]), (1, 1), tuple(e["h"]))
        c = update_cache(allocate_cache(g, MarchingSquaresCutCell), linear_phi(e), g)
        m = {k: c.read(k)[0] for k in COLS}
        check_entry(e, ms_cell_quantities(m, tuple(e["h"])), L["isapprox_rtol"])


def test_literal_flush_cube_on_the_device():
    import oracle as orc
    from helpers import MC_BITS, cube_nodes
    e = literals()["mc_flush_cube"]
    vals = np.empty(8)
    for p, (i, j, k) in enumerate(MC_BITS):  # corner p of MC_NODE_BITS -> node (i, j, k), i fastest
        vals[i + 2 * j + 4 * k] = e["corners"][p]
    g = CartesianGrid((0.0, 0.0, 0.0), (1, 1, 1), (1.0, 1.0, 1.0))
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), vals, g)
    assert c.read("kind")[0] == 2
    ff = c.read("face_fraction")
    assert np.array_equal(readers.closure_residual(ff, g.d)[0], e["exact"]["closure_residual"])
    m = orc.mc_cell(cube_nodes(), e["corners"], (1.0, 1.0, 1.0))
    for k in ("volume_fraction", "centroid", "face_fraction", "face_centroid_local", "interface_centroid"):
        assert np.array_equal(np.asarray(c.read(k)[0]), np.asarray(m[k])), k


def test_literal_plic_counts_on_the_device():
    L = literals()
    for e in L["plic_counts"]:
        g = CartesianIsoGrid(tuple(e["lo"]), tuple(e["hi"]), sz=e["sz"])
        c = update_cache(allocate_cache(g, PLICCutCell), SDFSphere(tuple(e["body"]["centre"]), e["body"]["radius"]), g)
        nodes, el = generate_mesh(c, g)
        got = {"nodes": len(nodes), "elements": len(el), "cut_cells": int((c.read("kind") == 2).sum())}
        if e["dim"] == 2:
            got["elements_min_frac_0"] = len(generate_mesh(c, g, min_tri_area_fraction=0.0)[1])
        check_entry(e, got, L["isapprox_rtol"])
