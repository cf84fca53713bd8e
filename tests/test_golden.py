"""The committed golden fixtures (tests/golden/, written by make_golden.py) against the oracle.

reference_literals.json holds the reference tests' own literal inputs and expectations: each is
checked here on the oracle (and, in test_gpu_golden.py, on the device), `approx` entries with
Julia's default isapprox, `exact` ones bit for bit.  oracle_fields.npz is a regression self-pin: the
inputs must regenerate bit for bit (the reference's xorshift, the sampled bodies) and the oracle must
still compute exactly the stored columns from them.
"""
import json
import os

import numpy as np
import pytest

import oracle as orc
from cutcell import CartesianIsoGrid, SDFSphere, readers
from helpers import sample, sample_vec
from refrng import randn_field, seeded

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def literals():
    with open(os.path.join(GOLDEN, "reference_literals.json")) as f:
        return json.load(f)


def fields():
    return np.load(os.path.join(GOLDEN, "oracle_fields.npz"))


def isapprox(x, y, rtol):
    """Julia's isapprox with atol = 0: ||x - y|| <= rtol * max(||x||, ||y||)."""
    x, y = np.atleast_1d(np.asarray(x, float)), np.atleast_1d(np.asarray(y, float))
    return np.linalg.norm(x - y) <= rtol * max(np.linalg.norm(x), np.linalg.norm(y))


def ms_cell_quantities(m, h):
    ff = np.asarray(m["face_fraction"], float)[None]
    return {"volume_fraction": m["volume_fraction"], "centroid": m["centroid"], "face_fraction": m["face_fraction"],
            "face_area": [readers.face_area_of(ff, k, h)[0] for k in range(4)],
            "face_centroid_local_slots_1_4": [m["face_centroid_local"][0][0], m["face_centroid_local"][3][0]],
            "interface_area": readers.interface_area(ff, h)[0], "interface_normal": readers.interface_normal(ff, h)[0],
            "interface_centroid": m["interface_centroid"], "kind": m["kind"], "ambiguous": m["ambiguous"]}


def check_entry(entry, got, rtol):
    for k, v in entry.get("exact", {}).items():
        assert np.array_equal(np.asarray(got[k]), np.asarray(v)), (entry["cite"], k, got[k], v)
    for k, v in entry.get("approx", {}).items():
        assert isapprox(got[k], v, rtol), (entry["cite"], k, got[k], v)


def linear_phi(entry):
    g, b = entry["phi"]["grad"], entry["phi"]["offset"]
    return lambda x: g[0] * x[0] + g[1] * x[1] + b


def test_literal_hand_cells_on_the_oracle():
    L = literals()
    for e in L["ms_hand_cells"]:
        h = tuple(e["h"])
        phi = sample(linear_phi(e), tuple(e["lo"]), h, (1, 1))
        m = {k: v[0] for k, v in orc.ms_cells(phi, tuple(e["lo"]), h, (1, 1)).items()}
        check_entry(e, ms_cell_quantities(m, h), L["isapprox_rtol"])


def test_literal_flush_cube_on_the_oracle():
    from test_oracle_mc import closure, cube, march_errors
    e = literals()["mc_flush_cube"]
    assert orc.mc_cell_case(e["corners"]) + 1 == e["exact"]["lewiner_case_1based"]
    m = cube(e["corners"])
    assert np.array_equal(closure(m), e["exact"]["closure_residual"])
    err = march_errors(e["corners"], m)
    assert err["area"] < e["below"]["march_area_error"] and err["volume"] < e["below"]["march_volume_error"]


def test_literal_plic_counts_and_sphere_area_on_the_oracle():
    L = literals()
    for e in L["plic_counts"]:
        g = CartesianIsoGrid(tuple(e["lo"]), tuple(e["hi"]), sz=e["sz"])
        cols, nrm, ic, _ = orc.plic(SDFSphere(tuple(e["body"]["centre"]), e["body"]["radius"]), e["dim"], g.x0, g.d, g.n)
        nodes, el = orc.plic_surface(e["dim"], g.x0, g.d, g.n, nrm, ic, 1e-4)
        got = {"nodes": len(nodes), "elements": len(el), "cut_cells": int((cols["kind"] == 2).sum())}
        if e["dim"] == 2:
            got["elements_min_frac_0"] = len(orc.plic_surface(2, g.x0, g.d, g.n, nrm, ic, 0.0)[1])
        check_entry(e, got, L["isapprox_rtol"])
    from test_oracle_mesh import grid3, sphere
    from helpers import tri_vector_areas
    a = L["mc_sphere_area"]
    x0, d, n = grid3(30)
    assert d[0] == a["h"]
    nodes, tris, _ = orc.mc_mesh(sample_vec(sphere, x0, d, n), x0, d, n)
    area = np.linalg.norm(tri_vector_areas(nodes, tris)[0], axis=1).sum()
    assert round(area, 4) == a["area_4_digits"]
    assert round(area / (4 * np.pi) - 1, 4) == a["relative_to_4pi"]


def test_fixture_inputs_regenerate():
    z = fields()
    assert np.array_equal(randn_field(seeded(0x2545F4914F6CDD1D), (9, 9, 9)), z["mc_rand/phi"])
    assert np.array_equal(randn_field(seeded(0x9E3779B97F4A7C15), (17, 13)), z["ms_rand/phi"])
    b = z["mc_sphere/body"]
    sph = SDFSphere(tuple(b[:3]), b[3])
    assert np.array_equal(orc.sample_nodes(sph, 3, z["mc_sphere/x0"], z["mc_sphere/d"], z["mc_sphere/n"]),
                          z["mc_sphere/phi"])


@pytest.mark.parametrize("case", ["mc_rand", "ms_rand", "mc_sphere"])
def test_oracle_still_computes_the_fixture(case):
    z = fields()
    fn = orc.mc_cells if case.startswith("mc") else orc.ms_cells
    cols = fn(z[case + "/phi"], tuple(z[case + "/x0"]), tuple(z[case + "/d"]), tuple(int(v) for v in z[case + "/n"]))
    for k, v in cols.items():
        want = z[case + "/" + k]
        assert v.dtype == want.dtype and np.array_equal(v.view(np.uint8), want.view(np.uint8)), (case, k)


def test_oracle_still_computes_the_plic_fixture():
    z = fields()
    x0, d, n = tuple(z["plic2d/x0"]), tuple(z["plic2d/d"]), tuple(int(v) for v in z["plic2d/n"])
    cols, nrm, ic, valid = orc.plic(SDFSphere((0.0, 0.0), 1.0), 2, x0, d, n)
    for k in ("kind", "volume_fraction", "face_fraction"):
        assert np.array_equal(cols[k], z["plic2d/" + k])
    assert np.array_equal(nrm, z["plic2d/normals"]) and np.array_equal(ic, z["plic2d/intercepts"])
    nodes, el = orc.plic_surface(2, x0, d, n, nrm, ic, 1e-4)
    assert np.array_equal(nodes, z["plic2d/surface_nodes"]) and np.array_equal(el, z["plic2d/surface_elements"])
