"""Writes the golden fixtures of this directory (run from the repository root after the oracle is
built; `python tests/golden/make_golden.py`).

reference_literals.json -- literal inputs and expected values of the reference's own tests, each
    entry citing the test file:line it is transcribed from.  Data only; the tests that read it run
    the oracle (test_golden.py) and the device (test_gpu_golden.py) on those inputs.
oracle_fields.npz -- regression self-pins: the reference tests' xorshift normal fields
    (test/test_mc_moments.jl:55-64, regenerated by tests/refrng.py), a sampled sphere and the 2D
    PLIC circle fixture (test/test_plic.jl:25-26), with every column the oracle computes for them.
    They pin nothing against the reference by themselves (the oracle is pinned by the tests listed in
    its header); they freeze the oracle's output so that a later change to it, or to the device path,
    shows up as a bit-level difference.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "cutcellmethods.jl_amd"))

ISAPPROX_RTOL = float(np.sqrt(np.finfo(np.float64).eps))  # Julia's default isapprox for Float64

LITERALS = {
    "about": "literal values of the reference's tests, transcribed; `approx` entries are compared "
             "with Julia's default isapprox (rtol = sqrt(eps)), `exact` entries bit for bit",
    "isapprox_rtol": ISAPPROX_RTOL,
    "ms_hand_cells": [
        {"cite": "test/test_moments.jl:113-128", "phi": {"grad": [0.0, 1.0], "offset": -0.3},
         "lo": [0.0, 0.0], "h": [1.0, 1.0],
         "exact": {"kind": 2, "ambiguous": 0},
         "approx": {"volume_fraction": 0.7, "centroid": [0.5, 0.65], "face_fraction": [0.7, 0.7, 0.0, 1.0],
                    "face_area": [0.7, 0.7, 0.0, 1.0], "face_centroid_local_slots_1_4": [0.65, 0.5],
                    "interface_area": 1.0, "interface_normal": [0.0, 1.0], "interface_centroid": [0.5, 0.3]}},
        {"cite": "test/test_moments.jl:130-135", "phi": {"grad": [1.0, 1.0], "offset": -0.5},
         "lo": [0.0, 0.0], "h": [1.0, 1.0],
         "approx": {"volume_fraction": 0.875, "interface_area": float(np.sqrt(0.5)),
                    "interface_normal": [float(1 / np.sqrt(2)), float(1 / np.sqrt(2))],
                    "interface_centroid": [0.25, 0.25]}},
        {"cite": "test/test_moments.jl:137-146", "phi": {"grad": [1.0, 0.0], "offset": -0.25},
         "lo": [0.0, 0.0], "h": [1.0, 4.0],
         "approx": {"volume_fraction": 0.75, "face_area": [0.0, 4.0, 0.75, 0.75],
                    "face_fraction": [0.0, 1.0, 0.75, 0.75], "interface_area": 4.0}},
    ],
    "mc_flush_cube": {
        "cite": "test/test_mc_moments.jl:486-494",
        "corners": [-0.6727807769355654, -0.23277645053557436, 1.0026954720119468, 0.22568989049603774,
                    1.8172940934743478, 0.02419684567652038, -0.7057160958113278, -2.13709186873607],
        "corner_order": "MC_NODE_BITS: (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1)",
        "exact": {"lewiner_case_1based": 11, "closure_residual": [0.0, 0.0, 0.0]},
        "below": {"march_area_error": 1e-14, "march_volume_error": 1e-14}},
    "plic_counts": [
        {"cite": "test/test_plic.jl:25-26, :118-136", "dim": 2, "lo": [-1.5, -1.5], "hi": [1.5, 1.5], "sz": 0.1,
         "body": {"centre": [0.0, 0.0], "radius": 1.0},
         "exact": {"elements": 84, "nodes": 168, "cut_cells": 84, "elements_min_frac_0": 900}},
        {"cite": "test/test_plic.jl:139-148", "dim": 3, "lo": [-1.5, -1.5, -1.5], "hi": [1.5, 1.5, 1.5], "sz": 0.15,
         "body": {"centre": [0.0, 0.0, 0.0], "radius": 1.0},
         "exact": {"nodes": 3384, "elements": 1688}},
    ],
    "mc_sphere_area": {"cite": "src/marching_cubes/surface.jl:63-66", "h": 0.1, "radius": 1.0,
                       "area_4_digits": 12.5266, "relative_to_4pi": -0.0032},
}


def fields():
    import oracle as orc
    from refrng import randn_field, seeded
    from cutcell import CartesianIsoGrid, SDFSphere
    out = {}
    # xorshift normals, the reference's own generator and seed convention
    rng = seeded(0x2545F4914F6CDD1D)
    n3 = (8, 8, 8)
    x0, d = (-0.5, -0.25, 0.125), (0.125, 0.0625, 0.1875)
    phi = randn_field(rng, tuple(m + 1 for m in n3))
    out["mc_rand/phi"] = phi
    out["mc_rand/x0"], out["mc_rand/d"], out["mc_rand/n"] = np.array(x0), np.array(d), np.array(n3)
    for k, v in orc.mc_cells(phi, x0, d, n3).items():
        out["mc_rand/" + k] = v
    rng = seeded(0x9E3779B97F4A7C15)
    n2 = (16, 12)
    x2, d2 = (-1.0, 0.5), (0.125, 0.25)
    phi2 = randn_field(rng, tuple(m + 1 for m in n2))
    out["ms_rand/phi"] = phi2
    out["ms_rand/x0"], out["ms_rand/d"], out["ms_rand/n"] = np.array(x2), np.array(d2), np.array(n2)
    for k, v in orc.ms_cells(phi2, x2, d2, n2).items():
        out["ms_rand/" + k] = v
    # a sampled sphere, phi and cells
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=12)
    sph = SDFSphere((0.05, -0.03, 0.02), 0.7)
    phs = orc.sample_nodes(sph, 3, g.x0, g.d, g.n)
    out["mc_sphere/phi"] = phs
    out["mc_sphere/x0"], out["mc_sphere/d"], out["mc_sphere/n"] = np.array(g.x0), np.array(g.d), np.array(g.n)
    out["mc_sphere/body"] = np.array([0.05, -0.03, 0.02, 0.7])
    for k, v in orc.mc_cells(phs, g.x0, g.d, g.n).items():
        out["mc_sphere/" + k] = v
    # the 2D PLIC fixture
    gp = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    cols, nrm, ic, valid = orc.plic(SDFSphere((0.0, 0.0), 1.0), 2, gp.x0, gp.d, gp.n)
    out["plic2d/x0"], out["plic2d/d"], out["plic2d/n"] = np.array(gp.x0), np.array(gp.d), np.array(gp.n)
    for k in ("kind", "volume_fraction", "face_fraction"):
        out["plic2d/" + k] = cols[k]
    out["plic2d/normals"], out["plic2d/intercepts"], out["plic2d/is_valid"] = nrm, ic, valid
    nodes, el = orc.plic_surface(2, gp.x0, gp.d, gp.n, nrm, ic, 1e-4)
    out["plic2d/surface_nodes"], out["plic2d/surface_elements"] = nodes, el
    return {k: np.ascontiguousarray(v) for k, v in out.items()}


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_literals.json"), "w") as f:
        json.dump(LITERALS, f, indent=1)
        f.write("\n")
    np.savez_compressed(os.path.join(HERE, "oracle_fields.npz"), **fields())
    print("wrote", os.listdir(HERE))
