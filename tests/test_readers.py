"""The host readers (src/cut_cell.jl:155-347) against closed forms: they are the checker of the device
diagnostics, so each is pinned here by a case whose answer the geometry fixes."""
import numpy as np

from cutcell import readers


def test_report_counts_and_smallest_fraction():
    # src/cut_cell.jl:302-317: a cut cell with a zero volume fraction does not set the minimum
    kind = np.array([0, 1, 2, 2, 2, 1], np.int8)
    amb = np.array([0, 0, 1, 0, 1, 0], np.uint8)
    vf = np.array([0.0, 1.0, 0.3, 0.0, 0.7, 1.0])
    r = readers.cut_cell_report(kind, amb, vf)
    assert r == {"n": 6, "cut": 3, "inside": 1, "outside": 2, "ambiguous": 2, "min_volume_fraction": 0.3}
    r = readers.cut_cell_report(np.array([0, 1], np.int8), np.zeros(2, np.uint8), np.array([0.0, 1.0]))
    assert r["cut"] == 0 and r["min_volume_fraction"] == 1.0


def test_interface_vector_area_of_a_half_open_cell():
    # A 2 x 3 cell whose body fills x > 0.5 of it: the -x face open, the +x face closed, the y faces
    # half open.  The interface is the plane x = 1 (length 3) with its normal out of the body along -x.
    ff = np.array([[1.0, 0.0, 0.5, 0.5]])
    v = readers.interface_normal_area(ff, (2.0, 3.0))
    assert np.array_equal(v, [[-3.0, 0.0]])
    assert readers.interface_area(ff, (2.0, 3.0))[0] == 3.0
    assert np.array_equal(readers.interface_normal(ff, (2.0, 3.0)), [[-1.0, 0.0]])
    # 3D, a 1 x 2 x 3 cell whose body is the slab z < 0.75 h_z: the -z face closed, the +z face open,
    # the four side faces a quarter open: the interface is the plane z = 0.75 h_z, area 1 x 2, along +z
    ff3 = np.array([[0.25, 0.25, 0.25, 0.25, 0.0, 1.0]])
    assert np.array_equal(readers.interface_normal_area(ff3, (1.0, 2.0, 3.0)), [[0.0, 0.0, 2.0]])
    assert readers.full_face_area((1.0, 2.0, 3.0), 0) == 6.0 and readers.full_face_area((1.0, 2.0, 3.0), 5) == 2.0
    assert readers.full_face_area((2.0, 3.0), 0) == 3.0 and readers.full_face_area((2.0, 3.0), 2) == 2.0


def test_closure_residual_is_exactly_zero():
    # src/cut_cell.jl:318-347: both sides are the same face_area_of products, so any stored fractions
    # close exactly; a residual built with one face left out must not
    rng = np.random.default_rng(5)
    for D, h in ((2, (0.3, 0.7)), (3, (0.1, 0.2, 0.35))):
        ff = rng.uniform(0, 1, (1000, 2 * D))
        assert np.all(readers.closure_residual(ff, h) == 0.0)
        acc = sum(readers.face_vector_area(ff, k, h) for k in range(2 * D - 1))  # a planted dropped term
        assert np.any(acc - readers.interface_normal_area(ff, h) != 0.0)
