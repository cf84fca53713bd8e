"""GPU checks of the device-side diagnostics (cut_cell_report, closure_residual, interface areas,
shared faces, a NaN audit; src/cut_cell.jl:190-347) and of the VTK / .npy writers, through the C ABI.

Bar: the report and every count, minimum and maximum equal to a host computation on the copied
columns; the interface totals to 1e-12 of the host's; a changed face fraction found as exactly one
mismatched face; the written files read back bit for bit.
"""
import ctypes
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from cutcell import (CartesianGrid, CartesianIsoGrid, MarchingCubesCutCell, MarchingSquaresCutCell,
                     PLICCutCell, SDFSphere, SDFTorus, Unsupported, allocate_cache, cut_cell_report,
                     diagnostics, generate_mesh, random_rotation, readers, update_cache, write_cache_vtk,
                     write_npy)
from helpers import GEO_CENTRE, GEO_RADIUS, banded_tree

pytestmark = pytest.mark.gpu


def host_report(c):
    kind, amb, vf = c.read("kind"), c.read("ambiguous"), c.read("volume_fraction")
    pos = vf[(kind == 2) & (vf > 0)]
    return {"n": kind.size, "cut": int((kind == 2).sum()), "inside": int((kind == 0).sum()),
            "outside": int((kind == 1).sum()), "ambiguous": int((amb != 0).sum()),
            "min_volume_fraction": min(1.0, float(pos.min())) if pos.size else 1.0}


def check(c, g):
    d = diagnostics(c, g)
    r = host_report(c)
    for k, v in r.items():
        assert d[k] == v, (k, d[k], v)
    assert d["closure_nonzero"] == 0 and d["closure_max"] == 0.0
    assert d["face_mismatches"] == 0 and d["centroid_mismatches"] == 0 and d["nonfinite"] == 0
    ff = c.read("face_fraction")
    ina = readers.interface_normal_area(ff, g.d)
    area = readers.interface_area(ff, g.d).sum()
    assert abs(d["interface_area"] - area) <= 1e-12 * max(area, 1.0)
    assert np.allclose(d["interface_normal_area"], ina.sum(axis=0), rtol=0, atol=1e-12 * max(area, 1.0))
    return d


def test_reports_and_diagnostics_of_every_method():
    g3 = CartesianGrid((-1.5,) * 3, (40, 36, 44), (0.075,) * 3)
    mc = update_cache(allocate_cache(g3, MarchingCubesCutCell), SDFSphere((0.01, 0.02, -0.03), 1.0), g3)
    d = check(mc, g3)
    assert d["cut"] > 1000
    # the interface of a closed surface inside the grid sums to zero vector area
    assert np.all(np.abs(d["interface_normal_area"]) < 1e-10)
    g2 = CartesianGrid((-1.3, -1.25), (48, 44), (0.055, 0.055))
    check(update_cache(allocate_cache(g2, MarchingSquaresCutCell), SDFSphere((0.013, -0.021), 0.83), g2), g2)
    gp = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=48)
    rng = np.random.default_rng(4)
    check(update_cache(allocate_cache(gp, PLICCutCell), SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15,
                                                                 random_rotation(rng)), gp), gp)
    # the report alone, for a cache on a tree's leaves
    tree = banded_tree()
    tc = update_cache(allocate_cache(tree, MarchingSquaresCutCell), SDFSphere(GEO_CENTRE, GEO_RADIUS), tree)
    assert cut_cell_report(tc) == host_report(tc)
    with pytest.raises(Unsupported):
        diagnostics(tc, g2)


def test_launch_shapes_and_float32():
    # 2 x 16 x 2 strips in 16 blocks (renumbered across the XCDs); 64 x 263 strips, more than the
    # 2048 blocks x 4 waves, so waves loop over several; the Float32 variant
    g = CartesianGrid((-1.0, -1.0, -0.5), (128, 128, 64), (1 / 64, 1 / 64, 1 / 64))
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), SDFSphere((0.01, 0.02, -0.03), 0.45), g)
    check(c, g)
    g2 = CartesianGrid((-1.0, -0.52), (4096, 2100), (1 / 2048, 1 / 2048))
    check(update_cache(allocate_cache(g2, MarchingSquaresCutCell), SDFSphere((0.001, 0.002), 0.4), g2), g2)
    c32 = update_cache(allocate_cache(g, MarchingCubesCutCell, np.float32), SDFSphere((0.01, 0.02, -0.03), 0.45), g)
    d = diagnostics(c32, g)
    r = host_report(c32)
    for k, v in r.items():
        assert d[k] == v, (k, d[k], v)
    assert d["closure_nonzero"] == 0 and d["face_mismatches"] == 0 and d["centroid_mismatches"] == 0
    area = readers.interface_area(c32.read("face_fraction").astype(np.float64), g.d).sum()
    assert abs(d["interface_area"] - area) <= 1e-5 * area
    assert cut_cell_report(c32) == r


def test_a_changed_face_is_found():
    g = CartesianGrid((-1.5,) * 3, (24,) * 3, (0.125,) * 3)
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), SDFSphere((0.01, 0.02, -0.03), 1.0), g)
    ptr, _ = c.column_ptr("face_fraction")
    cell = int(np.flatnonzero(c.read("kind") == 2)[0])
    assert cell % 24 + 1 < 24  # the +x neighbour exists
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library links
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    v = np.array([0.123456], dtype=np.float64)
    assert hip.hipMemcpy(ctypes.c_void_p(ptr + (6 * cell + 1) * 8), v.ctypes.data, 8, 1) == 0  # host to device
    assert hip.hipDeviceSynchronize() == 0  # a pageable copy may return before it lands
    d = diagnostics(c, g)
    assert d["face_mismatches"] == 1 and d["centroid_mismatches"] == 0 and d["nonfinite"] == 0
    # a NaN volume fraction in an uncut cell: the audit counts it, nothing else moves
    other = int(np.flatnonzero(c.read("kind") == 1)[0])
    vptr, _ = c.column_ptr("volume_fraction")
    nan = np.array([np.nan])
    assert hip.hipMemcpy(ctypes.c_void_p(vptr + other * 8), nan.ctypes.data, 8, 1) == 0
    assert hip.hipDeviceSynchronize() == 0
    d2 = diagnostics(c, g)
    assert d2["nonfinite"] == 1 and d2["face_mismatches"] == 1
    assert {k: d2[k] for k in ("cut", "inside", "outside", "min_volume_fraction")} == \
        {k: d[k] for k in ("cut", "inside", "outside", "min_volume_fraction")}


def test_planted_faces_at_every_seam_are_found():
    # one changed face fraction or face centroid per seam of the pass's march, each a single mismatch:
    # +x inside a strip and from lane 63 to the next strip; +y inside a strip, between the strips of
    # a band (wave to wave) and past a band (16 rows, read again); +z inside a band of layers and past
    # it (64 layers); rows and lanes past the grid's edges on a 70 x 40 x 140 grid
    g = CartesianGrid((-1.05, -0.6, -2.1), (70, 40, 140), (0.03, 0.03, 0.03))
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), SDFSphere((0.0, 0.0, 0.0), 0.5), g)
    assert check(c, g)["face_mismatches"] == 0
    n0, n1 = g.n[0], g.n[1]
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library links
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    ffp, _ = c.column_ptr("face_fraction")
    fcp, _ = c.column_ptr("face_centroid_local")
    ff = c.read("face_fraction")
    fc = c.read("face_centroid_local").reshape(-1, 12)

    def cell(i, j, k):
        return i + n0 * (j + n1 * k)

    def plant(ptr, offset, old):
        v = np.array([old + 0.123456], dtype=np.float64)
        assert hip.hipMemcpy(ctypes.c_void_p(ptr + offset * 8), v.ctypes.data, 8, 1) == 0

    faces = [(10, 20, 70, 1), (63, 20, 70, 1), (30, 5, 70, 3), (30, 3, 71, 3), (30, 7, 72, 2),
             (30, 15, 73, 3), (31, 35, 74, 3), (30, 20, 10, 5), (30, 21, 63, 5), (30, 22, 64, 4),
             (69, 39, 139, 0)]
    for i, j, k, slot in faces:
        c0 = cell(i, j, k)
        plant(ffp, 6 * c0 + slot, ff[c0, slot])
    cents = [(62, 11, 50, 1), (40, 11, 50, 3), (41, 31, 127, 5)]  # +x, wave seam +y, band-of-layers +z
    for i, j, k, slot in cents:
        c0 = cell(i, j, k)
        plant(fcp, 12 * c0 + 2 * slot + 1, fc[c0, 2 * slot + 1])
    assert hip.hipDeviceSynchronize() == 0  # a pageable copy may return before it lands
    d = diagnostics(c, g)
    assert d["face_mismatches"] == len(faces) and d["centroid_mismatches"] == len(cents), d


def read_vtk(path):
    """The arrays of a VTK XML file with raw appended data and UInt64 block headers, by name (the
    points of an unstructured grid as "points"); the header must parse as XML."""
    raw = open(path, "rb").read()
    head, rest = raw.split(b"<AppendedData", 1)
    root = ET.fromstring(head.decode() + "</VTKFile>")
    assert root.get("header_type") == "UInt64" and root.get("byte_order") == "LittleEndian"
    data = rest[rest.index(b"_") + 1:]
    dts = {"Float64": np.float64, "Float32": np.float32, "Int32": np.int32, "Int64": np.int64, "UInt8": np.uint8}
    out = {}
    for el in root.iter("DataArray"):
        off = int(el.get("offset"))
        nb = int(np.frombuffer(data[off:off + 8], np.uint64)[0])
        arr = np.frombuffer(data[off + 8:off + 8 + nb], dts[el.get("type")])
        ncomp = int(el.get("NumberOfComponents", "1"))
        out[el.get("Name", "points")] = arr.reshape(-1, ncomp) if ncomp > 1 else arr
    return root, out


def test_vtk_and_npy_read_back(tmp_path):
    g = CartesianGrid((-1.5, -1.4, -1.3), (20, 22, 24), (0.15, 0.13, 0.11))
    c = update_cache(allocate_cache(g, MarchingCubesCutCell), SDFSphere((0.01, 0.02, -0.03), 1.0), g)
    files = write_cache_vtk(str(tmp_path / "mc"), c, g)
    root, cells = read_vtk(files[0])
    img = root.find("ImageData")
    assert img.get("WholeExtent") == "0 20 0 22 0 24"
    assert tuple(float(v) for v in img.get("Spacing").split()) == g.d
    assert tuple(float(v) for v in img.get("Origin").split()) == g.x0
    assert np.array_equal(cells["volume_fraction"], c.read("volume_fraction"))
    assert np.array_equal(cells["kind"], c.read("kind").astype(np.int32))
    ff = c.read("face_fraction")
    for k, nm in enumerate(("-x", "+x", "-y", "+y", "-z", "+z")):
        assert np.array_equal(cells["face_fraction_" + nm], ff[:, k])
    assert np.array_equal(cells["centroid"], c.read("centroid"))
    assert np.array_equal(cells["phi"], c.read("phi"))
    _, surf = read_vtk(files[1])
    mesh = generate_mesh(c, g)
    assert np.array_equal(surf["points"], mesh.nodes)
    assert np.array_equal(surf["connectivity"].reshape(-1, 3), mesh.elements - 1)
    assert np.all(surf["types"] == 5) and np.array_equal(surf["offsets"], 3 * np.arange(1, len(mesh.elements) + 1))
    # .npy in the reference's layout
    p = write_npy(c, "face_fraction", tmp_path / "ff.npy")
    arr = np.load(p)
    assert arr.shape == (6, 20, 22, 24) and np.isfortran(arr)
    assert np.array_equal(arr.reshape(6, -1, order="F").T, ff)
    fcl = np.load(write_npy(c, "face_centroid_local", tmp_path / "fcl.npy"))
    assert fcl.shape == (2, 6, 20, 22, 24)
    assert np.array_equal(fcl.reshape(2, 6, -1, order="F").transpose(2, 1, 0), c.read("face_centroid_local"))
    phi = np.load(write_npy(c, "phi", tmp_path / "phi.npy"))
    assert phi.shape == (21, 23, 25) and np.array_equal(phi.reshape(-1, order="F"), c.read("phi"))
    kind = np.load(write_npy(c, "kind", tmp_path / "kind.npy"))
    assert kind.dtype == np.int8 and np.array_equal(kind.reshape(-1, order="F"), c.read("kind"))
    # a 2D PLIC cache: planes as cell data, line shards
    gp = CartesianIsoGrid((-1.5, -1.5), (1.5, 1.5), sz=0.1)
    cp = update_cache(allocate_cache(gp, PLICCutCell), SDFSphere((0.0, 0.0), 1.0), gp)
    pf = write_cache_vtk(str(tmp_path / "plic"), cp, gp)
    _, pc = read_vtk(pf[0])
    assert np.array_equal(pc["normal"], cp.normals) and np.array_equal(pc["intercept"], cp.intercepts)
    _, ps = read_vtk(pf[1])
    nodes, el = generate_mesh(cp, gp)
    assert len(ps["types"]) == len(el) == 84 and np.all(ps["types"] == 3)
    assert np.array_equal(ps["points"][:, :2], nodes) and np.all(ps["points"][:, 2] == 0)
    # a tree cache writes .npy (leaf order), not image data
    tree = banded_tree()
    tc = update_cache(allocate_cache(tree, MarchingSquaresCutCell), SDFSphere(GEO_CENTRE, GEO_RADIUS), tree)
    p4 = np.load(write_npy(tc, "phi", tmp_path / "p4.npy"))
    assert p4.shape == (4, tree.nleaves) and np.array_equal(p4.T, tc.phi)
    with pytest.raises(Unsupported):
        import cutcell._lib as L
        L.check(L.lib().cc_write_cache_vtk(tc.handle, ctypes.byref(L.make_grid(gp)), str(tmp_path / "t").encode(), 0))
