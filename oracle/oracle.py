"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.

The oracle (``cutcell_oracle.c``) is a plain-C restatement of CutCellMethods.jl's Float64
``update_cache!`` for MarchingSquaresCutCell, MarchingCubesCutCell and PLICCutCell.  Only the tests,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this module, and only as the
checker.  It never stands in for the HIP path.

Arrays follow the reference's storage: every per-cell field is its own column, column-major over the
grid (i fastest), vector fields interleaved per cell (``src/cut_cell.jl:112-120``,
``src/cache.jl:101-104``).  Columns come back flat over the cell linear index with the per-cell
components as trailing axes.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CUTCELL_ORACLE_LIB selects another build of the same sources (the sanitizer build of
# run_sanitized.sh); the default is the optimised liborc.so beside this file.
LIB_PATH = os.environ.get("CUTCELL_ORACLE_LIB", os.path.join(HERE, "liborc.so"))

FIELDS = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
          "face_centroid_local", "interface_centroid")

GEO_EMPTY, GEO_SPHERE, GEO_TORUS, GEO_BOX, GEO_PLANE, GEO_SPHERE_UNION, GEO_UNION = range(7)


class OrcCols(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_void_p), ("ambiguous", ctypes.c_void_p),
                ("volume_fraction", ctypes.c_void_p), ("centroid", ctypes.c_void_p),
                ("face_fraction", ctypes.c_void_p), ("face_centroid_local", ctypes.c_void_p),
                ("interface_centroid", ctypes.c_void_p)]


class OrcGeometry(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n_prims", ctypes.c_int32),
                ("p", ctypes.c_double * 24), ("prims", ctypes.c_void_p)]


_lib = None


def build():
    """Compile liborc.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p
        L.orc_sample_nodes.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.orc_ms_cells.argtypes = [vp, vp, vp, vp, vp]
        L.orc_ms_contour.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_ms_contour.restype = ctypes.c_int
        L.orc_mc_cells.argtypes = [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int64, vp]
        L.orc_mc_cell_case.argtypes = [vp]
        L.orc_mc_cell_case.restype = ctypes.c_int
        L.orc_mc_cell_triangles.argtypes = [vp, vp]
        L.orc_mc_cell_triangles.restype = ctypes.c_int
        L.orc_mc_edge_crossings.argtypes = [vp, vp, vp]
        L.orc_mc_cell.argtypes = [vp] * 10
        L.orc_ms_leaves.argtypes = [vp, vp, vp, ctypes.c_int64, vp, vp, vp, vp]
        L.orc_mc_mesh.argtypes = [vp] * 9
        L.orc_mc_mesh.restype = ctypes.c_int
        L.orc_volume_fraction.argtypes = [ctypes.c_int, vp, ctypes.c_double]
        L.orc_volume_fraction.restype = ctypes.c_double
        L.orc_plic_fit.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_plic_faces.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.orc_sdf_value.argtypes = [vp, ctypes.c_int, vp]
        L.orc_sdf_value.restype = ctypes.c_double
        L.orc_get_sdf.argtypes = [vp, ctypes.c_int, vp, vp]
        L.orc_get_sdf.restype = ctypes.c_double
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_calc_volume.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.orc_plic_surface.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_double, vp, vp, vp, vp]
        L.orc_plic_surface.restype = ctypes.c_int
        L.orc_calc_volume.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if n is not None and a.size < n:
        a = np.concatenate([a, np.zeros(n - a.size)])
    return a


def _i64(x, n=3):
    a = list(int(v) for v in x) + [1] * (n - len(x))
    return np.asarray(a, dtype=np.int64)


def make_geometry(geo):
    """An OrcGeometry from any object exposing ``kind``, ``params`` (<= 24 floats) and ``prims``
    (None, an (m, 4) array of sphere centres and radii, or an (m, 16) array of union children).  The prims array is kept alive on the
    returned struct."""
    g = OrcGeometry()
    g.kind = int(geo.kind)
    params = list(float(v) for v in geo.params)
    for i, v in enumerate(params[:24]):
        g.p[i] = v
    prims = getattr(geo, "prims", None)
    if prims is not None:
        arr = np.asarray(prims, dtype=np.float64)
        arr = np.ascontiguousarray(arr if arr.ndim == 2 else arr.reshape(-1, 4))
        g.n_prims = arr.shape[0]
        g.prims = arr.ctypes.data
        g._keep = arr
    else:
        g.n_prims = 0
        g.prims = None
    return g


def alloc_cols(ncell, dim, nan=False):
    fill = np.nan if nan else 0.0
    return {
        "kind": np.zeros(ncell, np.int8),
        "ambiguous": np.zeros(ncell, np.uint8),
        "volume_fraction": np.full(ncell, fill),
        "centroid": np.full((ncell, dim), fill),
        "face_fraction": np.full((ncell, 2 * dim), fill),
        "face_centroid_local": np.full((ncell, 2 * dim, dim - 1), fill),
        "interface_centroid": np.full((ncell, dim), fill),
    }


def _cols_struct(cols):
    s = OrcCols()
    for f in FIELDS:
        setattr(s, f, cols[f].ctypes.data)
    return s


def sample_nodes(geo, dim, x0, d, n, out=None):
    """``phi`` at every node (src/marching_squares/cache.jl:67-70), flat, column-major (into ``out``
    when given)."""
    nn = [int(v) + 1 for v in n[:dim]]
    phi = np.empty(int(np.prod(nn)), np.float64) if out is None else out
    g = make_geometry(geo)
    lib().orc_sample_nodes(ctypes.byref(g), dim, _p(_f64(x0, 3)), _p(_f64(d, 3)),
                           _p(_i64(n)), _p(phi))
    return phi


def sdf_value(geo, dim, x):
    g = make_geometry(geo)
    return lib().orc_sdf_value(ctypes.byref(g), dim, _p(_f64(x, 3)))


def get_sdf(geo, dim, x):
    g = make_geometry(geo)
    nrm = np.zeros(3)
    dist = lib().orc_get_sdf(ctypes.byref(g), dim, _p(_f64(x, 3)), _p(nrm))
    return dist, nrm[:dim]


def ms_cells(phi, x0, d, n):
    """Every cell of a 2D grid (src/marching_squares/cache.jl:75-78)."""
    ncell = int(n[0]) * int(n[1])
    cols = alloc_cols(ncell, 2)
    s = _cols_struct(cols)
    phi = _f64(phi)
    lib().orc_ms_cells(_p(phi), _p(_f64(x0, 3)), _p(_f64(d, 3)), _p(_i64(n)), ctypes.byref(s))
    return cols


def ms_contour(phi, x0, d, n):
    """generate_mesh over a 2D grid (src/marching_squares/surface.jl:59-86): (nodes (m, 2), lines
    (k, 2) 1-based)."""
    phi = _f64(phi)
    args = (_p(phi), _p(_f64(x0, 3)), _p(_f64(d, 3)), _p(_i64(n)))
    nv, nl = ctypes.c_int64(0), ctypes.c_int64(0)
    if lib().orc_ms_contour(*args, ctypes.byref(nv), ctypes.byref(nl), None, None) != 0:
        raise MemoryError("orc_ms_contour")
    nodes = np.zeros((nv.value, 2))
    lines = np.zeros((nl.value, 2), np.int64)
    if lib().orc_ms_contour(*args, ctypes.byref(nv), ctypes.byref(nl), _p(nodes), _p(lines)) != 0:
        raise MemoryError("orc_ms_contour")
    return nodes, lines


def ms_leaves(geo, x0, d, level, coord):
    """A marching-squares cache over quadtree leaves (src/marching_squares/cache.jl:129-137):
    (phi (n, 4) corner values, columns with one entry per leaf)."""
    level = np.ascontiguousarray(np.asarray(level, dtype=np.int32))
    coord = np.ascontiguousarray(np.asarray(coord, dtype=np.int32).reshape(-1, 2))
    n = level.size
    cols = alloc_cols(n, 2)
    s = _cols_struct(cols)
    phi4 = np.zeros((n, 4))
    g = make_geometry(geo)
    lib().orc_ms_leaves(ctypes.byref(g), _p(_f64(x0, 2)), _p(_f64(d, 2)), n, _p(level), _p(coord),
                        _p(phi4), ctypes.byref(s))
    return phi4, cols


def mc_cells(phi, x0, d, n, k0=0, k1=None, cols=None):
    """Cells with k in [k0, k1) of a 3D grid (src/marching_squares/cache.jl:75-78); ``cols`` (from
    alloc_cols) is written in place when given."""
    if k1 is None:
        k1 = int(n[2])
    ncell = int(n[0]) * int(n[1]) * (k1 - k0)
    if cols is None:
        cols = alloc_cols(ncell, 3)
    s = _cols_struct(cols)
    phi = _f64(phi)
    lib().orc_mc_cells(_p(phi), _p(_f64(x0, 3)), _p(_f64(d, 3)), _p(_i64(n)), k0, k1,
                       ctypes.byref(s))
    return cols


def mc_cell(nodes, phi_raw, cellsize):
    """One cell's CutCellData (src/marching_cubes/moments.jl:275-346) as a dict."""
    nodes = _f64(nodes).reshape(8, 3)
    phi_raw = _f64(phi_raw)
    kind = np.zeros(1, np.int8)
    amb = np.zeros(1, np.uint8)
    vf = np.zeros(1)
    cen, ff, fcl, ic = np.zeros(3), np.zeros(6), np.zeros(12), np.zeros(3)
    lib().orc_mc_cell(_p(nodes), _p(phi_raw), _p(_f64(cellsize)), _p(kind), _p(amb), _p(vf),
                      _p(cen), _p(ff), _p(fcl), _p(ic))
    return {"kind": int(kind[0]), "ambiguous": bool(amb[0]), "volume_fraction": float(vf[0]),
            "centroid": cen, "face_fraction": ff, "face_centroid_local": fcl.reshape(6, 2),
            "interface_centroid": ic}


def mc_cell_case(phi_raw):
    return lib().orc_mc_cell_case(_p(_f64(phi_raw)))


def mc_cell_triangles(phi_raw):
    codes = np.zeros(36, np.int8)
    n = lib().orc_mc_cell_triangles(_p(_f64(phi_raw)), _p(codes))
    return n, codes[:3 * max(n, 0)].reshape(-1, 3)


def mc_edge_crossings(local_nodes, phi_raw):
    pts = np.zeros(39)
    lib().orc_mc_edge_crossings(_p(_f64(local_nodes)), _p(_f64(phi_raw)), _p(pts))
    return pts.reshape(13, 3)


def mc_mesh(phi, x0, d, n):
    """The whole-grid surface (src/marching_cubes/surface.jl:82-112): (nodes (nv, 3), triangles
    (nt, 3) 1-based, owner (nt,) 0-based cell of each triangle)."""
    phi = _f64(phi)
    args = (_p(phi), _p(_f64(x0, 3)), _p(_f64(d, 3)), _p(_i64(n)))
    nv, nt = ctypes.c_int64(0), ctypes.c_int64(0)
    if lib().orc_mc_mesh(*args, ctypes.byref(nv), ctypes.byref(nt), None, None, None) != 0:
        raise MemoryError("orc_mc_mesh")
    nodes = np.zeros((nv.value, 3))
    tris = np.zeros((nt.value, 3), np.int64)
    owner = np.zeros(nt.value, np.int64)
    if lib().orc_mc_mesh(*args, ctypes.byref(nv), ctypes.byref(nt), _p(nodes), _p(tris),
                         _p(owner)) != 0:
        raise MemoryError("orc_mc_mesh")
    return nodes, tris, owner


def volume_fraction(normal, a):
    nv = _f64(normal, 3)
    return lib().orc_volume_fraction(len(normal), _p(nv), float(a))


def plic(geo, dim, x0, d, n, out=None):
    """A PLIC cache update (src/plic/cache.jl:77-92): returns (cols, normals, intercepts,
    is_valid); the centroid columns are NaN and ``ambiguous`` is false as allocate_cache seeds
    them (src/plic/cache.jl:60-75).  ``out``, a tuple of that form, is written in place."""
    ncell = int(np.prod([int(v) for v in n[:dim]]))
    if out is None:
        cols = alloc_cols(ncell, dim, nan=True)
        normals = np.zeros((ncell, dim))
        intercepts = np.zeros(ncell)
        valid = np.zeros(ncell, np.uint8)
    else:
        cols, normals, intercepts, valid = out
    g = make_geometry(geo)
    x0a, da, na = _f64(x0, 3), _f64(d, 3), _i64(n)
    lib().orc_plic_fit(ctypes.byref(g), dim, _p(x0a), _p(da), _p(na), _p(cols["volume_fraction"]),
                       _p(cols["kind"]), _p(normals), _p(intercepts), _p(valid))
    lib().orc_plic_faces(dim, _p(da), _p(na), _p(normals), _p(intercepts), _p(valid),
                         _p(cols["volume_fraction"]), _p(cols["face_fraction"]))
    return cols, normals, intercepts, valid


def num_threads():
    return lib().orc_num_threads()


def calc_volume(geo, dim, x0, d, n):
    """calc_volume (src/volume.jl:29-40), summed in cell order."""
    g = make_geometry(geo)
    return lib().orc_calc_volume(ctypes.byref(g), dim, _p(_f64(x0, 3)), _p(_f64(d, 3)), _p(_i64(n)))


def plic_surface(dim, x0, d, n, normals, intercepts, min_frac=1e-4):
    """The PLIC surface shards (src/plic/surface.jl): (nodes (m, dim), elements (k, dim) 1-based)."""
    normals = _f64(normals)
    intercepts = _f64(intercepts)
    args = (dim, _p(_f64(x0, 3)), _p(_f64(d, 3)), _p(_i64(n)), _p(normals), _p(intercepts), float(min_frac))
    nv, ne = ctypes.c_int64(0), ctypes.c_int64(0)
    lib().orc_plic_surface(*args, ctypes.byref(nv), ctypes.byref(ne), None, None)
    nodes = np.zeros((nv.value, dim))
    elems = np.zeros((ne.value, dim), np.int64)
    lib().orc_plic_surface(*args, ctypes.byref(nv), ctypes.byref(ne), _p(nodes), _p(elems))
    return nodes, elems
