"""The reference's cache API over libcutcell_hip -- ``allocate_cache`` / ``update_cache`` /
``generate_mesh`` and the uniform readers (``src/cache.jl:15-121``).

* ``allocate_cache(grid, method, dtype=np.float64)`` allocates every column on the device and seeds
  the cache as "no body" (``src/marching_squares/cache.jl:41-47``, ``src/marching_cubes/cache.jl:30-36``,
  ``src/plic/cache.jl:60-75``).
* ``update_cache(cache, geo, grid)`` is ``update_cache!``: same ``n``, any origin; blocks until the
  device has finished (``src/cache.jl:31-45``).  ``geo`` is a ``cutcell.geometry`` descriptor, or a
  nodal array / callable for the nodal methods (the reference's "any callable ``x -> phi``").
* ``cache.cells`` reads columns back as numpy arrays (``Adapt.adapt(Array, cache.cells)``): flat
  over the cell linear index (i fastest), with the per-cell components as trailing axes.
* ``generate_mesh(cache, grid)`` is the marching-squares contour (``src/marching_squares/surface.jl``)
  or the marching-cubes surface, marched on the device (``src/marching_cubes/surface.jl``);
  ``cell_indices(mesh, grid)`` names each triangle's cell.

Errors keep the reference's types: ``DimensionMismatch`` for a wrong ``n``, ``ArgumentError`` for an
anisotropic PLIC grid.
"""
import ctypes
import warnings

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, Unsupported, check, lib, make_geometry, make_grid
from .grid import CartesianGrid
from .tree import AdaptiveMesh


class MarchingSquaresCutCell:
    """src/marching_squares/marching_squares.jl:57"""
    code = _lib.CC_MARCHING_SQUARES


class MarchingCubesCutCell:
    """src/marching_cubes/marching_cubes.jl:54"""
    code = _lib.CC_MARCHING_CUBES


class PLICCutCell:
    """src/plic/plic.jl:30"""
    code = _lib.CC_PLIC


def _method_code(method):
    if isinstance(method, type):
        return method.code
    return method.code


class Cells:
    """Host views of a cache's CutCellData columns, read on access."""

    def __init__(self, cache):
        self._c = cache

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        if name in ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
                    "face_centroid_local", "interface_centroid"):
            return self._c.read(name)
        raise AttributeError(name)

    def __len__(self):
        return self._c.grid.ncells


class Cache:
    """A device cache of one method over one grid's ``n`` (MarchingSquaresCutCellCache,
    MarchingCubesCutCellCache or PLICCutCellCache)."""

    def __init__(self, handle, grid, method, dtype):
        self._h = ctypes.c_void_p(handle)
        self.grid = grid
        self.method = method
        self.dtype = np.dtype(dtype)
        self.cells = Cells(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().cc_cache_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- columns -----------------------------------------------------------------------------
    @property
    def on_leaves(self):
        """Whether the cache is laid out on a quadtree's leaves (allocated on an AdaptiveMesh)."""
        return isinstance(self.grid, AdaptiveMesh)

    def _shape(self, name):
        D = self.grid.dim
        N = self.grid.ncells
        if name == "phi":
            return (N, 4) if self.on_leaves else (int(np.prod(self.grid.nodegrid_size())),)
        return {
            "kind": (N,), "ambiguous": (N,), "volume_fraction": (N,), "centroid": (N, D),
            "face_fraction": (N, 2 * D), "face_centroid_local": (N, 2 * D, D - 1),
            "interface_centroid": (N, D), "normals": (N, D), "intercepts": (N,), "is_valid": (N,),
        }[name]

    def _dtype(self, name):
        if name == "kind":
            return np.dtype(np.int8)
        if name in ("ambiguous", "is_valid"):
            return np.dtype(np.uint8)
        return self.dtype

    def column_ptr(self, name):
        """(device pointer, byte size) of one column: the cache's own storage."""
        p = ctypes.c_void_p()
        nb = ctypes.c_int64()
        check(lib().cc_column(self._h, _lib.FIELDS[name], ctypes.byref(p), ctypes.byref(nb)))
        return p.value, nb.value

    def probe_fill(self, name, iters=5):
        """GB/s of a plain 16-byte-per-lane fill of column ``name`` -- the write rate this
        allocation reaches.  Overwrites the column: update the cache before reading it again."""
        ms = ctypes.c_float()
        check(lib().cc_probe_fill_column(self._h, _lib.FIELDS[name], int(iters), ctypes.byref(ms)))
        nb = self.column_ptr(name)[1] // 16 * 16
        return nb / (ms.value * 1e-3) / 1e9

    def read(self, name, start=0, stop=None):
        """Column ``name`` copied to the host, optionally only leading-axis rows [start, stop)."""
        shape = self._shape(name)
        dt = self._dtype(name)
        stop = shape[0] if stop is None else stop
        per = int(np.prod(shape[1:])) * dt.itemsize
        out = np.empty((stop - start,) + shape[1:], dtype=dt)
        check(lib().cc_copy_to_host(self._h, _lib.FIELDS[name], start * per, (stop - start) * per,
                                    out.ctypes.data))
        return out

    @property
    def phi(self):
        return self.read("phi")

    @property
    def normals(self):
        return self.read("normals")

    @property
    def intercepts(self):
        return self.read("intercepts")

    @property
    def is_valid(self):
        return self.read("is_valid")

    def cut_count(self):
        v = ctypes.c_int64()
        check(lib().cc_cut_count(self._h, ctypes.byref(v)))
        return v.value

    def stream(self):
        s = ctypes.c_void_p()
        check(lib().cc_stream(self._h, ctypes.byref(s)))
        return s.value

    def set_tuning(self, variant=2, kz=4, nontemporal=False):
        """Launch shape of the marching-cubes stream pass (0 flat, 1 z-march, 2 fused z-march), or
        of a 3D PLIC update (0 two passes, 1 fused sweep on 32 x 8 tiles, 2 on 64 x 8 tiles)."""
        check(lib().cc_set_tuning(self._h, int(variant), int(kz), 1 if nontemporal else 0))

    def get_tuning(self):
        v, k, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().cc_get_tuning(self._h, ctypes.byref(v), ctypes.byref(k), ctypes.byref(nt)))
        return v.value, k.value, bool(nt.value)

    PHI_STAT_NAMES = ("nodes", "exact", "cert_evals", "kept_sum", "uncertified", "band",
                      "sign_fail", "bricks")

    def set_phi_source(self, source, surrogate=None, coarse=8, band_cells=3.0):
        """How a sphere union's nodal phi is produced: ``"naive"`` (every node folds every
        primitive), ``"culled"`` (each brick folds only the primitives it cannot exclude; bitwise
        the same) or ``"surrogate"`` (the network on a lattice of spacing ``coarse`` cells, exact
        wherever the sign is not certified, under ``band_cells`` cells, or contradicted)."""
        code = {"naive": _lib.CC_PHI_NAIVE, "culled": _lib.CC_PHI_CULLED,
                "surrogate": _lib.CC_PHI_SURROGATE}[source]
        net = surrogate.handle if surrogate is not None else None
        check(lib().cc_set_phi_source(self._h, code, net, int(coarse), float(band_cells)))
        self._surrogate = surrogate  # keep the network alive while the cache uses it

    def phi_stats(self):
        """Counters of the last culled/surrogate update."""
        a = (ctypes.c_int64 * _lib.CC_PHI_STATS)()
        check(lib().cc_phi_stats(self._h, a))
        return dict(zip(self.PHI_STAT_NAMES, list(a)))

    def set_contour_weld(self, mode="auto"):
        """Where generate_mesh welds a marching-squares contour: ``"auto"`` (the device from 4096
        cut cells), ``"host"`` or ``"device"``; the nodes and lines are the same in every mode."""
        code = {"auto": 0, "host": 1, "device": 2}[mode]
        check(lib().cc_set_contour_weld(self._h, code))

    def set_profiling(self, on=True):
        check(lib().cc_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self):
        """[(kernel name, ms)] of the last update (profiling must be on)."""
        ms = (ctypes.c_float * 8)()
        names = (ctypes.c_char_p * 8)()
        n = ctypes.c_int()
        check(lib().cc_kernel_times(self._h, 8, ms, names, ctypes.byref(n)))
        return [(names[i].decode(), float(ms[i])) for i in range(n.value)]


def allocate_cache(grid, method, dtype=np.float64):
    """allocate_cache(grid, method; backend) -- seeded as no body.  On an ``AdaptiveMesh`` (marching
    squares only, src/marching_squares/cache.jl:100-106) the cache is laid out on the tree's leaves."""
    dt = np.dtype(dtype)
    if dt not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ArgumentError("dtype must be float64 or float32")
    h = ctypes.c_void_p()
    if isinstance(grid, AdaptiveMesh):
        if _method_code(method) != _lib.CC_MARCHING_SQUARES:
            raise Unsupported("a quadtree cache is a marching-squares cache")
        t = grid.leaves()
        check(lib().cc_leaf_cache_create(_lib.CC_FLOAT32 if dt == np.float32 else _lib.CC_FLOAT64,
                                         ctypes.byref(t), ctypes.byref(h)))
        return Cache(h.value, grid, method, dt)
    g = make_grid(grid)
    check(lib().cc_cache_create(_method_code(method),
                                _lib.CC_FLOAT32 if dt == np.float32 else _lib.CC_FLOAT64,
                                ctypes.byref(g), ctypes.byref(h)))
    return Cache(h.value, grid, method, dt)


def _update_leaves(cache, geo, mesh):
    """update_cache!(cache, geo, mesh::AdaptiveMesh{2}) (src/marching_squares/cache.jl:115-137)."""
    t = mesh.leaves()
    if hasattr(geo, "kind") and hasattr(geo, "params"):
        gg = make_geometry(geo)
        check(lib().cc_update_leaves(cache.handle, ctypes.byref(gg), ctypes.byref(t)))
    else:
        if callable(geo):  # cell_values(field, mesh, c) with a callable field: each leaf's corners
            X = mesh.cell_nodes().reshape(-1, 2)
            vals = np.array([float(geo(x)) for x in X])
        else:
            vals = np.asarray(geo)
        vals = np.ascontiguousarray(vals, dtype=cache.dtype).reshape(-1)
        if vals.size != 4 * mesh.nleaves:
            raise DimensionMismatch(f"`vals` has {vals.size} values, but `mesh` has {mesh.nleaves} "
                                    "leaves of four corners each")
        check(lib().cc_update_leaves_from_phi(cache.handle, vals.ctypes.data, 0, ctypes.byref(t)))
    cache.grid = mesh
    return cache


def _check_layout(cache, domain):
    if isinstance(domain, AdaptiveMesh) != cache.on_leaves:
        raise Unsupported("a tree cache is updated and contoured over its tree, a grid cache over "
                          "its grid")


def _check_same_n(cache, grid):
    _check_layout(cache, grid)
    if tuple(grid.n) != tuple(cache.grid.n):
        raise DimensionMismatch(
            f"the cache holds {cache.grid.n} cells, but `grid` has {grid.n}; a cache is allocated "
            "for one grid's `n`, and only its origin may move between updates")


def update_cache(cache, geo, grid, sync=True):
    """update_cache!(cache, geo, grid) -> cache.

    ``geo`` is a geometry descriptor, a nodal array (one value per node, column-major or shaped as
    the node grid), or a callable ``f(x) -> phi`` sampled at every node on the host.  ``grid`` may be
    an ``AdaptiveMesh`` for a cache allocated on one (then an array holds four corner values per leaf).
    """
    if isinstance(grid, AdaptiveMesh):
        _check_layout(cache, grid)
        return _update_leaves(cache, geo, grid)
    _check_same_n(cache, grid)
    g = make_grid(grid)
    if hasattr(geo, "kind") and hasattr(geo, "params"):
        gg = make_geometry(geo)
        fn = lib().cc_update if sync else lib().cc_update_async
        check(fn(cache.handle, ctypes.byref(gg), ctypes.byref(g)))
        cache.grid = grid
        return cache
    nn = grid.nodegrid_size()
    if callable(geo):
        from itertools import product
        pts = product(*[range(m) for m in reversed(nn)])
        vals = np.array([float(geo(grid.get_node(tuple(reversed(ix))))) for ix in pts])
    else:
        vals = np.asarray(geo)
        if vals.ndim > 1:
            if vals.shape != tuple(nn):
                raise DimensionMismatch(
                    f"`vals` is {vals.shape}, but `grid` has {nn} nodes ({grid.n} elements per "
                    "axis); the nodal construction reads one value per grid node")
            vals = vals.reshape(-1, order="F")
    vals = np.ascontiguousarray(vals, dtype=cache.dtype)
    if vals.size != int(np.prod(nn)):
        raise DimensionMismatch(f"`vals` has {vals.size} values, but `grid` has {nn} nodes")
    check(lib().cc_update_from_phi(cache.handle, vals.ctypes.data, 0, ctypes.byref(g)))
    cache.grid = grid
    return cache


def update_cache_from_device_phi(cache, phi_ptr, grid):
    """update_cache! from a field already in device memory (a HIP pointer to one value per node)."""
    _check_same_n(cache, grid)
    g = make_grid(grid)
    check(lib().cc_update_from_phi(cache.handle, ctypes.c_void_p(phi_ptr), 1, ctypes.byref(g)))
    cache.grid = grid
    return cache


class Surface(tuple):
    """A marching-cubes surface, ``(nodes (m, 3), triangles (k, 3) 1-based)`` -- the reference's
    ``Mesh{3}`` of ``Tri`` -- that remembers which march of which cache made it, so that
    ``cell_indices`` can run on the cache's device copy."""

    def __new__(cls, nodes, triangles, cache, serial):
        obj = super().__new__(cls, (nodes, triangles))
        obj.cache = cache
        obj.serial = serial
        return obj

    @property
    def nodes(self):
        return self[0]

    @property
    def elements(self):
        return self[1]


def _mesh_serial(cache):
    serial = ctypes.c_int64()
    check(lib().cc_mesh_device(cache.handle, None, None, None, None, ctypes.byref(serial)))
    return serial.value


def generate_mesh(cache, grid, warn=True, min_tri_area_fraction=1e-4):
    """generate_mesh(cache, grid; warn=true).

    Marching squares (``src/marching_squares/surface.jl:59-86``): the contour of the cache's phi as
    ``(nodes (m, 2), lines (k, 2) 1-based)``.  Marching cubes (``src/marching_cubes/surface.jl:82-112``):
    the cache's phi marched on the device into a ``Surface`` of shared vertices and outward-wound
    triangles, in the cache's element type.  PLIC (``src/plic/surface.jl:28-40``): every cell's stored
    plane clipped to its cell on the device, ``(nodes (m, D), elements (k, D) 1-based)``, segments or
    triangles smaller than ``min_tri_area_fraction`` of a cell edge / face dropped.  An empty result
    warns unless ``warn`` is false."""
    nn = ctypes.c_int64()
    nl = ctypes.c_int64()
    if isinstance(grid, AdaptiveMesh):  # src/marching_squares/surface.jl:92-110
        _check_layout(cache, grid)
        t = grid.leaves()
        check(lib().cc_generate_mesh_leaves(cache.handle, ctypes.byref(t), None, ctypes.byref(nn), None,
                                            ctypes.byref(nl)))
        nodes = np.empty((nn.value, 2))
        lines = np.empty((nl.value, 2), dtype=np.int64)
        check(lib().cc_generate_mesh_leaves(cache.handle, ctypes.byref(t), nodes.ctypes.data,
                                            ctypes.byref(nn), lines.ctypes.data, ctypes.byref(nl)))
        if warn and nl.value == 0:
            warnings.warn("generate_mesh: the field has no sign change on any leaf of `mesh`, so the "
                          "contour is empty (the tree misses the body, or lies wholly inside it)")
        return nodes, lines
    _check_same_n(cache, grid)
    g = make_grid(grid)
    if _method_code(cache.method) == _lib.CC_PLIC:
        D = grid.dim
        frac = float(min_tri_area_fraction)
        check(lib().cc_generate_mesh_plic(cache.handle, ctypes.byref(g), frac, None, ctypes.byref(nn), None,
                                          ctypes.byref(nl)))
        nodes = np.empty((nn.value, D), dtype=cache.dtype)
        elems = np.empty((nl.value, D), dtype=np.int64)
        check(lib().cc_generate_mesh_plic(cache.handle, ctypes.byref(g), frac, nodes.ctypes.data,
                                          ctypes.byref(nn), elems.ctypes.data, ctypes.byref(nl)))
        if warn and nl.value == 0:
            warnings.warn("generate_mesh: no cut cell anywhere in `grid`, so the surface is empty (the grid "
                          "misses the body, or lies wholly inside it)")
        return nodes, elems
    if _method_code(cache.method) == _lib.CC_MARCHING_CUBES:
        check(lib().cc_generate_mesh_mc(cache.handle, ctypes.byref(g), None, ctypes.byref(nn), None,
                                        ctypes.byref(nl)))
        nodes = np.empty((nn.value, 3), dtype=cache.dtype)
        tris = np.empty((nl.value, 3), dtype=np.int64)
        check(lib().cc_generate_mesh_mc(cache.handle, ctypes.byref(g), nodes.ctypes.data,
                                        ctypes.byref(nn), tris.ctypes.data, ctypes.byref(nl)))
        if warn and nl.value == 0:
            warnings.warn("generate_mesh: the field has no sign change anywhere in `grid`, so the "
                          "surface is empty (the grid misses the body, or lies wholly inside it)")
        return Surface(nodes, tris, cache, _mesh_serial(cache))
    check(lib().cc_generate_mesh_ms(cache.handle, ctypes.byref(g), None, ctypes.byref(nn), None,
                                    ctypes.byref(nl)))
    nodes = np.empty((nn.value, 2))
    lines = np.empty((nl.value, 2), dtype=np.int64)
    check(lib().cc_generate_mesh_ms(cache.handle, ctypes.byref(g), nodes.ctypes.data,
                                    ctypes.byref(nn), lines.ctypes.data, ctypes.byref(nl)))
    if warn and nl.value == 0:
        warnings.warn("generate_mesh: the field has no sign change anywhere in `grid`, so the "
                      "contour is empty (the grid misses the body, or lies wholly inside it)")
    return nodes, lines


def extract_surface_plic(normals, intercepts, cell_indices, grid, min_tri_area_fraction=1e-4):
    """extract_surface_plic(normals, intercepts, cell_indices, grid) (``src/plic/surface.jl:44-182``):
    the PLIC surface of planes the caller holds -- entry q is the plane (normals[q], intercepts[q]) in
    the unit-cell frame of cell ``cell_indices[q]`` (0-based linear indices, or (i, j[, k]) rows) --
    clipped on the device, in the given order.  The element type follows ``normals``."""
    normals = np.asarray(normals)
    dt = np.dtype(np.float32) if normals.dtype == np.float32 else np.dtype(np.float64)
    D = grid.dim
    nv = np.ascontiguousarray(normals, dtype=dt).reshape(-1, D)
    ic = np.ascontiguousarray(intercepts, dtype=dt).reshape(-1)
    cells = np.asarray(cell_indices, dtype=np.int64)
    if cells.ndim == 2:  # (i, j[, k]) rows
        lin = cells[:, 0].copy()
        stride = 1
        for a in range(1, D):
            stride *= int(grid.n[a - 1])
            lin += stride * cells[:, a]
        cells = lin
    cells = np.ascontiguousarray(cells.reshape(-1))
    if not (len(cells) == len(ic) == len(nv)):
        raise DimensionMismatch("one normal, one intercept and one cell per entry")
    g = make_grid(grid)
    code = _lib.CC_FLOAT32 if dt == np.float32 else _lib.CC_FLOAT64
    nn, ne = ctypes.c_int64(), ctypes.c_int64()
    args = (code, ctypes.byref(g), len(cells), cells.ctypes.data, nv.ctypes.data, ic.ctypes.data,
            float(min_tri_area_fraction))
    check(lib().cc_extract_surface_plic(*args, None, ctypes.byref(nn), None, ctypes.byref(ne)))
    nodes = np.empty((nn.value, D), dtype=dt)
    elems = np.empty((ne.value, D), dtype=np.int64)
    check(lib().cc_extract_surface_plic(*args, nodes.ctypes.data, ctypes.byref(nn), elems.ctypes.data,
                                        ctypes.byref(ne)))
    return nodes, elems


def cell_indices(mesh, grid):
    """cell_indices(mesh, grid) (``src/marching_cubes/surface.jl:128-141``): for each triangle of a
    marching-cubes surface, the 1-based linear index of the cell of ``grid`` its centroid floors
    into, clamped to the grid.  Runs on the device copy of ``mesh``, which must be the last surface
    its cache marched."""
    cache = getattr(mesh, "cache", None)
    if cache is None:
        raise ArgumentError("cell_indices takes a surface returned by generate_mesh(cache, grid)")
    if _mesh_serial(cache) != mesh.serial:
        raise ArgumentError("the cache has marched a newer surface since this one; call "
                            "generate_mesh again")
    out = np.empty(len(mesh[1]), dtype=np.int64)
    check(lib().cc_mesh_cell_indices(cache.handle, ctypes.byref(make_grid(grid)), out.ctypes.data, 0))
    return out


def calc_volume(geo, grid, dtype=np.float64):
    """calc_volume(geo, grid) (src/volume.jl:29-40): the measure where the body's per-cell planes are
    solid, over the grid's bounds, computed on the device in `dtype` (the grid's element type)."""
    dt = np.dtype(dtype)
    if dt not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ArgumentError("dtype must be float64 or float32")
    out = ctypes.c_double()
    gg = make_geometry(geo)
    check(lib().cc_calc_volume(_lib.CC_FLOAT32 if dt == np.float32 else _lib.CC_FLOAT64, ctypes.byref(gg),
                               ctypes.byref(make_grid(grid)), ctypes.byref(out)))
    return dt.type(out.value)


def _report_dict(r, dtype):
    return {"n": r.n, "cut": r.cut, "inside": r.inside, "outside": r.outside, "ambiguous": r.ambiguous,
            "min_volume_fraction": np.dtype(dtype).type(r.min_volume_fraction)}


def cut_cell_report(cache, *rest):
    """cut_cell_report(moments) (src/cut_cell.jl:302-317): counts by kind, ambiguous cells and the
    smallest nonzero cut volume fraction -- reduced on the device for a cache; given host columns
    (kind, ambiguous, volume_fraction) it is the host reader."""
    if not isinstance(cache, Cache):
        from . import readers
        return readers.cut_cell_report(cache, *rest)
    r = _lib.CCReport()
    check(lib().cc_cut_cell_report(cache.handle, ctypes.byref(r)))
    return _report_dict(r, cache.dtype)


def diagnostics(cache, grid):
    """The report plus, over every cell (on the device): cells with a nonzero closure_residual and its
    largest component; shared faces whose fractions or centroids differ bit for bit between their two
    cells; total interface_area and interface_normal_area (src/cut_cell.jl:190-226, :340-347); cells whose
    volume fraction or a face fraction is NaN or infinite (an audit with no reference counterpart)."""
    _check_same_n(cache, grid)
    dg = _lib.CCDiagnostics()
    check(lib().cc_cache_diagnostics(cache.handle, ctypes.byref(make_grid(grid)), ctypes.byref(dg)))
    out = _report_dict(dg.report, cache.dtype)
    out.update(closure_nonzero=dg.closure_nonzero, closure_max=dg.closure_max,
               face_mismatches=dg.face_mismatches, centroid_mismatches=dg.centroid_mismatches,
               interface_area=dg.interface_area,
               interface_normal_area=np.array(dg.interface_normal_area[:grid.dim]), nonfinite=dg.nonfinite)
    return out


def write_cache_vtk(prefix, cache, grid, surface=True):
    """write_cache_vtk(prefix, cache, grid; surface=true) -> files written: ``prefix_cells.vti`` (the
    cells, with phi or the planes) and, with ``surface``, ``prefix_surface.vtu`` (generate_mesh), the
    layout of the reference's write_cache_vtk (src/tri_clipping/surface.jl:214-252)."""
    _check_same_n(cache, grid)
    check(lib().cc_write_cache_vtk(cache.handle, ctypes.byref(make_grid(grid)), str(prefix).encode(),
                                   1 if surface else 0))
    return [str(prefix) + "_cells.vti"] + ([str(prefix) + "_surface.vtu"] if surface else [])


def write_npy(cache, name, path):
    """One column as a .npy file in the reference's layout (components first, then the cells, column
    major): ``numpy.load(path)`` gives e.g. face_fraction as (2D, n0, n1[, n2])."""
    check(lib().cc_write_npy(cache.handle, _lib.FIELDS[name], str(path).encode()))
    return str(path)


def volume_fractions(cache):
    return cache.read("volume_fraction")


def kinds(cache):
    return cache.read("kind")


def face_fractions(cache):
    return cache.read("face_fraction")


def is_cell_open(cache):
    """Per cell: whether any face is open (src/cache.jl:95-96), read off face_fraction alone."""
    return np.any(face_fractions(cache) != 0, axis=1)


def probe_bandwidth(nbytes, iters=5):
    """(fill GB/s, copy GB/s counting read + write) of a plain 16-byte-per-lane fill and copy of
    ``nbytes`` on the current device: the achievable-bandwidth yardstick."""
    f = ctypes.c_float()
    c = ctypes.c_float()
    check(lib().cc_probe_bandwidth(int(nbytes), int(iters), ctypes.byref(f), ctypes.byref(c)))
    nb = int(nbytes) & ~4095
    return nb / (f.value * 1e-3) / 1e9, 2 * nb / (c.value * 1e-3) / 1e9


def probe_write(nbytes, variant=0, blocks=16384, iters=5):
    """GB/s of a fill of ``nbytes`` with one of the probe's store patterns."""
    t = ctypes.c_float()
    check(lib().cc_probe_write(int(nbytes), int(iters), int(variant), int(blocks), ctypes.byref(t)))
    return (int(nbytes) & ~4095) / (t.value * 1e-3) / 1e9
