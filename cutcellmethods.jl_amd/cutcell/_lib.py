"""ctypes binding of libcutcell_hip.so (the C ABI declared in include/cutcell_hip.h).

The library is built in-tree by ``cutcellmethods.jl_amd/build.py``; nothing here falls back to a
CPU path: if the library is missing or no HIP device is present, calls raise.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libcutcell_hip.so")

CC_OK = 0
CC_ERR_DIMENSION_MISMATCH = 1
CC_ERR_ARGUMENT = 2
CC_ERR_HIP = 3
CC_ERR_UNSUPPORTED = 4
CC_ERR_OUT_OF_MEMORY = 5

CC_MARCHING_SQUARES, CC_MARCHING_CUBES, CC_PLIC = 1, 2, 3
CC_FLOAT64, CC_FLOAT32 = 0, 1
CC_PHI_NAIVE, CC_PHI_CULLED, CC_PHI_SURROGATE = 0, 1, 2
CC_PHI_STATS = 8

FIELDS = {
    "kind": 0, "ambiguous": 1, "volume_fraction": 2, "centroid": 3, "face_fraction": 4,
    "face_centroid_local": 5, "interface_centroid": 6, "phi": 7, "normals": 8, "intercepts": 9,
    "is_valid": 10,
}

EXPORTS = ("cc_cache_create", "cc_cache_destroy", "cc_update", "cc_update_from_phi",
           "cc_update_async", "cc_synchronize", "cc_column", "cc_copy_to_host", "cc_cut_count",
           "cc_stream", "cc_set_tuning", "cc_get_tuning", "cc_set_profiling", "cc_kernel_times", "cc_generate_mesh_ms",
           "cc_generate_mesh_mc", "cc_mesh_device", "cc_mesh_cell_indices",
           "cc_calc_volume", "cc_cut_cell_report", "cc_cache_diagnostics", "cc_write_cache_vtk", "cc_write_npy", "cc_generate_mesh_plic", "cc_extract_surface_plic", "cc_leaf_cache_create", "cc_update_leaves", "cc_update_leaves_from_phi", "cc_generate_mesh_leaves",
           "cc_probe_bandwidth", "cc_probe_write", "cc_probe_fill_column", "cc_set_contour_weld", "cc_last_error", "cc_version",
           "cc_mlp_create", "cc_mlp_destroy", "cc_mlp_num_params", "cc_mlp_get_params",
           "cc_mlp_set_params", "cc_mlp_eval", "cc_mlp_loss_grad", "cc_mlp_train",
           "cc_set_phi_source", "cc_phi_stats")


class CCGeometry(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n_prims", ctypes.c_int32),
                ("p", ctypes.c_double * 24), ("prims", ctypes.c_void_p)]


class CCGrid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("n", ctypes.c_int64 * 3), ("x0", ctypes.c_double * 3), ("d", ctypes.c_double * 3)]


class CCLeaves(ctypes.Structure):
    """A quadtree's leaf list (include/cutcell_hip.h, cc_leaves)."""
    _fields_ = [("n", ctypes.c_int64), ("level", ctypes.c_void_p), ("coord", ctypes.c_void_p),
                ("x0", ctypes.c_double * 2), ("d", ctypes.c_double * 2), ("on_device", ctypes.c_int32),
                ("version", ctypes.c_int32)]


class CCReport(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("cut", ctypes.c_int64), ("inside", ctypes.c_int64),
                ("outside", ctypes.c_int64), ("ambiguous", ctypes.c_int64),
                ("min_volume_fraction", ctypes.c_double)]


class CCDiagnostics(ctypes.Structure):
    _fields_ = [("report", CCReport), ("closure_nonzero", ctypes.c_int64), ("closure_max", ctypes.c_double),
                ("face_mismatches", ctypes.c_int64), ("centroid_mismatches", ctypes.c_int64),
                ("interface_area", ctypes.c_double), ("interface_normal_area", ctypes.c_double * 3),
                ("nonfinite", ctypes.c_int64)]


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch (src/cache.jl:107-121)."""


class ArgumentError(ValueError):
    """The reference's ArgumentError (src/plic/surface.jl:41-45)."""


class HIPError(RuntimeError):
    """A HIP runtime failure inside libcutcell_hip."""


class Unsupported(NotImplementedError):
    pass


_lib = None


def lib():
    """Load libcutcell_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run cutcellmethods.jl_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
        L.cc_cache_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(CCGrid),
                                      ctypes.POINTER(vp)]
        L.cc_cache_destroy.argtypes = [vp]
        L.cc_update.argtypes = [vp, ctypes.POINTER(CCGeometry), ctypes.POINTER(CCGrid)]
        L.cc_update_async.argtypes = [vp, ctypes.POINTER(CCGeometry), ctypes.POINTER(CCGrid)]
        L.cc_update_from_phi.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(CCGrid)]
        L.cc_synchronize.argtypes = [vp]
        L.cc_column.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp), i64p]
        L.cc_copy_to_host.argtypes = [vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, vp]
        L.cc_cut_count.argtypes = [vp, i64p]
        L.cc_stream.argtypes = [vp, ctypes.POINTER(vp)]
        L.cc_set_profiling.argtypes = [vp, ctypes.c_int]
        L.cc_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int)
        L.cc_get_tuning.argtypes = [vp, ip, ip, ip]
        L.cc_kernel_times.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                      ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_int)]
        L.cc_generate_mesh_ms.argtypes = [vp, ctypes.POINTER(CCGrid), vp, i64p, vp, i64p]
        L.cc_generate_mesh_mc.argtypes = [vp, ctypes.POINTER(CCGrid), vp, i64p, vp, i64p]
        L.cc_calc_volume.argtypes = [ctypes.c_int, ctypes.POINTER(CCGeometry), ctypes.POINTER(CCGrid),
                                     ctypes.POINTER(ctypes.c_double)]
        L.cc_generate_mesh_plic.argtypes = [vp, ctypes.POINTER(CCGrid), ctypes.c_double, vp, i64p, vp, i64p]
        L.cc_extract_surface_plic.argtypes = [ctypes.c_int, ctypes.POINTER(CCGrid), ctypes.c_int64, vp, vp, vp,
                                              ctypes.c_double, vp, i64p, vp, i64p]
        L.cc_cut_cell_report.argtypes = [vp, ctypes.POINTER(CCReport)]
        L.cc_write_cache_vtk.argtypes = [vp, ctypes.POINTER(CCGrid), ctypes.c_char_p, ctypes.c_int]
        L.cc_write_npy.argtypes = [vp, ctypes.c_int, ctypes.c_char_p]
        L.cc_cache_diagnostics.argtypes = [vp, ctypes.POINTER(CCGrid), ctypes.POINTER(CCDiagnostics)]
        lp = ctypes.POINTER(CCLeaves)
        L.cc_leaf_cache_create.argtypes = [ctypes.c_int, lp, ctypes.POINTER(vp)]
        L.cc_update_leaves.argtypes = [vp, ctypes.c_void_p, lp]
        L.cc_update_leaves_from_phi.argtypes = [vp, vp, ctypes.c_int, lp]
        L.cc_generate_mesh_leaves.argtypes = [vp, lp, vp, i64p, vp, i64p]
        L.cc_mesh_device.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), i64p, i64p, i64p]
        L.cc_mesh_cell_indices.argtypes = [vp, ctypes.POINTER(CCGrid), vp, ctypes.c_int]
        L.cc_probe_bandwidth.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_float)]
        L.cc_probe_write.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float)]
        L.cc_probe_fill_column.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.cc_set_contour_weld.argtypes = [vp, ctypes.c_int]
        fp = ctypes.POINTER(ctypes.c_float)
        L.cc_mlp_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_uint64, ctypes.POINTER(vp)]
        L.cc_mlp_destroy.argtypes = [vp]
        L.cc_mlp_num_params.argtypes = [vp, i64p, i64p]
        L.cc_mlp_get_params.argtypes = [vp, vp, vp]
        L.cc_mlp_set_params.argtypes = [vp, vp]
        L.cc_mlp_eval.argtypes = [vp, vp, ctypes.c_int64, vp]
        L.cc_mlp_loss_grad.argtypes = [vp, vp, vp, ctypes.c_int64, fp, vp]
        L.cc_mlp_train.argtypes = [vp, ctypes.POINTER(CCGeometry), ctypes.c_int64, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_uint64, vp]
        L.cc_set_phi_source.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double]
        L.cc_phi_stats.argtypes = [vp, i64p]
        L.cc_last_error.restype = ctypes.c_char_p
        L.cc_version.restype = ctypes.c_char_p
        for name in EXPORTS:
            if name not in ("cc_last_error", "cc_version"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc):
    if rc == CC_OK:
        return
    msg = lib().cc_last_error().decode(errors="replace")
    if rc == CC_ERR_DIMENSION_MISMATCH:
        raise DimensionMismatch(msg)
    if rc == CC_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if rc == CC_ERR_UNSUPPORTED:
        raise Unsupported(msg)
    if rc == CC_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise HIPError(msg)


def make_grid(grid):
    g = CCGrid()
    g.dim = grid.dim
    for a in range(3):
        g.n[a] = grid.n[a] if a < grid.dim else 1
        g.x0[a] = grid.x0[a] if a < grid.dim else 0.0
        g.d[a] = grid.d[a] if a < grid.dim else 1.0
    return g


def make_geometry(geo):
    import numpy as np
    g = CCGeometry()
    g.kind = int(geo.kind)
    for i, v in enumerate(list(geo.params)[:24]):
        g.p[i] = float(v)
    prims = getattr(geo, "prims", None)
    if prims is not None:
        arr = np.asarray(prims, dtype=np.float64)
        arr = np.ascontiguousarray(arr if arr.ndim == 2 else arr.reshape(-1, 4))
        g.n_prims = arr.shape[0]
        g.prims = arr.ctypes.data
        g._keep = arr
    return g
