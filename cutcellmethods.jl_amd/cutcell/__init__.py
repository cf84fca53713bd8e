"""cutcell -- the host side of libcutcell_hip, an MI355X (gfx950) implementation of
CutCellMethods.jl's ``update_cache!`` hot path (nodal marching squares, nodal marching cubes, PLIC).

The names mirror the reference's API (``src/CutCellMethods.jl:3-8``): method tags, ``allocate_cache``,
``update_cache`` (the reference's ``update_cache!``), ``generate_mesh`` and the readers.
"""
from ._lib import ArgumentError, DimensionMismatch, HIPError, Unsupported
from .cache import (Cache, MarchingCubesCutCell, MarchingSquaresCutCell, PLICCutCell, Surface,
                    allocate_cache, calc_volume, cell_indices, cut_cell_report, diagnostics,
                    extract_surface_plic, face_fractions, generate_mesh, is_cell_open, kinds,
                    update_cache, update_cache_from_device_phi, volume_fractions, write_cache_vtk,
                    write_npy)
from .geometry import (EmptyGeo, SDFBounded, SDFBox, SDFCircle, SDFPlane, SDFRectangle, SDFSphere,
                       SDFSphereUnion, SDFTorus, SDFUnion, random_rotation)
from .grid import CartesianGrid, CartesianIsoGrid
from .readers import (CELL_CUT, CELL_INSIDE, CELL_OUTSIDE, closure_residual, face_area_of,
                      face_centroid, full_face_area, interface_area, interface_normal,
                      interface_normal_area)
from .surrogate import Surrogate
from .tree import AdaptiveMesh

__all__ = [n for n in dir() if not n.startswith("_")]
