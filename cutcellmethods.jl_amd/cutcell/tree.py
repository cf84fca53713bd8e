"""A quadtree over a 2D CartesianGrid: the host side of CartesianMeshes' ``AdaptiveMesh{2}`` (the
reference's tree type, a dependency absent here), as far as the marching-squares leaf cache needs it
(``src/marching_squares/cache.jl:91-147``, ``src/marching_squares/surface.jl:92-110``).

The tree is its leaf list: leaf ``c`` has a refinement level (0 = the base grid's cells) and the
integer coordinates of its lower corner on that level's lattice, so its corners are
``x0 + (coord + bit) * d / 2^level`` (``cell_nodes(mesh, c)``,
``src/marching_squares/marching_squares.jl:113-121``).  ``d / 2^level`` is exact, which makes a corner
that a coarse and a fine leaf share one floating-point point.

Leaf order here: the base cells column-major (i fastest); refining a leaf replaces it, in place, by
its four children in Z order ((0,0), (1,0), (0,1), (1,1)).  Every entry point of the library works per
leaf, in whatever order the list has, so a tree with CartesianMeshes' own order is taken as it is.
"""
import ctypes
import itertools

import numpy as np

from ._lib import ArgumentError, CCLeaves

_versions = itertools.count(1)  # one tag per state of any tree in this process


class AdaptiveMesh:
    """AdaptiveMesh(base; max_level, initial_level): every base cell refined uniformly to
    ``initial_level``; ``refine`` / ``refine_where`` split leaves up to ``max_level``."""

    dim = 2

    def __init__(self, base, max_level=1, initial_level=0):
        if base.dim != 2:
            raise ArgumentError("a quadtree is built over a 2D grid")
        if not 0 <= initial_level <= max_level:
            raise ArgumentError("need 0 <= initial_level <= max_level")
        self.base = base
        self.max_level = int(max_level)
        n0, n1 = base.n
        jj, ii = np.meshgrid(np.arange(n1), np.arange(n0), indexing="ij")
        self.level = np.zeros(n0 * n1, dtype=np.int32)
        self.coord = np.ascontiguousarray(np.stack([ii.ravel(), jj.ravel()], axis=1).astype(np.int32))
        self.version = next(_versions)
        for _ in range(initial_level):
            self.refine(np.ones(self.nleaves, dtype=bool))

    # -- the leaf list ---------------------------------------------------------------------------
    @property
    def nleaves(self):
        return int(self.level.size)

    @property
    def ncells(self):
        return self.nleaves

    def leaf_level(self, i):
        return int(self.level[i])

    def elem_size(self, i=None):
        """Leaf size(s): d / 2^level, exact."""
        lv = self.level if i is None else self.level[i]
        d = np.asarray(self.base.d, dtype=np.float64)
        return np.stack([np.ldexp(d[0], -lv), np.ldexp(d[1], -lv)], axis=-1)

    def cell_nodes(self, i=None):
        """Corners (MS_NODE_BITS order, counter-clockwise from the lower corner): (4, 2), or (n, 4, 2)."""
        co = self.coord if i is None else self.coord[i]
        h = self.elem_size(i)
        x0 = np.asarray(self.base.x0, dtype=np.float64)
        bits = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])
        out = (co[..., None, :] + bits).astype(np.float64) * h[..., None, :] + x0
        return out

    def elem_centroid(self, i=None):
        co = self.coord if i is None else self.coord[i]
        h = self.elem_size(i)
        return np.asarray(self.base.x0, dtype=np.float64) + (co + 0.5) * h

    # -- refinement ------------------------------------------------------------------------------
    def refine(self, flags):
        """Split every flagged leaf below max_level into its four children, in place (Z order)."""
        flags = np.asarray(flags, dtype=bool) & (self.level < self.max_level)
        if not flags.any():
            return self
        reps = np.where(flags, 4, 1)
        lv = np.repeat(self.level, reps)
        co = np.repeat(self.coord, reps, axis=0)
        first = np.concatenate([[0], np.cumsum(reps)[:-1]])
        kids = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=np.int32)
        for q in range(4):
            at = first[flags] + q
            co[at] = 2 * self.coord[flags] + kids[q]
            lv[at] = self.level[flags] + 1
        self.level = np.ascontiguousarray(lv.astype(np.int32))
        self.coord = np.ascontiguousarray(co.astype(np.int32))
        self.version = next(_versions)
        return self

    def touch(self):
        """Declare ``level`` / ``coord`` edited by hand: the next update copies the leaves again."""
        self.version = next(_versions)
        return self

    def refine_where(self, pred):
        """Split the leaves i for which pred(self, i) holds."""
        return self.refine([bool(pred(self, i)) for i in range(self.nleaves)])

    # -- the ABI form ----------------------------------------------------------------------------
    def leaves(self):
        """The tree as a ``cc_leaves`` (keeps its arrays alive)."""
        t = CCLeaves()
        t.n = self.nleaves
        t.level = self.level.ctypes.data_as(ctypes.c_void_p)
        t.coord = self.coord.ctypes.data_as(ctypes.c_void_p)
        for a in range(2):
            t.x0[a] = float(self.base.x0[a])
            t.d[a] = float(self.base.d[a])
        t.on_device = 0
        t.version = self.version & 0x7FFFFFFF  # changes whenever the leaf arrays do
        t._keep = (self.level, self.coord)
        return t
