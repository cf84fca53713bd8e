"""The Cartesian grid a cache is allocated for (CartesianMeshes' ``CartesianGrid``, un-vendored).

A grid is an origin ``x0``, a cell count ``n`` and a cell size ``d`` per axis.  Node ``i`` (0-based)
sits at ``x0 + i*d`` per axis -- the reference's ``get_node(grid, i+1) = x0 + T(i)*d``, the lattice
expression every shared corner is built from (``src/marching_squares/marching_squares.jl:25-28,
106-111``).  A moving body is a grid whose origin moves (``src/cache.jl:36-37``); ``n`` is fixed for a
cache's lifetime (``src/cache.jl:107-112``).
"""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class CartesianGrid:
    x0: tuple
    n: tuple
    d: tuple

    def __post_init__(self):
        object.__setattr__(self, "x0", tuple(float(v) for v in self.x0))
        object.__setattr__(self, "n", tuple(int(v) for v in self.n))
        object.__setattr__(self, "d", tuple(float(v) for v in self.d))
        if not (len(self.x0) == len(self.n) == len(self.d)) or len(self.n) not in (2, 3):
            raise ValueError("a grid is 2D or 3D, with x0, n and d of one length")
        if any(v < 1 for v in self.n):
            raise ValueError(f"every axis needs at least one cell, got n = {self.n}")

    @property
    def dim(self):
        return len(self.n)

    @property
    def ncells(self):
        return int(np.prod(self.n))

    def nodegrid_size(self):
        return tuple(v + 1 for v in self.n)

    def is_isotropic(self):
        return all(v == self.d[0] for v in self.d)

    def moved(self, shift):
        """The same lattice with its origin shifted -- the body moving the other way."""
        return CartesianGrid(tuple(a + float(b) for a, b in zip(self.x0, shift)), self.n, self.d)

    def get_node(self, idx):
        """Coordinates of node ``idx`` (0-based), ``x0 + i*d`` per axis."""
        return np.array([self.x0[a] + float(idx[a]) * self.d[a] for a in range(self.dim)])


def CartesianIsoGrid(mincorner, maxcorner, sz=None, n=None):
    """An isotropic grid over ``[mincorner, maxcorner]``: from a cell size ``sz`` (``n`` rounded per
    axis) or a cell count ``n`` (the same on every axis, ``d = span/n``)."""
    lo = np.asarray(mincorner, dtype=np.float64)
    hi = np.asarray(maxcorner, dtype=np.float64)
    span = hi - lo
    if n is not None:
        nn = np.broadcast_to(np.asarray(n), lo.shape).astype(int)
        d = span[0] / nn[0]
        return CartesianGrid(tuple(lo), tuple(nn), tuple([d] * lo.size))
    nn = np.rint(span / float(sz)).astype(int)
    return CartesianGrid(tuple(lo), tuple(nn), tuple([float(sz)] * lo.size))
