"""Batched bodies: many 3D caches (marching cubes and PLIC mixed) updated by one hipGraph launch
per step (``cc_batch_*`` in include/cutcell_hip.h).  No counterpart in the reference, which updates
one cache per ``update_cache!`` call (``src/marching_squares/cache.jl:55-64``)."""
import ctypes

import numpy as np

from ._lib import CCGeometry, CCGrid, check, lib, make_geometry, make_grid


def _bind():
    L = lib()
    if not getattr(L, "_batch_bound", False):
        vp = ctypes.c_void_p
        L.cc_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.cc_batch_destroy.argtypes = [vp]
        L.cc_batch_step.argtypes = [vp, ctypes.POINTER(CCGeometry), ctypes.POINTER(CCGrid),
                                    ctypes.POINTER(ctypes.c_float)]
        L.cc_batch_set_kz.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.cc_batch_work_items.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_int64)]
        for n in ("cc_batch_create", "cc_batch_destroy", "cc_batch_step", "cc_batch_set_kz",
                  "cc_batch_work_items"):
            getattr(L, n).restype = ctypes.c_int
        L._batch_bound = True
    return L


class Batch:
    """B caches updated together; ``step(geos, grids)`` updates every body and blocks."""

    def __init__(self, caches):
        L = _bind()
        self.caches = list(caches)
        arr = (ctypes.c_void_p * len(self.caches))(*[c.handle.value for c in self.caches])
        h = ctypes.c_void_p()
        check(L.cc_batch_create(len(self.caches), arr, ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().cc_batch_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_kz(self, body, kz):
        check(_bind().cc_batch_set_kz(self._h, int(body), int(kz)))

    def work_items(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(_bind().cc_batch_work_items(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def step(self, geos, grids):
        """Update every body; returns the graph's device time in ms."""
        n = len(self.caches)
        G = (CCGeometry * n)()
        R = (CCGrid * n)()
        keep = []
        for i, (geo, grid) in enumerate(zip(geos, grids)):
            g = make_geometry(geo)
            keep.append(g)
            ctypes.memmove(ctypes.byref(G[i]), ctypes.byref(g), ctypes.sizeof(CCGeometry))
            R[i] = make_grid(grid)
        ms = ctypes.c_float()
        check(_bind().cc_batch_step(self._h, G, R, ctypes.byref(ms)))
        for c, grid in zip(self.caches, grids):
            c.grid = grid
        return ms.value
