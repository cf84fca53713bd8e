"""The allocation-free calc_volume entry point (cc_calc_volume_noalloc) through the C ABI, against
cc_calc_volume on the same inputs: bit for bit, on two calls in a row whose second has a smaller grid
and a smaller primitive table, in Float64 and Float32, 2D and 3D, on the null stream and on a
stream of the caller's; a primitive table beyond the scratch array's limit refused; and no device
memory taken by a call (hipMemGetInfo unchanged across it)."""
import ctypes
import gc

import numpy as np
import pytest

from cutcell import (CartesianIsoGrid, MarchingCubesCutCell, SDFBox, SDFCircle, SDFRectangle,
                     SDFSphere, SDFSphereUnion, SDFTorus, SDFUnion, allocate_cache, random_rotation)
from cutcell import _lib
from cutcell._lib import CCGrid, CCGeometry, lib, make_geometry, make_grid

pytestmark = pytest.mark.gpu

MAX_PRIM_DOUBLES = (256 * 1024 - (2048 * 8 + 256)) // 8


def noalloc(dtype, geo, grid, stream=None):
    L = lib()
    L.cc_calc_volume_noalloc.argtypes = [ctypes.c_int, ctypes.POINTER(CCGeometry), ctypes.POINTER(CCGrid),
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    out = ctypes.c_double(np.nan)
    g = make_geometry(geo)
    rc = L.cc_calc_volume_noalloc(dtype, ctypes.byref(g), ctypes.byref(make_grid(grid)), stream, ctypes.byref(out))
    return rc, out.value


def allocating(dtype, geo, grid):
    out = ctypes.c_double(np.nan)
    g = make_geometry(geo)
    rc = lib().cc_calc_volume(dtype, ctypes.byref(g), ctypes.byref(make_grid(grid)), ctypes.byref(out))
    assert rc == 0, lib().cc_last_error()
    return out.value


def same(a, b):
    return np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64)


def cases():
    rng = np.random.default_rng(71)
    big_c, big_r = rng.uniform(-0.6, 0.6, (300, 3)), rng.uniform(0.05, 0.15, 300)
    small_c, small_r = big_c[:40] * 0.9, big_r[:40]
    return [
        # (dim, first (geo, grid), second: smaller grid and smaller primitive table)
        (3, (SDFSphereUnion(big_c, big_r, 0.02), CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=72)),
            (SDFSphereUnion(small_c, small_r, 0.02), CartesianIsoGrid((-0.9,) * 3, (0.9,) * 3, n=33))),
        (3, (SDFUnion(SDFSphere((0.1, 0.0, 0.0), 0.4), SDFBox((-0.2, 0.1, 0.0), (0.2, 0.3, 0.25), random_rotation(rng)),
                      SDFTorus((0.0, 0.0, 0.1), 0.5, 0.12, random_rotation(rng)), smoothing=0.03),
             CartesianIsoGrid((-0.8,) * 3, (0.8,) * 3, n=64)),
            (SDFUnion(SDFSphere((0.1, 0.0, 0.0), 0.4), smoothing=0.0), CartesianIsoGrid((-0.8,) * 3, (0.8,) * 3, n=17))),
        (2, (SDFUnion(SDFCircle((0.02, 0.03), 0.31), SDFRectangle((0.25, 0.10), (0.18, 0.18))),
             CartesianIsoGrid((-0.6, -0.6), (0.6, 0.6), n=300)),
            (SDFCircle((0.0, 0.0), 0.3), CartesianIsoGrid((-0.6, -0.6), (0.6, 0.6), n=24))),
    ]


@pytest.mark.parametrize("dtype", [_lib.CC_FLOAT64, _lib.CC_FLOAT32])
def test_noalloc_matches_calc_volume_on_two_calls(dtype):
    for dim, first, second in cases():
        for geo, grid in (first, second):
            rc, v = noalloc(dtype, geo, grid)
            assert rc == 0, lib().cc_last_error()
            want = allocating(dtype, geo, grid)
            assert same(v, want), (dim, type(geo).__name__, grid.n, v, want)
            assert v > 0


def test_noalloc_on_a_callers_stream_and_takes_no_device_memory():
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    grid = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=40)
    cache = allocate_cache(grid, MarchingCubesCutCell)
    st = ctypes.c_void_p()
    assert lib().cc_stream(cache.handle, ctypes.byref(st)) == 0
    geo = SDFSphere((0.05, -0.02, 0.01), 0.6)
    rc, v0 = noalloc(_lib.CC_FLOAT64, geo, grid, st)  # the first call on this device looks up the array
    assert rc == 0, lib().cc_last_error()
    # caches of earlier tests freed now, and no collection inside the window, so that the device's
    # free memory moves only if these calls allocate
    gc.collect()
    gc.disable()
    try:
        assert hip.hipDeviceSynchronize() == 0
        free0, total = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(free0), ctypes.byref(total)) == 0
        for stream in (st, None, st):
            rc, v = noalloc(_lib.CC_FLOAT64, geo, grid, stream)
            assert rc == 0 and same(v, v0)
        free1 = ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(free1), ctypes.byref(total)) == 0
    finally:
        gc.enable()
    assert free1.value == free0.value
    assert same(v0, allocating(_lib.CC_FLOAT64, geo, grid))


def test_noalloc_refuses_a_table_beyond_its_scratch():
    n = MAX_PRIM_DOUBLES // 4 + 1
    rng = np.random.default_rng(72)
    geo = SDFSphereUnion(rng.uniform(-0.5, 0.5, (n, 3)), np.full(n, 0.05), 0.02)
    rc, _ = noalloc(_lib.CC_FLOAT64, geo, CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=8))
    assert rc == _lib.CC_ERR_ARGUMENT
    assert b"scratch" in lib().cc_last_error()
    ok = SDFSphereUnion(geo.centers[:n - 1], geo.radii[:n - 1], 0.02)  # exactly the limit
    rc, v = noalloc(_lib.CC_FLOAT64, ok, CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=8))
    assert rc == 0 and same(v, allocating(_lib.CC_FLOAT64, ok, CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=8)))
