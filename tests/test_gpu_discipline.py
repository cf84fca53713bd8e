"""Device-memory discipline of a steady-state update (the reference's "kernel discipline",
``test/test_moments.jl:712-723``: an update allocates only launch overhead), and the ordering of a
batched step after updates a caller left running on its member caches.

After warm-up, repeated updates of every kind of cache -- marching squares on a grid and on a
quadtree's leaves, marching cubes (analytic body, sphere union naive and culled), PLIC fused and
two-pass -- and batched steps leave the device's free memory exactly where it was: no update grows
or reallocates a buffer once its sizes are known."""
import ctypes
import gc

import numpy as np
import pytest

from cutcell import (AdaptiveMesh, CartesianGrid, CartesianIsoGrid, MarchingCubesCutCell,
                     MarchingSquaresCutCell, PLICCutCell, SDFCircle, SDFSphere, SDFSphereUnion,
                     SDFTorus, allocate_cache, random_rotation, update_cache)
from test_gpu_parity import COLS, same_bits

pytestmark = pytest.mark.gpu


def free_device_memory():
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    assert hip.hipDeviceSynchronize() == 0
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def steady_state_takes_nothing(step, warm=2, reps=5):
    for s in range(warm):
        step(s)
    gc.collect()
    gc.disable()
    try:
        before = free_device_memory()
        for s in range(reps):
            step(warm + s)
        after = free_device_memory()
    finally:
        gc.enable()
    assert after == before, f"{before - after} bytes taken by {reps} steady-state updates"


def moving(grid, step):
    return grid.moved((-0.37 * grid.d[0] * step, 0.0, 0.0)[:grid.dim])


def test_nodal_and_plic_updates_take_no_device_memory():
    rng = np.random.default_rng(81)
    g3 = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=96)
    cases = [
        (CartesianIsoGrid((-1.0, -1.0), (1.0, 1.0), sz=0.01), MarchingSquaresCutCell, SDFCircle((0.0, 0.0), 0.5), None),
        (g3, MarchingCubesCutCell, SDFSphere((0.05, -0.02, 0.01), 0.6), None),
        (g3, MarchingCubesCutCell, SDFSphereUnion(rng.uniform(-0.6, 0.6, (40, 3)), rng.uniform(0.05, 0.2, 40), 0.02), "naive"),
        (g3, MarchingCubesCutCell, SDFSphereUnion(rng.uniform(-0.6, 0.6, (40, 3)), rng.uniform(0.05, 0.2, 40), 0.02), "culled"),
        (g3, PLICCutCell, SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15, random_rotation(rng)), (2, 16)),
        (g3, PLICCutCell, SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15, random_rotation(rng)), (0, 1)),
    ]
    for grid, method, geo, how in cases:
        c = allocate_cache(grid, method)
        if isinstance(how, str):
            c.set_phi_source(how)
        elif how is not None:
            c.set_tuning(how[0], how[1], False)
        steady_state_takes_nothing(lambda s: update_cache(c, geo, moving(grid, s)))
        del c


def test_leaf_updates_take_no_device_memory():
    # test/test_moments.jl:716-720: a uniform tree of 256^2 leaves, the circle
    mesh = AdaptiveMesh(CartesianGrid((-2.0, -2.0), (8, 8), (0.5, 0.5)), max_level=5, initial_level=5)
    assert mesh.nleaves == 256 * 256
    c = allocate_cache(mesh, MarchingSquaresCutCell)
    steady_state_takes_nothing(lambda s: update_cache(c, SDFCircle((0.01 * s, 0.0), 1.0), mesh))


def test_batched_steps_take_no_device_memory_and_wait_for_async_updates():
    from cutcell.batch import Batch
    from cutcell.workloads import MovingBodies
    wl = MovingBodies(seed=9, nbodies=8, nmin=24, nmax=96)
    meth = {"mc": MarchingCubesCutCell, "plic": PLICCutCell}
    caches = [allocate_cache(wl.grid(b, 0), meth[B["method"]]) for b, B in enumerate(wl.bodies)]
    batch = Batch(caches)
    nb = len(caches)

    def step(s):
        batch.step([wl.geometry(b, s) for b in range(nb)], [wl.grid(b, s) for b in range(nb)])

    steady_state_takes_nothing(step)
    # an update left running on every member (cc_update_async, another body) is ordered before the
    # batch writes the same columns: the batch's result is exactly what it gives alone
    solo = [allocate_cache(wl.grid(b, 0), meth[B["method"]]) for b, B in enumerate(wl.bodies)]
    for s in (3, 4):
        for b in range(nb):
            update_cache(caches[b], SDFSphere((0.3, -0.2, 0.1), 0.3 + 0.01 * b), wl.grid(b, s), sync=False)
        step(s)
        for b in range(nb):
            update_cache(solo[b], wl.geometry(b, s), wl.grid(b, s))
            for name in COLS:
                assert same_bits(caches[b].read(name), solo[b].read(name)), (s, b, name)
