"""GPU parity at the benchmarked sizes, through the C ABI.

Bar: bit-exact against the oracle at full size, as at the small sizes of the other parity tests,
on the benchmark's own inputs (its default seed, 20261203):
- the 512^3 Float64 marching-cubes workload (sphere r = 0.4, centre U[-0.1, 0.1]^3 from the seed) at
  the position of its last timed step (step 35, the body 17.5 cells along +x) and under the stream
  shape the autotuner picks for it (fused z-march, one layer per work item, ordinary stores): phi
  at all 513^3 nodes, then every column of all 512^3 cells, slab by slab;
- the 384^3 Float64 PLIC workload (torus R = 0.5, r = 0.15, tilted by the rotation drawn from the
  seed): every column, plane and valid flag;
- the 512^3 composite body (the smooth union of 512 spheres drawn from the seed) at the grid
  position the benchmark compares: the culled fold's phi and every column bitwise the plain
  fold's at every node and cell, three slabs of four layers through the body bitwise the oracle's
  (the lattice is dyadic, so a slab's own origin reproduces the nodes exactly), and the surrogate
  (a network trained briefly here) with every column bitwise the plain fold's, phi bitwise at
  every corner of every cut cell and of the same sign everywhere;
- the batched workload (64 moving bodies from the seed, 64^3-256^3 cells each, marching cubes and
  PLIC mixed) at its last step after the batch autotuner: every body's columns from the one graph
  launch bitwise what cc_update gives that body alone, and the largest marching-cubes and PLIC
  bodies bitwise the oracle's.
The reference's invariants on the device output at the same sizes: shared faces bitwise
single-valued, closure exactly zero (src/cut_cell.jl:302-317; test/test_cache.jl:187-190).
"""
import gc

import numpy as np
import pytest

import oracle as orc
from cutcell import (CartesianIsoGrid, MarchingCubesCutCell, PLICCutCell, SDFSphere, SDFTorus,
                     Surrogate, allocate_cache, closure_residual, random_rotation, update_cache)
from cutcell.workloads import csg_spheres
from helpers import shared_face_mismatches

pytestmark = pytest.mark.gpu

SEED = 20261203  # the benchmark's default seed

COLS = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
        "face_centroid_local", "interface_centroid")


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and bool(np.array_equal(a.view(np.uint8), b.view(np.uint8)))


def test_marching_cubes_512_sphere_full_size_bitwise():
    rng = np.random.default_rng(SEED)
    centre = tuple(rng.uniform(-0.1, 0.1, 3))
    geo = SDFSphere(centre, 0.4)
    base = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=512)
    g = base.moved((-0.5 * base.d[0] * 35, 0.0, 0.0))
    c = allocate_cache(g, MarchingCubesCutCell)
    c.set_tuning(2, 1, False)
    update_cache(c, geo, g)
    print(f"512^3 sphere at {centre}, x0 {g.x0}", flush=True)
    phi = orc.sample_nodes(geo, 3, g.x0, g.d, g.n)
    assert same_bits(c.phi, phi)
    plane, slab, cut = 512 * 512, 32, 0
    for k0 in range(0, 512, slab):
        ref = orc.mc_cells(phi, g.x0, g.d, g.n, k0, k0 + slab)
        for name in COLS:
            assert same_bits(c.read(name, k0 * plane, (k0 + slab) * plane), ref[name]), (k0, name)
        cut += int((ref["kind"] == 2).sum())
        del ref
    assert cut == c.cut_count() and cut > 150_000
    del phi
    ff = c.cells.face_fraction
    assert shared_face_mismatches(ff, g.n, 3) == 0
    assert np.count_nonzero(closure_residual(ff, g.d)) == 0
    del ff, c
    gc.collect()


def test_plic_384_torus_full_size_bitwise():
    rng = np.random.default_rng(SEED)
    geo = SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15, random_rotation(rng))
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=384)
    c = update_cache(allocate_cache(g, PLICCutCell), geo, g)
    print("384^3 torus", flush=True)
    cols, normals, intercepts, valid = orc.plic(geo, 3, g.x0, g.d, g.n)
    assert same_bits(c.normals, normals.reshape(c.normals.shape))
    assert same_bits(c.intercepts, intercepts) and np.array_equal(c.is_valid, valid)
    del normals, intercepts, valid
    for name in ("kind", "volume_fraction", "face_fraction"):
        assert same_bits(c.read(name), cols[name]), name
    assert int((cols["kind"] == 2).sum()) > 100_000
    del cols
    ff = c.cells.face_fraction
    assert shared_face_mismatches(ff, g.n, 3) == 0
    assert np.count_nonzero(closure_residual(ff, g.d)) == 0
    del ff, c
    gc.collect()


def test_composite_body_512_full_size_bitwise():
    body = csg_spheres(SEED, 512)
    base = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=512)
    d = base.d[0]
    g = base.moved((-0.5 * d * 20, 0.0, 0.0))
    plain = update_cache(allocate_cache(g, MarchingCubesCutCell), body, g)
    culled = allocate_cache(g, MarchingCubesCutCell)
    culled.set_phi_source("culled")
    update_cache(culled, body, g)
    net = Surrogate(nfreq=64, width=256, depth=4, sigma=8.0, seed=SEED & 0xFFFF)
    net.train(body, nsamples=1 << 21, lo=-1.0, hi=1.0, band=0.05, epochs=8, batch=16384, lr=5e-3,
              seed=SEED)
    sur = allocate_cache(g, MarchingCubesCutCell)
    sur.set_phi_source("surrogate", net, coarse=8, band_cells=3.0)
    update_cache(sur, body, g)
    print(f"512^3 composite at x0 {g.x0}; surrogate {sur.phi_stats()}", flush=True)
    n0 = n1 = n2 = 512
    m0, m1 = n0 + 1, n1 + 1
    plane, layers, cut, differing = n0 * n1, 32, 0, 0
    for k0 in range(0, n2, layers):
        k1 = k0 + layers
        for name in COLS:
            a = plain.read(name, k0 * plane, k1 * plane)
            assert same_bits(culled.read(name, k0 * plane, k1 * plane), a), (k0, name)
            assert same_bits(sur.read(name, k0 * plane, k1 * plane), a), (k0, name)
        lo, hi = m0 * m1 * k0, m0 * m1 * (k1 + 1)
        pe = plain.read("phi", lo, hi)
        assert same_bits(culled.read("phi", lo, hi), pe), k0
        ps = sur.read("phi", lo, hi)
        assert np.array_equal(pe > 0, ps > 0), k0
        body_nodes = pe.size if k1 == n2 else m0 * m1 * layers  # each node layer once
        differing += int((pe[:body_nodes].view(np.uint64) != ps[:body_nodes].view(np.uint64)).sum())
        cells = np.flatnonzero(plain.read("kind", k0 * plane, k1 * plane) == 2)
        cut += cells.size
        i, j, k = cells % n0, (cells // n0) % n1, cells // plane
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    idx = (i + di) + m0 * ((j + dj) + m1 * (k + dk))
                    assert same_bits(ps[idx], pe[idx]), k0
    assert cut == plain.cut_count() and cut > 1_000_000
    assert differing > 0  # the network's values are kept away from the surface
    for k0 in (160, 256, 352):
        x0 = (g.x0[0], g.x0[1], g.x0[2] + k0 * d)
        phi = orc.sample_nodes(body, 3, x0, g.d, (n0, n1, 4))
        assert same_bits(plain.read("phi", m0 * m1 * k0, m0 * m1 * (k0 + 5)), phi), k0
        ref = orc.mc_cells(phi, x0, g.d, (n0, n1, 4))
        assert int((ref["kind"] == 2).sum()) > 10_000
        for name in COLS:
            assert same_bits(plain.read(name, k0 * plane, (k0 + 4) * plane), ref[name]), (k0, name)
    del plain, culled, sur
    net.close()
    gc.collect()


def test_batched_64_bodies_full_size_bitwise():
    from cutcell.autotune import tune_batch
    from cutcell.batch import Batch
    from cutcell.workloads import MovingBodies
    wl = MovingBodies(seed=SEED, nbodies=64)
    meth = {"mc": MarchingCubesCutCell, "plic": PLICCutCell}
    nb = len(wl.bodies)
    caches = [allocate_cache(wl.grid(b, 0), meth[B["method"]]) for b, B in enumerate(wl.bodies)]
    batch = Batch(caches)
    step = 31
    geos = [wl.geometry(b, step) for b in range(nb)]
    grids = [wl.grid(b, step) for b in range(nb)]
    kz = tune_batch(batch, [wl.geometry(b, 0) for b in range(nb)], [wl.grid(b, 0) for b in range(nb)])
    batch.step(geos, grids)
    print(f"64 bodies, {wl.ncells()} cells, depths {sorted(set(kz.values()))}", flush=True)
    for b in range(nb):
        solo = update_cache(allocate_cache(grids[b], meth[wl.bodies[b]["method"]]), geos[b], grids[b])
        for name in COLS:
            assert same_bits(caches[b].read(name), solo.read(name)), (b, name)
        solo.close()
    for m in ("mc", "plic"):
        b = max((i for i in range(nb) if wl.bodies[i]["method"] == m), key=lambda i: grids[i].ncells)
        g = grids[b]
        if m == "mc":
            ref = orc.mc_cells(orc.sample_nodes(geos[b], 3, g.x0, g.d, g.n), g.x0, g.d, g.n)
            names = COLS
        else:
            ref, *_ = orc.plic(geos[b], 3, g.x0, g.d, g.n)
            names = ("kind", "volume_fraction", "face_fraction")
        assert int((ref["kind"] == 2).sum()) > 1000, (m, b)
        for name in names:
            assert same_bits(caches[b].read(name), ref[name]), (m, b, name)
    batch.close()
    del caches
    gc.collect()
