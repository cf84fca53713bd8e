#!/usr/bin/env python3
"""Benchmark: MarchingCubesCutCell update_cache! at 512^3, Float64, on one MI355X.

Workload (one "step" = one blocking update_cache! over the whole grid, inputs resident in HBM):
a sphere of radius 0.4, centre drawn uniformly in [-0.1, 0.1]^3 from the seed, on the isotropic
grid [-1, 1]^3 with 512^3 cells; each step moves the body 0.5 cells along +x (the grid origin
shifts by -0.5 d), so every step recomputes the nodal field and every cell.

Prints one JSON line: cells/s for the whole update, the roofline of the dominant kernel (its
algorithmic bytes / its HIP-event duration on the cache's stream, against 8 TB/s), and the CPU
oracle (the C restatement of the reference's Float64 update, OpenMP over all host cores) timed
on a bounded slab of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cutcellmethods.jl_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 202  # kind 1 + ambiguous 1 + vf 8 + centroid 24 + face_fraction 48 + fcl 96 + ic 24
BYTES_PER_NODE = 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time budget of the oracle baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", default="auto",
                    help="stream-pass launch shape: variant,kz,nontemporal, or 'auto' (the library's "
                         "autotuner picks among the fused z-march shapes on this GPU before warm-up)")
    ap.add_argument("--quick", action="store_true",
                    help="only the headline workload (skip the secondary workloads)")
    ap.add_argument("--only", default=None,
                    choices=["csg", "batched", "secondary", "surface", "quadtree", "plic_surface", "diagnostics"],
                    help="run one secondary workload alone and print its JSON")
    ap.add_argument("--train-epochs", type=int, default=24,
                    help="surrogate training epochs of the composite-body workload")
    ap.add_argument("--train-samples", type=int, default=1 << 22,
                    help="surrogate training samples of the composite-body workload")
    ap.add_argument("--sweep", action="store_true",
                    help="also time every stream-pass launch shape and report them")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "mc_stream_512_traffic.json"))
    return ap.parse_args()


def workload(n, seed):
    from cutcell import CartesianIsoGrid, SDFSphere
    rng = np.random.default_rng(seed)
    centre = tuple(rng.uniform(-0.1, 0.1, 3))
    grid = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=n)
    return grid, SDFSphere(centre, 0.4), centre


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    return orc


def host_cpu():
    """(model name, logical CPUs) of the host the CPU baselines run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def _cpu_info(orc=None, threads=None):
    """cores (the threads the baseline used: the oracle's OpenMP threads unless given), the host
    CPU's model and its logical CPUs."""
    model, nproc = host_cpu()
    return {"cores": orc.num_threads() if threads is None else threads, "cpu_model": model, "nproc": nproc}


def own_fill(cache, names, iters=3):
    """The plain fill kernel over each of the cache's own columns in ``names`` (overwriting them):
    GB/s per column and over all of them together -- the write rate these allocations reach."""
    per, tb, tms = {}, 0, 0.0
    for name in names:
        gbs = cache.probe_fill(name, iters)
        nb = cache.column_ptr(name)[1] // 16 * 16
        per[name] = round(gbs, 1)
        tb += nb
        tms += nb / gbs / 1e6
    return per, tb / tms / 1e6


def cpu_baseline(grid, geo, budget_s):
    """The oracle (all host cores, OpenMP) over the same workload in slabs of 8 k-layers, from k = 0
    upward (wrapping), until about budget_s seconds of CPU update time have been spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    n = grid.n
    chunk = 8
    spent, cells, k, slabs = 0.0, 0, 0, 0
    while spent < budget_s:
        k1 = min(n[2], k + chunk)
        x0 = (grid.x0[0], grid.x0[1], grid.x0[2] + k * grid.d[2])
        sub = (n[0], n[1], k1 - k)
        t = time.perf_counter()
        phi = orc.sample_nodes(geo, 3, x0, grid.d, sub)
        orc.mc_cells(phi, x0, grid.d, sub)
        spent += time.perf_counter() - t
        cells += n[0] * n[1] * (k1 - k)
        slabs += 1
        k = 0 if k1 >= n[2] else k1
    return {"value": cells / spent, "unit": "cells/s", **_cpu_info(orc), "kind": "port",
            "sample": f"oracle update (sample_nodes + cell moments) of {slabs} slabs of 8 k-layers "
                      f"of the {n[0]}x{n[1]}x{n[2]} workload from k=0 up = {cells} cells in "
                      f"{spent:.2f} s"}


def _timed(fn, reps):
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), ts


def secondary(seed, cpu=True):
    """The other workloads, each at its stated size: returns a dict of measurements."""
    from cutcell import (CartesianIsoGrid, MarchingCubesCutCell, MarchingSquaresCutCell, PLICCutCell,
                         SDFSphere, SDFTorus, allocate_cache, generate_mesh, random_rotation,
                         update_cache)
    out = {}
    # marching squares, circle r=0.5 on [-1,1]^2 at sz=0.01 (200x200), then the contour
    g2 = CartesianIsoGrid((-1.0, -1.0), (1.0, 1.0), sz=0.01)
    c2 = allocate_cache(g2, MarchingSquaresCutCell)
    geo2 = SDFSphere((0.0, 0.0), 0.5)
    for _ in range(5):
        update_cache(c2, geo2, g2)
    t_up, _ = _timed(lambda: update_cache(c2, geo2, g2), 50)
    t_mesh, _ = _timed(lambda: generate_mesh(c2, g2), 10)
    nodes, lines = generate_mesh(c2, g2)
    out["ms_200"] = {"update_us": round(t_up * 1e6, 1), "generate_mesh_ms": round(t_mesh * 1e3, 3),
                     "contour_segments": int(lines.shape[0]), "cells": g2.ncells}
    if cpu:
        orc = _oracle()
        phi_buf = np.empty(int(np.prod(g2.nodegrid_size())))

        def cpu_update():
            orc.ms_cells(orc.sample_nodes(geo2, 2, g2.x0, g2.d, g2.n, out=phi_buf), g2.x0, g2.d, g2.n)

        cpu_update()
        t_cu, _ = _timed(cpu_update, 20)
        phi2 = orc.sample_nodes(geo2, 2, g2.x0, g2.d, g2.n)
        t_cm, _ = _timed(lambda: orc.ms_contour(phi2, g2.x0, g2.d, g2.n), 10)
        cn, cl = orc.ms_contour(phi2, g2.x0, g2.d, g2.n)
        out["ms_200"]["cpu_baseline"] = {
            "update_us": round(t_cu * 1e6, 1), "generate_mesh_ms": round(t_cm * 1e3, 3), **_cpu_info(orc),
            "kind": "port", "sample": "the oracle's update (sample_nodes + cell moments, OpenMP) and its "
                                      "contour (threaded screen, serial weld) of the same 200x200 grid",
            "same_contour": bool(np.array_equal(cn, nodes) and np.array_equal(cl, lines))}
    del c2
    # PLIC, torus R=0.5 r=0.15, seeded tilt, 384^3
    rng = np.random.default_rng(seed)
    g3 = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=384)
    tor = SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15, random_rotation(rng))
    c3 = allocate_cache(g3, PLICCutCell)
    # the write rate of the columns the sweep writes (the centroid columns keep their NaN seed)
    fill_cols, fill_all = own_fill(c3, ("kind", "volume_fraction", "face_fraction", "normals", "intercepts",
                                        "is_valid"))
    c3.set_profiling(True)
    from cutcell.autotune import tune_stream
    # the sweep's tile shape and z-depth, tuned on this GPU like the headline's stream shape
    ptune = tune_stream(c3, tor, g3, candidates=[(1, 16, False), (2, 8, False), (2, 16, False), (2, 32, False)],
                        repeats=2, use_cache=False)
    for _ in range(3):
        update_cache(c3, tor, g3)
    kts = []

    def plic_step():
        update_cache(c3, tor, g3)
        kts.append(dict(c3.kernel_times()))

    t_p, _ = _timed(plic_step, 10)
    kt = {k: float(np.median([v[k] for v in kts])) for k in kts[0]}
    n3 = g3.ncells
    two_pass, fused = 131 * n3, 90 * n3
    sweep_ms = kt.get("plic_fused")
    out["plic_384"] = {"update_ms": round(t_p * 1e3, 3), "cells_per_s": n3 / t_p,
                       "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
                       "sweep_tuning": [ptune["variant"], ptune["kz"], ptune["nontemporal"]],
                       "algorithmic_bytes_fused": fused, "algorithmic_bytes_two_pass": two_pass,
                       "hbm_frac_two_pass": round(two_pass / t_p / 1e9 / HBM_PEAK_GBS, 4),
                       "cut_cells": int((c3.read("kind") == 2).sum()),
                       "own_columns_fill_GBps": round(fill_all, 1), "column_fill_GBps": fill_cols}
    if sweep_ms:
        gbs = fused / (sweep_ms * 1e-3) / 1e9
        out["plic_384"]["roofline_fused"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                             "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                                             "frac_of_own_fill": round(gbs / fill_all, 4)}
    if cpu:
        orc = _oracle()
        layers = 16
        k0 = g3.n[2] // 2 - layers // 2
        x0 = (g3.x0[0], g3.x0[1], g3.x0[2] + k0 * g3.d[2])
        sub = (g3.n[0], g3.n[1], layers)
        ns = sub[0] * sub[1] * layers
        buf = (orc.alloc_cols(ns, 3, nan=True), np.zeros((ns, 3)), np.zeros(ns), np.zeros(ns, np.uint8))
        orc.plic(tor, 3, x0, g3.d, sub, out=buf)
        t_c, _ = _timed(lambda: orc.plic(tor, 3, x0, g3.d, sub, out=buf), 3)
        out["plic_384"]["cpu_baseline"] = {
            "value": round(t_c * g3.n[2] / layers * 1e3, 1), "unit": "ms per update", **_cpu_info(orc),
            "kind": "port", "sample": f"the oracle's fit and face passes (OpenMP) over the {layers} middle "
                                      f"k-layers ({ns} cells), scaled to the {g3.n[2]} layers"}
    del c3
    return out


def surface(seed, n=512, reps=8, cpu=True):
    """generate_mesh on the marching-cubes workload (the 512^3 sphere, moving 0.5 cell per step):
    the device march (HIP events: count + emit kernels, the host round trip for the output sizes
    excluded), the whole call with the copy of the mesh to the host, and the oracle's serial march
    of the same field on one host core (the reference's generate_mesh is a serial host march)."""
    from cutcell import MarchingCubesCutCell, allocate_cache, cell_indices, generate_mesh, update_cache
    grid, geo, _ = workload(n, seed)
    c = allocate_cache(grid, MarchingCubesCutCell)
    c.set_profiling(True)
    d = grid.d[0]
    dev, wall, ci = [], [], []
    for s in range(reps + 1):
        g = grid.moved((-0.5 * d * s, 0.0, 0.0))
        update_cache(c, geo, g)
        t = time.perf_counter()
        mesh = generate_mesh(c, g, warn=False)
        t1 = time.perf_counter()
        kt = dict(c.kernel_times())
        t2 = time.perf_counter()
        cell_indices(mesh, g)
        t3 = time.perf_counter()
        if s == 0:
            continue  # first march allocates the work arrays
        dev.append(kt["mesh_count"] + kt["mesh_emit"])
        wall.append((t1 - t) * 1e3)
        ci.append((t3 - t2) * 1e3)
    nv, nt = len(mesh.nodes), len(mesh.elements)
    out = {"workload": f"generate_mesh of the {n}^3 marching-cubes sphere cache",
           "vertices": nv, "triangles": nt, "cut_cells": c.cut_count(),
           "device_ms": round(float(np.median(dev)), 4), "call_ms_with_host_copy": round(float(np.median(wall)), 3),
           "cell_indices_ms": round(float(np.median(ci)), 3),
           "triangles_per_s_device": nt / (float(np.median(dev)) * 1e-3),
           "kernel_ms": {k: round(v, 4) for k, v in kt.items()}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        phi = c.phi
        g = grid.moved((-0.5 * d * reps, 0.0, 0.0))
        t = time.perf_counter()
        nodes, tris, _ = orc.mc_mesh(phi, g.x0, g.d, g.n)
        t_cpu = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(t_cpu * 1e3 / 2, 1), "unit": "ms per march", **_cpu_info(threads=1), "kind": "port",
                               "sample": "the oracle's serial march of the same field (it marches twice, counting "
                                         "then filling; half the time is reported)",
                               "same_mesh": bool(len(nodes) == nv and len(tris) == nt)}
    del c
    return out


def quadtree(seed, level=12, reps=20, cpu=True):
    """Marching squares on a quadtree's leaves (the reference's AdaptiveMesh{2} cache): the circle of
    its tree tests (centre (0.13, -0.07), r 0.9) in a band refined to `level` round the circle
    (base [-2, 2]^2 of 8 x 8, band half-width 2.5 leaf diagonals), about 10^6 leaves at level 12."""
    from cutcell import AdaptiveMesh, CartesianGrid, MarchingSquaresCutCell, SDFCircle, allocate_cache
    from cutcell import generate_mesh, update_cache
    centre, radius = (0.13, -0.07), 0.9
    base = CartesianGrid((-2.0, -2.0), (8, 8), (0.5, 0.5))
    mesh = AdaptiveMesh(base, max_level=level, initial_level=1)
    for _ in range(level):
        h = mesh.elem_size()
        c = mesh.elem_centroid() - np.asarray(centre)
        dist = np.abs(np.sqrt((c * c).sum(axis=1)) - radius)
        flags = (mesh.level < level) & (dist <= 2.5 * np.sqrt((h * h).sum(axis=1)))
        if not flags.any():
            break
        mesh.refine(flags)
    geo = SDFCircle(centre, radius)
    cache = allocate_cache(mesh, MarchingSquaresCutCell)
    cache.set_profiling(True)
    for _ in range(3):
        update_cache(cache, geo, mesh)
    ts, ks = [], []
    for _ in range(reps):
        t = time.perf_counter()
        update_cache(cache, geo, mesh)
        ts.append(time.perf_counter() - t)
        ks.append(dict(cache.kernel_times())["leaf_cells"])
    t_up, k_ms = float(np.median(ts)), float(np.median(ks))
    n = mesh.nleaves
    alg = n * (106 + 32)  # each leaf's record (2D, FP64) and its four corner values
    tm = []
    for _ in range(12):  # the first call also sizes the contour's work arrays
        t = time.perf_counter()
        nodes, lines = generate_mesh(cache, mesh)
        tm.append(time.perf_counter() - t)
    t_mesh = float(np.median(tm[2:]))
    out = {"workload": f"MarchingSquaresCutCell on {n} quadtree leaves (levels 1-{level}), circle r={radius}",
           "leaves": n, "cut_leaves": cache.cut_count(), "update_ms": round(t_up * 1e3, 4),
           "leaves_per_s": n / t_up, "kernel_ms": round(k_ms, 4),
           "kernel_hbm_frac": round(alg / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "generate_mesh_ms": round(t_mesh * 1e3, 2), "generate_mesh_first_call_ms": round(tm[0] * 1e3, 2),
           "contour_segments": int(lines.shape[0])}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        orc.ms_leaves(geo, base.x0, base.d, mesh.level[:1000], mesh.coord[:1000])
        t = time.perf_counter()
        orc.ms_leaves(geo, base.x0, base.d, mesh.level, mesh.coord)
        t_cpu = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(t_cpu * 1e3, 2), "unit": "ms per update", **_cpu_info(orc),
                               "kind": "port", "sample": "the oracle's leaf pass over the same tree (OpenMP)"}
    del cache
    return out


def plic_surface(seed, n=384, reps=5, cpu=True):
    """generate_mesh on the PLIC cache (the torus R=0.5 r=0.15 with a seeded tilt at 384^3, every cell's
    plane clipped on the device) and calc_volume of the same body on the same grid, each beside the
    oracle's serial restatement."""
    from cutcell import (CartesianIsoGrid, PLICCutCell, SDFTorus, allocate_cache, calc_volume,
                         generate_mesh, random_rotation, update_cache)
    rng = np.random.default_rng(seed)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=n)
    tor = SDFTorus((0.0, 0.0, 0.0), 0.5, 0.15, random_rotation(rng))
    c = allocate_cache(g, PLICCutCell)
    c.set_profiling(True)
    dev, wall = [], []
    for r in range(reps + 1):
        update_cache(c, tor, g)  # a new update: the next call clips again
        t = time.perf_counter()
        nodes, el = generate_mesh(c, g)
        t1 = time.perf_counter()
        kt = dict(c.kernel_times())
        if r:
            dev.append(sum(v for k, v in kt.items() if "host_sync" not in k))  # the kernels, not the round trips
            wall.append((t1 - t) * 1e3)
    vt = []
    for r in range(reps + 1):
        t = time.perf_counter()
        vol = calc_volume(tor, g)
        if r:
            vt.append((time.perf_counter() - t) * 1e3)
    out = {"workload": f"PLICCutCell surface and calc_volume, torus at {n}^3", "cells": g.ncells,
           "triangles": int(len(el)), "vertices": int(len(nodes)),
           "surface_device_ms": round(float(np.median(dev)), 4),
           "surface_kernel_ms": {k: round(v, 4) for k, v in kt.items()},
           "surface_call_ms_with_host_copy": round(float(np.median(wall)), 3),
           "calc_volume_ms": round(float(np.median(vt)), 3), "volume": float(vol)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        t = time.perf_counter()
        v_ref = orc.calc_volume(tor, 3, g.x0, g.d, g.n)
        t_v = time.perf_counter() - t
        nrm, ic = c.normals, c.intercepts
        t = time.perf_counter()
        orc.plic_surface(3, g.x0, g.d, g.n, nrm, ic)
        t_s = (time.perf_counter() - t) / 2  # it clips twice, counting then filling
        out["cpu_baseline"] = {"calc_volume_ms": round(t_v * 1e3, 1), "surface_ms": round(t_s * 1e3, 1),
                               **_cpu_info(threads=1), "kind": "port",
                               "sample": "the oracle's serial calc_volume and surface over the same grid",
                               "volume_rel_diff": abs(float(vol) - v_ref) / v_ref}
    del c
    return out


DIAG_BYTES_PER_CELL_3D_F64 = 1 + 1 + 8 + 6 * 8 + 12 * 8  # kind, ambiguous, volume, faces, face centroids


def diagnostics_pass(seed, n=512, reps=8, cpu=True, layers=16):
    """The diagnostics pass (report, closure residual, interface areas, shared faces) over the
    marching-cubes workload's cache at n^3: the fused device reduction (HIP events), the whole call,
    and the host's numpy readers over `layers` k-layers of the same columns, scaled to the grid."""
    from cutcell import MarchingCubesCutCell, allocate_cache, diagnostics, readers, update_cache
    grid, geo, _ = workload(n, seed)
    c = allocate_cache(grid, MarchingCubesCutCell)
    update_cache(c, geo, grid)
    c.set_profiling(True)
    dev, wall = [], []
    for r in range(reps + 1):
        t = time.perf_counter()
        d = diagnostics(c, grid)
        t1 = time.perf_counter()
        if r:
            dev.append(dict(c.kernel_times())["diagnostics"])
            wall.append((t1 - t) * 1e3)
    ms = float(np.median(dev))
    nbytes = grid.ncells * DIAG_BYTES_PER_CELL_3D_F64
    out = {"workload": f"diagnostics of the {n}^3 marching-cubes sphere cache", "cells": grid.ncells,
           "cut": d["cut"], "face_mismatches": d["face_mismatches"], "closure_nonzero": d["closure_nonzero"],
           "interface_area": d["interface_area"], "device_ms": round(ms, 4),
           "call_ms": round(float(np.median(wall)), 4),
           "roofline": {"bound": "hbm", "achieved": round(nbytes / (ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(nbytes / (ms * 1e-3) / 8e12, 4),
                        "bytes_per_cell": DIAG_BYTES_PER_CELL_3D_F64}}
    if cpu:
        mid = grid.n[2] // 2 - layers // 2
        lo, hi = grid.n[0] * grid.n[1] * mid, grid.n[0] * grid.n[1] * (mid + layers)
        cols = {k: c.read(k, lo, hi) for k in ("kind", "ambiguous", "volume_fraction", "face_fraction",
                                               "face_centroid_local")}
        t = time.perf_counter()
        readers.cut_cell_report(cols["kind"], cols["ambiguous"], cols["volume_fraction"])
        res = readers.closure_residual(cols["face_fraction"], grid.d)
        int(np.any(res != 0, axis=1).sum())
        float(readers.interface_area(cols["face_fraction"], grid.d).sum())
        readers.interface_normal_area(cols["face_fraction"], grid.d).sum(axis=0)
        ff = cols["face_fraction"].reshape(layers, grid.n[1], grid.n[0], 6)
        fcl = cols["face_centroid_local"].reshape(layers, grid.n[1], grid.n[0], 6, 2)
        int((ff[:, :, :-1, 1] != ff[:, :, 1:, 0]).sum() + (ff[:, :-1, :, 3] != ff[:, 1:, :, 2]).sum() +
            (ff[:-1, :, :, 5] != ff[1:, :, :, 4]).sum())
        int(np.any(fcl[:, :, :-1, 1] != fcl[:, :, 1:, 0], axis=-1).sum() +
            np.any(fcl[:, :-1, :, 3] != fcl[:, 1:, :, 2], axis=-1).sum() +
            np.any(fcl[:-1, :, :, 5] != fcl[1:, :, :, 4], axis=-1).sum())
        t_cpu = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(t_cpu * grid.n[2] / layers * 1e3, 1), "unit": "ms per pass",
                               **_cpu_info(threads=1), "kind": "host numpy readers",
                               "sample": f"{layers} middle k-layers ({hi - lo} cells) already on the host, "
                                         f"scaled to the {n}^3 grid"}
    del c
    return out


def fp32_error(grid, geo, cache64, tune=None):
    """The Float32 variant: its update time and its error against the Float64 cache (bitwise the
    oracle) on a slab through the body.  With ``tune`` "auto" its stream shape is tuned on the GPU
    the same way as the Float64 cache's; otherwise ``tune`` is the (variant, kz, nontemporal) used."""
    from cutcell import MarchingCubesCutCell, allocate_cache, update_cache
    c32 = allocate_cache(grid, MarchingCubesCutCell, np.float32)
    if tune == "auto":
        from cutcell.autotune import tune_stream
        cands = [(2, kz, nt) for kz in (1, 2, 4, 8) for nt in (False, True)]
        rec = tune_stream(c32, geo, grid, candidates=cands, repeats=3, use_cache=False)
        c32.set_tuning(rec["variant"], rec["kz"], bool(rec["nontemporal"]))
    elif tune is not None:
        c32.set_tuning(*tune)
    for _ in range(3):
        update_cache(c32, geo, grid)
    t32, _ = _timed(lambda: update_cache(c32, geo, grid), 8)
    update_cache(cache64, geo, grid)
    n = grid.n
    k0 = n[2] // 2 - 8
    lo, hi = n[0] * n[1] * k0, n[0] * n[1] * (k0 + 16)
    err = {}
    k64 = cache64.read("kind", lo, hi)
    k32 = c32.read("kind", lo, hi)
    err["kind_mismatch"] = int((k64 != k32).sum())
    for name in ("volume_fraction", "face_fraction", "centroid", "interface_centroid"):
        a = cache64.read(name, lo, hi)
        b = c32.read(name, lo, hi).astype(np.float64)
        err[name + "_max_abs"] = float(np.max(np.abs(a - b)))
    # the centroid is the first moment divided by the volume: ill-conditioned for slivers
    vf = cache64.read("volume_fraction", lo, hi)
    dc = np.max(np.abs(cache64.read("centroid", lo, hi) - c32.read("centroid", lo, hi).astype(np.float64)),
                axis=1)
    for lim in (1e-6, 1e-3):
        sel = (k64 == 2) & (vf >= lim) & (vf <= 1.0 - lim)
        err[f"centroid_max_abs_cut_vf_{lim:g}_to_1-{lim:g}"] = float(dc[sel].max()) if sel.any() else 0.0
    err["cell_size"] = float(grid.d[0])
    tuning = list(c32.get_tuning())
    del c32
    return {"update_ms": round(t32 * 1e3, 3), "cells_per_s": grid.ncells / t32, "stream_tuning": tuning,
            "error_vs_fp64_on_16_layers": err}


MLP_FLOP_PER_POINT = 2 * (128 * 256 + 3 * 256 * 256 + 256)  # 64 freqs -> 4 x 256 -> 1, one forward


def _compare_caches(exact, sur, grid, layers=16):
    """Slab by slab on the host: every CutCellData column bitwise equal, phi equal at every corner
    of every cut cell, the sign of phi equal everywhere; returns counts."""
    n0, n1, n2 = grid.n
    m0, m1 = n0 + 1, n1 + 1
    res = {"columns_bitwise_equal": True, "cut_corner_phi_equal": True, "sign_equal": True,
           "phi_nodes_differing": 0, "cut_cells": 0}
    cols = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
            "face_centroid_local", "interface_centroid")
    for k0 in range(0, n2, layers):
        k1 = min(n2, k0 + layers)
        lo, hi = n0 * n1 * k0, n0 * n1 * k1
        for name in cols:
            a, b = exact.read(name, lo, hi), sur.read(name, lo, hi)
            if not np.array_equal(a.view(np.uint8), b.view(np.uint8)):
                res["columns_bitwise_equal"] = False
        kind = exact.read("kind", lo, hi)
        pe = exact.read("phi", m0 * m1 * k0, m0 * m1 * (k1 + 1))
        ps = sur.read("phi", m0 * m1 * k0, m0 * m1 * (k1 + 1))
        if not np.array_equal(pe > 0, ps > 0):
            res["sign_equal"] = False
        last = k1 == n2
        body = pe.size if last else m0 * m1 * (k1 - k0)  # count each node layer once
        res["phi_nodes_differing"] += int((pe[:body].view(np.uint64) != ps[:body].view(np.uint64)).sum())
        cut = np.flatnonzero(kind == 2)
        res["cut_cells"] += int(cut.size)
        i, j, k = cut % n0, (cut // n0) % n1, cut // (n0 * n1)
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    idx = (i + di) + m0 * ((j + dj) + m1 * (k + dk))
                    if not np.array_equal(pe[idx].view(np.uint64), ps[idx].view(np.uint64)):
                        res["cut_corner_phi_equal"] = False
    return res


def csg(seed, n=512, train_epochs=24, train_samples=1 << 22, cpu=True):
    """The composite body (smooth union, k = 0.02, of 512 spheres, centres U[-0.7, 0.7]^3, radii
    U[0.05, 0.15]) at n^3, moving 0.5 cell per step: the exact path (every node folds every sphere),
    the culled exact path (bitwise the same phi) and the learned surrogate (network trained here on
    exact samples; exact fallback where it cannot be trusted).  Update times are medians of
    blocking updates; the network's training time is reported on its own."""
    from cutcell import CartesianIsoGrid, MarchingCubesCutCell, Surrogate, allocate_cache, update_cache
    from cutcell.workloads import csg_spheres
    body = csg_spheres(seed, 512)
    grid = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=n)
    d = grid.d[0]

    def at(step):
        return grid.moved((-0.5 * d * step, 0.0, 0.0))

    out = {"body": "smooth union k=0.02 of 512 spheres, centres U[-0.7,0.7]^3, radii U[0.05,0.15]",
           "cells": grid.ncells}
    net = Surrogate(nfreq=64, width=256, depth=4, sigma=8.0, seed=seed & 0xFFFF)
    t = time.perf_counter()
    hist = net.train(body, nsamples=train_samples, lo=-1.0, hi=1.0, band=0.05, epochs=train_epochs,
                     batch=16384, lr=5e-3, seed=seed)
    tt = time.perf_counter() - t
    tflop = 3 * MLP_FLOP_PER_POINT * train_samples * train_epochs / 1e12
    out["training"] = {"seconds": round(tt, 2), "samples": train_samples, "epochs": train_epochs,
                       "batch": 16384, "final_loss": float(hist[-1]), "first_epoch_loss": float(hist[0]),
                       "approx_tflop": round(tflop, 2), "approx_tflops": round(tflop / tt, 1)}
    from cutcell.autotune import tune_stream
    # the fused z-march evaluating the culled or surrogate phi from the bricks, or a phi pass and a
    # stream pass reading it (the z-march reads the phi pass's sign bitmap, so shallow depths are
    # cheap; flat tiles read phi)
    cands = [(2, 8, False), (2, 16, False), (2, 8, True), (2, 16, True),
             (1, 2, False), (1, 4, False), (1, 8, False), (1, 2, True), (0, 16, False)]
    cache = allocate_cache(grid, MarchingCubesCutCell)
    cache.set_profiling(True)
    res = {}
    for src, reps in (("naive", 3), ("culled", 6)):
        cache.set_phi_source(src)
        tune = tune_stream(cache, body, at(0), candidates=cands, repeats=2, use_cache=False)
        update_cache(cache, body, at(0))
        ts, kts = [], []
        for s in range(reps):
            t = time.perf_counter()
            update_cache(cache, body, at(1 + s))
            ts.append(time.perf_counter() - t)
            kts.append(dict(cache.kernel_times()))
        res[src] = {"update_ms": round(1e3 * float(np.median(ts)), 3),
                    "kernel_ms": {k: round(float(np.median([kt[k] for kt in kts])), 4) for k in kts[0]},
                    "stream_tuning": [tune["variant"], tune["kz"], tune["nontemporal"]],
                    "tuning_trials_ms": {f"{r['variant']},{r['kz']},{int(r['nontemporal'])}": round(r["ms"], 3)
                                         for r in tune["trials"]}}
        if src == "culled":
            st = cache.phi_stats()
            res[src]["mean_kept_spheres_per_brick"] = round(st["kept_sum"] / st["bricks"], 2)
    # the surrogate on the same cache allocation (where an allocation lands moves the stream pass)
    cache.set_phi_source("surrogate", net, coarse=8, band_cells=3.0)
    stune = tune_stream(cache, body, at(0), candidates=cands, repeats=2, use_cache=False)
    update_cache(cache, body, at(0))
    ts, kts = [], []
    for s in range(8):
        t = time.perf_counter()
        update_cache(cache, body, at(1 + s))
        ts.append(time.perf_counter() - t)
        kts.append(dict(cache.kernel_times()))
    st = cache.phi_stats()
    km = {k: round(float(np.median([kt[k] for kt in kts])), 4) for k in kts[0]}
    nc = n // 8 + 2
    lat_points = nc ** 3
    nn_tflops = lat_points * MLP_FLOP_PER_POINT / (km["nn_lattice"] * 1e-3) / 1e12
    res["surrogate"] = {
        "update_ms": round(1e3 * float(np.median(ts)), 3), "kernel_ms": km,
        "stream_tuning": [stune["variant"], stune["kz"], stune["nontemporal"]],
        "tuning_trials_ms": {f"{r['variant']},{r['kz']},{int(r['nontemporal'])}": round(r["ms"], 3)
                             for r in stune["trials"]},
        "coarse_lattice": f"{nc}^3 (8 cells)", "band": "3h",
        "exact_node_fraction": round(st["exact"] / st["nodes"], 4),
        "uncertified_fraction": round(st["uncertified"] / st["nodes"], 4),
        "band_fraction": round(st["band"] / st["nodes"], 4),
        "sign_contradicted_fraction": round(st["sign_fail"] / st["nodes"], 6),
        "certificate_evals": st["cert_evals"],
        "nn_lattice_tflops": round(nn_tflops, 1),
        "mfma_utilisation_bf16": round(nn_tflops / 2500.0, 4)}
    # identical results: the same grid position through the exact path (a second cache) and the
    # surrogate
    sur = allocate_cache(grid, MarchingCubesCutCell)
    update_cache(cache, body, at(20))
    update_cache(sur, body, at(20))
    res["agreement"] = _compare_caches(sur, cache, grid)
    # the analytic headline body (the sphere) on the same cache allocation, its own tuned shape: what
    # the surrogate update costs beyond an update whose phi is free
    _, sphere, _ = workload(n, seed)
    cache.set_phi_source("naive")
    atune = tune_stream(cache, sphere, at(0), candidates=[(2, 1, True), (2, 1, False), (2, 4, True), (2, 4, False)],
                        repeats=2, use_cache=False)
    ts = []
    for s in range(8):
        t = time.perf_counter()
        update_cache(cache, sphere, at(1 + s))
        ts.append(time.perf_counter() - t)
    res["analytic_sphere_same_cache"] = {"update_ms": round(1e3 * float(np.median(ts)), 3),
                                         "stream_tuning": [atune["variant"], atune["kz"], atune["nontemporal"]]}
    res["surrogate_minus_analytic_ms"] = round(res["surrogate"]["update_ms"]
                                               - res["analytic_sphere_same_cache"]["update_ms"], 3)
    res["speedup_vs_naive"] = round(res["naive"]["update_ms"] / res["surrogate"]["update_ms"], 2)
    res["speedup_vs_culled"] = round(res["culled"]["update_ms"] / res["surrogate"]["update_ms"], 3)
    res["timed_on"] = "one cache allocation for naive, culled and surrogate; a second for the agreement check"
    if cpu:
        # the oracle's exact fold (every node folds all 512 spheres) and cell pass on a slab of
        # `layers` k-layers through the body, scaled to the grid
        orc = _oracle()
        layers = 2
        k0 = n // 2 - layers // 2
        g0 = at(0)
        x0 = (g0.x0[0], g0.x0[1], g0.x0[2] + k0 * g0.d[2])
        sub = (n, n, layers)
        phi = np.empty((n + 1) * (n + 1) * (layers + 1))
        cols = orc.alloc_cols(n * n * layers, 3)
        t = time.perf_counter()
        orc.mc_cells(orc.sample_nodes(body, 3, x0, g0.d, sub, out=phi), x0, g0.d, sub, cols=cols)
        t_c = time.perf_counter() - t
        res["cpu_baseline"] = {"value": round(t_c * n / layers * 1e3, 1), "unit": "ms per update", **_cpu_info(orc),
                               "kind": "port", "sample": f"the oracle's exact update of {layers} middle k-layers "
                                                         f"({n * n * layers} cells), scaled to the {n} layers"}
    out.update(res)
    del cache, sur
    net.close()
    return out


def batched(seed, steps=32, cpu=True, cpu_budget_s=10.0):
    from cutcell import MarchingCubesCutCell, PLICCutCell, allocate_cache
    from cutcell.autotune import tune_batch
    from cutcell.batch import Batch
    from cutcell.workloads import MovingBodies
    wl = MovingBodies(seed=seed, nbodies=64)
    meth = {"mc": MarchingCubesCutCell, "plic": PLICCutCell}
    caches = [allocate_cache(wl.grid(b, 0), meth[B["method"]]) for b, B in enumerate(wl.bodies)]
    batch = Batch(caches)
    nb = len(caches)
    geos0 = [wl.geometry(b, 0) for b in range(nb)]
    grids0 = [wl.grid(b, 0) for b in range(nb)]
    for _ in range(3):
        batch.step(geos0, grids0)
    kz = tune_batch(batch, geos0, grids0)
    lat, dev = [], []
    for s in range(steps):
        geos = [wl.geometry(b, s) for b in range(nb)]
        grids = [wl.grid(b, s) for b in range(nb)]
        t = time.perf_counter()
        dev.append(batch.step(geos, grids))
        lat.append(time.perf_counter() - t)
    lat = np.array(lat) * 1e3
    ncell = wl.ncells()
    mc_cells = sum(int(np.prod(B["n"])) for B in wl.bodies if B["method"] == "mc")
    mc_nodes = sum(int(np.prod([na + 1 for na in B["n"]])) for B in wl.bodies if B["method"] == "mc")
    # what each body's update writes at least: a marching-cubes body its 202 B per cell and phi, a
    # PLIC body (the fused sweep) its 90 B per cell
    alg = BYTES_PER_CELL * mc_cells + BYTES_PER_NODE * mc_nodes + 90 * (ncell - mc_cells)
    out = {"bodies": nb, "cells_per_step": ncell, "mc_cells": mc_cells,
           "cells_per_s": ncell / (np.mean(lat) * 1e-3),
           "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                               "p99": round(float(np.percentile(lat, 99)), 3)},
           "graph_device_ms_p50": round(float(np.percentile(dev, 50)), 3),
           "algorithmic_bytes": alg,
           "hbm_frac": round(alg / (float(np.percentile(dev, 50)) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "kz_by_body_class": sorted(set(kz.values()))}
    del batch, caches
    if cpu:
        # the oracle updates the bodies of the last step one after another (outputs allocated
        # beforehand, untimed) until the budget is spent; the step time is the sum over bodies
        orc = _oracle()
        s = steps - 1
        spent, cells, done = 0.0, 0, 0
        for b in range(nb):
            if spent >= cpu_budget_s:
                break
            g, geo = wl.grid(b, s), wl.geometry(b, s)
            nc = g.ncells
            if wl.bodies[b]["method"] == "mc":
                phi = np.empty(int(np.prod(g.nodegrid_size())))
                cols = orc.alloc_cols(nc, 3)
                t = time.perf_counter()
                orc.mc_cells(orc.sample_nodes(geo, 3, g.x0, g.d, g.n, out=phi), g.x0, g.d, g.n, cols=cols)
            else:
                buf = (orc.alloc_cols(nc, 3, nan=True), np.zeros((nc, 3)), np.zeros(nc), np.zeros(nc, np.uint8))
                t = time.perf_counter()
                orc.plic(geo, 3, g.x0, g.d, g.n, out=buf)
            spent += time.perf_counter() - t
            cells += nc
            done += 1
        rate = cells / spent
        out["cpu_baseline"] = {"value": rate, "unit": "cells/s", "step_ms_estimate": round(ncell / rate * 1e3, 1),
                               **_cpu_info(orc), "kind": "port",
                               "sample": f"the oracle's update of bodies 0..{done - 1} of step {s} ({cells} cells "
                                         f"in {spent:.2f} s); the step estimate scales to its {ncell} cells"}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not offered (the nodal update would "
              "shard by z-slab with no exchange, each GPU recomputing its boundary node layer)",
              file=sys.stderr)
        sys.exit(2)
    from cutcell import MarchingCubesCutCell, allocate_cache, update_cache

    if a.only:
        fn = {"csg": lambda: csg(a.seed, train_epochs=a.train_epochs, train_samples=a.train_samples,
                                 cpu=not a.no_cpu),
              "batched": lambda: batched(a.seed, cpu=not a.no_cpu),
              "secondary": lambda: secondary(a.seed, cpu=not a.no_cpu),
              "surface": lambda: surface(a.seed, cpu=not a.no_cpu),
              "quadtree": lambda: quadtree(a.seed, cpu=not a.no_cpu),
              "plic_surface": lambda: plic_surface(a.seed, cpu=not a.no_cpu),
              "diagnostics": lambda: diagnostics_pass(a.seed, cpu=not a.no_cpu)}[a.only]
        print(json.dumps({a.only: fn()}))
        return
    grid, geo, centre = workload(a.n, a.seed)
    cache = allocate_cache(grid, MarchingCubesCutCell)
    cache.set_profiling(True)
    d = grid.d[0]

    def grid_at(step):
        return grid.moved((-0.5 * d * step, 0.0, 0.0))

    from cutcell.cache import probe_bandwidth
    fill_gbs, copy_gbs = probe_bandwidth(8 << 30, 5)  # also brings the memory clocks up
    # the write rate of the cache's own columns (where an allocation lands moves it); every one of
    # them is rewritten by the updates below
    col_fill, own_fill_gbs = own_fill(cache, ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
                                              "face_centroid_local", "interface_centroid", "phi"))
    tuning = None
    if a.tune == "auto":
        from cutcell.autotune import tune_stream
        cands = [(2, kz, nt) for kz in (1, 2, 4, 8) for nt in (False, True)]
        tuning = tune_stream(cache, geo, grid_at(0), candidates=cands, repeats=4, use_cache=False)
        tv, tk, tn = tuning["variant"], tuning["kz"], int(tuning["nontemporal"])
    else:
        tv, tk, tn = (int(x) for x in a.tune.split(","))
    cache.set_tuning(tv, tk, bool(tn))
    for s in range(a.warmup):
        update_cache(cache, geo, grid_at(s))
    per_kernel = {}
    t0 = time.perf_counter()
    for s in range(a.steps):
        update_cache(cache, geo, grid_at(a.warmup + s))  # blocks: the reference's contract
        for name, ms in cache.kernel_times():
            per_kernel.setdefault(name, []).append(ms)
    dt = time.perf_counter() - t0
    ncell = int(np.prod(grid.n))
    nnode = int(np.prod(grid.nodegrid_size()))
    ms_per_step = 1e3 * dt / a.steps
    value = ncell * a.steps / dt

    avg = {k: float(np.mean(v)) for k, v in per_kernel.items()}
    dom = max(avg, key=avg.get)
    if dom == "mc_stream":
        # cells written + phi once (written when the pass samples the body, read otherwise)
        alg = BYTES_PER_CELL * ncell + BYTES_PER_NODE * nnode
    elif dom == "sample_nodes":
        alg = BYTES_PER_NODE * nnode
    else:
        alg = BYTES_PER_CELL * cache.cut_count()
    achieved = alg / (avg[dom] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(a.traffic):
        try:
            traffic = json.load(open(a.traffic)).get("bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "traffic_source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this kernel on this workload, "
                                  "scaled by the same counters on plain copy/fill kernels "
                                  f"({os.path.relpath(a.traffic, ROOT)})" if traffic is not None else None,
                "algorithmic_bytes": alg,
                "kernel_ms": {k: round(v, 4) for k, v in avg.items()}}
    update_bytes = BYTES_PER_CELL * ncell + BYTES_PER_NODE * nnode
    roofline["probe_fill_GBps"] = round(fill_gbs, 1)
    roofline["probe_copy_GBps"] = round(copy_gbs, 1)
    roofline["own_columns_fill_GBps"] = round(own_fill_gbs, 1)
    roofline["face_centroid_local_fill_GBps"] = col_fill["face_centroid_local"]
    roofline["frac_of_own_fill"] = round(achieved / own_fill_gbs, 4)
    sweep = None
    if a.sweep:
        sweep = []
        for v, kz, nt in [(0, 16, 0), (0, 16, 1), (1, 8, 0), (1, 16, 0), (1, 16, 1), (2, 4, 0),
                          (2, 8, 0), (2, 16, 0), (2, 32, 0), (2, 16, 1), (2, 64, 1)]:
            cache.set_tuning(v, kz, bool(nt))
            update_cache(cache, geo, grid_at(0))
            ts = []
            t1 = time.perf_counter()
            for s in range(4):
                update_cache(cache, geo, grid_at(s + 1))
                ts.append(sum(ms for _, ms in cache.kernel_times()))
            wall = (time.perf_counter() - t1) / 4 * 1e3
            sweep.append({"variant": v, "kz": kz, "nontemporal": nt, "device_ms": round(min(ts), 4),
                          "wall_ms": round(wall, 3), "kernels": cache.kernel_times()})
        cache.set_tuning(tv, tk, bool(tn))
    cpu = None if a.no_cpu else cpu_baseline(grid, geo, a.cpu_seconds)
    out = {
        "metric": "update_cache! cells/s @512^3 MC FP64",
        "value": value, "unit": "cells/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"MarchingCubesCutCell update_cache! {a.n}^3 Float64, sphere r=0.4 "
                               f"centre {tuple(round(float(c), 4) for c in centre)}, moving 0.5 cell/step",
                   "cells": ncell, "cut_cells": cache.cut_count(),
                   "stream_tuning": {"variant": tv, "kz": tk, "nontemporal": bool(tn),
                                     "chosen_by": "autotuner" if tuning else "--tune"},
                   "update_hbm_fraction": round(update_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if sweep is not None:
        out["tuning_sweep"] = sweep
    if not a.quick:
        out["fp32"] = fp32_error(grid, geo, cache, "auto" if a.tune == "auto" else (tv, tk, bool(tn)))
        del cache
        out.update(secondary(a.seed, cpu=not a.no_cpu))
        out["mc_surface_512"] = surface(a.seed, cpu=not a.no_cpu)
        out["ms_quadtree"] = quadtree(a.seed, cpu=not a.no_cpu)
        out["plic_surface_384"] = plic_surface(a.seed, cpu=not a.no_cpu)
        out["diagnostics_512"] = diagnostics_pass(a.seed, cpu=not a.no_cpu)
        out["batched_64_bodies"] = batched(a.seed, cpu=not a.no_cpu)
        out["csg_512_spheres"] = csg(a.seed, train_epochs=a.train_epochs, train_samples=a.train_samples,
                                     cpu=not a.no_cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
