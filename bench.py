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
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time budget of the oracle baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", default="2,16,0",
                    help="stream-pass launch shape: variant,kz,nontemporal")
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
    return {"value": cells / spent, "unit": "cells/s", "cores": orc.num_threads(), "kind": "port",
            "sample": f"oracle update (sample_nodes + cell moments) of {slabs} slabs of 8 k-layers "
                      f"of the {n[0]}x{n[1]}x{n[2]} workload from k=0 up = {cells} cells in "
                      f"{spent:.2f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not offered (the nodal update shards "
              "by z-slab with no exchange -- see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from cutcell import MarchingCubesCutCell, allocate_cache, update_cache

    grid, geo, centre = workload(a.n, a.seed)
    cache = allocate_cache(grid, MarchingCubesCutCell)
    cache.set_profiling(True)
    tv, tk, tn = (int(x) for x in a.tune.split(","))
    cache.set_tuning(tv, tk, bool(tn))
    d = grid.d[0]

    def grid_at(step):
        return grid.moved((-0.5 * d * step, 0.0, 0.0))

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
                "traffic": traffic, "algorithmic_bytes": alg,
                "kernel_ms": {k: round(v, 4) for k, v in avg.items()}}
    update_bytes = BYTES_PER_CELL * ncell + BYTES_PER_NODE * nnode
    from cutcell.cache import probe_bandwidth
    fill_gbs, copy_gbs = probe_bandwidth(8 << 30, 5)
    roofline["probe_fill_GBps"] = round(fill_gbs, 1)
    roofline["probe_copy_GBps"] = round(copy_gbs, 1)
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
                               f"centre {tuple(round(c, 4) for c in centre)}, moving 0.5 cell/step",
                   "cells": ncell, "cut_cells": cache.cut_count(),
                   "update_hbm_fraction": round(update_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if sweep is not None:
        out["tuning_sweep"] = sweep
    print(json.dumps(out))


if __name__ == "__main__":
    main()
