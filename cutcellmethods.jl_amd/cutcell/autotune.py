"""The project's autotuner: picks launch shapes by measurement and remembers them.

* ``tune_stream(cache, geo, grid)`` -- the marching-cubes stream pass of one cache: variant (flat
  tiles, z-march reading phi, z-march sampling the body), slab depth ``kz`` and non-temporal stores,
  timed with the library's HIP-event kernel timer over a few updates; the winner is set on the cache.
* ``tune_batch(batch, geos, grids)`` -- the slab depth of every marching-cubes body of a batch, per
  size class (bodies grouped by cell count), by coordinate descent on the whole graph's device time.

Results are cached in a JSON file (``$CUTCELL_TUNING``, default ``~/.cache/cutcell/tuning.json``)
keyed by method, element type and grid shape, so a shape is measured once per machine.  Every
launch shape gives bit-identical results; only the time differs.
"""
import json
import os

import numpy as np

DEFAULT_STREAM_CANDIDATES = [(2, 1, False), (2, 2, False), (2, 4, False), (2, 8, False), (2, 16, False),
                             (2, 1, True), (2, 4, True), (1, 4, False), (1, 8, False), (1, 16, False),
                             (0, 16, False), (0, 16, True)]


def _path():
    return os.environ.get("CUTCELL_TUNING",
                          os.path.join(os.path.expanduser("~"), ".cache", "cutcell", "tuning.json"))


def _load():
    try:
        with open(_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(db):
    p = _path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def _key(cache, kind):
    return f"{kind}|{cache.method.__name__ if isinstance(cache.method, type) else cache.method}|" \
           f"{cache.dtype.name}|{'x'.join(str(v) for v in cache.grid.n)}"


def tune_stream(cache, geo, grid, candidates=None, repeats=3, use_cache=True):
    """Pick the fastest stream-pass launch shape for this cache and body; returns
    ``{"variant", "kz", "nontemporal", "ms", "trials"}``."""
    from .cache import update_cache
    db = _load() if use_cache else {}
    key = _key(cache, "stream")
    if key in db:
        r = db[key]
        cache.set_tuning(r["variant"], r["kz"], r["nontemporal"])
        return r
    was = cache.get_tuning()
    cache.set_profiling(True)
    trials = []
    for v, kz, nt in candidates or DEFAULT_STREAM_CANDIDATES:
        cache.set_tuning(v, kz, nt)
        update_cache(cache, geo, grid)  # warm this shape
        ts = []
        for _ in range(repeats):
            update_cache(cache, geo, grid)
            ts.append(sum(ms for _, ms in cache.kernel_times()))
        trials.append({"variant": v, "kz": kz, "nontemporal": bool(nt), "ms": float(min(ts))})
    best = min(trials, key=lambda r: r["ms"])
    cache.set_tuning(best["variant"], best["kz"], best["nontemporal"])
    rec = dict(best, trials=trials)
    if use_cache:
        db[key] = rec
        _save(db)
    del was
    return rec


def tune_batch(batch, geos, grids, kz_candidates=(1, 2, 4, 8, 16, 32), classes=3, repeats=2):
    """Per size class of the batch's marching-cubes bodies, and for its PLIC bodies as one more class,
    the slab depth that minimises the graph's device time (coordinate descent over the classes).
    Returns the chosen kz per body."""
    from .cache import MarchingCubesCutCell, PLICCutCell
    sizes = np.array([c.grid.ncells for c in batch.caches])
    mc = [i for i, c in enumerate(batch.caches) if c.method is MarchingCubesCutCell]
    plic = [i for i, c in enumerate(batch.caches) if c.method is PLICCutCell]
    kz = {i: 8 for i in mc}
    kz.update({i: 4 for i in plic})
    if not kz:
        return kz
    edges = np.quantile(sizes[mc], np.linspace(0, 1, classes + 1)[1:-1]) if classes > 1 and mc else []
    cls = {i: int(np.searchsorted(edges, sizes[i], side="right")) for i in mc}
    cls.update({i: classes for i in plic})

    def timed():
        ts = []
        for _ in range(repeats + 1):
            ts.append(batch.step(geos, grids))
        return min(ts[1:])

    for i in kz:
        batch.set_kz(i, kz[i])
    best_t = timed()
    for c in range(classes + 1):
        members = [i for i in kz if cls[i] == c]
        if not members:
            continue
        best_k = kz[members[0]]
        for k in kz_candidates:
            if k == best_k:
                continue
            for i in members:
                batch.set_kz(i, k)
            t = timed()
            if t < best_t:
                best_t, best_k = t, k
        for i in members:
            kz[i] = best_k
            batch.set_kz(i, best_k)
    return kz
