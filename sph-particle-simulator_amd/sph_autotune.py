"""Autotuner for the launch shapes of the two neighbour walks.

Each walk kernel exists in several launch shapes (LDS words per segment x wavefronts per
workgroup; `sphk.walk_shape_names()`).  Results are identical in every shape: a particle's
sums run in the same order whatever the shape.  Only speed changes with the particle
distribution: how many particles share a cell, how many columns the grid has.

`autotune(engine, ...)` times every shape of each walk on the engine's current state with
HIP events (reset-and-step: the state is saved, then restored before each timed step), picks
the fastest shape per walk, sets it on the engine and returns the timings.  Choices are
cached in a JSON file keyed by particle count, grid shape and an occupancy summary, so a later
engine with the same workload skips the search.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional

import numpy as np

import sphk

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "sph_amd_walk_shapes.json")


def workload_key(engine: "sphk.Engine") -> str:
    """Particle count, grid dims and the occupancy distribution of non-empty cells."""
    _, cells = engine.sorted_order()
    _, dims, _ = engine.grid()
    if cells.size:
        _, counts = np.unique(cells, return_counts=True)
        q = np.percentile(counts, [50, 90, 99]).astype(int).tolist()
        mx = int(counts.max())
    else:
        q, mx = [0, 0, 0], 0
    return f"n{engine.size()}_g{dims[0]}x{dims[1]}x{dims[2]}_occ{q[0]}-{q[1]}-{q[2]}-{mx}"


def _time_phases(engine, dt, steps):
    engine.reset_stats()
    engine.set_timing(True)
    for _ in range(steps):
        engine.snapshot_restore()
        engine.step(dt)
    engine.synchronize()
    st = engine.stats()
    engine.set_timing(False)
    k = max(st.timed_steps, 1)
    return st.density_computation_time / k, st.force_computation_time / k


def autotune(engine: "sphk.Engine", dt: float, steps: int = 5,
             cache_path: Optional[str] = DEFAULT_CACHE, use_cache: bool = True) -> Dict:
    """Times every walk shape on the engine's current state and keeps the fastest.

    The engine's particle state is restored afterwards (snapshot/restore on the device)."""
    key = workload_key(engine)
    cache = {}
    if cache_path and use_cache and os.path.exists(cache_path):
        try:
            with open(cache_path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        d, f = cache[key]["density"], cache[key]["force"]
        engine.set_walk_shapes(d, f)
        return dict(cache[key], key=key, cached=True)

    names = sphk.walk_shape_names()
    engine.snapshot_save()
    base_d, base_f = engine.walk_shapes()
    dens_ms, force_ms = [], []
    for v in range(len(names)):
        engine.set_walk_shapes(v, base_f)
        _time_phases(engine, dt, 1)  # warm-up of this shape
        td, _ = _time_phases(engine, dt, steps)
        dens_ms.append(td * 1e3)
    best_d = int(np.argmin(dens_ms))
    for v in range(len(names)):
        engine.set_walk_shapes(best_d, v)
        _time_phases(engine, dt, 1)
        _, tf = _time_phases(engine, dt, steps)
        force_ms.append(tf * 1e3)
    best_f = int(np.argmin(force_ms))
    engine.set_walk_shapes(best_d, best_f)
    engine.snapshot_restore()
    engine.synchronize()
    result = {"density": best_d, "force": best_f, "names": names,
              "density_ms": dens_ms, "force_ms": force_ms}
    if cache_path:
        cache[key] = result
        try:
            os.makedirs(os.path.dirname(cache_path), exist_ok=True)
            with open(cache_path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return dict(result, key=key, cached=False)
