"""Tile-shape autotuner for the fused path.

The push+deposit kernel trades LDS footprint (halo overhead, per-wave current tiles,
occupancy) against the number of particles that leave their tile each step (tile
perimeter).  The best tile depends on the order, the particles per cell and the drift,
so it is measured: each candidate runs a few warm-up steps and then a few steps timed
with HIP events (Simulation.time_kernels), and the fastest full step wins.  Results are
cached per workload key in a JSON file.
"""
from __future__ import annotations

import json
import os

DEFAULT_2D = ((16, 16), (32, 16), (8, 8))
DEFAULT_3D = ((8, 8, 4), (8, 8, 8), (4, 4, 4))
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "osiris_amd", "autotune.json")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(data, f, indent=1)
    except OSError:
        pass


def tune_tile(make_sim, key, candidates=None, dims=2, warmup=3, steps=5, cache=CACHE,
              use_cache=True):
    """Return (best_tile, {tile: ms_per_step}).

    make_sim(tile) must return a populated osiris_amd.Simulation (closed by the tuner).
    """
    db = _load(cache) if cache else {}
    if use_cache and key in db:
        ent = db[key]
        return tuple(ent["tile"]), {tuple(map(int, k.split("x"))): v for k, v in ent["ms"].items()}
    cands = candidates or (DEFAULT_2D if dims == 2 else DEFAULT_3D)
    timings = {}
    for t in cands:
        try:
            sim = make_sim(t)
        except Exception:  # tile not usable for this grid (e.g. LDS too large)
            continue
        try:
            sim.run(warmup)
            timings[tuple(t)] = sim.time_kernels(steps)["step"]
        finally:
            sim.close()
    if not timings:
        raise RuntimeError("autotune: no candidate tile worked")
    best = min(timings, key=timings.get)
    if cache:
        db[key] = {"tile": list(best), "ms": {"x".join(map(str, k)): v for k, v in timings.items()}}
        _save(cache, db)
    return best, timings
