"""Launch-configuration autotuner for the fused path.

The push+deposit kernel trades LDS footprint (halo overhead, per-wave current tiles,
occupancy) against the number of particles that leave their tile each step (tile
perimeter), and the waves of a workgroup share its staged field and current tiles.  The
best configuration depends on the order, the particles per cell and the drift, so it is
measured: each candidate (tile shape, waves per workgroup, push-kernel variant) runs a
few warm-up steps and then a few steps timed with HIP events (Simulation.time_kernels);
the fastest full step wins.  Every candidate gives bit-identical results, so the choice
is purely one of speed.  Results are cached per workload key in a JSON file.

The reference has no counterpart; its nearest knob is the fixed particle-chunk size of
its CPU push (p_cache_size, source/os-param.f90:227).
"""
from __future__ import annotations

import json
import os

# (tile, waves per workgroup, kernel): kernel 0 = compile-time tile variant where one is
# instantiated, 1 = generic kernel (runtime tile shape, any tile that fits the LDS)
DEFAULT_2D = (
    {"tile": (16, 16), "push_waves": 4, "kernel": 0},
    {"tile": (16, 16), "push_waves": 8, "kernel": 0},
    {"tile": (32, 16), "push_waves": 4, "kernel": 0},
    {"tile": (32, 16), "push_waves": 8, "kernel": 0},
    {"tile": (8, 8), "push_waves": 4, "kernel": 0},
)
DEFAULT_3D = (
    {"tile": (8, 8, 4), "push_waves": 4, "kernel": 0},
    {"tile": (8, 8, 4), "push_waves": 4, "kernel": 1},
    {"tile": (4, 4, 4), "push_waves": 4, "kernel": 1},
    {"tile": (8, 4, 4), "push_waves": 4, "kernel": 1},
    {"tile": (8, 8, 8), "push_waves": 4, "kernel": 1},
)
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "osiris_amd", "autotune.json")


def label(cfg) -> str:
    """'16x16/w8/k0' for a candidate configuration."""
    return (f"{'x'.join(map(str, cfg['tile']))}/w{cfg.get('push_waves', 4)}"
            f"/k{cfg.get('kernel', 0)}")


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


def tune(make_sim, key, candidates=None, dims=2, warmup=3, steps=5, cache=CACHE,
         use_cache=True):
    """Return (best_config, {label: ms_per_step}).

    make_sim(config) must return a populated osiris_amd.Simulation built with the
    config's tile, push_waves and kernel (the tuner closes it).  A candidate whose
    simulation cannot be created (tile larger than the grid or the LDS, no kernel with
    that many waves) is skipped.
    """
    db = _load(cache) if cache else {}
    if use_cache and key in db:
        ent = db[key]
        best = dict(ent["best"])
        best["tile"] = tuple(best["tile"])
        return best, dict(ent["ms"])
    cands = candidates or (DEFAULT_2D if dims == 2 else DEFAULT_3D)
    timings, by_label = {}, {}
    for cfg in cands:
        try:
            sim = make_sim(cfg)
        except Exception:  # noqa: BLE001 - candidate not usable for this workload
            continue
        try:
            sim.run(warmup)
            timings[label(cfg)] = sim.time_kernels(steps)["step"]
            by_label[label(cfg)] = cfg
        finally:
            sim.close()
    if not timings:
        raise RuntimeError("autotune: no candidate configuration worked")
    best_label = min(timings, key=timings.get)
    best = dict(by_label[best_label])
    best["tile"] = tuple(best["tile"])
    if cache:
        db[key] = {"best": {**best, "tile": list(best["tile"])}, "ms": timings}
        _save(cache, db)
    return best, timings
