#!/usr/bin/env python3
"""Headline benchmark: particle-updates/s of the full PIC time step on one MI355X.

Workload (the reference's headline configuration): 2D periodic counter-streaming
electron beams, 1024 x 1024 cells, dx = dy = 0.05, dt = 0.035, 64 particles per cell
split equally into two beams drifting at u_x = +/-0.2 with u_th = 0.01 (uniform random
in-cell positions, seeded), quadratic splines, Boris pusher, FP32.  A step is one full
time step: field gather + push + charge-conserving deposit + current fold + Yee +
energy diagnostics, all resident in HBM.

Prints ONE JSON line with the metric, the roofline of the dominant kernel (push +
deposit, measured with HIP events on the simulation stream) and a CPU baseline: the
reference's own AVX2 FP32 kernel (simd/os-spec-push-avx2.c) on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s (push+deposit+Yee) 2D quad 64ppc; ensemble p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes (FP32 SoA): per particle read ix2+x2+u3+q = 32 B, write
# ix2+x2+u3 = 28 B; per cell the gather reads E,B (24 B) and the deposit writes J (12 B).
BYTES_PER_PARTICLE = 60.0
BYTES_PER_CELL_PUSH = 36.0
# Yee per cell: B half step 36 B (read E, B; write B), E step 48 B (read E, B; write E;
# J counted with the deposit), B half step 36 B
BYTES_PER_CELL_YEE = 120.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def oracle_check(oa, steps=30):
    """Error of the GPU path against the FP64 oracle, reported beside the CPU baseline:
    the reference's CPU-runnable configuration (2D periodic thermal plasma 128 x 128,
    dx = 0.1, dt = 0.07, 16 ppc, u_th = 0.05, linear splines, Boris) on the same seeded
    particles, energy histories compared row by row."""
    from oracle import oracle as O

    rng = np.random.default_rng(2024)
    nx, dx, dt, ppc = (128, 128), (0.1, 0.1), 0.07, 16
    cells = np.stack(np.meshgrid(np.arange(1, nx[0] + 1), np.arange(1, nx[1] + 1),
                                 indexing="ij"), -1).reshape(-1, 2)
    ix = np.repeat(cells, ppc, axis=0).astype(np.int32)
    n = len(ix)
    x = rng.uniform(-0.5, 0.5, (n, 2)).astype(np.float32)
    u = rng.normal(0.0, 0.05, (n, 3)).astype(np.float32)
    q = np.full(n, -1.0 / ppc, np.float32)
    with oa.Simulation(2, nx, dx, dt, 1) as sim:
        sid = sim.add_species(-1.0, oa.OSR_BORIS)
        sim.set_particles(sid, ix, x, u, q)
        sim.run(steps)
        fld, par = sim.history()
    ref = O.Sim(2, nx, dx, dt, 1)
    ref.add_species(-1.0, 0, ix, x, u, q)
    ref.step(steps)
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    fscale = np.maximum(rfld.sum(1, keepdims=True), 1e-300)
    return {"workload": "2D periodic thermal plasma 128x128, 16 ppc, u_th=0.05, linear, Boris",
            "steps": steps, "reference": "FP64 CPU oracle (restatement of the reference)",
            "max_rel_err_particle_energy": float(np.max(np.abs(par - rpar) / np.abs(rpar))),
            "max_rel_err_field_energy": float(np.max(np.abs(fld - rfld) / fscale)),
            "tolerance": 1e-4}


def cpu_baseline(nx, ppc_total, n_total, order, dx, dt, seconds, threads):
    """The reference's AVX2 FP32 fused kernel on a strip of the same workload, host
    threads with private currents (the reference's OpenMP scheme) + FP32 Yee on the full
    grid; extrapolated to the full particle count."""
    from oracle import oracle as O

    if not O.ref_available("s"):
        return None
    # the -Ofast build (the reference's production optimisation level) when present
    fast = O.ref_available("s_fast")
    flags = "gcc -Ofast -mavx2 -mfma" if fast else "gcc -O2 -mavx2 -mfma"
    rng = np.random.default_rng(12345)
    rows = max(1, min(nx[1], 4_194_304 // (nx[0] * ppc_total)))
    g = O.make_grid(2, nx, order)
    glo = g.gcl[0]
    e = g.zeros(3, np.float32)
    b = g.zeros(3, np.float32)
    e[...] = rng.normal(0, 1e-3, e.shape)
    b[...] = rng.normal(0, 1e-3, b.shape)
    cells = np.stack(np.meshgrid(np.arange(1, nx[0] + 1), np.arange(1, rows + 1), indexing="ij"),
                     -1).reshape(-1, 2)
    ix = np.repeat(cells, ppc_total, axis=0).astype(np.int32)
    n = len(ix)
    x = rng.uniform(-0.5, 0.5, (n, 2)).astype(np.float32)
    u = rng.normal(0, 0.01, (n, 3)).astype(np.float32)
    u[: n // 2, 0] += 0.2
    u[n // 2:, 0] -= 0.2
    q = np.full(n, -1.0 / ppc_total, np.float32)
    timer = O.RefSimdTimer(order, ix, x, u, q, -1.0, e, b, glo, dx, dt, threads, fast=fast)
    timer.step()  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        timer.step()
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    t_push = el / reps  # push + deposit + reduce for n particles
    # FP32 Yee of the reference restatement on the full grid
    j = g.zeros(3, np.float32)
    t1 = time.perf_counter()
    ny = 3
    for _ in range(ny):
        O.yee_advance(g, dx, dt, e, b, j, prec="f32")
    t_yee = (time.perf_counter() - t1) / ny
    t_step = t_push * (n_total / n) + t_yee
    return {
        "value": n_total / t_step,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"reference FNAME(vadvance_deposit_2d_s{order}) AVX2 FP32 "
                   f"(simd/os-spec-push-avx2.c, {flags}) on {n} particles "
                   f"({nx[0]}x{rows} cells x {ppc_total} ppc, two beams), {threads} threads "
                   f"with private J + reduction, {reps} steps in {el:.1f} s; plus FP32 Yee "
                   f"(oracle) on the full {nx[0]}x{nx[1]} grid ({t_yee * 1e3:.1f} ms); "
                   f"extrapolated to {n_total} particles"),
    }


BYTES_PER_PARTICLE_3D = 76.0  # read ix3+x3+u3+q = 40 B, write ix3+x3+u3 = 36 B


def cpu_baseline_c3(seconds):
    """The FP32 oracle (a single-threaded C restatement; the reference's own SIMD kernels
    have no Vay pusher) on a 32^3 sample of the c3 workload: push + deposit of every
    particle, repeated for about `seconds` seconds."""
    from oracle import oracle as O

    rng = np.random.default_rng(7)
    nx = (32, 32, 32)
    g = O.make_grid(3, nx, 3)
    e = g.zeros(3, np.float32)
    b = g.zeros(3, np.float32)
    b[..., 2] = 0.5
    cells = np.stack(np.meshgrid(*[np.arange(1, 33)] * 3, indexing="ij"), -1).reshape(-1, 3)
    ix = np.repeat(cells, 8, axis=0).astype(np.int32)
    n = len(ix)
    x = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    u = rng.normal(0, 0.1, (n, 3)).astype(np.float32)
    q = np.full(n, -1.0 / 8, np.float32)
    j = g.zeros(3, np.float32)
    reps, t0 = 0, time.perf_counter()
    while True:
        ixc = ix.copy()  # every repetition starts from the same cells (no boundary pass)
        O.advance_deposit(g, 3, 1, -1.0, (0.1,) * 3, 0.05, e, b, j, ixc, x, u, q, prec="f32")
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": n * reps / el, "unit": "particle-updates/s", "cores": 1, "kind": "port",
            "sample": (f"FP32 oracle (oracle/pic_oracle.c, one thread) push + deposit of {n} "
                       f"particles (32^3 cells x 8 ppc, cubic, Vay, B0 = 0.5 z), {reps} "
                       f"repetitions in {el:.1f} s")}


def run_c3(args, oa):
    """3D periodic magnetized thermal plasma: 256^3 cells, dx = 0.1, dt = 0.05, 8 ppc
    (2x2x2, uniform random), u_th = 0.1, uniform B0 = 0.5 z, cubic splines, Vay pusher."""
    nx, dx, dt = (256, 256, 256), (0.1, 0.1, 0.1), 0.05
    steps = min(args.steps, 100)

    def make(cfg):
        s = oa.Simulation(3, nx, dx, dt, 3, tile=tuple(cfg["tile"]), reorder=args.reorder,
                          kernel=cfg.get("kernel", 0))
        sid = s.add_species(-1.0, oa.OSR_VAY)
        s.generate(sid, (2, 2, 2), density=1.0, uth=(0.1, 0.1, 0.1), seed=3, random_pos=True)
        s.fill_fields((0.0, 0.0, 0.0), (0.0, 0.0, 0.5))
        return s

    tune = None
    if len(args.tile) == 3 or not args.autotune:
        cfg = {"tile": tuple(args.tile) if len(args.tile) == 3 else (8, 8, 4), "kernel": 0}
    else:
        from osiris_amd.autotune import label, tune as autotune

        cfg, tune = autotune(make, "3d-o3-vay-256^3-ppc8", dims=3, warmup=2, steps=3,
                             use_cache=False)
        log("autotune (ms/step): " + ", ".join(f"{k}: {v:.3f}" for k, v in tune.items())
            + f" -> {label(cfg)}")
    sim = make(cfg)
    lc = sim.launch_config
    tile = lc["tile"]
    sid = 0
    n_total = sim.count(sid)
    cells = int(np.prod(nx))
    log(f"workload c3: {n_total} particles, footprint {sim.footprint / 2**30:.2f} GiB")
    sim.run(args.warmup)
    t0 = time.perf_counter()
    sim.run(steps)
    el = time.perf_counter() - t0
    value = n_total * steps / el
    kt = sim.time_kernels(args.timing_steps)
    log(f"timed: {steps} steps in {el:.3f} s -> {el * 1e3 / steps:.3f} ms/step, {value:.4e} upd/s; "
        "kernel ms/step: " + ", ".join(f"{k} {v:.4f}" for k, v in kt.items()))
    t_launch = kt["push_deposit"] * 1e-3
    bytes_launch = n_total * BYTES_PER_PARTICLE_3D + cells * BYTES_PER_CELL_PUSH
    achieved = bytes_launch / t_launch / 1e9
    traffic = None
    tj_path = os.path.join(ROOT, "profiles", "traffic_push_deposit_3d_c3.json")
    if os.path.exists(tj_path):
        try:
            with open(tj_path) as f:
                tj = json.load(f)
            # counters measured on the same kernel instantiation and launch shape only
            if (tuple(tj.get("tile", ())) == tuple(tile) and lc["compiled_tile"]
                    and tj.get("push_waves") == lc["push_waves"]):
                traffic = tj.get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    relayouts = sim.relayouts
    sim.close()
    out = {
        "metric": "particle-updates/s (push+deposit+Yee) 3D cubic Vay 8ppc",
        "value": value, "unit": "particle-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": el * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded counter-based RNG on the device)",
        "config": {"workload": ("3D periodic magnetized thermal plasma 256^3, dx=0.1, "
                                "dt=0.05, 8 ppc, u_th=0.1, B0=0.5 z, cubic splines, Vay"),
                   "particles": n_total, "cells": cells, "tile": list(tile),
                   "push_waves": lc["push_waves"], "compiled_tile": lc["compiled_tile"],
                   "autotune_ms_per_step": tune, "reorder": args.reorder,
                   "relayouts": relayouts},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": (f"push_deposit_3d_t<3,1,{tile[0]},{tile[1]},{tile[2]}>"
                                if lc["compiled_tile"] else "push_deposit_3d<3,1,256>"),
                     "kernel_ms": t_launch * 1e3,
                     "algorithmic_bytes": bytes_launch},
        "kernels_ms_per_step": kt,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_c3(args.cpu_seconds)
    print(json.dumps(out), flush=True)


# c4: laser-plasma interaction between open x walls (emf_bound "lindman", spec_bound
# "absorbing", zpulse_wall from x-min), periodic y.  Units: the laser frequency is the
# normalising frequency (omega0 = 1), so lengths are in c/omega0, times in 1/omega0 and a
# density of 0.01 is 0.01 n_c (plasma frequency 0.1 omega0).  dx = 0.2 resolves the
# wavelength 2 pi with 31 cells and dt = 0.14 the period with 45 steps.  The pulse:
# a0 = 2, linear polarisation in the simulation plane (E2), gaussian temporal envelope of
# duration 30 (exp(-2 ((t - 90) / 30)^2), cut at +-3 durations), gaussian beam of waist 20
# centred on the box (y = 128) and focused at the top of the ramp (x = 20).  Electrons:
# 8 ppc cold, density rising linearly from the x-min wall over 100 cells, then flat.
C4 = dict(dx=(0.2, 0.5), dt=0.14, density=0.01, ramp_cells=100, gap_cells=0, ppc=(4, 2),
          laser=dict(a0=2.0, omega0=1.0, pol=0.0, tenv_type="gaussian", tenv_duration=30.0,
                     tenv_range=180.0, per_type="gaussian", per_w0=20.0, per_focus=20.0,
                     per_center=(128.0,)))


def c4_particles(nx, gap, ramp, ppc, density):
    """Cold electrons on a regular lattice (ppc[0] x ppc[1] per cell, os-psource-std.f03:
    846-866) in cells x > gap, density rising linearly over `ramp` cells to `density`;
    charge per particle -n(x)/ppc."""
    xs = np.arange(gap + 1, nx[0] + 1)
    cells = np.stack(np.meshgrid(xs, np.arange(1, nx[1] + 1), indexing="ij"), -1).reshape(-1, 2)
    k = np.arange(ppc[0] * ppc[1])
    lx = (k // ppc[1]).astype(np.float32)
    ly = (k % ppc[1]).astype(np.float32)
    off = np.stack([(2 * (lx + 1) - 1 - ppc[0]) * (0.5 / ppc[0]),
                    (2 * (ly + 1) - 1 - ppc[1]) * (0.5 / ppc[1])], 1).astype(np.float32)
    ix = np.repeat(cells, len(k), axis=0).astype(np.int32)
    x = np.tile(off, (len(cells), 1))
    u = np.zeros((len(ix), 3), np.float32)
    dens = density * np.clip((ix[:, 0] - gap) / float(ramp), 0.0, 1.0)
    q = (-dens / len(k)).astype(np.float32)
    return ix, x, u, q


def oracle_check_c4(oa, steps=256, xwall="lindman"):
    """The c4 physics (cold ramped plasma from the x-min wall, a0 = 2 laser with gaussian
    temporal envelope and gaussian beam from x-min, omega0 = 1, Lindman or VPML x walls,
    absorbing particles) on a 256 x 32 grid (51.2 x 16 c/omega0) with a 40-cell ramp, a
    shorter pulse (duration 4, range 24) and waist 4 centred on the box: GPU against the
    FP64 oracle on the same particles, over the steps in which the laser enters the plasma
    (the FP32 oracle's own deviation is reported beside)."""
    from oracle import oracle as O

    nx = (256, 32)
    dx, dt = C4["dx"], C4["dt"]
    laser = dict(C4["laser"], tenv_duration=4.0, tenv_range=24.0, per_w0=4.0, per_focus=4.0,
                 per_center=(8.0,))
    ix, x, u, q = c4_particles(nx, 0, 40, C4["ppc"], C4["density"])
    with oa.Simulation(2, nx, dx, dt, 2, tile=(32, 16), xbnd=xwall) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        sim.add_zpulse_wall(**laser)
        sim.run(steps)
        fld, par = sim.history()
        n_gpu = sim.count(sid)
    refs = {}
    for prec in ("f64", "f32"):
        ref = O.Sim(2, nx, dx, dt, 2, prec=prec, open_x=True, xwall=xwall)
        ref.add_species(-1.0, 0, ix, x, u, q)
        ref.add_zpulse_wall(**laser)
        ref.step(steps)
        refs[prec] = (np.array(ref.fld_hist), np.array(ref.par_hist), len(ref.species[0].q))
    rfld, rpar, n_ref = refs["f64"]
    fscale = np.maximum(rfld.sum(1, keepdims=True), 1e-300)
    # the cold plasma has no kinetic energy until the laser reaches it: rows relative to
    # themselves once the energy exceeds 1e-3 of its peak; the FP32 oracle's own
    # deviation from the FP64 one is reported beside the GPU's
    peak = np.abs(rpar[:, 0]).max()
    live = np.abs(rpar[:, 0]) > 1e-3 * peak

    def rel(a):
        return float(np.max(np.abs(a[live] - rpar[live]) / np.abs(rpar[live])))

    err_gpu, err_f32 = rel(par), rel(refs["f32"][1])
    err_fld = float(np.max(np.abs(fld - rfld) / fscale))
    return {"workload": ("c4 physics on 256x32 (omega0=1, ramp 40 cells from the x-min wall, "
                         f"gaussian envelope duration 4 range 24, w0=4, {xwall} walls), "
                         f"{len(q)} particles"),
            "steps": steps, "reference": "FP64 CPU oracle (restatement of the reference)",
            "particles_left": {"gpu": int(n_gpu), "oracle": int(n_ref)},
            "max_rel_err_particle_energy": err_gpu,
            "fp32_oracle_max_rel_err_particle_energy": err_f32,
            "particle_energy_rows_compared": int(live.sum()),
            "max_rel_err_field_energy": err_fld,
            "tolerance": 1e-4,
            "pass": bool(err_gpu <= 1e-4 and err_fld <= 1e-4 and n_gpu == n_ref)}


def run_c4(args, oa):
    """Laser-plasma interaction: 2048 x 512 cells, dx = 0.2, dy = 0.5 c/omega0, dt = 0.14
    /omega0, cold electrons (8 ppc) at 0.01 n_c behind a 100-cell linear ramp from the
    x-min wall, a0 = 2 laser (omega0 = 1, gaussian envelope of duration 30, gaussian beam
    of waist 20 focused at x = 20) launched from x-min, Lindman (or, with --xwall vpml,
    8-cell VPML) x walls with absorbing particles, periodic y, quadratic, Boris."""
    nx = (2048, 512)
    dx, dt = C4["dx"], C4["dt"]
    tile = tuple(args.tile) if args.tile != [16, 16] else (32, 16)
    # the job is 1500 steps: the warm-up steps are its first ones, the rest are timed
    steps = min(args.steps or 1500, 1500 - args.warmup)
    ix, x, u, q = c4_particles(nx, C4["gap_cells"], C4["ramp_cells"], C4["ppc"], C4["density"])
    sim = oa.Simulation(2, nx, dx, dt, 2, tile=tile, xbnd=args.xwall)
    sid = sim.add_species(-1.0)
    sim.set_particles(sid, ix, x, u, q)
    sim.add_zpulse_wall(**C4["laser"])
    del ix, x, u, q
    cells = nx[0] * nx[1]
    log(f"workload c4: {sim.count(sid)} particles, footprint {sim.footprint / 2**30:.2f} GiB")
    sim.run(args.warmup)
    n0 = sim.count(sid)
    t0 = time.perf_counter()
    sim.run(steps)
    el = time.perf_counter() - t0
    n1 = sim.count(sid)
    n_mean = 0.5 * (n0 + n1)
    value = n_mean * steps / el
    kt = sim.time_kernels(args.timing_steps)
    log(f"timed: {steps} steps in {el:.3f} s -> {el * 1e3 / steps:.3f} ms/step, {value:.4e} upd/s "
        f"({n0} -> {n1} particles); kernel ms/step: " +
        ", ".join(f"{k} {v:.4f}" for k, v in kt.items()))
    t_launch = kt["push_deposit"] * 1e-3
    bytes_launch = n1 * BYTES_PER_PARTICLE + cells * BYTES_PER_CELL_PUSH
    achieved = bytes_launch / t_launch / 1e9
    fld, par = sim.history()
    relayouts = sim.relayouts
    waves = sim.push_waves
    sim.close()
    # the push's HBM counters (profiles/kernels_c4.md) are from the first steps, whose
    # particles are not the late run's, so no traffic figure is claimed for this launch
    traffic = None
    out = {
        "metric": "particle-updates/s (push+deposit+Yee+walls+laser) 2D quadratic 8ppc",
        "value": value, "unit": "particle-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": el * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (lattice-loaded cold plasma, host generated)",
        "config": {"workload": ("2D laser-plasma 2048x512, dx=0.2, dy=0.5 c/w0, dt=0.14/w0, "
                                "0.01 n_c with 100-cell ramp, 8 ppc cold, a0=2 zpulse_wall "
                                "from x-min (w0=1, gaussian envelope 30/w0, waist 20 c/w0), "
                                f"{'Lindman' if args.xwall == 'lindman' else 'VPML (8 cells)'} "
                                "x walls + absorbing particles, periodic y, quadratic"),
                   "particles": [n0, n1], "cells": cells, "tile": list(tile),
                   "relayouts": relayouts,
                   "field_energy_last": float(fld[-1].sum()),
                   "particle_energy_last": float(par[-1].sum())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "push_deposit_2d_t<2,0,%d,%d,%d>" % (*tile, waves),
                     "kernel_ms": t_launch * 1e3,
                     "algorithmic_bytes": bytes_launch},
        "kernels_ms_per_step": kt,
    }
    if not args.no_cpu:
        cb = cpu_baseline(nx, 8, n1, 2, dx, dt, args.cpu_seconds,
                          args.cpu_threads or min(16, os.cpu_count() or 1))
        if cb is not None:
            cb["oracle_check"] = oracle_check_c4(oa, xwall=args.xwall)
            out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


def run_c5(args, ens=None):
    """Ensemble serving: a stream of independent 2D periodic thermal-plasma jobs (Poisson
    arrivals at 20 jobs/s; side 64-512, 4-64 ppc, u_th log-uniform in [0.01, 0.2], order 1
    or 2, 100-2000 steps) replayed in real time (latency = completion - arrival, p50 and
    p99) and saturated (all jobs at t = 0: throughput), against the same jobs run one after
    the other as plain simulations.  Prints one JSON line and returns the exit status;
    when the library refuses ensembles the line carries its message and the status is 1."""
    if ens is None:
        from osiris_amd import ensemble as ens
    stream = ens.poisson_jobs(n=args.jobs)
    total = float(sum(j.particles * j.nsteps for _, j in stream))
    base = {"metric": METRIC, "unit": "particle-updates/s", "n_gpus": 1,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": (f"ensemble serving: {len(stream)} independent 2D periodic "
                                    "thermal plasmas, Poisson arrivals 20 jobs/s, side "
                                    "64-512, 4-64 ppc, order 1-2, 100-2000 steps"),
                       "jobs": len(stream), "particle_updates": total}}
    try:
        realtime = ens.serve(stream, realtime=True)
        saturated = ens.serve(stream, realtime=False)
    except Exception as e:  # the library's refusal or a failed job: no number
        base["error"] = str(e)
        print(json.dumps(base), flush=True)
        return 1
    seq = ens.run_sequential(stream)
    base.update({"value": saturated["particle_updates_per_s"],
                 "steps": None, "warmup": 0,
                 "ms_per_step": None,
                 "ensemble": {"latency_p50_s": realtime["latency_p50_s"],
                              "latency_p99_s": realtime["latency_p99_s"],
                              "realtime_wall_s": realtime["wall_s"],
                              "jobs_per_s": saturated["jobs_per_s"],
                              "sequential_jobs_per_s": seq["jobs_per_s"],
                              "speedup_vs_sequential": saturated["jobs_per_s"] / seq["jobs_per_s"],
                              "sequential_particle_updates_per_s": seq["particle_updates_per_s"]}})
    print(json.dumps(base), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: c2 200, c3 100, c4 the rest of the 1500-step "
                         "job after the warm-up)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, nargs=2, default=[1024, 1024])
    ap.add_argument("--ppc", type=int, nargs=2, default=[8, 4],
                    help="particles per cell per beam along x and y (2 beams)")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--tile", type=int, nargs="+", default=[16, 16])
    ap.add_argument("--workload", choices=("c2", "c3", "c4", "c5"), default="c2",
                    help="c2: the headline 2D two-stream step (default); c3: 3D periodic "
                         "magnetized thermal plasma 256^3, 8 ppc, cubic, Vay; c4: 2D "
                         "laser-plasma 2048x512 between open x walls; c5: ensemble serving "
                         "of a Poisson stream of independent jobs (latency and throughput "
                         "against sequential runs)")
    ap.add_argument("--jobs", type=int, default=256, help="c5: jobs in the stream")
    ap.add_argument("--xwall", choices=("lindman", "vpml"), default="lindman",
                    help="c4: field walls of the open x boundaries")
    ap.add_argument("--reorder", type=int, default=0,
                    help="in-tile particle order restored every N steps (0 = library default: "
                         "4 in 3D, off in 2D; negative = off)")
    ap.add_argument("--push-waves", type=int, default=0,
                    help="waves per push workgroup without autotuning (0 = the tile's default)")
    ap.add_argument("--autotune", action="store_true", default=True)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--plain-steps", type=int, default=20,
                    help="steps of the one-kernel-per-phase GPU path (0 = skip)")
    ap.add_argument("--timing-steps", type=int, default=10,
                    help="steps measured per kernel with HIP events (roofline)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles",
                                                      "traffic_push_deposit_2d_c2.json"))
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run the plain copy kernel over GIB GiB (traffic-counter "
                         "calibration and bandwidth probe) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU path "
                                   "(spatial domain decomposition) is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import osiris_amd as oa

    if oa.device_count() < 1:
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)

    if args.copy_probe > 0:
        nbytes = int(args.copy_probe * 2**30)
        ms = oa.copy_probe(nbytes, 5)
        print(json.dumps({"copy_probe_bytes": nbytes, "ms": ms,
                          "GB/s (read+write)": 2 * nbytes / ms / 1e6}), flush=True)
        return

    if args.steps is None and args.workload != "c4":
        args.steps = 100 if args.workload == "c3" else 200
    if args.workload == "c3":
        run_c3(args, oa)
        return
    if args.workload == "c4":
        run_c4(args, oa)
        return
    if args.workload == "c5":
        sys.exit(run_c5(args))

    nx = tuple(args.nx)
    dx, dt = (0.05, 0.05), 0.035
    ppc_beam = int(np.prod(args.ppc))

    def make(cfg, plain=False):
        s = oa.Simulation(2, nx, dx, dt, args.order, tile=tuple(cfg["tile"]), plain=plain,
                          push_waves=0 if plain else cfg.get("push_waves", 0),
                          kernel=0 if plain else cfg.get("kernel", 0))
        for k, ufl in enumerate((0.2, -0.2)):
            sid = s.add_species(-1.0, oa.OSR_BORIS)
            s.generate(sid, args.ppc, density=0.5, ufl=(ufl, 0.0, 0.0), uth=(0.01, 0.01, 0.01),
                       seed=1000 + k, random_pos=True)
        return s

    cfg = {"tile": tuple(args.tile[:2]), "push_waves": args.push_waves, "kernel": 0}
    tune = None
    if args.autotune:
        from osiris_amd.autotune import label, tune as autotune

        key = f"2d-o{args.order}-boris-{nx[0]}x{nx[1]}-ppc{2 * ppc_beam}-beams"
        cfg, tune = autotune(make, key, dims=2, use_cache=False)
        log("autotune (ms/step): " + ", ".join(f"{k}: {v:.3f}" for k, v in tune.items())
            + f" -> {label(cfg)}")
    tile = cfg["tile"]
    sim = make(cfg)
    waves = sim.push_waves
    n_total = sum(sim.count(s) for s in range(2))
    cells = nx[0] * nx[1]
    log(f"workload: {nx[0]}x{nx[1]} cells, {n_total} particles, footprint "
        f"{sim.footprint / 2**30:.2f} GiB")

    sim.run(args.warmup)
    t0 = time.perf_counter()
    sim.run(args.steps)  # returns after the device finished (status read back)
    el = time.perf_counter() - t0
    ms_step = el * 1e3 / args.steps
    value = n_total * args.steps / el
    log(f"timed: {args.steps} steps in {el:.3f} s -> {ms_step:.3f} ms/step, {value:.4e} upd/s")

    # per-kernel timing with HIP events on the simulation stream
    kt = sim.time_kernels(args.timing_steps)
    nsp = 2
    # both species (same pusher) are pushed by one launch
    t_launch = kt["push_deposit"] * 1e-3
    bytes_launch = n_total * BYTES_PER_PARTICLE + cells * BYTES_PER_CELL_PUSH
    achieved = bytes_launch / t_launch / 1e9
    log("kernel ms/step: " + ", ".join(f"{k} {v:.4f}" for k, v in kt.items()))
    kname = f"push_deposit_2d_t<{args.order},0,{tile[0]},{tile[1]},{waves}>"
    traffic = None
    if os.path.exists(args.traffic):
        try:
            with open(args.traffic) as f:
                tj = json.load(f)
            # counters measured on the same kernel instantiation and launch shape only
            if (tuple(tj.get("tile", ())) == tuple(tile) and tj.get("push_waves") == waves
                    and tj.get("species_per_launch") == nsp):
                traffic = tj.get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    fld, par = sim.history()
    relayouts = sim.relayouts
    sim.close()

    plain = None
    if args.plain_steps > 0:
        ps = make(cfg, plain=True)
        ps.run(2)
        t0 = time.perf_counter()
        ps.run(args.plain_steps)
        pel = time.perf_counter() - t0
        pkt = ps.time_kernels(3)
        ps.close()
        plain = {"value": n_total * args.plain_steps / pel,
                 "ms_per_step": pel * 1e3 / args.plain_steps,
                 "kernels_ms_per_step": pkt,
                 "speedup_fused_over_plain": (n_total * args.plain_steps / pel) and
                 value / (n_total * args.plain_steps / pel)}
        log(f"plain one-kernel-per-phase path: {plain['ms_per_step']:.3f} ms/step, "
            f"fused/plain = {plain['speedup_fused_over_plain']:.2f}x")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded counter-based RNG on the device)",
        "config": {
            "workload": (f"2D periodic counter-streaming electron beams {nx[0]}x{nx[1]}, "
                         f"dx=0.05, dt=0.035, {2 * ppc_beam} ppc (2 beams, u_x=+/-0.2, "
                         f"u_th=0.01), order {args.order} splines, Boris"),
            "particles": n_total,
            "cells": cells,
            "tile": list(tile),
            "push_waves": waves,
            "autotune_ms_per_step": tune,
            "relayouts": relayouts,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": kname + " (Boris; both species in one launch)",
            "kernel_ms": t_launch * 1e3,
            "algorithmic_bytes": bytes_launch,
        },
        "kernels_ms_per_step": kt,
        "kernels_roofline": {
            # algorithmic bytes per step (HBM-bound kernels); measured traffic of every
            # kernel is in profiles/kernels_c2.md
            "push_deposit": {"bytes": bytes_launch, "ms": kt["push_deposit"],
                             "GB/s": bytes_launch / kt["push_deposit"] / 1e6,
                             "frac": bytes_launch / kt["push_deposit"] / 1e6 / HBM_PEAK_GBS},
            "yee": {"bytes": cells * BYTES_PER_CELL_YEE, "ms": kt["yee"],
                    "GB/s": cells * BYTES_PER_CELL_YEE / kt["yee"] / 1e6,
                    "frac": cells * BYTES_PER_CELL_YEE / kt["yee"] / 1e6 / HBM_PEAK_GBS},
        },
        "plain_path": plain,
    }
    if not args.no_cpu:
        # at most 16 threads: the host cores allotted to one GPU (os.cpu_count() reports
        # the whole node)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cb = cpu_baseline(nx, 2 * ppc_beam, n_total, args.order, dx, dt, args.cpu_seconds,
                          threads)
        if cb is not None:
            cb["oracle_check"] = oracle_check(oa)
            log("oracle check: " + json.dumps(cb["oracle_check"]))
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
