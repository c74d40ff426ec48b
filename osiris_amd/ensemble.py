"""Ensemble serving of independent simulation jobs on one GPU (continuous batching).

``Ensemble`` wraps the osr_ens_* entry points: jobs are submitted at any time, admitted
in submission order while their device footprint fits the budget, and advanced together
with the other admitted jobs of the same kind (dims, order, tile shape): one kernel
launch per phase and time step covers every job, each step being a captured hipGraph.
A job's results are bit-identical to running it alone.

``poisson_jobs`` draws the stream of 2D periodic thermal-plasma jobs of the serving
workload; ``serve`` replays such a stream in real time and reports job latencies
(completion - arrival) and throughput; ``run_sequential`` runs the same jobs one at a
time through ``Simulation`` for comparison.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import time

import numpy as np

from ._lib import (OSR_BORIS, OSR_JOB_DONE, OSR_JOB_FAILED, EnsDesc, JobDesc, JobInfo,
                   OsrError, check, lib)


@dataclasses.dataclass
class SpeciesSpec:
    rqm: float = -1.0
    pusher: int = OSR_BORIS
    ppc: tuple = (2, 2)
    density: float = 1.0
    ufl: tuple = (0.0, 0.0, 0.0)
    uth: tuple = (0.05, 0.05, 0.05)
    seed: int = 1
    random_pos: bool = True


@dataclasses.dataclass
class Job:
    nx: tuple
    dx: tuple
    dt: float
    order: int
    nsteps: int
    species: list
    tile: tuple = (0, 0, 0)
    e0: tuple = (0.0, 0.0, 0.0)
    b0: tuple = (0.0, 0.0, 0.0)
    kernel: int = 0

    @property
    def dims(self):
        return len(self.nx)

    @property
    def particles(self):
        ncell = int(np.prod(self.nx))
        return sum(ncell * int(np.prod(s.ppc[: self.dims])) for s in self.species)

    def to_c(self) -> JobDesc:
        j = JobDesc()
        d = j.sim
        d.dims = self.dims
        for k in range(self.dims):
            d.nx[k] = int(self.nx[k])
            d.dx[k] = float(self.dx[k])
            d.tile[k] = int(self.tile[k]) if k < len(self.tile) else 0
        d.dt = float(self.dt)
        d.order = int(self.order)
        d.path = 0
        d.kernel = int(self.kernel)
        j.nsteps = int(self.nsteps)
        for k in range(3):
            j.e0[k] = float(self.e0[k])
            j.b0[k] = float(self.b0[k])
        if not 1 <= len(self.species) <= 8:
            raise ValueError("1..8 species")
        j.nspecies = len(self.species)
        for i, s in enumerate(self.species):
            c = j.species[i]
            c.rqm = float(s.rqm)
            c.pusher = int(s.pusher)
            for k in range(3):
                c.ppc[k] = int(s.ppc[k]) if k < len(s.ppc) else 1
                c.ufl[k] = float(s.ufl[k])
                c.uth[k] = float(s.uth[k])
            c.density = float(s.density)
            c.seed = int(s.seed)
            c.random_pos = 1 if s.random_pos else 0
        return j


class Ensemble:
    """Continuous-batching job server on the first GPU."""

    def __init__(self, hbm_budget: int = 0, max_active: int = 0, chunk_steps: int = 0):
        d = EnsDesc()
        d.hbm_budget = int(hbm_budget)
        d.max_active = int(max_active)
        d.chunk_steps = int(chunk_steps)
        self._h = C.c_void_p()
        check(lib().osr_ens_create(C.byref(d), C.byref(self._h)))
        self._nspec = {}

    def close(self):
        if self._h:
            check(lib().osr_ens_destroy(self._h))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, job: Job) -> int:
        jid = C.c_int64()
        check(lib().osr_ens_submit(self._h, C.byref(job.to_c()), C.byref(jid)))
        self._nspec[jid.value] = len(job.species)
        return jid.value

    def progress(self):
        """One non-blocking scheduling pass; returns (active, pending) job counts."""
        a, p = C.c_int(), C.c_int()
        check(lib().osr_ens_progress(self._h, C.byref(a), C.byref(p)))
        return a.value, p.value

    def wait(self):
        check(lib().osr_ens_wait(self._h))

    def info(self, jid: int) -> dict:
        i = JobInfo()
        check(lib().osr_ens_job_info(self._h, int(jid), C.byref(i)))
        return {f: getattr(i, f) for f, _ in JobInfo._fields_}

    def history(self, jid: int):
        """(fld[nsteps, 6], par[nsteps, nspecies]) of a finished job."""
        inf = self.info(jid)
        n, ns = inf["steps"], inf["nspecies"]
        fld = np.zeros((n, 6))
        par = np.zeros((n, ns))
        check(lib().osr_ens_job_history(self._h, int(jid), 0, n,
                                        fld.ctypes.data_as(C.c_void_p),
                                        par.ctypes.data_as(C.c_void_p)))
        return fld, par

    def release(self, jid: int):
        check(lib().osr_ens_job_release(self._h, int(jid)))

    @property
    def usage(self) -> int:
        b = C.c_int64()
        check(lib().osr_ens_usage(self._h, C.byref(b)))
        return b.value


def poisson_jobs(n=256, rate=20.0, seed=2024, sides=(64, 128, 256, 512), ppcs=(4, 16, 64),
                 uth=(0.01, 0.2), orders=(1, 2), steps=(100, 2000), dx=0.1, dt=0.07):
    """The serving workload: n independent 2D periodic thermal electron plasmas arriving
    as a Poisson process of the given rate (jobs/s).  Per job: grid side uniform in
    `sides`, particles per cell uniform in `ppcs` (a square lattice count per cell,
    uniform-random in-cell positions), u_th log-uniform in `uth`, spline order uniform in
    `orders`, steps uniform in `steps`.  Returns [(arrival_s, Job)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        side = int(rng.choice(sides))
        ppc = int(rng.choice(ppcs))
        p1 = int(round(math.sqrt(ppc)))
        u = float(math.exp(rng.uniform(math.log(uth[0]), math.log(uth[1]))))
        order = int(rng.choice(orders))
        ns = int(rng.integers(steps[0], steps[1] + 1))
        sp = SpeciesSpec(rqm=-1.0, ppc=(p1, p1), density=1.0, uth=(u, u, u),
                         seed=int(rng.integers(1, 2**62)))
        out.append((t, Job(nx=(side, side), dx=(dx, dx), dt=dt, order=order, nsteps=ns,
                           species=[sp])))
    return out


def serve(stream, hbm_budget=0, max_active=0, chunk_steps=0, realtime=True, keep=False):
    """Replays [(arrival_s, Job)] through an Ensemble.  With realtime, each job is submitted
    when its arrival time has passed on the wall clock; otherwise all are submitted at
    once (saturation).  Returns a dict of latency / throughput statistics (and the
    per-job histories when keep is set)."""
    ens = Ensemble(hbm_budget, max_active, chunk_steps)
    try:
        order = sorted(range(len(stream)), key=lambda k: stream[k][0])
        ids = [None] * len(stream)
        nxt = 0
        t0 = time.perf_counter()
        arrival = [stream[k][0] if realtime else 0.0 for k in range(len(stream))]
        while True:
            now = time.perf_counter() - t0
            while nxt < len(order) and arrival[order[nxt]] <= now:
                k = order[nxt]
                ids[k] = ens.submit(stream[k][1])
                nxt += 1
            active, pending = ens.progress()
            if nxt == len(order) and active == 0 and pending == 0:
                break
            if active == 0 and pending == 0 and nxt < len(order):
                time.sleep(max(0.0, min(0.002, arrival[order[nxt]] - now)))
            else:
                time.sleep(0.0002)
        wall = time.perf_counter() - t0
        lat, upd, hist = [], 0.0, {}
        ens_t0 = None
        for k, jid in enumerate(ids):
            inf = ens.info(jid)
            if inf["state"] == OSR_JOB_FAILED:
                raise OsrError(f"job {k} failed")
            assert inf["state"] == OSR_JOB_DONE
            lat.append(inf["t_done"] - inf["t_submit"])
            upd += inf["particles"] * inf["steps"]
            if keep:
                hist[k] = ens.history(jid)
            ens_t0 = inf["t_submit"] if ens_t0 is None else min(ens_t0, inf["t_submit"])
        lat = np.array(lat)
        res = {
            "jobs": len(stream),
            "wall_s": wall,
            "jobs_per_s": len(stream) / wall,
            "particle_updates_per_s": upd / wall,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max()),
        }
        if keep:
            res["histories"] = hist
        return res
    finally:
        ens.close()


def run_sequential(stream, keep=False):
    """Runs each job alone, one after the other, through Simulation (graph-captured steps
    of the same kernels).  Returns the wall time, throughput and optional histories."""
    from .sim import Simulation

    hist = {}
    upd = 0.0
    t0 = time.perf_counter()
    for k, (_, job) in enumerate(stream):
        tile = tuple(job.tile[: job.dims]) if any(job.tile) else None
        with Simulation(job.dims, job.nx, job.dx, job.dt, job.order, tile=tile,
                        kernel=job.kernel) as s:
            for sp in job.species:
                sid = s.add_species(sp.rqm, sp.pusher)
                s.generate(sid, sp.ppc, density=sp.density, ufl=sp.ufl, uth=sp.uth,
                           seed=sp.seed, random_pos=sp.random_pos)
            if any(job.e0) or any(job.b0):
                s.fill_fields(job.e0, job.b0)
            s.run(job.nsteps)
            upd += job.particles * job.nsteps
            if keep:
                hist[k] = s.history()
    wall = time.perf_counter() - t0
    res = {"jobs": len(stream), "wall_s": wall, "jobs_per_s": len(stream) / wall,
           "particle_updates_per_s": upd / wall}
    if keep:
        res["histories"] = hist
    return res
