"""Ensemble serving (continuous batching) on the GPU: every job's result is the same
alone, in a mix admitted at different times, and run as a plain Simulation; admission
respects the memory budget; the serving loop reports latencies."""
import numpy as np
import pytest

from osiris_amd.ensemble import Ensemble, Job, SpeciesSpec, poisson_jobs, run_sequential, serve
from osiris_amd._lib import OsrError

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="ensemble execution is switched off: a 256-job mixed serving run "
                            "ended in a GPU memory fault; the two causes found (uploads on the "
                            "null stream not ordered with the batch streams, a missing barrier "
                            "in the tile fill) are fixed, but the fixes have not had their GPU "
                            "run yet; the scheduler is covered on the CPU by "
                            "tests/test_ensemble_host.py"),
]


def _job(side, order, ppc, steps, seed, uth=0.05, drift=0.0, kernel=0):
    sp = SpeciesSpec(rqm=-1.0, ppc=(ppc, ppc), density=1.0, ufl=(drift, 0.0, 0.0),
                     uth=(uth, uth, uth), seed=seed)
    return Job(nx=(side, side), dx=(0.1, 0.1), dt=0.07, order=order, nsteps=steps,
               species=[sp], kernel=kernel)


def _mix():
    return [
        _job(64, 1, 2, 40, 11),
        _job(128, 2, 4, 75, 12, drift=0.1),
        _job(96, 2, 2, 33, 13, uth=0.2),
        _job(64, 2, 3, 120, 14),
        _job(160, 1, 2, 50, 15, uth=0.01),
        _job(128, 2, 4, 75, 16),
    ]


def _run_alone(job):
    with Ensemble() as ens:
        jid = ens.submit(job)
        ens.wait()
        return ens.history(jid)


def test_ensemble_alone_equals_mix_equals_simulation():
    jobs = _mix()
    target = 1
    fa, pa = _run_alone(jobs[target])
    # mix: half the jobs up front, the rest submitted while the first chunks run
    with Ensemble(chunk_steps=8) as ens:
        ids = [ens.submit(j) for j in jobs[:3]]
        ens.progress()
        ens.progress()
        ids += [ens.submit(j) for j in jobs[3:]]
        ens.wait()
        hist = [ens.history(i) for i in ids]
        infos = [ens.info(i) for i in ids]
    for j, inf, (f, p) in zip(jobs, infos, hist):
        assert inf["steps"] == j.nsteps
        assert f.shape == (j.nsteps, 6) and np.all(np.isfinite(f)) and np.all(p > 0)
    fm, pm = hist[target]
    assert np.array_equal(fa, fm) and np.array_equal(pa, pm)
    # the same kernels driven by a single simulation
    seq = run_sequential([(0.0, jobs[target])], keep=True)
    fs, ps = seq["histories"][0]
    assert np.array_equal(fa, fs) and np.array_equal(pa, ps)


def test_ensemble_budget_bounds_admission():
    jobs = [_job(128, 2, 4, 30, 100 + k) for k in range(6)]
    with Ensemble() as probe:
        jid = probe.submit(jobs[0])
        probe.progress()
        one = probe.info(jid)["footprint"]
        probe.wait()
    budget = int(2.5 * one)
    peak = 0
    with Ensemble(hbm_budget=budget, chunk_steps=4) as ens:
        ids = [ens.submit(j) for j in jobs]
        while True:
            a, p = ens.progress()
            peak = max(peak, ens.usage)
            assert a <= 2
            if a == 0 and p == 0:
                break
        assert peak <= budget
        for i in ids:
            inf = ens.info(i)
            assert inf["steps"] == 30
            assert inf["t_admit"] >= inf["t_submit"] and inf["t_done"] > inf["t_admit"]
        # too large for the budget: rejected at submission, the server goes on
        with pytest.raises(OsrError):
            ens.submit(_job(512, 2, 8, 10, 1))


def test_ensemble_serve_reports_latency():
    stream = poisson_jobs(n=12, rate=40.0, seed=7, sides=(64, 128), ppcs=(4, 16),
                          steps=(20, 60))
    res = serve(stream, realtime=True)
    assert res["jobs"] == 12
    assert 0 < res["latency_p50_s"] <= res["latency_p99_s"] <= res["latency_max_s"]
    assert res["particle_updates_per_s"] > 0
