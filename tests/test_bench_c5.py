"""bench.py --workload c5 (ensemble serving) on the CPU with a stand-in for the ensemble
module: the JSON line carries the real-time p50/p99 latencies, the saturated throughput
and the ratio to sequential runs; when the library refuses ensembles the line carries
its message instead of a number and the status is 1."""
import json
import types

import bench


class _Job:
    def __init__(self, n, steps):
        self.particles, self.nsteps = n, steps


def _fake(refuse=False):
    stream = [(0.05 * k, _Job(1000 * (k + 1), 10 + k)) for k in range(4)]

    def serve(s, realtime=True):
        if refuse:
            raise RuntimeError("osiris_amd error 4: ensemble execution is switched off")
        return {"wall_s": 2.0 if realtime else 1.0, "jobs_per_s": len(s) / 1.0,
                "particle_updates_per_s": 5.0e6, "latency_p50_s": 0.25,
                "latency_p99_s": 0.75}

    return types.SimpleNamespace(
        poisson_jobs=lambda n: stream[:n], serve=serve,
        run_sequential=lambda s: {"jobs_per_s": len(s) / 4.0, "particle_updates_per_s": 1.2e6})


def test_c5_line(capsys):
    rc = bench.run_c5(types.SimpleNamespace(jobs=4), ens=_fake())
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert rc == 0 and line["value"] == 5.0e6 and line["unit"] == "particle-updates/s"
    e = line["ensemble"]
    assert e["latency_p50_s"] == 0.25 and e["latency_p99_s"] == 0.75
    assert e["jobs_per_s"] == 4.0 and e["sequential_jobs_per_s"] == 1.0
    assert e["speedup_vs_sequential"] == 4.0
    assert line["config"]["jobs"] == 4
    assert line["config"]["particle_updates"] == sum(1000 * (k + 1) * (10 + k) for k in range(4))


def test_c5_refused(capsys):
    rc = bench.run_c5(types.SimpleNamespace(jobs=4), ens=_fake(refuse=True))
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert rc == 1 and "switched off" in line["error"] and "value" not in line
