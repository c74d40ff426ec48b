"""The ensemble scheduler's host logic on the CPU, over a validating mock of the HIP
runtime (tests/hostcheck): admission, batching, chunked graph replay, in-flight joins,
re-layouts after (injected) tile overflows and retirement must never hand a kernel a
descriptor that points to freed or too-small device memory, must release every device
allocation with the call that matches how it was made, and must return each job's full
history."""
import json
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HC = os.path.join(ROOT, "tests", "hostcheck")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def mock_lib():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out_dir = os.path.join(HC, "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libosiris_hostcheck.so")
    srcs = [os.path.join(HC, "capi_mock.hip"), os.path.join(HC, "mock_hip.h")] + [
        os.path.join(ROOT, "osiris_amd", "csrc", f)
        for f in os.listdir(os.path.join(ROOT, "osiris_amd", "csrc"))]
    if not os.path.exists(out) or os.path.getmtime(out) < max(os.path.getmtime(s) for s in srcs):
        # each process builds its own file and renames it into place, so parallel test
        # workers never load a library another one is still writing
        tmp = f"{out}.{os.getpid()}.tmp"
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC",
                        "-shared", "-rdynamic", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-o", tmp,
                        os.path.join(HC, "capi_mock.hip")], check=True)
        os.replace(tmp, out)
    return out


def _run(lib, n, overflow_p, notready_p, sides, ppcs, progress_p=0.3, mode="saturated",
         seed=1):
    r = subprocess.run([sys.executable, os.path.join(HC, "run_stream.py"), lib, str(n),
                        str(overflow_p), str(notready_p), sides, ppcs, str(progress_p), mode,
                        str(seed)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_ensemble_host_logic_with_relayouts(mock_lib):
    res = _run(mock_lib, 40, 0.01, 0.5, "64,128", "4,16")
    assert res == {"jobs": 40, "errors": 0, "order_errors": 0, "live": 0, "bad_histories": 0}


def test_ensemble_host_logic_large_jobs(mock_lib):
    res = _run(mock_lib, 24, 0.002, 0.3, "128,256", "4,16,64")
    assert res == {"jobs": 24, "errors": 0, "order_errors": 0, "live": 0, "bad_histories": 0}


def test_stream_order_checker_catches_planted_mistakes(mock_lib):
    """The mock runtime reports each planted stream-order mistake (a pool block freed while
    another stream still has queued work on it, a host-synchronous copy racing queued
    work, a pool block used by another stream before its allocation executed, a kernel on
    a stream reading what a null-stream copy just wrote)."""
    r = subprocess.run([sys.executable, "-c",
                        "import ctypes as C, sys; L = C.CDLL(sys.argv[1]); "
                        "L.mock_selftest_planted.restype = C.c_longlong; "
                        "print(L.mock_selftest_planted())", mock_lib],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert int(r.stdout.strip().splitlines()[-1]) == 4


def test_ensemble_host_logic_realtime_arrivals(mock_lib):
    """Arrivals spread over the run (admissions while chunks are queued) with slow stream
    progress, so that frees, re-layouts and graph rebuilds meet queued work."""
    res = _run(mock_lib, 48, 0.01, 0.7, "64,128", "4,16", progress_p=0.05, mode="realtime")
    assert res == {"jobs": 48, "errors": 0, "order_errors": 0, "live": 0, "bad_histories": 0}


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ensemble_host_logic_random_bursts(mock_lib, seed):
    """Submissions in seeded random bursts between scheduling passes, with injected tile
    overflows (device re-layouts, and host re-layouts where the repeated fill overflows
    again) and slow stream progress."""
    res = _run(mock_lib, 48, 0.02, 0.7, "64,128", "4,16", progress_p=0.05, mode="bursts",
               seed=seed)
    assert res == {"jobs": 48, "errors": 0, "order_errors": 0, "live": 0, "bad_histories": 0}


@pytest.mark.parametrize("seed", [1, 2])
def test_simulation_host_logic_with_relayouts(mock_lib, seed):
    """A single simulation's own path over the mock runtime (periodic, and Lindman walls
    with a laser wall): runs through the captured step graphs, device re-layouts after
    injected overflows, per-kernel timing, history and particle read-back.  No kernel or copy touches a dead or unordered block, every
    step is counted once (history rows), and all memory is returned."""
    r = subprocess.run([sys.executable, os.path.join(HC, "run_sim.py"), mock_lib, "0.02", "0.3",
                        str(seed)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["steps"] == 4 * (5 + 13 + 40 + 3), res  # runs + the timed steps, per sim
    assert res["relayouts"] > 0, res
    assert {k: res[k] for k in ("bad_rows", "errors", "order_errors", "live")} == {
        "bad_rows": 0, "errors": 0, "order_errors": 0, "live": 0}, res


def test_ensemble_host_logic_256_job_stream(mock_lib):
    """The serving workload's job count (256 Poisson arrivals; smaller grids and step
    counts than the GPU stream, so that the mock's host memory holds them): hundreds of
    admissions, retirements, re-layouts and batch rebuilds while other batches have queued
    work, with no stale or unordered access and every history complete."""
    res = _run(mock_lib, 256, 0.01, 0.5, "64,128", "4,16", progress_p=0.1, mode="realtime")
    assert res == {"jobs": 256, "errors": 0, "order_errors": 0, "live": 0, "bad_histories": 0}
