"""Runs a serving stream through the library's host logic over the mock runtime and
prints one JSON line of results (invoked by tests/test_ensemble_host.py in a subprocess,
so that the mock library and the real one never share a process)."""
import ctypes as C
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import osiris_amd._lib as L  # noqa: E402


def bursts(stream, seed):
    """Submissions in random bursts between random numbers of scheduling passes (seeded):
    admissions land at every point of the chunk cycle, independently of the wall clock."""
    from osiris_amd.ensemble import Ensemble

    rng = np.random.default_rng(seed)
    ens = Ensemble(0, 0, 8)
    ids, nxt = [], 0
    try:
        while True:
            for _ in range(int(rng.integers(0, 4))):
                if nxt < len(stream):
                    ids.append(ens.submit(stream[nxt][1]))
                    nxt += 1
            for _ in range(int(rng.integers(1, 4))):
                active, pending = ens.progress()
            if nxt == len(stream) and active == 0 and pending == 0:
                break
        return {"jobs": len(ids), "histories": {k: ens.history(j) for k, j in enumerate(ids)}}
    finally:
        ens.close()


def main():
    lib_path, n, overflow_p, notready_p = sys.argv[1], int(sys.argv[2]), float(sys.argv[3]), float(sys.argv[4])
    progress_p = float(sys.argv[7]) if len(sys.argv) > 7 else 0.3
    realtime = len(sys.argv) > 8 and sys.argv[8] == "realtime"
    sides = tuple(int(v) for v in sys.argv[5].split(","))
    ppcs = tuple(int(v) for v in sys.argv[6].split(","))
    L.LIB_PATH = lib_path
    L.lib()
    raw = C.CDLL(lib_path)
    C.c_double.in_dll(raw, "g_mock_overflow_p").value = overflow_p
    C.c_double.in_dll(raw, "g_mock_notready_p").value = notready_p
    C.c_double.in_dll(raw, "g_mock_progress_p").value = progress_p
    raw.mock_errors.restype = C.c_longlong
    raw.mock_order_errors.restype = C.c_longlong
    raw.mock_live.restype = C.c_longlong
    from osiris_amd.ensemble import poisson_jobs, serve

    stream = poisson_jobs(n=n, sides=sides, ppcs=ppcs, steps=(20, 400))
    if len(sys.argv) > 8 and sys.argv[8] == "bursts":
        res = bursts(stream, int(sys.argv[9]) if len(sys.argv) > 9 else 1)
    else:
        if realtime:  # arrivals compressed 100x so that admissions land between chunks
            stream = [(t / 100.0, j) for t, j in stream]
        res = serve(stream, realtime=realtime, chunk_steps=8, keep=True)
    bad_hist = 0
    for k, (fld, par) in res["histories"].items():
        nsteps = stream[k][1].nsteps
        rows = np.arange(nsteps, dtype=float)
        # the emulated finalisation writes step + column into each history row
        if fld.shape != (nsteps, 6) or not np.array_equal(fld[:, 0], rows) or \
                not np.array_equal(par[:, 0], rows + 6):
            bad_hist += 1
    print(json.dumps({"jobs": res["jobs"], "errors": raw.mock_errors(),
                      "order_errors": raw.mock_order_errors(), "live": raw.mock_live(),
                      "bad_histories": bad_hist}))


if __name__ == "__main__":
    main()
