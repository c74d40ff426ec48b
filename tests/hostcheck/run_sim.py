"""Runs single simulations through the library's host logic over the mock runtime and
prints one JSON line of results (invoked by tests/test_ensemble_host.py in a subprocess):
creation, particle generation, runs of several lengths through the captured step graphs,
device re-layouts after injected tile overflows, the per-kernel timing, history and
particle read-back, and destruction; periodic grids, and Lindman walls with a laser wall
(its table refilled before every batch of steps).  The mock's step end writes history row r, column c
as r + c, so the rows read back show whether every step was counted once."""
import ctypes as C
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import osiris_amd._lib as L  # noqa: E402


def main():
    lib_path, overflow_p, progress_p, seed = (sys.argv[1], float(sys.argv[2]),
                                              float(sys.argv[3]), int(sys.argv[4]))
    L.LIB_PATH = lib_path
    L.lib()
    raw = C.CDLL(lib_path)
    C.c_double.in_dll(raw, "g_mock_overflow_p").value = overflow_p
    C.c_double.in_dll(raw, "g_mock_progress_p").value = progress_p
    for f in ("mock_errors", "mock_order_errors", "mock_live"):
        getattr(raw, f).restype = C.c_longlong
    from osiris_amd.sim import Simulation

    rng = np.random.default_rng(seed)
    bad_rows = relayouts = steps = 0
    for nx, order, tile, xbnd in (((64, 48), 2, (16, 16), "periodic"),
                                  ((96, 64), 1, (32, 16), "periodic"),
                                  ((64, 64), 3, (16, 16), "periodic"),
                                  ((64, 32), 2, (16, 16), "lindman")):
        with Simulation(2, nx, (0.1, 0.1), 0.07, order, tile=tile, xbnd=xbnd) as s:
            for _ in range(2):
                sid = s.add_species(-1.0)
                s.generate(sid, (2, 2), density=1.0, uth=(0.05,) * 3,
                           seed=int(rng.integers(1, 1 << 30)))
            if xbnd != "periodic":  # a laser wall: its table is refilled before each batch
                s.add_zpulse_wall(a0=1.0, omega0=10.0, tenv_type="sin2", tenv_rise=2.0,
                                  tenv_flat=0.0, tenv_fall=2.0)
            for n in (5, 13, 40):
                s.run(n)
            s.time_kernels(3)
            fld, par = s.history()
            want = np.arange(s.steps)[:, None] + np.arange(8)[None, :]
            bad_rows += int(np.sum(np.any(np.hstack([fld, par]) != want[:, :6 + par.shape[1]],
                                          axis=1)))
            for k in range(2):
                s.get_particles(k)
            relayouts += s.relayouts
            steps += s.steps
    print(json.dumps({"steps": steps, "relayouts": relayouts, "bad_rows": bad_rows,
                      "errors": int(raw.mock_errors()), "order_errors": int(raw.mock_order_errors()),
                      "live": int(raw.mock_live())}))


if __name__ == "__main__":
    main()
