"""HIST and ZDF output of a resident simulation on the GPU: the files hold exactly the
device history and fields (ES23.16 prints 17 significant digits, which round-trip every
double; the ZDF dataset is the raw float32 plane)."""
import os
import struct

import numpy as np
import pytest

import osiris_amd as oa
from tests.pic_inputs import thermal_plasma

pytestmark = pytest.mark.gpu


def _zdf_records(path):
    """Minimal reader of a ZDF file: list of (id, name, payload bytes)."""
    b = open(path, "rb").read()
    assert b[:4] == b"ZDF1"
    i, out = 4, []
    while i < len(b):
        rid, = struct.unpack_from("<I", b, i)
        n, = struct.unpack_from("<I", b, i + 4)
        name = b[i + 8:i + 8 + n].decode()
        i += 8 + (n + 3) // 4 * 4
        length, = struct.unpack_from("<Q", b, i)
        i += 8
        out.append((rid, name, b[i:i + length]))
        i += length
    return out


def test_hist_and_field_dump(tmp_path):
    rng = np.random.default_rng(70)
    nx, dx, dt = (40, 24), (0.1, 0.1), 0.07
    with oa.Simulation(2, nx, dx, dt, 2) as sim:
        for k in range(2):
            ix, x, u, q = thermal_plasma(rng, 2, nx, 4, 0.05 * (k + 1))
            sid = sim.add_species(-1.0)
            sim.set_particles(sid, ix, x, u, q)
        sim.run(12)
        fld, par = sim.history()
        sim.write_hist(tmp_path / "HIST", names=["electrons", "hot electrons"])
        e, b = sim.get_fields()
        sim.save_field(0, tmp_path / "MS", xmin=(0.0, -1.2))
        sim.save_field(5, tmp_path / "MS")
    rows = [l.split() for l in (tmp_path / "HIST" / "fld_ene").read_text().splitlines()[2:]]
    assert len(rows) == 12
    got = np.array([[float(v) for v in r[2:]] for r in rows])
    assert np.array_equal(got, fld)
    assert [int(r[0]) for r in rows] == list(range(12))
    prow = [l.split() for l in (tmp_path / "HIST" / "par02_ene").read_text().splitlines()[2:]]
    assert np.array_equal(np.array([float(r[3]) for r in prow]), par[:, 1])
    assert all(int(r[2]) == 40 * 24 * 4 for r in prow)
    recs = _zdf_records(tmp_path / "MS" / "FLD" / "e1" / "e1-000012.zdf")
    kinds = [r[0] for r in recs]
    assert kinds == [0x00030000, 0x00210001, 0x00200001, 0x00110000, 0x00120000, 0x00130000]
    chunk = recs[4][2]
    data = np.frombuffer(chunk[4 + 24 * 2:], np.float32).reshape(nx[1], nx[0])
    assert np.array_equal(data, e[..., 0])
    it = recs[2][2]
    n, t = struct.unpack_from("<id", it, 0)
    assert n == 12 and t == 12 * dt
    recs_b = _zdf_records(tmp_path / "MS" / "FLD" / "b3" / "b3-000012.zdf")
    data = np.frombuffer(recs_b[4][2][4 + 48:], np.float32).reshape(nx[1], nx[0])
    assert np.array_equal(data, b[..., 2])
