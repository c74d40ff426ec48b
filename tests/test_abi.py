"""CPU checks of the C ABI: the library loads without a GPU, exports every function the
header declares, the ctypes signatures cover them, and calls fail cleanly (status, no
crash) when no GPU is present or arguments are invalid."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import osiris_amd as oa
from osiris_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "osiris_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(?:int|void|const char\s*\*)\s*\*?\s*(\w+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_expected_entry_points():
    names = header_functions()
    for n in ("osr_advance_deposit", "osr_sim_create", "osr_sim_run", "osr_yee_advance",
              "hip_vadvance_deposit_2d_s2_", "hip_vadvance_deposit_3d_s3_"):
        assert n in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    L = oa.lib()
    for n in header_functions():
        assert hasattr(L, n), n
    assert set(header_functions()) == set(_lib.SIGNATURES)


def test_errors_without_gpu_are_status_codes():
    L = oa.lib()
    if L.osr_device_count() > 0:
        pytest.skip("GPU present")
    d = _lib.SimDesc()
    d.dims = 2
    d.nx[0] = d.nx[1] = 16
    d.dx[0] = d.dx[1] = 0.1
    d.dt = 0.05
    d.order = 2
    h = C.c_void_p()
    st = L.osr_sim_create(C.byref(d), C.byref(h))
    assert st == 2  # OSR_ERR_HIP
    assert "GPU" in L.osr_last_error().decode()


def test_argument_validation():
    L = oa.lib()
    d = _lib.SimDesc()
    d.dims = 4
    h = C.c_void_p()
    assert L.osr_sim_create(C.byref(d), C.byref(h)) == 1
    d.dims = 2
    d.nx[0] = d.nx[1] = 16
    d.dx[0] = d.dx[1] = 0.1
    d.order = 2
    d.dt = 0.2  # violates the Courant condition of the Yee scheme
    assert L.osr_sim_create(C.byref(d), C.byref(h)) == 1
    assert "stability" in L.osr_last_error().decode()
    assert L.osr_advance_deposit(5, 2, 0, *([None] * 4), 1, 0, -1.0, *([None] * 8), 0.1,
                                 None) == 1


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of the header's structs have the C sizes and field offsets."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"osr_sim_desc": _lib.SimDesc, "osr_ens_desc": _lib.EnsDesc,
               "osr_job_species": _lib.JobSpecies, "osr_job_desc": _lib.JobDesc,
               "osr_job_info": _lib.JobInfo, "osr_zpulse_wall": _lib.ZpulseWall,
               "osr_zpulse": _lib.ZpulseBox}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "osiris_amd.h"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = {}
    for ln in out.split("\n"):
        if ln:
            s, f, v = ln.split()
            got[(s, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_noalloc_rejects_oversized_current_before_touching_the_gpu():
    """osr_advance_deposit_noalloc refuses work whose current does not fit its fixed device
    array, with a status and a message, before any device call (so this runs without a
    GPU; the pointers are never dereferenced)."""
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    big = np.array([300, 300, 300], np.int32)  # 3 * 300^3 int64 values: more than 64 MiB
    off = np.array([2, 2, 2], np.int32)
    dxa = np.array([0.1] * 3, np.float64)
    ene = np.zeros(1, np.float64)
    dummy = C.c_void_p(16)
    st = oa.lib().osr_advance_deposit_noalloc(3, 2, oa.OSR_BORIS, dummy, dummy, dummy, dummy, 1,
                                              10, -1.0, dummy, dummy, P(big), P(off), dummy,
                                              P(big), P(off), P(dxa), 0.04, P(ene))
    assert st == 3  # OSR_ERR_NOMEM
    assert "more than its fixed" in oa.lib().osr_last_error().decode()
    assert ene[0] == 0.0
