"""The Fortran host binding (osiris_amd/fortran/m_osiris_amd.f03) driving the GPU path:
the drop-in push called the way spec/os-spec-vpush.f03:144-169 calls the reference's
native kernel, the periodic Yee advance, and a resident run."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "osiris_amd", "lib", "push_example")


@pytest.mark.gpu
def test_fortran_push_example():
    if not os.path.exists(EXE):
        pytest.skip("Fortran example not built (no Fortran compiler at build time)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "push_example: OK" in r.stdout
