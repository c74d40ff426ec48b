"""The multiply-high division of the field kernels (osr::FastDiv in
osiris_amd/csrc/osr_types.hpp) equals the integer division: tests/checks/fastdiv_check.cpp,
built here with g++, checks every divisor up to 70000 (every grid extent the library
accepts) and 200000 larger ones up to 2^31 at the edge dividends (0, multiples of the
divisor and their neighbours, 2^32 - 1) and at random ones.  With OSR_FASTDIV_FULL=1 it
also checks every dividend below 2^32 for seven divisors (about a minute)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fastdiv_equals_integer_division(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "fastdiv_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "checks", "fastdiv_check.cpp")], check=True)
    args = [str(exe)] + (["full"] if os.environ.get("OSR_FASTDIV_FULL") == "1" else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
