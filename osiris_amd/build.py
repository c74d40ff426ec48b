"""Build recipe of the native parts (run from anywhere; everything is built in-tree).

* ``osiris_amd/lib/libosiris_amd.so``: the HIP kernels and the C ABI, compiled by hipcc
  for gfx950 as one translation unit.
* ``osiris_amd/lib/libosiris_amd_f.so``: the ISO_C_BINDING Fortran module over the C ABI
  (only when a Fortran compiler is present).
"""
from __future__ import annotations

import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # every fused multiply-add in the kernels is written explicitly
    "-ffp-contract=off",
    # IEEE division and square root, as the reference's AVX2 kernels
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_hip(verbose: bool = False) -> str:
    src = os.path.join(ROOT, "osiris_amd", "csrc", "capi.hip")
    out_dir = os.path.join(ROOT, "osiris_amd", "lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libosiris_amd.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_fortran(verbose: bool = False) -> str | None:
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None:
        return None
    fdir = os.path.join(ROOT, "osiris_amd", "fortran")
    out_dir = os.path.join(ROOT, "osiris_amd", "lib")
    mod_dir = os.path.join(out_dir, "mod")
    os.makedirs(mod_dir, exist_ok=True)
    src = os.path.join(fdir, "m_osiris_amd.f03")
    if not os.path.exists(src):
        return None
    out = os.path.join(out_dir, "libosiris_amd_f.so")
    cmd = [fc, "-O2", "-fPIC", "-shared", "-J", mod_dir, "-o", out, src,
           "-L" + out_dir, "-losiris_amd", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    for name in ("push_example", "smooth_setup_example", "zpulse_layout_example"):
        drv = os.path.join(fdir, name + ".f03")
        if not os.path.exists(drv):
            continue
        exe = os.path.join(out_dir, name)
        cmd = [fc, "-O2", "-I", mod_dir, "-o", exe, drv, "-L" + out_dir, "-losiris_amd_f",
               "-losiris_amd", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return out


def build_all(verbose: bool = False):
    build_hip(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
