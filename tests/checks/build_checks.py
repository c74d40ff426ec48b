"""Builds the device check library of tests/checks (test infrastructure) with the
product's HIP flags: tests/checks/_build/librn_check.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "_build", "librn_check.so")


def build(verbose: bool = False) -> str:
    from osiris_amd.build import HIPCC, HIP_FLAGS

    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp,
           os.path.join(HERE, "rn_check.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
