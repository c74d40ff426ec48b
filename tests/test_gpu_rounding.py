"""The push path's reciprocal, quotient and square root sequences (rcp_rn, div_rn, sqrt_rn
in osiris_amd/csrc/osr_device.hpp) are correctly rounded: on the device each is compared
bit for bit with the IEEE operation the compiler emits under
-fhip-fp32-correctly-rounded-divide-sqrt, which the reference's AVX2 kernels also use
(_mm256_div_ps, _mm256_sqrt_ps).

* rcp_rn: every normal b whose reciprocal is normal, both signs (exhaustive).
* sqrt_rn: every x from 2^-100 to the largest float (exhaustive); the push path takes
  the root of 1 + u^2 >= 1 only.
* div_rn: every b mantissa against an even spread of 8192 a mantissas here; with
  OSR_RN_EXHAUSTIVE=1, every one of the 2^46 mantissa pairs.  A pair of mantissas stands
  for every pair of exponents whose quotient and remainder stay normal, since the
  sequence only scales with the exponents there.  Zeros, signs and a spread of exponents
  are checked as explicit pairs.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "checks", "_build", "librn_check.so")


@pytest.fixture(scope="module")
def rn():
    if not os.path.exists(LIB):
        pytest.fail("tests/checks/_build/librn_check.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    u64p = C.POINTER(C.c_ulonglong)
    lib.rn_check_rcp.argtypes = [C.c_uint32, C.c_uint64, C.c_int, u64p, u64p]
    lib.rn_check_sqrt.argtypes = [C.c_uint32, C.c_uint64, u64p, u64p]
    lib.rn_check_div.argtypes = [C.c_uint32, C.c_uint32, u64p, u64p]
    lib.rn_check_div_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, u64p, u64p]
    for f in (lib.rn_check_rcp, lib.rn_check_sqrt, lib.rn_check_div, lib.rn_check_div_pairs):
        f.restype = C.c_int
    return lib


def _call(f, *args):
    bad, first = C.c_ulonglong(), C.c_ulonglong()
    assert f(*args, C.byref(bad), C.byref(first)) == 0
    return bad.value, first.value


def _bits(x):
    return np.array([x], np.uint32).view(np.float32)[0]


def test_rcp_rn_every_normal_input(rn):
    # b from 2^-125 to just below 2^126: 1/b stays normal
    first = (2 << 23)
    n = (253 << 23) - first
    bad, idx = _call(rn.rn_check_rcp, first, n, 2)
    assert bad == 0, f"rcp_rn differs from 1/b at b = {_bits(first + idx)!r} ({bad} inputs)"


def test_sqrt_rn_every_input_from_2_to_the_minus_100(rn):
    # below about 2^-102 the remainder x - g^2 underflows and the sequence is not
    # correctly rounded; the push path takes the root of 1 + u^2 >= 1 only
    first = 27 << 23  # 2^-100
    n = (255 << 23) - first
    bad, idx = _call(rn.rn_check_sqrt, first, n)
    assert bad == 0, f"sqrt_rn differs from sqrtf at x = {_bits(first + idx)!r} ({bad} inputs)"


def test_div_rn_mantissa_pairs(rn):
    exhaustive = os.environ.get("OSR_RN_EXHAUSTIVE") == "1"
    total = 1 << 23
    chunk = 2048 if exhaustive else 1
    starts = range(0, total, chunk) if exhaustive else range(0, total, total // 8192)
    bad_all = 0
    for k, a0 in enumerate(starts):
        bad, idx = _call(rn.rn_check_div, a0, chunk)
        if bad:
            ia, ib = idx >> 23, idx & (total - 1)
            a, b = _bits(0x3F800000 | ia), _bits(0x3F800000 | ib)
            pytest.fail(f"div_rn({a!r}, {b!r}) differs from the IEEE quotient ({bad} pairs)")
        bad_all += bad
        if exhaustive and k % 256 == 0:
            print(f"div_rn: a mantissas {a0} .. {a0 + chunk} of {total}, 0 mismatches",
                  flush=True)
    assert bad_all == 0


def test_div_rn_zeros_signs_and_exponents(rn):
    rng = np.random.default_rng(11)
    n = 1 << 20
    # the push path's operand ranges and beyond: numerators from 2^-25 to 2^20 and zero
    # (the path divides only normal or zero numerators by div_rn), denominators from
    # 2^-24 to 2^20, all signs
    a = (rng.uniform(1, 2, n) * 2.0 ** rng.integers(-25, 21, n)).astype(np.float32)
    b = (rng.uniform(1, 2, n) * 2.0 ** rng.integers(-24, 21, n)).astype(np.float32)
    a *= np.where(rng.random(n) < 0.5, -1, 1).astype(np.float32)
    b *= np.where(rng.random(n) < 0.5, -1, 1).astype(np.float32)
    # +0 over either sign (-0 numerators give +0, not the IEEE -0: the path's numerators
    # are +-1/2 - x0, whose exact cancellation is +0, tem, 1 and 2, never -0)
    a[::7] = 0.0
    a[2::13] = b[2::13]  # exact quotient 1
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    bad, idx = _call(rn.rn_check_div_pairs, a.ctypes.data, b.ctypes.data, n)
    assert bad == 0, f"div_rn({a[idx]!r}, {b[idx]!r}) differs ({bad} pairs)"


if __name__ == "__main__":
    sys.exit(pytest.main([__file__, "-v", "-s"]))
