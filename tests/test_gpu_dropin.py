"""GPU parity of the stateless drop-in entry points.

The push/position arithmetic follows the reference's AVX2 single precision kernel
operation by operation, so cell indices, positions and momenta must be BIT-IDENTICAL to
the reference's own simd/os-spec-push-avx2.c compiled from source (oracle/_ref).  The
deposited current is summed in a different order (exact int64 fixed point on the GPU,
sequential float adds in the reference), so J is compared with a tolerance of 2e-6 of
max|J| (a few FP32 ulps of the node sums).
"""
import numpy as np
import pytest

import osiris_amd as oa
from oracle import oracle as O
from tests.pic_inputs import random_fields, random_particles

pytestmark = pytest.mark.gpu

CASES = [(d, o) for d in (2, 3) for o in (1, 2, 3)]


def _setup(dims, order, n, seed):
    rng = np.random.default_rng(seed)
    nx = (12,) * dims
    g = O.make_grid(dims, nx, order)
    e, b = random_fields(rng, g)
    lo, hi = O.guard_cells(order)
    ix, x, u, q = random_particles(rng, dims, nx, n)
    return g, e, b, ix, x, u, q, lo


@pytest.mark.parametrize("dims,order", CASES)
def test_boris_bitexact_vs_reference_simd(dims, order):
    if not O.ref_available("s"):
        pytest.skip("reference AVX2 kernels not built (oracle/_ref)")
    g, e, b, ix, x, u, q, lo = _setup(dims, order, 3000, 100 + 10 * dims + order)
    dx = (0.1,) * dims
    dt = 0.05 if dims == 2 else 0.04
    print(f"case dims={dims} order={order} n={len(q)}", flush=True)
    j = g.zeros(3, np.float32)
    gix, gx, gu, gene = oa.advance_deposit(dims, order, ix.copy(), x.copy(), u.copy(), q, -1.0,
                                           e, b, lo, j, lo, dx, dt)
    rix, rx, ru, rj, rene = O.ref_vadvance_deposit(dims, order, ix, x, u, q, -1.0, e, b, lo,
                                                   g.zeros(3, np.float32), lo, dx, dt, prec="s")
    assert np.array_equal(gix, rix)
    assert np.array_equal(gx.view(np.uint32), rx.view(np.uint32))
    assert np.array_equal(gu.view(np.uint32), ru.view(np.uint32))
    if dims == 3 and order == 3:
        # the reference's FP32 AVX2 3D cubic deposit strides z planes by Dy
        # (simd/os-spec-current-avx2.c:1097); compare J with its FP64 twin instead
        _, _, _, rj, _ = O.ref_vadvance_deposit(dims, order, ix, x.astype(np.float64),
                                                u.astype(np.float64), q.astype(np.float64),
                                                -1.0, e.astype(np.float64), b.astype(np.float64),
                                                lo, g.zeros(3, np.float64), lo, dx, dt, prec="d")
    scale = np.abs(rj).max()
    assert np.abs(j - rj).max() <= 2e-6 * scale
    assert abs(gene - rene) <= 1e-6 * abs(rene)


@pytest.mark.parametrize("dims,order", CASES)
def test_vay_vs_oracle(dims, order):
    """Vay has no SIMD reference kernel: compare with the FP32 and FP64 restatements of
    spec/os-spec-push.f03:636-796."""
    g, e, b, ix, x, u, q, lo = _setup(dims, order, 3000, 200 + 10 * dims + order)
    dx = (0.1,) * dims
    dt = 0.05 if dims == 2 else 0.04
    j = g.zeros(3, np.float32)
    gix, gx, gu, gene = oa.advance_deposit(dims, order, ix.copy(), x.copy(), u.copy(), q, -1.0,
                                           e, b, lo, j, lo, dx, dt, pusher=oa.OSR_VAY)
    e64, b64 = e.astype(np.float64), b.astype(np.float64)
    oix, ox, ou, oq = ix.copy(), x.astype(np.float64), u.astype(np.float64), q.astype(np.float64)
    oj = g.zeros(3)
    oene = O.advance_deposit(g, order, 1, -1.0, dx, dt, e64, b64, oj, oix, ox, ou, oq)
    # identical cells except particles that end within FP32 rounding of a cell edge
    bad = np.any(gix != oix, axis=1)
    near = np.any(np.abs(np.abs(ox) - 0.5) < 1e-5, axis=1) | np.any(np.abs(gx) > 0.4999, axis=1)
    assert not np.any(bad & ~near)
    assert np.abs(gu - ou).max() <= 2e-6 * max(1.0, np.abs(ou).max())
    assert np.abs(j - oj).max() <= 1e-5 * np.abs(oj).max()
    assert abs(gene - oene) <= 1e-5 * abs(oene)


def test_fortran_named_dropin_matches_core():
    """hip_vadvance_deposit_2d_s2_ (all arguments by reference) = osr_advance_deposit."""
    import ctypes as C

    g, e, b, ix, x, u, q, lo = _setup(2, 2, 1000, 5)
    dx = np.array([0.1, 0.1])
    dt = C.c_double(0.05)
    j1 = g.zeros(3, np.float32)
    a = oa.advance_deposit(2, 2, ix.copy(), x.copy(), u.copy(), q, -1.0, e, b, lo, j1, lo,
                           tuple(dx), 0.05)
    ix2, x2, u2 = ix.copy(), x.copy(), u.copy()
    j2 = g.zeros(3, np.float32)
    ef, bf = np.ascontiguousarray(e, np.float32), np.ascontiguousarray(b, np.float32)
    size = np.array([e.shape[1], e.shape[0]], np.int32)
    off = np.array([lo, lo], np.int32)
    i0, i1 = C.c_int(1), C.c_int(len(q))
    rqm = C.c_float(-1.0)
    ene = C.c_double(0.0)
    p = lambda arr: arr.ctypes.data_as(C.c_void_p)  # noqa: E731
    oa.lib().hip_vadvance_deposit_2d_s2_(p(ix2), p(x2), p(u2), p(q), C.byref(i0), C.byref(i1),
                                         C.byref(rqm), p(ef), p(bf), p(size), p(off), p(j2),
                                         p(size), p(off), p(dx), C.byref(dt), C.byref(ene))
    assert np.array_equal(a[0], ix2) and np.array_equal(a[1], x2) and np.array_equal(a[2], u2)
    assert np.array_equal(j1, j2)
    assert ene.value == a[3]


@pytest.mark.parametrize("dims", (2, 3))
def test_yee_dropin_vs_oracle(dims):
    rng = np.random.default_rng(11 + dims)
    nx = (16, 12, 10)[:dims]
    g = O.make_grid(dims, nx, 1)
    e, b = random_fields(rng, g, 1.0, 1.0, np.float32)
    jay = g.zeros(3, np.float32)
    jay[...] = rng.normal(0, 0.1, jay.shape)
    O.update_periodic(g, jay, 0, "f32")
    dx = (0.1, 0.12, 0.15)[:dims]
    dt = 0.04
    ge, gb = e.copy(), b.copy()
    oa.yee_advance(dims, nx, g.gcl, dx, dt, ge, gb, jay)
    oe, ob = e.copy(), b.copy()
    O.update_periodic(g, oe, 0, "f32")
    O.update_periodic(g, ob, 0, "f32")
    O.yee_advance(g, dx, dt, oe, ob, jay, prec="f32")
    # same FP32 operation order, no contraction: bit-identical
    assert np.array_equal(ge, oe)
    assert np.array_equal(gb, ob)
