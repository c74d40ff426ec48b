"""GPU check of osr_advance_deposit_noalloc, the entry point that allocates no device
memory (its working memory is the library's fixed device array): on arrays already in
device memory it must give what osr_advance_deposit gives on the same host arrays --
cell indices, positions and momenta bit for bit, the current within 1e-6 of max|J| (the
fixed-point scale is set from the charge sum, which the two entry points add in a
different order; it is the same unless that sum lies within round-off of a power of two),
the energy within 1e-12 relative -- on two calls in a row, the second one smaller, so
that the second call works in what the first one left in the shared array."""
import ctypes as C

import numpy as np
import pytest

import osiris_amd as oa
from osiris_amd._lib import check, lib
from oracle import oracle as O
from tests.pic_inputs import random_fields, random_particles

pytestmark = pytest.mark.gpu


def _case(dims, order, nx, n, seed):
    rng = np.random.default_rng(seed)
    g = O.make_grid(dims, nx, order)
    e, b = random_fields(rng, g)
    lo, _ = O.guard_cells(order)
    ix, x, u, q = random_particles(rng, dims, nx, n)
    return g, e, b, ix, x, u, q, lo


class _Dev:
    """Device copies of host arrays through the HIP runtime (the test allocates; the entry
    point under test does not)."""

    def __init__(self):
        try:
            self.hip = C.CDLL("libamdhip64.so")
        except OSError:
            self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 4))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(dims, order, pusher, ix, x, u, q, e, b, lo, j, dx, dt):
    dev = _Dev()
    try:
        t = {k: dev.put(v) for k, v in dict(ix=ix, x=x, u=u, q=q, e=e, b=b, j=j).items()}
        emf_size = np.array(list(reversed(e.shape[:-1])), np.int32)
        j_size = np.array(list(reversed(j.shape[:-1])), np.int32)
        off = np.array([lo] * dims, np.int32)
        dxa = np.array(list(dx), np.float64)
        ene = np.zeros(1, np.float64)
        P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        check(lib().osr_advance_deposit_noalloc(dims, order, pusher, t["ix"], t["x"], t["u"],
                                                t["q"], 1, len(q), -1.0, t["e"], t["b"],
                                                P(emf_size), P(off), t["j"], P(j_size), P(off),
                                                P(dxa), float(dt), P(ene)))
        return (dev.get(t["ix"], ix), dev.get(t["x"], x), dev.get(t["u"], u), dev.get(t["j"], j),
                float(ene[0]))
    finally:
        dev.free()


@pytest.mark.parametrize("dims,order,pusher", [(2, 2, oa.OSR_BORIS), (2, 1, oa.OSR_VAY),
                                               (3, 3, oa.OSR_BORIS)])
def test_noalloc_matches_advance_deposit_twice(dims, order, pusher):
    dx = (0.1,) * dims
    dt = 0.05 if dims == 2 else 0.04
    sizes = [((64,) * dims if dims == 2 else (24,) * 3, 40000),
             ((20,) * dims if dims == 2 else (10,) * 3, 3000)]
    for call, (nx, n) in enumerate(sizes):
        print(f"case noalloc dims={dims} order={order} pusher={pusher} call={call} nx={nx} "
              f"n={n}", flush=True)
        g, e, b, ix, x, u, q, lo = _case(dims, order, nx, n, 1000 + 10 * dims + order + call)
        j0 = np.ascontiguousarray(np.random.default_rng(5).normal(0, 1e-3, g.zeros(3, np.float32)
                                                                  .shape).astype(np.float32))
        jr = j0.copy()
        rix, rx, ru, rene = oa.advance_deposit(dims, order, ix.copy(), x.copy(), u.copy(), q, -1.0,
                                               e, b, lo, jr, lo, dx, dt, pusher=pusher)
        nix, nx_, nu, nj, nene = _noalloc(dims, order, pusher, ix, x, u, q, e, b, lo, j0, dx, dt)
        assert np.array_equal(nix, rix)
        assert np.array_equal(nx_.view(np.uint32), rx.view(np.uint32))
        assert np.array_equal(nu.view(np.uint32), ru.view(np.uint32))
        assert np.abs(nj - jr).max() <= 1e-6 * np.abs(jr).max()
        assert abs(nene - rene) <= 1e-12 * abs(rene)
