"""The initial electrostatic field (init_fields -> es_solver_2d, emf/os-emf-es-solver.f03:
125-236) in the CPU oracle, against what the mathematics fixes.  The reference sums the
free-space field of line charges, charge * r / r^2 without the 1/(2 pi) of the 2D Green
function, so its field satisfies div E = 2 pi rho:

* a smooth charge blob: the discrete divergence at the cell corners equals 2 pi rho to
  the lattice error (1%), and the discrete curl vanishes;
* a point charge: E1 and E2 are exactly antisymmetric about it on the Yee lattice, and
  far away the field is the monopole Q r / r^2;
* a deliberately shifted lattice (the E1 offset applied to E2) fails the Gauss check.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

N, DX = 64, (0.1, 0.1)
DXA = np.array([DX[0], DX[1], 1.0])


def _solve(rho_int, prec="f64"):
    g = O.make_grid(2, (N, N), 2)
    L = O._lib(prec)
    r = g.zeros(1, L._real)
    g.interior(r)[..., 0] = rho_int
    ef = g.zeros(3, L._real)
    L.orc_es_solver_2d(C.byref(g.c()), O._p(DXA), O._p(r), O._p(ef))
    return g.interior(ef)


def _blob(sigma=4.0):
    x = np.arange(1, N + 1)
    X, Y = np.meshgrid(x, x, indexing="xy")
    c = N / 2 + 0.5
    return np.exp(-((X - c) ** 2 + (Y - c) ** 2) / (2 * sigma ** 2))


def _div(E):
    return ((E[1:, 1:, 0] - E[1:, :-1, 0]) / DX[0] + (E[1:, 1:, 1] - E[:-1, 1:, 1]) / DX[1])


def test_gauss_law_and_zero_curl():
    rho = _blob()
    E = _solve(rho)
    core = rho[1:, 1:] > 0.5
    ratio = _div(E)[core] / (2 * np.pi * rho[1:, 1:][core])
    assert np.all(np.abs(ratio - 1) < 0.01), (ratio.min(), ratio.max())
    curl = ((E[:-1, 1:, 1] - E[:-1, :-1, 1]) / DX[0] - (E[1:, :-1, 0] - E[:-1, :-1, 0]) / DX[1])
    assert np.abs(curl).max() < 1e-3 * np.abs(E).max() / DX[0]
    assert not E[..., 2].any()


def test_shifted_lattice_fails_gauss():
    """The same sum with E2 sampled at E1's offsets: the check above rejects it."""
    rho = _blob()
    x = np.arange(1, N + 1)
    ei, ri = np.meshgrid(x, x, indexing="ij")
    d = (x[:, None] - x[None, :]).astype(np.float64)  # target - source, per axis
    E1 = np.zeros((N, N))
    E2 = np.zeros((N, N))
    for (sj, si) in zip(*np.nonzero(rho > 1e-12)):
        q = rho[sj, si]
        a1 = (d[:, si] + 0.5) * DX[0]  # x offsets of E1 (x + 1/2)
        b1 = d[:, sj] * DX[1]
        r2 = a1[None, :] ** 2 + b1[:, None] ** 2
        E1 += q * a1[None, :] / r2
        E2 += q * b1[:, None] / r2  # wrong: E2 belongs at (x, y + 1/2)
    E = np.stack([E1, E2], -1) * DX[0] * DX[1]
    core = rho[1:, 1:] > 0.5
    ratio = _div(E)[core] / (2 * np.pi * rho[1:, 1:][core])
    assert np.abs(ratio - 1).max() > 0.05


def test_point_charge_symmetry_and_far_field():
    rho = np.zeros((N, N))
    s = 20  # source at x = y = 21 (1-based)
    rho[s, s] = 3.0
    E = _solve(rho)
    # E1 lives at x + 1/2: E1(s + k) = -E1(s - 1 - k) along x, even along y about y = s
    k = np.arange(0, 15)
    assert np.array_equal(E[s, s + k, 0], -E[s, s - 1 - k, 0])
    assert np.array_equal(E[s + k, s, 1], -E[s - 1 - k, s, 1])
    assert np.array_equal(E[s + 5, :, 0], E[s - 5, :, 0])
    # far field: E = Q dV r / r^2 exactly for one source (the sum has one term)
    Q = 3.0 * DX[0] * DX[1]
    j, i = 60, 50  # target E1 point (x + 1/2, y)
    rx, ry = (i - s + 0.5) * DX[0], (j - s) * DX[1]
    assert abs(E[j, i, 0] / (Q * rx / (rx * rx + ry * ry)) - 1) < 1e-12


def test_fp32_accumulation_tracks_fp64():
    rho = _blob(3.0)
    a = _solve(rho, "f64")
    b = _solve(rho.astype(np.float32), "f32")
    assert np.abs(b - a).max() < 2e-5 * np.abs(a).max()


@pytest.mark.parametrize("u1", [0.5, -2.0])
def test_beam_field_is_gauss_and_beta_e(u1):
    """es_solver_2d_beam: the field of charged sheets drifting along x.  After the
    solve, (E2(j) - E2(j-1)) / dy = rho(j) up to rounding between interior rows (discrete
    Gauss law along y), and B3 = beta * (E2(i) + E2(i+1)) / 2 with beta = u1 / gamma."""
    nx, ny = 16, 40
    g = O.make_grid(2, (nx, ny), 2)
    L = O._lib("f64")
    rng = np.random.default_rng(9)
    r = g.zeros(1, np.float64)
    rho = np.zeros((ny, nx))
    rho[10:30, :] = rng.uniform(0.5, 1.5, (20, nx))
    g.interior(r)[..., 0] = rho
    ef = g.zeros(3, np.float64)
    bf = g.zeros(3, np.float64)
    L.orc_es_solver_2d_beam(C.byref(g.c()), O._p(DXA), O._p(r), float(np.float32(u1)), 0,
                            O._p(ef), O._p(bf))
    E = g.interior(ef)
    B = g.interior(bf)
    # interior rows (the guard rows hold periodic images, which a finite column of sheets
    # does not have)
    gauss = (E[1:, :, 1] - E[:-1, :, 1]) / DX[1]
    assert np.allclose(gauss, rho[1:], rtol=0, atol=1e-12 * np.abs(rho).max())
    gamma = np.sqrt(float(np.float32(u1) ** 2) + 1.0)  # u1^2 in float32, the rest in FP64
    e2p = np.concatenate([E[:, 1:, 1], E[:, :1, 1]], axis=1)  # periodic x + 1
    want = np.float32(u1) / gamma * 0.5 * (E[..., 1] + e2p)
    assert np.allclose(B[..., 2], want, rtol=1e-12, atol=1e-14)
    assert not E[..., 0].any() and not E[..., 2].any() and not B[..., :2].any()
